"""TEST INFRASTRUCTURE ONLY: ctypes wrapper of the FP64 CPU oracle.

Used by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg as
the checker.  The product (umwm_amd) never imports this module.
"""
import ctypes as C
import os
import subprocess

import numpy as np

from umwm_amd import abi

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "liboracle.so")

CFLAGS = ["-O2", "-ffp-contract=off", "-fno-fast-math", "-fopenmp", "-fPIC",
          "-std=c11", "-Wall", "-Wextra", "-shared"]


def build(force=False):
    """Compiles oracle/umwm_oracle.c into oracle/_build/liboracle.so."""
    src = os.path.join(HERE, "umwm_oracle.c")
    os.makedirs(os.path.dirname(LIB_PATH), exist_ok=True)
    if (not force and os.path.exists(LIB_PATH)
            and os.path.getmtime(LIB_PATH) >= os.path.getmtime(src)
            and os.path.getmtime(LIB_PATH) >= os.path.getmtime(os.path.join(HERE, "umwm_oracle.h"))):
        return LIB_PATH
    tmp = LIB_PATH + ".tmp"
    subprocess.check_call(["gcc", *CFLAGS, "-o", tmp, src, "-lm"])
    os.replace(tmp, LIB_PATH)
    return LIB_PATH


_lib = None


def lib():
    global _lib
    if _lib is None:
        build()  # no-op when the library is newer than its sources
        L = C.CDLL(LIB_PATH)
        P = C.POINTER
        L.oracle_create.restype = C.c_void_p
        L.oracle_create.argtypes = [P(abi.Params), P(abi.GridInput), P(abi.ForcingInput),
                                    C.c_char_p, C.c_int]
        L.oracle_free.argtypes = [C.c_void_p]
        L.oracle_dims.argtypes = [C.c_void_p, P(C.c_int)]
        L.oracle_field.restype = C.c_void_p
        L.oracle_field.argtypes = [C.c_void_p, C.c_char_p, P(C.c_long), P(C.c_int)]
        L.oracle_get_scalar.restype = C.c_double
        L.oracle_get_scalar.argtypes = [C.c_void_p, C.c_char_p]
        L.oracle_set_scalar.argtypes = [C.c_void_p, C.c_char_p, C.c_double]
        for fn in ("forcinginterpolate", "sin_d12", "sds_d12", "snl_d12", "s_ice",
                   "source", "propagation", "refraction", "update", "stress_atm", "stress_ocn",
                   "diag", "substep"):
            getattr(L, "oracle_" + fn).argtypes = [C.c_void_p]
        L.oracle_forcinginput.argtypes = [C.c_void_p, P(abi.ForcingInput)]
        L.oracle_stokes_drift.argtypes = [C.c_void_p, C.c_int]
        L.oracle_init_jonswap.argtypes = [C.c_void_p, P(abi.Jonswap)]
        L.oracle_run_dtg.argtypes = [C.c_void_p]
        L.oracle_run_dtg.restype = C.c_int
        L.oracle_set_threads.argtypes = [C.c_int]
        for fn in ("sheltering_coare35",):
            getattr(L, "oracle_" + fn).restype = C.c_double
            getattr(L, "oracle_" + fn).argtypes = [C.c_double]
        L.oracle_stress_tail.restype = C.c_double
        L.oracle_stress_tail.argtypes = [C.c_double, C.c_double]
        L.oracle_friction_velocity_skin.restype = C.c_double
        L.oracle_friction_velocity_skin.argtypes = [C.c_double] * 4
        L.oracle_drag_coefficient_skin.restype = C.c_double
        L.oracle_drag_coefficient_skin.argtypes = [C.c_double] * 6
        L.oracle_distance_haversine.restype = C.c_double
        L.oracle_distance_haversine.argtypes = [C.c_double] * 4
        _lib = L
    return _lib


class Oracle:
    """One oracle model.  Fields come back as numpy views in reference shape:
    spectra (im, pm, om) indexed [i-1, p-1, o-1] (ghost dropped), tables
    (im, om), per-point (im,), unrolled (imm,), 2-D (nm, mm)."""

    def __init__(self, params, grid=None, forcing=None, threads=1):
        L = lib()
        L.oracle_set_threads(int(threads))
        g, self._gk = grid if grid is not None else (abi.GridInput(), [])
        f, self._fk = forcing if forcing is not None else (None, [])
        err = C.create_string_buffer(256)
        self.p = params
        h = L.oracle_create(C.byref(params), C.byref(g),
                            C.byref(f) if f is not None else None, err, 256)
        if not h:
            raise ValueError("oracle_create: " + err.value.decode())
        self.h = h
        d = (C.c_int * 8)()
        L.oracle_dims(h, d)
        self.im, self.imm, self.om, self.pm, self.mm, self.nm, self.lm = list(d)[:7]

    def close(self):
        if getattr(self, "h", None):
            lib().oracle_free(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def raw(self, name):
        n = C.c_long()
        isint = C.c_int()
        ptr = lib().oracle_field(self.h, name.encode(), C.byref(n), C.byref(isint))
        if not ptr:
            raise KeyError(name)
        ct = C.c_int if isint.value else C.c_double
        return np.ctypeslib.as_array(C.cast(ptr, C.POINTER(ct)), shape=(n.value,))

    SPECTRA = {"e", "ef", "ssin", "sds", "snl", "dummy", "rotl", "rotr"}
    TABLES = {"k", "kdk", "k3dk", "k4", "oneoverk4", "fkovg", "logl2overz", "cothkd",
              "invcp0", "sbf", "sdv", "sdt", "snl_arg", "bf1_renorm", "bf2_renorm",
              "dwn", "l2", "sice", "cp0", "cg0"}
    UNROLLED = {"mi", "ni", "iw", "ie", "is", "in", "iiw", "iie", "iis", "iin", "dx",
                "dy", "d", "wspd", "wdir", "uc", "vc", "rhoa", "rhow", "rhorat", "fice"}
    TWOD = {"lon", "lat", "dx_2d", "dy_2d", "curv", "d_2d", "ar_2d", "mask", "ii"}

    def get(self, name, ghost=False):
        a = self.raw(name)
        im, om, pm = self.im, self.om, self.pm
        if name in self.SPECTRA:
            a = a.reshape(im + 1, pm, om)
            return a if ghost else a[1:]
        if name in self.TABLES:
            a = a.reshape(im + 1, om)
            return a if ghost else a[1:]
        if name in self.UNROLLED:
            return a[1:]
        if name in self.TWOD:
            return a.reshape(self.nm, self.mm)
        if name in ("cth_curv", "sth_curv"):
            return a.reshape(im + 1, pm)[1:]
        if name in ("us", "vs"):
            return a.reshape(max(self.lm, 1), im)
        if name in ("f", "dom", "th", "cth", "sth", "cth2", "pl", "pr"):
            return a
        if name == "cth2pp":
            return a.reshape(pm, pm)
        return a[1:]  # per sea point (im + 1 with unused 0)

    def scalar(self, name):
        return lib().oracle_get_scalar(self.h, name.encode())

    def set_scalar(self, name, v):
        if lib().oracle_set_scalar(self.h, name.encode(), float(v)):
            raise KeyError(name)

    def __getattr__(self, name):
        if name.startswith("_") or name in ("h",):
            raise AttributeError(name)
        fn = getattr(lib(), "oracle_" + name, None)
        if fn is None:
            raise AttributeError(name)
        return lambda *a: fn(self.h, *a)

    def stokes_drift(self, init=False):
        lib().oracle_stokes_drift(self.h, int(init))

    def init_jonswap(self, js):
        lib().oracle_init_jonswap(self.h, C.byref(js))

    def forcinginput(self, forcing):
        f, self._fk2 = forcing
        lib().oracle_forcinginput(self.h, C.byref(f))

    def run_dtg(self):
        return lib().oracle_run_dtg(self.h)


def sheltering_coare35(w):
    return lib().oracle_sheltering_coare35(float(w))


def stress_tail(w, kmax):
    return lib().oracle_stress_tail(float(w), float(kmax))


def friction_velocity_skin(w, z, nu, kappa):
    return lib().oracle_friction_velocity_skin(float(w), float(z), float(nu), float(kappa))


def drag_coefficient_skin(cd_form, w, wrel, z, nu, kappa):
    return lib().oracle_drag_coefficient_skin(float(cd_form), float(w), float(wrel),
                                              float(z), float(nu), float(kappa))


def distance_haversine(lon1, lon2, lat1, lat2):
    return lib().oracle_distance_haversine(float(lon1), float(lon2), float(lat1), float(lat2))
