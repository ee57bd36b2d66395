"""ctypes binding of libumwm_hip.so (include/umwm_hip.h).

Every call goes through the C-ABI; nothing here computes model physics.
Import fails loudly when the library has not been built: there is no CPU
fallback.
"""
import ctypes as C
import os

import numpy as np

from . import abi

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libumwm_hip.so")

# exported entry points declared in include/umwm_hip.h
EXPORTS = [
    "umwm_hip_error_string", "umwm_hip_last_error", "umwm_hip_device_count",
    "umwm_hip_create", "umwm_hip_destroy", "umwm_hip_dims", "umwm_hip_init_jonswap",
    "umwm_hip_stokes_drift", "umwm_hip_forcinginterpolate", "umwm_hip_forcinginput",
    "umwm_hip_holland_forcinginput",
    "umwm_hip_sin_d12", "umwm_hip_sds_d12", "umwm_hip_snl_d12", "umwm_hip_s_ice",
    "umwm_hip_source", "umwm_hip_propagation", "umwm_hip_refraction", "umwm_hip_update",
    "umwm_hip_stress", "umwm_hip_diag", "umwm_hip_end_first", "umwm_hip_substep", "umwm_hip_substeps",
    "umwm_hip_substeps_noalloc",
    "umwm_hip_run_dtg", "umwm_hip_get_field", "umwm_hip_set_field", "umwm_hip_get_scalar",
    "umwm_hip_set_scalar", "umwm_hip_field_size", "umwm_hip_set_timing",
    "umwm_hip_get_timing", "umwm_hip_synchronize", "umwm_hip_batch_create",
    "umwm_hip_batch_destroy", "umwm_hip_batch_run_dtg", "umwm_hip_batch_tile",
    "umwm_hip_batch_substeps",
    "umwm_hip_set_tile", "umwm_hip_get_tile", "umwm_hip_batch_set_tile",
    "umwm_hip_write_output", "umwm_hip_write_restart", "umwm_hip_read_restart",
    "umwm_hip_surrogate_create", "umwm_hip_surrogate_destroy", "umwm_hip_surrogate_dims",
    "umwm_hip_surrogate_get_layer", "umwm_hip_surrogate_set_layer", "umwm_hip_surrogate_forward",
    "umwm_hip_surrogate_gradients", "umwm_hip_surrogate_collect", "umwm_hip_surrogate_train",
    "umwm_hip_surrogate_eval", "umwm_hip_surrogate_get_norm", "umwm_hip_surrogate_attach",
    "umwm_hip_surrogate_stats",
    "umwm_hip_surrogate_infer", "umwm_hip_surrogate_fields",
]


class UmwmError(RuntimeError):
    pass


_lib = None


def lib():
    """Loads the library; raises if it is missing (no fallback path)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise UmwmError(f"{LIB_PATH} not found: build it with `python -m umwm_amd.build`")
    L = C.CDLL(LIB_PATH)
    P = C.POINTER
    vp = C.c_void_p
    L.umwm_hip_error_string.restype = C.c_char_p
    L.umwm_hip_error_string.argtypes = [C.c_int]
    L.umwm_hip_last_error.restype = C.c_char_p
    L.umwm_hip_last_error.argtypes = []
    L.umwm_hip_device_count.argtypes = [P(C.c_int)]
    L.umwm_hip_create.argtypes = [P(abi.Params), P(abi.GridInput), P(abi.ForcingInput), C.c_int,
                                  P(vp)]
    L.umwm_hip_destroy.argtypes = [vp]
    L.umwm_hip_dims.argtypes = [vp, P(C.c_int32)]
    L.umwm_hip_init_jonswap.argtypes = [vp, P(abi.Jonswap)]
    L.umwm_hip_stokes_drift.argtypes = [vp, C.c_int]
    L.umwm_hip_forcinginput.argtypes = [vp, P(abi.ForcingInput)]
    L.umwm_hip_holland_forcinginput.argtypes = [vp, P(abi.Holland)]
    for fn in ("forcinginterpolate", "sin_d12", "sds_d12", "snl_d12", "s_ice", "source",
               "propagation", "refraction", "update", "diag", "end_first", "synchronize"):
        getattr(L, "umwm_hip_" + fn).argtypes = [vp]
    L.umwm_hip_stress.argtypes = [vp, C.c_char_p]
    for fn in ("write_output", "write_restart", "read_restart"):
        getattr(L, "umwm_hip_" + fn).argtypes = [vp, C.c_char_p]
    L.umwm_hip_substep.argtypes = [vp, P(C.c_double), P(C.c_double)]
    L.umwm_hip_substeps.argtypes = [vp, C.c_int32, P(C.c_double), P(C.c_double), P(C.c_int32)]
    L.umwm_hip_substeps_noalloc.argtypes = [vp, C.c_int32, P(C.c_double), P(C.c_double),
                                            P(C.c_int32)]
    L.umwm_hip_run_dtg.argtypes = [vp, P(C.c_int32)]
    L.umwm_hip_get_field.argtypes = [vp, C.c_char_p, P(C.c_double), C.c_int64]
    L.umwm_hip_set_field.argtypes = [vp, C.c_char_p, P(C.c_double), C.c_int64]
    L.umwm_hip_get_scalar.argtypes = [vp, C.c_char_p, P(C.c_double)]
    L.umwm_hip_set_scalar.argtypes = [vp, C.c_char_p, C.c_double]
    L.umwm_hip_field_size.restype = C.c_int64
    L.umwm_hip_field_size.argtypes = [vp, C.c_char_p]
    L.umwm_hip_set_timing.argtypes = [vp, C.c_int]
    L.umwm_hip_get_timing.argtypes = [vp, P(C.c_double), P(C.c_int64), C.c_int]
    L.umwm_hip_batch_create.argtypes = [P(vp), C.c_int32, P(vp)]
    L.umwm_hip_batch_destroy.argtypes = [vp]
    L.umwm_hip_batch_run_dtg.argtypes = [vp, C.c_int, P(C.c_int32)]
    L.umwm_hip_batch_substeps.argtypes = [vp, C.c_int32, C.c_int, P(C.c_int32)]
    L.umwm_hip_batch_tile.argtypes = [vp, P(C.c_int32)]
    L.umwm_hip_set_tile.argtypes = [vp, P(C.c_int32)]
    L.umwm_hip_get_tile.argtypes = [vp, P(C.c_int32)]
    L.umwm_hip_batch_set_tile.argtypes = [vp, P(C.c_int32)]
    fp = P(C.c_float)
    L.umwm_hip_surrogate_create.argtypes = [P(abi.SurrogateConfig), P(vp)]
    L.umwm_hip_surrogate_destroy.argtypes = [vp]
    L.umwm_hip_surrogate_dims.argtypes = [vp, P(C.c_int32)]
    L.umwm_hip_surrogate_get_layer.argtypes = [vp, C.c_int32, fp, fp]
    L.umwm_hip_surrogate_set_layer.argtypes = [vp, C.c_int32, fp, fp]
    L.umwm_hip_surrogate_forward.argtypes = [vp, fp, C.c_int32, fp]
    L.umwm_hip_surrogate_gradients.argtypes = [vp, fp, fp, C.c_int32, P(C.c_double), P(fp),
                                               P(fp)]
    L.umwm_hip_surrogate_collect.argtypes = [vp, vp, P(C.c_int64)]
    L.umwm_hip_surrogate_train.argtypes = [vp, C.c_int32, C.c_double, C.c_int64, P(C.c_double),
                                           P(C.c_double)]
    L.umwm_hip_surrogate_eval.argtypes = [vp, C.c_int64, C.c_int64, P(C.c_double)]
    L.umwm_hip_surrogate_get_norm.argtypes = [vp, P(C.c_float), P(C.c_float), P(C.c_float),
                                              P(C.c_float)]
    L.umwm_hip_surrogate_attach.argtypes = [vp, vp, P(abi.SurrogatePolicy)]
    L.umwm_hip_surrogate_stats.argtypes = [vp, P(C.c_double), C.c_int]
    L.umwm_hip_surrogate_infer.argtypes = [vp, fp, C.c_int32, fp]
    L.umwm_hip_surrogate_fields.argtypes = [vp, P(C.c_double), P(C.c_double)]
    _lib = L
    return L


def check(rc):
    if rc != 0:
        L = lib()
        raise UmwmError(f"{L.umwm_hip_error_string(rc).decode()} ({rc}): "
                        f"{L.umwm_hip_last_error().decode()}")
    return rc


def device_count():
    n = C.c_int(0)
    lib().umwm_hip_device_count(C.byref(n))
    return n.value


SPECTRA = {"e", "ef", "es", "ssin", "sds", "snl", "dummy", "rotl", "rotr"}
TABLES = {"k", "kdk", "k3dk", "k4", "oneoverk4", "cp0", "cg0", "invcp0", "fkovg", "logl2overz",
          "cothkd", "sbf", "sdv", "sdt", "snl_arg", "bf1_renorm", "bf2_renorm", "dwn", "sice"}


class Domain:
    """One device-resident model domain (umwm_domain)."""

    def __init__(self, params, grid=None, forcing=None, precision=8):
        L = lib()
        g, self._gk = grid if grid is not None else (abi.GridInput(), [])
        f, self._fk = forcing if forcing is not None else (None, [])
        h = C.c_void_p()
        check(L.umwm_hip_create(C.byref(params), C.byref(g),
                                C.byref(f) if f is not None else None, int(precision),
                                C.byref(h)))
        self.h = h
        self.precision = precision
        d = (C.c_int32 * 8)()
        check(L.umwm_hip_dims(h, d))
        self.im, self.imm, self.om, self.pm, self.mm, self.nm, self.lm = list(d)[:7]

    def close(self):
        if getattr(self, "h", None):
            lib().umwm_hip_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- state exchange ------------------------------------------------------
    def get(self, name):
        L = lib()
        n = L.umwm_hip_field_size(self.h, name.encode())
        if n <= 0:
            raise KeyError(name)
        buf = np.empty(n, dtype=np.float64)
        check(L.umwm_hip_get_field(self.h, name.encode(),
                                   buf.ctypes.data_as(C.POINTER(C.c_double)), n))
        return self._shape(name, buf)

    def _shape(self, name, buf):
        if name in SPECTRA:
            return buf.reshape(self.im, self.pm, self.om)
        if name in TABLES:
            return buf.reshape(self.im, self.om)
        if name in ("us", "vs"):
            return buf.reshape(max(self.lm, 1), self.im)
        if name in ("mask", "ii", "dx_2d", "dy_2d", "curv", "d_2d", "lon", "lat"):
            return buf.reshape(self.nm, self.mm)
        return buf

    def set(self, name, arr):
        arr = np.ascontiguousarray(np.asarray(arr, dtype=np.float64).reshape(-1))
        check(lib().umwm_hip_set_field(self.h, name.encode(),
                                       arr.ctypes.data_as(C.POINTER(C.c_double)), arr.size))

    def scalar(self, name):
        v = C.c_double()
        check(lib().umwm_hip_get_scalar(self.h, name.encode(), C.byref(v)))
        return v.value

    def set_scalar(self, name, v):
        check(lib().umwm_hip_set_scalar(self.h, name.encode(), float(v)))

    # -- reference operators ------------------------------------------------
    def forcinginterpolate(self):
        check(lib().umwm_hip_forcinginterpolate(self.h))

    def forcinginput(self, forcing):
        f, self._fk2 = forcing
        check(lib().umwm_hip_forcinginput(self.h, C.byref(f)))

    def holland_forcinginput(self, storm):
        """The next forcing level's winds generated on the device from an
        abi.Holland storm (planar grids, winds the only forcing)."""
        check(lib().umwm_hip_holland_forcinginput(self.h, C.byref(storm)))

    def sin_d12(self):
        check(lib().umwm_hip_sin_d12(self.h))

    def sds_d12(self):
        check(lib().umwm_hip_sds_d12(self.h))

    def snl_d12(self):
        check(lib().umwm_hip_snl_d12(self.h))

    def s_ice(self):
        check(lib().umwm_hip_s_ice(self.h))

    def source(self):
        check(lib().umwm_hip_source(self.h))

    def propagation(self):
        check(lib().umwm_hip_propagation(self.h))

    def refraction(self):
        check(lib().umwm_hip_refraction(self.h))

    def update(self):
        check(lib().umwm_hip_update(self.h))

    def stress(self, option="atm"):
        check(lib().umwm_hip_stress(self.h, option.encode()))

    def diag(self):
        check(lib().umwm_hip_diag(self.h))

    def end_first(self):
        check(lib().umwm_hip_end_first(self.h))

    def stokes_drift(self, init=False):
        check(lib().umwm_hip_stokes_drift(self.h, int(init)))

    def init_jonswap(self, js):
        check(lib().umwm_hip_init_jonswap(self.h, C.byref(js)))

    def operator_substep(self):
        """One inner-loop iteration through the operator entry points, in
        the reference's call order (src/umwm_top.F90:91-141)."""
        self.forcinginterpolate()
        self.sin_d12()
        self.sds_d12()
        self.snl_d12()
        self.s_ice()
        self.source()
        self.propagation()
        self.refraction()
        self.update()
        self.stress("atm")
        if self.scalar("first"):
            self.diag()
            self.end_first()

    # -- fused path -----------------------------------------------------------
    def substep(self):
        s, d = C.c_double(), C.c_double()
        check(lib().umwm_hip_substep(self.h, C.byref(s), C.byref(d)))
        return s.value, d.value

    def write_output(self, path):
        """Gridded output (the reference's output_grid_nc variables)."""
        check(lib().umwm_hip_write_output(self.h, str(path).encode()))

    def write_restart(self, path):
        check(lib().umwm_hip_write_restart(self.h, str(path).encode()))

    def read_restart(self, path):
        check(lib().umwm_hip_read_restart(self.h, str(path).encode()))

    def substeps(self, n):
        """n substeps back to back (one host wait); returns (sumt, dts,
        substeps taken this dtg)."""
        s, d, k = C.c_double(), C.c_double(), C.c_int32()
        check(lib().umwm_hip_substeps(self.h, int(n), C.byref(s), C.byref(d), C.byref(k)))
        return s.value, d.value, k.value

    def substeps_noalloc(self, n):
        """substeps(n) with the fused kernels' scratch in the library's
        shared device array: no device allocation (see umwm_hip.h)."""
        s, d, k = C.c_double(), C.c_double(), C.c_int32()
        check(lib().umwm_hip_substeps_noalloc(self.h, int(n), C.byref(s), C.byref(d),
                                              C.byref(k)))
        return s.value, d.value, k.value

    def run_dtg(self):
        n = C.c_int32()
        check(lib().umwm_hip_run_dtg(self.h, C.byref(n)))
        return n.value

    def set_tile(self, D1, W1, D2, W2, NB=2):
        t = (C.c_int32 * 5)(int(D1), int(W1), int(D2), int(W2), int(NB))
        check(lib().umwm_hip_set_tile(self.h, t))

    def tile(self):
        t = (C.c_int32 * 5)()
        check(lib().umwm_hip_get_tile(self.h, t))
        return tuple(t)

    def set_timing(self, on=True):
        check(lib().umwm_hip_set_timing(self.h, int(on)))

    def timing(self, reset=False):
        """Mean-free accumulators: (ms, counts) of rates, integrate, advect,
        stress and the surrogate's inference."""
        ms = (C.c_double * 5)()
        cnt = (C.c_int64 * 5)()
        check(lib().umwm_hip_get_timing(self.h, ms, cnt, int(reset)))
        return list(ms), list(cnt)

    def synchronize(self):
        check(lib().umwm_hip_synchronize(self.h))


class Batch:
    """Many domains advanced together (umwm_batch)."""

    def __init__(self, domains):
        self.domains = list(domains)
        arr = (C.c_void_p * len(self.domains))(*[d.h.value for d in self.domains])
        h = C.c_void_p()
        check(lib().umwm_hip_batch_create(arr, len(self.domains), C.byref(h)))
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            lib().umwm_hip_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def run_dtg(self, use_graph=True):
        n = (C.c_int32 * len(self.domains))()
        check(lib().umwm_hip_batch_run_dtg(self.h, int(use_graph), n))
        return list(n)

    def substeps(self, n, use_graph=True):
        """n substeps of every member back to back; returns each member's
        substeps taken this dtg."""
        out = (C.c_int32 * len(self.domains))()
        check(lib().umwm_hip_batch_substeps(self.h, int(n), int(use_graph), out))
        return list(out)

    def tile(self):
        t = (C.c_int32 * 5)()
        check(lib().umwm_hip_batch_tile(self.h, t))
        return tuple(t)

    def set_tile(self, D1, W1, D2, W2, NB=2):
        t = (C.c_int32 * 5)(int(D1), int(W1), int(D2), int(W2), int(NB))
        check(lib().umwm_hip_batch_set_tile(self.h, t))


def _f32(a):
    return np.ascontiguousarray(np.asarray(a, dtype=np.float32))


def _fptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


class Surrogate:
    """The learned source-term surrogate (umwm_surrogate): a BF16 MFMA
    network trained on the device from samples of the plain source
    kernels."""

    def __init__(self, config):
        h = C.c_void_p()
        check(lib().umwm_hip_surrogate_create(C.byref(config), C.byref(h)))
        self.h = h
        d = (C.c_int32 * 6)()
        check(lib().umwm_hip_surrogate_dims(h, d))
        self.nin, self.nout, self.hidden, self.layers = list(d)[:4]
        self.widths = [self.nin] + [self.hidden] * self.layers + [self.nout]

    def close(self):
        if getattr(self, "h", None):
            lib().umwm_hip_surrogate_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def info(self):
        d = (C.c_int32 * 6)()
        check(lib().umwm_hip_surrogate_dims(self.h, d))
        return {"samples": d[4], "trained": bool(d[5])}

    def get_layer(self, l):
        K, N = self.widths[l], self.widths[l + 1]
        w = np.empty((N, K), np.float32)
        b = np.empty(N, np.float32)
        check(lib().umwm_hip_surrogate_get_layer(self.h, l, _fptr(w), _fptr(b)))
        return w, b

    def set_layer(self, l, w, b):
        w, b = _f32(w), _f32(b)
        assert w.shape == (self.widths[l + 1], self.widths[l]) and b.shape == (w.shape[0],)
        check(lib().umwm_hip_surrogate_set_layer(self.h, l, _fptr(w), _fptr(b)))

    def forward(self, x):
        x = _f32(x)
        y = np.empty((x.shape[0], self.nout), np.float32)
        check(lib().umwm_hip_surrogate_forward(self.h, _fptr(x), x.shape[0], _fptr(y)))
        return y

    def infer(self, x):
        """Outputs through the inference kernels (last layer transposed,
        bfloat16)."""
        x = _f32(x)
        y = np.empty((x.shape[0], self.nout), np.float32)
        check(lib().umwm_hip_surrogate_infer(self.h, _fptr(x), x.shape[0], _fptr(y)))
        return y

    def gradients(self, x, t):
        x, t = _f32(x), _f32(t)
        n = x.shape[0]
        gws = [np.empty((self.widths[l + 1], self.widths[l]), np.float32)
               for l in range(self.layers + 1)]
        gbs = [np.empty(self.widths[l + 1], np.float32) for l in range(self.layers + 1)]
        pw = (C.POINTER(C.c_float) * len(gws))(*[_fptr(g) for g in gws])
        pb = (C.POINTER(C.c_float) * len(gbs))(*[_fptr(g) for g in gbs])
        loss = C.c_double()
        check(lib().umwm_hip_surrogate_gradients(self.h, _fptr(x), _fptr(t), n, C.byref(loss),
                                                 pw, pb))
        return loss.value, gws, gbs

    def collect(self, dom):
        n = C.c_int64()
        check(lib().umwm_hip_surrogate_collect(self.h, dom.h, C.byref(n)))
        return n.value

    def train(self, steps, lr=1e-3, ntrain=0):
        a, b = C.c_double(), C.c_double()
        check(lib().umwm_hip_surrogate_train(self.h, int(steps), float(lr), int(ntrain),
                                             C.byref(a), C.byref(b)))
        return a.value, b.value

    def get_norm(self):
        """Input means and inverse standard deviations, output (ln dm^2)
        means and standard deviations of the last training."""
        a, b = np.empty(self.nin, np.float32), np.empty(self.nin, np.float32)
        c, d = np.empty(self.nout, np.float32), np.empty(self.nout, np.float32)
        check(lib().umwm_hip_surrogate_get_norm(self.h, _fptr(a), _fptr(b), _fptr(c), _fptr(d)))
        return a, b, c, d

    def eval(self, first, n):
        v = C.c_double()
        check(lib().umwm_hip_surrogate_eval(self.h, int(first), int(n), C.byref(v)))
        return v.value

    def attach(self, dom, policy=None):
        check(lib().umwm_hip_surrogate_attach(dom.h, self.h,
                                              C.byref(policy) if policy is not None else None))

    @staticmethod
    def detach(dom):
        check(lib().umwm_hip_surrogate_attach(dom.h, None, None))

    @staticmethod
    def fields(dom):
        """The network's spreading factor dm^2 of the last substep (im, pm,
        om), before the kernels clamp it to its bounds, and each point's
        choice (im,)."""
        n = dom.im * dom.pm * dom.om
        sds = np.empty(n)
        use = np.empty(dom.im)
        check(lib().umwm_hip_surrogate_fields(dom.h, sds.ctypes.data_as(C.POINTER(C.c_double)),
                                              use.ctypes.data_as(C.POINTER(C.c_double))))
        return sds.reshape(dom.im, dom.pm, dom.om), use

    @staticmethod
    def stats(dom, reset=False):
        out = (C.c_double * 7)()
        check(lib().umwm_hip_surrogate_stats(dom.h, out, int(reset)))
        keys = ("points", "used", "fallback_range", "fallback_residual", "fallback_audit",
                "audited", "audit_error")
        return dict(zip(keys, list(out)))

