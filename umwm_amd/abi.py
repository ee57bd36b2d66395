"""ctypes mirrors of the C-ABI structs declared in include/umwm_hip.h.

The field order follows the C declarations exactly; both the library and the
oracle read these structs.  Namelist defaults are the values of the
reference's namelists/main.nml (read by src/umwm_init.F90:67-92).
"""
import ctypes as C

import numpy as np

I32 = C.c_int32
F64 = C.c_double


class Params(C.Structure):
    """umwm_params: namelist groups DOMAIN, PHYSICS, GRID, FORCING,
    FORCING_CONSTANT, OUTPUT, STOKES."""
    _fields_ = [
        ("mm", I32), ("nm", I32), ("om", I32), ("pm", I32),
        ("isglobal", I32), ("restart", I32),
        ("gridfromfile", I32), ("topofromfile", I32),
        ("fillestuaries", I32), ("filllakes", I32),
        ("winds", I32), ("currents", I32), ("air_density", I32),
        ("water_density", I32), ("seaice", I32),
        ("stokes", I32), ("lm", I32), ("reserved0", I32),
        ("fmin", F64), ("fmax", F64), ("fprog", F64), ("dtg", F64),
        ("g", F64), ("nu_air", F64), ("nu_water", F64), ("sfct", F64),
        ("kappa", F64), ("z", F64), ("gustiness", F64), ("dmin", F64),
        ("explim", F64),
        ("sin_fac", F64), ("sin_diss1", F64), ("sin_diss2", F64),
        ("sds_fac", F64), ("sds_power", F64), ("mss_fac", F64),
        ("snl_fac", F64), ("sdt_fac", F64), ("sbf_fac", F64), ("sbp_fac", F64),
        ("delx", F64), ("dely", F64), ("dpt", F64),
        ("wspd0", F64), ("wdir0", F64), ("uc0", F64), ("vc0", F64),
        ("rhoa0", F64), ("rhow0", F64), ("fice0", F64), ("fice_lth", F64),
        ("fice_uth", F64),
        ("depths", F64 * 100),
    ]


class GridInput(C.Structure):
    _fields_ = [
        ("lon", C.POINTER(F64)), ("lat", C.POINTER(F64)), ("z", C.POINTER(F64)),
        ("exclude", C.POINTER(I32)), ("nexclude", I32), ("reserved0", I32),
    ]


class ForcingInput(C.Structure):
    _fields_ = [
        ("uw", C.POINTER(F64)), ("vw", C.POINTER(F64)),
        ("uc", C.POINTER(F64)), ("vc", C.POINTER(F64)),
        ("rhoa", C.POINTER(F64)), ("rhow", C.POINTER(F64)),
        ("fice", C.POINTER(F64)),
    ]


class Jonswap(C.Structure):
    _fields_ = [
        ("alpha", F64), ("gamma", F64), ("sigma_a", F64), ("sigma_b", F64),
        ("fp", F64), ("fp_from_wind", I32), ("dir_from_wind", I32),
        ("thw", F64),
    ]


class SurrogateConfig(C.Structure):
    _fields_ = [
        ("om", I32), ("pm", I32), ("hidden", I32), ("layers", I32), ("batch", I32),
        ("chunk", I32), ("capacity", C.c_int64), ("seed", C.c_uint64),
        ("fmin", F64), ("fmax", F64),
    ]


class Holland(C.Structure):
    _fields_ = [(n, F64) for n in ("xc", "yc", "vmax", "rmax", "b", "rho", "dp", "trans_u",
                                   "trans_v", "lat_deg", "inflow_deg")]


def make_holland(xc, yc, vmax, rmax, b=1.5, rho=1.15, dp=0.0, trans=(0.0, 0.0), lat_deg=20.0,
                 inflow_deg=20.0):
    return Holland(xc=xc, yc=yc, vmax=vmax, rmax=rmax, b=b, rho=rho, dp=dp, trans_u=trans[0],
                   trans_v=trans[1], lat_deg=lat_deg, inflow_deg=inflow_deg)


class SurrogatePolicy(C.Structure):
    _fields_ = [
        ("audit_stride", I32), ("reserved0", I32), ("resid_tol", F64), ("audit_tol", F64),
        ("range_margin", F64), ("zmax", F64),
    ]


def make_surrogate_config(om=37, pm=36, hidden=512, layers=3, batch=4096, chunk=65536,
                          capacity=1 << 20, seed=1, fmin=0.0313, fmax=2.0):
    return SurrogateConfig(om=om, pm=pm, hidden=hidden, layers=layers, batch=batch, chunk=chunk,
                           capacity=capacity, seed=seed, fmin=fmin, fmax=fmax)


def make_surrogate_policy(audit_stride=16, resid_tol=0.05, audit_tol=0.25, range_margin=0.1,
                          zmax=8.0):
    return SurrogatePolicy(audit_stride=audit_stride, resid_tol=resid_tol, audit_tol=audit_tol,
                           range_margin=range_margin, zmax=zmax)


# Stokes depth levels of the reference namelist (namelists/main.nml:72).
DEFAULT_DEPTHS = [0.1, 0.2, 0.3, 0.4, 0.5, 0.6, 0.7, 0.8, 0.9, 1.0, 1.2, 1.4,
                  1.6, 1.8, 2.0, 2.4, 2.8, 3.2, 3.6, 4.0, 4.5, 5.0, 6, 7, 8, 9,
                  10, 12, 14, 16, 18, 20, 24, 28, 32, 40, 50, 60, 70, 80, 90,
                  100]

DEFAULTS = dict(
    # &DOMAIN
    isglobal=0, mm=21, nm=11, om=37, pm=36, fmin=0.0313, fmax=2.0, fprog=2.0,
    dtg=3600.0, restart=0,
    # &PHYSICS
    g=9.80665, nu_air=1.56e-5, nu_water=0.90e-6, sfct=0.07, kappa=0.4,
    z=10.0, gustiness=0.0, dmin=10.0, explim=0.9, sin_fac=0.11,
    sin_diss1=0.10, sin_diss2=0.001, sds_fac=42.0, sds_power=2.4,
    mss_fac=360.0, snl_fac=5.0, sdt_fac=0.002, sbf_fac=0.003, sbp_fac=0.003,
    # &GRID
    gridfromfile=0, delx=10000.0, dely=10000.0, topofromfile=0, dpt=1000.0,
    fillestuaries=0, filllakes=0,
    # &FORCING
    winds=0, currents=0, air_density=0, water_density=0, seaice=0,
    # &FORCING_CONSTANT
    wspd0=10.0, wdir0=0.0, uc0=0.0, vc0=0.0, rhoa0=1.2, rhow0=1030.0,
    fice0=0.0, fice_lth=0.30, fice_uth=0.75,
    # &OUTPUT
    stokes=1,
)


def make_params(**kw):
    """Builds a Params struct from the reference defaults plus overrides.
    `depths` may be a sequence of Stokes levels (count(depths >= 0) used)."""
    vals = dict(DEFAULTS)
    depths = kw.pop("depths", DEFAULT_DEPTHS)
    unknown = set(kw) - set(vals)
    if unknown:
        raise KeyError(f"unknown namelist keys: {sorted(unknown)}")
    vals.update(kw)
    p = Params()
    for k, v in vals.items():
        ft = dict(Params._fields_)[k]
        setattr(p, k, int(v) if ft is I32 else float(v))
    arr = [-1.0] * 100
    for q, d in enumerate(depths):
        arr[q] = float(d)
    p.depths[:] = arr
    p.lm = sum(1 for d in arr if d >= 0)
    return p


def params_dict(p):
    out = {}
    for name, _ in Params._fields_:
        if name in ("depths", "reserved0"):
            continue
        out[name] = getattr(p, name)
    out["depths"] = [d for d in p.depths if d >= 0]
    return out


def dptr(a):
    """Pointer to a contiguous float64 numpy array (or NULL for None)."""
    if a is None:
        return C.POINTER(F64)()
    assert a.dtype == np.float64 and (a.flags["C_CONTIGUOUS"] or a.flags["F_CONTIGUOUS"])
    return a.ctypes.data_as(C.POINTER(F64))


def make_grid_input(lon=None, lat=None, z=None, exclude=None):
    """Returns (GridInput, keepalive).  2-D arrays are given as numpy
    (mm, nm) arrays and passed in Fortran order (m fastest)."""
    keep = []

    def f2(a):
        if a is None:
            return None
        b = np.asfortranarray(np.asarray(a, dtype=np.float64))
        keep.append(b)
        return b

    g = GridInput()
    g.lon = dptr(f2(lon))
    g.lat = dptr(f2(lat))
    g.z = dptr(f2(z))
    if exclude is not None and len(exclude):
        ex = np.ascontiguousarray(np.asarray(exclude, dtype=np.int32).reshape(-1))
        keep.append(ex)
        g.exclude = ex.ctypes.data_as(C.POINTER(I32))
        g.nexclude = len(ex) // 2
    else:
        g.exclude = C.POINTER(I32)()
        g.nexclude = 0
    return g, keep


def make_forcing_input(**fields):
    """ForcingInput from (mm, nm) arrays keyed uw vw uc vc rhoa rhow fice."""
    keep = []
    fi = ForcingInput()
    for name, _ in ForcingInput._fields_:
        a = fields.get(name)
        if a is None:
            setattr(fi, name, C.POINTER(F64)())
        else:
            b = np.asfortranarray(np.asarray(a, dtype=np.float64))
            keep.append(b)
            setattr(fi, name, b.ctypes.data_as(C.POINTER(F64)))
    return fi, keep


def make_jonswap(alpha=0.01, gamma=3.3, sigma_a=0.07, sigma_b=0.09, fp=0.3,
                 fp_from_wind=False, dir_from_wind=True, thw=0.0):
    js = Jonswap()
    js.alpha, js.gamma, js.sigma_a, js.sigma_b = alpha, gamma, sigma_a, sigma_b
    js.fp = fp
    js.fp_from_wind = int(fp_from_wind)
    js.dir_from_wind = int(dir_from_wind)
    js.thw = thw
    return js
