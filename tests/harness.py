"""Shared helpers of the parity tests: build the oracle and the HIP domain
from the same workload and compare fields."""
import numpy as np

from oracle import oracle_py
from umwm_amd import configs, hip


def make_pair(wl, precision=8, threads=4, jonswap=True):
    """Oracle and HIP domain from one workload, both started from the same
    JONSWAP spectrum with stokes_drift('init') recomputed afterwards."""
    o = oracle_py.Oracle(wl.params, wl.grid_input(), wl.forcing_input(0), threads=threads)
    d = hip.Domain(wl.params, wl.grid_input(), wl.forcing_input(0), precision=precision)
    if jonswap and wl.jonswap is not None:
        o.init_jonswap(wl.jonswap)
        d.init_jonswap(wl.jonswap)
        o.stokes_drift(True)
        d.stokes_drift(True)
    if wl.snapshots is not None:
        # umwm_run's first outer iteration: forcinginput(start + dtg)
        o.forcinginput(wl.forcing_input(1))
        d.forcinginput(wl.forcing_input(1))
    return o, d


def relerr(a, b):
    """Normwise relative error max|a-b| / max|b| (0 when both vanish)."""
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    den = np.max(np.abs(b)) if b.size else 0.0
    num = np.max(np.abs(a - b)) if b.size else 0.0
    if den == 0:
        return float(num)
    return float(num / den)


def pointwise_relerr(a, b, floor=0.0):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return float(np.max(np.abs(a - b) / np.maximum(np.abs(b), floor + 1e-300)))


def hs_rmse(hs, hs_ref):
    """RMS of the Hs difference relative to the RMS of the reference Hs."""
    hs = np.asarray(hs)
    hs_ref = np.asarray(hs_ref)
    return float(np.sqrt(np.mean((hs - hs_ref) ** 2)) / np.sqrt(np.mean(hs_ref ** 2)))


TINY_CASES = {
    "const_depth": dict(land=False, shelf=False),
    "land_shelf": dict(land=True, shelf=True),
    "global": dict(land=True, shelf=True, isglobal=True, mm=16, nm=10),
    "winds_currents": dict(land=True, shelf=True, winds=True, currents=True),
}


def tiny_case(name, seed=0, **kw):
    args = dict(TINY_CASES[name])
    args.update(kw)
    return configs.tiny(seed=seed, **args)


def spectrum_pointwise(a, b, floor=1e-6):
    """Largest relative error over the bins that hold at least `floor` of
    their point's largest bin: unlike the normwise relerr it also judges
    the spectral tail and the swell far below each point's peak."""
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    peak = np.max(np.abs(b), axis=tuple(range(1, b.ndim)), keepdims=True)
    m = np.abs(b) >= floor * peak
    m &= peak > 0
    if not m.any():
        return 0.0
    return float(np.max(np.abs(a[m] - b[m]) / np.abs(b[m])))


def angle_err(a, b):
    """Largest difference of two angle fields, wrapped to [0, pi]."""
    d = np.asarray(a, dtype=np.float64) - np.asarray(b, dtype=np.float64)
    return float(np.max(np.abs(np.angle(np.exp(1j * d))))) if d.size else 0.0
