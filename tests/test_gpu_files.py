"""Gridded output and restart files through the C-ABI: the variables of the
reference's output_grid_nc (src/umwm_io.F90:317-854) and
restart_write/restart_read (src/umwm_restart.F90:18-145), written to the
raw container of umwm_amd/csrc/rawio.hpp and read back with
umwm_amd/rawfile.py."""
import copy

import numpy as np
import pytest

from umwm_amd import hip, rawfile
from tests.harness import make_pair, relerr, tiny_case

pytestmark = pytest.mark.gpu

# output_grid_nc's variables in the order it defines them (stokes = .true.)
OUTPUT_VARS = [
    "z", "u_stokes", "v_stokes", "d_stokes", "frequency", "theta", "lon", "lat", "seamask",
    "depth", "wspd", "wdir", "uc", "vc", "rhoa", "rhow", "fice", "psim", "momx", "momy",
    "cgmxx", "cgmxy", "cgmyy", "shelt", "epsx_atm", "epsy_atm", "epsx_ocn", "epsy_ocn",
    "taux_form", "tauy_form", "taux_form_1", "tauy_form_1", "taux_form_2", "tauy_form_2",
    "taux_form_3", "tauy_form_3", "taux_skin", "tauy_skin", "taux_diag", "tauy_diag",
    "taux_ocn", "tauy_ocn", "taux_bot", "tauy_bot", "taux_snl", "tauy_snl", "tailatmx",
    "tailatmy", "tailocnx", "tailocny", "cd", "ust", "swh", "mss", "mwp", "mwl", "mwd", "dwp",
    "dwl", "dwd", "dcp0", "dcg0", "dcp", "dcg", "physics_time_step"]


def _domain(wl, precision=8):
    d = hip.Domain(wl.params, wl.grid_input(), wl.forcing_input(0), precision=precision)
    if wl.jonswap is not None:
        d.init_jonswap(wl.jonswap)
    d.stokes_drift(True)
    return d


def test_gridded_output(tmp_path):
    wl = tiny_case("land_shelf", seed=4)
    d = _domain(wl)
    try:
        d.set_scalar("keep_sources", 1)
        d.run_dtg()
        d.stress("ocn")
        path = tmp_path / "umwmout_2012-01-01_01:00:00.bin"
        d.write_output(path)
        f = rawfile.read(path)
        assert list(f) == OUTPUT_VARS
        mm, nm, im = d.mm, d.nm, d.im
        mi, ni = d.get("mi")[:im].astype(int) - 1, d.get("ni")[:im].astype(int) - 1
        mask = d.get("mask")
        for name, src in (("swh", "ht"), ("ust", "ustar"), ("mwp", "mwp"), ("wspd", "wspd"),
                          ("taux_ocn", "taux_ocntop"), ("taux_form_2", "taux2"),
                          ("physics_time_step", "physics_time_step"), ("d_stokes", "ds")):
            v = f[name]
            assert v.dims == ["x", "y", "time"] and v.data.shape == (mm, nm, 1), name
            g = v.data[:, :, 0]
            ref = d.get(src)[:im].astype(np.float32)
            assert np.array_equal(g[mi, ni], ref), name
            assert np.all(np.isnan(g[mask.T == 0])), name
        us = d.get("us")
        u = f["u_stokes"]
        assert u.dims == ["x", "y", "z", "time"] and u.data.shape == (mm, nm, d.lm, 1)
        for l in range(d.lm):
            assert np.array_equal(u.data[mi, ni, l, 0], us[l].astype(np.float32))
        assert np.all(f["z"].data < 0)
        assert np.array_equal(f["seamask"].data[:, :, 0], mask.T)
        assert f["swh"].units == "m" and f["taux_ocn"].units == "n/m^2"
        assert np.all(f["psim"].data[mi, ni, 0] == 0)
    finally:
        d.close()


def test_restart_round_trip(tmp_path, oracle_lib):
    """A run written to a restart after one hour and read into a domain
    created with restart = 1: the spectrum and ustar come back in single
    precision (as the reference's nf90_float restart), the next substep is
    not cold, and the continued hour matches the FP64 oracle started from
    the same single-precision state."""
    wl = tiny_case("winds_currents", seed=6)
    a = _domain(wl)
    try:
        a.run_dtg()
        path = tmp_path / "umwmrst_2012-01-01_01:00:00.bin"
        a.write_restart(path)
        f = rawfile.read(path)
        assert list(f) == ["lon", "lat", "frequency", "theta", "ust", "wavenumber", "F"]
        assert f["F"].dims == ["f", "th", "x"]
        assert f["F"].data.shape == (a.om, a.pm, a.im)
        e32 = a.get("e").astype(np.float32)
        assert np.array_equal(f["F"].data, e32.reshape(a.im, a.pm, a.om).T)
        wl_r = copy.copy(wl)
        wl_r.params = copy.copy(wl.params)
        wl_r.params.restart = 1
        o, b = make_pair(wl_r, jonswap=False)
        b.stokes_drift(True)
        b.read_restart(path)
        assert np.array_equal(b.get("e"), e32.astype(np.float64))
        assert np.array_equal(b.get("ustar"), a.get("ustar").astype(np.float32).astype(np.float64))
        o.stokes_drift(True)
        o.get("e")[:] = b.get("e")
        o.get("ustar")[:] = b.get("ustar")
        sumt, dts = b.substep()
        assert dts > 0  # restart: no cold substep
        o.substep()
        assert o.scalar("dts") == pytest.approx(dts, rel=1e-12)
        o.run_dtg()
        b.run_dtg()
        for name in ("ht", "mwp", "ustar"):
            assert relerr(b.get(name), o.get(name)) <= 1e-10, name
    finally:
        a.close()
