"""GPU parity: the HIP library (operator path and fused path) against the
FP64 CPU oracle on identical seeded inputs.  All calls go through the
C-ABI of libumwm_hip.so."""
import numpy as np
import pytest

from umwm_amd import configs
from tests.harness import hs_rmse, make_pair, pointwise_relerr, relerr, tiny_case

pytestmark = pytest.mark.gpu

# FP64 tolerances: the device math library (exp, pow, cos) differs from
# glibc by a few ulp and the fused kernels reassociate sums, so a single
# operator agrees to ~1e-14 and a multi-substep run to ~1e-12.  The bars
# below are the 1e-10 the FP64 plain path must meet, with room to spare
# for single operators.
TOL_OP = 1e-11
TOL_RUN = 1e-10

MAP_FIELDS = ["mask", "ii", "mi", "ni", "iw", "ie", "is", "in", "iiw", "iie", "iis", "iin"]
TABLE_FIELDS = ["k", "kdk", "k3dk", "k4", "oneoverk4", "cp0", "cg0", "invcp0", "fkovg",
                "logl2overz", "cothkd", "sbf", "sdv", "snl_arg", "bf1_renorm", "bf2_renorm",
                "dwn"]


@pytest.mark.parametrize("case", ["const_depth", "land_shelf", "global", "winds_currents"])
def test_setup_maps_bit_exact(case, oracle_lib):
    wl = tiny_case(case)
    o, d = make_pair(wl, jonswap=False)
    assert (o.im, o.imm) == (d.im, d.imm)
    for name in MAP_FIELDS:
        a, b = d.get(name), o.get(name)
        np.testing.assert_array_equal(np.asarray(a).reshape(-1),
                                      np.asarray(b).reshape(-1)[: np.asarray(a).size],
                                      err_msg=name)
    for name in TABLE_FIELDS:
        assert relerr(d.get(name), o.get(name)) <= 1e-15, name
    for name in ("dtamin", "cfllim", "dlnf", "dth", "bf1", "bf2"):
        assert abs(d.scalar(name) - o.scalar(name)) <= 1e-15 * abs(o.scalar(name)), name


OCN_PAIRS = [("tailocnx", "tailocny"), ("taux_ocntop", "tauy_ocntop"),
             ("taux_ocnbot", "tauy_ocnbot"), ("taux_snl", "tauy_snl"), ("epsx_ocn", "epsy_ocn"),
             ("epsx_atm", "epsy_atm"), ("taux1", "tauy1"), ("taux2", "tauy2"), ("taux3", "tauy3")]


def _vector_compare(o, d, pairs, tol, tag):
    """Vector fields judged against the vector's magnitude (a component of
    a field symmetric about an axis is round-off)."""
    for x, y in pairs:
        ox, oy = o.get(x), o.get(y)
        scale = max(np.max(np.abs(ox)), np.max(np.abs(oy)))
        for name, ref in ((x, ox), (y, oy)):
            err = float(np.max(np.abs(d.get(name) - ref)))
            assert err <= tol * scale, f"{tag}: {name} abs err {err:.3e} vs scale {scale:.3e}"


def _op_compare(o, d, names, tol, tag):
    for name in names:
        err = relerr(d.get(name), o.get(name))
        assert err <= tol, f"{tag}: {name} rel err {err:.3e}"


@pytest.mark.parametrize("case", ["const_depth", "land_shelf", "global", "winds_currents"])
def test_operators_fp64(case, oracle_lib):
    """Each reference operator, called in the driver's order for two
    substeps, against the oracle's module arrays."""
    wl = tiny_case(case)
    o, d = make_pair(wl)
    for step in range(3):
        tag = f"{case} substep {step}"
        o.forcinginterpolate()
        d.forcinginterpolate()
        _op_compare(o, d, ["wspd", "wdir", "rhorat", "uc", "vc"], 1e-15, tag)
        o.sin_d12()
        d.sin_d12()
        np.testing.assert_array_equal(d.get("oc"), o.get("oc"), err_msg=tag)
        _op_compare(o, d, ["ssin", "shelt", "fcutoff"], TOL_OP, tag)
        o.sds_d12()
        d.sds_d12()
        _op_compare(o, d, ["dummy", "sds"], TOL_OP, tag)
        o.snl_d12()
        d.snl_d12()
        _op_compare(o, d, ["snl", "sds", "sdt"], TOL_OP, tag)
        o.s_ice()
        d.s_ice()
        _op_compare(o, d, ["sice"], TOL_OP, tag)
        o.source()
        d.source()
        for s in ("dts", "sumt"):
            assert abs(d.scalar(s) - o.scalar(s)) <= TOL_OP * max(abs(o.scalar(s)), 1.0), (tag, s)
        _op_compare(o, d, ["ef", "e"], TOL_OP, tag)
        o.propagation()
        d.propagation()
        _op_compare(o, d, ["ef"], TOL_OP, tag)
        o.refraction()
        d.refraction()
        _op_compare(o, d, ["rotl", "rotr", "ef"], TOL_OP, tag)
        assert abs(d.scalar("dtr") - o.scalar("dtr")) <= TOL_OP * max(abs(o.scalar("dtr")), 1.0)
        o.update()
        d.update()
        _op_compare(o, d, ["e"], TOL_OP, tag)
        o.stress_atm()
        d.stress("atm")
        _op_compare(o, d, ["ustar", "cd", "taux", "tauy", "taux_form", "tauy_form",
                           "taux_skin", "tauy_skin", "taux_diag", "tauy_diag"], TOL_OP, tag)
        o.stress_ocn()
        d.stress("ocn")
        _vector_compare(o, d, OCN_PAIRS, TOL_OP, tag)
        if step == 0:
            o.diag()
            d.diag()
            _op_compare(o, d, ["ht", "mwp", "mss", "mwl", "dwl", "dwp"], TOL_OP, tag)
            # vector components are judged against the vector's magnitude:
            # a component of a spectrum symmetric about an axis is ~1e-15
            for x, y in (("momx", "momy"), ("cgmxx", "cgmxy"), ("cgmyy", "cgmxy")):
                ox, oy = o.get(x), o.get(y)
                scale = max(np.max(np.abs(ox)), np.max(np.abs(oy)))
                for name, ref in ((x, ox), (y, oy)):
                    err = np.max(np.abs(d.get(name) - ref)) / scale
                    assert err <= TOL_OP, f"{tag}: {name} rel err {err:.3e}"
            o.set_scalar("first", 0)
            d.end_first()


@pytest.mark.parametrize("case", ["const_depth", "land_shelf", "global", "winds_currents"])
def test_fused_fp64_substeps(case, oracle_lib):
    """The fused three-kernel substep against the oracle's operator
    sequence, substep by substep."""
    wl = tiny_case(case)
    o, d = make_pair(wl)
    for step in range(6):
        o.substep()
        s_d, dts_d = d.substep()
        assert abs(dts_d - o.scalar("dts")) <= TOL_RUN * max(o.scalar("dts"), 1.0), step
        assert abs(s_d - o.scalar("sumt")) <= TOL_RUN * max(o.scalar("sumt"), 1.0), step
        err = relerr(d.get("e"), o.get("e"))
        assert err <= TOL_RUN, f"{case} step {step}: e rel err {err:.3e}"
        for name in ("ustar", "cd", "taux", "tauy"):
            err = pointwise_relerr(d.get(name), o.get(name), floor=1e-12)
            assert err <= TOL_RUN, f"{case} step {step}: {name} rel err {err:.3e}"


def test_fused_matches_operators_fp64(oracle_lib):
    """Operator path and fused path of the library agree with each other."""
    wl = tiny_case("land_shelf", seed=3)
    _, a = make_pair(wl)
    _, b = make_pair(wl)
    for _ in range(4):
        a.operator_substep()
        b.substep()
    assert relerr(b.get("e"), a.get("e")) <= TOL_RUN
    assert relerr(b.get("ustar"), a.get("ustar")) <= TOL_RUN


def test_run_dtg_fp64_tiny(oracle_lib):
    """One dtg of the inner loop, Stokes drift and diag."""
    wl = tiny_case("land_shelf", seed=5)
    o, d = make_pair(wl)
    n_o = o.run_dtg()
    n_d = d.run_dtg()
    assert n_o == n_d
    for name in ("ht", "mwp", "ustar"):
        err = pointwise_relerr(d.get(name), o.get(name), floor=1e-12)
        assert err <= TOL_RUN, f"{name} {err:.3e}"
    assert relerr(d.get("e"), o.get("e")) <= TOL_RUN
    for name in ("us", "vs"):
        assert relerr(d.get(name), o.get(name)) <= TOL_RUN, name
        # per level, pointwise above 1e-6 of the level's largest drift
        for lev, (x, x_ref) in enumerate(zip(d.get(name), o.get(name))):
            err = pointwise_relerr(x, x_ref, floor=1e-6 * np.max(np.abs(x_ref)))
            assert err <= 1e-9, f"{name} level {lev}: {err:.3e}"
    assert relerr(d.get("ds"), o.get("ds")) <= TOL_RUN


def test_c1_fp64_one_hour(oracle_lib):
    """Config 1 (64x48, 10 km, 4000 m, 15 m/s, JONSWAP), 1 h of dynamic
    substeps: Hs, Tm, ustar and the Stokes drift to relative 1e-10."""
    wl = configs.c1()
    o, d = make_pair(wl, threads=8)
    n_o = o.run_dtg()
    n_d = d.run_dtg()
    assert n_o == n_d
    for name in ("ht", "mwp", "ustar"):
        err = pointwise_relerr(d.get(name), o.get(name), floor=1e-12)
        assert err <= TOL_RUN, f"{name} {err:.3e}"
    for name in ("us", "vs"):
        assert relerr(d.get(name), o.get(name)) <= TOL_RUN, name


def test_c1_fp32_vs_fp64(oracle_lib):
    """FP32 path on config 1: Hs within 0.5 %, ustar within 3 %."""
    wl = configs.c1()
    o, d = make_pair(wl, precision=4, threads=8)
    o.run_dtg()
    d.run_dtg()
    hs, hs_ref = d.get("ht"), o.get("ht")
    assert hs_rmse(hs, hs_ref) <= 0.005
    assert pointwise_relerr(hs, hs_ref) <= 0.005
    assert pointwise_relerr(d.get("ustar"), o.get("ustar")) <= 0.03


@pytest.mark.parametrize("case", ["land_shelf", "winds_currents"])
def test_fp32_tiny_tracks_fp64(case, oracle_lib):
    wl = tiny_case(case, seed=7)
    o, d = make_pair(wl, precision=4)
    o.run_dtg()
    d.run_dtg()
    assert hs_rmse(d.get("ht"), o.get("ht")) <= 0.005


def test_fp32_tracks_fp64_below_one_metre(oracle_lib):
    """A shelf whose shallow edge lies between 0.3 and 0.8 m with dmin =
    0.2 m (the reference requires only dmin > 0, src/umwm_init.F90:147):
    the FP32 sweeps derive their tables from the true depth there, as the
    FP64 records and the oracle do."""
    wl = configs.tiny(seed=7, winds=True, dmin=0.2, dshelf=0.6)
    z = np.asarray(wl.grid["z"])
    assert (-z[z < 0]).min() < 1.0  # sea points shallower than 1 m exist
    o, d = make_pair(wl, precision=4)
    o.run_dtg()
    d.run_dtg()
    assert hs_rmse(d.get("ht"), o.get("ht")) <= 0.005


def test_stress_ocn_needs_operator_sources():
    """After a fused substep the source arrays of stress('ocn') were never
    materialised: without keep_sources the call reports a state error
    instead of reading stale arrays."""
    from umwm_amd import hip
    wl = tiny_case("land_shelf")
    d = hip.Domain(wl.params, wl.grid_input(), wl.forcing_input(0), precision=8)
    d.init_jonswap(wl.jonswap)
    d.substep()
    with pytest.raises(hip.UmwmError, match="stress\\('ocn'\\)"):
        d.stress("ocn")
    # keep_sources set mid-dtg: nothing is saved until the dtg ends
    d.set_scalar("keep_sources", 1)
    with pytest.raises(hip.UmwmError, match="keep_sources"):
        d.stress("ocn")
    d.close()


@pytest.mark.parametrize("case", ["land_shelf", "winds_currents"])
def test_stress_ocn_after_fused_hour(case, oracle_lib):
    """With keep_sources, the fused run_dtg saves E^n and ustar of the
    substep that reaches dtg; stress('ocn') at the hour end then equals the
    oracle's output-time stress('ocn') after the same hour (FP64)."""
    wl = tiny_case(case, seed=9)
    o, d = make_pair(wl)
    d.set_scalar("keep_sources", 1)
    for hour in range(2):
        if hour and wl.snapshots is not None:  # the next level (make_pair loaded hour 1)
            o.forcinginput(wl.forcing_input(hour + 1))
            d.forcinginput(wl.forcing_input(hour + 1))
        assert o.run_dtg() == d.run_dtg()
        # forcing as the last active substep interpolated it (substeps run
        # ahead past the dtg must not re-interpolate it at sumt = dtg)
        _op_compare(o, d, ["wspd", "wdir", "uc", "vc"], TOL_RUN, f"{case} hour {hour}")
        o.stress_ocn()
        d.stress("ocn")
        _vector_compare(o, d, OCN_PAIRS, TOL_RUN, f"{case} hour {hour}")
        _op_compare(o, d, ["ssin", "sds", "snl"], TOL_RUN, f"{case} hour {hour}")


def _fresh(wl, precision=8):
    from umwm_amd import hip
    d = hip.Domain(wl.params, wl.grid_input(), wl.forcing_input(0), precision=precision)
    if wl.jonswap is not None:
        d.init_jonswap(wl.jonswap)
    d.stokes_drift(True)
    return d


@pytest.mark.parametrize("precision", [8, 4])
def test_substeps_back_to_back_equal_single_substeps(precision):
    """umwm_hip_substeps(n) launches n substeps with one host wait: the
    spectrum, sumt and substep count equal n single substeps bit for bit;
    running past the end of the dtg does nothing more."""
    wl = tiny_case("winds_currents", seed=3)
    a, b = _fresh(wl, precision), _fresh(wl, precision)
    try:
        sa = [a.substep() for _ in range(6)]
        sumt, dts, n = b.substeps(6)
        assert n == 6 and sumt == sa[-1][0] and dts == sa[-1][1]
        assert np.array_equal(a.get("e"), b.get("e"))
        assert np.array_equal(a.get("ustar"), b.get("ustar"))
        # to the end of the hour and beyond: extra substeps are no-ops
        while a.scalar("sumt") < wl.params.dtg:
            a.substep()
        n_hour = int(a.scalar("nsub"))
        sumt, _, n = b.substeps(n_hour - 6 + 5)
        assert n == n_hour and sumt == wl.params.dtg
        assert np.array_equal(a.get("e"), b.get("e"))
    finally:
        a.close()
        b.close()


def test_run_dtg_runs_ahead_like_single_substeps():
    """run_dtg launches runs of substeps between host waits; its hours match
    hours of single substeps bit for bit, substep counts included."""
    wl = tiny_case("land_shelf", seed=11)
    a, b = _fresh(wl), _fresh(wl)
    try:
        for hour in range(2):
            f = wl.forcing_input(hour) if hour else None
            if f is not None:
                a.forcinginput(f)
                b.forcinginput(f)
            n_b = b.run_dtg()
            a.set_scalar("sumt", 0.0)
            n_a = 0
            while a.scalar("sumt") < wl.params.dtg:
                a.substep()
                n_a += 1
            a.stokes_drift(False)
            a.diag()
            assert n_a == n_b
            assert np.array_equal(a.get("e"), b.get("e"))
            assert np.array_equal(a.get("ht"), b.get("ht"))
    finally:
        a.close()
        b.close()


# Tile shapes of the two sweep kernels at pm = 36 and pm = 24: each wave
# holding exactly DMAX directions (with directions paired with their
# opposite when DMAX is even, unpaired when it is odd), waves holding fewer
# directions than DMAX, and a wave holding none.
TILES = {
    36: [(36, 1, 18, 2, 4), (18, 2, 12, 3, 2), (6, 6, 6, 6, 4), (9, 4, 9, 4, 2),
         (36, 2, 18, 3, 2), (9, 5, 6, 7, 2)],
    24: [(6, 4, 6, 4, 2), (18, 2, 12, 2, 4), (36, 1, 3, 8, 2)],
}


@pytest.mark.parametrize("pm,tile", [(pm, t) for pm, ts in TILES.items() for t in ts])
def test_fused_tiles_fp64(pm, tile, oracle_lib):
    """The fused substep under each tile shape against the oracle."""
    wl = tiny_case("winds_currents", seed=13, pm=pm)
    o, d = make_pair(wl)
    d.set_tile(*tile)
    for step in range(4):
        o.substep()
        d.substep()
        err = relerr(d.get("e"), o.get("e"))
        assert err <= TOL_RUN, f"tile {tile} step {step}: e rel err {err:.3e}"
        err = pointwise_relerr(d.get("ustar"), o.get("ustar"), floor=1e-12)
        assert err <= TOL_RUN, f"tile {tile} step {step}: ustar rel err {err:.3e}"
    d.close()


@pytest.mark.parametrize("tile", TILES[36][:3])
def test_fused_tiles_fp32(tile):
    """FP32 under other tiles against FP32 under the default tile: only the
    order of the directional sums differs."""
    wl = tiny_case("land_shelf", seed=17, pm=36)
    a, b = _fresh(wl, 4), _fresh(wl, 4)
    try:
        b.set_tile(*tile)
        for _ in range(4):
            a.substep()
            b.substep()
        assert relerr(b.get("e"), a.get("e")) <= 1e-5
        assert pointwise_relerr(b.get("ustar"), a.get("ustar"), floor=1e-12) <= 1e-5
    finally:
        a.close()
        b.close()


@pytest.mark.parametrize("precision", [4, 8])
def test_dts_is_the_exact_grid_min(precision):
    """The rates kernel's grid-wide reduction over about 700 workgroups
    spread over every XCD: dts = min(min_i pts(i), dtamin, dtg - sumt)
    exactly, with pts the per-point physics_time_step the same substep
    stored (src/umwm_physics.F90:41-58)."""
    wl = configs.c2(seed=4, mm=360, nm=180, hours=1)
    d = _fresh(wl, precision)
    try:
        assert d.im > 40000
        d.substep()  # cold first substep: dts = 0
        assert d.scalar("dts") == 0.0
        for _ in range(3):
            sumt0 = d.scalar("sumt")
            d.substep()
            pts = np.asarray(d.get("physics_time_step"))[: d.im]
            want = min(float(pts.min()), d.scalar("dtamin"), wl.params.dtg - sumt0)
            assert d.scalar("dts") == want
            assert d.scalar("sumt") == sumt0 + want
    finally:
        d.close()
