"""Batched ensembles: members of uneven grid and spectral sizes advanced
together (one launch per kernel per substep, replayed as a hipGraph) give
exactly the results of advancing each member alone with the same tile, and
each member keeps its own dts and substep count."""
import numpy as np
import pytest

from umwm_amd import configs, hip

MEMBERS = [  # (seed, mm, nm, om, pm, land fraction, shelf, wind)
    (11, 24, 18, 25, 24, 0.2, True, 9.0),
    (12, 30, 14, 31, 28, 0.0, False, 17.0),
    (13, 18, 26, 37, 36, 0.35, True, 24.0),
    (14, 36, 20, 29, 32, 0.1, False, 6.0),
]


def make(prec):
    out = []
    for seed, mm, nm, om, pm, lf, shelf, w in MEMBERS:
        wl = configs.regional_random(seed, mm, nm, om, pm, dx=1e4, land_frac=lf, shelf=shelf,
                                     wspd=w, hours=1)
        d = hip.Domain(wl.params, wl.grid_input(), wl.forcing_input(0), precision=prec)
        if wl.jonswap is not None:
            d.init_jonswap(wl.jonswap)
        d.stokes_drift(True)
        out.append(d)
    return out


@pytest.mark.gpu
@pytest.mark.parametrize("prec", [4, 8])
@pytest.mark.parametrize("use_graph", [True, False])
@pytest.mark.parametrize("D", [9, 6])  # 6: pm = 24 and 36 take the mask-free, paired sweeps
def test_batch_equals_members_alone(prec, use_graph, D):
    batch_doms = make(prec)
    alone = make(prec)
    b = hip.Batch(batch_doms)
    try:
        # members are grouped by pm; each group runs the batch's D with
        # W = ceil(pm / D) waves, so a member alone gets that same shape
        pmax = max(d.pm for d in batch_doms)
        b.set_tile(D, -(-pmax // D), D, -(-pmax // D), 2)
        for d in alone:
            w = -(-d.pm // D)
            d.set_tile(D, w, D, w, 2)
        for hour in range(2):
            nsub = b.run_dtg(use_graph=use_graph)
            n_alone = [d.run_dtg() for d in alone]
            assert nsub == n_alone, (hour, nsub, n_alone)
            for k, (db, da) in enumerate(zip(batch_doms, alone)):
                for name in ("e", "ustar", "ht"):
                    a, c = db.get(name), da.get(name)
                    assert np.array_equal(a, c), f"member {k} hour {hour} {name}"
                assert db.scalar("sumt") == da.scalar("sumt")
        # members differ: the batch really held uneven, independent domains
        assert len({d.pm for d in batch_doms}) == 4
        assert len(set(nsub)) > 1
    finally:
        b.close()
        for d in batch_doms + alone:
            d.close()


# Members with hourly forcing snapshots (winds; winds and currents) next to
# a member with constant forcing: (tiny() arguments, keeps its sources)
MIXED = [
    (dict(seed=21, mm=14, nm=11, om=25, pm=24, winds=True), True),
    (dict(seed=22, mm=16, nm=12, om=31, pm=28, winds=True, currents=True), False),
    (dict(seed=23, mm=12, nm=13, om=25, pm=24, land=True, shelf=True), False),
    (dict(seed=24, mm=13, nm=10, om=29, pm=28, winds=True, land=False, shelf=True), False),
]


def make_mixed(prec):
    wls, doms = [], []
    for kw, keep in MIXED:
        wl = configs.tiny(**kw)
        d = hip.Domain(wl.params, wl.grid_input(), wl.forcing_input(0), precision=prec)
        d.init_jonswap(wl.jonswap)
        d.stokes_drift(True)
        if wl.snapshots is not None:
            d.forcinginput(wl.forcing_input(1))
        if keep:
            d.set_scalar("keep_sources", 1)
        wls.append(wl)
        doms.append(d)
    return wls, doms


@pytest.mark.gpu
@pytest.mark.parametrize("prec", [4, 8])
@pytest.mark.parametrize("use_graph", [True, False])
def test_batch_forcing_currents_keep_equal_alone(prec, use_graph):
    """Three hours with a forcing advance before each later hour: the
    replayed graph must follow the swapped forcing levels; a member with
    currents refreshes its rotation every substep while the others keep the
    first substep's; a member keeping its sources gives the stress('ocn') of
    the member run alone.  Every member equals itself run alone, bit for
    bit."""
    wls, batch_doms = make_mixed(prec)
    _, alone = make_mixed(prec)
    b = hip.Batch(batch_doms)
    try:
        for hour in range(3):
            if hour:
                for wl, db, da in zip(wls, batch_doms, alone):
                    if wl.snapshots is not None:
                        db.forcinginput(wl.forcing_input(hour + 1))
                        da.forcinginput(wl.forcing_input(hour + 1))
            nsub = b.run_dtg(use_graph=use_graph)
            n_alone = [d.run_dtg() for d in alone]
            assert nsub == n_alone, (hour, nsub, n_alone)
            for k, (db, da) in enumerate(zip(batch_doms, alone)):
                for name in ("e", "ustar", "ht", "wspd", "uc", "mwd", "dwd"):
                    assert np.array_equal(db.get(name), da.get(name)), \
                        f"member {k} hour {hour} {name}"
                for name in ("dtr_temp", "maxrot", "sumt"):
                    assert db.scalar(name) == da.scalar(name), f"member {k} hour {hour} {name}"
            db, da = batch_doms[0], alone[0]
            db.stress("ocn")
            da.stress("ocn")
            for name in ("taux_ocntop", "tauy_ocntop", "epsx_ocn", "taux_snl"):
                assert np.array_equal(db.get(name), da.get(name)), f"hour {hour} {name}"
        # the member with currents refreshed its rotation limit after the
        # first substep, the others did not: dtr_temp differs in kind
        assert batch_doms[1].scalar("zerocurrents") == 0
    finally:
        b.close()
        for d in batch_doms + alone:
            d.close()
