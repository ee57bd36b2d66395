"""GPU: the learned source-term surrogate.

  * the BF16 MFMA kernels (forward, loss, backward with split-K weight
    gradients) against the FP64 numpy restatement oracle/mlp_ref.py, which
    finite differences pin (tests/test_surrogate_ref.py), on a small network
    and on the full 1335 -> 512 -> 512 -> 512 -> 1332 one;
  * training on the device from samples of the plain FP32 source kernels on
    seeded config-1 grids, judged on samples of a held-out seed;
  * the surrogate path of the fused substep on a fresh config-1 seed against
    the FP64 oracle (Hs within 2 %, ustar within 3 %), its fallback rate, and
    what it costs against the plain FP32 path;
  * what the learned part gains over a per-bin mean spreading factor (the
    network with its last layer zeroed), both against the FP64 oracle;
  * the fallbacks: an untrained network under a full audit falls back
    everywhere (and the run equals the plain FP32 one), and inputs outside
    the training range fall back on the range check;
  * the domain-wide audit gate: closed by an audit above the tolerance, it
    sends every point to the plain rates in the next substeps.
"""
import os
import time

import numpy as np
import pytest

from oracle import mlp_ref
from tests.harness import hs_rmse, make_pair, pointwise_relerr
from umwm_amd import abi, configs, hip

pytestmark = pytest.mark.gpu

THREADS = min(os.cpu_count() or 1, 16)


def make_net(om, pm, hidden, layers, batch, chunk=None, capacity=0, seed=1):
    cfg = abi.make_surrogate_config(om=om, pm=pm, hidden=hidden, layers=layers, batch=batch,
                                    chunk=chunk or max(batch, 64), capacity=capacity, seed=seed)
    return hip.Surrogate(cfg)


def ref_arrays(net):
    """The network's weights as the device multiplies them (bf16)."""
    ws, bs = [], []
    for l in range(net.layers + 1):
        w, b = net.get_layer(l)
        ws.append(mlp_ref.bf16_round(w))
        bs.append(b.astype(np.float64))
    return ws, bs


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a, np.float64) - b) / np.linalg.norm(b))


def cosine(a, b):
    a, b = np.asarray(a, np.float64).ravel(), np.asarray(b, np.float64).ravel()
    return float(a @ b / (np.linalg.norm(a) * np.linalg.norm(b)))


@pytest.mark.parametrize("shape", ["small", "full"])
def test_kernels_match_fp64_restatement(shape):
    """Forward outputs and every layer's gradients.  The device rounds
    inputs, weights and activations to bfloat16 (8 significant bits) and
    accumulates in FP32; the reference gets the same bf16 inputs and
    weights and computes in FP64, so the bound is a few bf16 roundings
    carried through the layers: 2 % normwise on outputs and gradients,
    cosine above 0.999."""
    if shape == "small":
        om, pm, hidden, layers, n = 4, 4, 64, 2, 64
    else:
        om, pm, hidden, layers, n = 37, 36, 512, 3, 256
    net = make_net(om, pm, hidden, layers, batch=n)
    try:
        rng = np.random.default_rng(7)
        if shape == "small":  # weights of our own, set through the ABI
            for l in range(layers + 1):
                K, N = net.widths[l], net.widths[l + 1]
                net.set_layer(l, rng.normal(0, 1 / np.sqrt(K), (N, K)), rng.normal(0, 0.1, N))
        x = rng.normal(size=(n, net.nin)).astype(np.float32)
        t = rng.normal(size=(n, net.nout)).astype(np.float32)
        ws, bs = ref_arrays(net)
        xb = mlp_ref.bf16_round(x)
        _, _, y_ref = mlp_ref.forward(ws, bs, xb)
        y = net.forward(x)
        assert rel(y, y_ref) <= 0.02, rel(y, y_ref)
        # the inference kernels of a substep: last layer transposed, output
        # rounded to bfloat16
        yi = net.infer(x)
        assert rel(yi, y_ref) <= 0.025, rel(yi, y_ref)
        L, gws, gbs = net.gradients(x, t)
        L_ref, gw_ref, gb_ref = mlp_ref.gradients(ws, bs, xb, t)
        assert abs(L - L_ref) <= 0.02 * L_ref, (L, L_ref)
        for l in range(layers + 1):
            assert rel(gws[l], gw_ref[l]) <= 0.02, (l, rel(gws[l], gw_ref[l]))
            assert cosine(gws[l], gw_ref[l]) >= 0.999, l
            assert rel(gbs[l], gb_ref[l]) <= 0.02, (l, rel(gbs[l], gb_ref[l]))
    finally:
        net.close()


# ---------------------------------------------------------------------------
# training on the plain kernel's samples and the surrogate path
# ---------------------------------------------------------------------------
# 117 seeds x 6 snapshots x 2852 sea points = 2.0 M training samples
TRAIN_SEEDS = list(range(1, 118))
HELDOUT_SEED = 118
EVAL_SEED = 1000


def fp32_domain(wl):
    d = hip.Domain(wl.params, wl.grid_input(), wl.forcing_input(0), precision=4)
    d.init_jonswap(wl.jonswap)
    d.stokes_drift(True)
    if wl.snapshots is not None:
        d.forcinginput(wl.forcing_input(1))
    return d


def harvest(net, seeds, per_seed=6):
    """`per_seed` snapshots of each seed's first hour, at the cold start and
    then as the model time passes each further 1/per_seed of the hour."""
    for s in seeds:
        d = fp32_domain(configs.c1_seeded(s))
        try:
            net.collect(d)
            taken = 1
            while d.scalar("sumt") < 3600.0 or taken == 1:
                d.substep()
                if taken < per_seed and d.scalar("sumt") >= taken * 3600.0 / per_seed:
                    net.collect(d)
                    taken += 1
        finally:
            d.close()


@pytest.fixture(scope="module")
def trained():
    net = make_net(37, 36, 512, 3, batch=2048, chunk=4096, capacity=2_020_000, seed=3)
    harvest(net, TRAIN_SEEDS)
    ntrain = net.info()["samples"]
    harvest(net, [HELDOUT_SEED])
    total = net.info()["samples"]
    assert total - ntrain >= 2852  # the held-out seed found room
    t0 = time.perf_counter()
    first, last = net.train(3000, lr=1e-3, ntrain=ntrain)
    last2 = net.train(1000, lr=2e-4, ntrain=ntrain)[1]
    t_train = time.perf_counter() - t0
    held = net.eval(ntrain, (total - ntrain) // 8 * 8)
    print(f"surrogate: {ntrain} training samples, {total - ntrain} held out; loss "
          f"{first:.4f} -> {last:.4f} -> {last2:.4f}, held-out {held:.4f}, {t_train:.1f} s")
    yield net, dict(ntrain=ntrain, first=first, last=last2, held=held)
    net.close()


def test_training_reduces_the_loss(trained):
    net, info = trained
    assert info["ntrain"] == len(TRAIN_SEEDS) * 6 * 2852 >= 2_000_000
    assert info["last"] <= 0.1 * info["first"]
    # the held-out seed is fitted about as well as the training seeds
    assert info["held"] <= 3.0 * info["last"] + 0.01


def test_network_dissipation_at_a_fresh_start(trained):
    """The dissipation the network gives a fresh seed's initial spectrum
    against the plain sds_d12 evaluation of the same state (energy-weighted
    relative error), through the substep's inference path: the network's
    spreading factor dm^2 times the closed form sds / dm^2 of the bin."""
    net, _ = trained
    wl = configs.c1_seeded(EVAL_SEED)
    a, b = fp32_domain(wl), fp32_domain(wl)
    try:
        net.attach(a, abi.make_surrogate_policy(audit_stride=1))
        a.substep()  # cold substep: inference on the initial spectrum
        dm_net, use = hip.Surrogate.fields(a)
        for op in ("forcinginterpolate", "sin_d12", "sds_d12", "snl_d12"):
            getattr(b, op)()
        sds, dm = b.get("sds"), b.get("dummy")
        assert np.all(np.isfinite(dm_net))
        sds_net = sds / dm * dm_net
        w = b.get("e") * b.get("kdk")[:, None, :]
        err = float(np.sum(np.abs(sds_net - sds) * w) / np.sum(sds * w))
        st = hip.Surrogate.stats(a)
        print(f"surrogate at a fresh start: energy-weighted sds error {err:.4f}, "
              f"points using it {np.mean(use):.3f}, stats {st}")
        assert err <= 0.1
    finally:
        a.close()
        b.close()


def test_surrogate_path_on_a_fresh_seed(trained):
    """One hour on a config-1 seed the network never saw: the FP32 path with
    the surrogate against the FP64 oracle (Hs 2 %, ustar 3 %), beside the
    plain FP32 path; the fallback rate and the time of both paths."""
    net, _ = trained
    wl = configs.c1_seeded(EVAL_SEED)
    o, plain = make_pair(wl, precision=4, threads=THREADS)
    o2, sur = make_pair(wl, precision=4, threads=1)
    o2.close()
    try:
        net.attach(sur, abi.make_surrogate_policy())
        o.run_dtg()
        t0 = time.perf_counter()
        plain.run_dtg()
        t_plain = time.perf_counter() - t0
        t0 = time.perf_counter()
        sur.run_dtg()
        t_sur = time.perf_counter() - t0
        st = hip.Surrogate.stats(sur)
        hs_ref = o.get("ht")
        e_sur = hs_rmse(sur.get("ht"), hs_ref)
        e_plain = hs_rmse(plain.get("ht"), hs_ref)
        u_sur = pointwise_relerr(sur.get("ustar"), o.get("ustar"))
        used = st["used"] / max(st["points"], 1)
        print(f"surrogate path: Hs RMSE {e_sur:.4f} (plain FP32 {e_plain:.2e}), ustar "
              f"{u_sur:.4f}, surrogate used on {used:.3f} of point-substeps, audit error "
              f"{st['audit_error']:.4f}, time {t_sur:.3f} s vs plain {t_plain:.3f} s; {st}")
        assert st["points"] > 0 and used > 0.5
        assert e_sur <= 0.02
        assert u_sur <= 0.03
    finally:
        for m in (o, plain, sur):
            m.close()


def test_spectrum_inference_matches_fp64_restatement(trained):
    """The substep's inference path, whose first layer forms its inputs from
    the spectrum itself (ln e, normalised, bf16) and whose last layer writes
    dm^2 - 1 = exp(sd y + mu) - 1 into the [bin][point] array, against the
    FP64 numpy forward on the same state: inputs built here from e, wind and
    depth, rounded to bf16, the network's bf16 weights.  The bound is that
    of the kernel test plus the bf16 rounding of dm^2 - 1 (0.4 %): 2 %
    normwise on ln dm^2 and on dm^2."""
    net, _ = trained
    wl = configs.c1_seeded(EVAL_SEED)
    a, b = fp32_domain(wl), fp32_domain(wl)
    try:
        net.attach(a, abi.make_surrogate_policy())
        e_a0 = a.get("e")
        a.substep()  # cold substep: inference on the initial spectrum
        dm_net, _ = hip.Surrogate.fields(a)
        b.forcinginterpolate()
        # the reference sees the state the substep's inference saw
        assert np.array_equal(e_a0, b.get("e"))
        im = b.im
        e = b.get("e")
        feats = [np.log(np.maximum(e.transpose(0, 2, 1).reshape(im, -1), 1e-30))]
        w = np.maximum(b.get("wspd")[:im], 1e-2)
        wd = b.get("wdir")[:im]
        feats.append(np.stack([w * np.cos(wd), w * np.sin(wd), np.log(b.get("d")[:im])], 1))
        x = np.concatenate(feats, 1)
        fmu, fis, tmu, tsd = (v.astype(np.float64) for v in net.get_norm())
        xb = mlp_ref.bf16_round((x - fmu) * fis)
        ws, bs = ref_arrays(net)
        y = mlp_ref.forward(ws, bs, xb)[2][:, :net.nout]
        ln_ref = np.minimum(tsd * y + tmu, 30.0)
        # fields are (im, pm, om); outputs are bins o pm + p
        ln_net = np.log(dm_net.transpose(0, 2, 1).reshape(im, -1))
        # the same rows through the bf16-input inference kernels
        ln_inf = np.minimum(tsd * net.infer(((x - fmu) * fis).astype(np.float32)) + tmu, 30.0)
        for f in ("wspd", "wdir"):
            assert np.array_equal(a.get(f)[:im], b.get(f)[:im])
        print(f"spectrum inference: rel(fields, numpy) {rel(ln_net, ln_ref):.4f}, "
              f"rel(infer, numpy) {rel(ln_inf, ln_ref):.4f}")
        assert rel(ln_net, ln_ref) <= 0.02, rel(ln_net, ln_ref)
        assert rel(np.exp(ln_net), np.exp(ln_ref)) <= 0.02
    finally:
        a.close()
        b.close()


def test_network_gains_over_the_mean_spreading_factor(trained):
    """What the learned part gains: one hour on the fresh config-1 seed with
    every point taking the network's spreading factor (no residual or audit
    fallback), against the same run with the network's last layer zeroed,
    which gives each bin the training set's mean ln dm^2 (a per-bin
    climatological factor, the method without the learned dependence on the
    spectrum and forcing).  Both against the FP64 oracle: the network's Hs
    error is under half of the mean factor's, and under 2 %."""
    net, _ = trained
    wl = configs.c1_seeded(EVAL_SEED)
    pol = abi.make_surrogate_policy(audit_stride=0, resid_tol=1e30)
    o, a = make_pair(wl, precision=4, threads=THREADS)
    o2, b = make_pair(wl, precision=4, threads=1)
    o2.close()
    L = net.layers
    w, bias = net.get_layer(L)
    try:
        o.run_dtg()
        net.attach(a, pol)
        a.run_dtg()
        used_a = hip.Surrogate.stats(a)
        net.set_layer(L, np.zeros_like(w), np.zeros_like(bias))
        net.attach(b, pol)
        b.run_dtg()
        used_b = hip.Surrogate.stats(b)
        hs_ref = o.get("ht")
        e_net, e_mean = hs_rmse(a.get("ht"), hs_ref), hs_rmse(b.get("ht"), hs_ref)
        u_net = pointwise_relerr(a.get("ustar"), o.get("ustar"))
        u_mean = pointwise_relerr(b.get("ustar"), o.get("ustar"))
        print(f"surrogate gain: Hs RMSE {e_net:.4f} with the network, {e_mean:.4f} with the "
              f"mean factor; ustar {u_net:.4f} vs {u_mean:.4f}; used {used_a['used']:.0f} / "
              f"{used_b['used']:.0f} of {used_a['points']:.0f}")
        assert used_a["used"] >= 0.9 * used_a["points"]
        assert e_net <= 0.5 * e_mean
        assert e_net <= 0.02
    finally:
        net.set_layer(L, w, bias)
        for m in (o, a, b):
            m.close()


def test_untrained_network_falls_back_under_full_audit():
    """Random weights (statistics from a few samples): with every workgroup
    audited each substep, the audit rejects the network at every point and
    the run is the plain FP32 run."""
    net = make_net(37, 36, 64, 1, batch=64, chunk=4096, capacity=20_000, seed=5)
    wl = configs.c1_seeded(2)
    a, b = fp32_domain(wl), fp32_domain(wl)
    try:
        net.collect(a)
        net.train(1, lr=1e-6)
        net.attach(b, abi.make_surrogate_policy(audit_stride=1, audit_tol=0.02, resid_tol=1e30))
        for _ in range(4):
            a.substep()
            b.substep()
        st = hip.Surrogate.stats(b)
        assert st["audited"] == st["points"] > 0
        assert st["used"] == 0
        assert st["fallback_audit"] + st["fallback_range"] == st["points"]
        assert pointwise_relerr(b.get("ht"), a.get("ht")) <= 1e-5
        assert np.max(np.abs(b.get("e") - a.get("e"))) <= 1e-5 * np.max(np.abs(a.get("e")))
    finally:
        for x in (a, b):
            x.close()
        net.close()


def test_out_of_range_inputs_fall_back(trained):
    """Winds of 40 m/s lie outside the training range (10-20 m/s): every
    point falls back on the range check."""
    net, _ = trained
    wl = configs.c1_seeded(EVAL_SEED, vary=False)
    wl.params.wspd0 = 40.0
    d = fp32_domain(wl)
    try:
        net.attach(d, abi.make_surrogate_policy())
        for _ in range(3):
            d.substep()
        st = hip.Surrogate.stats(d)
        assert st["points"] > 0 and st["fallback_range"] == st["points"]
    finally:
        d.close()


def test_audit_gate_closes_and_reopens(trained):
    """The domain-wide gate: a quarter of the workgroups audited each
    substep.  With a zero audit tolerance the first substep's audit closes
    the gate and from the second substep on no point takes the network,
    audited or not; with the tolerance above every error it never closes,
    and the network serves as without the gate."""
    net, _ = trained
    wl = configs.c1_seeded(HELDOUT_SEED)
    a, b = fp32_domain(wl), fp32_domain(wl)
    try:
        a.substep()
        b.substep()
        net.attach(a, abi.make_surrogate_policy(audit_stride=4, audit_tol=0.0, resid_tol=1e30))
        net.attach(b, abi.make_surrogate_policy(audit_stride=4, audit_tol=1e30, resid_tol=1e30))
        a.substep()
        b.substep()
        s_a = hip.Surrogate.stats(a, reset=True)
        s_b = hip.Surrogate.stats(b, reset=True)
        # first substep with the surrogate: the gate is open in both, the
        # zero tolerance rejects only the audited points of `a`
        assert s_a["audited"] > 0 and s_a["used"] > 0
        assert s_a["used"] + s_a["fallback_audit"] == s_b["used"]
        for _ in range(3):
            a.substep()
            b.substep()
        s_a = hip.Surrogate.stats(a)
        s_b = hip.Surrogate.stats(b)
        assert s_a["used"] == 0
        assert s_a["fallback_audit"] + s_a["fallback_range"] == s_a["points"]
        assert s_b["used"] >= 0.9 * (s_b["points"] - s_b["fallback_range"]) > 0
        # reattaching resets the gate
        net.attach(a, abi.make_surrogate_policy(audit_stride=4, audit_tol=1e30, resid_tol=1e30))
        hip.Surrogate.stats(a, reset=True)
        a.substep()
        assert hip.Surrogate.stats(a)["used"] > 0
    finally:
        a.close()
        b.close()
