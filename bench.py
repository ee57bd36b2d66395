"""Benchmark of the UMWM timestep hot path on one MI355X.

Headline workload: the hurricane ensemble (1500x1200 grid at 2 km, om = 49,
pm = 72, shelf from 5 m over 200 km from a seeded coastline, translating
Holland vortex delivered as hourly snapshots, refraction active), whose 16
members run in waves sized to HBM.  One step is one substep of every member
of a wave.  The chosen path is FP32 + batched + surrogate: the members of a
wave advance through the batch API, one captured hipGraph per substep, with
the learned source-term surrogate (a BF16 MFMA network trained in the same
run from the plain FP32 kernel's samples on reduced hurricane grids) where
its checks pass.  The plain path is FP64, one member at a time.  After the
first hour the chosen path's Hs and ustar are compared with the FP64 run.
The members are independent and every wave does the same work, so the
throughput of one wave is the ensemble's.

Secondary legs (each optional): the C2 global 0.5 deg workload in FP32 and
FP64, the C3 batched ensemble over its 3 hours, and the CPU baseline (the
FP64 oracle, serial and on the box's cores) on a crop of the hurricane.

Prints one JSON line.  --gpus N with N > 1 exits: the build targets one GPU.
"""
import argparse
import json
import math
import os
import re
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

from umwm_amd import abi, autotune, configs, hip  # noqa: E402

METRIC = "spectral-bin updates/s (source+advect+refract), chosen vs plain at Hs RMSE ≤2%"
HBM_PEAK_GBS = 8000.0
BF16_PEAK_TFLOPS = 2500.0

KERNELS = ("k_rates", "k_integrate", "k_advect", "k_stress")
# algorithmic HBM words per bin-update of each fused kernel (one word = the
# spectrum's precision): k_rates reads E^n; k_integrate reads E^n, writes E*;
# k_advect reads E* and E^n, writes E^{n+1}; k_stress works on O(om) per point
ALG_WORDS = {"k_rates": 1, "k_integrate": 2, "k_advect": 3, "k_stress": 0}
PROFILE = os.path.join(ROOT, "profiles", "k_advect_c5_fp32_surrogate.md")
PROFILE64 = os.path.join(ROOT, "profiles", "k_integrate_c5_fp64.md")
# tile of the fused kernels for the 49 x 72 members in FP32 (the autotuner's
# choice on a full-size member, profiles/k_advect_c5_fp32_surrogate.md)
C5_TILE32 = (36, 2, 18, 4, 2)
# and in FP64 (the plain path; the autotuner's choice on a full-size member)
C5_TILE64 = (18, 4, 9, 8, 2)


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def profile_traffic(path):
    """HBM bytes per launch of each fused kernel on this workload, measured
    with rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE (scaled by a 1 GiB copy) and
    recorded in a committed profile summary as
    `traffic_bytes_per_launch[kernel]: bytes`; {} when absent."""
    try:
        txt = open(path).read()
    except OSError:
        return {}
    return {k: float(v) for k, v in
            re.findall(r"traffic_bytes_per_launch\[(\w+)\]:\s*([0-9.eE+]+)", txt)}


def roofline(kms, bins, word, traffic=None):
    """Roofline of the dominant fused kernel (the longest of the four): its
    algorithmic bytes per launch over its mean HIP-event duration, with its
    measured traffic per launch from `traffic` (kernel -> bytes) if known."""
    kd = dict(zip(KERNELS, kms))
    dom = max(kd, key=kd.get)
    alg = ALG_WORDS[dom] * word * bins
    ach = alg / (kd[dom] * 1e-3) / 1e9
    return {"bound": "hbm", "kernel": dom, "achieved": round(ach, 1), "peak": HBM_PEAK_GBS,
            "unit": "GB/s", "frac": round(ach / HBM_PEAK_GBS, 4),
            "traffic": (traffic or {}).get(dom), "alg_bytes_per_launch": alg,
            "ms": round(kd[dom], 4)}


def make_domain(wl, precision, hour=1):
    d = hip.Domain(wl.params, wl.grid_input(), wl.forcing_input(0), precision=precision)
    d.init_jonswap(wl.jonswap)
    d.stokes_drift(True)
    if wl.snapshots is not None:
        d.forcinginput(wl.forcing_input(hour))
    return d


def kernel_ms(d, n):
    """Mean per-kernel milliseconds over n substeps (HIP events on the
    domain's stream): rates, integrate, advect, stress, surrogate."""
    d.set_timing(True)
    d.timing(reset=True)
    for _ in range(n):
        d.substep()
    ms, cnt = d.timing(reset=True)
    d.set_timing(False)
    return [m / max(c, 1) for m, c in zip(ms, cnt)]


def timed_substeps(d, steps):
    """`steps` substeps back to back (one host wait); all must be active."""
    d.synchronize()
    n0 = int(d.scalar("nsub"))
    t0 = time.perf_counter()
    _, _, n1 = d.substeps(steps)
    wall = time.perf_counter() - t0
    if n1 - n0 != steps:
        raise RuntimeError(f"only {n1 - n0} of {steps} timed substeps were inside the hour")
    return wall


def hour_fields(d):
    """The fields compared with FP64 at the end of an hour: Hs, Tm, ustar
    and the surface Stokes drift speed."""
    f = {k: d.get(k) for k in ("ht", "mwp", "ustar")}
    f["us"] = np.hypot(d.get("us")[0], d.get("vs")[0])
    return f


def field_errors(h32, ref):
    """Relative errors of hour_fields() against the FP64 ones."""
    def rmse(a, b):
        return float(np.sqrt(np.mean((a - b) ** 2)) / max(np.sqrt(np.mean(b ** 2)), 1e-30))
    e = {"hs_rmse_rel": rmse(h32["ht"], ref["ht"]),
         "hs_max_rel": float(np.max(np.abs(h32["ht"] - ref["ht"]) / np.maximum(ref["ht"], 1e-3))),
         "tm_max_rel": float(np.max(np.abs(h32["mwp"] - ref["mwp"]) / np.maximum(ref["mwp"], 1e-3))),
         "ustar_max_rel": float(np.max(np.abs(h32["ustar"] - ref["ustar"])
                                       / np.maximum(np.abs(ref["ustar"]), 1e-3))),
         "ustar_rmse_rel": rmse(h32["ustar"], ref["ustar"]),
         "stokes_surface_rmse_rel": rmse(h32["us"], ref["us"]),
         "nonfinite_points": int(np.sum(~np.isfinite(h32["ht"])))}
    return e


def finish_hour(d):
    while d.scalar("sumt") < 3600.0:
        d.substeps(8)
    d.stokes_drift(False)
    d.diag()


# ---------------------------------------------------------------------------
# hurricane: plain FP64, surrogate, chosen path
# ---------------------------------------------------------------------------
def plain_fp64(wl, steps, warmup, tile):
    t0 = time.perf_counter()
    d = make_domain(wl, 8)
    log(f"bench: c5 fp64 member created in {time.perf_counter() - t0:.1f} s, "
        f"{d.im} sea points x {d.om}x{d.pm}")
    try:
        if tile:
            d.set_tile(*tile)
        d.substep()  # cold first substep (dts = 0), not a bin update
        for _ in range(warmup):
            d.substep()
        kms = kernel_ms(d, 3)
        wall = timed_substeps(d, steps)
        bins = d.im * d.om * d.pm
        rate = steps * bins / wall
        log(f"bench: c5 fp64 {rate / 1e9:.2f} G bin-updates/s ({1e3 * wall / steps:.1f} ms/substep)")
        finish_hour(d)
        fields = hour_fields(d)
        return {"value": rate, "dtype": "f64", "ms_per_step": 1e3 * wall / steps,
                "kernel_ms": dict(zip(KERNELS, [round(x, 3) for x in kms[:4]])),
                "tile": d.tile(), "roofline": roofline(kms[:4], bins, 8, profile_traffic(PROFILE64)),
                "substeps_hour1": int(d.scalar("nsub"))}, fields
    finally:
        d.close()


def train_surrogate(members, samples, steps, hours=1):
    """The 49 x 72 network, trained on samples the plain FP32 kernel produces
    on reduced hurricane grids (100 x 80 cells of 9 km, the same generator:
    shelf, coastline, storm, winds advanced each hour): one sample per sea
    point of snapshots spread evenly over `hours` hours of `members`
    ensemble members, so that the sea states of the whole storm are seen;
    the last member's samples are held out."""
    cfg = abi.make_surrogate_config(om=49, pm=72, hidden=512, layers=3, batch=1024, chunk=65536,
                                    capacity=samples, seed=11)
    net = hip.Surrogate(cfg)
    t0 = time.perf_counter()
    ntrain = 0
    for m in range(members):
        if m == members - 1:
            ntrain = net.info()["samples"]
        wl = configs.c5(member=m, mm=100, nm=80, dx=9000.0, hours=hours)
        d = make_domain(wl, 4)
        try:
            per = max(2, samples // (members * d.im))
            span = hours * 3600.0
            net.collect(d)
            taken = 1
            for h in range(hours):
                if h:
                    d.holland_forcinginput(wl.storm(h + 1))
                    d.set_scalar("sumt", 0.0)
                while d.scalar("sumt") < 3600.0:
                    d.substep()
                    t = h * 3600.0 + d.scalar("sumt")
                    if taken < per and t >= taken * span / per:
                        net.collect(d)
                        taken += 1
        finally:
            d.close()
    n = net.info()["samples"]
    t_collect = time.perf_counter() - t0
    t0 = time.perf_counter()
    first, last = net.train(steps, lr=1e-3, ntrain=ntrain)
    last = net.train(max(steps // 3, 1), lr=2e-4, ntrain=ntrain)[1]
    t_train = time.perf_counter() - t0
    held = net.eval(ntrain, (n - ntrain) // 8 * 8)
    info = {"samples": n, "train_samples": ntrain, "steps": steps + max(steps // 3, 1),
            "loss_first": first, "loss_last": last, "loss_heldout": held,
            "collect_s": round(t_collect, 1), "train_s": round(t_train, 1),
            "network": "3531-512-512-512-3528 GELU, BF16 weights/activations, FP32 accumulate",
            "source": f"{members} hurricane members on 100x80 cells of 9 km, snapshots over "
                      f"{hours} h; the last member held out"}
    log(f"bench: surrogate {n} samples ({t_collect:.1f} s), loss {first:.4f} -> {last:.4f}, "
        f"held-out {held:.4f} ({t_train:.1f} s)")
    return net, info


def chosen_path(members, steps, warmup, net, fields64, tile, graph=True):
    """A wave of members through the batch API (graph replay) with the
    surrogate attached: hour 1 warms up (cold substep, capture) and gives
    the fields compared with FP64; hour 2 is timed, then timed again with
    the surrogate detached (FP32 batched alone)."""
    t0 = time.perf_counter()
    wls = [configs.c5(member=m, hours=2) for m in range(members)]
    doms = [make_domain(wl, 4) for wl in wls]
    log(f"bench: c5 wave of {members} fp32 members created in {time.perf_counter() - t0:.1f} s")
    b = None
    try:
        pol = abi.make_surrogate_policy(audit_stride=16)
        for d in doms:
            d.set_tile(*tile)
            if net is not None:
                net.attach(d, pol)
        b = hip.Batch(doms)
        b.set_tile(*tile)
        t0 = time.perf_counter()
        n1 = b.run_dtg(use_graph=graph)
        log(f"bench: c5 hour 1 of the wave in {time.perf_counter() - t0:.1f} s, substeps {n1}")
        h32 = hour_fields(doms[0])
        stats1 = hip.Surrogate.stats(doms[0], reset=True) if net is not None else None
        # hour 2: next snapshot, then warmup and `steps` timed substeps
        for wl, d in zip(wls, doms):
            d.forcinginput(wl.forcing_input(2))
            d.set_scalar("sumt", 0.0)
        bins = [d.im * d.om * d.pm for d in doms]
        b.substeps(warmup, use_graph=graph)
        before = [int(d.scalar("nsub")) for d in doms]
        t0 = time.perf_counter()
        after = b.substeps(steps, use_graph=graph)
        wall = time.perf_counter() - t0
        if any(a - s != steps for a, s in zip(after, before)):
            raise RuntimeError(f"timed substeps left the hour: {before} -> {after}")
        rate = steps * sum(bins) / wall
        log(f"bench: c5 chosen {rate / 1e9:.2f} G bin-updates/s ({1e3 * wall / steps:.1f} ms per "
            f"substep of {members} members)")
        stats2 = hip.Surrogate.stats(doms[0], reset=True) if net is not None else None
        # the same members without the surrogate
        for d in doms:
            hip.Surrogate.detach(d)
        b.close()
        b = hip.Batch(doms)
        b.set_tile(*tile)
        b.substeps(2, use_graph=graph)
        before = [int(d.scalar("nsub")) for d in doms]
        n_plain = min(steps, 40)
        t0 = time.perf_counter()
        after = b.substeps(n_plain, use_graph=graph)
        wall_b = time.perf_counter() - t0
        if any(a - s != n_plain for a, s in zip(after, before)):
            raise RuntimeError("FP32 batched substeps left the hour")
        rate_b = n_plain * sum(bins) / wall_b
        log(f"bench: c5 fp32 batched without surrogate {rate_b / 1e9:.2f} G bin-updates/s")
        b.close()
        b = None
        # per-kernel times of one member, surrogate attached again
        if net is not None:
            net.attach(doms[0], pol)
        kms = kernel_ms(doms[0], 3)
        out = {"members_per_wave": members, "bins_per_member_step": bins[0],
               "sea_points": doms[0].im, "ms_per_step": 1e3 * wall / steps, "wall_s": wall,
               "fp32_batched_no_surrogate": {"value": rate_b, "steps": n_plain,
                                             "ms_per_step": 1e3 * wall_b / n_plain},
               "kernel_ms": dict(zip(KERNELS + ("surrogate",), [round(x, 3) for x in kms])),
               "tile": tile, "substeps_hour1": n1}
        if net is not None:
            used = stats2["used"] / max(stats2["points"], 1)
            out["surrogate_stats"] = {"hour1": stats1, "timed": stats2, "used_fraction": used}
        # errors of the chosen path against FP64 after the first hour
        e = field_errors(h32, fields64)
        out["errors_vs_fp64_after_1h"] = e
        log("bench: c5 errors vs fp64 after 1 h " + json.dumps({k: round(v, 5) for k, v in e.items()}))
        return rate, kms, out
    finally:
        if b is not None:
            b.close()
        for d in doms:
            d.close()


# ---------------------------------------------------------------------------
# secondary legs
# ---------------------------------------------------------------------------
def c2_leg(steps, warmup):
    """C2 (720x360 global 0.5 deg, om = 37, pm = 36), FP32 and FP64 fused
    single-domain paths, tiles autotuned."""
    wl = configs.c2(seed=2, hours=1)
    out = {"workload": "c2: 720x360 global 0.5 deg lat-lon, om=37 pm=36, GRF land 30%, "
                       "log-normal depth, GRF winds 0-25 m/s"}
    for prec, tag in ((4, "fp32"), (8, "fp64")):
        d = make_domain(wl, prec)
        try:
            d.substep()
            tile, _ = autotune.autotune(d, steps=2)
            for _ in range(warmup):
                d.substep()
            kms = kernel_ms(d, 3)
            wall = timed_substeps(d, steps)
            bins = d.im * d.om * d.pm
            r = steps * bins / wall
            out[tag] = {"value": r, "ms_per_step": 1e3 * wall / steps, "tile": tile,
                        "kernel_ms": dict(zip(KERNELS, [round(x, 4) for x in kms[:4]])),
                        "roofline": roofline(kms[:4], bins, prec)}
            out["bins_per_step"] = bins
            log(f"bench: c2 {tag} {r / 1e9:.2f} G bin-updates/s")
        finally:
            d.close()
    return out


def c3_leg(n, hours=3, seed=3, scale=1.0):
    """C3: n uneven regional domains (100x100..600x400, om 25-49, pm 24-72;
    `scale` shrinks the grids) advanced together through all `hours` hours
    (one launch per kernel per pm class per substep, hipGraph replay), then
    the same members one domain at a time for the same hours."""
    wls = configs.c3(seed=seed, n=n, hours=hours, scale=scale)

    def make(wl):
        d = hip.Domain(wl.params, wl.grid_input(), wl.forcing_input(0), precision=4)
        d.init_jonswap(wl.jonswap)
        d.stokes_drift(True)
        return d

    t0 = time.perf_counter()
    doms = [make(wl) for wl in wls]
    bins = [d.im * d.om * d.pm for d in doms]
    log(f"bench: c3 {n} members, {sum(bins) / 1e9:.2f} G bins per substep set, "
        f"set-up {time.perf_counter() - t0:.1f} s")
    b = hip.Batch(doms)
    try:
        t0 = time.perf_counter()
        nsub = [b.run_dtg(use_graph=True) for _ in range(hours)]
        wall = time.perf_counter() - t0
    finally:
        b.close()
        for d in doms:
            d.close()
    # the cold substep of hour 1 updates no bin
    upd = sum(sum(ns * bb for ns, bb in zip(h, bins)) for h in nsub) - sum(bins)
    out = {"value": upd / wall, "unit": "bin-updates/s", "members": n, "hours": hours,
           "grid_scale": scale, "wall_s": wall,
           "substeps_per_hour": [int(np.mean(h)) for h in nsub]}
    log(f"bench: c3 batched {upd / wall / 1e9:.2f} G bin-updates/s over {wall:.1f} s")
    doms = [make(wl) for wl in wls]
    try:
        t0 = time.perf_counter()
        for d in doms:
            for _ in range(hours):
                d.run_dtg()
        wall_s = time.perf_counter() - t0
    finally:
        for d in doms:
            d.close()
    out["sequential"] = {"value": upd / wall_s, "wall_s": wall_s}
    out["batched_vs_sequential"] = wall_s / wall
    log(f"bench: c3 one domain at a time {upd / wall_s / 1e9:.2f} G bin-updates/s")
    return out


def crop(wl, m0, m1, n0, n1):
    """The same workload on a rectangle of its grid (forcing snapshots
    cropped alike)."""
    kw = abi.params_dict(wl.params)
    depths = kw.pop("depths")
    kw.pop("lm")
    kw.update(mm=m1 - m0, nm=n1 - n0)
    p = abi.make_params(depths=depths, **kw)
    g = {k: np.ascontiguousarray(v[m0:m1, n0:n1]) for k, v in wl.grid.items()}
    snap = wl.snapshots

    def snaps(h):
        return {k: np.ascontiguousarray(v[m0:m1, n0:n1]) for k, v in snap(h).items()}

    return configs.Workload(wl.name + "_crop", p, grid=g, forcing=snaps(0), jonswap=wl.jonswap,
                            hours=wl.hours, snapshots=snaps)


def cpu_baseline(budget_s=12.0):
    """The FP64 oracle on the box's host cores, serial and all cores, on a
    120x100 crop of the hurricane member around the storm (about 12,000
    sea points x 49 x 72 bins), timed over substeps after the cold one."""
    try:
        from oracle import oracle_py
    except Exception as e:  # pragma: no cover
        return {"value": None, "error": str(e)}
    full = configs.c5(member=0, hours=1)
    m0, n0 = 690, 900
    wl = crop(full, m0, m0 + 120, n0, n0 + 100)
    nproc = os.cpu_count() or 1
    # the box grants a one-GPU job 16 host cores, whatever the OS reports
    cores = min(nproc, 16)
    out = {}
    for tag, threads in (("serial", 1), ("all_cores", cores)):
        o = oracle_py.Oracle(wl.params, wl.grid_input(), wl.forcing_input(0), threads=threads)
        try:
            o.init_jonswap(wl.jonswap)
            o.stokes_drift(True)
            o.forcinginput(wl.forcing_input(1))
            o.substep()  # cold
            bins = o.im * o.om * o.pm
            n, t0 = 0, time.perf_counter()
            while True:
                o.substep()
                n += 1
                if time.perf_counter() - t0 > budget_s or n >= 40:
                    break
            dt = time.perf_counter() - t0
            out[tag] = {"value": n * bins / dt, "threads": threads, "substeps": n, "seconds": dt}
            log(f"bench: cpu {tag} ({threads} threads) {n * bins / dt / 1e6:.1f} M bin-updates/s")
        finally:
            o.close()
    return {"value": out["all_cores"]["value"], "unit": "bin-updates/s", "cores": cores,
            "kind": "port", "nproc": nproc, "serial": out["serial"], "all_cores": out["all_cores"],
            "sample": f"FP64 oracle (C restatement, -O2, OpenMP) on a 120x100 crop of hurricane "
                      f"member 0 around the storm at 2 km ({o.im} sea points x 49x72 bins), "
                      f"substeps after the cold one; serial and {cores} threads"}


def counters_run(steps):
    """For rocprofv3 --pmc: a briefly trained network, one FP32 hurricane
    member with the bench's tile, two plain substeps, then `steps`
    substeps with the surrogate attached, one at a time (no batch, no
    graph)."""
    net, _ = train_surrogate(2, 20000, 100, hours=1)
    d = make_domain(configs.c5(member=0, hours=1), 4)
    try:
        d.set_tile(*C5_TILE32)
        d.substep()
        d.substep()
        net.attach(d, abi.make_surrogate_policy())
        for _ in range(steps):
            d.substep()
        d.synchronize()
        log(f"bench: counters run, {steps} substeps with the surrogate")
    finally:
        d.close()
        net.close()


def ensemble_run(hours, members, sur_members, sur_samples, sur_steps, sur_hours):
    """The hurricane ensemble simulated for `hours` hours: member 0 in FP64
    (the plain path; its fields held on the host hour by hour, since it and
    a wave of FP32 members do not fit in HBM together), then a wave of
    `members` members on the chosen path (FP32, batched with graph replay,
    the surrogate attached).  Both take the storm's winds generated on the
    device each hour.  Per hour: member 0's errors against FP64, the
    surrogate's share of point-substeps, the wave's bin-update rate."""
    t_start = time.perf_counter()
    wl0 = configs.c5(member=0, hours=hours)
    ref = []
    d = make_domain(wl0, 8)
    try:
        d.set_tile(*C5_TILE64)
        for h in range(hours):
            if h:
                d.holland_forcinginput(wl0.storm(h + 1))
            n = d.run_dtg()
            ref.append(hour_fields(d))
            log(f"bench: ensemble fp64 member 0 hour {h + 1}: {n} substeps "
                f"({time.perf_counter() - t_start:.0f} s)")
    finally:
        d.close()
    net, sur_info = train_surrogate(sur_members, sur_samples, sur_steps, sur_hours)
    wls = [configs.c5(member=m, hours=hours) for m in range(members)]
    doms = [make_domain(wl, 4) for wl in wls]
    bins = [x.im * x.om * x.pm for x in doms]
    b = None
    per_hour = []
    try:
        pol = abi.make_surrogate_policy(audit_stride=16)
        for x in doms:
            x.set_tile(*C5_TILE32)
            net.attach(x, pol)
        b = hip.Batch(doms)
        b.set_tile(*C5_TILE32)
        for h in range(hours):
            if h:
                for wl, x in zip(wls, doms):
                    x.holland_forcinginput(wl.storm(h + 1))
            t0 = time.perf_counter()
            nsub = b.run_dtg(use_graph=True)
            wall = time.perf_counter() - t0
            st = hip.Surrogate.stats(doms[0], reset=True)
            upd = sum(n * bb for n, bb in zip(nsub, bins)) - (sum(bins) if h == 0 else 0)
            row = {"hour": h + 1, "substeps": [int(n) for n in nsub], "wall_s": round(wall, 2),
                   "rate": upd / wall, "surrogate_used": st["used"] / max(st["points"], 1),
                   "fallback_audit": st["fallback_audit"] / max(st["points"], 1),
                   "audit_error": st["audit_error"]}
            row.update(field_errors(hour_fields(doms[0]), ref[h]))
            per_hour.append(row)
            log(f"bench: ensemble wave hour {h + 1}: {row['rate'] / 1e9:.1f} G/s, used "
                f"{row['surrogate_used']:.3f} (audit error {row['audit_error']:.3f}, audit "
                f"fallbacks {row['fallback_audit']:.3f}), hs rmse {row['hs_rmse_rel']:.4f}, ustar rmse "
                f"{row['ustar_rmse_rel']:.4f} ({time.perf_counter() - t_start:.0f} s)")
    finally:
        if b is not None:
            b.close()
        for x in doms:
            x.close()
        net.close()
    worst = {k: max(r[k] for r in per_hour) for k in
             ("hs_rmse_rel", "ustar_rmse_rel", "stokes_surface_rmse_rel", "nonfinite_points")}
    return {"mode": "ensemble", "workload": "c5: hurricane ensemble, 1500x1200 at 2 km, om=49 "
            "pm=72, device-generated Holland winds each hour", "hours": hours,
            "members": members, "surrogate": sur_info, "per_hour": per_hour,
            "worst_hour": worst,
            "mean_rate": float(np.mean([r["rate"] for r in per_hour])),
            "wall_s": round(time.perf_counter() - t_start, 1)}


def counters_plain_run(steps):
    """For rocprofv3 --pmc: one FP64 hurricane member (the plain path) with
    the bench's tile, one cold substep, then `steps` substeps one at a time."""
    d = make_domain(configs.c5(member=0, hours=1), 8)
    try:
        d.set_tile(*C5_TILE64)
        d.substep()
        for _ in range(steps):
            d.substep()
        d.synchronize()
        log(f"bench: counters run, {steps} FP64 substeps")
    finally:
        d.close()


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100, help="timed substeps of the chosen path")
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--members", type=int, default=3, help="hurricane members per HBM wave")
    ap.add_argument("--plain-steps", type=int, default=30, help="timed FP64 substeps")
    ap.add_argument("--sur-members", type=int, default=6)
    ap.add_argument("--sur-samples", type=int, default=300_000)
    ap.add_argument("--sur-hours", type=int, default=1,
                    help="hours of the reduced-grid storms the training samples span")
    ap.add_argument("--sur-steps", type=int, default=3000)
    ap.add_argument("--no-surrogate", action="store_true")
    ap.add_argument("--no-c2", action="store_true")
    ap.add_argument("--c3-members", type=int, default=64, help="0 skips the C3 leg")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--only-chosen", action="store_true",
                    help="profiling runs: the surrogate and the chosen wave only")
    ap.add_argument("--no-graph", action="store_true",
                    help="batch without hipGraph replay (profiler runs)")
    ap.add_argument("--copy-probe", action="store_true",
                    help="only copy 1 GiB device to device three times: the known "
                         "traffic that scales the FETCH_SIZE / WRITE_SIZE counters")
    ap.add_argument("--counters", action="store_true",
                    help="counter runs: one hurricane member with the surrogate (trained "
                         "briefly) through single-domain substeps, no batch")
    ap.add_argument("--ensemble-hours", type=int, default=0,
                    help="instead of the bench line: the hurricane simulated for this many "
                         "hours (member 0 in FP64, then a wave on the chosen path), errors "
                         "per hour")
    ap.add_argument("--counters-plain", action="store_true",
                    help="counter runs: one FP64 hurricane member (the plain path) through "
                         "single-domain substeps")
    args = ap.parse_args()
    if args.ensemble_hours > 0:
        print(json.dumps(ensemble_run(args.ensemble_hours, args.members, args.sur_members,
                                      args.sur_samples, args.sur_steps, args.sur_hours)),
              flush=True)
        return
    if args.counters_plain:
        counters_plain_run(args.steps)
        return
    if args.counters:
        counters_run(args.steps)
        return
    if args.copy_probe:
        import torch
        a = torch.ones(1 << 28, device="cuda", dtype=torch.float32)
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        log("bench: copy probe, 3 copies of 1 GiB")
        return
    if args.gpus != 1:
        print("bench.py: this build runs on one GPU (how the path would shard over "
              "several is a design note only), so --gpus must be 1", file=sys.stderr)
        sys.exit(2)
    t_start = time.perf_counter()
    tile = C5_TILE32
    wl0 = configs.c5(member=0, hours=1)
    plain, f64 = (None, None)
    if not args.only_chosen:
        log("bench: c5 plain fp64 member")
        plain, f64 = plain_fp64(wl0, args.plain_steps, args.warmup, C5_TILE64)
    net, sur_info = None, None
    if not args.no_surrogate:
        log("bench: surrogate training")
        net, sur_info = train_surrogate(args.sur_members, args.sur_samples, args.sur_steps,
                                        args.sur_hours)
    if f64 is None:  # profiling: compare with the FP32 path itself
        f64 = {k: np.zeros(1) + 1 for k in ("ht", "mwp", "ustar", "us")}
    log("bench: c5 chosen path")
    rate, kms, chosen = chosen_path(args.members, args.steps, args.warmup, net, f64, tile,
                                    graph=not args.no_graph)
    if net is not None:
        net.close()
    bins = chosen["bins_per_member_step"]
    traffic = profile_traffic(PROFILE)
    out = {
        "metric": METRIC,
        "value": rate,
        "unit": "bin-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": chosen["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "c5: hurricane ensemble, 1500x1200 at 2 km, om=49 pm=72, shelf "
                               "5-3000 m over 200 km, seeded coastline, translating Holland "
                               "vortex (Rmax 30 km, Vmax 60 m/s, B 1.5, 5 m/s), 16 members "
                               "(heading +-10 deg, Vmax +-10 %) in HBM waves",
                   "members_per_wave": args.members,
                   "waves_for_16_members": math.ceil(16 / args.members),
                   "sea_points_per_member": chosen["sea_points"],
                   "bins_per_member_step": bins,
                   "chosen": "fp32 + batched (one hipGraph per substep) + BF16 surrogate",
                   "plain": "fp64 fused, one member"},
        "chosen": chosen,
        "roofline": roofline(kms[:4], bins, 4, traffic),
    }
    if len(kms) > 4 and kms[4] > 0:
        # MFMA roofline of the surrogate's network on one member per substep
        flop = 2 * (3531 * 512 + 512 * 512 * 2 + 512 * 3528) * chosen["sea_points"]
        out["roofline_surrogate"] = {"bound": "mfma", "achieved": round(flop / (kms[4] * 1e-3) / 1e12, 1),
                                     "peak": BF16_PEAK_TFLOPS, "unit": "TFLOP/s",
                                     "frac": round(flop / (kms[4] * 1e-3) / 1e12 / BF16_PEAK_TFLOPS, 4),
                                     "flop_per_substep": flop, "ms": round(kms[4], 3),
                                     "note": "network and its input kernels together"}
    if sur_info is not None:
        out["surrogate"] = sur_info
    if plain is not None:
        out["plain"] = plain
        out["chosen_vs_plain"] = rate / plain["value"]
        out["errors_vs_fp64_after_1h"] = chosen["errors_vs_fp64_after_1h"]
    if not args.only_chosen:
        if not args.no_c2:
            log("bench: c2 leg")
            out["c2"] = c2_leg(20, 3)
        if args.c3_members > 0:
            log("bench: c3 leg")
            out["batched_c3"] = c3_leg(args.c3_members)
            # the same generator at a quarter of the grid size: members too
            # small to fill the GPU alone, where batching is what pays
            log("bench: c3 leg, quarter-scale grids")
            out["batched_c3_quarter_scale"] = c3_leg(args.c3_members, scale=0.25)
        if not args.no_cpu:
            log("bench: cpu baseline")
            out["cpu_baseline"] = cpu_baseline()
    out["bench_wall_s"] = round(time.perf_counter() - t_start, 1)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
