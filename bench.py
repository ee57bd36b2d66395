"""Benchmark of the UMWM timestep hot path on one MI355X.

Workload: the 720x360 global 0.5 deg lat-lon grid (GRF land 30 %, log-normal
depth, GRF winds 0-25 m/s, JONSWAP start from the local wind), om = 37,
pm = 36.  A "step" is one substep (source + advect + refract + stress) over
the whole grid.  The chosen path (FP32) and the plain path (FP64) run side
by side on the same inputs; after the timed steps both finish the 6 h run
and the FP32 Hs / Tm / ustar fields are compared with the FP64 ones.

Prints one JSON line.  --gpus N with N > 1 exits: the build targets one GPU.
"""
import argparse
import json
import os
import re
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

from umwm_amd import autotune, configs, hip  # noqa: E402

METRIC = "spectral-bin updates/s (source+advect+refract), chosen vs plain at Hs RMSE ≤2%"
HBM_PEAK_GBS = 8000.0


KERNELS = ("k_rates", "k_integrate", "k_advect", "k_stress")
# algorithmic HBM words per bin-update of each fused kernel (one word = the
# spectrum's precision): k_rates reads E^n; k_integrate reads E^n, writes E*;
# k_advect reads E* and E^n, writes E^{n+1}; k_stress works on O(om) per point
ALG_WORDS = {"k_rates": 1, "k_integrate": 2, "k_advect": 3, "k_stress": 0}
PROFILE = os.path.join(ROOT, "profiles", "k_advect_c2_fp32.md")


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def profile_traffic():
    """HBM bytes per k_advect launch on this workload, measured with
    rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE (scaled by a 1 GiB copy) and
    recorded in the committed profile summary; None when absent."""
    try:
        txt = open(PROFILE).read()
    except OSError:
        return None
    m = re.search(r"traffic_bytes_per_launch:\s*([0-9.eE+]+)", txt)
    return float(m.group(1)) if m else None


def roofline(r, word, traffic):
    """Roofline of the dominant kernel (the longest of the four): its
    algorithmic bytes per launch over its mean HIP-event duration."""
    kms = dict(zip(KERNELS, r["kms"]))
    dom = max(kms, key=kms.get)
    alg = ALG_WORDS[dom] * word * r["bins"]
    ach = alg / (kms[dom] * 1e-3) / 1e9
    return {"bound": "hbm", "kernel": dom, "achieved": round(ach, 1), "peak": HBM_PEAK_GBS,
            "unit": "GB/s", "frac": round(ach / HBM_PEAK_GBS, 4),
            "traffic": traffic if dom == "k_advect" and word == 4 else None,
            "alg_bytes_per_launch": alg, "ms": round(kms[dom], 4)}


def roofline_substep(r, word):
    """The whole substep against B_alg = 3 words per bin-update, the lower
    bound for any implementation of the path (one read of E^n for the rates
    and the global dts, one read to integrate and advect, one write)."""
    t = sum(r["kms"]) * 1e-3
    ach = 3 * word * r["bins"] / t / 1e9
    return {"bound": "hbm", "achieved": round(ach, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(ach / HBM_PEAK_GBS, 4), "b_alg_bytes_per_bin": 3 * word,
            "ms": round(t * 1e3, 4)}


def make_domain(wl, precision):
    d = hip.Domain(wl.params, wl.grid_input(), wl.forcing_input(0), precision=precision)
    d.init_jonswap(wl.jonswap)
    d.stokes_drift(True)
    d.forcinginput(wl.forcing_input(1))
    return d


def timed_steps(d, warmup, steps, tune=True, kernel_steps=5):
    """Cold substep, autotuning and warmup, then `kernel_steps` substeps
    with per-kernel HIP-event timing, then `steps` substeps launched back
    to back (one host wait, as run_dtg runs them) for the wall-clock rate.
    Returns wall seconds, per-kernel mean ms, and the tile chosen."""
    d.substep()  # cold first substep (dts = 0), not counted
    tile = None
    if tune:
        tile, _ = autotune.autotune(d, steps=2, verbose=log)
    for _ in range(warmup):
        d.substep()
    d.set_timing(True)
    d.timing(reset=True)
    for _ in range(kernel_steps):
        d.substep()
    ms, cnt = d.timing(reset=True)
    d.set_timing(False)
    d.synchronize()
    n0 = int(d.scalar("nsub"))
    t0 = time.perf_counter()
    _, _, n1 = d.substeps(steps)
    wall = time.perf_counter() - t0
    # every timed substep must have advanced the model: one that starts at
    # sumt >= dtg does nothing and would inflate the rate
    if n1 - n0 != steps:
        raise RuntimeError(f"only {n1 - n0} of {steps} timed substeps were inside the hour; "
                           "lower --steps/--warmup")
    log(f"bench: {n1} substeps taken so far this hour")
    return wall, [m / max(c, 1) for m, c in zip(ms, cnt)], tile


def finish_run(d, wl, hours):
    """Completes the current hour and the remaining hours of the run."""
    done = 0
    while d.scalar("sumt") < wl.params.dtg:
        d.substep()
    d.stokes_drift(False)
    d.diag()
    done = 1
    for h in range(done, hours):
        d.forcinginput(wl.forcing_input(h + 1))
        d.run_dtg()


def cpu_baseline(wl_full, seconds_budget=20.0):
    """The FP64 oracle on a longitude slab of the same grid, all host cores
    (capped at 16), timed over substeps after the cold one."""
    try:
        from oracle import oracle_py
    except Exception as e:  # pragma: no cover
        return {"value": None, "error": str(e)}
    mm_s = 120
    g = dict(wl_full.grid)
    for key in ("lon", "lat", "z"):
        g[key] = np.ascontiguousarray(g[key][:mm_s])
    f = {k: np.ascontiguousarray(v[:mm_s]) for k, v in wl_full.forcing.items()}
    from umwm_amd import abi
    kw = abi.params_dict(wl_full.params)
    depths = kw.pop("depths")
    for k in ("lm",):
        kw.pop(k)
    kw.update(mm=mm_s, isglobal=0)
    p = abi.make_params(depths=depths, **kw)
    threads = min(os.cpu_count() or 1, 16)
    o = oracle_py.Oracle(p, abi.make_grid_input(**g), abi.make_forcing_input(**f), threads=threads)
    o.init_jonswap(wl_full.jonswap)
    o.stokes_drift(True)
    fi = abi.make_forcing_input(**f)
    o.forcinginput(fi)
    o.substep()  # cold
    bins = o.im * o.om * o.pm
    n, t0 = 0, time.perf_counter()
    while True:
        o.substep()
        n += 1
        if time.perf_counter() - t0 > seconds_budget or n >= 20:
            break
    dt = time.perf_counter() - t0
    return {"value": n * bins / dt, "unit": "bin-updates/s", "cores": threads, "kind": "port",
            "sample": f"FP64 oracle, {n} substeps of a {mm_s}x{wl_full.params.nm} longitude "
                      f"slab of the same grid ({o.im} sea points x {o.om}x{o.pm} bins), "
                      f"{threads} OpenMP threads"}


def make_member(wl, precision):
    d = hip.Domain(wl.params, wl.grid_input(), wl.forcing_input(0), precision=precision)
    d.init_jonswap(wl.jonswap)
    d.stokes_drift(True)
    return d


def batched_leg(n, seed=3, precision=4, sequential=True, scale=1.0):
    """C3: n uneven regional domains (100x100..600x400, om 25-49, pm 24-72)
    advanced together, one launch per kernel per substep replayed as a
    hipGraph; the first hour (cold substep, graph capture) is untimed and the
    second is timed.  With `sequential`, the same members are then advanced
    one domain at a time for the same hour (each with the batch's tile) to
    show what batching gains."""
    t0 = time.perf_counter()
    wls = configs.c3(seed=seed, n=n, hours=2, scale=scale)
    doms = [make_member(wl, precision) for wl in wls]
    setup = time.perf_counter() - t0
    bins = [d.im * d.om * d.pm for d in doms]
    log(f"bench: c3 {n} members, {sum(bins) / 1e9:.2f} G bins per substep set, set-up {setup:.1f} s")
    # autotune on a probe built like the largest member
    big = int(np.argmax(bins))
    probe = make_member(wls[big], precision)
    probe.substep()
    tile, _ = autotune.autotune(probe, steps=2, verbose=log)
    probe.close()
    b = hip.Batch(doms)
    b.set_tile(*tile)
    b.run_dtg(use_graph=True)
    t0 = time.perf_counter()
    nsub = b.run_dtg(use_graph=True)
    wall = time.perf_counter() - t0
    upd = sum(n_ * b_ for n_, b_ in zip(nsub, bins))
    out = {"value": upd / wall, "unit": "bin-updates/s", "members": n, "grid_scale": scale,
           "bins_per_member_substep": [int(x) for x in bins], "substeps": nsub,
           "wall_s": wall, "tile": tile, "graph": True, "dtype": "f32" if precision == 4 else "f64"}
    log(f"bench: c3 batched {upd / wall / 1e9:.2f} G bin-updates/s over {wall:.2f} s")
    b.close()
    for d in doms:
        d.close()
    if sequential:
        doms = [make_member(wl, precision) for wl in wls]
        for d in doms:
            d.set_tile(*tile)
            d.run_dtg()
        for d in doms:
            d.synchronize()
        t0 = time.perf_counter()
        nseq = [d.run_dtg() for d in doms]
        wall_s = time.perf_counter() - t0
        upd_s = sum(n_ * b_ for n_, b_ in zip(nseq, bins))
        out["sequential"] = {"value": upd_s / wall_s, "wall_s": wall_s, "substeps": nseq}
        out["batched_vs_sequential"] = (upd / wall) / (upd_s / wall_s)
        log(f"bench: c3 one domain at a time {upd_s / wall_s / 1e9:.2f} G bin-updates/s")
        for d in doms:
            d.close()
    return out


def hurricane_leg(steps, warmup, member=0):
    """C5 member: 1500x1200 at 2 km, om = 49, pm = 72, shelf and coastline,
    translating Holland vortex.  FP32 and FP64 each take a cold substep, the
    autotuner, `warmup` and `steps` timed substeps, then finish the first hour
    so that their Hs and ustar can be compared at the same time."""
    wl = configs.c5(member=member, hours=1)
    out, fields = {"member": member, "note": wl.note}, {}
    for prec, tag in ((4, "fp32"), (8, "fp64")):
        t0 = time.perf_counter()
        d = make_domain(wl, prec)
        log(f"bench: c5 {tag} created in {time.perf_counter() - t0:.1f} s, "
            f"{d.im} sea points x {d.om}x{d.pm}")
        bins = d.im * d.om * d.pm
        wall, kms, tile = timed_steps(d, warmup, steps)
        rate = steps * bins / wall
        log(f"bench: c5 {tag} {rate / 1e9:.2f} G bin-updates/s")
        while d.scalar("sumt") < wl.params.dtg:
            d.substep()
        d.diag()
        fields[tag] = (d.get("ht"), d.get("ustar"))
        out[tag] = {"value": rate, "ms_per_step": 1e3 * wall / steps,
                    "kernel_ms": dict(zip(KERNELS, [round(x, 4) for x in kms])), "tile": tile}
        out["sea_points"], out["bins_per_step"] = d.im, bins
        d.close()
    (h32, u32), (h64, u64) = fields["fp32"], fields["fp64"]
    out["fp32_vs_fp64"] = out["fp32"]["value"] / out["fp64"]["value"]
    out["errors_vs_fp64_after_1h"] = {
        "hs_rmse_rel": float(np.sqrt(np.mean((h32 - h64) ** 2)) / np.sqrt(np.mean(h64 ** 2))),
        "ustar_max_rel": float(np.max(np.abs(u32 - u64) / np.maximum(np.abs(u64), 1e-3)))}
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=40)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--hours", type=int, default=6)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--small", action="store_true", help="quarter-size grid for smoke runs")
    ap.add_argument("--only", choices=["fp32", "fp64"], default=None,
                    help="run one path only (profiling runs); prints that path's line")
    ap.add_argument("--no-tune", action="store_true", help="skip the tile autotuner")
    ap.add_argument("--tile", type=int, nargs=5, default=None,
                    metavar=("D1", "W1", "D2", "W2", "NB"),
                    help="fixed tile shape (implies --no-tune)")
    ap.add_argument("--no-finish", action="store_true",
                    help="skip finishing the simulated hours (profiling runs)")
    ap.add_argument("--no-c5", action="store_true", help="skip the hurricane member leg")
    ap.add_argument("--c3-members", type=int, default=64,
                    help="members of the batched ensemble leg (0 skips it)")
    ap.add_argument("--copy-probe", action="store_true",
                    help="run a 1 GiB device-to-device copy first (scales HBM counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build runs on one GPU (how the path would shard over "
              "several is a design note only), so --gpus must be 1", file=sys.stderr)
        sys.exit(2)
    if args.copy_probe:
        import torch
        a = torch.ones(1 << 28, dtype=torch.float32, device="cuda")
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        log("bench: copy probe 3 x 1 GiB done")
    mm, nm = (720, 360) if not args.small else (180, 90)
    log(f"bench: building c2 {mm}x{nm}")
    wl = configs.c2(seed=2, mm=mm, nm=nm, hours=args.hours)

    res = {}
    paths = ((8, "fp64"), (4, "fp32"))
    if args.only:
        paths = tuple(x for x in paths if x[1] == args.only)
    for prec, tag in paths:
        log(f"bench: {tag} create")
        d = make_domain(wl, prec)
        bins = d.im * d.om * d.pm
        log(f"bench: {tag} timed {args.steps} substeps of {bins} bins")
        if args.tile:
            d.set_tile(*args.tile)
        wall, kms, tile = timed_steps(d, args.warmup, args.steps,
                                      tune=not (args.no_tune or args.tile))
        tile = tile or (tuple(args.tile) if args.tile else None)
        rate = args.steps * bins / wall
        # prognostic bins (o <= oc, the rest is the diagnostic tail), for
        # information: sum over sea points of (oc + 1) pm, oc 0-based
        oc = np.asarray(d.get("oc"))[: d.im].astype(np.int64)
        prog = int((oc + 1).sum()) * d.pm
        log(f"bench: {tag} {rate / 1e9:.2f} G bin-updates/s, kernels ms "
            + ", ".join(f"{k} {v:.3f}" for k, v in zip(KERNELS, kms)))
        if not args.no_finish:
            log(f"bench: {tag} finishing {args.hours} h")
            t0 = time.perf_counter()
            finish_run(d, wl, args.hours)
            log(f"bench: {tag} run done in {time.perf_counter() - t0:.1f} s")
        res[tag] = dict(rate=rate, wall=wall, kms=kms, bins=bins, im=d.im, prog=prog,
                        hs=d.get("ht"), tm=d.get("mwp"), ust=d.get("ustar"),
                        us=d.get("us")[0], vs=d.get("vs")[0], tile=tile)
        d.close()

    if args.only:
        r = res[args.only]
        word = 4 if args.only == "fp32" else 8
        print(json.dumps({"path": args.only, "value": r["rate"], "unit": "bin-updates/s",
                          "kernel_ms": dict(zip(KERNELS, r["kms"])), "tile": r["tile"],
                          "roofline": roofline(r, word, None)}), flush=True)
        return

    r32, r64 = res["fp32"], res["fp64"]
    hs_ref = r64["hs"]
    hs_rmse = float(np.sqrt(np.mean((r32["hs"] - hs_ref) ** 2)) / np.sqrt(np.mean(hs_ref ** 2)))

    def maxrel(a, b, floor):
        return float(np.max(np.abs(a - b) / np.maximum(np.abs(b), floor)))

    errs = {
        "hs_rmse_rel": hs_rmse,
        "hs_max_rel": maxrel(r32["hs"], hs_ref, 1e-3),
        "tm_max_rel": maxrel(r32["tm"], r64["tm"], 1e-3),
        "ustar_max_rel": maxrel(r32["ust"], r64["ust"], 1e-3),
        "stokes_surface_rmse_rel": float(
            np.sqrt(np.mean((np.hypot(r32["us"], r32["vs"]) - np.hypot(r64["us"], r64["vs"])) ** 2))
            / max(np.sqrt(np.mean(np.hypot(r64["us"], r64["vs"]) ** 2)), 1e-30)),
    }

    out = {
        "metric": METRIC,
        "value": r32["rate"],
        "unit": "bin-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * r32["wall"] / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": ("c2: 720x360 global 0.5 deg lat-lon, om=37 pm=36, GRF land 30%, "
                                "log-normal depth, GRF winds 0-25 m/s, JONSWAP start, "
                                f"{args.hours} h") if not args.small else "c2 quarter size",
                   "sea_points": r32["im"], "bins_per_step": r32["bins"],
                   "prognostic_bins_per_step": r32["prog"],
                   "chosen": "fp32 fused", "plain": "fp64 fused"},
        "kernel_ms": dict(zip(KERNELS, [round(x, 4) for x in r32["kms"]])),
        "tile": r32["tile"],
        "plain": {"value": r64["rate"], "dtype": "f64",
                  "ms_per_step": 1e3 * r64["wall"] / args.steps,
                  "kernel_ms": dict(zip(KERNELS, [round(x, 4) for x in r64["kms"]])),
                  "tile": r64["tile"], "roofline": roofline(r64, 8, None)},
        "chosen_vs_plain": r32["rate"] / r64["rate"],
        "errors_vs_fp64": errs,
        "roofline": roofline(r32, 4, profile_traffic()),
        "roofline_substep": roofline_substep(r32, 4),
    }
    if args.c3_members > 0:
        log("bench: batched ensemble leg")
        out["batched_c3"] = batched_leg(args.c3_members)
        # the same draw with grids shrunk 4x per side: members too small to
        # fill the chip alone, where one launch for all of them pays
        log("bench: batched ensemble leg, small members")
        out["batched_c3_small"] = batched_leg(args.c3_members, scale=0.25)
    if not args.no_c5:
        log("bench: hurricane member leg")
        out["hurricane_c5"] = hurricane_leg(10, 2)
    if not args.no_cpu:
        log("bench: cpu baseline")
        out["cpu_baseline"] = cpu_baseline(wl)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
