"""Tile-shape autotuner for the fused kernels.

k_rates and k_integrate give each thread D consecutive directions of one
sea point and W = ceil(pm / D) waves per workgroup; k_advect keeps NB
directions of gathers in flight per wave.  Larger D means fewer
cross-wave exchanges and more work per thread but more registers; larger
NB means more loads in flight but fewer resident waves.  The best shape
depends on pm, the precision and the grid size, so it is measured: each
candidate runs a few live substeps with HIP-event timing and the fastest
shape per kernel is kept.  Candidates only change the order of
directional sums, so the model state advances normally while tuning.
"""
import math

RATES_D = (6, 9, 18, 36)
INTEGRATE_D = (3, 6, 9, 12, 18)
ADVECT_NB = (2, 4)


def candidates(pm, ds):
    out = []
    for D in ds:
        W = math.ceil(pm / D)
        if 1 <= W <= 8:
            out.append((D, W))
    return out


def autotune(dom, steps=3, verbose=None):
    """Chooses (D1, W1, D2, W2, NB) for `dom` by timing live substeps;
    returns the chosen tuple and the per-candidate mean kernel times (ms:
    rates, integrate, advect, stress)."""
    pm = dom.pm
    c1 = candidates(pm, RATES_D)
    c2 = candidates(pm, INTEGRATE_D)
    n = max(len(c1), len(c2), len(ADVECT_NB))
    table = []
    dom.set_timing(True)
    for k in range(n):
        a = c1[min(k, len(c1) - 1)]
        b = c2[min(k, len(c2) - 1)]
        nb = ADVECT_NB[min(k, len(ADVECT_NB) - 1)]
        dom.set_tile(a[0], a[1], b[0], b[1], nb)
        dom.timing(reset=True)
        for _ in range(steps):
            dom.substep()
        ms, cnt = dom.timing(reset=True)
        t = [m / max(c, 1) for m, c in zip(ms, cnt)]
        table.append((a, b, nb, t))
        if verbose:
            verbose(f"autotune: rates {a} integrate {b} advect NB={nb}: "
                    + ", ".join(f"{x:.3f}" for x in t) + " ms")
    dom.set_timing(False)
    best1 = min(c1, key=lambda a: min(t[0] for x, _, _, t in table if x == a))
    best2 = min(c2, key=lambda b: min(t[1] for _, y, _, t in table if y == b))
    best3 = min(ADVECT_NB, key=lambda nb: min(t[2] for _, _, z, t in table if z == nb))
    choice = (best1[0], best1[1], best2[0], best2[1], best3)
    dom.set_tile(*choice)
    return choice, table


def autotune_batch(batch, probe, steps=3, verbose=None):
    """Tunes a batch through one representative member `probe` (a Domain
    built like the largest member) and applies its choice to the batch."""
    choice, table = autotune(probe, steps=steps, verbose=verbose)
    batch.set_tile(*choice)
    return choice, table
