"""Host logic of bench.py (no GPU): the per-kernel traffic it reads from the
committed rocprofv3 summaries, the roofline of the dominant kernel, and the
error measures of a path's fields against FP64."""
import importlib.util
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def load_bench():
    spec = importlib.util.spec_from_file_location("bench", os.path.join(ROOT, "bench.py"))
    b = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(b)
    return b


def test_profile_traffic_reads_both_summaries():
    b = load_bench()
    t32 = b.profile_traffic(b.PROFILE)
    t64 = b.profile_traffic(b.PROFILE64)
    for t in (t32, t64):
        assert set(t) == {"k_advect", "k_integrate", "k_rates"}
    # measured traffic is at least the algorithmic bytes of each kernel
    bins = 5574405816
    for t, word in ((t32, 4), (t64, 8)):
        assert t["k_advect"] >= 3 * word * bins
        assert t["k_rates"] >= word * bins
    assert b.profile_traffic(os.path.join(ROOT, "no_such_profile.md")) == {}


def test_roofline_takes_the_dominant_kernel():
    b = load_bench()
    r = b.roofline([10.0, 20.0, 15.0, 1.0], 1_000_000_000, 4, {"k_integrate": 9e9})
    assert r["kernel"] == "k_integrate" and r["traffic"] == 9e9
    # k_integrate moves two words per bin: 8e9 bytes in 20 ms
    assert abs(r["achieved"] - 400.0) < 1e-9 and abs(r["frac"] - 0.05) < 1e-12
    r = b.roofline([10.0, 20.0, 25.0, 1.0], 1_000_000_000, 8)
    assert r["kernel"] == "k_advect" and r["traffic"] is None


def test_field_errors_against_closed_forms():
    b = load_bench()
    n = 1000
    ref = {"ht": np.full(n, 2.0), "mwp": np.full(n, 8.0), "ustar": np.full(n, 0.5),
           "us": np.full(n, 0.1)}
    f = {k: v * 1.01 for k, v in ref.items()}
    f["ht"][0] = np.nan
    e = b.field_errors(f, ref)
    assert abs(e["ustar_rmse_rel"] - 0.01) < 1e-12
    assert abs(e["stokes_surface_rmse_rel"] - 0.01) < 1e-12
    assert abs(e["tm_max_rel"] - 0.01) < 1e-12
    assert e["nonfinite_points"] == 1
