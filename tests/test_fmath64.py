"""The FP64 exp and log of the fused sweeps (umwm_amd/csrc/fmath64.hpp,
Tang's table-driven methods) built for the host from the same source and
checked against 80-bit long double (libm expl/logl, 11 bits more than
double): within 1 ulp of the correctly rounded result over the ranges the
sweeps use and beyond, special values as C's exp/log give them, and the
committed tables equal to what their generator computes from 50-digit
decimal arithmetic."""
import ctypes
import os
import subprocess
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "umwm_amd", "csrc")

SRC = r"""
#include "fmath64.hpp"
using namespace umwm::fm64;
static const double ET[kExpTabLen] = {UMWM_FM64_EXP_TABLE};
static const double LT[kLogTabLen] = {UMWM_FM64_LOG_TABLE};
static const double IT[kInvTabLen] = {UMWM_FM64_INV_TABLE};
extern "C" void fm_exp(const double* x, double* y, long n) {
  for (long i = 0; i < n; ++i) y[i] = exp_t(x[i], ET);
}
extern "C" void fm_log(const double* x, double* y, long n) {
  for (long i = 0; i < n; ++i) y[i] = log_t(x[i], LT, IT);
}
extern "C" void fm_exp_fast(const double* x, double* y, long n) {
  for (long i = 0; i < n; ++i) y[i] = exp_t<false>(x[i], ET);
}
extern "C" void fm_log_fast(const double* x, double* y, long n) {
  for (long i = 0; i < n; ++i) y[i] = log_t<false>(x[i], LT, IT);
}
"""


@pytest.fixture(scope="module")
def fm(tmp_path_factory):
    d = tmp_path_factory.mktemp("fm64")
    src, so = d / "fm.cpp", d / "libfm.so"
    src.write_text(SRC)
    subprocess.run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-I", CSRC, "-o", str(so),
                    str(src)], check=True)
    L = ctypes.CDLL(str(so))
    P = ctypes.POINTER(ctypes.c_double)

    def call(name, x):
        x = np.ascontiguousarray(x, dtype=np.float64)
        y = np.empty_like(x)
        getattr(L, name)(x.ctypes.data_as(P), y.ctypes.data_as(P), ctypes.c_long(x.size))
        return y
    return call


def ulps(y, ref):
    """|y - ref| in units of the last place of the double nearest ref."""
    u = np.spacing(np.abs(ref.astype(np.float64))).astype(np.longdouble)
    return np.abs((y.astype(np.longdouble) - ref) / u)


EXP_RANGES = [(-1.0, 1.0), (-40.0, 40.0), (-700.0, 700.0), (-1e-6, 1e-6), (-708.0, -700.0),
              (700.0, 709.7)]
LOG_RANGES = [(0.5, 2.0), (0.99, 1.01), (1 - 1e-3, 1 + 1e-3), (1e-10, 1e10), (1e-300, 1e-200),
              (1e200, 1e300)]


@pytest.mark.parametrize("lo,hi", EXP_RANGES)
def test_exp_within_one_ulp(fm, lo, hi):
    x = np.random.default_rng(11).uniform(lo, hi, 400_000)
    err = ulps(fm("fm_exp", x), np.exp(x.astype(np.longdouble)))
    assert err.max() < 1.0, err.max()


@pytest.mark.parametrize("lo,hi", LOG_RANGES)
def test_log_within_one_ulp(fm, lo, hi):
    x = np.exp(np.random.default_rng(12).uniform(np.log(lo), np.log(hi), 400_000))
    ref = np.log(x.astype(np.longdouble))
    keep = ref != 0
    err = ulps(fm("fm_log", x)[keep], ref[keep])
    assert err.max() < 1.0, err.max()


def test_exact_points(fm):
    # log of powers of two and of 1 is e ln2 rounded once; exp(0) = 1
    e = np.arange(-1074, 1024)
    x = np.ldexp(1.0, e)
    ref = (e.astype(np.longdouble) * np.log(np.longdouble(2))).astype(np.float64)
    assert np.array_equal(fm("fm_log", x), ref)
    assert fm("fm_log", np.array([1.0]))[0] == 0.0
    assert fm("fm_exp", np.array([0.0, -0.0])).tolist() == [1.0, 1.0]


def test_special_values(fm):
    xl = np.array([0.0, -0.0, -1.0, -np.inf, np.inf, np.nan, 5e-324, 2.2e-308])
    with np.errstate(all="ignore"):
        ref = np.log(xl)
    got = fm("fm_log", xl)
    assert np.array_equal(got, ref, equal_nan=True), (got, ref)
    xe = np.array([-np.inf, np.inf, np.nan, -746.0, -745.2, 709.79, 709.8, 1e4, -1e4])
    with np.errstate(all="ignore"):
        ref = np.exp(xe)
    got = fm("fm_exp", xe)
    assert np.array_equal(got, ref, equal_nan=True), (got, ref)


def test_exp_of_log_round_trip(fm):
    # the pair the sweeps use for powers: exp(p log x) against x^p
    x = np.exp(np.random.default_rng(13).uniform(-60, 10, 200_000))
    p = 2.4
    got = fm("fm_exp", p * fm("fm_log", x))
    ref = x.astype(np.longdouble) ** np.longdouble(p)
    rel = np.abs((got.astype(np.longdouble) - ref) / ref).astype(np.float64)
    # p |log x| ulp of the exponent, plus the two roundings
    assert np.all(rel <= (p * np.abs(np.log(x)) + 2) * 2.3e-16)


def test_tables_match_generator():
    sys.path.insert(0, CSRC)
    try:
        import gen_fmath64
    finally:
        sys.path.pop(0)
    with open(os.path.join(CSRC, "fmath64_tables.inc")) as fh:
        assert fh.read() == gen_fmath64.render(gen_fmath64.tables())


def test_unchecked_variants_on_their_domain(fm):
    # the sweeps' variants: the same bits wherever they are defined (exp of
    # anything but NaN, log of finite x >= 0), log 0 = -inf and exp(-inf) = 0
    rng = np.random.default_rng(14)
    xe = np.concatenate([rng.uniform(-800, 800, 100_000), [-np.inf, np.inf, 0.0]])
    assert np.array_equal(fm("fm_exp_fast", xe), fm("fm_exp", xe))
    xl = np.concatenate([np.exp(rng.uniform(-745, 709, 100_000)), [0.0, 5e-324, 1.0]])
    with np.errstate(all="ignore"):
        assert np.array_equal(fm("fm_log_fast", xl), fm("fm_log", xl))
    assert fm("fm_exp_fast", fm("fm_log_fast", np.array([0.0])) * 2.4)[0] == 0.0
