// Learned source-term surrogate: a network (mlp.hpp) that maps one sea
// point's spectrum and forcing to the one non-local factor of the combined
// source rate src/umwm_physics.F90:59-68 integrates, R = ssin - sds - sbf.
// Whitecap dissipation is sds = dm^2 cf coth (e k^4)^p, a closed form of
// the bin itself times the directional spreading factor dm^2 = (1 +
// mss_fac mss)^2, where mss is the cos^2-weighted convolution of the whole
// spectrum below the bin's frequency (src/umwm_source_functions.F90:85-105);
// ssin and sbf are closed forms of the bin and the forcing.  The network
// learns ln dm^2 for every bin, and the kernels form R from it and the
// closed forms.  Learning R itself, or ln sds, was measured to fail: R is
// dominated by the wind input where sds is a small remainder, and ln sds
// spans so many decades per bin that a small normalised error is a large
// relative one (0.43 energy-weighted error in sds on a fresh seed).
//
// Inputs per point (om pm + 3): ln max(e, 1e-30) of every bin (bin b =
// o pm + p), the wind components u10, v10 and ln depth, each normalised by
// the mean and standard deviation of the training samples.  Outputs (om
// pm): ln dm^2, normalised the same way; the inference epilogue undoes both
// and stores dm^2 - 1 as bf16.
//
// Training samples come from the plain FP32 (or FP64) source kernels of
// domains the caller runs (umwm_hip_surrogate_collect): the operator
// kernels evaluate sds_d12's spreading factor on the current spectrum.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

#include "../../include/umwm_hip.h"
#include "mlp.hpp"

namespace umwm {

template <typename T>
struct DomainImpl;

constexpr float kSurLogFloor = 1e-30f;
// Floor of a target column's standard deviation in the normalisation:
// 0.005 in ln dm^2, a 0.5 % change of the spreading factor.  Columns that
// barely vary (the lowest frequencies, whose prefix holds no energy) then
// do not weigh their noise into the loss; the energetic bins about the
// spectral peak, whose dm^2 stays within a few per cent of 1, keep their
// weight.
constexpr double kSurTargetSdFloor = 0.005;
// Floor of an input column's standard deviation: 0.1 in ln e, a 10 % change
// of a bin's energy.  The spectral inputs that vary at all vary by units of
// ln e; below the floor a column's spread is rounding.
constexpr double kSurInputSdFloor = 0.1;
// Relative widening of the spreading-factor bounds of the residual check
// (fused.hip k_rates): the network's dm^2 may stray this far outside [1,
// (1 + 2 mss_fac m0)^2] in a bin before the bin counts as violating.
constexpr float kSurBoundTol = 0.25f;

// Device statistics of one attached domain, accumulated over substeps: the
// counts (SS_POINTS .. SS_AUDITED) as unsigned 64-bit integers, the two
// audit sums as doubles, in one array of SS_N 8-byte words.
enum SurStat {
  SS_POINTS = 0,      // point-substeps evaluated
  SS_USED,            // of which took the surrogate's rates
  SS_FALL_RANGE,      // fell back: inputs outside the training range
  SS_FALL_RESID,      // fell back: implied spreading factor outside its bounds
  SS_FALL_AUDIT,      // fell back: audited and found wrong beyond tolerance
  SS_AUDITED,         // point-substeps audited by the plain kernel
  SS_AUDIT_NUM,       // sum over audited points of sum_b |sds_net - sds| e kdk
  SS_AUDIT_DEN,       // ... of sum_b sds e kdk (sds_net = dm^2 from the network)
  SS_N
};

struct SurrogateImpl {
  int om = 0, pm = 0, nin = 0, nout = 0, hidden = 0, layers = 0;
  double fmin = 0, fmax = 0;  // frequency grid the network serves
  int batch = 0, chunk = 0;
  int64_t capacity = 0, count = 0;
  uint64_t seed = 0;
  Mlp net;
  hipStream_t stream = nullptr;
  std::vector<void*> allocs;
  // training set: raw inputs fp32 [capacity][nin], raw targets fp32
  // [capacity][nout]
  float *feat = nullptr, *targ = nullptr;
  // normalisation: input and target mean and 1/std (fp32, per column)
  float *fmu = nullptr, *fis = nullptr, *tmu = nullptr, *tsd = nullptr;
  bool stats_ready = false;
  int64_t stats_rows = 0;
  // training range of the forcing and of Hs (min, max of wind speed, depth,
  // Hs), for the out-of-range check
  float range[6] = {0, 0, 0, 0, 0, 0};
  float* range_d = nullptr;
  // inference: the normalised forcing inputs (u10, v10, ln depth) of one
  // chunk of points, fp32 [3][chunk]
  float* ext = nullptr;
  // training minibatch: row indices, normalised targets
  int32_t* idx = nullptr;
  float* tbatch = nullptr;
  double* loss_hist = nullptr;
  int loss_cap = 0;
  ~SurrogateImpl();
};

// Normalisation statistics and training ranges over the first n samples.
int surrogate_stats(SurrogateImpl& S, int64_t n);
// Appends one training sample per sea point of D at its current state
// (the operator kernels evaluate the source arrays first).
template <typename T>
int surrogate_collect(SurrogateImpl& S, DomainImpl<T>& D, int64_t* total);

}  // namespace umwm

struct umwm_surrogate {
  umwm::SurrogateImpl* impl = nullptr;
};
