// Learned source-term surrogate: a network (mlp.hpp) that maps one sea
// point's spectrum and forcing to the combined rate of wind input, whitecap
// dissipation and bottom friction, R = ssin - sds - sbf (1/s per bin; sds
// after its cothkd factor, as src/umwm_physics.F90:59-68 combines them).
//
// Inputs per point (om pm + 3): ln max(e, 1e-30) of every bin (bin b =
// o pm + p), the wind components u10, v10 and ln depth, each normalised by
// the mean and standard deviation of the training samples.  Outputs (om
// pm): y_b = asinh(R_b / (2 pi f_o s0)), s0 = 1e-7, normalised the same
// way; the inference epilogue undoes both and stores R itself.
//
// Training samples come from the plain FP32 (or FP64) source kernels of
// domains the caller runs (umwm_hip_surrogate_collect): the operator
// kernels evaluate ssin, sds and the sbf table on the current spectrum.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

#include "../../include/umwm_hip.h"
#include "mlp.hpp"

namespace umwm {

template <typename T>
struct DomainImpl;

constexpr double kSurS0 = 1e-7;
constexpr float kSurLogFloor = 1e-30f;

// Device statistics of one attached domain, accumulated over substeps.
enum SurStat {
  SS_POINTS = 0,      // point-substeps evaluated
  SS_USED,            // of which took the surrogate's rates
  SS_FALL_RANGE,      // fell back: inputs outside the training range
  SS_FALL_RESID,      // fell back: implied dissipation negative beyond tolerance
  SS_FALL_AUDIT,      // fell back: audited and found wrong beyond tolerance
  SS_AUDITED,         // point-substeps audited by the plain kernel
  SS_AUDIT_NUM,       // sum over audited points of sum_b |R - R_plain| e kdk
  SS_AUDIT_DEN,       // ... of sum_b (|ssin| + sds_plain + sbf) e kdk
  SS_N
};

struct SurrogateImpl {
  int om = 0, pm = 0, nin = 0, nout = 0, hidden = 0, layers = 0;
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
  // frequencies of the served grid (fp32 [om]) and per output 2 pi f s0
  float *freq = nullptr, *cscale = nullptr;
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
