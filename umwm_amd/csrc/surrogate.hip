// Learned source-term surrogate (surrogate.hpp): the device training set
// filled from the plain source kernels, normalisation, the training loop on
// the BF16 MFMA kernels of mlp.hip, the per-substep inference that leaves
// the network's rates where the fused kernels read them, and the C-ABI of
// all of it.  No counterpart in the reference, whose source terms are the
// closed forms of src/umwm_source_functions.F90:16-105 combined in
// src/umwm_physics.F90:59-68; the surrogate stands in for Sin + Sds + Sbf
// there and the plain kernels judge it (fused.hip, k_rates).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "domain.hpp"
#include "surrogate.hpp"

namespace umwm {

SurrogateImpl::~SurrogateImpl() {
  if (stream) (void)hipStreamSynchronize(stream);
  for (void* p : allocs) (void)hipFree(p);
  if (stream) (void)hipStreamDestroy(stream);
}

namespace {

template <typename U>
U* salloc(SurrogateImpl& S, size_t n) {
  void* p = nullptr;
  if (hipMalloc(&p, std::max<size_t>(n, 1) * sizeof(U)) != hipSuccess) return nullptr;
  S.allocs.push_back(p);
  return static_cast<U*>(p);
}

inline dim3 g1(int64_t n, int t = 256) { return dim3((unsigned)((n + t - 1) / t)); }

__global__ void k_fill_f32(float* p, int64_t n, float v) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}
__global__ void k_fill_f64(double* p, int64_t n, double v) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}
__global__ void k_fill_i32(int32_t* p, int64_t n, int32_t v) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

// One training sample per sea point: the raw inputs (ln e of every bin,
// u10, v10, ln depth) and the raw targets ln dm^2, the directional
// spreading factor (1 + mss_fac mss)^2 that sds_d12 just left in `dummy`
// (src/umwm_source_functions.F90:92-99).
template <typename T>
__global__ void k_sur_collect(DomainDev<T> d, float* __restrict__ feat, float* __restrict__ targ,
                              int64_t row0, int64_t cap, int nin, int nout) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= d.im) return;
  const int64_t row = row0 + j;
  if (row >= cap) return;
  const int nb = d.om * d.pm;
  const size_t ldi = d.ldi;
  float* f = feat + row * nin;
  float* t = targ + row * nout;
  for (int b = 0; b < nb; ++b) {
    const size_t q = (size_t)b * ldi + j;
    f[b] = logf(fmaxf((float)d.e[q], kSurLogFloor));
    t[b] = (float)log(fmax((double)d.dummy[q], 1.0));
  }
  const double w = d.wspd[j], wd = d.wdir[j];
  f[nb] = (float)(w * cos(wd));
  f[nb + 1] = (float)(w * sin(wd));
  f[nb + 2] = (float)log((double)d.depth[j]);
}

// Column sums of a [rows][ld] matrix shifted by the column's first value,
// sum (x - x0) and sum (x - x0)^2, a chunk of rows per block row y
// (accumulated in double).  The shift keeps the variance free of the
// cancellation that E[x^2] - E[x]^2 suffers when the mean is large beside
// the spread: over 2 M rows of a column pinned at ln 1e-30 = -69.08 it left
// a spurious standard deviation of 6e-6, whose inverse (1.6e5) turned half
// a float32 ulp of ln e into 0.6 of a normalised input.
__global__ void k_colsums(const float* __restrict__ a, int64_t rows, int cols, int ld,
                          int64_t chunk, double* __restrict__ s1, double* __restrict__ s2) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= cols) return;
  const int64_t r0 = (int64_t)blockIdx.y * chunk, r1 = std::min(rows, r0 + chunk);
  const double x0 = a[c];
  double a1 = 0, a2 = 0;
  for (int64_t r = r0; r < r1; ++r) {
    const double v = (double)a[r * ld + c] - x0;
    a1 += v;
    a2 += v * v;
  }
  atomicAdd(s1 + c, a1);
  atomicAdd(s2 + c, a2);
}

// Mean and standard deviation (or its inverse) per column from the shifted
// sums of k_colsums.  A standard deviation below `floor` is raised to it,
// and a column that does not vary at all gets 1: an input column whose ln e
// barely varies (a bin at the ln 1e-30 floor in nearly every sample) would
// otherwise amplify rounding into its normalised input, and a target
// column that barely varies (ln dm^2 of the lowest frequencies, whose
// prefix holds little energy) would weigh its noise into the loss as much
// as the columns that matter.
__global__ void k_colfinal(const float* __restrict__ a, const double* __restrict__ s1,
                           const double* __restrict__ s2, int cols, int64_t rows,
                           float* __restrict__ mu, float* __restrict__ sd, int inv, double floor) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= cols) return;
  const double d = s1[c] / (double)rows;
  const double v = s2[c] / (double)rows - d * d;
  double s = v > 0 ? sqrt(v) : 0;
  if (s == 0) s = 1;
  s = fmax(s, floor);
  mu[c] = (float)((double)a[c] + d);
  sd[c] = (float)(inv ? 1 / s : s);
}

__device__ __forceinline__ uint64_t splitmix(uint64_t x) {
  x += 0x9E3779B97F4A7C15ull;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
  return x ^ (x >> 31);
}

// Training or evaluation rows: row r of the batch is sample base + r
// (evaluation, hashed == 0) or a hashed draw from [0, nrows) (training).
__global__ void k_sur_gather(const float* __restrict__ feat, const float* __restrict__ targ,
                             int rows, int nin, int nout, int K0, int N3, const float* fmu,
                             const float* fis, const float* tmu, const float* tsd,
                             bf16* __restrict__ x, float* __restrict__ tb, int64_t base,
                             int hashed, uint64_t seed, int64_t nrows) {
  const int r = blockIdx.y;
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= rows) return;
  const int64_t row = hashed ? (int64_t)(splitmix(seed ^ (uint64_t)r * 0x632BE59BD9B4E019ull) %
                                         (uint64_t)nrows)
                             : base + r;
  if (c < K0)
    x[(int64_t)r * K0 + c] = c < nin ? (bf16)((feat[row * nin + c] - fmu[c]) * fis[c]) : (bf16)0.f;
  if (c < N3)
    tb[(int64_t)r * N3 + c] = c < nout ? (targ[row * nout + c] - tmu[c]) / tsd[c] : 0.f;
}

// Inference inputs, per point: wind and depth (normalised) into ext
// [3][ldx], and the range verdict on wind speed and depth.  The spectral
// inputs are formed by the first layer's GEMM from the spectrum itself.
template <typename T>
__global__ void k_sur_point(DomainDev<T> d, int j0, int rows, int nb,
                            const float* __restrict__ fmu, const float* __restrict__ fis,
                            const float* __restrict__ range, float margin,
                            float* __restrict__ ext, int64_t ldx, int32_t* __restrict__ ok) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= rows) return;
  const int j = j0 + r;
  float w = (float)d.wspd[j];
  w = w > 1e-2f ? w : 1e-2f;  // as sin_d12 floors it (src/umwm_source_functions.F90:22)
  const float wd = (float)d.wdir[j], dep = (float)d.depth[j];
  const float in3[3] = {w * cosf(wd), w * sinf(wd), logf(dep)};
  for (int c = 0; c < 3; ++c) ext[c * ldx + r] = (in3[c] - fmu[nb + c]) * fis[nb + c];
  const float lo = 1.f - margin, hi = 1.f + margin;
  ok[j] = (w >= range[0] * lo && w <= range[1] * hi && dep >= range[2] * lo &&
           dep <= range[3] * hi) ? 1 : 0;
}

__global__ void k_store_loss(double* acc, double* out) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    *out = *acc;
    *acc = 0;
  }
}

__global__ void k_f32_to_bf16_pad(const float* __restrict__ a, int rows, int cols, int ld,
                                  bf16* __restrict__ x) {
  const int r = blockIdx.y, c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c < ld && r < rows) x[(int64_t)r * ld + c] = c < cols ? (bf16)a[(int64_t)r * cols + c] : (bf16)0.f;
}
__global__ void k_f32_pad(const float* __restrict__ a, int rows, int cols, int ld,
                          float* __restrict__ t) {
  const int r = blockIdx.y, c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c < ld && r < rows) t[(int64_t)r * ld + c] = c < cols ? a[(int64_t)r * cols + c] : 0.f;
}

int pad8(int v) { return (v + 7) / 8 * 8; }

}  // namespace

// Normalisation statistics and training ranges over the first n samples.
int surrogate_stats(SurrogateImpl& S, int64_t n) {
  if (n < 2) return fail(UMWM_ERR_STATE, "surrogate: fewer than two training samples collected");
  hipStream_t s = S.stream;
  const int cols = std::max(S.nin, S.nout);
  std::vector<void*> tmp;
  double *s1 = nullptr, *s2 = nullptr;
  UMWM_HIP_TRY(hipMalloc(&s1, cols * sizeof(double)));
  UMWM_HIP_TRY(hipMalloc(&s2, cols * sizeof(double)));
  const int64_t chunk = 4096, nch = (n + chunk - 1) / chunk;
  auto one = [&](const float* a, int nc, float* mu, float* sd, int inv, double floor) -> int {
    hipLaunchKernelGGL(k_fill_f64, g1(nc), dim3(256), 0, s, s1, (int64_t)nc, 0.0);
    hipLaunchKernelGGL(k_fill_f64, g1(nc), dim3(256), 0, s, s2, (int64_t)nc, 0.0);
    hipLaunchKernelGGL(k_colsums, dim3((nc + 255) / 256, (unsigned)nch), dim3(256), 0, s, a, n, nc,
                       nc, chunk, s1, s2);
    hipLaunchKernelGGL(k_colfinal, g1(nc), dim3(256), 0, s, a, (const double*)s1,
                       (const double*)s2, nc, n, mu, sd, inv, floor);
    return hipGetLastError() == hipSuccess ? 0 : fail(UMWM_ERR_HIP, "surrogate: statistics launch");
  };
  int rc = one(S.feat, S.nin, S.fmu, S.fis, 1, kSurInputSdFloor);
  if (!rc) rc = one(S.targ, S.nout, S.tmu, S.tsd, 0, kSurTargetSdFloor);
  // training range of wind speed and depth (host pass over two columns)
  std::vector<float> uv((size_t)n * 3);
  if (!rc) {
    UMWM_HIP_TRY(hipStreamSynchronize(s));
    UMWM_HIP_TRY(hipMemcpy2D(uv.data(), 3 * sizeof(float), S.feat + (S.nin - 3),
                             S.nin * sizeof(float), 3 * sizeof(float), n, hipMemcpyDeviceToHost));
  }
  (void)hipFree(s1);
  (void)hipFree(s2);
  if (rc) return rc;
  float wlo = INFINITY, whi = -INFINITY, dlo = INFINITY, dhi = -INFINITY;
  for (int64_t r = 0; r < n; ++r) {
    const float w = std::hypot(uv[3 * r], uv[3 * r + 1]), d = std::exp(uv[3 * r + 2]);
    wlo = std::min(wlo, w);
    whi = std::max(whi, w);
    dlo = std::min(dlo, d);
    dhi = std::max(dhi, d);
  }
  S.range[0] = wlo; S.range[1] = whi; S.range[2] = dlo; S.range[3] = dhi;
  S.stats_ready = true;
  S.stats_rows = n;
  return 0;
}

int launch_surrogate_inference(DomainImpl<float>& D, hipStream_t s) {
  SurrogateImpl& S = *D.sur;
  const int nb = S.nout, im = D.h.im;
  for (int j0 = 0; j0 < im; j0 += S.chunk) {
    const int rows = std::min(S.chunk, im - j0);
    hipLaunchKernelGGL(k_sur_point<float>, g1(rows), dim3(256), 0, s, D.h, j0, rows, nb,
                       (const float*)S.fmu, (const float*)S.fis, (const float*)S.range_d,
                       (float)D.sur_margin, S.ext, (int64_t)S.chunk, D.h.sur_ok);
    UMWM_HIP_TRY(hipGetLastError());
    // the first layer reads the spectrum itself (ln e, normalised, with the
    // zmax check), the last writes dm^2 into the [bin][point] array
    SpecOperand a;
    a.e = D.h.e + j0;
    a.lde = (int64_t)D.h.ldi;
    a.cols = (int)(D.h.ldi - j0);
    a.nb = nb;
    a.mu = S.fmu;
    a.is = S.fis;
    a.ext = S.ext;
    a.ldx = S.chunk;
    a.zmax = (float)D.sur_zmax;
    a.ok = D.h.sur_ok + j0;
    bf16* out = reinterpret_cast<bf16*>(const_cast<uint16_t*>(D.h.sur_r)) + j0;
    if (int rc = mlp_infer_spectrum(S.net, a, rows, out, D.h.ldi, s, S.tmu, S.tsd)) return rc;
  }
  return 0;
}

// Appends one sample per sea point of the domain at its current state.
template <typename T>
int surrogate_collect(SurrogateImpl& S, DomainImpl<T>& D, int64_t* total) {
  if (D.h.om != S.om || D.h.pm != S.pm || D.S.P.fmin != S.fmin || D.S.P.fmax != S.fmax)
    return fail(UMWM_ERR_ARGUMENT,
                "surrogate_collect: the domain's spectral grid (om, pm, fmin, fmax) differs "
                "from the network's");
  if (!D.ops_ready) return fail(UMWM_ERR_STATE, "surrogate_collect: operator arrays missing");
  int rc = launch_op_sin(D);
  if (!rc) rc = launch_op_sds(D);
  if (!rc) rc = launch_op_snl(D);
  if (rc) return rc;
  const int64_t take = std::min<int64_t>(D.h.im, S.capacity - S.count);
  if (take > 0) {
    hipLaunchKernelGGL(k_sur_collect<T>, g1(D.h.im), dim3(256), 0, D.stream, D.h, S.feat, S.targ,
                       S.count, S.capacity, S.nin, S.nout);
    UMWM_HIP_TRY(hipGetLastError());
    S.count += take;
    S.stats_ready = false;
  }
  UMWM_HIP_TRY(hipStreamSynchronize(D.stream));
  if (total) *total = S.count;
  return 0;
}
template int surrogate_collect<float>(SurrogateImpl&, DomainImpl<float>&, int64_t*);
template int surrogate_collect<double>(SurrogateImpl&, DomainImpl<double>&, int64_t*);

}  // namespace umwm

// =====================================================================
// extern "C"
// =====================================================================
using namespace umwm;

namespace {

int sur_check(umwm_surrogate* s) {
  return (s && s->impl) ? 0 : fail(UMWM_ERR_ARGUMENT, "null surrogate");
}

// Copies layer l's padded device weights to or from unpadded host arrays.
int layer_copy(SurrogateImpl& S, int l, float* w, float* b, const float* wi, const float* bi) {
  if (l < 0 || l >= (int)S.net.L.size()) return fail(UMWM_ERR_ARGUMENT, "surrogate: no such layer");
  const MlpLayerDev& L = S.net.L[l];
  std::vector<int> dims = {S.nin};
  for (int k = 0; k < S.layers; ++k) dims.push_back(S.hidden);
  dims.push_back(S.nout);
  const int K = dims[l], N = dims[l + 1];
  UMWM_HIP_TRY(hipStreamSynchronize(S.stream));
  if (w || b) {
    std::vector<float> hw((size_t)L.N * L.K), hb(L.N);
    UMWM_HIP_TRY(hipMemcpy(hw.data(), L.w, hw.size() * 4, hipMemcpyDeviceToHost));
    UMWM_HIP_TRY(hipMemcpy(hb.data(), L.b, hb.size() * 4, hipMemcpyDeviceToHost));
    if (w)
      for (int n = 0; n < N; ++n) std::memcpy(w + (size_t)n * K, hw.data() + (size_t)n * L.K, K * 4);
    if (b) std::memcpy(b, hb.data(), N * 4);
  }
  if (wi || bi) {
    std::vector<float> hw((size_t)L.N * L.K, 0.f), hb(L.N, 0.f);
    if (wi)
      for (int n = 0; n < N; ++n) std::memcpy(hw.data() + (size_t)n * L.K, wi + (size_t)n * K, K * 4);
    if (bi) std::memcpy(hb.data(), bi, N * 4);
    if (wi) UMWM_HIP_TRY(hipMemcpy(L.w, hw.data(), hw.size() * 4, hipMemcpyHostToDevice));
    if (bi) UMWM_HIP_TRY(hipMemcpy(L.b, hb.data(), hb.size() * 4, hipMemcpyHostToDevice));
    if (int rc = mlp_sync_bf16(S.net, S.stream)) return rc;
    UMWM_HIP_TRY(hipStreamSynchronize(S.stream));
  }
  return 0;
}

// Raw inputs (n x nin, fp32 host) into the network's bf16 input rows.
int put_inputs(SurrogateImpl& S, const float* x, int n) {
  float* tmp = nullptr;
  UMWM_HIP_TRY(hipMalloc(&tmp, (size_t)n * S.nin * sizeof(float)));
  hipError_t e = hipMemcpy(tmp, x, (size_t)n * S.nin * sizeof(float), hipMemcpyHostToDevice);
  if (e == hipSuccess) {
    const int K0 = S.net.L[0].K;
    hipLaunchKernelGGL(k_f32_to_bf16_pad, dim3((K0 + 255) / 256, n), dim3(256), 0, S.stream,
                       (const float*)tmp, n, S.nin, K0, S.net.x);
    e = hipStreamSynchronize(S.stream);
  }
  (void)hipFree(tmp);
  if (e != hipSuccess) return fail(UMWM_ERR_HIP, std::string("surrogate inputs: ") + hipGetErrorString(e));
  return 0;
}

}  // namespace

extern "C" {

int umwm_hip_surrogate_create(const umwm_surrogate_config* cfg, umwm_surrogate** out) {
  if (!cfg || !out) return fail(UMWM_ERR_ARGUMENT, "null argument");
  if (cfg->om < 3 || cfg->pm < 4 || cfg->hidden < 8 || cfg->layers < 1 || cfg->layers > 8 ||
      cfg->batch < 8 || cfg->batch % 8 || cfg->capacity < 0 || cfg->chunk < 64 ||
      cfg->chunk % 64)
    return fail(UMWM_ERR_ARGUMENT, "surrogate config: om >= 3, pm >= 4, hidden >= 8, layers "
                                   "1..8, batch a positive multiple of 8, chunk a positive "
                                   "multiple of 64");
  auto* S = new SurrogateImpl();
  S->om = cfg->om;
  S->pm = cfg->pm;
  S->nout = cfg->om * cfg->pm;
  S->nin = S->nout + 3;
  S->hidden = cfg->hidden;
  S->layers = cfg->layers;
  S->batch = cfg->batch;
  S->chunk = cfg->chunk;
  S->capacity = cfg->capacity;
  S->seed = cfg->seed;
  S->fmin = cfg->fmin > 0 ? cfg->fmin : 0.0313;
  S->fmax = cfg->fmax > 0 ? cfg->fmax : 2.0;
  auto bail = [&](int rc) {
    delete S;
    return rc;
  };
  if (hipStreamCreateWithFlags(&S->stream, hipStreamNonBlocking) != hipSuccess)
    return bail(fail(UMWM_ERR_HIP, "hipStreamCreate failed"));
  std::vector<int> dims = {S->nin};
  for (int k = 0; k < S->layers; ++k) dims.push_back(S->hidden);
  dims.push_back(S->nout);
  if (int rc = mlp_create(dims, cfg->seed, std::max(S->batch, S->chunk), S->net)) return bail(rc);
  const size_t cap = (size_t)std::max<int64_t>(S->capacity, 1);
  S->feat = salloc<float>(*S, cap * S->nin);
  S->targ = salloc<float>(*S, cap * S->nout);
  S->fmu = salloc<float>(*S, S->nin);
  S->fis = salloc<float>(*S, S->nin);
  S->tmu = salloc<float>(*S, S->nout);
  S->tsd = salloc<float>(*S, S->nout);
  S->range_d = salloc<float>(*S, 8);
  S->tbatch = salloc<float>(*S, (size_t)S->batch * S->net.L.back().N);
  S->ext = salloc<float>(*S, (size_t)3 * S->chunk);
  if (!S->tbatch || !S->ext) return bail(fail(UMWM_ERR_NOMEM, "surrogate: device allocation failed"));
  *out = new umwm_surrogate{S};
  return 0;
}

int umwm_hip_surrogate_destroy(umwm_surrogate* s) {
  if (!s) return 0;
  delete s->impl;
  delete s;
  return 0;
}

int umwm_hip_surrogate_dims(const umwm_surrogate* s, int32_t out[6]) {
  if (!s || !s->impl || !out) return fail(UMWM_ERR_ARGUMENT, "null argument");
  const SurrogateImpl& S = *s->impl;
  out[0] = S.nin; out[1] = S.nout; out[2] = S.hidden; out[3] = S.layers;
  out[4] = (int32_t)std::min<int64_t>(S.count, INT32_MAX);
  out[5] = S.stats_ready ? 1 : 0;
  return 0;
}

int umwm_hip_surrogate_get_layer(umwm_surrogate* s, int32_t layer, float* w, float* b) {
  if (int rc = sur_check(s)) return rc;
  return layer_copy(*s->impl, layer, w, b, nullptr, nullptr);
}

int umwm_hip_surrogate_set_layer(umwm_surrogate* s, int32_t layer, const float* w, const float* b) {
  if (int rc = sur_check(s)) return rc;
  return layer_copy(*s->impl, layer, nullptr, nullptr, w, b);
}

int umwm_hip_surrogate_forward(umwm_surrogate* s, const float* x, int32_t n, float* y) {
  if (int rc = sur_check(s)) return rc;
  SurrogateImpl& S = *s->impl;
  if (!x || !y || n < 1 || n > S.net.cap_rows)
    return fail(UMWM_ERR_ARGUMENT, "surrogate_forward: 1 <= n <= max(batch, chunk) rows");
  if (int rc = put_inputs(S, x, n)) return rc;
  if (int rc = mlp_forward(S.net, n, S.stream)) return rc;
  const int N3 = S.net.L.back().N;
  UMWM_HIP_TRY(hipStreamSynchronize(S.stream));
  UMWM_HIP_TRY(hipMemcpy2D(y, S.nout * sizeof(float), S.net.y, N3 * sizeof(float),
                           S.nout * sizeof(float), n, hipMemcpyDeviceToHost));
  return 0;
}

int umwm_hip_surrogate_infer(umwm_surrogate* s, const float* x, int32_t n, float* y) {
  if (int rc = sur_check(s)) return rc;
  SurrogateImpl& S = *s->impl;
  if (!x || !y || n < 1 || n > S.net.cap_rows)
    return fail(UMWM_ERR_ARGUMENT, "surrogate_infer: 1 <= n <= max(batch, chunk) rows");
  if (int rc = put_inputs(S, x, n)) return rc;
  const int ldo = (n + 7) / 8 * 8;
  bf16* out = nullptr;
  UMWM_HIP_TRY(hipMalloc(&out, (size_t)S.nout * ldo * sizeof(bf16)));
  int rc = mlp_infer_transposed(S.net, n, out, ldo, S.stream);
  std::vector<uint16_t> h((size_t)S.nout * ldo);
  if (!rc && hipStreamSynchronize(S.stream) != hipSuccess) rc = fail(UMWM_ERR_HIP, "surrogate_infer");
  if (!rc && hipMemcpy(h.data(), out, h.size() * 2, hipMemcpyDeviceToHost) != hipSuccess)
    rc = fail(UMWM_ERR_HIP, "surrogate_infer: download");
  (void)hipFree(out);
  if (rc) return rc;
  for (int r = 0; r < n; ++r)
    for (int o = 0; o < S.nout; ++o) {
      const uint32_t bits = (uint32_t)h[(size_t)o * ldo + r] << 16;
      std::memcpy(y + (size_t)r * S.nout + o, &bits, 4);
    }
  return 0;
}

int umwm_hip_surrogate_gradients(umwm_surrogate* s, const float* x, const float* t, int32_t n,
                                 double* loss, float* const* gw, float* const* gb) {
  if (int rc = sur_check(s)) return rc;
  SurrogateImpl& S = *s->impl;
  if (!x || !t || n < 8 || n % 8 || n > S.net.cap_rows)
    return fail(UMWM_ERR_ARGUMENT, "surrogate_gradients: n a multiple of 8 within the buffers");
  if (int rc = put_inputs(S, x, n)) return rc;
  const int N3 = S.net.L.back().N;
  float *tt = nullptr, *tp = nullptr;
  UMWM_HIP_TRY(hipMalloc(&tt, (size_t)n * S.nout * 4));
  UMWM_HIP_TRY(hipMalloc(&tp, (size_t)n * N3 * 4));
  int rc = 0;
  if (hipMemcpy(tt, t, (size_t)n * S.nout * 4, hipMemcpyHostToDevice) != hipSuccess)
    rc = fail(UMWM_ERR_HIP, "surrogate_gradients: upload");
  if (!rc) {
    hipLaunchKernelGGL(k_f32_pad, dim3((N3 + 255) / 256, n), dim3(256), 0, S.stream,
                       (const float*)tt, n, S.nout, N3, tp);
    hipLaunchKernelGGL(k_fill_f64, dim3(1), dim3(64), 0, S.stream, S.net.loss_acc, (int64_t)1, 0.0);
    rc = mlp_forward(S.net, n, S.stream);
  }
  if (!rc) rc = mlp_loss(S.net, tp, nullptr, n, S.stream);
  if (!rc) rc = mlp_backward(S.net, n, S.stream);
  if (!rc && hipStreamSynchronize(S.stream) != hipSuccess) rc = fail(UMWM_ERR_HIP, "surrogate_gradients");
  if (!rc && loss) (void)hipMemcpy(loss, S.net.loss_acc, sizeof(double), hipMemcpyDeviceToHost);
  std::vector<int> dims = {S.nin};
  for (int k = 0; k < S.layers; ++k) dims.push_back(S.hidden);
  dims.push_back(S.nout);
  for (int l = 0; !rc && l < (int)S.net.L.size(); ++l) {
    const MlpLayerDev& L = S.net.L[l];
    const int K = dims[l], N = dims[l + 1];
    if (gw && gw[l] &&
        hipMemcpy2D(gw[l], K * 4, L.gw, L.K * 4, K * 4, N, hipMemcpyDeviceToHost) != hipSuccess)
      rc = fail(UMWM_ERR_HIP, "surrogate_gradients: download");
    if (gb && gb[l] && hipMemcpy(gb[l], L.gb, N * 4, hipMemcpyDeviceToHost) != hipSuccess)
      rc = fail(UMWM_ERR_HIP, "surrogate_gradients: download");
  }
  (void)hipFree(tt);
  (void)hipFree(tp);
  return rc;
}

int umwm_hip_surrogate_train(umwm_surrogate* s, int32_t steps, double lr, int64_t ntrain,
                             double* loss_first, double* loss_last) {
  if (int rc = sur_check(s)) return rc;
  SurrogateImpl& S = *s->impl;
  if (ntrain <= 0 || ntrain > S.count) ntrain = S.count;
  if (steps < 1 || !(lr > 0)) return fail(UMWM_ERR_ARGUMENT, "surrogate_train: steps >= 1, lr > 0");
  if (!S.stats_ready || S.stats_rows != ntrain)
    if (int rc = surrogate_stats(S, ntrain)) return rc;
  UMWM_HIP_TRY(hipMemcpy(S.range_d, S.range, sizeof(S.range), hipMemcpyHostToDevice));
  if (S.loss_cap < steps) {
    if (S.loss_hist) (void)hipFree(S.loss_hist);
    S.loss_hist = nullptr;
    UMWM_HIP_TRY(hipMalloc(&S.loss_hist, steps * sizeof(double)));
    S.loss_cap = steps;
  }
  hipStream_t st = S.stream;
  const int K0 = S.net.L[0].K, N3 = S.net.L.back().N, B = S.batch;
  hipLaunchKernelGGL(k_fill_f64, g1(steps), dim3(256), 0, st, S.loss_hist, (int64_t)steps, 0.0);
  hipLaunchKernelGGL(k_fill_f64, dim3(1), dim3(64), 0, st, S.net.loss_acc, (int64_t)1, 0.0);
  for (int k = 0; k < steps; ++k) {
    const uint64_t seed = S.seed * 0x9E3779B97F4A7C15ull + (uint64_t)S.net.step * 0xD1B54A32D192ED03ull;
    hipLaunchKernelGGL(k_sur_gather, dim3((std::max(K0, N3) + 255) / 256, B), dim3(256), 0, st,
                       (const float*)S.feat, (const float*)S.targ, B, S.nin, S.nout, K0, N3,
                       (const float*)S.fmu, (const float*)S.fis, (const float*)S.tmu,
                       (const float*)S.tsd, S.net.x, S.tbatch, (int64_t)0, 1, seed, ntrain);
    int rc = mlp_forward(S.net, B, st);
    if (!rc) rc = mlp_loss(S.net, S.tbatch, nullptr, B, st);
    if (!rc) {
      hipLaunchKernelGGL(k_store_loss, dim3(1), dim3(64), 0, st, S.net.loss_acc, S.loss_hist + k);
      rc = mlp_backward(S.net, B, st);
    }
    if (!rc) rc = mlp_adam(S.net, (float)lr, 0.9f, 0.999f, 1e-8f, 0.f, st);
    if (rc) return rc;
  }
  UMWM_HIP_TRY(hipStreamSynchronize(st));
  std::vector<double> h(steps);
  UMWM_HIP_TRY(hipMemcpy(h.data(), S.loss_hist, steps * sizeof(double), hipMemcpyDeviceToHost));
  const int w = std::max(1, std::min(steps / 4, 20));
  double a = 0, b = 0;
  for (int k = 0; k < w; ++k) {
    a += h[k];
    b += h[steps - 1 - k];
  }
  if (loss_first) *loss_first = a / w;
  if (loss_last) *loss_last = b / w;
  return 0;
}

int umwm_hip_surrogate_get_norm(umwm_surrogate* s, float* in_mean, float* in_inv_sd,
                                float* out_mean, float* out_sd) {
  if (int rc = sur_check(s)) return rc;
  SurrogateImpl& S = *s->impl;
  if (!S.stats_ready) return fail(UMWM_ERR_STATE, "surrogate_get_norm: train first");
  const std::pair<float*, const float*> cp[4] = {
      {in_mean, S.fmu}, {in_inv_sd, S.fis}, {out_mean, S.tmu}, {out_sd, S.tsd}};
  for (int i = 0; i < 4; ++i) {
    if (!cp[i].first) continue;
    const size_t n = i < 2 ? S.nin : S.nout;
    UMWM_HIP_TRY(hipMemcpy(cp[i].first, cp[i].second, n * sizeof(float), hipMemcpyDeviceToHost));
  }
  return 0;
}

int umwm_hip_surrogate_eval(umwm_surrogate* s, int64_t first, int64_t n, double* loss) {
  if (int rc = sur_check(s)) return rc;
  SurrogateImpl& S = *s->impl;
  if (!S.stats_ready) return fail(UMWM_ERR_STATE, "surrogate_eval: train first");
  if (first < 0 || n < 8 || first + n > S.count)
    return fail(UMWM_ERR_ARGUMENT, "surrogate_eval: rows outside the collected samples");
  hipStream_t st = S.stream;
  const int K0 = S.net.L[0].K, N3 = S.net.L.back().N;
  double total = 0;
  int64_t done = 0;
  while (done + 8 <= n) {
    const int rows = (int)std::min<int64_t>(S.batch, (n - done) / 8 * 8);
    hipLaunchKernelGGL(k_fill_f64, dim3(1), dim3(64), 0, st, S.net.loss_acc, (int64_t)1, 0.0);
    hipLaunchKernelGGL(k_sur_gather, dim3((std::max(K0, N3) + 255) / 256, rows), dim3(256), 0, st,
                       (const float*)S.feat, (const float*)S.targ, rows, S.nin, S.nout, K0, N3,
                       (const float*)S.fmu, (const float*)S.fis, (const float*)S.tmu,
                       (const float*)S.tsd, S.net.x, S.tbatch, first + done, 0, (uint64_t)0,
                       S.count);
    int rc = mlp_forward(S.net, rows, st);
    if (!rc) rc = mlp_loss(S.net, S.tbatch, nullptr, rows, st);
    if (rc) return rc;
    double l = 0;
    UMWM_HIP_TRY(hipStreamSynchronize(st));
    UMWM_HIP_TRY(hipMemcpy(&l, S.net.loss_acc, sizeof(double), hipMemcpyDeviceToHost));
    total += l * rows;
    done += rows;
  }
  if (loss) *loss = total / (double)done;
  return 0;
}

}  // extern "C"
