// Fused substep kernels for gfx950.
//
// One substep of the reference (src/umwm_top.F90:95-115) is four kernels:
//   k_rates      sin + sds + rate + per-point max|rate| -> physics_time_step,
//                then a grid-wide min (atomic max on order-preserving keys,
//                finished by the last block) -> dts, dta, sumt, dtr
//                (src/umwm_physics.F90:22-58, src/umwm_advection.F90:329-365).
//                Reads E^n once and writes the exclusive frequency prefix of
//                the directional cos^2 convolution (3 values per frequency
//                and point).
//   k_integrate  sin + sds + snl + exponential integration + diagnostic
//                tail (src/umwm_physics.F90:60-84); writes only
//                E* = (E^n + E_s)/2 (es).
//   k_advect     upwind propagation on E* (src/umwm_advection.F90:18-99),
//                refraction (:102-180) and e = ef, one wave per
//                (frequency, 64 sea points), with E_s recovered in place as
//                2 E* - E^n from the e it overwrites (one spectrum stream
//                fewer than storing E_s; the round-off is |E^n + E_s| ulp);
//                also the per-frequency partial sums of
//                stress_vector_integral (src/umwm_stress.F90:313-331).
//   k_stress     per point: frequency sums, tail and the FP64 form/skin
//                drag chain of stress('atm') (src/umwm_stress.F90:216-264).
//
// k_rates and k_integrate sweep frequencies in order inside the kernel: a
// workgroup owns 64 consecutive sea points (one per lane, so every
// spectrum/table access is a coalesced 64-wide stream of the
// i-contiguous layout) and W waves, wave w owning directions
// p = w*D .. w*D+D-1.  The prefix over frequency of
// sum_pp e(o,pp) cos^2(th_pp - th_p) (src/umwm_source_functions.F90:92-97)
// uses cos^2(a-b) = (1 + cos2a cos2b + sin2a sin2b)/2: three directional
// moments per frequency, exchanged between waves through LDS when W > 1,
// which turns the O(pm^2) convolution into O(pm).  Each sweep prefetches
// the next frequency's spectrum (and, in FP32, its tables) while computing
// the current one.  Launches whose waves all hold exactly DMAX directions
// take a variant without per-slot masks, in which k_rates also pairs each
// direction with its opposite.
//
// snl(o) = bf1r(o) X(o+1) + bf2r(o) X(o+2) - snl_fac X(o), X = sds e
// (src/umwm_source_functions.F90:165-171) is folded into the integration
// as each X becomes known: k_integrate keeps per direction only the
// partial E* = (E^n + partial E_s)/2 of the two previous frequencies.
//
// Wind input per bin (src/umwm_source_functions.F90:38-80) is evaluated
// as y = |x| x times one of three per-frequency factors (sheltering for
// y > 0, sin_diss1 against the wind, sin_diss1 fieldscale2 for swell
// overrunning it), with x = A cos th + B sin th - cp0 from the angle-sum
// identity (A, B per frequency) - the reference's expression regrouped,
// equal to it up to rounding.
//
// All per-point streams go through raw buffer loads and stores (bufio.hpp):
// one descriptor per frequency in scalar registers, 32-bit lane offsets
// computed once per thread, directions reached with a scalar offset.  The
// FP32 sweeps derive their (frequency, point) tables from k and the depth
// (TabGen); FP64 reads a 12-value record per lane (three 16-byte loads).
// The prefix moments and stress partial sums are 4-value records.
#include <hip/hip_runtime.h>

#include <type_traits>

#include "bufio.hpp"
#include "domain.hpp"
#include "fmath64.hpp"
#include "physics.hpp"
#include "surrogate.hpp"

namespace umwm {
namespace {

template <typename U>
__device__ __forceinline__ U pload(const U* a, size_t n, int j) {
  return bld<U>(make_rsrc(a, (uint32_t)(n * sizeof(U))), (uint32_t)j * (uint32_t)sizeof(U));
}
template <typename U>
__device__ __forceinline__ void pstore(U* a, size_t n, int j, U v) {
  bst(make_rsrc(a, (uint32_t)(n * sizeof(U))), (uint32_t)j * (uint32_t)sizeof(U), v);
}

// Per-point inputs of the source terms (src/umwm_source_functions.F90:21-52,
// :174-177).
template <typename T>
struct PointState {
  T wspd, a;   // floored wind speed, 2.5 ustar
  T pre;       // (1 - fice) 2 pi rhorat
  T shelt;     // sheltering coefficient
  T uc, vc;
  T cw, sw;    // cos, sin of wdir
  T csdt;      // sdt_fac sqrt(rhorat) ustar
  T ice;       // Kohout attenuation coefficient (0 below fice_lth)
  int oc;      // 0-based last prognostic bin
};

// Fresh evaluation (k_rates): wind floor, cut-off bin, sheltering; wave 0
// stores them for the later kernels of the substep.
template <typename T>
__device__ inline PointState<T> point_fresh(const DomainDev<T>& d, int j, bool write) {
  const size_t ldi = d.ldi, ldq = d.ldq;
  PointState<T> s;
  T w = pload(d.wspd, ldq, j);
  w = w > T(1e-2) ? w : T(1e-2);
  T fc = T(0.53) * T(d.g) / w;
  if (fc > T(d.fprog)) fc = T(d.fprog);
  int oc = 0;
  for (int o = d.om - 3; o >= 0; --o) {  // src/umwm_source_functions.F90:30-36
    oc = o;
    if (fc > cld(d.f, o)) break;
  }
  const T wdir = pload(d.wdir, ldq, j);
  const T rr = pload(d.rhorat, ldq, j);
  const T fice = pload(d.fice, ldq, j);
  const T ust = T(pload(d.ustar, ldi, j));
  s.wspd = w;
  s.a = T(2.5) * ust;
  s.pre = (T(1) - fice) * T(kTwoPiD) * rr;
  s.shelt = sheltering_coare35(w);
  s.uc = pload(d.uc, ldq, j);
  s.vc = pload(d.vc, ldq, j);
  s.cw = cos(wdir);
  s.sw = sin(wdir);
  s.csdt = T(d.sdt_fac) * sqrt(rr) * ust;
  s.ice = fice > T(d.fice_lth) ? T(sice_coef(pload(d.hs_ice, ldi, j))) : T(0);
  s.oc = oc;
  if (write) {
    pstore(d.wspd, ldq, j, w);
    pstore(d.fcutoff, ldi, j, fc);
    pstore(d.oc, ldi, j, (int32_t)oc);
    pstore(d.shelt, ldi, j, s.shelt);
    pstore(d.cwdir, ldi, j, s.cw);
    pstore(d.swdir, ldi, j, s.sw);
  }
  return s;
}

// Later kernels of the substep read back what k_rates stored.
template <typename T>
__device__ inline PointState<T> point_stored(const DomainDev<T>& d, int j) {
  const size_t ldi = d.ldi, ldq = d.ldq;
  PointState<T> s;
  const T rr = pload(d.rhorat, ldq, j);
  const T fice = pload(d.fice, ldq, j);
  const T ust = T(pload(d.ustar, ldi, j));
  s.wspd = pload(d.wspd, ldq, j);
  s.a = T(2.5) * ust;
  s.pre = (T(1) - fice) * T(kTwoPiD) * rr;
  s.shelt = pload(d.shelt, ldi, j);
  s.uc = pload(d.uc, ldq, j);
  s.vc = pload(d.vc, ldq, j);
  s.cw = pload(d.cwdir, ldi, j);
  s.sw = pload(d.swdir, ldi, j);
  s.csdt = T(d.sdt_fac) * sqrt(rr) * ust;
  s.ice = fice > T(d.fice_lth) ? T(sice_coef(pload(d.hs_ice, ldi, j))) : T(0);
  s.oc = pload(d.oc, ldi, j);
  return s;
}

// Per-frequency factors of the wind input: ssin = y * (y > 0 ? pos :
// (overrun ? neg2 : neg1)), y = |x| x, x = A cth + B sth - cp0.
template <typename T>
struct SinFreq {
  T A, B, cp0, pos, neg1, neg2;
};

template <typename T>
__device__ __forceinline__ SinFreq<T> sin_freq(const PointState<T>& s, T logl2, T cp0, T fkovg,
                                               T sin_diss1, T fs2) {
  SinFreq<T> r;
  const T ba = s.wspd + s.a * logl2;
  r.A = ba * s.cw - s.uc;
  r.B = ba * s.sw - s.vc;
  r.cp0 = cp0;
  const T g = s.pre * fkovg;
  r.pos = g * s.shelt;
  r.neg1 = g * sin_diss1;
  r.neg2 = r.neg1 * fs2;
  return r;
}

template <typename T>
__device__ __forceinline__ T sin_bin(const SinFreq<T>& q, const PointState<T>& s, T ct, T st,
                                     bool tail) {
  const T x = q.A * ct + q.B * st - q.cp0;
  const T y = fabs(x) * x;
  const T cwt = s.cw * ct + s.sw * st;  // cos(wdir - th)
  // select between values (a select between members would become an
  // address select and spill the struct to scratch)
  const T pos = q.pos, neg1 = q.neg1, neg2 = q.neg2;
  const T neg = cwt > T(0) ? neg2 : neg1;
  T r = y * (y > T(0) ? pos : neg);
  if (tail) r = r < T(0) ? T(0) : r;  // src/umwm_source_functions.F90:73-80
  return r;
}

// Rotation tendency of src/umwm_advection.F90:335-338 for one direction:
// 0.5 ((dcx sth + dvc) odx - (dcy cth + duc) ody).
template <typename T>
__device__ __forceinline__ T rot_dir(T dcx, T dcy, T sth, T cth, T dvc, T duc, T odx, T ody) {
  return T(0.5) * ((dcx * sth + dvc) * odx - (dcy * cth + duc) * ody);
}

// Per-frequency, per-point tables of the source terms: one record of
// DomainDev::ftab.
template <typename T>
struct FreqTab {
  T v[FT_N];
  __device__ __forceinline__ T operator[](int i) const { return v[i]; }
};

// Descriptor of frequency o's slice of a [om][rows][ldi] spectrum (rows =
// pm; a direction is then the scalar offset p * ldi * sizeof(T)) or of a
// [om][ldi][rec] record block (rec values per point).
template <typename T>
__device__ __forceinline__ rsrc_t freq_rsrc(const T* base, int o, int rows, size_t ldi) {
  return row_rsrc(base, (size_t)o, (size_t)rows * ldi);
}

// Frequency o's slice of the surrogate's bf16 rate rows, [om][pm][ldi].
__device__ __forceinline__ rsrc_t sur_rsrc(const uint16_t* base, int o, int pm, size_t ldi) {
  return row_rsrc(base, (size_t)o, (size_t)pm * ldi);
}

// First N slots of the record of (o, j); vr = j * FT_N * sizeof(T).
template <int N, typename T>
__device__ __forceinline__ void load_tab(const T* ftab, size_t ldi, int o, uint32_t vr,
                                         FreqTab<T>& r) {
  T tmp[N];
  bldn<T, N>(freq_rsrc(ftab, o, FT_N, ldi), vr, tmp);
#pragma unroll
  for (int i = 0; i < N; ++i) r.v[i] = tmp[i];
#pragma unroll
  for (int i = N; i < FT_N; ++i) r.v[i] = T(0);
}

// FP32 sweeps derive the (frequency, point) tables from k and the depth
// (src/umwm_init.F90:1480-1533) rather than load them: the sweeps are
// bound by memory traffic, and one k per (frequency, point) replaces a
// 9-value record.  With u = e^{-2 kd}, kd = min(k d, 20):
//   sinh(2 kd) = (1/u - u)/2,  cosh(kd) = (1/sqrt(u) + sqrt(u))/2,
//   coth(0.2 kd) = (1 + w)/(1 - w) with w = e^{-0.4 kd},
//   cp0 = 2 pi f / k, cg0 = cp0 (1/2 + k d / sinh(2 kd) + s k^2/(rho g + s k^2)),
//   kdk = k dom / |cg0|, bf1r(o) = snl_fac bf1 kdk(o+1)/kdk(o) (o <= om-3),
//   log(l2/z) with l2 = cp0/(2 f) capped at 20 m.
// It keeps k of the next three frequencies and kdk of the next two, so
// bf1r and bf2r need no look-back.  FP64 keeps the record tables: its
// library exp would cost more than the traffic saved.
struct TabGen {
  const float* kb;
  const float *fr, *domr;
  size_t ldi;
  uint32_t vj;
  int om;
  float dep, sgr, sbf, sbp, s1, s2, l2z, lz;
  float k0, k1, k2, kn;   // k(o), k(o+1), k(o+2), k(o+3) (loaded ahead)
  float kdk0, kdk1, kdk2;

  __device__ float load_k(int o) const {
    return bld<float>(row_rsrc(kb, (size_t)min(o, om - 1), ldi), vj);
  }
  __device__ float kdk_of(float k, int o) const {
    o = min(o, om - 1);
    const float kd = fminf(k * dep, 20.f);
    const float u = __builtin_amdgcn_exp2f(-2.f * 1.4426950408889634f * kd);
    const float sinh2 = 0.5f * (__builtin_amdgcn_rcpf(u) - u);
    const float cp = 6.283185307179586f * cld(fr, o) * __builtin_amdgcn_rcpf(k);
    const float sk2 = sgr * k * k;
    const float cg = cp * (0.5f + k * dep * __builtin_amdgcn_rcpf(sinh2) +
                           sk2 * __builtin_amdgcn_rcpf(1.f + sk2));
    return k * cld(domr, o) * __builtin_amdgcn_rcpf(fabsf(cg));
  }
  template <typename D>
  __device__ void init(const D& d, int j, uint32_t vj_) {
    kb = d.k;
    fr = d.f;
    domr = d.dom;
    ldi = d.ldi;
    vj = vj_;
    om = d.om;
    // lanes past im read a depth of 0; they take 1 m so that their tables
    // stay finite (the reference requires only dmin > 0 of sea points)
    dep = j < d.im ? pload(d.depth, ldi, j) : 1.f;
    sgr = float(d.sfct / (d.rhow0 * d.g));
    sbf = float(d.sbf_fac);
    sbp = float(d.sbp_fac);
    s1 = float(d.snl_fac * d.bf1);
    s2 = float(d.snl_fac * d.bf2);
    lz = float(log(d.z));
    l2z = float(log(20.0 / d.z));
    k0 = load_k(0);
    k1 = load_k(1);
    k2 = load_k(2);
    kn = load_k(3);
    kdk0 = kdk_of(k0, 0);
    kdk1 = kdk_of(k1, 1);
    kdk2 = kdk_of(k2, 2);
  }
  // slots of frequency o (the FreqTab layout of the record tables)
  __device__ void table(int o, FreqTab<float>& r) const {
    const float k = k0, fo = cld(fr, o);
    const float kd = fminf(k * dep, 20.f);
    const float u = __builtin_amdgcn_exp2f(-2.f * 1.4426950408889634f * kd);
    const float iu = __builtin_amdgcn_rcpf(u);
    const float isinh2 = __builtin_amdgcn_rcpf(0.5f * (iu - u));
    const float v = __builtin_amdgcn_sqrtf(u);
    const float ch = 0.5f * (__builtin_amdgcn_rcpf(v) + v);
    const float w = __builtin_amdgcn_exp2f(-0.4f * 1.4426950408889634f * kd);
    const float cp = 6.283185307179586f * fo * __builtin_amdgcn_rcpf(k);
    const float sk2 = sgr * k * k;
    const float l2 = 0.5f * cp * __builtin_amdgcn_rcpf(fo);
    r.v[FT_CP0] = cp;
    r.v[FT_LOGL2] = l2 > 20.f ? l2z : 0.6931471805599453f * __builtin_amdgcn_logf(l2) - lz;
    r.v[FT_COTH] = (1.f + w) * __builtin_amdgcn_rcpf(1.f - w);
    r.v[FT_SBF] = sbf * k * isinh2 + sbp * k * __builtin_amdgcn_rcpf(ch * ch);
    r.v[FT_K] = k;
    r.v[FT_CG0] = cp * (0.5f + k * dep * isinh2 + sk2 * __builtin_amdgcn_rcpf(1.f + sk2));
    const bool bo = o <= om - 3;
    const float ik = __builtin_amdgcn_rcpf(kdk0);
    r.v[FT_B1] = bo ? s1 * kdk1 * ik : 0.f;
    r.v[FT_B2] = bo ? s2 * kdk2 * ik : 0.f;
    r.v[FT_KDK] = kdk0;
    for (int q = FT_PAD; q < FT_N; ++q) r.v[q] = 0.f;
  }
  // moves to frequency o+1; `next` is k(o+4), loaded during iteration o
  __device__ void advance(int o, float next) {
    k0 = k1;
    k1 = k2;
    k2 = kn;
    kn = next;
    kdk0 = kdk1;
    kdk1 = kdk2;
    kdk2 = kdk_of(k2, o + 3);
  }
};

// Transcendentals of the sweeps.  FP32 works in base 2 on the hardware
// log2/exp2 (exponents pre-scaled by log2 e where a natural exponential is
// meant).  FP64 uses the table-driven natural log and exp of fmath64.hpp
// (within 1 ulp, about 0.65x the instructions of the device library's) and
// a reciprocal from the hardware estimate refined by two Newton steps (its
// arguments are positive and normal: the spreading factor >= 1 and
// cf coth).  The tables (4.2 KB) are staged into LDS by every block of the
// FP64 sweeps: read from global memory they were 3.5 KB of divergent
// vector loads per wave and bin, more than the spectrum stream itself.
constexpr int kFm64Len = fm64::kExpTabLen + fm64::kLogTabLen + fm64::kInvTabLen;
__device__ const double kFm64Tab[kFm64Len] = {UMWM_FM64_EXP_TABLE, UMWM_FM64_LOG_TABLE,
                                              UMWM_FM64_INV_TABLE};
__shared__ __attribute__((aligned(16))) double sFm64[kFm64Len];

// Fills the LDS tables; every thread of the block must call it.
template <typename T>
__device__ __forceinline__ void stage_fm64() {
  if constexpr (std::is_same_v<T, double>) {
    for (int k = threadIdx.x; k < kFm64Len; k += blockDim.x) sFm64[k] = kFm64Tab[k];
    __syncthreads();
  }
}

template <typename T>
struct Tr;
template <>
struct Tr<float> {
  static constexpr float kExpScale = 1.4426950408889634f;  // log2(e)
  __device__ static __forceinline__ float lg(float x) { return __builtin_amdgcn_logf(x); }
  __device__ static __forceinline__ float ex(float x) { return __builtin_amdgcn_exp2f(x); }
  __device__ static __forceinline__ float lg_nn(float x) { return __builtin_amdgcn_logf(x); }
  __device__ static __forceinline__ float ex_nn(float x) { return __builtin_amdgcn_exp2f(x); }
  __device__ static __forceinline__ float rcp(float x) { return __builtin_amdgcn_rcpf(x); }
};
template <>
struct Tr<double> {
  static constexpr double kExpScale = 1.0;
  __device__ static __forceinline__ double lg(double x) {
    return fm64::log_t(x, sFm64 + fm64::kExpTabLen, sFm64 + fm64::kExpTabLen + fm64::kLogTabLen);
  }
  __device__ static __forceinline__ double ex(double x) { return fm64::exp_t(x, sFm64); }
  // k_rates' forms: its logarithms are of |e|, k and cf (finite, >= 0) and a
  // NaN spectral value reaches its rate through e itself, so the special
  // cases reduce to log 0 = -inf (0.9 ms less per hurricane member-substep;
  // in k_integrate the same change added spills and cost 2.7 ms)
  __device__ static __forceinline__ double lg_nn(double x) {
    return fm64::log_t<false>(x, sFm64 + fm64::kExpTabLen,
                              sFm64 + fm64::kExpTabLen + fm64::kLogTabLen);
  }
  __device__ static __forceinline__ double ex_nn(double x) { return fm64::exp_t<false>(x, sFm64); }
  __device__ static __forceinline__ double rcp(double x) {
    double r = __builtin_amdgcn_rcp(x);
    r = fma(r, fma(-x, r, 1.0), r);
    return fma(r, fma(-x, r, 1.0), r);
  }
};

// Per-frequency wind-input factors in the form the sweeps use: ssin =
// y * (y > 0 ? pos : neg1 * m), y = |x| x, x = A cth + B sth - cp0, with
// m = fieldscale2 where the wind overruns the direction (cos(wdir - th) >
// 0), else 1 (a per-lane, per-direction constant).
template <typename T>
__device__ __forceinline__ T sin_y(const SinFreq<T>& q, T ct, T st) {
  const T x = fma(q.A, ct, fma(q.B, st, -q.cp0));
  return fabs(x) * x;
}
template <typename T>
__device__ __forceinline__ T sin_m(const SinFreq<T>& q, T y, T m) {
  const T pos = q.pos, neg = q.neg1 * m;
  return y * (y > T(0) ? pos : neg);
}

// ---------------------------------------------------------------------
// k_rates
// ---------------------------------------------------------------------
template <typename T, int DMAX, bool FULL, bool SUR>
__global__ __launch_bounds__(512) void k_rates(const DomainDev<T>* __restrict__ doms,
                                               const BlockTask* __restrict__ tasks, int first,
                                               int rot_all) {
  extern __shared__ double smem[];
  stage_fm64<T>();
  const BlockTask tk = tasks[blockIdx.x];
  const DomainDev<T>& d = doms[tk.dom];
  // rotation tendency on the first substep of the launch (rot_all) and on
  // every substep of a domain with currents (src/umwm_advection.F90:117,
  // currents .or. first): the choice is per domain, so that batch members
  // with and without currents each follow the reference
  const bool rotation = rot_all || d.currents;
  Scalars* sc = d.sc;
  const int W = blockDim.x >> 6, lane = threadIdx.x & 63, w = wave_id();
  const int j = tk.i0 + lane;
  const bool active = sc->sumt < d.dtg;
  const bool valid = active && j < d.im;
  // substeps taken so far: the rotating audit group (read before the last
  // block counts this substep)
  [[maybe_unused]] const int nsub0 = sc->nsub;
  const int pm = d.pm, om = d.om;
  const size_t ldi = d.ldi, ldq = d.ldq;
  const int D = FULL ? DMAX : (pm + W - 1) / W, p0 = w * D;
  // directions of this wave (uniform); FULL launches have W * DMAX = pm
  const int nd = FULL ? DMAX : max(0, min(D, pm - p0));
  const uint32_t vj = (uint32_t)j * (uint32_t)sizeof(T);
  const uint32_t vr = (uint32_t)j * (uint32_t)(FT_N * sizeof(T));
  const uint32_t rb = (uint32_t)ldi * (uint32_t)sizeof(T);  // bytes of one direction row
  // the surrogate's bf16 rate rows
  [[maybe_unused]] const uint32_t vj2 = (uint32_t)j * 2u, rb2 = (uint32_t)ldi * 2u;
  // Arrays the frequency loop uses, held in registers: buffer stores are
  // opaque to the compiler, which would otherwise reload these pointers
  // from the domain record after every store.
  const T* const eb = d.e;
  T* const pfxb = d.pfx;
  const T* const ftab = d.ftab;
  const T* const fr = d.f;
  const T* const dir4 = d.dir4;
  double pts = INFINITY, rotmax = -INFINITY;

  if (active) {
    const PointState<T> s = point_fresh(d, j, valid && w == 0);
    const int oc = s.oc;
    if (rotation && valid && w == 0) pstore(d.oc_rot, ldi, j, (int32_t)oc);
    uint32_t vie = 0, viw = 0, vin = 0, vis = 0;
    T dvc = 0, duc = 0, odx = 0, ody = 0;
    if (rotation) {
      vie = (uint32_t)pload(d.ie, ldi, j) * (uint32_t)sizeof(T);
      viw = (uint32_t)pload(d.iw, ldi, j) * (uint32_t)sizeof(T);
      vin = (uint32_t)pload(d.in_, ldi, j) * (uint32_t)sizeof(T);
      vis = (uint32_t)pload(d.is_, ldi, j) * (uint32_t)sizeof(T);
      dvc = pload(d.vc, ldq, pload(d.iie, ldi, j)) - pload(d.vc, ldq, pload(d.iiw, ldi, j));
      duc = pload(d.uc, ldq, pload(d.iin, ldi, j)) - pload(d.uc, ldq, pload(d.iis, ldi, j));
      odx = pload(d.oneovdx, ldi, j);
      ody = pload(d.oneovdy, ldi, j);
    }
    const T sin_diss1 = T(d.sin_fac * d.fieldscale1), fs2 = T(d.fieldscale2);
    const T tpsf = T(d.twopisds_fac), spow = T(d.sds_power), mssf = T(d.mss_fac);
    const T nu4 = T(4 * d.nu_water), ig = T(1 / d.g);
    const double hdth = 0.5 * d.dth;
    // Slots q >= nd are computed but never used: their spectrum loads go
    // through a zero-range descriptor and read 0, their direction records
    // are dir4's zero padding, and they are masked out of the rate
    // maximum.  With no per-slot branch the compiler can interleave the
    // directions.
    const rsrc_t rz = make_rsrc(eb, 0u);
    // FULL launches with an even DMAX pair each direction p with p + pm/2
    // (theta + pi): cos and sin change sign while cos 2theta and sin 2theta
    // do not, so a pair shares its direction record, the mss factor and one
    // update of the moments.  Wave w then holds pairs w DMAX/2 + h in slots
    // h and h + DMAX/2, which also halves the direction records the loop
    // keeps in scalar registers.
    constexpr bool PAIR = FULL && DMAX % 2 == 0;
    constexpr int H = DMAX / 2;
    const int ph = w * H, half = pm / 2;
    auto dirp = [&](int q) { return PAIR ? (q < H ? ph + q : ph + q - H + half) : p0 + q; };
    // overrun factor of slot q: fieldscale2 where the wind overruns the
    // direction, else 1; the paired loop reads it as bit q of mb (one
    // register instead of DMAX)
    using MB = std::conditional_t<(DMAX > 32), uint64_t, uint32_t>;
    MB mb = 0;
    T mq[DMAX];
#pragma unroll
    for (int q = 0; q < DMAX; ++q) {
      T ct, st, c2, s2;
      cld4(dir4, dirp(q), ct, st, c2, s2);
      const bool over = s.cw * ct + s.sw * st > T(0);
      mq[q] = over ? fs2 : T(1);
      if (over) mb |= MB(1) << q;
    }
    // Surrogate (FP32): lanes whose inputs passed the range check take the
    // network's spreading factor dm^2 in place of the one the moment
    // prefix gives, so that sds = dm^2 s with s = cf coth (e k^4)^p the
    // closed form of the bin.  The residual check holds the network's dm^2
    // against the bounds the prefix sets, [1, (1 + 2 mss_fac m0)^2] (m0 the
    // frequency prefix of the zeroth moment, which bounds the cos^2-weighted
    // ones), and the value used is clamped to them; a lane whose
    // energy-weighted share of bins outside the bounds (widened by
    // kSurBoundTol) exceeds the residual tolerance, or whose audited
    // dissipation error exceeds the audit tolerance, falls back to the plain
    // rates, which these waves always evaluate.  All waves of a workgroup
    // hold the same 64 points, so these choices agree across waves.
    // The domain-wide gate: when the previous substep's audited workgroups
    // together found the network's dissipation error above the audit
    // tolerance, every point takes the plain rates this substep; the
    // audited workgroups still evaluate the network, so that the audit can
    // open the gate again.
    [[maybe_unused]] int sok = 0;
    [[maybe_unused]] bool aud = false, gate = false;
    bool do_sur = false;
    [[maybe_unused]] const uint16_t* const rsur = d.sur_r;
    if constexpr (SUR) {
      if (rsur) {
        sok = valid ? pload(d.sur_ok, ldi, j) : 0;
        aud = d.sur_audit_stride > 0 && ((tk.local + nsub0) % d.sur_audit_stride) == 0;
        gate = sc->sur_closed == 0;
        do_sur = __any(sok != 0) && (gate || aud);
      }
    }
    T rmax = 0;
    [[maybe_unused]] T rmax_s = 0, neg = 0, tot = 0, anum = 0, aden = 0;
    // One sweep over frequencies: `ds` also evaluates the network's
    // dissipation; `pass0` stores the prefix moments and the rotation
    // maximum.
    auto sweep = [&](bool ds, bool pass0) {
    double A0 = 0, AC = 0, AS = 0;
    T ec[DMAX];
    // the network's dm^2 - 1 of this and the next frequency as raw bf16
    // bits; the direction's row offset joins the lane offset (a vector add)
    // rather than holding DMAX more scalar registers
    [[maybe_unused]] uint32_t rc[DMAX];
    FreqTab<T> tc;
    {
      const rsrc_t ep = freq_rsrc(eb, 0, pm, ldi);
#pragma unroll
      for (int q = 0; q < DMAX; ++q) ec[q] = bld<T>(FULL || q < nd ? ep : rz, vj, dirp(q) * rb);
      if constexpr (SUR) {
        if (ds) {
          const rsrc_t rp = sur_rsrc(rsur, 0, pm, ldi);
#pragma unroll
          for (int q = 0; q < DMAX; ++q) rc[q] = bld_u16(FULL || q < nd ? rp : rz, vj2 + dirp(q) * rb2);
        }
      }
    }
    // FP32 derives the tables from k (TabGen); FP64 loads the record
    constexpr bool kDerive = sizeof(T) == 4;
    TabGen tg;
    if constexpr (kDerive) tg.init(d, j, vj);
    for (int o = 0; o < om; ++o) {
      // issue the next frequency's loads before computing this one
      const int on = o + 1 < om ? o + 1 : o;
      T en[DMAX];
      const rsrc_t epn = freq_rsrc(eb, on, pm, ldi);
#pragma unroll
      for (int q = 0; q < DMAX; ++q) en[q] = bld<T>(FULL || q < nd ? epn : rz, vj, dirp(q) * rb);
      [[maybe_unused]] uint32_t rn[DMAX];
      if constexpr (SUR) {
        if (ds) {
          const rsrc_t rp = sur_rsrc(rsur, on, pm, ldi);
#pragma unroll
          for (int q = 0; q < DMAX; ++q) rn[q] = bld_u16(FULL || q < nd ? rp : rz, vj2 + dirp(q) * rb2);
        }
      }
      float knext = 0.f;
      if constexpr (kDerive) {
        knext = tg.load_k(o + 4);
        tg.table(o, tc);
      } else {
        load_tab<kRatesSlots>(ftab, ldi, o, vr, tc);
      }

      const T k = tc[FT_K], k2 = k * k, fo = cld(fr, o);
      const T sdt = s.csdt * k;
      const T sice = T(2) * tc[FT_CG0] * s.ice;
      const T crest = tc[FT_SBF] + sdt + nu4 * k2 - sice;  // sbf + sdt + sdv - sice
      const SinFreq<T> sf = sin_freq(s, tc[FT_LOGL2], tc[FT_CP0], fo * k * ig, sin_diss1, fs2);
      const T pa0 = T(hdth * A0), pac = T(hdth * AC), pas = T(hdth * AS);
      if (pass0 && valid && w == 0) {
        if constexpr (sizeof(T) == 4) {
          bst16(freq_rsrc(pfxb, o, 4, ldi), (uint32_t)j * 16u,
                typename Vec16<T>::type{pa0, pac, pas, T(0)});
        } else {
          const rsrc_t pf = freq_rsrc(pfxb, o, 4, ldi);
          bst16(pf, (uint32_t)j * 32u, typename Vec16<T>::type{pa0, pac});
          bst16(pf, (uint32_t)j * 32u + 16u, typename Vec16<T>::type{pas, T(0)});
        }
      }
      // sds (after cothkd) = cf coth dm^2 (e k4)^p = dm^2 ex(p lg e + Lk)
      const T cfo = tpsf * fo * tc[FT_COTH];
      const T Lk = T(4) * spow * Tr<T>::lg_nn(k) + Tr<T>::lg_nn(cfo);  // lg(k^4) = 4 lg k
      const T cs = T(1) - (tc[FT_B1] + tc[FT_B2]);                // snl_arg
      const T q0 = T(1) + mssf * pa0, qc = mssf * pac, qs = mssf * pas;
      [[maybe_unused]] const T sbfv = tc[FT_SBF], kdkv = tc[FT_KDK];
      T m0 = 0, mc = 0, ms = 0, rq = 0;
      [[maybe_unused]] T rqs = 0;
      // sds = dm^2 sat with sat = ex(p lg |e| + Lk) = cf coth (e k4)^p: of
      // |e|, so that a bin emptied by round-off to about -1e-30 gives a
      // vanishing sds, as the 0 the reference's power of a non-positive
      // base is taken to be here, without a clamp instruction
      // the network's dm^2 of one bin: its rate, the residual of the
      // spreading-factor bounds and the audit sums (sat = cf coth (e k^4)^p,
      // sdp = the plain sds = dm^2 sat)
      [[maybe_unused]] const T qhi = (T(1) + T(2) * mssf * pa0) * (T(1) + T(2) * mssf * pa0);
      auto sur_bin = [&](T ss, T qn, T e, T sat, T sdp, bool use_slot) {
        const T dmn = T(1) + qn;  // the network stores dm^2 - 1
        const T sdn = fmin(fmax(dmn, T(1)), qhi) * sat;
        const T r2 = ss - fma(sdn, cs, crest);
        rqs = fmax(rqs, use_slot ? fabs(r2) : T(0));
        const T wv = e * kdkv;
        const bool inb = dmn >= T(1 - kSurBoundTol) && dmn <= qhi * T(1 + kSurBoundTol);
        neg += inb ? T(0) : wv;  // NaN or Inf fail the bounds too
        tot += wv;
        if (aud) {
          anum = fma(fabs(sdn - sdp), wv, anum);
          aden = fma(sdp, wv, aden);
        }
      };
      // an opaque copy per frequency (an empty asm statement): without it the
      // compiler hoists the DMAX bit tests out of the frequency loop into
      // DMAX registers, which costs a wave per SIMD at DMAX = 36
      MB mbl = mb;
      asm volatile("" : "+v"(mbl));
      if constexpr (PAIR) {
#pragma unroll
        for (int h = 0; h < H; ++h) {
          T ct, st, c2, s2;
          cld4(dir4, ph + h, ct, st, c2, s2);
          const T ea = ec[h], eb2 = ec[h + H];
          const T u = fma(sf.A, ct, sf.B * st);  // x of theta + pi is -u - cp0
          const T xa = u - sf.cp0, xb = -u - sf.cp0;
          const T ya = fabs(xa) * xa, yb = fabs(xb) * xb;
          const T na = ((mbl >> h) & 1) ? sf.neg2 : sf.neg1;
          const T nb = ((mbl >> (h + H)) & 1) ? sf.neg2 : sf.neg1;
          const T ssa = ya * (ya > T(0) ? sf.pos : na), ssb = yb * (yb > T(0) ? sf.pos : nb);
          T dm = fma(qc, c2, fma(qs, s2, q0));
          dm = dm * dm;
          const T sata = Tr<T>::ex_nn(fma(spow, Tr<T>::lg_nn(fabs(ea)), Lk));
          const T satb = Tr<T>::ex_nn(fma(spow, Tr<T>::lg_nn(fabs(eb2)), Lk));
          const T spa = dm * sata, spb = dm * satb;
          const T ra = ssa - fma(spa, cs, crest);
          const T rbb = ssb - fma(spb, cs, crest);
          rq = fmax(rq, fmax(fabs(ra), fabs(rbb)));
          if constexpr (SUR) {
            if (ds) {
              sur_bin(ssa, bf16_bits_to_f(rc[h]), ea, sata, spa, true);
              sur_bin(ssb, bf16_bits_to_f(rc[h + H]), eb2, satb, spb, true);
            }
          }
          const T ep2 = ea + eb2;
          m0 += ep2;
          mc = fma(ep2, c2, mc);
          ms = fma(ep2, s2, ms);
        }
      } else {
#pragma unroll
        for (int q = 0; q < DMAX; ++q) {
          T ct, st, c2, s2;
          cld4(dir4, p0 + q, ct, st, c2, s2);
          const T e = ec[q];
          const T ss = sin_m(sf, sin_y(sf, ct, st), mq[q]);
          T dm = fma(qc, c2, fma(qs, s2, q0));
          dm = dm * dm;
          const T sat = Tr<T>::ex_nn(fma(spow, Tr<T>::lg_nn(fabs(e)), Lk));
          const T sp = dm * sat;
          const T r = ss - fma(sp, cs, crest);
          rq = fmax(rq, FULL || q < nd ? fabs(r) : T(0));
          if constexpr (SUR) {
            if (ds) sur_bin(ss, bf16_bits_to_f(rc[q]), e, sat, sp, FULL || q < nd);
          }
          m0 += e;
          mc = fma(e, c2, mc);
          ms = fma(e, s2, ms);
        }
      }
      if (o <= oc) {  // rates of prognostic bins only
        rmax = fmax(rmax, rq);
        if constexpr (SUR) rmax_s = fmax(rmax_s, rqs);
      }
      if (rotation && pass0) {
        // rotation tendency of the first substep (src/umwm_advection.F90:329-349)
        const rsrc_t cpr = row_rsrc(d.cp0, o, ldi);
        const T dcx = bld<T>(cpr, vie) - bld<T>(cpr, viw);
        const T dcy = bld<T>(cpr, vin) - bld<T>(cpr, vis);
        double rm = 0;  // bins above oc hold 0 and join the signed max
        if (o <= oc) {
          rm = -INFINITY;
          for (int q = 0; q < nd; ++q) {
            const int p = dirp(q);
            rm = fmax(rm, (double)rot_dir(dcx, dcy, cld(d.sth, p), cld(d.cth, p), dvc, duc, odx,
                                          ody));
          }
        }
        if (nd > 0) rotmax = fmax(rotmax, rm);
      }
      double s0 = m0, s2c = mc, s2s = ms;
      if (W > 1) {
        double* b = smem + (size_t)(o & 1) * 3 * W * 64;
        b[(0 * W + w) * 64 + lane] = s0;
        b[(1 * W + w) * 64 + lane] = s2c;
        b[(2 * W + w) * 64 + lane] = s2s;
        __syncthreads();
        s0 = s2c = s2s = 0;
        for (int q = 0; q < W; ++q) {
          s0 += b[(0 * W + q) * 64 + lane];
          s2c += b[(1 * W + q) * 64 + lane];
          s2s += b[(2 * W + q) * 64 + lane];
        }
      }
      const double k3 = k2 * tc[FT_KDK];  // k^3 dk
      A0 += k3 * s0;
      AC += k3 * s2c;
      AS += k3 * s2s;
#pragma unroll
      for (int q = 0; q < DMAX; ++q) ec[q] = en[q];
      if constexpr (SUR) {
        if (ds) {
#pragma unroll
          for (int q = 0; q < DMAX; ++q) rc[q] = rn[q];
        }
      }
      if constexpr (kDerive) tg.advance(o, knext);
    }
    };
    sweep(do_sur, true);
    // the surrogate's verdict per lane (the same in every wave)
    [[maybe_unused]] bool use = false;
    if constexpr (SUR) {
      if (rsur) {
        double sneg = neg, stot = tot, snum = anum, sden = aden;
        if (W > 1) {
          __syncthreads();
          smem[(0 * W + w) * 64 + lane] = sneg;
          smem[(1 * W + w) * 64 + lane] = stot;
          smem[(2 * W + w) * 64 + lane] = snum;
          smem[(3 * W + w) * 64 + lane] = sden;
          __syncthreads();
          sneg = stot = snum = sden = 0;
          for (int q = 0; q < W; ++q) {
            sneg += smem[(0 * W + q) * 64 + lane];
            stot += smem[(1 * W + q) * 64 + lane];
            snum += smem[(2 * W + q) * 64 + lane];
            sden += smem[(3 * W + q) * 64 + lane];
          }
        }
        // written as negated <= so that NaN sums count as failures
        const bool rfail = !(sneg <= (double)d.sur_resid_tol * stot);
        const bool afail = aud && !(snum <= (double)d.sur_audit_tol * sden);
        use = valid && sok && !rfail && !afail && gate;
        if (w == 0) {
          if (valid) pstore(d.sur_use, ldi, j, (int32_t)(use ? 1 : 0));
          // counts as 64-bit integer atomics (ballot counts), the audit sums
          // as FP64 atomics from audited workgroups only: FP64 atomic adds
          // from every workgroup to the same eight words serialised and cost
          // several times the sweep itself
          const bool cnt[SS_AUDIT_NUM] = {valid,
                                          use,
                                          valid && !sok,
                                          valid && sok && rfail,
                                          valid && sok && !rfail && (afail || !gate),
                                          valid && aud};
          auto* ctr = reinterpret_cast<unsigned long long*>(d.sur_stats);
#pragma unroll
          for (int q = 0; q < SS_AUDIT_NUM; ++q) {
            const unsigned long long c = __popcll(__ballot(cnt[q]));
            if (lane == 0 && c) atomicAdd(ctr + q, c);
          }
          if (aud) {
            const double a = wave_sum(valid ? snum : 0.0), b = wave_sum(valid ? sden : 0.0);
            if (lane == 0) {
              atomicAdd(d.sur_stats + SS_AUDIT_NUM, a);
              atomicAdd(d.sur_stats + SS_AUDIT_DEN, b);
              // this substep's sums for the gate (thread 0 of the block,
              // which waits for its atomics before the arrival count)
              atomicAdd(&sc->sur_anum, a);
              atomicAdd(&sc->sur_aden, b);
            }
          }
        }
      }
    }
    // max |rate| of the point over all waves
    __syncthreads();
    T rsel = rmax;
    if constexpr (SUR) rsel = use ? rmax_s : rmax;
    smem[w * 64 + lane] = (double)rsel;
    __syncthreads();
    if (w == 0) {
      double mx = 0;
      for (int q = 0; q < W; ++q) mx = fmax(mx, smem[q * 64 + lane]);
      pts = d.explim / mx;
      if (valid) pstore(d.pts, ldi, j, pts);
      if (!valid) pts = INFINITY;
    }
    if (!valid) rotmax = -INFINITY;
  }
  // Block reductions, then the grid-wide ones through two agent-scope
  // atomics per block on order-preserving integer keys (okey) kept in the
  // domain's Scalars: an atomic max of ~okey(min pts) and of okey(max
  // rotation rate).  The last block to arrive reads and clears them.  There
  // are no per-block partials in memory, hence no release fence: on this
  // multi-XCD GPU an agent-scope release writes back and invalidates the
  // L2 and cost about a fifth of the kernel.  Device atomics are performed
  // at the device's coherence point; the wait for their completion before
  // the arrival count orders them ahead of it.
  __shared__ double bred[2][16];
  __shared__ int is_last;
  __syncthreads();
  const double wmin = wave_min(pts), wmax = wave_max(rotmax);
  if (lane == 0) {
    bred[0][w] = wmin;
    bred[1][w] = wmax;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double mn = INFINITY, mx = -INFINITY;
    for (int q = 0; q < W; ++q) {
      mn = fmin(mn, bred[0][q]);
      mx = fmax(mx, bred[1][q]);
    }
    atomicMax(&sc->pts_key, ~okey(mn));
    atomicMax(&sc->rot_key, okey(mx));
    // both atomics complete (s_waitcnt 0) before the arrival count
    __builtin_amdgcn_s_waitcnt(0);
    const unsigned tkt = atomicAdd(&sc->counter, 1u);
    is_last = (tkt == (unsigned)d.nblocks - 1);
  }
  __syncthreads();
  if (!is_last) return;
  if (threadIdx.x == 0) {
    const unsigned long long kmin = atomicExch(&sc->pts_key, 0ull);
    const unsigned long long kmax = atomicExch(&sc->rot_key, 0ull);
    const double mn = kmin ? okey_inv(~kmin) : INFINITY;
    const double mx = kmax ? okey_inv(kmax) : -INFINITY;
    if (active) {
      // src/umwm_physics.F90:44-58
      double dts = fmin(fmin(mn, d.dtamin), d.dtg - sc->sumt);
      if (!d.restart && first) dts = 0;
      sc->dts = dts;
      sc->dta = dts;
      sc->sumt = sc->sumt + dts;
      // src/umwm_advection.F90:352-365 (signed maxval)
      if (rotation) {
        sc->maxrot = mx;
        sc->dtr_temp = (mx <= d.tiny) ? dts : d.dth / mx;
      }
      sc->dtr = fmin(sc->dtr_temp, dts);
      sc->nsub = sc->nsub + 1;
      sc->last = sc->sumt >= d.dtg ? 1 : 0;
      if constexpr (SUR) {
        if (d.sur_r) {
          // the gate for the next substep: closed while the sampled error
          // exceeds the tolerance; a substep with no audited point keeps it
          const double an = __longlong_as_double(
              (long long)atomicExch(reinterpret_cast<unsigned long long*>(&sc->sur_anum), 0ull));
          const double ad = __longlong_as_double(
              (long long)atomicExch(reinterpret_cast<unsigned long long*>(&sc->sur_aden), 0ull));
          if (ad > 0) sc->sur_closed = an <= (double)d.sur_audit_tol * ad ? 0 : 1;
        }
      }
    }
    sc->active = active ? 1 : 0;
    sc->counter = 0;
  }
}

// ---------------------------------------------------------------------
// k_integrate
// ---------------------------------------------------------------------
template <typename T>
struct Prefix {
  T a0, ac, as;
};

template <typename T>
__device__ __forceinline__ Prefix<T> load_prefix(const T* pfx, size_t ldi, int o, int j) {
  const rsrc_t pf = freq_rsrc(pfx, o, 4, ldi);
  Prefix<T> r;
  if constexpr (sizeof(T) == 4) {
    const auto v = bld16<T>(pf, (uint32_t)j * 16u);
    r.a0 = v[0];
    r.ac = v[1];
    r.as = v[2];
  } else {
    const auto v = bld16<T>(pf, (uint32_t)j * 32u);
    r.a0 = v[0];
    r.ac = v[1];
    r.as = bld<T>(pf, (uint32_t)j * 32u + 16u);
  }
  return r;
}

template <typename T, int DMAX, bool FULL, bool SUR>
__global__ __launch_bounds__(512) void k_integrate(const DomainDev<T>* __restrict__ doms,
                                                   const BlockTask* __restrict__ tasks) {
  const BlockTask tk = tasks[blockIdx.x];
  const DomainDev<T>& d = doms[tk.dom];
  const Scalars* sc = d.sc;
  if (!sc->active) return;  // uniform over the block
  stage_fm64<T>();
  const int W = blockDim.x >> 6, w = wave_id();
  const int j = tk.i0 + (threadIdx.x & 63);
  const bool valid = j < d.im;
  const int pm = d.pm, om = d.om;
  const size_t ldi = d.ldi;
  const int D = FULL ? DMAX : (pm + W - 1) / W, p0 = w * D;
  const int nd = FULL ? DMAX : max(0, min(D, pm - p0));
  const T dts = T(sc->dts);
  const T dtx = dts * Tr<T>::kExpScale;  // dts in the exponential's base
  const uint32_t vj = (uint32_t)j * (uint32_t)sizeof(T);
  const uint32_t vr = (uint32_t)j * (uint32_t)(FT_N * sizeof(T));
  const uint32_t rb = (uint32_t)ldi * (uint32_t)sizeof(T);
  // Arrays the frequency loop uses, held in registers (see k_rates).
  const T* const eb = d.e;
  T* const esb = d.es;
  const T* const ftab = d.ftab;
  const T* const pfxb = d.pfx;
  const T* const fr = d.f;

  const PointState<T> s = point_stored(d, j);
  const int oc = s.oc;
  const T sin_diss1 = T(d.sin_fac * d.fieldscale1), fs2 = T(d.fieldscale2);
  const T tpsf = T(d.twopisds_fac), spow = T(d.sds_power), ipow = T(d.inv_sds_power);
  const T mssf = T(d.mss_fac), dsnl = dts * T(d.snl_fac);
  const T nu4 = T(4 * d.nu_water), ig = T(1 / d.g);
  // direction constants of this wave, loaded once.  Slots q >= nd are
  // computed but never used: their loads and stores go through a
  // zero-range descriptor (loads read 0, stores are dropped), so the
  // direction loop has no per-slot branch and the compiler can interleave
  // the directions.
  const rsrc_t rz = make_rsrc(eb, 0u);
  T mq[DMAX], ctq[DMAX], stq[DMAX];
#pragma unroll
  for (int q = 0; q < DMAX; ++q) {
    T ct, st, c2, s2;
    cld4(d.dir4, p0 + q, ct, st, c2, s2);
    mq[q] = s.cw * ct + s.sw * st > T(0) ? fs2 : T(1);
    ctq[q] = ct;
    stq[q] = st;
  }
  // per direction, the partial E* = (E^n + partial E_s) / 2 of the two
  // previous frequencies, still waiting for their snl gains
  T H1[DMAX], H2[DMAX];
#pragma unroll
  for (int q = 0; q < DMAX; ++q) H1[q] = H2[q] = 0;
  T b1m1 = 0, b2m2 = 0, b2m1 = 0;  // bf1r(o-1), bf2r(o-2), bf2r(o-1)
  // Surrogate (FP32): the lanes k_rates let take the network's rates use
  // its spreading factor dm^2, clamped to [1, (1 + 2 mss_fac m0)^2] as
  // there, wherever the moment prefix's dm^2 would enter: the dissipation
  // in the integration, in snl and in the diagnostic tail.
  [[maybe_unused]] int use = 0;
  [[maybe_unused]] bool do_sur = false;
  [[maybe_unused]] const uint16_t* const rsur = d.sur_r;
  [[maybe_unused]] const uint32_t vj2 = (uint32_t)j * 2u, rb2 = (uint32_t)ldi * 2u;
  if constexpr (SUR) {
    if (rsur) {
      use = valid ? pload(d.sur_use, ldi, j) : 0;
      do_sur = __any(use != 0);
    }
  }
  T ec[DMAX];
  [[maybe_unused]] uint32_t rc[DMAX];  // raw bf16 bits of dm^2 - 1 (see k_rates)
  {
    const rsrc_t ep = freq_rsrc(eb, 0, pm, ldi);
#pragma unroll
    for (int q = 0; q < DMAX; ++q) ec[q] = bld<T>(FULL || q < nd ? ep : rz, vj, (p0 + q) * rb);
    if constexpr (SUR) {
      if (do_sur) {
        const rsrc_t rp = sur_rsrc(rsur, 0, pm, ldi);
#pragma unroll
        for (int q = 0; q < DMAX; ++q) rc[q] = bld_u16(FULL || q < nd ? rp : rz, vj2 + (p0 + q) * rb2);
      }
    }
  }
  FreqTab<T> tc;
  Prefix<T> pc{};
  // FP32 derives the tables from k (TabGen) and prefetches the prefix;
  // FP64 loads both when needed (as in k_rates)
  constexpr bool kDerive = sizeof(T) == 4;
  TabGen tg;
  if constexpr (kDerive) {
    tg.init(d, j, vj);
    pc = load_prefix(pfxb, ldi, 0, j);
  }
  for (int o = 0; o < om; ++o) {
    const int on = o + 1 < om ? o + 1 : o;
    T en[DMAX];
    const rsrc_t epn = freq_rsrc(eb, on, pm, ldi);
#pragma unroll
    for (int q = 0; q < DMAX; ++q) en[q] = bld<T>(FULL || q < nd ? epn : rz, vj, (p0 + q) * rb);
    [[maybe_unused]] uint32_t rn[DMAX];
    if constexpr (SUR) {
      if (do_sur) {
        const rsrc_t rp = sur_rsrc(rsur, on, pm, ldi);
#pragma unroll
        for (int q = 0; q < DMAX; ++q) rn[q] = bld_u16(FULL || q < nd ? rp : rz, vj2 + (p0 + q) * rb2);
      }
    }
    Prefix<T> pn{};
    float knext = 0.f;
    if constexpr (kDerive) {
      knext = tg.load_k(o + 4);
      pn = load_prefix(pfxb, ldi, on, j);
      tg.table(o, tc);
    } else {
      load_tab<kIntegrateSlots>(ftab, ldi, o, vr, tc);
      pc = load_prefix(pfxb, ldi, o, j);
    }

    const T k = tc[FT_K], fo = cld(fr, o);
    const T sdt = s.csdt * k;
    const T sice = T(2) * tc[FT_CG0] * s.ice;
    const T ctail = sdt + nu4 * (k * k) - sice;  // sdt + sdv - sice
    const T crest = tc[FT_SBF] + ctail;
    const SinFreq<T> sf = sin_freq(s, tc[FT_LOGL2], tc[FT_CP0], fo * k * ig, sin_diss1, fs2);
    const T cf = tpsf * fo;
    const T coth = tc[FT_COTH];
    const T lk4 = T(4) * Tr<T>::lg(k);        // lg(k^4)
    const T Lk = spow * lk4 + Tr<T>::lg(cf);  // sds before cothkd
    const T c0 = -dtx * crest, dtc = dtx * coth;
    const T icfc = Tr<T>::rcp(cf * coth);
    // the spreading factor's upper bound, from the zeroth-moment prefix
    [[maybe_unused]] const T qhi = (T(1) + T(2) * mssf * pc.a0) * (T(1) + T(2) * mssf * pc.a0);
    // dm = q0 + qc cos 2th + qs sin 2th = (q0 - qc) + ct (2 qc ct + 2 qs st),
    // with cos 2th = 2 ct^2 - 1 and sin 2th = 2 ct st: the sweep keeps
    // only ct and st per direction
    const T qc = mssf * pc.ac;
    const T q0m = T(1) + mssf * pc.a0 - qc, qc2 = T(2) * qc, qs2 = T(2) * mssf * pc.as;
    const bool prog = o <= oc;
    const bool in1 = (o >= 1) && (o - 1 <= oc);  // bin o-1 is prognostic
    const bool fin = valid && (o >= 2) && (o - 2 <= oc);  // bin o-2 completes now
    const T hdb1 = in1 ? T(0.5) * dts * b1m1 : T(0), hdb2 = T(0.5) * dts * b2m2;
    const uint32_t vtail = (valid && !prog) ? vj : kDropOffset;
    const uint32_t vfin = fin ? vj : kDropOffset;
    const bool anyfin = __any(fin), anytail = __any(valid && !prog);
    const rsrc_t rs0 = freq_rsrc(esb, o, pm, ldi);
    // max(o, 2) - 2 rather than a conditional: the latter becomes a vector
    // saturating subtract, which moves the descriptor to vector registers
    const rsrc_t rs2 = freq_rsrc(esb, max(o, 2) - 2, pm, ldi);
    // One pass over this wave's directions; PROG / TAIL / FIN say whether
    // any lane of the wave integrates a prognostic bin, fills a tail bin
    // or completes bin o-2 at this frequency (wave-uniform, so each
    // combination is a straight-line loop).
    auto sweep = [&](auto PROG, auto TAIL, auto FIN) {
#pragma unroll
      for (int q = 0; q < DMAX; ++q) {
        const uint32_t so = (uint32_t)(p0 + q) * rb;
        const T ct = ctq[q], st = stq[q];
        const T e = ec[q];
        const T ss = sin_m(sf, sin_y(sf, ct, st), mq[q]);
        T dm = fma(ct, fma(qc2, ct, qs2 * st), q0m);
        dm = dm * dm;
        if constexpr (SUR) {
          if (do_sur && use) dm = fmin(fmax(T(1) + bf16_bits_to_f(rc[q]), T(1)), qhi);  // dm^2 - 1
        }
        // sds before cothkd; |e|: see k_rates
        const T sdsp = dm * Tr<T>::ex(fma(spow, Tr<T>::lg(fabs(e)), Lk));
        const T X = sdsp * e;
        T G = 0;
        if constexpr (decltype(PROG)::value) {
          // E_s = e exp(dts (ssin - sds - sbf - sdt - sdv + sice)) + dts snl
          const T a = fma(-dtc, sdsp, fma(dtx, ss, c0));
          G = prog ? fma(-dsnl, X, e * Tr<T>::ex(a)) : T(0);
        }
        if constexpr (decltype(TAIL)::value) {
          // diagnostic tail, src/umwm_physics.F90:70-80
          const T num = fmax(ss, T(0)) - ctail;
          const T t = fmax(num, T(0)) * (icfc * Tr<T>::rcp(dm));
          const T es_ = Tr<T>::ex(fma(ipow, Tr<T>::lg(t), -lk4));  // k^-4 t^(1/p)
          bst(FULL || q < nd ? rs0 : rz, vtail, T(0.5) * (e + es_), so);
        }
        const T h1 = fma(hdb1, X, H1[q]);  // bf1r(o-1) X(o) into bin o-1
        if constexpr (decltype(FIN)::value)  // bf2r(o-2) X(o) completes bin o-2
          bst(FULL || q < nd ? rs2 : rz, vfin, fma(hdb2, X, H2[q]), so);
        H2[q] = h1;
        H1[q] = T(0.5) * (e + G);
      }
    };
    using Y = std::true_type;
    using N = std::false_type;
    const bool anyprog = __any(valid && prog);
    if (anyprog) {
      if (anytail) {
        if (anyfin) sweep(Y{}, Y{}, Y{}); else sweep(Y{}, Y{}, N{});
      } else {
        if (anyfin) sweep(Y{}, N{}, Y{}); else sweep(Y{}, N{}, N{});
      }
    } else {
      if (anytail) {
        if (anyfin) sweep(N{}, Y{}, Y{}); else sweep(N{}, Y{}, N{});
      } else {
        if (anyfin) sweep(N{}, N{}, Y{}); else sweep(N{}, N{}, N{});
      }
    }
    b2m2 = b2m1;
    b2m1 = tc[FT_B2];
    b1m1 = tc[FT_B1];
#pragma unroll
    for (int q = 0; q < DMAX; ++q) ec[q] = en[q];
    if constexpr (SUR) {
      if (do_sur) {
#pragma unroll
        for (int q = 0; q < DMAX; ++q) rc[q] = rn[q];
      }
    }
    if constexpr (kDerive) {
      tg.advance(o, knext);
      pc = pn;
    }
  }
}

// ---------------------------------------------------------------------
// k_advect: propagation + refraction + e = ef, one wave per (frequency o,
// 64 sea points); workgroups of FW waves take FW consecutive frequencies.
// Directions are processed in blocks of NB (pm is a multiple of 4) with
// the next block's loads in flight while the current block is computed.
//
// The upwind flux of src/umwm_advection.F90:28-44 is evaluated by sign:
// both x-edge speeds cg(i)+cg(nbr) times cos(th + curv) share the sign of
// cos(th + curv), so the x flux is c (Ae E*(upwind east) - Aw E*(upwind
// west)), Ae = 2 (cg(i) + cg(ie)) dye, and likewise in y; refraction's
// upwind flux (:166-172) is r (E* upwind) per edge.  Regrouped, equal to
// the reference up to rounding.
// ---------------------------------------------------------------------
template <typename T>
struct AdvLoads {
  T eo, ee, ew, en, es;  // E* at the point and its four neighbours
  T en0;                 // E^n at the point
};

template <typename T, int NB>
__global__ __launch_bounds__(256) void k_advect(const DomainDev<T>* __restrict__ doms,
                                                const BlockTask* __restrict__ tasks) {
  const BlockTask tk = tasks[blockIdx.x];
  const DomainDev<T>& d = doms[tk.dom];
  const int o = (int)blockIdx.y * (int)(blockDim.x >> 6) + wave_id();
  const Scalars* sc = d.sc;
  if (o >= d.om || !sc->active) return;
  const int j = tk.i0 + (threadIdx.x & 63);
  const bool valid = j < d.im;
  const int pm = d.pm;
  const size_t ldi = d.ldi, ldq = d.ldq;
  const uint32_t vj = (uint32_t)j * (uint32_t)sizeof(T);
  const uint32_t vst = valid ? vj : kDropOffset;
  const uint32_t rb = (uint32_t)ldi * (uint32_t)sizeof(T);
  constexpr uint32_t E = sizeof(T);
  const T kdta = T(0.25) * T(sc->dta), dtr = T(sc->dtr);
  // wave-uniform in a scalar register, so that the current terms below are
  // a scalar branch rather than an exec-masked block issued for every bin
  const bool anycur = __builtin_amdgcn_readfirstlane(sc->anycur) != 0;

  const PointState<T> s = point_stored(d, j);
  const bool prog = o <= s.oc, rot = o <= pload(d.oc_rot, ldi, j);
  const T sin_diss1 = T(d.sin_fac * d.fieldscale1), fs2 = T(d.fieldscale2);
  // neighbour gathers read the frequency-major tables (one plane per
  // table), where 64 neighbours share a few cache lines
  const rsrc_t cpr = row_rsrc(d.cp0, o, ldi), cgr = row_rsrc(d.cg0, o, ldi);
  const T cp0 = bld<T>(cpr, vj);
  T logl2, fkovg;
  if constexpr (sizeof(T) == 4) {
    // FP32 forms both from cp0 = 2 pi f / k: f k / g = 2 pi f^2 / (g cp0),
    // log(l2/z) with l2 = cp0 / (2 f) capped at 20 m
    // (src/umwm_init.F90:1502,1507,1529-1533)
    const float fo = cld(d.f, o), icp = __builtin_amdgcn_rcpf(cp0);
    fkovg = float(kTwoPiD / d.g) * fo * fo * icp;
    const float l2 = 0.5f * cp0 * __builtin_amdgcn_rcpf(fo);
    logl2 = l2 > 20.f ? float(log(20.0 / d.z))
                      : 0.6931471805599453f * __builtin_amdgcn_logf(l2) - float(log(d.z));
  } else {
    logl2 = bld<T>(row_rsrc(d.logl2overz, o, ldi), vj);
    fkovg = bld<T>(row_rsrc(d.fkovg, o, ldi), vj);
  }
  const SinFreq<T> sf = sin_freq(s, logl2, cp0, fkovg, sin_diss1, fs2);
  const rsrc_t pe = freq_rsrc(d.e, o, pm, ldi);    // E^n in, E^{n+1} out
  const rsrc_t ps = freq_rsrc(d.es, o, pm, ldi);   // E*
  // the stress partials' cp0(o, i-1) (src/umwm_stress.F90:219,316) and kdk,
  // loaded here so that their latency hides behind the direction loop
  const T cp_prev = bld<T>(cpr, (uint32_t)((j == 0) ? d.im : j - 1) * E);
  const T kdk = bld<T>(row_rsrc(d.kdk, o, ldi), vj);
  // direction tables in registers (see k_rates)
  const T* const cthb = d.cth;
  const T* const sthb = d.sth;
  // stress_vector_integral partial sums over the directions of this
  // frequency (FP32 within the frequency in the FP32 path, FP64 across)
  T tx = 0, ty = 0, hs = 0;
  auto stress_acc = [&](T ct, T st, T v) {
    const T m = s.cw * ct + s.sw * st > T(0) ? fs2 : T(1);
    T ss = sin_m(sf, sin_y(sf, ct, st), m);
    if (!prog) ss = fmax(ss, T(0));
    const T wv = v * ss;
    tx = fma(wv, ct, tx);
    ty = fma(wv, st, ty);
    hs += v;
  };
  if (__all(!prog || !valid)) {
    // above the cut-off only the diagnostic tail: e = ef = E_s
    for (int pb = 0; pb < pm; pb += NB) {
      T v[NB];
#pragma unroll
      for (int q = 0; q < NB; ++q)
        v[q] = fma(T(2), bld<T>(ps, vj, (pb + q) * rb), -bld_nt<T>(pe, vj, (pb + q) * rb));
#pragma unroll
      for (int q = 0; q < NB; ++q) {
        bst_stream(pe, vst, v[q], (pb + q) * rb);
        stress_acc(cld(cthb, pb + q), cld(sthb, pb + q), v[q]);
      }
    }
  } else {
    const int ie = pload(d.ie, ldi, j), iw = pload(d.iw, ldi, j);
    const int in = pload(d.in_, ldi, j), is = pload(d.is_, ldi, j);
    const uint32_t vie = (uint32_t)ie * E, viw = (uint32_t)iw * E;
    const uint32_t vin = (uint32_t)in * E, vis = (uint32_t)is * E;
    const T dye = pload(d.dye, ldi, j), dyw = pload(d.dyw, ldi, j);
    const T dxn = pload(d.dxn, ldi, j), dxs = pload(d.dxs, ldi, j);
    const T kk = kdta * pload(d.oneovar, ldi, j);
    // ice above fice_uth zeroes the update (src/umwm_advection.F90:93-96)
    const T keep = pload(d.fice, ldq, j) > T(d.fice_uth) ? T(0) : T(1);
    T feup = 0, fedn = 0, fwup = 0, fwdn = 0, fnup = 0, fndn = 0, fsup = 0, fsdn = 0;
    if (anycur) {
      const int iie = pload(d.iie, ldi, j), iiw = pload(d.iiw, ldi, j);
      const int iin = pload(d.iin, ldi, j), iis = pload(d.iis, ldi, j);
      const T ue = s.uc + pload(d.uc, ldq, iie), uw = s.uc + pload(d.uc, ldq, iiw);
      const T vn = s.vc + pload(d.vc, ldq, iin), vs = s.vc + pload(d.vc, ldq, iis);
      feup = (ue + fabs(ue)) * dye;
      fedn = (ue - fabs(ue)) * dye;
      fwup = (uw + fabs(uw)) * dyw;
      fwdn = (uw - fabs(uw)) * dyw;
      fnup = (vn + fabs(vn)) * dxn;
      fndn = (vn - fabs(vn)) * dxn;
      fsup = (vs + fabs(vs)) * dxs;
      fsdn = (vs - fabs(vs)) * dxs;
    }
    // rotation: r(p)/2 = rx sth + ry cth + r0 (zero above the first
    // substep's cut-off), src/umwm_advection.F90:119-156
    T rx = 0, ry = 0, r0 = 0;
    if (rot) {
      const T dcx = bld<T>(cpr, vie) - bld<T>(cpr, viw);
      const T dcy = bld<T>(cpr, vin) - bld<T>(cpr, vis);
      const T dvc = pload(d.vc, ldq, pload(d.iie, ldi, j)) - pload(d.vc, ldq, pload(d.iiw, ldi, j));
      const T duc = pload(d.uc, ldq, pload(d.iin, ldi, j)) - pload(d.uc, ldq, pload(d.iis, ldi, j));
      const T odx = pload(d.oneovdx, ldi, j), ody = pload(d.oneovdy, ldi, j);
      rx = T(0.25) * dcx * odx;
      ry = T(-0.25) * dcy * ody;
      r0 = T(0.25) * (dvc * odx - duc * ody);
    }
    const T cgi = bld<T>(cgr, vj);
    const T Ae = T(2) * (cgi + bld<T>(cgr, vie)) * dye, Aw = T(2) * (cgi + bld<T>(cgr, viw)) * dyw;
    const T As = T(2) * (cgi + bld<T>(cgr, vis)) * dxs, An = T(2) * (cgi + bld<T>(cgr, vin)) * dxn;
    const T ccu = pload(d.ccurv, ldi, j), scu = pload(d.scurv, ldi, j);
    const T kr = dtr * T(d.oneovdth);
    auto rotp = [&](int p) { return fma(rx, cld(sthb, p), fma(ry, cld(cthb, p), r0)); };
    auto load_block = [&](int pb, AdvLoads<T>* L) {
#pragma unroll
      for (int q = 0; q < NB; ++q) {
        const uint32_t so = (pb + q) * rb;
        L[q].eo = bld<T>(ps, vj, so);
        L[q].ee = bld<T>(ps, vie, so);
        L[q].ew = bld<T>(ps, viw, so);
        L[q].en = bld<T>(ps, vin, so);
        L[q].es = bld<T>(ps, vis, so);
        L[q].en0 = bld_nt<T>(pe, vj, so);
      }
    };
    AdvLoads<T> A[NB], B[NB];
    const T efirst = bld<T>(ps, vj, 0);  // E*(0), periodic neighbour of pm-1
    T eprev = bld<T>(ps, vj, (pm - 1) * rb);
    load_block(0, A);
    T rprev = rotp(pm - 1), rcur = rotp(0);
    for (int pb = 0; pb < pm; pb += NB) {
      const bool more = pb + NB < pm;
      if (more) load_block(pb + NB, B);
#pragma unroll
      for (int q = 0; q < NB; ++q) {
        const int p = pb + q;
        const T enext = q + 1 < NB ? A[(q + 1) < NB ? q + 1 : 0].eo : (more ? B[0].eo : efirst);
        const int pn = p + 1 < pm ? p + 1 : 0;
        const T rnext = rotp(pn);
        const T ct = cld(cthb, p), st = cld(sthb, p);
        const T cc = fma(ct, ccu, -st * scu);  // cos(th + curv)
        const T sn = fma(st, ccu, ct * scu);   // sin(th + curv)
        const T e0 = A[q].eo, ee = A[q].ee, ew = A[q].ew, en = A[q].en, es = A[q].es;
        const bool px = cc > T(0), py = sn > T(0);
        T flux = cc * fma(Ae, px ? e0 : ee, -Aw * (px ? ew : e0)) +
                 sn * fma(An, py ? e0 : en, -As * (py ? es : e0));
        if (anycur)
          flux = flux + (feup * e0 + fedn * ee - fwup * ew - fwdn * e0 + fnup * e0 + fndn * en -
                         fsup * es - fsdn * e0);
        const T Es = fma(T(2), e0, -A[q].en0);  // E_s = 2 E* - E^n
        const T rl = rcur + rnext, rr = rcur + rprev;  // edge rotation rates
        const T fr = fma(rl, rl > T(0) ? e0 : enext, -rr * (rr > T(0) ? eprev : e0));
        T v = fma(-kr, fr, fma(-kk, flux, Es));
        v = prog ? v * keep : Es;
        bst_stream(pe, vst, v, p * rb);
        stress_acc(ct, st, v);
        eprev = e0;
        rprev = rcur;
        rcur = rnext;
      }
#pragma unroll
      for (int q = 0; q < NB; ++q) A[q] = B[q];
    }
  }
  if (valid) {
    const T icp = T(1) / cp_prev;
    const rsrc_t sp = freq_rsrc(d.spart, o, 4, ldi);
    if constexpr (sizeof(T) == 4) {
      bst16(sp, (uint32_t)j * 16u, f32x4{tx * icp, ty * icp, hs * kdk, T(0)});
    } else {
      bst16(sp, (uint32_t)j * 32u, f64x2{tx * icp, ty * icp});
      bst(sp, (uint32_t)j * 32u + 16u, hs * kdk);
    }
  }
}

// ---------------------------------------------------------------------
// k_keep: on the substep that reaches dtg (and only when the domain keeps
// them), saves E^n and the incoming ustar, from which stress('ocn')
// rebuilds that substep's ssin, sds and snl at output time.  Launched
// between k_integrate (which reads E^n) and k_advect (which overwrites it).
// ---------------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(64) void k_keep(const DomainDev<T>* __restrict__ doms,
                                             const BlockTask* __restrict__ tasks) {
  const BlockTask tk = tasks[blockIdx.x];
  const DomainDev<T>& d = doms[tk.dom];
  if (!d.sc->active || !d.sc->last || !d.e_keep) return;
  const int j = tk.i0 + (int)threadIdx.x;
  if (j >= d.im) return;
  const size_t ldi = d.ldi, rows = (size_t)d.om * d.pm;
  for (size_t r = 0; r < rows; ++r) d.e_keep[r * ldi + j] = d.e[r * ldi + j];
  d.ustar_keep[j] = d.ustar[j];
}

// ---------------------------------------------------------------------
// k_stress: frequency sums + stress('atm') chain per point
// ---------------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(64) void k_stress(const DomainDev<T>* __restrict__ doms,
                                               const BlockTask* __restrict__ tasks) {
  const BlockTask tk = tasks[blockIdx.x];
  const DomainDev<T>& d = doms[tk.dom];
  if (!d.sc->active) return;
  const int j = tk.i0 + (int)threadIdx.x;
  if (j >= d.im) return;
  const size_t ldi = d.ldi, ldq = d.ldq;
  const int om = d.om;
  constexpr uint32_t E = sizeof(T);
  T w = pload(d.wspd, ldq, j);
  w = w > T(1e-2) ? w : T(1e-2);
  const int oc = pload(d.oc, ldi, j);
  StressIn si;
  si.sx = si.sy = si.dx = si.dy = 0;
  double hs = 0;
  for (int o = 0; o < om; ++o) {
    const double kdk = bld<T>(row_rsrc(d.kdk, o, ldi), (uint32_t)j * E);
    const rsrc_t sp = freq_rsrc(d.spart, o, 4, ldi);
    double tx, ty, hso;
    if constexpr (sizeof(T) == 4) {
      const f32x4 v = bld16<float>(sp, (uint32_t)j * 16u);
      tx = v[0];
      ty = v[1];
      hso = v[2];
    } else {
      const f64x2 v = bld16<double>(sp, (uint32_t)j * 32u);
      tx = v[0];
      ty = v[1];
      hso = bld<double>(sp, (uint32_t)j * 32u + 16u);
    }
    si.sx += tx * kdk;
    si.sy += ty * kdk;
    if (o >= oc) {
      si.dx += tx * kdk;
      si.dy += ty * kdk;
    }
    if (o == om - 1) {
      si.tomx = tx;
      si.tomy = ty;
    }
    hs += hso;
  }
  si.wspd = w;
  si.wdir = pload(d.wdir, ldq, j);
  si.uc = pload(d.uc, ldq, j);
  si.vc = pload(d.vc, ldq, j);
  si.us1 = pload(d.us1, ldi, j);
  si.vs1 = pload(d.vs1, ldi, j);
  si.rhoa = pload(d.rhoa, ldq, j);
  si.rhow = pload(d.rhow, ldq, j);
  si.kom = bld<T>(row_rsrc(d.k, om - 1, ldi), (uint32_t)j * E);
  store_stress(d, j, stress_atm_point(si, d.dthg, d.z, d.nu_air, d.kappa));
  pstore(d.hs_ice, ldi, j, 4 * sqrt(hs * d.dth));
}

template <typename T, int D1, int D2, bool SUR>
void launch_sweeps(const DomainDev<T>* doms, const BlockTask* tasks, dim3 grid, size_t sh1,
                   int first, int rotation, Tile tile, hipStream_t s, hipEvent_t* ev) {
  // a launch whose waves all hold DMAX directions (W D = pm) takes the
  // variant with no per-slot masks
  if (tile.W1 * D1 == tile.pmax)
    hipLaunchKernelGGL((k_rates<T, D1, true, SUR>), grid, dim3(64 * tile.W1), sh1, s, doms, tasks,
                       first, rotation);
  else
    hipLaunchKernelGGL((k_rates<T, D1, false, SUR>), grid, dim3(64 * tile.W1), sh1, s, doms, tasks,
                       first, rotation);
  if (ev) (void)hipEventRecord(ev[1], s);
  if (tile.W2 * D2 == tile.pmax)
    hipLaunchKernelGGL((k_integrate<T, D2, true, SUR>), grid, dim3(64 * tile.W2), 0, s, doms, tasks);
  else
    hipLaunchKernelGGL((k_integrate<T, D2, false, SUR>), grid, dim3(64 * tile.W2), 0, s, doms,
                       tasks);
}

template <typename T, int D1, int D2>
int launch_all(const DomainDev<T>* doms, const BlockTask* tasks, int ntasks, int omax, int first,
               int rotation, Tile tile, bool keep, hipStream_t s, hipEvent_t* ev, bool sur) {
  const dim3 grid(ntasks);
  // the moments' double buffer, and the surrogate's four cross-wave sums
  const size_t sh1 = (size_t)std::max(2 * 3 * tile.W1, 4 * tile.W1) * 64 * sizeof(double);
  if (ev) (void)hipEventRecord(ev[0], s);
  if constexpr (std::is_same_v<T, float>) {
    if (sur)
      launch_sweeps<T, D1, D2, true>(doms, tasks, grid, sh1, first, rotation, tile, s, ev);
    else
      launch_sweeps<T, D1, D2, false>(doms, tasks, grid, sh1, first, rotation, tile, s, ev);
  } else {
    if (sur) return fail(UMWM_ERR_ARGUMENT, "the surrogate serves the FP32 path only");
    launch_sweeps<T, D1, D2, false>(doms, tasks, grid, sh1, first, rotation, tile, s, ev);
  }
  if (keep) hipLaunchKernelGGL((k_keep<T>), grid, dim3(64), 0, s, doms, tasks);
  if (ev) (void)hipEventRecord(ev[2], s);
  // advect: one wave per (64 points, frequency), FW frequencies per
  // workgroup (2 measured 1-2 % faster than 4 or 8 on C2)
  const int fw = 2;
  const dim3 ga(ntasks, (omax + fw - 1) / fw);
  if (tile.NB == 4)
    hipLaunchKernelGGL((k_advect<T, 4>), ga, dim3(64 * fw), 0, s, doms, tasks);
  else
    hipLaunchKernelGGL((k_advect<T, 2>), ga, dim3(64 * fw), 0, s, doms, tasks);
  if (ev) (void)hipEventRecord(ev[3], s);
  hipLaunchKernelGGL((k_stress<T>), grid, dim3(64), 0, s, doms, tasks);
  if (ev) (void)hipEventRecord(ev[4], s);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail(UMWM_ERR_HIP, std::string("fused launch: ") + hipGetErrorString(e));
  return 0;
}

template <typename T, int D1>
int launch_d2(const DomainDev<T>* doms, const BlockTask* tasks, int ntasks, int omax, int first,
              int rotation, Tile tile, bool keep, hipStream_t s, hipEvent_t* ev, bool sur) {
#define UMWM_D2(D2_)                                                                         \
  case D2_:                                                                                  \
    return launch_all<T, D1, D2_>(doms, tasks, ntasks, omax, first, rotation, tile, keep, s, ev, sur);
  switch (tile.D2) {
    UMWM_D2(3)
    UMWM_D2(6)
    UMWM_D2(9)
    UMWM_D2(12)
    UMWM_D2(18)
  }
#undef UMWM_D2
  return fail(UMWM_ERR_ARGUMENT, "unsupported integrate tile D");
}

}  // namespace

bool tile_supported(int D) { return D == 3 || D == 6 || D == 9 || D == 12 || D == 18 || D == 36; }

template <typename T>
int launch_fused(const DomainDev<T>* doms, const BlockTask* tasks, int ntasks, int omax,
                 int first, int rotation, Tile tile, bool keep, hipStream_t s, hipEvent_t* ev,
                 bool sur) {
  if (ntasks <= 0) return 0;
  if (tile.W1 < 1 || tile.W1 > 8 || tile.W2 < 1 || tile.W2 > 8)
    return fail(UMWM_ERR_ARGUMENT, "tile W must be 1..8");
  if (sur) {
    // the surrogate variants hold the network's dm^2 of this and the next
    // frequency beside the spectrum: more waves of fewer directions each,
    // the same W D, keep them off scratch (k_rates above 18 directions and
    // k_integrate above 9 spill)
    while (tile.D1 > 18 && tile.D1 % 2 == 0 && tile.W1 * 2 <= 8 && tile_supported(tile.D1 / 2)) {
      tile.D1 /= 2;
      tile.W1 *= 2;
    }
    while (tile.D2 > 9 && tile.D2 % 2 == 0 && tile.W2 * 2 <= 8 && tile_supported(tile.D2 / 2)) {
      tile.D2 /= 2;
      tile.W2 *= 2;
    }
  }
  switch (tile.D1) {
    case 6: return launch_d2<T, 6>(doms, tasks, ntasks, omax, first, rotation, tile, keep, s, ev, sur);
    case 9: return launch_d2<T, 9>(doms, tasks, ntasks, omax, first, rotation, tile, keep, s, ev, sur);
    case 18: return launch_d2<T, 18>(doms, tasks, ntasks, omax, first, rotation, tile, keep, s, ev, sur);
    case 36: return launch_d2<T, 36>(doms, tasks, ntasks, omax, first, rotation, tile, keep, s, ev, sur);
  }
  return fail(UMWM_ERR_ARGUMENT, "unsupported rates tile D");
}

template int launch_fused<float>(const DomainDev<float>*, const BlockTask*, int, int, int, int,
                                 Tile, bool, hipStream_t, hipEvent_t*, bool);
template int launch_fused<double>(const DomainDev<double>*, const BlockTask*, int, int, int, int,
                                  Tile, bool, hipStream_t, hipEvent_t*, bool);

}  // namespace umwm
