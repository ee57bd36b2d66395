// Fused substep kernels for gfx950.
//
// One substep of the reference (src/umwm_top.F90:95-115) is four kernels:
//   k_rates      sin + sds + rate + per-point max|rate| -> physics_time_step,
//                then a grid-wide min (last-block reduction) -> dts, dta,
//                sumt, dtr (src/umwm_physics.F90:22-58,
//                src/umwm_advection.F90:329-365).  Reads E^n once and
//                writes the exclusive frequency prefix of the directional
//                cos^2 convolution (3 values per frequency and point).
//   k_integrate  sin + sds + snl + exponential integration + diagnostic
//                tail (src/umwm_physics.F90:60-84); writes E_s (ef) and
//                E* = (E^n + E_s)/2 (es).
//   k_advect     upwind propagation on E* (src/umwm_advection.F90:18-99),
//                refraction (:102-180) and e = ef, one thread per
//                (frequency, sea point); also the per-frequency partial
//                sums of stress_vector_integral (src/umwm_stress.F90:313-331).
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
// the next frequency's spectrum and tables while computing the current.
//
// snl(o) = bf1r(o) X(o+1) + bf2r(o) X(o+2) - snl_fac X(o), X = sds e
// (src/umwm_source_functions.F90:165-171) is folded into the integration
// as each X becomes known: k_integrate keeps per direction only the
// partial E_s of the two previous frequencies and their E^n.
//
// Per-direction angles use the angle-sum identity with per-point cos/sin
// (cos(wdir - th) = cos wdir cos th + sin wdir sin th, and the same for the
// curvature-rotated ray directions of src/umwm_init.F90:1304-1309).
#include <hip/hip_runtime.h>

#include "domain.hpp"
#include "physics.hpp"

namespace umwm {
namespace {

// Per-point prologue of sin_d12 (src/umwm_source_functions.F90:21-52):
// wind floor, cut-off bin, sheltering.  k_rates evaluates it and wave 0
// stores the results; the later kernels of the substep read them back
// (load_point_fast) instead of repeating the search and the exp.
template <typename T>
__device__ inline void load_point(const DomainDev<T>& d, int j, SinPoint<T>& s, T& csdt,
                                  double& ice, int& oc, T& cw, T& sw, bool write) {
  T w = d.wspd[j];
  w = w > T(1e-2) ? w : T(1e-2);
  T fc = T(0.53) * T(d.g) / w;
  if (fc > T(d.fprog)) fc = T(d.fprog);
  oc = cutoff_bin(fc, d.f, d.om);
  const T shelt = sheltering_coare35(w);
  const T ust = T(d.ustar[j]);
  const T rr = d.rhorat[j];
  s.wspd = w;
  s.a = T(2.5) * ust;
  s.pre = (T(1) - d.fice[j]) * T(kTwoPiD) * rr;
  s.shelt = shelt;
  s.uc = d.uc[j];
  s.vc = d.vc[j];
  s.oc = oc;
  csdt = T(d.sdt_fac) * sqrt(rr) * ust;
  ice = (d.fice[j] > T(d.fice_lth)) ? sice_coef(d.hs_ice[j]) : 0.0;
  cw = cos(d.wdir[j]);
  sw = sin(d.wdir[j]);
  if (write) {
    d.wspd[j] = w;
    d.fcutoff[j] = fc;
    d.oc[j] = oc;
    d.shelt[j] = shelt;
    d.cwdir[j] = cw;
    d.swdir[j] = sw;
  }
}

template <typename T>
__device__ inline void load_point_fast(const DomainDev<T>& d, int j, SinPoint<T>& s, T& csdt,
                                       double& ice, int& oc, T& cw, T& sw) {
  const T ust = T(d.ustar[j]);
  const T rr = d.rhorat[j];
  oc = d.oc[j];
  s.wspd = d.wspd[j];
  s.a = T(2.5) * ust;
  s.pre = (T(1) - d.fice[j]) * T(kTwoPiD) * rr;
  s.shelt = d.shelt[j];
  s.uc = d.uc[j];
  s.vc = d.vc[j];
  s.oc = oc;
  csdt = T(d.sdt_fac) * sqrt(rr) * ust;
  ice = (d.fice[j] > T(d.fice_lth)) ? sice_coef(d.hs_ice[j]) : 0.0;
  cw = d.cwdir[j];
  sw = d.swdir[j];
}

// Rotation tendency of src/umwm_advection.F90:335-338 for one direction.
template <typename T>
__device__ inline T rot_dir(T dcx, T dcy, T sth, T cth, T vce, T vcw, T ucn, T ucs, T odx,
                            T ody) {
  return T(0.5) * (((dcx * sth + vce) - vcw) * odx - ((dcy * cth + ucn) - ucs) * ody);
}

// Per-frequency, per-point inputs of the source terms.
template <typename T>
struct FreqTab {
  T cp0, logl2, fkovg, k4, coth, sbf, sdv, sarg, k, cg0, k3dk;
};

template <typename T>
__device__ inline void load_tab(const DomainDev<T>& d, size_t t, FreqTab<T>& r) {
  r.cp0 = d.cp0[t];
  r.logl2 = d.logl2overz[t];
  r.fkovg = d.fkovg[t];
  r.k4 = d.k4[t];
  r.coth = d.cothkd[t];
  r.sbf = d.sbf[t];
  r.sdv = d.sdv[t];
  r.sarg = d.snl_arg[t];
  r.k = d.k[t];
  r.cg0 = d.cg0[t];
  r.k3dk = d.k3dk[t];
}

// ---------------------------------------------------------------------
// k_rates
// ---------------------------------------------------------------------
template <typename T, int DMAX>
__global__ __launch_bounds__(512) void k_rates(const DomainDev<T>* __restrict__ doms,
                                               const BlockTask* __restrict__ tasks, int first,
                                               int rotation) {
  extern __shared__ double smem[];
  const BlockTask tk = tasks[blockIdx.x];
  const DomainDev<T>& d = doms[tk.dom];
  Scalars* sc = d.sc;
  const int W = blockDim.x >> 6, lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int j = tk.i0 + lane;
  const bool active = sc->sumt < d.dtg;
  const bool valid = active && j < d.im;
  const int pm = d.pm, om = d.om;
  const size_t ldi = d.ldi, sl = (size_t)pm * ldi;
  const int D = (pm + W - 1) / W, p0 = w * D;
  const int nd = max(0, min(D, pm - p0));  // directions of this wave (uniform)
  double pts = INFINITY, rotmax = -INFINITY;

  if (active) {
    SinPoint<T> s;
    T csdt;
    double ice;
    int oc;
    T cw, sw;
    load_point(d, j, s, csdt, ice, oc, cw, sw, valid && w == 0);
    if (rotation && valid && w == 0) d.oc_rot[j] = oc;
    int ie = 0, iw = 0, in = 0, is = 0;
    T vce = 0, vcw = 0, ucn = 0, ucs = 0, odx = 0, ody = 0;
    if (rotation) {
      ie = d.ie[j]; iw = d.iw[j]; in = d.in_[j]; is = d.is_[j];
      vce = d.vc[d.iie[j]]; vcw = d.vc[d.iiw[j]]; ucn = d.uc[d.iin[j]]; ucs = d.uc[d.iis[j]];
      odx = d.oneovdx[j];
      ody = d.oneovdy[j];
    }
    const T sin_fac = T(d.sin_fac), fs1 = T(d.fieldscale1), fs2 = T(d.fieldscale2);
    const T tpsf = T(d.twopisds_fac), spow = T(d.sds_power), mssf = T(d.mss_fac);
    const double hdth = 0.5 * d.dth;
    T* __restrict__ PF = d.pfx + j;
    const T* __restrict__ EN = d.e + j;
    const size_t pstr = (size_t)om * ldi;
    double A0 = 0, AC = 0, AS = 0;
    T rmax = 0;
    // FP32 prefetches the next frequency's tables as well as its spectrum;
    // FP64 prefetches only the spectrum to stay within the register file
    constexpr bool kPrefetchTables = sizeof(T) == 4;
    T ec[DMAX];
    FreqTab<T> tc;
#pragma unroll
    for (int q = 0; q < DMAX; ++q) ec[q] = (q < nd) ? EN[(size_t)(p0 + q) * ldi] : T(0);
    if constexpr (kPrefetchTables) load_tab(d, j, tc);
    for (int o = 0; o < om; ++o) {
      // issue the next frequency's loads before computing this one
      const int on = o + 1 < om ? o + 1 : o;
      T en[DMAX];
      FreqTab<T> tn;
#pragma unroll
      for (int q = 0; q < DMAX; ++q)
        en[q] = (q < nd) ? EN[(size_t)on * sl + (size_t)(p0 + q) * ldi] : T(0);
      if constexpr (kPrefetchTables)
        load_tab(d, (size_t)on * ldi + j, tn);
      else
        load_tab(d, (size_t)o * ldi + j, tc);
      T dcx = 0, dcy = 0;
      if (rotation) {
        const T* cp = d.cp0 + (size_t)o * ldi;
        dcx = cp[ie] - cp[iw];
        dcy = cp[in] - cp[is];
      }
      const T sdt = csdt * tc.k;
      const T sice = ice != 0.0 ? T(2 * (double)tc.cg0 * ice) : T(0);
      const T fo = d.f[o];
      const T ba = s.wspd + s.a * tc.logl2;
      const T Aw = ba * cw - s.uc, Bw = ba * sw - s.vc;
      const T pa0 = T(hdth * A0), pac = T(hdth * AC), pas = T(hdth * AS);
      if (valid && w == 0) {
        PF[(size_t)o * ldi] = pa0;
        PF[pstr + (size_t)o * ldi] = pac;
        PF[2 * pstr + (size_t)o * ldi] = pas;
      }
      const T cf = tpsf * fo;
      const bool prog = o <= oc;
      T m0 = 0, mc = 0, ms = 0;
#pragma unroll
      for (int q = 0; q < DMAX; ++q) {
        if (q < nd) {
          const int p = p0 + q;
          const T e = ec[q];
          const T ct = d.cth[p], st = d.sth[p], c2 = d.c2th[p], s2 = d.s2th[p];
          const T cwt = cw * ct + sw * st;
          const T ss = ssin_ab(s, Aw, Bw, cwt, tc.cp0, ct, st, tc.fkovg, sin_fac, fs1, fs2, o);
          const T mss = pa0 + pac * c2 + pas * s2;
          T dm = T(1) + mssf * mss;
          dm = dm * dm;
          const T sds = cf * dm * pw(e * tc.k4, spow) * tc.coth;
          if (prog) {
            const T r = ss - sds * tc.sarg - tc.sbf - sdt - tc.sdv + sice;
            const T a = fabs(r);
            rmax = a > rmax ? a : rmax;
            if (rotation)
              rotmax = fmax(rotmax, (double)rot_dir(dcx, dcy, st, ct, vce, vcw, ucn, ucs, odx, ody));
          } else if (rotation) {
            rotmax = fmax(rotmax, 0.0);
          }
          m0 += e;
          mc += e * c2;
          ms += e * s2;
        }
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
      const double k3 = tc.k3dk;
      A0 += k3 * s0;
      AC += k3 * s2c;
      AS += k3 * s2s;
#pragma unroll
      for (int q = 0; q < DMAX; ++q) ec[q] = en[q];
      if constexpr (kPrefetchTables) tc = tn;
    }
    // max |rate| of the point over all waves
    __syncthreads();
    smem[w * 64 + lane] = (double)rmax;
    __syncthreads();
    if (w == 0) {
      double mx = 0;
      for (int q = 0; q < W; ++q) mx = fmax(mx, smem[q * 64 + lane]);
      pts = d.explim / mx;
      if (valid) d.pts[j] = pts;
      if (!valid) pts = INFINITY;
    }
    if (!valid) rotmax = -INFINITY;
  }
  // block reductions -> per-domain partials; the last block finishes
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
    d.part_min[tk.local] = mn;
    d.part_max[tk.local] = mx;
    __threadfence();
    const unsigned tkt = atomicAdd(&sc->counter, 1u);
    is_last = (tkt == (unsigned)d.nblocks - 1);
  }
  __syncthreads();
  if (!is_last) return;
  __threadfence();
  double mn = INFINITY, mx = -INFINITY;
  for (int q = threadIdx.x; q < d.nblocks; q += blockDim.x) {
    mn = fmin(mn, __hip_atomic_load(&d.part_min[q], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
    mx = fmax(mx, __hip_atomic_load(&d.part_max[q], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
  }
  mn = wave_min(mn);
  mx = wave_max(mx);
  __syncthreads();
  if (lane == 0) {
    bred[0][w] = mn;
    bred[1][w] = mx;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    mn = INFINITY;
    mx = -INFINITY;
    for (int q = 0; q < W; ++q) {
      mn = fmin(mn, bred[0][q]);
      mx = fmax(mx, bred[1][q]);
    }
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
    }
    sc->active = active ? 1 : 0;
    sc->counter = 0;
  }
}

// ---------------------------------------------------------------------
// k_integrate
// ---------------------------------------------------------------------
template <typename T, int DMAX>
__global__ __launch_bounds__(512) void k_integrate(const DomainDev<T>* __restrict__ doms,
                                                   const BlockTask* __restrict__ tasks) {
  const BlockTask tk = tasks[blockIdx.x];
  const DomainDev<T>& d = doms[tk.dom];
  const Scalars* sc = d.sc;
  if (!sc->active) return;
  const int W = blockDim.x >> 6, lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int j = tk.i0 + lane;
  const bool valid = j < d.im;
  const int pm = d.pm, om = d.om;
  const size_t ldi = d.ldi, sl = (size_t)pm * ldi;
  const int D = (pm + W - 1) / W, p0 = w * D;
  const int nd = max(0, min(D, pm - p0));
  const T dts = T(sc->dts);

  SinPoint<T> s;
  T csdt;
  double ice;
  int oc;
  T cw, sw;
  load_point_fast(d, j, s, csdt, ice, oc, cw, sw);
  const T sin_fac = T(d.sin_fac), fs1 = T(d.fieldscale1), fs2 = T(d.fieldscale2);
  const T tpsf = T(d.twopisds_fac), spow = T(d.sds_power), ipow = T(d.inv_sds_power);
  const T mssf = T(d.mss_fac), dsnl = dts * T(d.snl_fac);
  const T* __restrict__ PF = d.pfx + j;
  const T* __restrict__ EN = d.e + j;
  const size_t pstr = (size_t)om * ldi;
  // per direction: partial E_s and E^n of the two previous frequencies
  T G1[DMAX], G2[DMAX], E1[DMAX], E2[DMAX];
#pragma unroll
  for (int q = 0; q < DMAX; ++q) G1[q] = G2[q] = E1[q] = E2[q] = 0;
  T b1m1 = 0, b2m2 = 0, b2m1 = 0;  // bf1r(o-1), bf2r(o-2), bf2r(o-1)
  T* __restrict__ EF = d.ef + j;
  T* __restrict__ ES = d.es + j;
  // FP32 prefetches the next frequency's tables as well as its spectrum;
  // FP64 prefetches only the spectrum to stay within the register file
  constexpr bool kPrefetchTables = sizeof(T) == 4;
  T ec[DMAX];
#pragma unroll
  for (int q = 0; q < DMAX; ++q) ec[q] = (q < nd) ? EN[(size_t)(p0 + q) * ldi] : T(0);
  FreqTab<T> tc;
  T bc1 = 0, bc2 = 0, okc = 0, pa0 = 0, pac = 0, pas = 0;
  if constexpr (kPrefetchTables) {
    load_tab(d, j, tc);
    bc1 = d.bf1r[j];
    bc2 = d.bf2r[j];
    okc = d.oneoverk4[j];
    pa0 = PF[0];
    pac = PF[pstr];
    pas = PF[2 * pstr];
  }
  for (int o = 0; o < om; ++o) {
    const int on = o + 1 < om ? o + 1 : o;
    const size_t tn_ = (size_t)on * ldi + j;
    T en[DMAX];
#pragma unroll
    for (int q = 0; q < DMAX; ++q)
      en[q] = (q < nd) ? EN[(size_t)on * sl + (size_t)(p0 + q) * ldi] : T(0);
    FreqTab<T> tn;
    T bn1 = 0, bn2 = 0, okn = 0, qa0 = 0, qac = 0, qas = 0;
    if constexpr (kPrefetchTables) {
      load_tab(d, tn_, tn);
      bn1 = d.bf1r[tn_];
      bn2 = d.bf2r[tn_];
      okn = d.oneoverk4[tn_];
      qa0 = PF[(size_t)on * ldi];
      qac = PF[pstr + (size_t)on * ldi];
      qas = PF[2 * pstr + (size_t)on * ldi];
    } else {
      const size_t tc_ = (size_t)o * ldi + j;
      load_tab(d, tc_, tc);
      bc1 = d.bf1r[tc_];
      bc2 = d.bf2r[tc_];
      okc = d.oneoverk4[tc_];
      pa0 = PF[(size_t)o * ldi];
      pac = PF[pstr + (size_t)o * ldi];
      pas = PF[2 * pstr + (size_t)o * ldi];
    }

    const T sdt = csdt * tc.k;
    const T sice = ice != 0.0 ? T(2 * (double)tc.cg0 * ice) : T(0);
    const T fo = d.f[o];
    const T ba = s.wspd + s.a * tc.logl2;
    const T Aw = ba * cw - s.uc, Bw = ba * sw - s.vc;
    const T cf = tpsf * fo;
    const bool prog = o <= oc;
    const bool in1 = (o >= 1) && (o - 1 <= oc);  // bin o-1 is prognostic
    const bool fin = (o >= 2) && (o - 2 <= oc);  // bin o-2 completes now
#pragma unroll
    for (int q = 0; q < DMAX; ++q) {
      if (q < nd) {
        const int p = p0 + q;
        const T e = ec[q];
        const T ct = d.cth[p], st = d.sth[p], c2 = d.c2th[p], s2 = d.s2th[p];
        const T cwt = cw * ct + sw * st;
        const T ss = ssin_ab(s, Aw, Bw, cwt, tc.cp0, ct, st, tc.fkovg, sin_fac, fs1, fs2, o);
        const T mss = pa0 + pac * c2 + pas * s2;
        T dm = T(1) + mssf * mss;
        dm = dm * dm;
        const T sdsp = cf * dm * pw(e * tc.k4, spow);  // sds before cothkd
        const T X = sdsp * e;
        T G = 0;
        if (prog) {
          G = e * ex(dts * (ss - sdsp * tc.coth - tc.sbf - sdt - tc.sdv + sice)) - dsnl * X;
        } else if (valid) {
          const T num = ss - sdt - tc.sdv + sice;
          T es_ = 0;
          if (num >= T(0)) es_ = okc * pw(num / (cf * dm * tc.coth), ipow);
          const size_t b = (size_t)o * sl + (size_t)p * ldi;
          EF[b] = es_;
          ES[b] = T(0.5) * (e + es_);
        }
        T g1 = G1[q];
        if (in1) g1 = g1 + dts * b1m1 * X;  // bf1r(o-1) X(o) into bin o-1
        if (fin && valid) {
          const T es_ = G2[q] + dts * b2m2 * X;  // bf2r(o-2) X(o) completes bin o-2
          const size_t b = (size_t)(o - 2) * sl + (size_t)p * ldi;
          EF[b] = es_;
          ES[b] = T(0.5) * (E2[q] + es_);
        }
        G2[q] = g1;
        G1[q] = G;
        E2[q] = E1[q];
        E1[q] = e;
      }
    }
    b2m2 = b2m1;
    b2m1 = bc2;
    b1m1 = bc1;
#pragma unroll
    for (int q = 0; q < DMAX; ++q) ec[q] = en[q];
    if constexpr (kPrefetchTables) {
      tc = tn;
      bc1 = bn1;
      bc2 = bn2;
      okc = okn;
      pa0 = qa0;
      pac = qac;
      pas = qas;
    }
  }
}

// ---------------------------------------------------------------------
// k_advect: propagation + refraction + e = ef, one thread per
// (frequency o, sea point); workgroups of FW waves take FW consecutive
// frequencies of 64 points.  Directions are processed in blocks of NB
// (pm is a multiple of 4) with the next block's loads in flight while the
// current block is computed, so each wave keeps 2*NB directions' worth of
// gathers outstanding.
// ---------------------------------------------------------------------
template <typename T>
struct AdvLoads {
  T eo, ee, ew, en, es, ef;
};

template <typename T, int NB>
__global__ __launch_bounds__(256) void k_advect(const DomainDev<T>* __restrict__ doms,
                                                const BlockTask* __restrict__ tasks) {
  const BlockTask tk = tasks[blockIdx.x];
  const DomainDev<T>& d = doms[tk.dom];
  const int o = blockIdx.y * (blockDim.x >> 6) + (threadIdx.x >> 6);
  const Scalars* sc = d.sc;
  if (o >= d.om || !sc->active) return;
  const int j = tk.i0 + (threadIdx.x & 63);
  const bool valid = j < d.im;
  const int pm = d.pm;
  const size_t ldi = d.ldi, sl = (size_t)pm * ldi;
  const T kdta = T(0.25) * T(sc->dta), dtr = T(sc->dtr);
  const bool anycur = sc->anycur != 0;

  SinPoint<T> s;
  T csdt;
  double ice;
  int oc;
  T cw, sw;
  load_point_fast(d, j, s, csdt, ice, oc, cw, sw);
  (void)csdt;
  (void)ice;
  const bool prog = o <= oc, rot = o <= d.oc_rot[j];
  const size_t t = (size_t)o * ldi + j;
  // e, ef and es are distinct arrays; restrict lets loads move past stores
  const T* __restrict__ ESo = d.es + (size_t)o * sl;
  const T* __restrict__ EFo = d.ef + (size_t)o * sl + j;
  T* __restrict__ Eo = d.e + (size_t)o * sl + j;
  const T sin_fac = T(d.sin_fac), fs1 = T(d.fieldscale1), fs2 = T(d.fieldscale2);
  const T* cp = d.cp0 + (size_t)o * ldi;
  const T ba = s.wspd + s.a * d.logl2overz[t];
  const T Aw = ba * cw - s.uc, Bw = ba * sw - s.vc;
  const T fkovg = d.fkovg[t], cp0 = cp[j];
  double tx = 0, ty = 0, hs = 0;
  auto stress_acc = [&](int p, T v) {
    const T ct = d.cth[p], st = d.sth[p];
    const T ss = ssin_ab(s, Aw, Bw, cw * ct + sw * st, cp0, ct, st, fkovg, sin_fac, fs1, fs2, o);
    const double es_ = (double)v * (double)ss;
    tx += es_ * (double)ct;
    ty += es_ * (double)st;
    hs += (double)v;
  };
  if (!prog) {
    // above the cut-off only the diagnostic tail: e = ef
    for (int p0 = 0; p0 < pm; p0 += NB) {
      T v[NB];
#pragma unroll
      for (int q = 0; q < NB; ++q) v[q] = EFo[(size_t)(p0 + q) * ldi];
#pragma unroll
      for (int q = 0; q < NB; ++q) {
        if (valid) Eo[(size_t)(p0 + q) * ldi] = v[q];
        stress_acc(p0 + q, v[q]);
      }
    }
  } else {
    const int ie = d.ie[j], iw = d.iw[j], in = d.in_[j], is = d.is_[j];
    const T dye = d.dye[j], dyw = d.dyw[j], dxn = d.dxn[j], dxs = d.dxs[j];
    const T oneovar = d.oneovar[j], odx = d.oneovdx[j], ody = d.oneovdy[j];
    const T vce = d.vc[d.iie[j]], vcw = d.vc[d.iiw[j]], ucn = d.uc[d.iin[j]], ucs = d.uc[d.iis[j]];
    const bool iced = d.fice[j] > T(d.fice_uth);
    T feup = 0, fedn = 0, fwup = 0, fwdn = 0, fnup = 0, fndn = 0, fsup = 0, fsdn = 0;
    if (anycur) {
      const T ue = s.uc + d.uc[d.iie[j]], uw = s.uc + d.uc[d.iiw[j]];
      const T vn = s.vc + d.vc[d.iin[j]], vs = s.vc + d.vc[d.iis[j]];
      feup = (ue + fabs(ue)) * dye;
      fedn = (ue - fabs(ue)) * dye;
      fwup = (uw + fabs(uw)) * dyw;
      fwdn = (uw - fabs(uw)) * dyw;
      fnup = (vn + fabs(vn)) * dxn;
      fndn = (vn - fabs(vn)) * dxn;
      fsup = (vs + fabs(vs)) * dxs;
      fsdn = (vs - fabs(vs)) * dxs;
    }
    const T* cg = d.cg0 + (size_t)o * ldi;
    const T cgi = cg[j], cge_ = cg[ie], cgw_ = cg[iw], cgs_ = cg[is], cgn_ = cg[in];
    const T dcx = cp[ie] - cp[iw], dcy = cp[in] - cp[is];
    const T ccu = d.ccurv[j], scu = d.scurv[j];
    const T odth = T(d.oneovdth);
    auto rotp = [&](int p) {
      return rot ? rot_dir(dcx, dcy, d.sth[p], d.cth[p], vce, vcw, ucn, ucs, odx, ody) : T(0);
    };
    auto load_block = [&](int p0, AdvLoads<T>* L) {
#pragma unroll
      for (int q = 0; q < NB; ++q) {
        const T* Ep = ESo + (size_t)(p0 + q) * ldi;
        L[q].eo = Ep[j];
        L[q].ee = Ep[ie];
        L[q].ew = Ep[iw];
        L[q].en = Ep[in];
        L[q].es = Ep[is];
        L[q].ef = EFo[(size_t)(p0 + q) * ldi];
      }
    };
    AdvLoads<T> A[NB], B[NB];
    const T efirst = ESo[j];  // E*(0) for the periodic neighbour of pm-1
    T eprev = ESo[(size_t)(pm - 1) * ldi + j];
    load_block(0, A);
    T rprev = rotp(pm - 1), rcur = rotp(0);
    for (int p0 = 0; p0 < pm; p0 += NB) {
      const bool more = p0 + NB < pm;
      if (more) load_block(p0 + NB, B);
#pragma unroll
      for (int q = 0; q < NB; ++q) {
        const int p = p0 + q;
        const T enext = q + 1 < NB ? A[(q + 1) < NB ? q + 1 : 0].eo : (more ? B[0].eo : efirst);
        const int pn = p + 1 < pm ? p + 1 : 0;
        const T rnext = rotp(pn);
        const T ct = d.cth[p], st = d.sth[p];
        const T cc = ct * ccu - st * scu;  // cos(th + curv)
        const T sn = st * ccu + ct * scu;  // sin(th + curv)
        const T cge = cgi * cc + cge_ * cc;
        const T cgw = cgi * cc + cgw_ * cc;
        const T cgs = cgi * sn + cgs_ * sn;
        const T cgn = cgi * sn + cgn_ * sn;
        const T e0 = A[q].eo, ee = A[q].ee, ew = A[q].ew, en = A[q].en, es = A[q].es;
        T flux = (cge + fabs(cge)) * dye * e0 + (cge - fabs(cge)) * dye * ee -
                 (cgw + fabs(cgw)) * dyw * ew - (cgw - fabs(cgw)) * dyw * e0 +
                 (cgn + fabs(cgn)) * dxn * e0 + (cgn - fabs(cgn)) * dxn * en -
                 (cgs + fabs(cgs)) * dxs * es - (cgs - fabs(cgs)) * dxs * e0;
        if (anycur)
          flux = flux + (feup * e0 + fedn * ee - fwup * ew - fwdn * e0 + fnup * e0 + fndn * en -
                         fsup * es - fsdn * e0);
        T v = A[q].ef - kdta * flux * oneovar;
        if (iced) v = 0;
        const T rl = T(0.5) * (rcur + rnext), rr = T(0.5) * (rcur + rprev);
        const T fr = T(0.5) *
                     ((rl + fabs(rl)) * e0 + (rl - fabs(rl)) * enext - (rr + fabs(rr)) * eprev -
                      (rr - fabs(rr)) * e0) *
                     odth;
        v = v - dtr * fr;
        if (iced) v = 0;
        if (valid) Eo[(size_t)p * ldi] = v;
        stress_acc(p, v);
        eprev = e0;
        rprev = rcur;
        rcur = rnext;
      }
#pragma unroll
      for (int q = 0; q < NB; ++q) A[q] = B[q];
    }
  }
  if (valid) {
    const size_t st3 = (size_t)d.om * ldi;
    const int jc = (j == 0) ? d.im : j - 1;  // cp0(o, i-1), src/umwm_stress.F90:219,316
    const double icp = 1.0 / (double)cp[jc];
    d.spart[t] = tx * icp;
    d.spart[st3 + t] = ty * icp;
    d.spart[2 * st3 + t] = hs * (double)d.kdk[t];
  }
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
  const size_t ldi = d.ldi, st3 = (size_t)d.om * ldi;
  const int om = d.om;
  T w = d.wspd[j];
  w = w > T(1e-2) ? w : T(1e-2);
  const int oc = d.oc[j];
  StressIn si;
  si.sx = si.sy = si.dx = si.dy = 0;
  double hs = 0;
  for (int o = 0; o < om; ++o) {
    const size_t t = (size_t)o * ldi + j;
    const double kdk = d.kdk[t];
    const double tx = d.spart[t], ty = d.spart[st3 + t];
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
    hs += d.spart[2 * st3 + t];
  }
  si.wspd = w;
  si.wdir = d.wdir[j];
  si.uc = d.uc[j];
  si.vc = d.vc[j];
  si.us1 = d.us1[j];
  si.vs1 = d.vs1[j];
  si.rhoa = d.rhoa[j];
  si.rhow = d.rhow[j];
  si.kom = d.k[(size_t)(om - 1) * ldi + j];
  store_stress(d, j, stress_atm_point(si, d.dthg, d.z, d.nu_air, d.kappa));
  d.hs_ice[j] = 4 * sqrt(hs * d.dth);
}

template <typename T, int D1, int D2>
int launch_all(const DomainDev<T>* doms, const BlockTask* tasks, int ntasks, int omax, int first,
               int rotation, Tile tile, hipStream_t s, hipEvent_t* ev) {
  const dim3 grid(ntasks);
  const size_t sh1 = tile.W1 > 1 ? (size_t)2 * 3 * tile.W1 * 64 * sizeof(double)
                                 : (size_t)64 * sizeof(double);
  if (ev) (void)hipEventRecord(ev[0], s);
  hipLaunchKernelGGL((k_rates<T, D1>), grid, dim3(64 * tile.W1), sh1, s, doms, tasks, first,
                     rotation);
  if (ev) (void)hipEventRecord(ev[1], s);
  hipLaunchKernelGGL((k_integrate<T, D2>), grid, dim3(64 * tile.W2), 0, s, doms, tasks);
  if (ev) (void)hipEventRecord(ev[2], s);
  // advect: one wave per (64 points, frequency), FW frequencies per workgroup
  const int fw = 4;
  const dim3 ga(ntasks, (omax + fw - 1) / fw);
  if (tile.NB == 4)
    hipLaunchKernelGGL((k_advect<T, 4>), ga, dim3(64 * fw), 0, s, doms, tasks);
  else
    hipLaunchKernelGGL((k_advect<T, 2>), ga, dim3(64 * fw), 0, s, doms, tasks);
  hipLaunchKernelGGL((k_stress<T>), grid, dim3(64), 0, s, doms, tasks);
  if (ev) (void)hipEventRecord(ev[3], s);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail(UMWM_ERR_HIP, std::string("fused launch: ") + hipGetErrorString(e));
  return 0;
}

template <typename T, int D1>
int launch_d2(const DomainDev<T>* doms, const BlockTask* tasks, int ntasks, int omax, int first,
              int rotation, Tile tile, hipStream_t s, hipEvent_t* ev) {
  switch (tile.D2) {
    case 3: return launch_all<T, D1, 3>(doms, tasks, ntasks, omax, first, rotation, tile, s, ev);
    case 6: return launch_all<T, D1, 6>(doms, tasks, ntasks, omax, first, rotation, tile, s, ev);
    case 9: return launch_all<T, D1, 9>(doms, tasks, ntasks, omax, first, rotation, tile, s, ev);
    case 12: return launch_all<T, D1, 12>(doms, tasks, ntasks, omax, first, rotation, tile, s, ev);
    case 18: return launch_all<T, D1, 18>(doms, tasks, ntasks, omax, first, rotation, tile, s, ev);
  }
  return fail(UMWM_ERR_ARGUMENT, "unsupported integrate tile D");
}

}  // namespace

bool tile_supported(int D) { return D == 3 || D == 6 || D == 9 || D == 12 || D == 18 || D == 36; }

template <typename T>
int launch_fused(const DomainDev<T>* doms, const BlockTask* tasks, int ntasks, int omax,
                 int first, int rotation, Tile tile, hipStream_t s, hipEvent_t* ev) {
  if (ntasks <= 0) return 0;
  if (tile.W1 < 1 || tile.W1 > 8 || tile.W2 < 1 || tile.W2 > 8)
    return fail(UMWM_ERR_ARGUMENT, "tile W must be 1..8");
  switch (tile.D1) {
    case 6: return launch_d2<T, 6>(doms, tasks, ntasks, omax, first, rotation, tile, s, ev);
    case 9: return launch_d2<T, 9>(doms, tasks, ntasks, omax, first, rotation, tile, s, ev);
    case 18: return launch_d2<T, 18>(doms, tasks, ntasks, omax, first, rotation, tile, s, ev);
    case 36: return launch_d2<T, 36>(doms, tasks, ntasks, omax, first, rotation, tile, s, ev);
  }
  return fail(UMWM_ERR_ARGUMENT, "unsupported rates tile D");
}

template int launch_fused<float>(const DomainDev<float>*, const BlockTask*, int, int, int, int,
                                 Tile, hipStream_t, hipEvent_t*);
template int launch_fused<double>(const DomainDev<double>*, const BlockTask*, int, int, int, int,
                                  Tile, hipStream_t, hipEvent_t*);

}  // namespace umwm
