// Fused substep kernels for gfx950.
//
// One substep of the reference (src/umwm_top.F90:95-115) is three kernels:
//   k_rates      sin + sds + rate + per-point max|rate| -> physics_time_step,
//                then a grid-wide min (last-block reduction) -> dts, dta,
//                sumt, dtr (src/umwm_physics.F90:22-58,
//                src/umwm_advection.F90:329-365).  Reads E^n once.
//   k_integrate  sin + sds + snl + exponential integration + diagnostic
//                tail (src/umwm_physics.F90:60-84); writes E_s (ef) and
//                E* = (E^n + E_s)/2 (es).
//   k_advect     upwind propagation on E* (src/umwm_advection.F90:18-99),
//                refraction (:102-180), e = ef, and stress('atm')
//                (src/umwm_stress.F90:216-264) fused as a per-point
//                reduction of the new spectrum.
//
// Thread mapping: a workgroup owns 64 consecutive sea points (one per
// lane, so every spectrum/table access is a coalesced 64-wide stream of
// the i-contiguous layout) and W waves; wave w owns directions
// p = w*D .. w*D+D-1.  Frequencies are swept in order inside the kernel:
// the exclusive prefix over frequency of the directional cos^2
// convolution (src/umwm_source_functions.F90:92-97) is evaluated through
// cos^2(a-b) = (1 + cos2a cos2b + sin2a sin2b)/2, i.e. three directional
// moments per frequency exchanged between waves through LDS, which turns
// the O(pm^2) convolution into O(pm).
//
// snl(o) = bf1r(o) X(o+1) + bf2r(o) X(o+2) - snl_fac X(o), X = sds e
// (src/umwm_source_functions.F90:165-171) is folded into the integration
// as each X becomes known: k_integrate keeps per direction only the
// partial E_s of the two previous frequencies and their E^n.
//
// Per-direction angles use the angle-sum identity with per-point cos/sin
// (cos(wdir - th) = cos wdir cos th + sin wdir sin th, and the same for the
// curvature-rotated ray directions of src/umwm_init.F90:1304-1309), which
// keeps the direction state out of registers.
#include <hip/hip_runtime.h>

#include "domain.hpp"
#include "physics.hpp"

namespace umwm {
namespace {

template <typename T>
__device__ inline void load_point(const DomainDev<T>& d, int j, SinPoint<T>& s, T& csdt,
                                  double& ice, int& oc, bool write) {
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
  if (write) {
    d.wspd[j] = w;
    d.fcutoff[j] = fc;
    d.oc[j] = oc;
    d.shelt[j] = shelt;
  }
}

// 0.5 * dth * (A0 + AC cos2th + AS sin2th) = prefix of the cos^2
// convolution, then dummy = (1 + mss_fac * mss)^2.
template <typename T>
__device__ inline T dummy_from(double A0, double AC, double AS, T c2, T s2, double hdth,
                               T mss_fac) {
  const T mss = T(hdth * (A0 + AC * (double)c2 + AS * (double)s2));
  T dm = T(1) + mss_fac * mss;
  return dm * dm;
}

// Exchanges the three directional moments of one frequency between the W
// waves of the workgroup and advances the running prefixes.  Double
// buffered, so one barrier per frequency suffices.
__device__ inline void moments_exchange(double* smem, int buf, int W, int w, int lane,
                                        double m0, double mc, double ms, double k3dk,
                                        double& A0, double& AC, double& AS) {
  double* b = smem + (size_t)buf * 3 * W * 64;
  b[(0 * W + w) * 64 + lane] = m0;
  b[(1 * W + w) * 64 + lane] = mc;
  b[(2 * W + w) * 64 + lane] = ms;
  __syncthreads();
  double s0 = 0, sc = 0, ss = 0;
  for (int q = 0; q < W; ++q) {
    s0 += b[(0 * W + q) * 64 + lane];
    sc += b[(1 * W + q) * 64 + lane];
    ss += b[(2 * W + q) * 64 + lane];
  }
  A0 += k3dk * s0;
  AC += k3dk * sc;
  AS += k3dk * ss;
}

// Rotation tendency of src/umwm_advection.F90:335-338 for one direction.
template <typename T>
__device__ inline T rot_dir(T dcx, T dcy, T sth, T cth, T vce, T vcw, T ucn, T ucs, T odx,
                            T ody) {
  return T(0.5) * (((dcx * sth + vce) - vcw) * odx - ((dcy * cth + ucn) - ucs) * ody);
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
  const size_t ldi = d.ldi;
  const int D = (pm + W - 1) / W, p0 = w * D;
  const int nd = max(0, min(D, pm - p0));  // directions of this wave (uniform)
  double pts = INFINITY, rotmax = -INFINITY;

  if (active) {
    SinPoint<T> s;
    T csdt;
    double ice;
    int oc;
    load_point(d, j, s, csdt, ice, oc, valid && w == 0);
    if (rotation && valid && w == 0) d.oc_rot[j] = oc;
    const T cw = cos(d.wdir[j]), sw = sin(d.wdir[j]);
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
    double A0 = 0, AC = 0, AS = 0;
    T rmax = 0;
    for (int o = 0; o < om; ++o) {
      const size_t t = (size_t)o * ldi + j;
      const T cp0 = d.cp0[t], logl2 = d.logl2overz[t], fkovg = d.fkovg[t], k4 = d.k4[t];
      const T coth = d.cothkd[t], sbf = d.sbf[t], sdv = d.sdv[t], sarg = d.snl_arg[t];
      const T sdt = csdt * d.k[t];
      const T sice = ice != 0.0 ? T(2 * (double)d.cg0[t] * ice) : T(0);
      const T fo = d.f[o];
      const bool prog = o <= oc;
      T dcx = 0, dcy = 0;
      if (rotation) {
        const T* cp = d.cp0 + (size_t)o * ldi;
        dcx = cp[ie] - cp[iw];
        dcy = cp[in] - cp[is];
      }
      const T* E = d.e + (size_t)o * pm * ldi + j;
      double m0 = 0, mc = 0, ms = 0;
#pragma unroll
      for (int q = 0; q < DMAX; ++q) {
        if (q < nd) {
          const int p = p0 + q;
          const T e = E[(size_t)p * ldi];
          const T ct = d.cth[p], st = d.sth[p], c2 = d.c2th[p], s2 = d.s2th[p];
          const T cwt = cw * ct + sw * st;
          const T ss = ssin_bin(s, logl2, cwt, cp0, ct, st, fkovg, sin_fac, fs1, fs2, o);
          const T dm = dummy_from(A0, AC, AS, c2, s2, hdth, mssf);
          const T sds = tpsf * fo * dm * pw(e * k4, spow) * coth;
          if (prog) {
            const T r = ss - sds * sarg - sbf - sdt - sdv + sice;
            const T a = fabs(r);
            rmax = a > rmax ? a : rmax;
            if (rotation)
              rotmax = fmax(rotmax, (double)rot_dir(dcx, dcy, st, ct, vce, vcw, ucn, ucs, odx, ody));
          } else if (rotation) {
            rotmax = fmax(rotmax, 0.0);
          }
          m0 += (double)e;
          mc += (double)(e * c2);
          ms += (double)(e * s2);
        }
      }
      moments_exchange(smem, o & 1, W, w, lane, m0, mc, ms, (double)d.k3dk[t], A0, AC, AS);
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
  extern __shared__ double smem[];
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
  load_point(d, j, s, csdt, ice, oc, false);
  const T cw = cos(d.wdir[j]), sw = sin(d.wdir[j]);
  const T sin_fac = T(d.sin_fac), fs1 = T(d.fieldscale1), fs2 = T(d.fieldscale2);
  const T tpsf = T(d.twopisds_fac), spow = T(d.sds_power), ipow = T(d.inv_sds_power);
  const T mssf = T(d.mss_fac), dsnl = dts * T(d.snl_fac);
  const double hdth = 0.5 * d.dth;
  double A0 = 0, AC = 0, AS = 0;
  // per direction: partial E_s and E^n of the two previous frequencies
  T G1[DMAX], G2[DMAX], E1[DMAX], E2[DMAX];
#pragma unroll
  for (int q = 0; q < DMAX; ++q) G1[q] = G2[q] = E1[q] = E2[q] = 0;
  T b1m1 = 0, b2m2 = 0, b2m1 = 0;  // bf1r(o-1), bf2r(o-2), bf2r(o-1)
  T* EF = d.ef + j;
  T* ES = d.es + j;
  for (int o = 0; o < om; ++o) {
    const size_t t = (size_t)o * ldi + j;
    const T cp0 = d.cp0[t], logl2 = d.logl2overz[t], fkovg = d.fkovg[t], k4 = d.k4[t];
    const T coth = d.cothkd[t], sbf = d.sbf[t], sdv = d.sdv[t];
    const T sdt = csdt * d.k[t];
    const T sice = ice != 0.0 ? T(2 * (double)d.cg0[t] * ice) : T(0);
    const T b1 = d.bf1r[t], b2 = d.bf2r[t];
    const T fo = d.f[o];
    const T* E = d.e + (size_t)o * sl + j;
    double m0 = 0, mc = 0, ms = 0;
    const bool prog = o <= oc;
    const bool in1 = (o >= 1) && (o - 1 <= oc);  // bin o-1 is prognostic
    const bool fin = (o >= 2) && (o - 2 <= oc);  // bin o-2 completes now
#pragma unroll
    for (int q = 0; q < DMAX; ++q) {
      if (q < nd) {
        const int p = p0 + q;
        const T e = E[(size_t)p * ldi];
        const T ct = d.cth[p], st = d.sth[p], c2 = d.c2th[p], s2 = d.s2th[p];
        const T cwt = cw * ct + sw * st;
        const T ss = ssin_bin(s, logl2, cwt, cp0, ct, st, fkovg, sin_fac, fs1, fs2, o);
        const T dm = dummy_from(A0, AC, AS, c2, s2, hdth, mssf);
        const T sdsp = tpsf * fo * dm * pw(e * k4, spow);  // sds before cothkd
        const T X = sdsp * e;
        T G = 0;
        if (prog) {
          G = e * ex(dts * (ss - sdsp * coth - sbf - sdt - sdv + sice)) - dsnl * X;
        } else if (valid) {
          const T num = ss - sdt - sdv + sice;
          T es_ = 0;
          if (num >= T(0)) es_ = d.oneoverk4[t] * pw(num / (tpsf * fo * dm * coth), ipow);
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
        m0 += (double)e;
        mc += (double)(e * c2);
        ms += (double)(e * s2);
      }
    }
    b2m2 = b2m1;
    b2m1 = b2;
    b1m1 = b1;
    moments_exchange(smem, o & 1, W, w, lane, m0, mc, ms, (double)d.k3dk[t], A0, AC, AS);
  }
}

// ---------------------------------------------------------------------
// k_advect: propagation + refraction + update + stress('atm')
// ---------------------------------------------------------------------
template <typename T, int DMAX>
__global__ __launch_bounds__(512) void k_advect(const DomainDev<T>* __restrict__ doms,
                                                const BlockTask* __restrict__ tasks) {
  extern __shared__ double smem[];
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
  const T kdta = T(0.25) * T(sc->dta), dtr = T(sc->dtr);
  const bool anycur = sc->anycur != 0;

  SinPoint<T> s;
  T csdt;
  double ice;
  int oc;
  load_point(d, j, s, csdt, ice, oc, false);
  const int ocr = d.oc_rot[j];
  const int ie = d.ie[j], iw = d.iw[j], in = d.in_[j], is = d.is_[j];
  const T dye = d.dye[j], dyw = d.dyw[j], dxn = d.dxn[j], dxs = d.dxs[j];
  const T oneovar = d.oneovar[j], odx = d.oneovdx[j], ody = d.oneovdy[j];
  const bool iced = d.fice[j] > T(d.fice_uth);
  const T vce = d.vc[d.iie[j]], vcw = d.vc[d.iiw[j]], ucn = d.uc[d.iin[j]], ucs = d.uc[d.iis[j]];
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
  const T cw = cos(d.wdir[j]), sw = sin(d.wdir[j]);
  const T ccu = d.ccurv[j], scu = d.scurv[j];
  // refraction neighbours outside this wave's directions (waves without
  // directions read direction 0, which stays in bounds and is unused)
  const int pprev = nd > 0 ? d.pr[p0] : 0, pnext = nd > 0 ? d.pl[p0 + nd - 1] : 0;
  const T sin_fac = T(d.sin_fac), fs1 = T(d.fieldscale1), fs2 = T(d.fieldscale2);
  const T odth = T(d.oneovdth);
  const int jc = (j == 0) ? d.im : j - 1;  // cp0(o, i-1), src/umwm_stress.F90:219,316
  double sx = 0, sy = 0, dgx = 0, dgy = 0, tomx = 0, tomy = 0, hs = 0;
  for (int o = 0; o < om; ++o) {
    const size_t t = (size_t)o * ldi + j;
    const T* cg = d.cg0 + (size_t)o * ldi;
    const T* cp = d.cp0 + (size_t)o * ldi;
    const T cgi = cg[j], cge_ = cg[ie], cgw_ = cg[iw], cgs_ = cg[is], cgn_ = cg[in];
    const T dcx = cp[ie] - cp[iw], dcy = cp[in] - cp[is];
    const T logl2 = d.logl2overz[t], fkovg = d.fkovg[t], cp0 = cp[j];
    const double kdk = d.kdk[t], cpq = cp[jc];
    const bool prog = o <= oc, rot = o <= ocr;
    const T* ESo = d.es + (size_t)o * sl;
    T ev[DMAX];
#pragma unroll
    for (int q = 0; q < DMAX; ++q) ev[q] = (q < nd) ? ESo[(size_t)(p0 + q) * ldi + j] : T(0);
    const T eprev = ESo[(size_t)pprev * ldi + j], enext = ESo[(size_t)pnext * ldi + j];
    double tx = 0, ty = 0;
#pragma unroll
    for (int q = 0; q < DMAX; ++q) {
      if (q < nd) {
        const int p = p0 + q;
        const T ct = d.cth[p], st = d.sth[p];
        const size_t b = (size_t)o * sl + (size_t)p * ldi + j;
        T v = d.ef[b];
        if (prog) {
          const T* Ep = ESo + (size_t)p * ldi;
          const T e0 = ev[q], ee = Ep[ie], ew = Ep[iw], en = Ep[in], es = Ep[is];
          const T cc = ct * ccu - st * scu;  // cos(th + curv)
          const T sn = st * ccu + ct * scu;  // sin(th + curv)
          const T cge = cgi * cc + cge_ * cc;
          const T cgw = cgi * cc + cgw_ * cc;
          const T cgs = cgi * sn + cgs_ * sn;
          const T cgn = cgi * sn + cgn_ * sn;
          T flux = (cge + fabs(cge)) * dye * e0 + (cge - fabs(cge)) * dye * ee -
                   (cgw + fabs(cgw)) * dyw * ew - (cgw - fabs(cgw)) * dyw * e0 +
                   (cgn + fabs(cgn)) * dxn * e0 + (cgn - fabs(cgn)) * dxn * en -
                   (cgs + fabs(cgs)) * dxs * es - (cgs - fabs(cgs)) * dxs * e0;
          if (anycur)
            flux = flux + (feup * e0 + fedn * ee - fwup * ew - fwdn * e0 + fnup * e0 + fndn * en -
                           fsup * es - fsdn * e0);
          v = v - kdta * flux * oneovar;
          if (iced) v = 0;
          T rl = 0, rr = 0;
          if (rot) {
            const int pl = d.pl[p], pr = d.pr[p];
            const T r0 = rot_dir(dcx, dcy, st, ct, vce, vcw, ucn, ucs, odx, ody);
            const T rL = rot_dir(dcx, dcy, d.sth[pl], d.cth[pl], vce, vcw, ucn, ucs, odx, ody);
            const T rR = rot_dir(dcx, dcy, d.sth[pr], d.cth[pr], vce, vcw, ucn, ucs, odx, ody);
            rl = T(0.5) * (r0 + rL);
            rr = T(0.5) * (r0 + rR);
          }
          const T epl = (q + 1 < nd) ? ev[(q + 1) < DMAX ? q + 1 : q] : enext;
          const T epr = (q > 0) ? ev[q > 0 ? q - 1 : 0] : eprev;
          const T fr = T(0.5) *
                       ((rl + fabs(rl)) * e0 + (rl - fabs(rl)) * epl - (rr + fabs(rr)) * epr -
                        (rr - fabs(rr)) * e0) *
                       odth;
          v = v - dtr * fr;
          if (iced) v = 0;
        }
        if (valid) d.e[b] = v;
        const T cwt = cw * ct + sw * st;
        const T ss = ssin_bin(s, logl2, cwt, cp0, ct, st, fkovg, sin_fac, fs1, fs2, o);
        const double es_ = (double)v * (double)ss;
        tx += es_ * (double)ct / cpq;
        ty += es_ * (double)st / cpq;
        hs += (double)v * kdk;
      }
    }
    sx += tx * kdk;
    sy += ty * kdk;
    if (o >= oc) {
      dgx += tx * kdk;
      dgy += ty * kdk;
    }
    if (o == om - 1) {
      tomx = tx;
      tomy = ty;
    }
  }
  // combine the W waves' partial integrals of each point
  double* r = smem;
  r[(0 * W + w) * 64 + lane] = sx;
  r[(1 * W + w) * 64 + lane] = sy;
  r[(2 * W + w) * 64 + lane] = dgx;
  r[(3 * W + w) * 64 + lane] = dgy;
  r[(4 * W + w) * 64 + lane] = tomx;
  r[(5 * W + w) * 64 + lane] = tomy;
  r[(6 * W + w) * 64 + lane] = hs;
  __syncthreads();
  if (w != 0 || !valid) return;
  double acc[7];
#pragma unroll
  for (int c = 0; c < 7; ++c) {
    double a = 0;
    for (int q = 0; q < W; ++q) a += r[(c * W + q) * 64 + lane];
    acc[c] = a;
  }
  StressIn si;
  si.sx = acc[0];
  si.sy = acc[1];
  si.dx = acc[2];
  si.dy = acc[3];
  si.tomx = acc[4];
  si.tomy = acc[5];
  si.wspd = s.wspd;
  si.wdir = d.wdir[j];
  si.uc = s.uc;
  si.vc = s.vc;
  si.us1 = d.us1[j];
  si.vs1 = d.vs1[j];
  si.rhoa = d.rhoa[j];
  si.rhow = d.rhow[j];
  si.kom = d.k[(size_t)(om - 1) * ldi + j];
  store_stress(d, j, stress_atm_point(si, d.dthg, d.z, d.nu_air, d.kappa));
  d.hs_ice[j] = 4 * sqrt(acc[6] * d.dth);
}

template <typename T, int DMAX>
int launch_all(const DomainDev<T>* doms, const BlockTask* tasks, int ntasks, int first,
               int rotation, int W, hipStream_t s, hipEvent_t* ev) {
  const dim3 grid(ntasks), block(64 * W);
  const size_t shA = (size_t)2 * 3 * W * 64 * sizeof(double);
  const size_t shC = (size_t)7 * W * 64 * sizeof(double);
  if (ev) (void)hipEventRecord(ev[0], s);
  hipLaunchKernelGGL((k_rates<T, DMAX>), grid, block, shA, s, doms, tasks, first, rotation);
  if (ev) (void)hipEventRecord(ev[1], s);
  hipLaunchKernelGGL((k_integrate<T, DMAX>), grid, block, shA, s, doms, tasks);
  if (ev) (void)hipEventRecord(ev[2], s);
  hipLaunchKernelGGL((k_advect<T, DMAX>), grid, block, shC, s, doms, tasks);
  if (ev) (void)hipEventRecord(ev[3], s);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail(UMWM_ERR_HIP, std::string("fused launch: ") + hipGetErrorString(e));
  return 0;
}

}  // namespace

template <typename T>
int launch_fused(const DomainDev<T>* doms, const BlockTask* tasks, int ntasks, int first,
                 int rotation, Tile tile, hipStream_t s, hipEvent_t* ev) {
  if (ntasks <= 0) return 0;
  if (tile.W < 1 || tile.W > 8) return fail(UMWM_ERR_ARGUMENT, "tile W must be 1..8");
  switch (tile.D) {
    case 3: return launch_all<T, 3>(doms, tasks, ntasks, first, rotation, tile.W, s, ev);
    case 6: return launch_all<T, 6>(doms, tasks, ntasks, first, rotation, tile.W, s, ev);
    case 9: return launch_all<T, 9>(doms, tasks, ntasks, first, rotation, tile.W, s, ev);
    case 12: return launch_all<T, 12>(doms, tasks, ntasks, first, rotation, tile.W, s, ev);
    case 18: return launch_all<T, 18>(doms, tasks, ntasks, first, rotation, tile.W, s, ev);
  }
  return fail(UMWM_ERR_ARGUMENT, "unsupported tile D");
}

template int launch_fused<float>(const DomainDev<float>*, const BlockTask*, int, int, int, Tile,
                                 hipStream_t, hipEvent_t*);
template int launch_fused<double>(const DomainDev<double>*, const BlockTask*, int, int, int,
                                  Tile, hipStream_t, hipEvent_t*);

}  // namespace umwm
