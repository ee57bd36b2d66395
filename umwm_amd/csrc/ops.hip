// Operator-path kernels: one launch sequence per reference subroutine, with
// the reference's module arrays (ssin, sds, snl, dummy, rotl, rotr, sdt,
// sice) materialised in HBM.  This is the drop-in behind the Fortran
// wrappers that keep the reference's call sequence; the fused path in
// fused.hip computes the same substep without materialising them.
#include <hip/hip_runtime.h>

#include <cfloat>

#include "domain.hpp"
#include "physics.hpp"

namespace umwm {
namespace {

constexpr int kBlock = 256;

inline dim3 grid1(int n) { return dim3((unsigned)((n + kBlock - 1) / kBlock)); }

// ---------------------------------------------------------------------
// src/umwm_forcing.F90:40-99 forcinginterpolate (gustiness = 0)
// ---------------------------------------------------------------------
// One launch interpolates the forcing of every listed domain (one domain, or
// all forced members of a batch): block row y serves domain list[y], whose
// forcing levels are read through its device-resident ForcingDev record, so
// a captured graph follows forcinginput's level swaps.  With `gate` (the
// fused substep) the fields are left alone once sumt >= dtg: a substep
// launched past the end of the dtg leaves the forcing as the last active
// substep interpolated it, as the reference's loop test would.  With forced
// currents the kernel also raises the domain's anycur flag where a current
// is nonzero (zerocurrents, src/umwm_advection.F90:259-265), on the fields
// as they stand after the launch; k_anycur_reset clears the flags first.
template <typename T>
__global__ void k_forcing(const DomainDev<T>* __restrict__ doms,
                          const ForcingDev<T>* __restrict__ fds,
                          const int32_t* __restrict__ list, int gate) {
  const int k = list[blockIdx.y];
  const DomainDev<T>& d = doms[k];
  const ForcingDev<T>& F = fds[k];
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= d.imm) return;
  const double s = d.sc->sumt, dtg = d.dtg;
  if (!gate || s < dtg) {
    const size_t q2 = (size_t)(F.ni[q] - 1) * d.mm + (F.mi[q] - 1);
    const T wb = T(1 - s / dtg), wf = T(s), tg = T(dtg);
    if (F.winds) {
      T uw = F.uwb[q2] * wb + F.uwf[q2] * wf / tg;
      T vw = F.vwb[q2] * wb + F.vwf[q2] * wf / tg;
      d.wspd[q] = sqrt(uw * uw + vw * vw);
      d.wdir[q] = atan2(vw, uw);
    }
    if (F.seaice) d.fice[q] = F.ficeb[q2] * wb + F.ficef[q2] * wf / tg;
    if (F.currents) {
      d.uc[q] = F.ucb[q2] * wb + F.ucf[q2] * wf / tg;
      d.vc[q] = F.vcb[q2] * wb + F.vcf[q2] * wf / tg;
    }
    if (F.air_density) d.rhoa[q] = F.rhoab[q] * wb + F.rhoaf[q] * wf / tg;
    if (F.water_density) d.rhow[q] = F.rhowb[q] * wb + F.rhowf[q] * wf / tg;
    d.rhorat[q] = d.rhoa[q] / d.rhow[q];
  }
  if (F.currents) {
    const int n = d.isglobal ? d.imm : d.im;
    if (q < n && (d.uc[q] != T(0) || d.vc[q] != T(0))) atomicOr(&d.sc->anycur, 1);
  }
}

// Clears anycur of the listed domains whose currents are forced, ahead of
// k_forcing.
template <typename T>
__global__ void k_anycur_reset(const DomainDev<T>* __restrict__ doms,
                               const ForcingDev<T>* __restrict__ fds,
                               const int32_t* __restrict__ list, int n) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n && fds[list[t]].currents) doms[list[t]].sc->anycur = 0;
}

// ---------------------------------------------------------------------
// src/umwm_source_functions.F90:16-82 sin_d12
// ---------------------------------------------------------------------
template <typename T>
__global__ void k_sin_prologue(DomainDev<T> d) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= d.imm) return;
  T w = d.wspd[q];
  w = w > T(1e-2) ? w : T(1e-2);
  d.wspd[q] = w;
  if (q < d.im) {
    T fc = T(0.53) * T(d.g) / w;
    if (fc > T(d.fprog)) fc = T(d.fprog);
    d.fcutoff[q] = fc;
    d.oc[q] = cutoff_bin(fc, d.f, d.om);
    d.shelt[q] = sheltering_coare35(w);
  }
}

template <typename T>
__device__ inline SinPoint<T> sin_point(const DomainDev<T>& d, int j) {
  SinPoint<T> s;
  s.wspd = d.wspd[j];
  s.a = T(2.5) * T(d.ustar[j]);
  s.pre = (T(1) - d.fice[j]) * T(kTwoPiD) * d.rhorat[j];
  s.shelt = d.shelt[j];
  s.uc = d.uc[j];
  s.vc = d.vc[j];
  s.oc = d.oc[j];
  return s;
}

template <typename T>
__global__ void k_sin_bins(DomainDev<T> d) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  const int op = blockIdx.y, o = op / d.pm, p = op % d.pm;
  if (j >= d.im) return;
  const SinPoint<T> s = sin_point(d, j);
  const size_t t = (size_t)o * d.ldi + j;
  const T cwt = cos(d.wdir[j] - d.th[p]);
  d.ssin[((size_t)o * d.pm + p) * d.ldi + j] =
      ssin_bin(s, d.logl2overz[t], cwt, d.cp0[t], d.cth[p], d.sth[p], d.fkovg[t], T(d.sin_fac),
               T(d.fieldscale1), T(d.fieldscale2), o);
}

// ---------------------------------------------------------------------
// src/umwm_source_functions.F90:85-105 sds_d12, directional convolution
// in the reference's order
// ---------------------------------------------------------------------
template <typename T>
__global__ void k_sds(DomainDev<T> d) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  const int p = blockIdx.y;
  if (j >= d.im) return;
  const int pm = d.pm;
  const size_t ldi = d.ldi;
  T raw = 0;
  for (int o = 0; o < d.om; ++o) {
    if (o > 0 && d.mss_fac > 0) {
      T acc = 0;
      const T* eo = d.e + (size_t)(o - 1) * pm * ldi + j;
      const T* w = d.cth2pp + (size_t)p * pm;
      for (int pp = 0; pp < pm; ++pp) acc += eo[pp * ldi] * w[pp];
      raw = raw + acc * d.k3dk[(size_t)(o - 1) * ldi + j];
    }
    T dm = T(1) + T(d.mss_fac) * raw;
    dm = dm * dm;
    const size_t b = ((size_t)o * pm + p) * ldi + j;
    d.dummy[b] = dm;
    d.sds[b] = T(d.twopisds_fac) * d.f[o] * dm *
               pw(d.e[b] * d.k4[(size_t)o * ldi + j], T(d.sds_power));
  }
}

// ---------------------------------------------------------------------
// src/umwm_source_functions.F90:156-183 snl_d12
// ---------------------------------------------------------------------
template <typename T>
__global__ void k_snl(DomainDev<T> d) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  const int op = blockIdx.y, o = op / d.pm, p = op % d.pm;
  if (j >= d.im) return;
  const size_t ldi = d.ldi, pl = (size_t)d.pm * ldi;
  const size_t b = ((size_t)o * d.pm + p) * ldi + j;
  T v = 0;
  if (o <= d.oc[j]) {
    const size_t t = (size_t)o * ldi + j;
    v = d.bf1r[t] * d.sds[b + pl] * d.e[b + pl] + d.bf2r[t] * d.sds[b + 2 * pl] * d.e[b + 2 * pl] -
        T(d.snl_fac) * d.sds[b] * d.e[b];
  }
  d.snl[b] = v;
}

template <typename T>
__global__ void k_snl_post(DomainDev<T> d) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  const int op = blockIdx.y, o = op / d.pm, p = op % d.pm;
  if (j >= d.im) return;
  const size_t t = (size_t)o * d.ldi + j;
  const size_t b = ((size_t)o * d.pm + p) * d.ldi + j;
  d.sds[b] = d.sds[b] * d.cothkd[t];
  if (p == 0) d.sdt[t] = T(d.sdt_fac) * sqrt(d.rhorat[j]) * T(d.ustar[j]) * d.k[t];
}

// ---------------------------------------------------------------------
// src/umwm_source_functions.F90:108-153 s_ice
// ---------------------------------------------------------------------
template <typename T>
__global__ void k_sice(DomainDev<T> d) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= d.im) return;
  const size_t ldi = d.ldi;
  if (d.fice[j] > T(d.fice_lth)) {
    double ht = 0;
    for (int p = 0; p < d.pm; ++p)
      for (int o = 0; o < d.om; ++o)
        ht += (double)(d.e[((size_t)o * d.pm + p) * ldi + j] * d.kdk[(size_t)o * ldi + j]);
    ht = 4 * sqrt(ht * d.dth);
    const double c = sice_coef(ht);
    for (int o = 0; o < d.om; ++o) d.sice[(size_t)o * ldi + j] = T(2 * (double)d.cg0[(size_t)o * ldi + j] * c);
  } else {
    for (int o = 0; o < d.om; ++o) d.sice[(size_t)o * ldi + j] = 0;
  }
}

// ---------------------------------------------------------------------
// src/umwm_physics.F90:9-86 source
// ---------------------------------------------------------------------
template <typename T>
__device__ inline void last_block_dts(const DomainDev<T>& d, double bmin, int first) {
  __shared__ double red[kBlock / 64];
  __shared__ int is_last;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  bmin = wave_min(bmin);
  if (lane == 0) red[w] = bmin;
  __syncthreads();
  // grid min through an atomic max of ~okey (as in the fused rates kernel:
  // no per-block partials, so no L2-flushing release fence)
  if (threadIdx.x == 0) {
    double m = red[0];
    for (int q = 1; q < (int)(blockDim.x >> 6); ++q) m = fmin(m, red[q]);
    atomicMax(&d.sc->pts_key, ~okey(m));
    __builtin_amdgcn_s_waitcnt(0);  // the atomic completes before the arrival count
    unsigned t = atomicAdd(&d.sc->counter, 1u);
    is_last = (t == gridDim.x - 1);
  }
  __syncthreads();
  if (!is_last) return;
  if (threadIdx.x == 0) {
    const unsigned long long k = atomicExch(&d.sc->pts_key, 0ull);
    const double mn = k ? okey_inv(~k) : INFINITY;
    Scalars* sc = d.sc;
    double dts = fmin(fmin(mn, d.dtamin), d.dtg - sc->sumt);
    if (!d.restart && first) dts = 0;
    sc->dts = dts;
    sc->dta = dts;
    sc->sumt = sc->sumt + dts;
    sc->counter = 0;
  }
}

template <typename T>
__global__ void k_src_rate(DomainDev<T> d, int first) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  double pts = INFINITY;
  if (j < d.im) {
    const size_t ldi = d.ldi;
    const int oc = d.oc[j];
    T mx = 0;
    for (int p = 0; p < d.pm; ++p)
      for (int o = 0; o < d.om; ++o) {
        const size_t b = ((size_t)o * d.pm + p) * ldi + j, t = (size_t)o * ldi + j;
        T r = 0;
        if (o <= oc)
          r = d.ssin[b] - d.sds[b] * d.snl_arg[t] - d.sbf[t] - d.sdt[t] - d.sdv[t] + d.sice[t];
        d.ef[b] = r;
        T a = fabs(r);
        mx = a > mx ? a : mx;
      }
    pts = d.explim / (double)mx;
    d.pts[j] = pts;
  }
  last_block_dts(d, pts, first);
}

template <typename T>
__global__ void k_src_int(DomainDev<T> d) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  const int op = blockIdx.y, o = op / d.pm, p = op % d.pm;
  if (j >= d.im) return;
  const T dts = T(d.sc->dts);
  const size_t t = (size_t)o * d.ldi + j;
  const size_t b = ((size_t)o * d.pm + p) * d.ldi + j;
  T ef = d.ef[b];
  if (o <= d.oc[j]) {
    ef = d.e[b] * ex(dts * (d.ssin[b] - d.sds[b] - d.sbf[t] - d.sdt[t] - d.sdv[t] + d.sice[t])) +
         dts * d.snl[b];
  } else {
    const T num = d.ssin[b] - d.sdt[t] - d.sdv[t] + d.sice[t];
    if (num >= T(0))
      ef = d.oneoverk4[t] *
           pw(num / (T(d.twopisds_fac) * d.f[o] * d.dummy[b] * d.cothkd[t]), T(d.inv_sds_power));
  }
  d.ef[b] = ef;
  d.e[b] = T(0.5) * (d.e[b] + ef);
}

// ---------------------------------------------------------------------
// src/umwm_advection.F90:18-99 propagation
// ---------------------------------------------------------------------
template <typename T>
__global__ void k_propagation(DomainDev<T> d) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  const int op = blockIdx.y, o = op / d.pm, p = op % d.pm;
  if (j >= d.im || o > d.oc[j]) return;
  const size_t ldi = d.ldi;
  const int ie = d.ie[j], iw = d.iw[j], is = d.is_[j], in = d.in_[j];
  const T* cg = d.cg0 + (size_t)o * ldi;
  const T cc = d.cth_curv[(size_t)p * ldi + j], sc = d.sth_curv[(size_t)p * ldi + j];
  const T cge = cg[j] * cc + cg[ie] * cc;
  const T cgw = cg[j] * cc + cg[iw] * cc;
  const T cgs = cg[j] * sc + cg[is] * sc;
  const T cgn = cg[j] * sc + cg[in] * sc;
  const T* E = d.e + ((size_t)o * d.pm + p) * ldi;
  const T e0 = E[j], ee = E[ie], ew = E[iw], en = E[in], es = E[is];
  const T dye = d.dye[j], dyw = d.dyw[j], dxn = d.dxn[j], dxs = d.dxs[j];
  T flux = (cge + fabs(cge)) * dye * e0 + (cge - fabs(cge)) * dye * ee -
           (cgw + fabs(cgw)) * dyw * ew - (cgw - fabs(cgw)) * dyw * e0 +
           (cgn + fabs(cgn)) * dxn * e0 + (cgn - fabs(cgn)) * dxn * en -
           (cgs + fabs(cgs)) * dxs * es - (cgs - fabs(cgs)) * dxs * e0;
  if (d.sc->anycur) {
    const T ue = d.uc[j] + d.uc[d.iie[j]], uw = d.uc[j] + d.uc[d.iiw[j]];
    const T vn = d.vc[j] + d.vc[d.iin[j]], vs = d.vc[j] + d.vc[d.iis[j]];
    const T feup = (ue + fabs(ue)) * dye, fedn = (ue - fabs(ue)) * dye;
    const T fwup = (uw + fabs(uw)) * dyw, fwdn = (uw - fabs(uw)) * dyw;
    const T fnup = (vn + fabs(vn)) * dxn, fndn = (vn - fabs(vn)) * dxn;
    const T fsup = (vs + fabs(vs)) * dxs, fsdn = (vs - fabs(vs)) * dxs;
    flux = flux + (feup * e0 + fedn * ee - fwup * ew - fwdn * e0 + fnup * e0 + fndn * en -
                   fsup * es - fsdn * e0);
  }
  const size_t b = ((size_t)o * d.pm + p) * ldi + j;
  T v = d.ef[b] - T(0.25) * T(d.sc->dta) * flux * d.oneovar[j];
  if (d.fice[j] > T(d.fice_uth)) v = 0;
  d.ef[b] = v;
}

// ---------------------------------------------------------------------
// src/umwm_advection.F90:102-180 refraction
// ---------------------------------------------------------------------
template <typename T>
__device__ inline T rotation(const DomainDev<T>& d, int o, int p, int j) {
  const size_t ldi = d.ldi;
  const T* cp = d.cp0 + (size_t)o * ldi;
  return T(0.5) * (((cp[d.ie[j]] - cp[d.iw[j]]) * d.sth[p] + d.vc[d.iie[j]] - d.vc[d.iiw[j]]) *
                       d.oneovdx[j] -
                   ((cp[d.in_[j]] - cp[d.is_[j]]) * d.cth[p] + d.uc[d.iin[j]] - d.uc[d.iis[j]]) *
                       d.oneovdy[j]);
}

template <typename T>
__global__ void k_rotation(DomainDev<T> d) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  double mx = -INFINITY;
  if (j < d.im) {
    const int oc = d.oc[j];
    d.oc_rot[j] = oc;
    for (int o = 0; o < d.om; ++o)
      for (int p = 0; p < d.pm; ++p) {
        const size_t b = ((size_t)o * d.pm + p) * d.ldi + j;
        T rl = 0, rr = 0;
        if (o <= oc) {
          const T r0 = rotation(d, o, p, j);
          rl = T(0.5) * (r0 + rotation(d, o, d.pl[p], j));
          rr = T(0.5) * (r0 + rotation(d, o, d.pr[p], j));
          mx = fmax(mx, (double)r0);
        } else {
          mx = fmax(mx, 0.0);
        }
        d.rotl[b] = rl;
        d.rotr[b] = rr;
      }
  }
  // block max -> last block sets dtr_temp (signed maxval, :352-361)
  __shared__ double red[kBlock / 64];
  __shared__ int is_last;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  mx = wave_max(mx);
  if (lane == 0) red[w] = mx;
  __syncthreads();
  if (threadIdx.x == 0) {
    double m = red[0];
    for (int q = 1; q < (int)(blockDim.x >> 6); ++q) m = fmax(m, red[q]);
    atomicMax(&d.sc->rot_key, okey(m));
    __builtin_amdgcn_s_waitcnt(0);  // the atomic completes before the arrival count
    unsigned t = atomicAdd(&d.sc->counter_rot, 1u);
    is_last = (t == gridDim.x - 1);
  }
  __syncthreads();
  if (!is_last || threadIdx.x != 0) return;
  const unsigned long long k = atomicExch(&d.sc->rot_key, 0ull);
  const double m = k ? okey_inv(k) : -INFINITY;
  Scalars* sc = d.sc;
  sc->maxrot = m;
  sc->dtr_temp = (m <= d.tiny) ? sc->dts : d.dth / m;
  sc->counter_rot = 0;
}

template <typename T>
__global__ void k_refraction(DomainDev<T> d) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  const int op = blockIdx.y, o = op / d.pm, p = op % d.pm;
  const double dtr = fmin(d.sc->dtr_temp, d.sc->dts);
  if (j == 0 && op == 0) d.sc->dtr = dtr;
  if (j >= d.im || o > d.oc[j]) return;
  const size_t ldi = d.ldi;
  const size_t b = ((size_t)o * d.pm + p) * ldi + j;
  const T* E = d.e + (size_t)o * d.pm * ldi + j;
  const T rl = d.rotl[b], rr = d.rotr[b];
  const T ep = E[(size_t)p * ldi], el = E[(size_t)d.pl[p] * ldi], er = E[(size_t)d.pr[p] * ldi];
  const T flux = T(0.5) * ((rl + fabs(rl)) * ep + (rl - fabs(rl)) * el - (rr + fabs(rr)) * er -
                           (rr - fabs(rr)) * ep) *
                 T(d.oneovdth);
  T v = d.ef[b] - T(dtr) * flux;
  if (d.fice[j] > T(d.fice_uth)) v = 0;
  d.ef[b] = v;
}

template <typename T>
__global__ void k_update(DomainDev<T> d) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  const int op = blockIdx.y;
  if (j >= d.im) return;
  const size_t b = (size_t)op * d.ldi + j;
  d.e[b] = d.ef[b];
}

// ---------------------------------------------------------------------
// src/umwm_stress.F90:216-264 stress('atm') from materialised ssin
// ---------------------------------------------------------------------
template <typename T>
__global__ void k_stress_atm(DomainDev<T> d) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= d.im) return;
  const size_t ldi = d.ldi;
  const int om = d.om, oc = d.oc[j];
  // stress_vector_integral divides by cp0(o, i-1) (src/umwm_stress.F90:219,316)
  const int jc = (j == 0) ? d.im : j - 1;
  StressIn s;
  s.sx = s.sy = s.dx = s.dy = s.tomx = s.tomy = 0;
  for (int o = 0; o < om; ++o) {
    const double cp = d.cp0[(size_t)o * ldi + jc];
    double tx = 0, ty = 0;
    for (int p = 0; p < d.pm; ++p) {
      const size_t b = ((size_t)o * d.pm + p) * ldi + j;
      const double es = (double)d.e[b] * (double)d.ssin[b];
      tx = tx + es * (double)d.cth[p] / cp;
      ty = ty + es * (double)d.sth[p] / cp;
    }
    const double kdk = d.kdk[(size_t)o * ldi + j];
    s.sx += tx * kdk;
    s.sy += ty * kdk;
    if (o >= oc) {
      s.dx += tx * kdk;
      s.dy += ty * kdk;
    }
    if (o == om - 1) {
      s.tomx = tx;
      s.tomy = ty;
    }
  }
  s.wspd = d.wspd[j];
  s.wdir = d.wdir[j];
  s.uc = d.uc[j];
  s.vc = d.vc[j];
  s.us1 = d.us1[j];
  s.vs1 = d.vs1[j];
  s.rhoa = d.rhoa[j];
  s.rhow = d.rhow[j];
  s.kom = d.k[(size_t)(om - 1) * ldi + j];
  store_stress(d, j, stress_atm_point(s, d.dthg, d.z, d.nu_air, d.kappa));
}

// ---------------------------------------------------------------------
// src/umwm_stress.F90:34-214 stress('ocn'): momentum into the ocean top
// and bottom, Snl momentum loss, energy fluxes into the ocean and from the
// air, and the form-drag partitions, from the current e and the source
// arrays (ssin, sds after cothkd, snl) of the last source evaluation.
// FP64 sums; one thread per sea point.
// ---------------------------------------------------------------------
template <typename T>
__global__ void k_stress_ocn(DomainDev<T> d, OcnDev out, double bf1, double bf2) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= d.im) return;
  const size_t ldi = d.ldi;
  const int om = d.om, pm = d.pm, oc = d.oc[j];
  const double rw = d.rhow[j], dthg = d.dthg;
  auto tab = [&](const T* a, int o) { return (double)a[(size_t)o * ldi + j]; };
  auto spec = [&](const T* a, int o, int p) { return (double)a[((size_t)o * pm + p) * ldi + j]; };
  double top_x = 0, top_y = 0, bot_x = 0, bot_y = 0, snl_x = 0, snl_y = 0;
  double eo_x = 0, eo_y = 0, ea_x = 0, ea_y = 0, tomx = 0, tomy = 0;
  double t1x = 0, t1y = 0, t2x = 0, t2y = 0, t3x = 0, t3y = 0;
  const double cwd = cos((double)d.wdir[j]), swd = sin((double)d.wdir[j]);
  for (int o = 0; o < om; ++o) {
    const double icp = tab(d.invcp0, o), cp = tab(d.cp0, o), kdk = tab(d.kdk, o);
    const double sdt = tab(d.sdt, o), sdv = tab(d.sdv, o), sbf = tab(d.sbf, o);
    // Snl weights of bins o-1, o-2 (o >= 2, 0-based)
    const double wtop = o >= 2 ? bf1 * cp * tab(d.invcp0, o - 1) + bf2 * cp * tab(d.invcp0, o - 2) : 0;
    const double wsnl =
        o >= 2 ? bf1 * (1 - cp * tab(d.invcp0, o - 1)) + bf2 * (1 - cp * tab(d.invcp0, o - 2)) : 0;
    double ux = 0, uy = 0, bx = 0, by = 0, nx = 0, ny = 0, qx = 0, qy = 0, ax = 0, ay = 0;
    for (int p = 0; p < pm; ++p) {
      const double ct = d.cth[p], st = d.sth[p];
      const double e = spec(d.e, o, p), sds = spec(d.sds, o, p), snl = spec(d.snl, o, p);
      const double ssin = spec(d.ssin, o, p);
      const double v = e * (sds + sdt + sdv);
      ux = ux + v * ct * icp;
      uy = uy + v * st * icp;
      if (o >= 2) {
        ux = ux - snl * wtop * ct * icp;
        uy = uy - snl * wtop * st * icp;
      }
      bx = bx + e * sbf * ct * icp;
      by = by + e * sbf * st * icp;
      if (o >= 2 && o <= oc) {
        nx = nx + snl * wsnl * ct * icp;
        ny = ny + snl * wsnl * st * icp;
      }
      qx = qx + e * sds * ct;
      qy = qy + e * sds * st;
      if (o >= 2) {
        qx = qx - snl * wtop * ct;
        qy = qy - snl * wtop * st;
      }
      ax = ax + e * ssin * ct;
      ay = ay + e * ssin * st;
      const double dm = e * ssin * icp * kdk;
      if (ssin > 0) {
        t1x += dm * ct;
        t1y += dm * st;
      } else if (cwd * ct + swd * st < 0) {  // cos(wdir - th) < 0
        t2x += dm * ct;
        t2y += dm * st;
      } else {
        t3x += dm * ct;
        t3y += dm * st;
      }
    }
    top_x += ux * kdk;
    top_y += uy * kdk;
    bot_x += bx * kdk;
    bot_y += by * kdk;
    snl_x += nx * kdk;
    snl_y += ny * kdk;
    eo_x += qx * kdk;
    eo_y += qy * kdk;
    ea_x += ax * kdk;
    ea_y += ay * kdk;
    if (o == om - 1) {
      tomx = ux;
      tomy = uy;
    }
  }
  double w = d.wspd[j];
  w = w > 1e-2 ? w : 1e-2;
  const double kom = tab(d.k, om - 1);
  const double tail = stress_tail(w, kom);
  const double tlx = tomx * kom * tail * rw * dthg, tly = tomy * kom * tail * rw * dthg;
  const double vals[kOcnFields] = {
      tlx, tly,
      top_x * rw * dthg + tlx + d.taux_skin[j], top_y * rw * dthg + tly + d.tauy_skin[j],
      bot_x * rw * dthg, bot_y * rw * dthg, snl_x * rw * dthg, snl_y * rw * dthg,
      eo_x * rw * dthg, eo_y * rw * dthg, ea_x * rw * dthg, ea_y * rw * dthg,
      t1x * rw * dthg, t1y * rw * dthg, t2x * rw * dthg, t2y * rw * dthg,
      t3x * rw * dthg, t3y * rw * dthg};
  for (int q = 0; q < kOcnFields; ++q) out.f[q][j] = vals[q];
}

// ---------------------------------------------------------------------
// src/umwm_physics.F90:89-205 diag
// ---------------------------------------------------------------------
// One pass over the point's spectrum, frequency by frequency: the tables of
// a frequency are read once and the directions of its spectrum rows (one
// coalesced row per (o, p) across the lanes) are summed in FP64.  The
// dominant bin is the reference's maxloc over e kdk (o fastest,
// src/umwm_physics.F90:157-160): of equal maxima the one first in (p, o)
// order.
template <typename T>
__global__ void k_diag(DomainDev<T> d, DiagDev g) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= d.im) return;
  const size_t ldi = d.ldi;
  const int om = d.om, pm = d.pm, oc = d.oc[j];
  double m0 = 0, m2 = 0, mk = 0, mx = 0, my = 0, gxx = 0, gxy = 0, gyy = 0, xc = 0, yc = 0;
  double best = -INFINITY;
  int opk = 0, ppk = 0;
  for (int o = 0; o < om; ++o) {
    const size_t t = (size_t)o * ldi + j;
    const double kdk = d.kdk[t], f = d.f[o], k3dk = d.k3dk[t];
    const double icp = d.invcp0[t], cg = d.cg0[t];
    const bool prog = o <= oc;
    const T* eo = d.e + (size_t)o * pm * ldi + j;
    double s0 = 0, sk = 0, sx = 0, sy = 0, sxx = 0, sxy = 0, syy = 0;
    for (int p = 0; p < pm; ++p) {
      const double c = d.cth[p], s = d.sth[p];
      const double e = eo[(size_t)p * ldi];
      const double ek = e * kdk;
      s0 += ek;
      sk += e * k3dk;
      sx += ek * c;
      sy += ek * s;
      if (prog) {
        sxx += ek * (c * c);
        sxy += ek * c * s;
        syy += ek * (s * s);
      }
      if (ek > best || (ek == best && (p < ppk || (p == ppk && o < opk)))) {
        best = ek;
        opk = o;
        ppk = p;
      }
    }
    m0 += s0;
    m2 += (f * f) * s0;
    mk += sk;
    xc += sx;
    yc += sy;
    if (prog) {
      mx += sx * icp;
      my += sy * icp;
      gxx += cg * icp * sxx;
      gxy += cg * icp * sxy;
      gyy += cg * icp * syy;
    }
  }
  const double tiny = 2.2250738585072014e-308, rw = d.rhow[j];
  g.mwp[j] = sqrt(m0 / (m2 + tiny));
  g.momx[j] = mx * rw * d.dthg;
  g.momy[j] = my * rw * d.dthg;
  g.cgmxx[j] = gxx * rw * d.dthg;
  g.cgmxy[j] = gxy * rw * d.dthg;
  g.cgmyy[j] = gyy * rw * d.dthg;
  g.ht[j] = 4 * sqrt(m0 * d.dth);
  g.mwd[j] = atan2(yc, xc);
  g.mss[j] = mk * d.dth;
  g.mwl[j] = kTwoPiD * sqrt(m0 / (mk + tiny));
  const size_t tp = (size_t)opk * ldi + j;
  g.dwd[j] = d.th[ppk];
  g.dwl[j] = kTwoPiD / (double)d.k[tp];
  g.dwp[j] = 1. / (double)d.f[opk];
  g.dcp0[j] = d.cp0[tp];
  g.dcg0[j] = d.cg0[tp];
  const double cur = (double)d.uc[j] * d.cth[ppk] + (double)d.vc[j] * d.sth[ppk];
  g.dcp[j] = g.dcp0[j] + cur;
  g.dcg[j] = g.dcg0[j] + cur;
}

// Significant wave height of E^n for the fused path's s_ice.
template <typename T>
__global__ void k_hs_ice(DomainDev<T> d) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= d.im) return;
  // four independent partial sums (o mod 4): one serial chain of FP64
  // adds per point made the kernel latency-bound
  double h4[4] = {0, 0, 0, 0};
  auto term = [&](int o, int p) {
    return (double)d.e[((size_t)o * d.pm + p) * d.ldi + j] * (double)d.kdk[(size_t)o * d.ldi + j];
  };
  for (int p = 0; p < d.pm; ++p) {
    int o = 0;
    for (; o + 3 < d.om; o += 4) {
#pragma unroll
      for (int r = 0; r < 4; ++r) h4[r] += term(o + r, p);
    }
    for (; o < d.om; ++o) h4[0] += term(o, p);
  }
  const double h = (h4[0] + h4[1]) + (h4[2] + h4[3]);
  d.hs_ice[j] = 4 * sqrt(h * d.dth);
}

// ---------------------------------------------------------------------
// src/umwm_stokes.F90:13-132 stokes_drift
// ---------------------------------------------------------------------
// us(l), vs(l) = sum over o of util(o,l) sum over p of e cos th, e sin th
// (src/umwm_stokes.F90:59-105) in one pass over the spectrum: per
// frequency the point's directional sums are formed once and added into
// the lm level accumulators it keeps in LDS (in o order at every level),
// which are written once at the end.
template <typename T>
__global__ __launch_bounds__(64) void k_stokes(DomainDev<T> d, StokesDev S) {
  extern __shared__ double acc[];  // [lm][2][points of the block]
  const int t = threadIdx.x, np = blockDim.x;
  const int j = blockIdx.x * np + t;
  const int lm = d.lm;
  double* au = acc + t;
  for (int l = 0; l < lm; ++l) {
    au[(2 * l) * np] = 0;
    au[(2 * l + 1) * np] = 0;
  }
  if (j >= d.im) return;  // no barrier follows: each lane owns its column
  const size_t ldi = d.ldi;
  const double dj = d.depth[j];
  for (int o = 0; o < d.om; ++o) {
    // four independent partial sums over the directions: one serial chain
    // of FP64 adds per point made the kernel latency-bound
    double a4[4] = {0, 0, 0, 0}, b4[4] = {0, 0, 0, 0};
    const T* eo = d.e + (size_t)o * d.pm * ldi + j;
    int p = 0;
    for (; p + 3 < d.pm; p += 4) {
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const double e = eo[(size_t)(p + r) * ldi];
        a4[r] += e * (double)d.cth[p + r];
        b4[r] += e * (double)d.sth[p + r];
      }
    }
    for (; p < d.pm; ++p) {
      const double e = eo[(size_t)p * ldi];
      a4[0] += e * (double)d.cth[p];
      b4[0] += e * (double)d.sth[p];
    }
    const double a = (a4[0] + a4[1]) + (a4[2] + a4[3]);
    const double b = (b4[0] + b4[1]) + (b4[2] + b4[3]);
    const size_t tq = (size_t)o * ldi + j;
    const double k = d.k[tq], kd = k * dj, dwn = d.dwn[tq];
    const double fo = kTwoPiD * (double)d.f[o];
    const double sh = sinh(kd);
    for (int l = 0; l < lm; ++l) {
      const double dep = S.depth[l];
      if (fabs(dep) > dj) continue;  // level below the sea bed: zero drift
      const double arg = 2 * k * (dep + dj);
      double ut;
      if (fabs(arg) > 50 || kd > 50) {
        ut = fo * 2 * (k * k) * exp(2 * k * dep) * dwn * d.dth;
      } else {
        ut = fo * (k * k) * cosh(2 * k * (dep + dj)) / (sh * sh) * dwn * d.dth;
      }
      au[(2 * l) * np] += ut * a;
      au[(2 * l + 1) * np] += ut * b;
    }
  }
  for (int l = 0; l < lm; ++l) {
    const double u = au[(2 * l) * np], v = au[(2 * l + 1) * np];
    S.us[(size_t)l * ldi + j] = u;
    S.vs[(size_t)l * ldi + j] = v;
    if (l == 0) {
      d.us1[j] = u;
      d.vs1[j] = v;
    }
  }
}

template <typename T>
__global__ void k_stokes_efold(DomainDev<T> d, StokesDev S) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= d.im) return;
  const size_t ldi = d.ldi;
  auto mag = [&](int l) {
    const double u = S.us[(size_t)l * ldi + j], v = S.vs[(size_t)l * ldi + j];
    return sqrt(u * u + v * v);
  };
  double ds = 0;
  const double m1 = mag(0);
  if (m1 != 0) {
    const double ef = m1 * (1 / exp(1.0));
    for (int l = 1; l < d.lm; ++l) {
      const double ma = mag(l - 1), mb = mag(l);
      if (mb < ef) {
        ds = (fabs(ma - ef) * S.depth[l] + fabs(mb - ef) * S.depth[l - 1]) / (ma - mb);
        break;
      }
    }
  }
  S.ds[j] = -ds;
}

// ---------------------------------------------------------------------
// JONSWAP x cos^2 initial spectrum (build-defined initializer)
// ---------------------------------------------------------------------
template <typename T>
__global__ void k_jonswap(DomainDev<T> d, umwm_jonswap js, double dlnf) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  const int op = blockIdx.y, o = op / d.pm, p = op % d.pm;
  if (j >= d.im) return;
  const double pi = 3.141592653589793, g = d.g;
  double fp = js.fp;
  if (js.fp_from_wind) {
    fp = 0.13 * g / fmax((double)d.wspd[j], 3.0);
    fp = fmin(fmax(fp, 0.05), 0.5);
  }
  const double thw = js.dir_from_wind ? (double)d.wdir[j] : js.thw;
  const double fo = d.f[o];
  const double sig = fo <= fp ? js.sigma_a : js.sigma_b;
  const double r = exp(-(fo - fp) * (fo - fp) / (2 * sig * sig * fp * fp));
  const double S = js.alpha * g * g * pow(2 * pi, -4.) * pow(fo, -5.) *
                   exp(-1.25 * pow(fp / fo, 4.)) * pow(js.gamma, r);
  const double c = cos((double)d.th[p] - thw);
  const double D = c > 0 ? (2 / pi) * c * c : 0;
  d.e[((size_t)o * d.pm + p) * d.ldi + j] =
      T(S * fo * dlnf * D / (double)d.kdk[(size_t)o * d.ldi + j]);
}

}  // namespace

// =====================================================================
// launchers
// =====================================================================
#define LAUNCH(kernel, grid, block, ...)                                     \
  do {                                                                       \
    hipLaunchKernelGGL(kernel, grid, block, 0, D.stream, __VA_ARGS__);       \
    hipError_t e_ = hipGetLastError();                                       \
    if (e_ != hipSuccess)                                                    \
      return fail(UMWM_ERR_HIP, std::string(#kernel) + ": " + hipGetErrorString(e_)); \
  } while (0)

template <typename T>
int launch_forcing_list(const DomainDev<T>* doms, const ForcingDev<T>* fds, const int32_t* list,
                        int n, int maximm, bool currents, bool gate, hipStream_t s) {
  if (n <= 0) return 0;
  if (currents)
    hipLaunchKernelGGL(k_anycur_reset<T>, dim3((n + 63) / 64), dim3(64), 0, s, doms, fds, list, n);
  hipLaunchKernelGGL(k_forcing<T>, dim3(grid1(maximm).x, n), dim3(kBlock), 0, s, doms, fds, list,
                     gate ? 1 : 0);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail(UMWM_ERR_HIP, std::string("k_forcing: ") + hipGetErrorString(e));
  return 0;
}

template <typename T>
int launch_forcinginterpolate(DomainImpl<T>& D, bool gate) {
  return launch_forcing_list<T>(D.d, D.fd, D.list0, 1, D.h.imm, D.S.P.currents != 0, gate,
                                D.stream);
}

template <typename T>
int launch_op_sin(DomainImpl<T>& D) {
  LAUNCH(k_sin_prologue<T>, grid1(D.h.imm), dim3(kBlock), D.h);
  LAUNCH(k_sin_bins<T>, dim3(grid1(D.h.im).x, D.h.om * D.h.pm), dim3(kBlock), D.h);
  return 0;
}

template <typename T>
int launch_op_sds(DomainImpl<T>& D) {
  LAUNCH(k_sds<T>, dim3(grid1(D.h.im).x, D.h.pm), dim3(kBlock), D.h);
  return 0;
}

template <typename T>
int launch_op_snl(DomainImpl<T>& D) {
  dim3 g(grid1(D.h.im).x, D.h.om * D.h.pm);
  LAUNCH(k_snl<T>, g, dim3(kBlock), D.h);
  LAUNCH(k_snl_post<T>, g, dim3(kBlock), D.h);
  return 0;
}

template <typename T>
int launch_op_sice(DomainImpl<T>& D) {
  LAUNCH(k_sice<T>, grid1(D.h.im), dim3(kBlock), D.h);
  return 0;
}

template <typename T>
int launch_op_source(DomainImpl<T>& D) {
  LAUNCH(k_src_rate<T>, grid1(D.h.im), dim3(kBlock), D.h, (int)D.first);
  LAUNCH(k_src_int<T>, dim3(grid1(D.h.im).x, D.h.om * D.h.pm), dim3(kBlock), D.h);
  return 0;
}

template <typename T>
int launch_op_propagation(DomainImpl<T>& D) {
  LAUNCH(k_propagation<T>, dim3(grid1(D.h.im).x, D.h.om * D.h.pm), dim3(kBlock), D.h);
  return 0;
}

template <typename T>
int launch_op_refraction(DomainImpl<T>& D) {
  if (D.S.P.currents || D.first || !D.rot_valid) {
    LAUNCH(k_rotation<T>, grid1(D.h.im), dim3(kBlock), D.h);
    D.rot_valid = true;
  }
  LAUNCH(k_refraction<T>, dim3(grid1(D.h.im).x, D.h.om * D.h.pm), dim3(kBlock), D.h);
  return 0;
}

template <typename T>
int launch_op_update(DomainImpl<T>& D) {
  LAUNCH(k_update<T>, dim3(grid1(D.h.im).x, D.h.om * D.h.pm), dim3(kBlock), D.h);
  return 0;
}

template <typename T>
int launch_op_stress_atm(DomainImpl<T>& D) {
  LAUNCH(k_stress_atm<T>, grid1(D.h.im), dim3(kBlock), D.h);
  return 0;
}

template <typename T>
int launch_op_stress_ocn(DomainImpl<T>& D) {
  LAUNCH(k_stress_ocn<T>, grid1(D.h.im), dim3(kBlock), D.h, D.ocn, D.S.bf1, D.S.bf2);
  return 0;
}

template <typename T>
int launch_diag(DomainImpl<T>& D) {
  LAUNCH(k_diag<T>, grid1(D.h.im), dim3(kBlock), D.h, D.diag);
  return 0;
}

template <typename T>
int launch_hs_ice(DomainImpl<T>& D) {
  LAUNCH(k_hs_ice<T>, grid1(D.h.im), dim3(kBlock), D.h);
  return 0;
}

template <typename T>
int launch_stokes(DomainImpl<T>& D) {
  if (D.h.lm < 1) return 0;
  // 64 points per block, fewer when the levels' accumulators would pass
  // 64 KB of LDS (lm > 64)
  const int np = D.h.lm <= 64 ? 64 : 32;
  const size_t sh = (size_t)2 * D.h.lm * np * sizeof(double);
  hipLaunchKernelGGL(k_stokes<T>, dim3((D.h.im + np - 1) / np), dim3(np), sh, D.stream, D.h,
                     D.stokes);
  hipError_t e_ = hipGetLastError();
  if (e_ != hipSuccess) return fail(UMWM_ERR_HIP, std::string("k_stokes: ") + hipGetErrorString(e_));
  LAUNCH(k_stokes_efold<T>, grid1(D.h.im), dim3(kBlock), D.h, D.stokes);
  return 0;
}

template <typename T>
int launch_jonswap(DomainImpl<T>& D, const umwm_jonswap& js) {
  LAUNCH(k_jonswap<T>, dim3(grid1(D.h.im).x, D.h.om * D.h.pm), dim3(kBlock), D.h, js, D.S.dlnf);
  return 0;
}

#define INST(T)                                                          \
  template int launch_forcinginterpolate<T>(DomainImpl<T>&, bool);       \
  template int launch_forcing_list<T>(const DomainDev<T>*, const ForcingDev<T>*, \
                                      const int32_t*, int, int, bool, bool, hipStream_t); \
  template int launch_op_sin<T>(DomainImpl<T>&);                         \
  template int launch_op_sds<T>(DomainImpl<T>&);                         \
  template int launch_op_snl<T>(DomainImpl<T>&);                         \
  template int launch_op_sice<T>(DomainImpl<T>&);                        \
  template int launch_op_source<T>(DomainImpl<T>&);                      \
  template int launch_op_propagation<T>(DomainImpl<T>&);                 \
  template int launch_op_refraction<T>(DomainImpl<T>&);                  \
  template int launch_op_update<T>(DomainImpl<T>&);                      \
  template int launch_op_stress_atm<T>(DomainImpl<T>&);                  \
  template int launch_op_stress_ocn<T>(DomainImpl<T>&);                  \
  template int launch_diag<T>(DomainImpl<T>&);                           \
  template int launch_hs_ice<T>(DomainImpl<T>&);                         \
  template int launch_stokes<T>(DomainImpl<T>&);                         \
  template int launch_jonswap<T>(DomainImpl<T>&, const umwm_jonswap&);
INST(float)
INST(double)

}  // namespace umwm
