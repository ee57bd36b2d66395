// Shared declarations of the UMWM gfx950 library: device-side domain
// layout, per-domain scalars, launch descriptors and error plumbing.
//
// Device layout (one domain):
//   spectra     T[o][p][ldi]  sea point j = 0..im-1 contiguous, ghost at
//                             j = im (holds tiny(e), never updated; the
//                             reference's e(:,:,0), src/umwm_init.F90:1176)
//   tables      T[o][ldi]     per (frequency, sea point); cp0/cg0 carry the
//                             ghost column (src/umwm_init.F90:1652-1656)
//   per (p, j)  T[p][ldi]     cth_curv, sth_curv
//   per point   [ldq]         all mm*nm points (sea first, then land) for
//                             the forcing fields the reference keeps on imm
//   per sea pt  [ldi]
// Index maps are 0-based on the device: aliased neighbours point at the
// ghost (im) where the reference points at 0, true neighbours index the
// full unrolled grid (0..imm-1).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>

#include "../../include/umwm_hip.h"

namespace umwm {

// Sticky per-thread error message, surfaced by umwm_hip_last_error().
void set_error(const std::string& msg);
int fail(int code, const std::string& msg);

#define UMWM_HIP_TRY(expr)                                                     \
  do {                                                                         \
    hipError_t e_ = (expr);                                                    \
    if (e_ != hipSuccess)                                                      \
      return ::umwm::fail(UMWM_ERR_HIP, std::string(#expr) + ": " +            \
                                            hipGetErrorString(e_));            \
  } while (0)

// Scalars that the reference keeps in module globals and that the device
// updates itself (src/umwm_module.F90:60,66; src/umwm_advection.F90:110).
struct Scalars {
  double sumt, dts, dta, dtr, dtr_temp, maxrot;
  int32_t active;    // this substep runs (sumt < dtg at its start)
  int32_t nsub;      // substeps taken since the last reset
  uint32_t counter;  // arrival counter of the rates kernel's last block
  uint32_t counter_rot;
  int32_t anycur;    // any current != 0 (zerocurrents = !anycur)
  int32_t last;      // the latest active substep reached sumt >= dtg
  // the rates kernel's running grid min of pts and max of the rotation
  // rate as order-preserving keys (0 = none yet; its last block clears them)
  unsigned long long pts_key, rot_key;
  // surrogate (FP32): this substep's sampled audit sums (error and weight)
  // and the domain's verdict on the previous substep's audit (0: the
  // network may serve, 1: every point takes the plain rates)
  double sur_anum, sur_aden;
  int32_t sur_closed, sur_pad;
};

// Everything a kernel needs to know about one domain.
template <typename T>
struct DomainDev {
  int32_t im, imm, om, pm, ldi, ldq, mm, nm, lm;
  int32_t isglobal, restart, currents, nblocks, block0;
  // namelist and derived constants (double; kernels cast to T)
  double g, explim, fprog, dtg, dtamin, dth, dthg, oneovdth, twopisds_fac;
  double sin_fac, fieldscale1, fieldscale2, sds_power, inv_sds_power;
  double mss_fac, snl_fac, sdt_fac, z, nu_air, kappa, fice_lth, fice_uth, nu_water;
  double tiny;  // tiny(e) of the working precision
  // dispersion-table constants for the FP32 sweeps, which derive their
  // (frequency, point) tables from k and the depth
  double sfct, rhow0, sbf_fac, sbp_fac, bf1, bf2;
  // spectra
  T *e, *ef, *es;
  T *ssin, *sds, *snl, *dummy, *rotl, *rotr;  // operator path only
  // (o, j) tables
  const T *k, *kdk, *k3dk, *k4, *oneoverk4, *cp0, *cg0, *invcp0, *fkovg;
  const T *logl2overz, *cothkd, *sbf, *sdv, *snl_arg, *bf1r, *bf2r, *dwn;
  T *sdt, *sice;  // operator path only
  // (p, j)
  const T *cth_curv, *sth_curv;
  // spectral constants
  const T *f, *th, *cth, *sth, *c2th, *s2th, *cth2pp;
  // per direction {cth, sth, cos 2th, sin 2th}, one 4-value record, then
  // kDirPad zero records so that a wave's unused direction slots index it
  // without a clamp
  const T* dir4;
  const T* dom;   // frequency increment d omega per bin
  const int32_t *pl, *pr;
  // per unrolled point (ldq)
  T *wspd, *wdir, *uc, *vc, *rhoa, *rhow, *rhorat, *fice;
  // per sea point (ldi)
  const T *oneovar, *oneovdx, *oneovdy, *dxn, *dxs, *dye, *dyw, *depth;
  const T *ccurv, *scurv;  // cos, sin of the grid rotation (src/umwm_init.F90:1304-1309)
  const int32_t *ie, *iw, *in_, *is_, *iie, *iiw, *iin, *iis;
  int32_t *oc, *oc_rot;
  T *shelt, *fcutoff;
  T *cwdir, *swdir;  // cos, sin of wdir, written by the fused rates kernel
  double *pts, *hs_ice;
  // stress('atm') outputs, FP64 in both precisions
  double *ustar, *cd, *taux, *tauy, *taux_form, *tauy_form, *taux_skin;
  // when set, the substep that reaches dtg saves its E^n and its incoming
  // ustar here, from which stress('ocn') rebuilds that substep's source
  // arrays (null: not kept)
  T* e_keep;
  double* ustar_keep;
  double *tauy_skin, *taux_diag, *tauy_diag, *tailatmx, *tailatmy;
  double *us1, *vs1;  // surface Stokes drift used by stress('atm')
  // fused-path copy of the (o, j) tables, [om][ldi][FT_N]: the FT_N values
  // of one (frequency, point) are contiguous, so a lane fetches them with a
  // few 16-byte loads and a wave reads one contiguous span; the exclusive
  // frequency prefix of the cos^2 directional convolution times dth/2,
  // [om][ldi][4] (3 used, written by k_rates); the per-frequency stress /
  // wave-height partial sums, [om][ldi][4] (3 used; summed over frequency
  // in FP64 by k_stress)
  T* ftab;
  T* pfx;
  T* spart;
  Scalars* sc;
  // learned surrogate of the FP32 path (null sur_r: none attached): the
  // network's spreading factor per bin as dm^2 - 1 in bfloat16 bits,
  // [o*pm+p][ldi]; per point the input-range verdict (sur_ok) and the
  // choice k_rates made for the current substep (sur_use, read by
  // k_integrate); statistics (SurStat, surrogate.hpp); the policy: every
  // sur_audit_stride-th workgroup of points (rotating with the substep) is
  // also evaluated by the plain kernel, and a point falls back where the
  // energy-weighted share of its bins whose dm^2 leaves the prefix bounds,
  // or its audited dissipation error, exceeds the tolerances
  const uint16_t* sur_r;
  int32_t *sur_ok, *sur_use;
  double* sur_stats;
  int32_t sur_audit_stride, sur_pad;
  float sur_resid_tol, sur_audit_tol;
};

// Slots of one (frequency, point) record of DomainDev::ftab: the tables the
// sweep kernels cannot rebuild from k cheaply (k^4, 1/k^4, sdv, f k/g,
// snl_arg and k^3 dk are formed from k, kdk, bf1r and bf2r on the fly).
// The integrate kernel reads the first eight, the rates kernel all nine.
enum FusedTable {
  FT_CP0, FT_LOGL2, FT_COTH, FT_SBF, FT_K, FT_CG0, FT_B1, FT_B2, FT_KDK, FT_PAD, FT_N = 12
};
constexpr int kRatesSlots = 12;
constexpr int kIntegrateSlots = 8;

// Zero records after the pm records of DomainDev::dir4.  A sweep wave w of
// W <= 8 reads slots w*ceil(pm/W) + q, q < DMAX <= 36, which stay below
// pm + W - 1 + 36 <= pm + 43.
constexpr int kDirPad = 48;

// One workgroup of a fused launch: which domain, which 64 sea points.
struct BlockTask {
  int32_t dom;
  int32_t i0;
  int32_t local;  // block index within the domain
  int32_t pad;
};

// Diagnostic outputs of diag() (src/umwm_physics.F90:89-205), FP64.
struct DiagDev {
  double *ht, *mwp, *mwd, *mss, *mwl, *momx, *momy, *cgmxx, *cgmxy, *cgmyy;
  double *dwd, *dwl, *dwp, *dcp0, *dcg0, *dcp, *dcg;
};

// stress('ocn') outputs (src/umwm_stress.F90:34-214), FP64, per sea point,
// in the order of kOcnFields.
constexpr int kOcnFields = 18;
struct OcnDev {
  double* f[kOcnFields];
};
inline const char* const* ocn_field_names() {
  static const char* const n[kOcnFields] = {"tailocnx", "tailocny", "taux_ocntop", "tauy_ocntop", "taux_ocnbot", "tauy_ocnbot", "taux_snl", "tauy_snl", "epsx_ocn", "epsy_ocn", "epsx_atm", "epsy_atm", "taux1", "tauy1", "taux2", "tauy2", "taux3", "tauy3"};
  return n;
}

// Stokes drift outputs (src/umwm_stokes.F90:13-132), FP64.
struct StokesDev {
  const double* depth;  // lm levels, negative down
  double *us, *vs, *ds; // us/vs [l][ldi]
};

// Forcing levels kept on the device (src/umwm_forcing.F90:11-99).
template <typename T>
struct ForcingDev {
  const T *uwb, *uwf, *vwb, *vwf;      // (mm,nm)
  const T *ucb, *ucf, *vcb, *vcf;      // (mm,nm)
  const T *ficeb, *ficef;              // (mm,nm)
  const T *rhoab, *rhoaf, *rhowb, *rhowf;  // unrolled (ldq)
  const int32_t *mi, *ni;              // unrolled -> 1-based (m,n)
  int32_t winds, currents, air_density, water_density, seaice, pad;
};

constexpr int kLanes = 64;  // wavefront width on CDNA

inline int round_up(int v, int m) { return (v + m - 1) / m * m; }

}  // namespace umwm
