// C-ABI of the UMWM gfx950 library (include/umwm_hip.h): domain lifetime,
// reference operator entry points, the fused substep and dtg loop, field
// exchange in the reference's layout, timing, and batched ensembles.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "domain.hpp"
#include "physics.hpp"
#include "rawio.hpp"
#include "surrogate.hpp"

// Scratch of umwm_hip_substeps_noalloc: the fused kernels' prefix moments
// and stress partials ([om][ldi][4] each in the working precision), which
// the other entry points keep in buffers allocated with the domain.  Every
// GPU the library runs on holds one copy until the process ends; in host
// code the name is only a stand-in, and the device address comes from
// hipGetSymbolAddress (umwm::noalloc_scratch).
constexpr size_t kNoAllocScratchBytes = size_t(32) << 20;
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

namespace umwm {

namespace {
thread_local std::string g_last_error;

// device address of `scratch` per device, looked up once per device with
// that device current
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
unsigned char* g_scratch_ptr[kMaxDevices] = {};
}

void set_error(const std::string& msg) { g_last_error = msg; }
int fail(int code, const std::string& msg) {
  g_last_error = msg;
  return code;
}

// The calling thread's current device's copy of `scratch`.
int noalloc_scratch(unsigned char** out) {
  int dev = 0;
  UMWM_HIP_TRY(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) return fail(UMWM_ERR_HIP, "device index out of range");
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    UMWM_HIP_TRY(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) return fail(UMWM_ERR_HIP, "hipGetSymbolAddress returned null for the scratch array");
    g_scratch_ptr[dev] = static_cast<unsigned char*>(p);
  }
  *out = g_scratch_ptr[dev];
  return 0;
}

namespace {

// ---------------------------------------------------------------------
// small device helpers
// ---------------------------------------------------------------------
template <typename T>
__global__ void k_fill(T* p, size_t n, T v) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  for (; i < n; i += (size_t)gridDim.x * blockDim.x) p[i] = v;
}

__global__ void k_begin_dtg(Scalars* sc) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    sc->sumt = 0;
    sc->nsub = 0;
  }
}

// Gathers (sumt, dts, active, nsub) of every batch member into one array so
// the host reads them with a single copy.
struct ScalarsView {
  double sumt, dts;
  int32_t active, nsub;
};
__global__ void k_gather_scalars(Scalars* const* scs, int n, ScalarsView* out) {
  int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n) {
    const Scalars* s = scs[k];
    out[k].sumt = s->sumt;
    out[k].dts = s->dts;
    out[k].active = s->active;
    out[k].nsub = s->nsub;
  }
}

template <typename T>
int dfill(T* p, size_t n, T v, hipStream_t s) {
  if (!n) return 0;
  size_t blocks = std::min<size_t>((n + 255) / 256, 65536);
  hipLaunchKernelGGL(k_fill<T>, dim3((unsigned)blocks), dim3(256), 0, s, p, n, v);
  UMWM_HIP_TRY(hipGetLastError());
  return 0;
}

template <typename T>
int upload(T* dst, const std::vector<double>& src, size_t n, hipStream_t s) {
  std::vector<T> tmp(n, T(0));
  for (size_t q = 0; q < std::min(n, src.size()); ++q) tmp[q] = T(src[q]);
  UMWM_HIP_TRY(hipMemcpyAsync(dst, tmp.data(), n * sizeof(T), hipMemcpyHostToDevice, s));
  UMWM_HIP_TRY(hipStreamSynchronize(s));
  return 0;
}

int upload_int(int32_t* dst, const std::vector<int>& src, size_t n, int pad, hipStream_t s) {
  std::vector<int32_t> tmp(n, pad);
  for (size_t q = 0; q < std::min(n, src.size()); ++q) tmp[q] = src[q];
  UMWM_HIP_TRY(hipMemcpyAsync(dst, tmp.data(), n * sizeof(int32_t), hipMemcpyHostToDevice, s));
  UMWM_HIP_TRY(hipStreamSynchronize(s));
  return 0;
}

// XCD-aware block order: the hardware deals workgroups round-robin to the
// 8 XCDs, so hand each XCD a contiguous range of point blocks; the x-
// neighbours of a point (one grid column away in the unrolled index) then
// sit in the same XCD's L2.
constexpr int kXcd = 8;
std::vector<int> xcd_order(int nb) {
  std::vector<int> pb(nb);
  const int per = nb / kXcd, rem = nb % kXcd;
  for (int b = 0; b < nb; ++b) {
    const int x = b % kXcd, k = b / kXcd;
    pb[b] = x * per + std::min(x, rem) + k;
  }
  return pb;
}

// Default tile shape before autotuning: D directions per thread, W =
// ceil(pm / D) waves (at most 8).
Tile choose_tile(int pm) {
  auto pick = [pm](std::initializer_list<int> ds, int& D, int& W) {
    for (int c : ds)
      if ((pm + c - 1) / c <= 8) {
        D = c;
        W = (pm + c - 1) / c;
        return;
      }
    D = 36;
    W = (pm + 35) / 36;
  };
  Tile t;
  pick({9, 18, 36}, t.D1, t.W1);
  pick({9, 12, 18}, t.D2, t.W2);
  t.pmax = pm;
  return t;
}

bool tile_ok(const Tile& t, int pm) {
  const bool d1 = t.D1 == 6 || t.D1 == 9 || t.D1 == 18 || t.D1 == 36;
  const bool d2 = t.D2 == 3 || t.D2 == 6 || t.D2 == 9 || t.D2 == 12 || t.D2 == 18;
  return d1 && d2 && t.W1 >= 1 && t.W1 <= 8 && t.W2 >= 1 && t.W2 <= 8 && t.D1 * t.W1 >= pm &&
         t.D2 * t.W2 >= pm && (t.NB == 2 || t.NB == 4);
}

Tile tile_from(const int32_t* a, int pmax) {
  Tile t;
  t.D1 = a[0];
  t.W1 = a[1];
  t.D2 = a[2];
  t.W2 = a[3];
  t.NB = a[4];
  t.pmax = pmax;
  return t;
}

void tile_to(const Tile& t, int32_t* a) {
  a[0] = t.D1;
  a[1] = t.W1;
  a[2] = t.D2;
  a[3] = t.W2;
  a[4] = t.NB;
}

}  // namespace

template <typename T>
DomainImpl<T>::~DomainImpl() {
  if (stream) (void)hipStreamSynchronize(stream);
  if (gexec) (void)hipGraphExecDestroy(gexec);
  if (graph) (void)hipGraphDestroy(graph);
  for (auto& e : ev)
    if (e) (void)hipEventDestroy(e);
  if (sc_host) (void)hipHostFree(sc_host);
  arena.release();
  if (stream) (void)hipStreamDestroy(stream);
}

template <typename T>
int ensure_ops(DomainImpl<T>& D) {
  if (D.ops_ready) return 0;
  const size_t ns = (size_t)D.h.om * D.h.pm * D.h.ldi, nt = (size_t)D.h.om * D.h.ldi;
  DomainDev<T>& h = D.h;
  h.ef = D.arena.template alloc<T>(ns);
  if (!h.ef) return fail(UMWM_ERR_NOMEM, "device allocation of operator arrays failed");
  // ef as the fused path leaves it: E^{n+1}, equal to e
  UMWM_HIP_TRY(hipMemcpyAsync(h.ef, h.e, ns * sizeof(T), hipMemcpyDeviceToDevice, D.stream));
  h.ssin = D.arena.template alloc<T>(ns);
  h.sds = D.arena.template alloc<T>(ns);
  h.snl = D.arena.template alloc<T>(ns);
  h.dummy = D.arena.template alloc<T>(ns);
  h.rotl = D.arena.template alloc<T>(ns);
  h.rotr = D.arena.template alloc<T>(ns);
  h.sdt = D.arena.template alloc<T>(nt);
  h.sice = D.arena.template alloc<T>(nt);
  for (auto& p : D.ocn.f) {
    p = D.arena.template alloc<double>(h.ldi);
    if (!p) return fail(UMWM_ERR_NOMEM, "device allocation of operator arrays failed");
    if (int rc = dfill<double>(p, h.ldi, 0.0, D.stream)) return rc;
  }
  if (!h.ssin || !h.sds || !h.snl || !h.dummy || !h.rotl || !h.rotr || !h.sdt || !h.sice)
    return fail(UMWM_ERR_NOMEM, "device allocation of operator arrays failed");
  for (T* p : {h.ssin, h.sds, h.snl, h.dummy, h.rotl, h.rotr})
    if (int rc = dfill<T>(p, ns, T(0), D.stream)) return rc;
  for (T* p : {h.sdt, h.sice})
    if (int rc = dfill<T>(p, nt, T(0), D.stream)) return rc;
  UMWM_HIP_TRY(hipMemcpyAsync(D.d, &D.h, sizeof(DomainDev<T>), hipMemcpyHostToDevice, D.stream));
  UMWM_HIP_TRY(hipStreamSynchronize(D.stream));
  D.ops_ready = true;
  return 0;
}

template <typename T>
int create_impl(const umwm_params& P, const umwm_grid_input* g, const umwm_forcing_input* f0,
                DomainImpl<T>*& out) {
  auto* D = new DomainImpl<T>();
  std::string err;
  int rc = build_setup(P, g, f0, D->S, err);
  if (rc) {
    delete D;
    return fail(rc, err);
  }
  const HostSetup& S = D->S;
  D->precision = sizeof(T);
  D->first = true;
  D->firstdtg = !P.restart;
  if (hipStreamCreateWithFlags(&D->stream, hipStreamNonBlocking) != hipSuccess) {
    delete D;
    return fail(UMWM_ERR_HIP, "hipStreamCreate failed");
  }
  hipStream_t s = D->stream;
  DomainDev<T>& h = D->h;
  DeviceArena& A = D->arena;
  h.im = S.im; h.imm = S.imm; h.om = S.om; h.pm = S.pm; h.ldi = S.ldi; h.ldq = S.ldq;
  h.mm = S.mm; h.nm = S.nm; h.lm = S.lm;
  h.isglobal = P.isglobal; h.restart = P.restart; h.currents = P.currents;
  h.g = P.g; h.explim = P.explim; h.fprog = P.fprog; h.dtg = P.dtg; h.dtamin = S.dtamin;
  h.dth = S.dth; h.dthg = S.dthg; h.oneovdth = S.oneovdth; h.twopisds_fac = S.twopisds_fac;
  h.sin_fac = P.sin_fac; h.fieldscale1 = S.fieldscale1; h.fieldscale2 = S.fieldscale2;
  h.sds_power = P.sds_power; h.inv_sds_power = S.inv_sds_power; h.mss_fac = P.mss_fac;
  h.snl_fac = P.snl_fac; h.sdt_fac = P.sdt_fac; h.z = P.z; h.nu_air = P.nu_air;
  h.kappa = P.kappa; h.fice_lth = P.fice_lth; h.fice_uth = P.fice_uth;
  h.nu_water = P.nu_water;
  h.tiny = sizeof(T) == 8 ? DBL_MIN : (double)FLT_MIN;
  h.sfct = P.sfct; h.rhow0 = P.rhow0; h.sbf_fac = P.sbf_fac; h.sbp_fac = P.sbp_fac;
  h.bf1 = S.bf1; h.bf2 = S.bf2;

  const int om = S.om, pm = S.pm, ldi = S.ldi, ldq = S.ldq;
  const size_t ns = (size_t)om * pm * ldi, nt = (size_t)om * ldi;
  // the fused kernels address one frequency of a spectrum (pm rows), of the
  // table block (FT_N rows) or of a per-point array with one buffer
  // descriptor each, and drop predicated stores with an offset of 2^31
  const size_t maxrows = std::max<size_t>({(size_t)pm, (size_t)FT_N, 4});
  if (maxrows * ldi * std::max(sizeof(T), sizeof(double)) >= (size_t(1) << 31) ||
      (size_t)ldq * sizeof(double) >= (size_t(1) << 31)) {
    delete D;
    return fail(UMWM_ERR_SIZE, "domain too large for 32-bit buffer descriptors");
  }
  // ef is an operator-path array (allocated with the others on first use):
  // the fused substep keeps E_s in place and never stores it
  h.e = A.template alloc<T>(ns);
  h.ef = nullptr;
  h.es = A.template alloc<T>(ns);
  if (!h.e || !h.es) {
    delete D;
    return fail(UMWM_ERR_NOMEM, "device allocation of the spectrum failed");
  }
  // tables
  struct TabUp { const T** dst; const std::vector<double>* src; };
  T *k = A.template alloc<T>(nt), *kdk = A.template alloc<T>(nt), *k3dk = A.template alloc<T>(nt);
  T *k4 = A.template alloc<T>(nt), *ok4 = A.template alloc<T>(nt), *cp0 = A.template alloc<T>(nt);
  T *cg0 = A.template alloc<T>(nt), *icp = A.template alloc<T>(nt), *fkg = A.template alloc<T>(nt);
  T *ll2 = A.template alloc<T>(nt), *cth = A.template alloc<T>(nt), *sbf = A.template alloc<T>(nt);
  T *sdv = A.template alloc<T>(nt), *sarg = A.template alloc<T>(nt), *b1 = A.template alloc<T>(nt);
  T *b2 = A.template alloc<T>(nt), *dwn = A.template alloc<T>(nt);
  T *ccv = A.template alloc<T>((size_t)pm * ldi), *scv = A.template alloc<T>((size_t)pm * ldi);
  T *f = A.template alloc<T>(om), *th = A.template alloc<T>(pm), *ct = A.template alloc<T>(pm);
  T *st = A.template alloc<T>(pm), *c2 = A.template alloc<T>(pm), *s2 = A.template alloc<T>(pm);
  T* c2pp = A.template alloc<T>((size_t)pm * pm);
  int32_t *pl = A.template alloc<int32_t>(pm), *pr = A.template alloc<int32_t>(pm);
  if (!dwn || !c2pp || !pr) {
    delete D;
    return fail(UMWM_ERR_NOMEM, "device allocation of tables failed");
  }
  const std::pair<T*, const std::vector<double>*> tabs[] = {
      {k, &S.k}, {kdk, &S.kdk}, {k3dk, &S.k3dk}, {k4, &S.k4}, {ok4, &S.oneoverk4},
      {cp0, &S.cp0}, {cg0, &S.cg0}, {icp, &S.invcp0}, {fkg, &S.fkovg}, {ll2, &S.logl2overz},
      {cth, &S.cothkd}, {sbf, &S.sbf}, {sdv, &S.sdv}, {sarg, &S.snl_arg}, {b1, &S.bf1r},
      {b2, &S.bf2r}, {dwn, &S.dwn}};
  for (auto& tb : tabs)
    if ((rc = upload<T>(tb.first, *tb.second, nt, s))) { delete D; return rc; }
  if ((rc = upload<T>(ccv, S.cth_curv, (size_t)pm * ldi, s)) ||
      (rc = upload<T>(scv, S.sth_curv, (size_t)pm * ldi, s)) || (rc = upload<T>(f, S.f, om, s)) ||
      (rc = upload<T>(th, S.th, pm, s)) || (rc = upload<T>(ct, S.cth, pm, s)) ||
      (rc = upload<T>(st, S.sth, pm, s)) || (rc = upload<T>(c2, S.c2th, pm, s)) ||
      (rc = upload<T>(s2, S.s2th, pm, s)) || (rc = upload<T>(c2pp, S.cth2pp, (size_t)pm * pm, s)) ||
      (rc = upload_int(pl, S.pl, pm, 0, s)) || (rc = upload_int(pr, S.pr, pm, 0, s))) {
    delete D;
    return rc;
  }
  h.k = k; h.kdk = kdk; h.k3dk = k3dk; h.k4 = k4; h.oneoverk4 = ok4; h.cp0 = cp0; h.cg0 = cg0;
  h.invcp0 = icp; h.fkovg = fkg; h.logl2overz = ll2; h.cothkd = cth; h.sbf = sbf; h.sdv = sdv;
  h.snl_arg = sarg; h.bf1r = b1; h.bf2r = b2; h.dwn = dwn; h.cth_curv = ccv; h.sth_curv = scv;
  h.f = f; h.th = th; h.cth = ct; h.sth = st; h.c2th = c2; h.s2th = s2; h.cth2pp = c2pp;
  {
    const size_t n4 = (size_t)4 * (pm + kDirPad);
    std::vector<double> d4(n4, 0.0);
    for (int p = 0; p < pm; ++p) {
      d4[4 * p] = S.cth[p];
      d4[4 * p + 1] = S.sth[p];
      d4[4 * p + 2] = S.c2th[p];
      d4[4 * p + 3] = S.s2th[p];
    }
    T* dd = A.template alloc<T>(n4);
    if (!dd) { delete D; return fail(UMWM_ERR_NOMEM, "device allocation failed"); }
    if ((rc = upload<T>(dd, d4, n4, s))) { delete D; return rc; }
    h.dir4 = dd;
    T* dw = A.template alloc<T>((size_t)om);
    if (!dw) { delete D; return fail(UMWM_ERR_NOMEM, "device allocation failed"); }
    if ((rc = upload<T>(dw, S.dom, (size_t)om, s))) { delete D; return rc; }
    h.dom = dw;
  }
  h.pl = pl; h.pr = pr;
  // per-(frequency, point) table records for the fused FP64 kernels,
  // [o][ldi][FT_N] (the FP32 kernels derive them from k)
  if (sizeof(T) == 8) {
    const std::vector<double>* ft[FT_PAD] = {&S.cp0, &S.logl2overz, &S.cothkd, &S.sbf, &S.k,
                                             &S.cg0, &S.bf1r, &S.bf2r, &S.kdk};
    std::vector<double> blk((size_t)FT_N * nt, 0.0);
    for (size_t q = 0; q < nt; ++q)
      for (int t = 0; t < FT_PAD; ++t) blk[q * FT_N + t] = (*ft[t])[q];
    h.ftab = A.template alloc<T>((size_t)FT_N * nt);
    if (!h.ftab) {
      delete D;
      return fail(UMWM_ERR_NOMEM, "device allocation of the table block failed");
    }
    if ((rc = upload<T>(h.ftab, blk, (size_t)FT_N * nt, s))) { delete D; return rc; }
  }
  // per unrolled point
  T** unr[] = {&h.wspd, &h.wdir, &h.uc, &h.vc, &h.rhoa, &h.rhow, &h.rhorat, &h.fice};
  for (T** p : unr) *p = A.template alloc<T>(ldq);
  std::vector<double> rhorat(ldq, 0.0);
  for (int q = 0; q < S.imm; ++q) rhorat[q] = S.rhoa[q] / S.rhow[q];
  if ((rc = upload<T>(h.wspd, S.wspd, ldq, s)) || (rc = upload<T>(h.wdir, S.wdir, ldq, s)) ||
      (rc = upload<T>(h.uc, S.uc, ldq, s)) || (rc = upload<T>(h.vc, S.vc, ldq, s)) ||
      (rc = upload<T>(h.rhoa, S.rhoa, ldq, s)) || (rc = upload<T>(h.rhow, S.rhow, ldq, s)) ||
      (rc = upload<T>(h.rhorat, rhorat, ldq, s)) || (rc = upload<T>(h.fice, S.fice, ldq, s))) {
    delete D;
    return rc;
  }
  // per sea point
  T *ovar = A.template alloc<T>(ldi), *ovdx = A.template alloc<T>(ldi), *ovdy = A.template alloc<T>(ldi);
  T *dxn = A.template alloc<T>(ldi), *dxs = A.template alloc<T>(ldi), *dye = A.template alloc<T>(ldi);
  T *dyw = A.template alloc<T>(ldi), *dep = A.template alloc<T>(ldi);
  T *ccu = A.template alloc<T>(ldi), *scu = A.template alloc<T>(ldi);
  std::vector<double> dsea(S.d.begin(), S.d.begin() + S.im);
  if ((rc = upload<T>(ovar, S.oneovar, ldi, s)) || (rc = upload<T>(ovdx, S.oneovdx, ldi, s)) ||
      (rc = upload<T>(ovdy, S.oneovdy, ldi, s)) || (rc = upload<T>(dxn, S.dxn, ldi, s)) ||
      (rc = upload<T>(dxs, S.dxs, ldi, s)) || (rc = upload<T>(dye, S.dye, ldi, s)) ||
      (rc = upload<T>(dyw, S.dyw, ldi, s)) || (rc = upload<T>(dep, dsea, ldi, s)) ||
      (rc = upload<T>(ccu, S.ccurv, ldi, s)) || (rc = upload<T>(scu, S.scurv, ldi, s))) {
    delete D;
    return rc;
  }
  h.oneovar = ovar; h.oneovdx = ovdx; h.oneovdy = ovdy; h.dxn = dxn; h.dxs = dxs; h.dye = dye;
  h.dyw = dyw; h.depth = dep; h.ccurv = ccu; h.scurv = scu;
  int32_t *ie = A.template alloc<int32_t>(ldi), *iw = A.template alloc<int32_t>(ldi);
  int32_t *in = A.template alloc<int32_t>(ldi), *is = A.template alloc<int32_t>(ldi);
  int32_t *iie = A.template alloc<int32_t>(ldi), *iiw = A.template alloc<int32_t>(ldi);
  int32_t *iin = A.template alloc<int32_t>(ldi), *iis = A.template alloc<int32_t>(ldi);
  // padding lanes point at the ghost (aliased) or at point 0 (true maps)
  if ((rc = upload_int(ie, S.ie, ldi, S.im, s)) || (rc = upload_int(iw, S.iw, ldi, S.im, s)) ||
      (rc = upload_int(in, S.in, ldi, S.im, s)) || (rc = upload_int(is, S.is, ldi, S.im, s)) ||
      (rc = upload_int(iie, S.iie, ldi, 0, s)) || (rc = upload_int(iiw, S.iiw, ldi, 0, s)) ||
      (rc = upload_int(iin, S.iin, ldi, 0, s)) || (rc = upload_int(iis, S.iis, ldi, 0, s))) {
    delete D;
    return rc;
  }
  h.ie = ie; h.iw = iw; h.in_ = in; h.is_ = is; h.iie = iie; h.iiw = iiw; h.iin = iin; h.iis = iis;
  h.oc = A.template alloc<int32_t>(ldi);
  h.oc_rot = A.template alloc<int32_t>(ldi);
  h.shelt = A.template alloc<T>(ldi);
  h.cwdir = A.template alloc<T>(ldi);
  h.swdir = A.template alloc<T>(ldi);
  h.fcutoff = A.template alloc<T>(ldi);
  h.pts = A.template alloc<double>(ldi);
  h.hs_ice = A.template alloc<double>(ldi);
  double** pd[] = {&h.ustar, &h.cd, &h.taux, &h.tauy, &h.taux_form, &h.tauy_form, &h.taux_skin,
                   &h.tauy_skin, &h.taux_diag, &h.tauy_diag, &h.tailatmx, &h.tailatmy,
                   &h.us1, &h.vs1};
  for (double** p : pd) *p = A.template alloc<double>(ldi);
  if (!h.vs1 || !h.hs_ice) {
    delete D;
    return fail(UMWM_ERR_NOMEM, "device allocation of point arrays failed");
  }
  UMWM_HIP_TRY(hipMemsetAsync(h.oc, 0, ldi * sizeof(int32_t), s));
  UMWM_HIP_TRY(hipMemsetAsync(h.oc_rot, 0, ldi * sizeof(int32_t), s));
  for (double** p : pd) UMWM_HIP_TRY(hipMemsetAsync(*p, 0, ldi * sizeof(double), s));
  UMWM_HIP_TRY(hipMemsetAsync(h.pts, 0, ldi * sizeof(double), s));
  UMWM_HIP_TRY(hipMemsetAsync(h.hs_ice, 0, ldi * sizeof(double), s));
  {
    std::vector<double> u(ldi, 0.0), c(ldi, 0.0);
    for (int j = 0; j < S.im; ++j) { u[j] = S.ustar[j]; c[j] = S.cd[j]; }
    UMWM_HIP_TRY(hipMemcpyAsync(h.ustar, u.data(), ldi * sizeof(double), hipMemcpyHostToDevice, s));
    UMWM_HIP_TRY(hipMemcpyAsync(h.cd, c.data(), ldi * sizeof(double), hipMemcpyHostToDevice, s));
    UMWM_HIP_TRY(hipStreamSynchronize(s));
  }
  if ((rc = dfill<T>(h.shelt, ldi, T(0), s)) || (rc = dfill<T>(h.fcutoff, ldi, T(0), s)) ||
      (rc = dfill<T>(h.cwdir, ldi, T(1), s)) || (rc = dfill<T>(h.swdir, ldi, T(0), s))) {
    delete D;
    return rc;
  }
  // spectrum: e = tiny(e) everywhere, ghost included (src/umwm_init.F90:1182)
  const T tiny = sizeof(T) == 8 ? T(DBL_MIN) : T(FLT_MIN);
  if ((rc = dfill<T>(h.e, ns, tiny, s)) || (rc = dfill<T>(h.es, ns, tiny, s))) {
    delete D;
    return rc;
  }
  // reduction scratch and scalars
  h.nblocks = (S.im + kLanes - 1) / kLanes;
  h.block0 = 0;
  h.pfx = A.template alloc<T>(4 * nt);
  h.spart = A.template alloc<T>(4 * nt);
  h.sc = A.template alloc<Scalars>(1);
  if (!h.sc || !h.pfx || !h.spart) {
    delete D;
    return fail(UMWM_ERR_NOMEM, "device allocation failed");
  }
  if ((rc = dfill<T>(h.pfx, 4 * nt, T(0), s)) || (rc = dfill<T>(h.spart, 4 * nt, T(0), s))) {
    delete D;
    return rc;
  }
  Scalars sc0{};
  sc0.anycur = (!P.currents && (P.uc0 != 0 || P.vc0 != 0)) ? 1 : 0;
  if (P.currents) {
    for (int q = 0; q < S.imm; ++q)
      if (S.uc[q] != 0 || S.vc[q] != 0) sc0.anycur = 1;
  }
  UMWM_HIP_TRY(hipMemcpyAsync(h.sc, &sc0, sizeof(Scalars), hipMemcpyHostToDevice, s));
  UMWM_HIP_TRY(hipHostMalloc((void**)&D->sc_host, sizeof(Scalars), hipHostMallocDefault));
  // diag and Stokes outputs
  double** dg[] = {&D->diag.ht, &D->diag.mwp, &D->diag.mwd, &D->diag.mss, &D->diag.mwl,
                   &D->diag.momx, &D->diag.momy, &D->diag.cgmxx, &D->diag.cgmxy, &D->diag.cgmyy,
                   &D->diag.dwd, &D->diag.dwl, &D->diag.dwp, &D->diag.dcp0, &D->diag.dcg0,
                   &D->diag.dcp, &D->diag.dcg};
  for (double** p : dg) {
    *p = A.template alloc<double>(ldi);
    UMWM_HIP_TRY(hipMemsetAsync(*p, 0, ldi * sizeof(double), s));
  }
  const int lm = std::max(S.lm, 1);
  double* dl = A.template alloc<double>(lm);
  std::vector<double> dlh(lm, 0.0);
  for (int l = 0; l < S.lm; ++l) dlh[l] = S.depth_l[l];
  UMWM_HIP_TRY(hipMemcpyAsync(dl, dlh.data(), lm * sizeof(double), hipMemcpyHostToDevice, s));
  D->stokes.depth = dl;
  const size_t nus = (size_t)lm * ldi;
  D->stokes.us = A.template alloc<double>(nus);
  D->stokes.vs = A.template alloc<double>(nus);
  D->stokes.ds = A.template alloc<double>(ldi);
  if (!D->stokes.ds) {
    delete D;
    return fail(UMWM_ERR_NOMEM, "device allocation of Stokes arrays failed");
  }
  UMWM_HIP_TRY(hipMemsetAsync(D->stokes.us, 0, nus * sizeof(double), s));
  UMWM_HIP_TRY(hipMemsetAsync(D->stokes.vs, 0, nus * sizeof(double), s));
  UMWM_HIP_TRY(hipMemsetAsync(D->stokes.ds, 0, ldi * sizeof(double), s));
  // forcing levels
  ForcingDev<T>& F = D->fh;
  const size_t n2 = (size_t)S.mm * S.nm;
  F.winds = P.winds; F.currents = P.currents; F.air_density = P.air_density;
  F.water_density = P.water_density; F.seaice = P.seaice;
  int32_t *mi = A.template alloc<int32_t>(ldq), *ni = A.template alloc<int32_t>(ldq);
  if ((rc = upload_int(mi, S.mi, ldq, 1, s)) || (rc = upload_int(ni, S.ni, ldq, 1, s))) {
    delete D;
    return rc;
  }
  F.mi = mi; F.ni = ni;
  const bool anyf = P.winds || P.currents || P.air_density || P.water_density || P.seaice;
  if (anyf) {
    D->uwb = A.template alloc<T>(n2); D->uwf = A.template alloc<T>(n2);
    D->vwb = A.template alloc<T>(n2); D->vwf = A.template alloc<T>(n2);
    D->ucb = A.template alloc<T>(n2); D->ucf = A.template alloc<T>(n2);
    D->vcb = A.template alloc<T>(n2); D->vcf = A.template alloc<T>(n2);
    D->ficeb = A.template alloc<T>(n2); D->ficef = A.template alloc<T>(n2);
    D->rhoab = A.template alloc<T>(ldq); D->rhoaf = A.template alloc<T>(ldq);
    D->rhowb = A.template alloc<T>(ldq); D->rhowf = A.template alloc<T>(ldq);
    if (!D->rhowf) {
      delete D;
      return fail(UMWM_ERR_NOMEM, "device allocation of forcing levels failed");
    }
    if ((rc = upload<T>(D->uwf, S.uwf, n2, s)) || (rc = upload<T>(D->vwf, S.vwf, n2, s)) ||
        (rc = upload<T>(D->ucf, S.ucf, n2, s)) || (rc = upload<T>(D->vcf, S.vcf, n2, s)) ||
        (rc = upload<T>(D->ficef, S.ficef, n2, s)) || (rc = upload<T>(D->rhoaf, S.rhoaf, ldq, s)) ||
        (rc = upload<T>(D->rhowf, S.rhowf, ldq, s)) || (rc = upload<T>(D->uwb, S.uwf, n2, s)) ||
        (rc = upload<T>(D->vwb, S.vwf, n2, s)) || (rc = upload<T>(D->ucb, S.ucf, n2, s)) ||
        (rc = upload<T>(D->vcb, S.vcf, n2, s)) || (rc = upload<T>(D->ficeb, S.ficef, n2, s)) ||
        (rc = upload<T>(D->rhoab, S.rhoaf, ldq, s)) || (rc = upload<T>(D->rhowb, S.rhowf, ldq, s))) {
      delete D;
      return rc;
    }
    F.uwb = D->uwb; F.uwf = D->uwf; F.vwb = D->vwb; F.vwf = D->vwf;
    F.ucb = D->ucb; F.ucf = D->ucf; F.vcb = D->vcb; F.vcf = D->vcf;
    F.ficeb = D->ficeb; F.ficef = D->ficef;
    F.rhoab = D->rhoab; F.rhoaf = D->rhoaf; F.rhowb = D->rhowb; F.rhowf = D->rhowf;
  }
  D->fd = A.template alloc<ForcingDev<T>>(1);
  D->list0 = A.template alloc<int32_t>(1);
  D->d_scr = A.template alloc<DomainDev<T>>(1);
  if ((rc = noalloc_scratch(&D->scr))) {
    delete D;
    return rc;
  }
  if (!D->fd || !D->list0 || !D->d_scr) {
    delete D;
    return fail(UMWM_ERR_NOMEM, "device allocation failed");
  }
  UMWM_HIP_TRY(hipMemcpyAsync(D->fd, &F, sizeof(ForcingDev<T>), hipMemcpyHostToDevice, s));
  UMWM_HIP_TRY(hipMemsetAsync(D->list0, 0, sizeof(int32_t), s));
  // device copy of the domain struct and the block task list
  D->d = A.template alloc<DomainDev<T>>(1);
  const int nb = h.nblocks;
  D->tasks = A.template alloc<BlockTask>(std::max(nb, 1));
  D->ntasks = nb;
  {
    std::vector<BlockTask> tk(nb);
    std::vector<int> ord = xcd_order(nb);
    for (int b = 0; b < nb; ++b) tk[b] = BlockTask{0, ord[b] * kLanes, ord[b], 0};
    UMWM_HIP_TRY(hipMemcpyAsync(D->tasks, tk.data(), nb * sizeof(BlockTask), hipMemcpyHostToDevice, s));
  }
  UMWM_HIP_TRY(hipMemcpyAsync(D->d, &h, sizeof(DomainDev<T>), hipMemcpyHostToDevice, s));
  D->tile = choose_tile(pm);
  for (auto& e : D->ev) UMWM_HIP_TRY(hipEventCreate(&e));
  UMWM_HIP_TRY(hipStreamSynchronize(s));
  // stokes_drift('init') on the cold-start spectrum, as umwm_initialize does
  if ((rc = launch_stokes(*D))) { delete D; return rc; }
  if ((rc = launch_hs_ice(*D))) { delete D; return rc; }
  UMWM_HIP_TRY(hipStreamSynchronize(s));
  out = D;
  return 0;
}

// ---------------------------------------------------------------------
// substeps
// ---------------------------------------------------------------------
template <typename T>
int read_scalars(DomainImpl<T>& D) {
  UMWM_HIP_TRY(hipMemcpyAsync(D.sc_host, D.h.sc, sizeof(Scalars), hipMemcpyDeviceToHost, D.stream));
  UMWM_HIP_TRY(hipStreamSynchronize(D.stream));
  return 0;
}

template <typename T>
bool any_forcing(const DomainImpl<T>& D) {
  const umwm_params& P = D.S.P;
  return P.winds || P.currents || P.air_density || P.water_density || P.seaice;
}

// Runs sin_d12, sds_d12 and snl_d12 (operator kernels) on the E^n and
// ustar that k_keep saved from the substep that reached dtg, leaving ssin,
// sds (after cothkd), snl and sdt as that substep's source evaluation left
// them (src/umwm_top.F90:95-98); the forcing and the cut-off have not
// changed since.
template <typename T>
int materialize_sources(DomainImpl<T>& D) {
  T* const e = D.h.e;
  double* const ust = D.h.ustar;
  D.h.e = D.h.e_keep;
  D.h.ustar = D.h.ustar_keep;
  int rc = launch_op_sin(D);
  if (!rc) rc = launch_op_sds(D);
  if (!rc) rc = launch_op_snl(D);
  D.h.e = e;
  D.h.ustar = ust;
  return rc;
}

// keep_sources: allocates the E^n / ustar save of the substep reaching dtg
template <typename T>
int enable_keep(DomainImpl<T>& D) {
  if (!D.h.e_keep) {
    const size_t ns = (size_t)D.h.om * D.h.pm * D.h.ldi;
    D.h.e_keep = D.arena.template alloc<T>(ns);
    D.h.ustar_keep = D.arena.template alloc<double>(D.h.ldi);
    if (!D.h.e_keep || !D.h.ustar_keep) {
      D.h.e_keep = nullptr;
      return fail(UMWM_ERR_NOMEM, "device allocation of the kept sources failed");
    }
    UMWM_HIP_TRY(hipMemcpyAsync(D.d, &D.h, sizeof(DomainDev<T>), hipMemcpyHostToDevice, D.stream));
  }
  // nothing is saved yet: clear the flag of an earlier dtg end
  if (int rc = read_scalars(D)) return rc;
  Scalars s = *D.sc_host;
  s.last = 0;
  UMWM_HIP_TRY(hipMemcpy(D.h.sc, &s, sizeof(Scalars), hipMemcpyHostToDevice));
  return 0;
}

// One fused substep; `dd` is the device domain record the fused kernels
// read (the domain's own by default, or the one whose scratch lives in the
// shared array).
// The attached surrogate's rates for the current spectrum (FP32 domains).
template <typename T>
int surrogate_rates(DomainImpl<T>& D, hipStream_t s, bool& on) {
  on = false;
  if constexpr (std::is_same_v<T, float>) {
    if (D.sur && D.h.sur_r) {
      on = true;
      return launch_surrogate_inference(D, s);
    }
  }
  return 0;
}

template <typename T>
int fused_substep(DomainImpl<T>& D, const DomainDev<T>* dd = nullptr) {
  int rc;
  if (any_forcing(D) && (rc = launch_forcinginterpolate(D, true))) return rc;
  bool sur = false;
  if (D.timing) UMWM_HIP_TRY(hipEventRecord(D.ev[5], D.stream));
  if ((rc = surrogate_rates(D, D.stream, sur))) return rc;
  if (D.timing) UMWM_HIP_TRY(hipEventRecord(D.ev[6], D.stream));
  const int rotation = (D.S.P.currents || D.first || !D.rot_valid) ? 1 : 0;
  rc = launch_fused<T>(dd ? dd : D.d, D.tasks, D.ntasks, D.h.om, D.first ? 1 : 0, rotation,
                       D.tile, D.h.e_keep != nullptr, D.stream, D.timing ? D.ev : nullptr, sur);
  if (rc) return rc;
  D.ef_mirrors_e = true;
  D.rot_valid = true;
  if (D.first) {
    if ((rc = launch_diag(D))) return rc;
    D.first = false;
  }
  if (D.timing) {
    UMWM_HIP_TRY(hipEventSynchronize(D.ev[4]));
    for (int q = 0; q < 4; ++q) {
      float ms = 0;
      UMWM_HIP_TRY(hipEventElapsedTime(&ms, D.ev[q], D.ev[q + 1]));
      D.ms[q] += ms;
      D.launches[q] += 1;
    }
    if (sur) {
      float ms = 0;
      UMWM_HIP_TRY(hipEventElapsedTime(&ms, D.ev[5], D.ev[6]));
      D.ms[4] += ms;
      D.launches[4] += 1;
    }
  }
  return 0;
}

// What a captured steady-state substep depends on beyond stable device
// pointers: the launch shape, the keep and surrogate switches and the
// surrogate's range settings (launch arguments by value).
template <typename T>
uint64_t graph_signature(const DomainImpl<T>& D) {
  uint64_t h = 1469598103934665603ull;
  auto mix = [&](uint64_t v) { h = (h ^ v) * 1099511628211ull; };
  const Tile& t = D.tile;
  for (int v : {t.D1, t.W1, t.D2, t.W2, t.NB, t.pmax}) mix((uint64_t)v);
  mix(D.h.e_keep != nullptr);
  mix((uint64_t)(uintptr_t)D.sur);
  mix((uint64_t)(uintptr_t)D.h.sur_r);
  uint64_t b;
  std::memcpy(&b, &D.sur_margin, 8);
  mix(b);
  std::memcpy(&b, &D.sur_zmax, 8);
  mix(b);
  return h | 1;  // never 0, the value of "nothing captured"
}

// One steady-state substep (not the first, no timing) replayed from a
// captured hipGraph: forcing interpolation, the surrogate's inference and
// the fused kernels, the same launches fused_substep makes.
template <typename T>
int graph_substep(DomainImpl<T>& D) {
  const uint64_t sig = graph_signature(D);
  if (!D.gexec || D.gsig != sig) {
    if (D.gexec) (void)hipGraphExecDestroy(D.gexec);
    if (D.graph) (void)hipGraphDestroy(D.graph);
    D.gexec = nullptr;
    D.graph = nullptr;
    UMWM_HIP_TRY(hipStreamBeginCapture(D.stream, hipStreamCaptureModeThreadLocal));
    const int rc = fused_substep(D);
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(D.stream, &g);
    if (rc) {
      if (g) (void)hipGraphDestroy(g);
      return rc;
    }
    if (e != hipSuccess) return fail(UMWM_ERR_HIP, std::string("graph capture: ") + hipGetErrorString(e));
    D.graph = g;
    UMWM_HIP_TRY(hipGraphInstantiate(&D.gexec, g, nullptr, nullptr, 0));
    D.gsig = sig;
  }
  UMWM_HIP_TRY(hipGraphLaunch(D.gexec, D.stream));
  return 0;
}

// n substeps back to back; the scalars are read once, after the last.
// Steady-state substeps replay the captured graph.
template <typename T>
int substeps_impl(DomainImpl<T>& D, int n) {
  for (int k = 0; k < n; ++k) {
    const bool steady = !D.first && D.rot_valid && !D.timing && D.use_graph;
    if (int rc = steady ? graph_substep(D) : fused_substep(D)) return rc;
  }
  return read_scalars(D);
}

// umwm_hip_substeps_noalloc: the same substeps with the prefix moments and
// stress partials in the device's shared scratch array.  Calls no
// allocating routine: the record copy and the scratch address were set up
// when the domain was created.
template <typename T>
int substeps_noalloc_impl(DomainImpl<T>& D, int n) {
  const size_t part = (size_t)4 * D.h.om * D.h.ldi * sizeof(T);
  const size_t off = (part + 255) / 256 * 256;
  if (off + part > kNoAllocScratchBytes)
    return fail(UMWM_ERR_SIZE, "substeps_noalloc: the domain needs " +
                                   std::to_string(off + part) + " bytes of scratch (4 om ldi "
                                   "values twice), more than the " +
                                   std::to_string(kNoAllocScratchBytes) +
                                   " bytes of the shared array; use umwm_hip_substeps");
  if (!D.scr || !D.d_scr) return fail(UMWM_ERR_STATE, "substeps_noalloc: no scratch set up");
  D.h_scr = D.h;
  D.h_scr.pfx = reinterpret_cast<T*>(D.scr);
  D.h_scr.spart = reinterpret_cast<T*>(D.scr + off);
  UMWM_HIP_TRY(hipMemcpyAsync(D.d_scr, &D.h_scr, sizeof(DomainDev<T>), hipMemcpyHostToDevice,
                              D.stream));
  // the array keeps the previous call's contents: start from zeros, as the
  // domain's own buffers did (lanes past im are read, never used)
  if (int rc = dfill<T>(reinterpret_cast<T*>(D.scr), (off + part) / sizeof(T), T(0), D.stream))
    return rc;
  for (int k = 0; k < n; ++k)
    if (int rc = fused_substep(D, (const DomainDev<T>*)D.d_scr)) return rc;
  return read_scalars(D);
}

template <typename T>
int run_dtg_impl(DomainImpl<T>& D, int32_t* nsub) {
  hipLaunchKernelGGL(k_begin_dtg, dim3(1), dim3(64), 0, D.stream, D.h.sc);
  UMWM_HIP_TRY(hipGetLastError());
  const double dtg = D.S.P.dtg;
  double sumt = 0, dts = 0;
  int n = 0, rc;
  for (;;) {
    // run ahead by the substeps the last dts says are left (at least one,
    // at most 64); those past the end of the dtg do nothing
    int ahead = 1;
    if (dts > 0) ahead = (int)std::min(64.0, std::max(1.0, std::ceil((dtg - sumt) / dts)));
    if ((rc = substeps_impl(D, ahead))) return rc;
    n += ahead;
    sumt = D.sc_host->sumt;
    dts = D.sc_host->dts;
    if (!(sumt < dtg)) break;
    if (n > 10000000) return fail(UMWM_ERR_STATE, "run_dtg: substep count runaway");
  }
  n = D.sc_host->nsub;
  if (D.firstdtg) D.firstdtg = false;
  if (D.S.P.stokes && (rc = launch_stokes(D))) return rc;
  if ((rc = launch_diag(D))) return rc;
  UMWM_HIP_TRY(hipStreamSynchronize(D.stream));
  if (nsub) *nsub = n;
  return 0;
}

// ---------------------------------------------------------------------
// field exchange
// ---------------------------------------------------------------------
enum FieldKind { F_SPEC, F_TAB, F_PT_T, F_PT_D, F_PT_I, F_UNR_T, F_HOST_MAP, F_HOST_2D, F_HOST_UNR, F_STOKES, F_HOST_SEA };

struct FieldRef {
  FieldKind kind;
  void* ptr = nullptr;
  const std::vector<int>* imap = nullptr;
  const std::vector<double>* dmap = nullptr;
};

template <typename T>
bool find_field(DomainImpl<T>& D, const std::string& n, FieldRef& r) {
  DomainDev<T>& h = D.h;
  const HostSetup& S = D.S;
  auto set = [&](FieldKind k, void* p) { r.kind = k; r.ptr = p; return true; };
  // spectra
  if (n == "e") return set(F_SPEC, h.e);
  if (n == "ef") {
    if (D.ef_mirrors_e) return set(F_SPEC, h.e);
    if (ensure_ops(D)) return false;
    return set(F_SPEC, h.ef);
  }
  if (n == "es") return set(F_SPEC, h.es);
  if (n == "ssin" || n == "sds" || n == "snl" || n == "dummy" || n == "rotl" || n == "rotr" ||
      n == "sdt" || n == "sice") {
    if (ensure_ops(D)) return false;
    if (n == "ssin") return set(F_SPEC, h.ssin);
    if (n == "sds") return set(F_SPEC, h.sds);
    if (n == "snl") return set(F_SPEC, h.snl);
    if (n == "dummy") return set(F_SPEC, h.dummy);
    if (n == "rotl") return set(F_SPEC, h.rotl);
    if (n == "rotr") return set(F_SPEC, h.rotr);
    if (n == "sdt") return set(F_TAB, h.sdt);
    if (n == "sice") return set(F_TAB, h.sice);
  }
  // tables
  const std::pair<const char*, const T*> tabs[] = {
      {"k", h.k}, {"kdk", h.kdk}, {"k3dk", h.k3dk}, {"k4", h.k4}, {"oneoverk4", h.oneoverk4},
      {"cp0", h.cp0}, {"cg0", h.cg0}, {"invcp0", h.invcp0}, {"fkovg", h.fkovg},
      {"logl2overz", h.logl2overz}, {"cothkd", h.cothkd}, {"sbf", h.sbf}, {"sdv", h.sdv},
      {"snl_arg", h.snl_arg}, {"bf1_renorm", h.bf1r}, {"bf2_renorm", h.bf2r}, {"dwn", h.dwn}};
  for (auto& t : tabs)
    if (n == t.first) return set(F_TAB, (void*)t.second);
  const std::pair<const char*, T*> unr[] = {
      {"wspd", h.wspd}, {"wdir", h.wdir}, {"uc", h.uc}, {"vc", h.vc}, {"rhoa", h.rhoa},
      {"rhow", h.rhow}, {"rhorat", h.rhorat}, {"fice", h.fice}};
  for (auto& t : unr)
    if (n == t.first) return set(F_UNR_T, t.second);
  const std::pair<const char*, const T*> ptt[] = {
      {"shelt", h.shelt}, {"fcutoff", h.fcutoff}, {"oneovar", h.oneovar}, {"oneovdx", h.oneovdx},
      {"oneovdy", h.oneovdy}, {"dxn", h.dxn}, {"dxs", h.dxs}, {"dye", h.dye}, {"dyw", h.dyw}};
  for (auto& t : ptt)
    if (n == t.first) return set(F_PT_T, (void*)t.second);
  const std::pair<const char*, double*> ptd[] = {
      {"ustar", h.ustar}, {"cd", h.cd}, {"taux", h.taux}, {"tauy", h.tauy},
      {"taux_form", h.taux_form}, {"tauy_form", h.tauy_form}, {"taux_skin", h.taux_skin},
      {"tauy_skin", h.tauy_skin}, {"taux_diag", h.taux_diag}, {"tauy_diag", h.tauy_diag},
      {"tailatmx", h.tailatmx}, {"tailatmy", h.tailatmy}, {"physics_time_step", h.pts},
      {"ht", D.diag.ht}, {"mwp", D.diag.mwp}, {"mwd", D.diag.mwd}, {"mss", D.diag.mss},
      {"mwl", D.diag.mwl}, {"momx", D.diag.momx}, {"momy", D.diag.momy},
      {"cgmxx", D.diag.cgmxx}, {"cgmxy", D.diag.cgmxy}, {"cgmyy", D.diag.cgmyy},
      {"dwd", D.diag.dwd}, {"dwl", D.diag.dwl}, {"dwp", D.diag.dwp}, {"dcp0", D.diag.dcp0},
      {"dcg0", D.diag.dcg0}, {"dcp", D.diag.dcp}, {"dcg", D.diag.dcg}, {"ds", D.stokes.ds},
      {"hs_ice", h.hs_ice}};
  for (auto& t : ptd)
    if (n == t.first) return set(F_PT_D, t.second);
  for (int q = 0; q < kOcnFields; ++q)
    if (n == ocn_field_names()[q]) {
      if (!D.ocn.f[q]) return false;  // before the operator arrays exist
      return set(F_PT_D, D.ocn.f[q]);
    }
  if (n == "oc") return set(F_PT_I, h.oc);
  if (n == "oc_rot") return set(F_PT_I, h.oc_rot);
  if (n == "us") return set(F_STOKES, D.stokes.us);
  if (n == "vs") return set(F_STOKES, D.stokes.vs);
  // host-side setup products
  const std::pair<const char*, const std::vector<int>*> maps[] = {
      {"iw", &S.iw}, {"ie", &S.ie}, {"is", &S.is}, {"in", &S.in},
      {"iiw", &S.iiw}, {"iie", &S.iie}, {"iis", &S.iis}, {"iin", &S.iin}};
  for (auto& t : maps)
    if (n == t.first) { r.kind = F_HOST_MAP; r.imap = t.second; return true; }
  if (n == "mi") { r.kind = F_HOST_UNR; r.imap = &S.mi; return true; }
  if (n == "ni") { r.kind = F_HOST_UNR; r.imap = &S.ni; return true; }
  if (n == "mask") { r.kind = F_HOST_2D; r.imap = &S.mask; return true; }
  if (n == "ii") { r.kind = F_HOST_2D; r.imap = &S.ii; return true; }
  const std::pair<const char*, const std::vector<double>*> d2[] = {
      {"dx_2d", &S.dx2}, {"dy_2d", &S.dy2}, {"curv", &S.curv2}, {"d_2d", &S.d2},
      {"lon", &S.lon}, {"lat", &S.lat}};
  for (auto& t : d2)
    if (n == t.first) { r.kind = F_HOST_2D; r.dmap = t.second; return true; }
  const std::pair<const char*, const std::vector<double>*> du[] = {
      {"dx", &S.dx}, {"dy", &S.dy}, {"d", &S.d}};
  for (auto& t : du)
    if (n == t.first) { r.kind = F_HOST_UNR; r.dmap = t.second; return true; }
  if (n == "ar") { r.kind = F_HOST_SEA; r.dmap = &S.ar; return true; }
  return false;
}

template <typename T>
int64_t field_size(const DomainImpl<T>& D, FieldKind k) {
  const auto& h = D.h;
  switch (k) {
    case F_SPEC: return (int64_t)h.om * h.pm * h.im;
    case F_TAB: return (int64_t)h.om * h.im;
    case F_PT_T: case F_PT_D: case F_PT_I: case F_HOST_MAP: case F_HOST_SEA: return h.im;
    case F_UNR_T: case F_HOST_UNR: return h.imm;
    case F_HOST_2D: return (int64_t)h.mm * h.nm;
    case F_STOKES: return (int64_t)h.im * std::max(h.lm, 1);
  }
  return 0;
}

template <typename T>
int get_field_impl(DomainImpl<T>& D, const std::string& name, double* buf, int64_t n) {
  FieldRef r;
  if (!find_field(D, name, r)) return fail(UMWM_ERR_UNKNOWN_FIELD, "unknown field: " + name);
  const int64_t want = field_size(D, r.kind);
  if (n != want)
    return fail(UMWM_ERR_SIZE, "field " + name + " has " + std::to_string(want) + " values");
  const auto& h = D.h;
  const size_t ldi = h.ldi, im = h.im;
  UMWM_HIP_TRY(hipStreamSynchronize(D.stream));
  switch (r.kind) {
    case F_SPEC: {
      // device [o][p][ldi] -> reference (om,pm,im) column-major
      const size_t np = (size_t)h.om * h.pm;
      std::vector<T> tmp(np * ldi);
      UMWM_HIP_TRY(hipMemcpy(tmp.data(), r.ptr, tmp.size() * sizeof(T), hipMemcpyDeviceToHost));
      // device op = o*pm + p; reference index o + om*(p + pm*j)
      for (size_t o = 0; o < (size_t)h.om; ++o)
        for (size_t p = 0; p < (size_t)h.pm; ++p) {
          const T* src = tmp.data() + (o * h.pm + p) * ldi;
          for (size_t j = 0; j < im; ++j) buf[j * np + p * h.om + o] = (double)src[j];
        }
      return 0;
    }
    case F_TAB: {
      std::vector<T> tmp((size_t)h.om * ldi);
      UMWM_HIP_TRY(hipMemcpy(tmp.data(), r.ptr, tmp.size() * sizeof(T), hipMemcpyDeviceToHost));
      for (size_t o = 0; o < (size_t)h.om; ++o)
        for (size_t j = 0; j < im; ++j) buf[j * h.om + o] = (double)tmp[o * ldi + j];
      return 0;
    }
    case F_PT_T: case F_UNR_T: {
      std::vector<T> tmp(n);
      UMWM_HIP_TRY(hipMemcpy(tmp.data(), r.ptr, n * sizeof(T), hipMemcpyDeviceToHost));
      for (int64_t q = 0; q < n; ++q) buf[q] = (double)tmp[q];
      return 0;
    }
    case F_PT_D:
      UMWM_HIP_TRY(hipMemcpy(buf, r.ptr, n * sizeof(double), hipMemcpyDeviceToHost));
      return 0;
    case F_PT_I: {
      std::vector<int32_t> tmp(n);
      UMWM_HIP_TRY(hipMemcpy(tmp.data(), r.ptr, n * sizeof(int32_t), hipMemcpyDeviceToHost));
      for (int64_t q = 0; q < n; ++q) buf[q] = tmp[q] + 1.0;  // 1-based bin index
      return 0;
    }
    case F_STOKES: {
      for (int l = 0; l < std::max(h.lm, 1); ++l)
        UMWM_HIP_TRY(hipMemcpy(buf + (size_t)l * im, (double*)r.ptr + (size_t)l * ldi,
                               im * sizeof(double), hipMemcpyDeviceToHost));
      return 0;
    }
    case F_HOST_MAP: {
      // aliased maps: ghost -> 0; true maps: 1-based
      const bool aliased = name.size() == 2;
      for (int64_t q = 0; q < n; ++q) {
        const int v = (*r.imap)[q];
        buf[q] = aliased ? (v == (int)im ? 0.0 : v + 1.0) : v + 1.0;
      }
      return 0;
    }
    case F_HOST_UNR: case F_HOST_2D: case F_HOST_SEA:
      for (int64_t q = 0; q < n; ++q) buf[q] = r.imap ? (double)(*r.imap)[q] : (*r.dmap)[q];
      return 0;
  }
  return fail(UMWM_ERR_UNKNOWN_FIELD, name);
}

template <typename T>
int set_field_impl(DomainImpl<T>& D, const std::string& name, const double* buf, int64_t n) {
  if (name == "ef") D.ef_mirrors_e = false;
  FieldRef r;
  if (!find_field(D, name, r)) return fail(UMWM_ERR_UNKNOWN_FIELD, "unknown field: " + name);
  const int64_t want = field_size(D, r.kind);
  if (n != want)
    return fail(UMWM_ERR_SIZE, "field " + name + " has " + std::to_string(want) + " values");
  const auto& h = D.h;
  const size_t ldi = h.ldi, im = h.im;
  UMWM_HIP_TRY(hipStreamSynchronize(D.stream));
  switch (r.kind) {
    case F_SPEC: {
      const size_t np = (size_t)h.om * h.pm;
      std::vector<T> tmp(np * ldi);
      UMWM_HIP_TRY(hipMemcpy(tmp.data(), r.ptr, tmp.size() * sizeof(T), hipMemcpyDeviceToHost));
      for (size_t o = 0; o < (size_t)h.om; ++o)
        for (size_t p = 0; p < (size_t)h.pm; ++p) {
          T* dst = tmp.data() + (o * h.pm + p) * ldi;
          for (size_t j = 0; j < im; ++j) dst[j] = T(buf[j * np + p * h.om + o]);
        }
      UMWM_HIP_TRY(hipMemcpy(r.ptr, tmp.data(), tmp.size() * sizeof(T), hipMemcpyHostToDevice));
      break;
    }
    case F_PT_T: case F_UNR_T: {
      std::vector<T> tmp(n);
      for (int64_t q = 0; q < n; ++q) tmp[q] = T(buf[q]);
      UMWM_HIP_TRY(hipMemcpy(r.ptr, tmp.data(), n * sizeof(T), hipMemcpyHostToDevice));
      break;
    }
    case F_PT_D:
      UMWM_HIP_TRY(hipMemcpy(r.ptr, buf, n * sizeof(double), hipMemcpyHostToDevice));
      break;
    default:
      return fail(UMWM_ERR_ARGUMENT, "field " + name + " is read-only");
  }
  if (name == "e") {
    int rc = launch_hs_ice(D);
    if (rc) return rc;
    UMWM_HIP_TRY(hipStreamSynchronize(D.stream));
  }
  return 0;
}

// ---------------------------------------------------------------------
// batch
// ---------------------------------------------------------------------
template <typename T>
struct BatchImpl {
  // members with the same pm share one task list and launch shape: W waves
  // of D directions each, W = ceil(pm / D) for the batch's D
  struct Group {
    int pm = 0, omax = 0, ntasks = 0;
    BlockTask* tasks = nullptr;
    Tile tile;
  };
  std::vector<DomainImpl<T>*> doms;
  DeviceArena arena;
  DomainDev<T>* d_doms = nullptr;
  std::vector<Group> groups;
  Tile tile;  // as set (D1, D2, NB are what the groups share)
  hipStream_t stream = nullptr;
  Scalars** d_scs = nullptr;
  ScalarsView* d_view = nullptr;
  ScalarsView* h_view = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  // forcing records of the members (refreshed from each member at the
  // start of every run, so that forcinginput's level swaps reach the
  // kernels and the captured graph) and the list of forced members
  ForcingDev<T>* d_fds = nullptr;
  int32_t* d_flist = nullptr;
  int nforced = 0, maximm = 0;
  bool forced_currents = false;
  // k_keep runs for the batch when any member keeps its sources; the
  // captured graph is rebuilt when that changes
  bool keep = false, keep_captured = false;
  int omax = 0;
  ~BatchImpl() {
    if (stream) (void)hipStreamSynchronize(stream);
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    if (h_view) (void)hipHostFree(h_view);
    arena.release();
    if (stream) (void)hipStreamDestroy(stream);
  }
};

}  // namespace umwm

// =====================================================================
// extern "C"
// =====================================================================
using namespace umwm;

struct umwm_domain {
  DomainBase* base = nullptr;
};

struct umwm_batch {
  int precision = 8;
  void* impl = nullptr;
};

namespace {

template <typename Fn>
int dispatch(umwm_domain* dom, Fn&& fn) {
  if (!dom || !dom->base) return fail(UMWM_ERR_ARGUMENT, "null domain");
  if (dom->base->precision == 4) return fn(*static_cast<DomainImpl<float>*>(dom->base));
  return fn(*static_cast<DomainImpl<double>*>(dom->base));
}

template <typename T>
int sync_and_check(DomainImpl<T>& D) {
  UMWM_HIP_TRY(hipStreamSynchronize(D.stream));
  return 0;
}

}  // namespace

namespace {
// Holland (1980) wind of a translating cyclone at every grid point of a
// planar domain, written into the next forcing level ([n][m], m fastest, as
// k_forcing reads it); FP64 throughout.
template <typename T>
__global__ void k_holland(T* __restrict__ uw, T* __restrict__ vw, int mm, int nm, double dx,
                          double dy, umwm_holland st, double dp, double f) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= mm * nm) return;
  const int m = q % mm, n = q / mm;
  const double x = m * dx - st.xc, y = n * dy - st.yc;
  const double r = sqrt(x * x + y * y) + 1.0;
  const double a = pow(st.rmax / r, st.b);
  const double hf = 0.5 * r * f;
  const double v = sqrt(st.b * dp / st.rho * a * exp(-a) + hf * hf) - hf;
  const double th = atan2(y, x) + 0.5 * M_PI + st.inflow_deg * (M_PI / 180.0);
  uw[q] = T(v * cos(th) + 0.5 * st.trans_u);
  vw[q] = T(v * sin(th) + 0.5 * st.trans_v);
}
}  // namespace

extern "C" {

const char* umwm_hip_error_string(int status) {
  switch (status) {
    case UMWM_OK: return "ok";
    case UMWM_ERR_ARGUMENT: return "bad argument";
    case UMWM_ERR_HIP: return "HIP runtime error";
    case UMWM_ERR_NOMEM: return "out of memory";
    case UMWM_ERR_STATE: return "call made in the wrong state";
    case UMWM_ERR_UNKNOWN_FIELD: return "unknown field";
    case UMWM_ERR_SIZE: return "buffer size mismatch";
    case UMWM_ERR_IO: return "file error";
  }
  return "unknown status";
}

const char* umwm_hip_last_error(void) { return g_last_error.c_str(); }

int umwm_hip_device_count(int* count) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) n = 0;
  if (count) *count = n;
  return 0;
}

int umwm_hip_create(const umwm_params* params, const umwm_grid_input* grid,
                    const umwm_forcing_input* forcing0, int precision, umwm_domain** out) {
  if (!params || !out) return fail(UMWM_ERR_ARGUMENT, "null argument");
  if (precision != UMWM_FP32 && precision != UMWM_FP64)
    return fail(UMWM_ERR_ARGUMENT, "precision must be 4 or 8");
  const std::string perr = check_params(*params);
  if (!perr.empty()) return fail(UMWM_ERR_ARGUMENT, perr);
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n < 1)
    return fail(UMWM_ERR_HIP, "no HIP device visible");
  auto* dom = new umwm_domain();
  int rc;
  if (precision == UMWM_FP32) {
    DomainImpl<float>* D = nullptr;
    rc = create_impl<float>(*params, grid, forcing0, D);
    dom->base = D;
  } else {
    DomainImpl<double>* D = nullptr;
    rc = create_impl<double>(*params, grid, forcing0, D);
    dom->base = D;
  }
  if (rc) {
    delete dom;
    return rc;
  }
  *out = dom;
  return 0;
}

int umwm_hip_destroy(umwm_domain* dom) {
  if (!dom) return 0;
  delete dom->base;
  delete dom;
  return 0;
}

int umwm_hip_dims(const umwm_domain* dom, int32_t out[8]) {
  if (!dom || !dom->base || !out) return fail(UMWM_ERR_ARGUMENT, "null argument");
  const HostSetup& S = dom->base->S;
  out[0] = S.im; out[1] = S.imm; out[2] = S.om; out[3] = S.pm;
  out[4] = S.mm; out[5] = S.nm; out[6] = S.lm; out[7] = dom->base->precision;
  return 0;
}

int umwm_hip_init_jonswap(umwm_domain* dom, const umwm_jonswap* js) {
  if (!js) return fail(UMWM_ERR_ARGUMENT, "null jonswap");
  return dispatch(dom, [&](auto& D) {
    if (js->fp <= 0 && !js->fp_from_wind) return fail(UMWM_ERR_ARGUMENT, "fp must be > 0");
    int rc = launch_jonswap(D, *js);
    if (!rc) rc = launch_hs_ice(D);
    if (!rc) rc = sync_and_check(D);
    return rc;
  });
}

int umwm_hip_stokes_drift(umwm_domain* dom, int init) {
  (void)init;  // transfer coefficients are evaluated on the fly
  return dispatch(dom, [&](auto& D) {
    int rc = launch_stokes(D);
    return rc ? rc : sync_and_check(D);
  });
}

int umwm_hip_forcinginterpolate(umwm_domain* dom) {
  return dispatch(dom, [&](auto& D) {
    int rc = launch_forcinginterpolate(D, false);
    return rc ? rc : sync_and_check(D);
  });
}

int umwm_hip_forcinginput(umwm_domain* dom, const umwm_forcing_input* nx) {
  if (!nx) return fail(UMWM_ERR_ARGUMENT, "null forcing");
  return dispatch(dom, [&](auto& D) {
    using T = std::remove_reference_t<decltype(*D.h.e)>;
    const umwm_params& P = D.S.P;
    if (!any_forcing(D)) return 0;
    if ((P.winds && (!nx->uw || !nx->vw)) || (P.currents && (!nx->uc || !nx->vc)) ||
        (P.air_density && !nx->rhoa) || (P.water_density && !nx->rhow) ||
        (P.seaice && !nx->fice))
      return fail(UMWM_ERR_ARGUMENT, "forcing snapshot lacks a switched field");
    UMWM_HIP_TRY(hipStreamSynchronize(D.stream));
    const size_t n2 = (size_t)D.S.mm * D.S.nm, ldq = D.S.ldq;
    // level n <- level n+1 (src/umwm_forcing.F90:11-37)
    std::swap(D.uwb, D.uwf); std::swap(D.vwb, D.vwf); std::swap(D.ucb, D.ucf);
    std::swap(D.vcb, D.vcf); std::swap(D.ficeb, D.ficef); std::swap(D.rhoab, D.rhoaf);
    std::swap(D.rhowb, D.rhowf);
    int rc = 0;
    load_forcing_level(D.S, *nx);
    if (P.winds && ((rc = upload<T>(D.uwf, D.S.uwf, n2, D.stream)) ||
                    (rc = upload<T>(D.vwf, D.S.vwf, n2, D.stream))))
      return rc;
    if (P.currents && ((rc = upload<T>(D.ucf, D.S.ucf, n2, D.stream)) ||
                       (rc = upload<T>(D.vcf, D.S.vcf, n2, D.stream))))
      return rc;
    if (P.seaice && (rc = upload<T>(D.ficef, D.S.ficef, n2, D.stream))) return rc;
    if (P.air_density && (rc = upload<T>(D.rhoaf, D.S.rhoaf, ldq, D.stream))) return rc;
    if (P.water_density && (rc = upload<T>(D.rhowf, D.S.rhowf, ldq, D.stream))) return rc;
    ForcingDev<T>& F = D.fh;
    F.uwb = D.uwb; F.uwf = D.uwf; F.vwb = D.vwb; F.vwf = D.vwf;
    F.ucb = D.ucb; F.ucf = D.ucf; F.vcb = D.vcb; F.vcf = D.vcf;
    F.ficeb = D.ficeb; F.ficef = D.ficef;
    F.rhoab = D.rhoab; F.rhoaf = D.rhoaf; F.rhowb = D.rhowb; F.rhowf = D.rhowf;
    // the kernels (and any graph that captured them) read the levels
    // through the device record
    UMWM_HIP_TRY(hipMemcpyAsync(D.fd, &F, sizeof(ForcingDev<T>), hipMemcpyHostToDevice, D.stream));
    UMWM_HIP_TRY(hipStreamSynchronize(D.stream));
    return 0;
  });
}

int umwm_hip_holland_forcinginput(umwm_domain* dom, const umwm_holland* st) {
  if (!st) return fail(UMWM_ERR_ARGUMENT, "null storm");
  return dispatch(dom, [&](auto& D) {
    using T = std::remove_reference_t<decltype(*D.h.e)>;
    const umwm_params& P = D.S.P;
    if (!P.winds || P.currents || P.seaice || P.air_density || P.water_density)
      return fail(UMWM_ERR_STATE, "holland_forcinginput: the wind must be the only forcing");
    if (P.gridfromfile || P.isglobal)
      return fail(UMWM_ERR_STATE, "holland_forcinginput: planar grids only");
    if (!(st->rmax > 0) || !(st->b > 0) || !(st->rho > 0) || !(st->vmax >= 0))
      return fail(UMWM_ERR_ARGUMENT, "holland_forcinginput: rmax, b, rho > 0, vmax >= 0");
    UMWM_HIP_TRY(hipStreamSynchronize(D.stream));
    // level n <- level n+1 (src/umwm_forcing.F90:11-37), the new level
    // generated in place
    std::swap(D.uwb, D.uwf);
    std::swap(D.vwb, D.vwf);
    const int mm = D.S.mm, nm = D.S.nm;
    const double dp = st->dp > 0 ? st->dp : st->rho * M_E * st->vmax * st->vmax / st->b;
    const double f = 2 * 7.292e-5 * std::sin(st->lat_deg * M_PI / 180.0);
    hipLaunchKernelGGL(k_holland<T>, dim3((unsigned)(((int64_t)mm * nm + 255) / 256)), dim3(256),
                       0, D.stream, D.uwf, D.vwf, mm, nm, P.delx, P.dely, *st, dp, f);
    UMWM_HIP_TRY(hipGetLastError());
    ForcingDev<T>& F = D.fh;
    F.uwb = D.uwb; F.uwf = D.uwf; F.vwb = D.vwb; F.vwf = D.vwf;
    UMWM_HIP_TRY(hipMemcpyAsync(D.fd, &F, sizeof(ForcingDev<T>), hipMemcpyHostToDevice, D.stream));
    UMWM_HIP_TRY(hipStreamSynchronize(D.stream));
    return 0;
  });
}

#define OP_ENTRY(name, launcher)                                   \
  int name(umwm_domain* dom) {                                     \
    return dispatch(dom, [&](auto& D) {                            \
      int rc = ensure_ops(D);                                      \
      if (!rc) rc = launcher(D);                                   \
      if (!rc) rc = sync_and_check(D);                             \
      return rc;                                                   \
    });                                                            \
  }

OP_ENTRY(umwm_hip_sin_d12, launch_op_sin)
OP_ENTRY(umwm_hip_sds_d12, launch_op_sds)
OP_ENTRY(umwm_hip_snl_d12, launch_op_snl)
OP_ENTRY(umwm_hip_s_ice, launch_op_sice)
OP_ENTRY(umwm_hip_propagation, launch_op_propagation)
OP_ENTRY(umwm_hip_update, launch_op_update)

int umwm_hip_source(umwm_domain* dom) {
  return dispatch(dom, [&](auto& D) {
    int rc = ensure_ops(D);
    D.ef_mirrors_e = false;
    if (!rc) rc = launch_op_source(D);
    if (!rc) rc = sync_and_check(D);
    return rc;
  });
}

int umwm_hip_refraction(umwm_domain* dom) {
  return dispatch(dom, [&](auto& D) {
    int rc = ensure_ops(D);
    if (!rc) rc = launch_op_refraction(D);
    if (!rc) rc = sync_and_check(D);
    return rc;
  });
}

int umwm_hip_stress(umwm_domain* dom, const char* option) {
  if (!option) return fail(UMWM_ERR_ARGUMENT, "null option");
  const std::string opt(option, strnlen(option, 3));
  if (opt != "atm" && opt != "ocn")
    return fail(UMWM_ERR_ARGUMENT, "stress option must be 'atm' or 'ocn'");
  return dispatch(dom, [&](auto& D) {
    int rc = 0;
    if (opt == "ocn" && D.ef_mirrors_e) {
      // after fused substeps: rebuild the sources of the substep that
      // reached dtg from its saved E^n and incoming ustar
      if (D.h.e_keep && (rc = read_scalars(D))) return rc;
      if (!D.h.e_keep || !D.sc_host->last)
        return fail(UMWM_ERR_STATE,
                    "stress('ocn') reads ssin, sds and snl of the last source evaluation: "
                    "after fused substeps it needs keep_sources set before the substep "
                    "that reaches dtg, or run that substep through the operator entry points");
      if ((rc = ensure_ops(D))) return rc;
      if ((rc = materialize_sources(D))) return rc;
    } else if (opt == "ocn" && !D.ops_ready) {
      return fail(UMWM_ERR_STATE,
                  "stress('ocn') reads ssin, sds and snl of the last source evaluation: "
                  "no source terms have been evaluated yet");
    }
    rc = ensure_ops(D);
    if (!rc) rc = opt == "atm" ? launch_op_stress_atm(D) : launch_op_stress_ocn(D);
    if (!rc) rc = sync_and_check(D);
    return rc;
  });
}

int umwm_hip_diag(umwm_domain* dom) {
  return dispatch(dom, [&](auto& D) {
    int rc = launch_diag(D);
    return rc ? rc : sync_and_check(D);
  });
}

int umwm_hip_end_first(umwm_domain* dom) {
  return dispatch(dom, [&](auto& D) {
    D.first = false;
    return 0;
  });
}

int umwm_hip_substep(umwm_domain* dom, double* sumt, double* dts) {
  return dispatch(dom, [&](auto& D) {
    int rc = fused_substep(D);
    if (!rc) rc = read_scalars(D);
    if (rc) return rc;
    if (sumt) *sumt = D.sc_host->sumt;
    if (dts) *dts = D.sc_host->dts;
    return 0;
  });
}

int umwm_hip_substeps(umwm_domain* dom, int32_t n, double* sumt, double* dts, int32_t* nsub) {
  if (n < 1) return fail(UMWM_ERR_ARGUMENT, "substeps: n must be >= 1");
  return dispatch(dom, [&](auto& D) {
    if (int rc = substeps_impl(D, n)) return rc;
    if (sumt) *sumt = D.sc_host->sumt;
    if (dts) *dts = D.sc_host->dts;
    if (nsub) *nsub = D.sc_host->nsub;
    return 0;
  });
}

int umwm_hip_substeps_noalloc(umwm_domain* dom, int32_t n, double* sumt, double* dts,
                              int32_t* nsub) {
  if (n < 1) return fail(UMWM_ERR_ARGUMENT, "substeps_noalloc: n must be >= 1");
  return dispatch(dom, [&](auto& D) {
    if (int rc = substeps_noalloc_impl(D, n)) return rc;
    if (sumt) *sumt = D.sc_host->sumt;
    if (dts) *dts = D.sc_host->dts;
    if (nsub) *nsub = D.sc_host->nsub;
    return 0;
  });
}

int umwm_hip_run_dtg(umwm_domain* dom, int32_t* nsub) {
  return dispatch(dom, [&](auto& D) { return run_dtg_impl(D, nsub); });
}

int umwm_hip_get_field(umwm_domain* dom, const char* name, double* buf, int64_t n) {
  if (!name || !buf) return fail(UMWM_ERR_ARGUMENT, "null argument");
  return dispatch(dom, [&](auto& D) { return get_field_impl(D, name, buf, n); });
}

}  // extern "C"

namespace {
// ---------------------------------------------------------------------
// gridded output and restart files (raw container, rawio.hpp)
// ---------------------------------------------------------------------
// A field of the domain as doubles in the reference's shape, or zeros when
// it does not exist yet (the reference's module arrays start at zero).
template <typename T>
int fetch_field(DomainImpl<T>& D, const std::string& name, int64_t n, std::vector<double>& out) {
  FieldRef r;
  out.assign((size_t)n, 0.0);
  if (!find_field(D, name, r)) return 0;
  const int64_t have = field_size(D, r.kind);
  std::vector<double> tmp((size_t)have);
  if (int rc = get_field_impl(D, name, tmp.data(), have)) return rc;
  std::copy(tmp.begin(), tmp.begin() + std::min(have, n), out.begin());
  return 0;
}

// src/umwm_io.F90:317-854 output_grid_nc: every field on the (x, y) grid,
// sea points through the unrolled index and NaN on land (gatherfield,
// :857-901); lon, lat, seamask and depth as 2-D arrays.
template <typename T>
int write_output_impl(DomainImpl<T>& D, const std::string& path) {
  const HostSetup& S = D.S;
  const int64_t mm = S.mm, nm = S.nm, im = S.im, lm = std::max(S.lm, 1);
  const double nan = std::nan("");
  std::vector<RawVar> vars;
  auto add = [&](const char* name, const char* units, const char* desc, int32_t type,
                 std::vector<std::string> dims, std::vector<int64_t> shape,
                 std::vector<double> data) {
    RawVar v;
    v.name = name;
    v.units = units;
    v.description = desc;
    v.type = type;
    v.dims = std::move(dims);
    v.shape = std::move(shape);
    v.data = std::move(data);
    vars.push_back(std::move(v));
  };
  auto to_grid = [&](const std::vector<double>& sea, size_t off) {
    std::vector<double> g((size_t)(mm * nm), nan);
    for (int64_t j = 0; j < im; ++j)
      g[(size_t)((S.ni[j] - 1) * mm + (S.mi[j] - 1))] = sea[off + (size_t)j];
    return g;
  };
  std::vector<double> buf;
  auto grid_var = [&](const char* name, const char* src, const char* units, const char* desc) {
    if (int rc = fetch_field(D, src, im, buf)) return rc;
    add(name, units, desc, RAW_F32, {"x", "y", "time"}, {mm, nm, 1}, to_grid(buf, 0));
    return 0;
  };
  const bool stokes = S.P.stokes != 0;
  if (stokes) {
    std::vector<double> z((size_t)lm);
    for (int64_t l = 0; l < lm; ++l) z[(size_t)l] = S.depth_l[(size_t)l];  // -depths
    add("z", "m", "depth", RAW_F32, {"z"}, {lm}, z);
    for (const char* c : {"us", "vs"}) {
      if (int rc = fetch_field(D, c, im * lm, buf)) return rc;
      std::vector<double> g;
      for (int64_t l = 0; l < lm; ++l) {
        const auto gl = to_grid(buf, (size_t)(l * im));
        g.insert(g.end(), gl.begin(), gl.end());
      }
      add(c[0] == 'u' ? "u_stokes" : "v_stokes", "m/s",
          c[0] == 'u' ? "Stokes drift, x" : "Stokes drift, y", RAW_F32, {"x", "y", "z", "time"},
          {mm, nm, lm, 1}, g);
    }
    if (int rc = grid_var("d_stokes", "ds", "m", "Stokes drift e-folding depth")) return rc;
  }
  add("frequency", "hz", "frequency", RAW_F32, {"f"}, {(int64_t)S.om}, S.f);
  add("theta", "rad", "direction", RAW_F32, {"th"}, {(int64_t)S.pm}, S.th);
  add("lon", "degrees east", "longitude", RAW_F32, {"x", "y", "time"}, {mm, nm, 1}, S.lon);
  add("lat", "degrees north", "latitude", RAW_F32, {"x", "y", "time"}, {mm, nm, 1}, S.lat);
  add("seamask", "non-dimensional", "sea mask", RAW_I32, {"x", "y", "time"}, {mm, nm, 1},
      std::vector<double>(S.mask.begin(), S.mask.end()));
  add("depth", "m", "water depth", RAW_F32, {"x", "y", "time"}, {mm, nm, 1}, S.d2);
  struct G { const char *name, *src, *units, *desc; };
  static const G grid[] = {
      {"wspd", "wspd", "m/s", "wind speed"},
      {"wdir", "wdir", "rad", "wind direction"},
      {"uc", "uc", "m/s", "ocean current, x"},
      {"vc", "vc", "m/s", "ocean current, y"},
      {"rhoa", "rhoa", "kg/m^3", "air density"},
      {"rhow", "rhow", "kg/m^3", "water density"},
      {"fice", "fice", "non-dimensional", "sea ice fraction"},
      {"psim", "psim", "non-dimensional", "stability function (coupled runs only)"},
      {"momx", "momx", "kgm/s", "wave momentum, x"},
      {"momy", "momy", "kgm/s", "wave momentum, y"},
      {"cgmxx", "cgmxx", "kgm^2/s^2", "momentum flux xx"},
      {"cgmxy", "cgmxy", "kgm^2/s^2", "momentum flux xy"},
      {"cgmyy", "cgmyy", "kgm^2/s^2", "momentum flux yy"},
      {"shelt", "shelt", "non-dimensional", "sheltering coefficient"},
      {"epsx_atm", "epsx_atm", "kg/s^3", "energy flux from air, x"},
      {"epsy_atm", "epsy_atm", "kg/s^3", "energy flux from air, y"},
      {"epsx_ocn", "epsx_ocn", "kg/s^3", "energy flux into ocean, x"},
      {"epsy_ocn", "epsy_ocn", "kg/s^3", "energy flux into ocean, y"},
      {"taux_form", "taux_form", "n/m^2", "form drag, x"},
      {"tauy_form", "tauy_form", "n/m^2", "form drag, y"},
      {"taux_form_1", "taux1", "n/m^2", "form drag, wind pushing waves, x"},
      {"tauy_form_1", "tauy1", "n/m^2", "form drag, wind pushing waves, y"},
      {"taux_form_2", "taux2", "n/m^2", "form drag, waves against wind, x"},
      {"tauy_form_2", "tauy2", "n/m^2", "form drag, waves against wind, y"},
      {"taux_form_3", "taux3", "n/m^2", "form drag, waves overrunning wind, x"},
      {"tauy_form_3", "tauy3", "n/m^2", "form drag, waves overrunning wind, y"},
      {"taux_skin", "taux_skin", "n/m^2", "skin drag, x"},
      {"tauy_skin", "tauy_skin", "n/m^2", "skin drag, y"},
      {"taux_diag", "taux_diag", "n/m^2", "form drag of the diagnostic tail, x"},
      {"tauy_diag", "tauy_diag", "n/m^2", "form drag of the diagnostic tail, y"},
      {"taux_ocn", "taux_ocntop", "n/m^2", "stress into ocean top, x"},
      {"tauy_ocn", "tauy_ocntop", "n/m^2", "stress into ocean top, y"},
      {"taux_bot", "taux_ocnbot", "n/m^2", "stress into ocean bottom, x"},
      {"tauy_bot", "tauy_ocnbot", "n/m^2", "stress into ocean bottom, y"},
      {"taux_snl", "taux_snl", "n/m^2", "momentum loss by Snl, x"},
      {"tauy_snl", "tauy_snl", "n/m^2", "momentum loss by Snl, y"},
      {"tailatmx", "tailatmx", "n/m^2", "form drag tail, x"},
      {"tailatmy", "tailatmy", "n/m^2", "form drag tail, y"},
      {"tailocnx", "tailocnx", "n/m^2", "ocean stress tail, x"},
      {"tailocny", "tailocny", "n/m^2", "ocean stress tail, y"},
      {"cd", "cd", "non-dimensional", "drag coefficient"},
      {"ust", "ustar", "m/s", "friction velocity"},
      {"swh", "ht", "m", "significant wave height"},
      {"mss", "mss", "non-dimensional", "mean squared slope"},
      {"mwp", "mwp", "s", "mean wave period"},
      {"mwl", "mwl", "m", "mean wavelength"},
      {"mwd", "mwd", "rad", "mean wave direction"},
      {"dwp", "dwp", "s", "dominant wave period"},
      {"dwl", "dwl", "m", "dominant wavelength"},
      {"dwd", "dwd", "rad", "dominant wave direction"},
      {"dcp0", "dcp0", "m/s", "dominant phase speed, intrinsic"},
      {"dcg0", "dcg0", "m/s", "dominant group speed, intrinsic"},
      {"dcp", "dcp", "m/s", "dominant phase speed"},
      {"dcg", "dcg", "m/s", "dominant group speed"},
      {"physics_time_step", "physics_time_step", "s", "physics time step"}};
  for (const G& g : grid)
    if (int rc = grid_var(g.name, g.src, g.units, g.desc)) return rc;
  std::string err;
  if (raw_write(path, vars, err)) return fail(UMWM_ERR_IO, err);
  return 0;
}

// src/umwm_restart.F90:55-145 restart_write: per sea point lon, lat, ust,
// wavenumber(f,x) and the spectrum F(f,th,x), single precision as there.
template <typename T>
int write_restart_impl(DomainImpl<T>& D, const std::string& path) {
  const HostSetup& S = D.S;
  const int64_t im = S.im, om = S.om, pm = S.pm;
  std::vector<RawVar> vars(7);
  std::vector<double> lon((size_t)im), lat((size_t)im);
  for (int64_t j = 0; j < im; ++j) {
    const size_t q = (size_t)((S.ni[j] - 1) * S.mm + (S.mi[j] - 1));
    lon[(size_t)j] = S.lon[q];
    lat[(size_t)j] = S.lat[q];
  }
  auto set = [&](int k, const char* name, const char* units, const char* desc,
                 std::vector<std::string> dims, std::vector<int64_t> shape,
                 std::vector<double> data) {
    vars[k].name = name;
    vars[k].units = units;
    vars[k].description = desc;
    vars[k].type = RAW_F32;
    vars[k].dims = std::move(dims);
    vars[k].shape = std::move(shape);
    vars[k].data = std::move(data);
  };
  std::vector<double> ust, kk, e;
  int rc = fetch_field(D, "ustar", im, ust);
  if (!rc) rc = fetch_field(D, "k", om * im, kk);
  if (!rc) rc = fetch_field(D, "e", om * pm * im, e);
  if (rc) return rc;
  set(0, "lon", "degrees east", "longitude", {"x"}, {im}, lon);
  set(1, "lat", "degrees north", "latitude", {"x"}, {im}, lat);
  set(2, "frequency", "hz", "frequency", {"f"}, {om}, S.f);
  set(3, "theta", "rad", "direction", {"th"}, {pm}, S.th);
  set(4, "ust", "m/s", "friction velocity", {"x"}, {im}, ust);
  set(5, "wavenumber", "rad/m", "wavenumber", {"f", "x"}, {om, im}, kk);
  set(6, "F", "m^2/rad/(rad/m)", "wave variance spectrum", {"f", "th", "x"}, {om, pm, im}, e);
  std::string err;
  if (raw_write(path, vars, err)) return fail(UMWM_ERR_IO, err);
  return 0;
}

// src/umwm_restart.F90:18-52 restart_read: e and ustar of the sea points.
template <typename T>
int read_restart_impl(DomainImpl<T>& D, const std::string& path) {
  const HostSetup& S = D.S;
  std::vector<RawVar> vars;
  std::string err;
  if (raw_read(path, vars, err)) return fail(UMWM_ERR_IO, err);
  const RawVar *F = nullptr, *u = nullptr;
  for (const RawVar& v : vars) {
    if (v.name == "F") F = &v;
    if (v.name == "ust") u = &v;
  }
  if (!F || !u) return fail(UMWM_ERR_IO, path + ": no F or ust variable");
  const std::vector<int64_t> want_f = {S.om, S.pm, S.im}, want_u = {S.im};
  if (F->shape != want_f || u->shape != want_u)
    return fail(UMWM_ERR_SIZE, path + ": F or ust does not match the domain's (om, pm, im)");
  int rc = set_field_impl(D, "e", F->data.data(), (int64_t)F->data.size());
  if (!rc) rc = set_field_impl(D, "ustar", u->data.data(), (int64_t)u->data.size());
  return rc;
}

}  // namespace

extern "C" {

int umwm_hip_write_output(umwm_domain* dom, const char* path) {
  if (!path) return fail(UMWM_ERR_ARGUMENT, "null path");
  return dispatch(dom, [&](auto& D) { return write_output_impl(D, path); });
}

int umwm_hip_write_restart(umwm_domain* dom, const char* path) {
  if (!path) return fail(UMWM_ERR_ARGUMENT, "null path");
  return dispatch(dom, [&](auto& D) { return write_restart_impl(D, path); });
}

int umwm_hip_read_restart(umwm_domain* dom, const char* path) {
  if (!path) return fail(UMWM_ERR_ARGUMENT, "null path");
  return dispatch(dom, [&](auto& D) { return read_restart_impl(D, path); });
}

int umwm_hip_set_field(umwm_domain* dom, const char* name, const double* buf, int64_t n) {
  if (!name || !buf) return fail(UMWM_ERR_ARGUMENT, "null argument");
  return dispatch(dom, [&](auto& D) { return set_field_impl(D, name, buf, n); });
}

int64_t umwm_hip_field_size(const umwm_domain* dom, const char* name) {
  if (!dom || !dom->base || !name) return 0;
  umwm_domain* m = const_cast<umwm_domain*>(dom);
  int64_t out = 0;
  dispatch(m, [&](auto& D) {
    FieldRef r;
    if (find_field(D, name, r)) out = field_size(D, r.kind);
    return 0;
  });
  return out;
}

int umwm_hip_get_scalar(umwm_domain* dom, const char* name, double* v) {
  if (!name || !v) return fail(UMWM_ERR_ARGUMENT, "null argument");
  return dispatch(dom, [&](auto& D) {
    const std::string n(name);
    const HostSetup& S = D.S;
    if (n == "dtamin") { *v = S.dtamin; return 0; }
    if (n == "cfllim") { *v = S.cfllim; return 0; }
    if (n == "cgmax") { *v = S.cgmax; return 0; }
    if (n == "mindelx") { *v = S.mindelx; return 0; }
    if (n == "dlnf") { *v = S.dlnf; return 0; }
    if (n == "dth") { *v = S.dth; return 0; }
    if (n == "bf1") { *v = S.bf1; return 0; }
    if (n == "bf2") { *v = S.bf2; return 0; }
    if (n == "first") { *v = D.first ? 1 : 0; return 0; }
    int rc = read_scalars(D);
    if (rc) return rc;
    const Scalars& s = *D.sc_host;
    if (n == "sumt") { *v = s.sumt; return 0; }
    if (n == "dts") { *v = s.dts; return 0; }
    if (n == "dta") { *v = s.dta; return 0; }
    if (n == "dtr") { *v = s.dtr; return 0; }
    if (n == "dtr_temp") { *v = s.dtr_temp; return 0; }
    if (n == "maxrot") { *v = s.maxrot; return 0; }
    if (n == "nsub") { *v = s.nsub; return 0; }
    if (n == "zerocurrents") { *v = s.anycur ? 0 : 1; return 0; }
    return fail(UMWM_ERR_UNKNOWN_FIELD, "unknown scalar: " + n);
  });
}

int umwm_hip_set_scalar(umwm_domain* dom, const char* name, double v) {
  if (!name) return fail(UMWM_ERR_ARGUMENT, "null argument");
  return dispatch(dom, [&](auto& D) {
    const std::string n(name);
    if (n == "first") { D.first = v != 0; return 0; }
    if (n == "keep_sources") return v != 0 ? enable_keep(D) : 0;
    int rc = read_scalars(D);
    if (rc) return rc;
    Scalars s = *D.sc_host;
    if (n == "sumt") s.sumt = v;
    else if (n == "dts") s.dts = v;
    else if (n == "dtr_temp") s.dtr_temp = v;
    else return fail(UMWM_ERR_UNKNOWN_FIELD, "scalar not settable: " + n);
    UMWM_HIP_TRY(hipMemcpy(D.h.sc, &s, sizeof(Scalars), hipMemcpyHostToDevice));
    return 0;
  });
}

int umwm_hip_set_timing(umwm_domain* dom, int enable) {
  return dispatch(dom, [&](auto& D) {
    D.timing = enable != 0;
    return 0;
  });
}

int umwm_hip_get_timing(umwm_domain* dom, double ms[5], int64_t counts[5], int reset) {
  return dispatch(dom, [&](auto& D) {
    for (int q = 0; q < 5; ++q) {
      if (ms) ms[q] = D.ms[q];
      if (counts) counts[q] = D.launches[q];
      if (reset) {
        D.ms[q] = 0;
        D.launches[q] = 0;
      }
    }
    return 0;
  });
}

int umwm_hip_synchronize(umwm_domain* dom) {
  return dispatch(dom, [&](auto& D) { return sync_and_check(D); });
}

int umwm_hip_set_tile(umwm_domain* dom, const int32_t tile[5]) {
  if (!tile) return fail(UMWM_ERR_ARGUMENT, "null tile");
  return dispatch(dom, [&](auto& D) {
    Tile t = tile_from(tile, D.S.pm);
    if (!tile_ok(t, D.S.pm))
      return fail(UMWM_ERR_ARGUMENT, "tile needs D1 in {6,9,18,36}, D2 in {3,6,9,12,18}, "
                                     "W in 1..8, D*W >= pm, NB in {2,4}");
    D.tile = t;
    return 0;
  });
}

int umwm_hip_get_tile(const umwm_domain* dom, int32_t tile[5]) {
  if (!dom || !dom->base || !tile) return fail(UMWM_ERR_ARGUMENT, "null argument");
  umwm_domain* m = const_cast<umwm_domain*>(dom);
  return dispatch(m, [&](auto& D) {
    tile_to(D.tile, tile);
    return 0;
  });
}

}  // extern "C"

// ---------------------------------------------------------------------
// batch entry points
// ---------------------------------------------------------------------
namespace {

template <typename T>
int batch_create_impl(umwm_domain* const* doms, int n, BatchImpl<T>*& out) {
  auto* B = new BatchImpl<T>();
  int pmmax = 0;
  for (int k = 0; k < n; ++k) {
    auto* D = static_cast<DomainImpl<T>*>(doms[k]->base);
    B->doms.push_back(D);
    pmmax = std::max(pmmax, D->S.pm);
    B->omax = std::max(B->omax, D->S.om);
  }
  B->tile = choose_tile(pmmax);
  if (hipStreamCreateWithFlags(&B->stream, hipStreamNonBlocking) != hipSuccess) {
    delete B;
    return fail(UMWM_ERR_HIP, "hipStreamCreate failed");
  }
  // domain structs; one task list per pm class, largest domains first and
  // XCD-interleaved within the class
  std::vector<DomainDev<T>> hd(n);
  for (int k = 0; k < n; ++k) hd[k] = B->doms[k]->h;
  std::vector<int> order(n);
  for (int k = 0; k < n; ++k) order[k] = k;
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) {
    return (size_t)B->doms[a]->S.im * B->doms[a]->S.om * B->doms[a]->S.pm >
           (size_t)B->doms[b]->S.im * B->doms[b]->S.om * B->doms[b]->S.pm;
  });
  std::vector<int> pms;
  for (int k : order)
    if (std::find(pms.begin(), pms.end(), B->doms[k]->S.pm) == pms.end())
      pms.push_back(B->doms[k]->S.pm);
  B->d_doms = B->arena.template alloc<DomainDev<T>>(n);
  B->d_scs = B->arena.template alloc<Scalars*>(n);
  B->d_view = B->arena.template alloc<ScalarsView>(n);
  B->d_fds = B->arena.template alloc<ForcingDev<T>>(n);
  B->d_flist = B->arena.template alloc<int32_t>(n);
  if (!B->d_view || !B->d_fds || !B->d_flist) {
    delete B;
    return fail(UMWM_ERR_NOMEM, "batch allocation failed");
  }
  for (int pm : pms) {
    typename BatchImpl<T>::Group g;
    g.pm = pm;
    std::vector<BlockTask> tk;
    for (int k : order) {
      if (B->doms[k]->S.pm != pm) continue;
      g.omax = std::max(g.omax, B->doms[k]->S.om);
      for (int b = 0; b < hd[k].nblocks; ++b) tk.push_back(BlockTask{k, b * kLanes, b, 0});
    }
    std::vector<int> ord = xcd_order((int)tk.size());
    std::vector<BlockTask> tk2(tk.size());
    for (size_t b = 0; b < tk.size(); ++b) tk2[b] = tk[ord[b]];
    g.ntasks = (int)tk2.size();
    g.tasks = B->arena.template alloc<BlockTask>(std::max(g.ntasks, 1));
    if (!g.tasks) {
      delete B;
      return fail(UMWM_ERR_NOMEM, "batch allocation failed");
    }
    UMWM_HIP_TRY(hipMemcpy(g.tasks, tk2.data(), tk2.size() * sizeof(BlockTask), hipMemcpyHostToDevice));
    g.tile = choose_tile(pm);
    B->groups.push_back(g);
  }
  std::vector<Scalars*> scs(n);
  for (int k = 0; k < n; ++k) scs[k] = hd[k].sc;
  std::vector<int32_t> fl;
  for (int k = 0; k < n; ++k)
    if (any_forcing(*B->doms[k])) {
      fl.push_back(k);
      B->maximm = std::max(B->maximm, B->doms[k]->S.imm);
      B->forced_currents |= B->doms[k]->S.P.currents != 0;
    }
  B->nforced = (int)fl.size();
  if (!fl.empty())
    UMWM_HIP_TRY(hipMemcpy(B->d_flist, fl.data(), fl.size() * sizeof(int32_t), hipMemcpyHostToDevice));
  UMWM_HIP_TRY(hipMemcpy(B->d_doms, hd.data(), n * sizeof(DomainDev<T>), hipMemcpyHostToDevice));
  UMWM_HIP_TRY(hipMemcpy(B->d_scs, scs.data(), n * sizeof(Scalars*), hipMemcpyHostToDevice));
  UMWM_HIP_TRY(hipHostMalloc((void**)&B->h_view, n * sizeof(ScalarsView), hipHostMallocDefault));
  out = B;
  return 0;
}

// Runs a per-domain launcher on another stream (batch members share the
// batch stream so a whole substep can be captured as one graph).
template <typename T>
int on_stream(DomainImpl<T>& D, hipStream_t s, int (*fn)(DomainImpl<T>&)) {
  hipStream_t keep = D.stream;
  D.stream = s;
  const int rc = fn(D);
  D.stream = keep;
  return rc;
}

// Uploads every member's current domain and forcing records (a member's
// record changes with keep_sources, its operator arrays and forcinginput)
// and drops the captured graph when the batch's keep flag changed.
template <typename T>
int batch_refresh(BatchImpl<T>& B) {
  const int n = (int)B.doms.size();
  std::vector<DomainDev<T>> hd(n);
  std::vector<ForcingDev<T>> hf(n);
  bool keep = false;
  for (int k = 0; k < n; ++k) {
    hd[k] = B.doms[k]->h;
    hf[k] = B.doms[k]->fh;
    keep |= B.doms[k]->h.e_keep != nullptr;
  }
  UMWM_HIP_TRY(hipMemcpyAsync(B.d_doms, hd.data(), n * sizeof(DomainDev<T>), hipMemcpyHostToDevice, B.stream));
  UMWM_HIP_TRY(hipMemcpyAsync(B.d_fds, hf.data(), n * sizeof(ForcingDev<T>), hipMemcpyHostToDevice, B.stream));
  UMWM_HIP_TRY(hipStreamSynchronize(B.stream));
  B.keep = keep;
  if (B.exec && B.keep_captured != keep) {
    (void)hipGraphExecDestroy(B.exec);
    (void)hipGraphDestroy(B.graph);
    B.exec = nullptr;
    B.graph = nullptr;
  }
  return 0;
}

// One substep of every member: the forcing of all forced members in one
// launch, then the fused kernels once per pm class.  `first` = 1 takes the
// cold first substep with the rotation tendency; afterwards the kernels
// refresh the rotation only for members with currents.
template <typename T>
int batch_launch_substep(BatchImpl<T>& B, int first) {
  if (int rc = launch_forcing_list<T>(B.d_doms, B.d_fds, B.d_flist, B.nforced, B.maximm,
                                      B.forced_currents, true, B.stream))
    return rc;
  bool any_sur = false;
  for (auto* D : B.doms) {
    bool on = false;
    if (int rc = surrogate_rates(*D, B.stream, on)) return rc;
    any_sur |= on;
  }
  for (const auto& g : B.groups)
    if (int rc = launch_fused<T>(B.d_doms, g.tasks, g.ntasks, g.omax, first, first, g.tile,
                                 B.keep, B.stream, nullptr, any_sur))
      return rc;
  for (auto* D : B.doms) D->ef_mirrors_e = true;
  const int n = (int)B.doms.size();
  hipLaunchKernelGGL(k_gather_scalars, dim3((n + 63) / 64), dim3(64), 0, B.stream,
                     (Scalars* const*)B.d_scs, n, B.d_view);
  UMWM_HIP_TRY(hipGetLastError());
  return 0;
}

// Members still in their first substep take it eagerly (cold dts = 0,
// rotation tendency); then the steady-state substep is captured once.
template <typename T>
int batch_prepare(BatchImpl<T>& B, int use_graph) {
  int rc;
  bool any_first = false;
  for (auto* D : B.doms) any_first |= D->first;
  if (any_first) {
    for (auto* D : B.doms)
      if (!D->first) return fail(UMWM_ERR_STATE, "batch members must start together");
    if ((rc = batch_launch_substep(B, 1))) return rc;
    for (auto* D : B.doms) {
      D->first = false;
      D->rot_valid = true;
      if ((rc = on_stream(*D, B.stream, launch_diag<T>))) return rc;
    }
  }
  if (use_graph && !B.exec) {
    UMWM_HIP_TRY(hipStreamBeginCapture(B.stream, hipStreamCaptureModeThreadLocal));
    rc = batch_launch_substep(B, 0);
    B.keep_captured = B.keep;
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(B.stream, &g);
    if (rc) return rc;
    if (e != hipSuccess) return fail(UMWM_ERR_HIP, std::string("graph capture: ") + hipGetErrorString(e));
    B.graph = g;
    UMWM_HIP_TRY(hipGraphInstantiate(&B.exec, g, nullptr, nullptr, 0));
  }
  return 0;
}

// n substeps of every member, launched back to back (graph replays) with
// one host wait at the end; a member whose sumt reached dtg does nothing.
// nsub[k] receives member k's substeps taken this dtg.
template <typename T>
int batch_substeps_impl(BatchImpl<T>& B, int n, int use_graph, int32_t* nsub) {
  const int nd = (int)B.doms.size();
  for (auto* D : B.doms) UMWM_HIP_TRY(hipStreamSynchronize(D->stream));
  int rc;
  if ((rc = batch_refresh(B))) return rc;
  if ((rc = batch_prepare(B, use_graph))) return rc;
  for (int c = 0; c < n; ++c) {
    if (use_graph) {
      UMWM_HIP_TRY(hipGraphLaunch(B.exec, B.stream));
    } else if ((rc = batch_launch_substep(B, 0))) {
      return rc;
    }
  }
  UMWM_HIP_TRY(hipMemcpyAsync(B.h_view, B.d_view, nd * sizeof(ScalarsView), hipMemcpyDeviceToHost, B.stream));
  UMWM_HIP_TRY(hipStreamSynchronize(B.stream));
  if (nsub)
    for (int k = 0; k < nd; ++k) nsub[k] = B.h_view[k].nsub;
  return 0;
}

template <typename T>
int batch_run_dtg_impl(BatchImpl<T>& B, int use_graph, int32_t* nsub) {
  const int n = (int)B.doms.size();
  for (auto* D : B.doms) UMWM_HIP_TRY(hipStreamSynchronize(D->stream));
  int rc;
  if ((rc = batch_refresh(B))) return rc;
  for (auto* D : B.doms)
    hipLaunchKernelGGL(k_begin_dtg, dim3(1), dim3(64), 0, B.stream, D->h.sc);
  UMWM_HIP_TRY(hipGetLastError());
  if ((rc = batch_prepare(B, use_graph))) return rc;
  // replay in chunks; finished members exit their blocks at once
  int chunk = 4, total = 0;
  for (;;) {
    for (int c = 0; c < chunk; ++c) {
      if (use_graph) {
        UMWM_HIP_TRY(hipGraphLaunch(B.exec, B.stream));
      } else if ((rc = batch_launch_substep(B, 0))) {
        return rc;
      }
    }
    total += chunk;
    UMWM_HIP_TRY(hipMemcpyAsync(B.h_view, B.d_view, n * sizeof(ScalarsView), hipMemcpyDeviceToHost, B.stream));
    UMWM_HIP_TRY(hipStreamSynchronize(B.stream));
    bool done = true;
    for (int k = 0; k < n; ++k) done &= !(B.h_view[k].sumt < B.doms[k]->S.P.dtg);
    if (done) break;
    if (total > 10000000) return fail(UMWM_ERR_STATE, "batch run_dtg: substep count runaway");
    chunk = std::min(chunk * 2, 32);
  }
  for (auto* D : B.doms) {
    if (D->firstdtg) D->firstdtg = false;
    if (D->S.P.stokes && (rc = on_stream(*D, B.stream, launch_stokes<T>))) return rc;
    if ((rc = on_stream(*D, B.stream, launch_diag<T>))) return rc;
  }
  UMWM_HIP_TRY(hipStreamSynchronize(B.stream));
  if (nsub)
    for (int k = 0; k < n; ++k) nsub[k] = B.h_view[k].nsub;
  return 0;
}

}  // namespace

extern "C" {

int umwm_hip_batch_create(umwm_domain* const* doms, int32_t n, umwm_batch** out) {
  if (!doms || n < 1 || !out) return fail(UMWM_ERR_ARGUMENT, "bad batch arguments");
  const int prec = doms[0] && doms[0]->base ? doms[0]->base->precision : 0;
  for (int k = 0; k < n; ++k)
    if (!doms[k] || !doms[k]->base || doms[k]->base->precision != prec)
      return fail(UMWM_ERR_ARGUMENT, "batch members must exist and share the precision");
  auto* b = new umwm_batch();
  b->precision = prec;
  int rc;
  if (prec == 4) {
    BatchImpl<float>* B = nullptr;
    rc = batch_create_impl<float>(doms, n, B);
    b->impl = B;
  } else {
    BatchImpl<double>* B = nullptr;
    rc = batch_create_impl<double>(doms, n, B);
    b->impl = B;
  }
  if (rc) {
    delete b;
    return rc;
  }
  *out = b;
  return 0;
}

int umwm_hip_batch_destroy(umwm_batch* b) {
  if (!b) return 0;
  if (b->precision == 4) delete static_cast<BatchImpl<float>*>(b->impl);
  else delete static_cast<BatchImpl<double>*>(b->impl);
  delete b;
  return 0;
}

int umwm_hip_batch_run_dtg(umwm_batch* b, int use_graph, int32_t* nsub) {
  if (!b) return fail(UMWM_ERR_ARGUMENT, "null batch");
  if (b->precision == 4) return batch_run_dtg_impl(*static_cast<BatchImpl<float>*>(b->impl), use_graph, nsub);
  return batch_run_dtg_impl(*static_cast<BatchImpl<double>*>(b->impl), use_graph, nsub);
}

int umwm_hip_batch_substeps(umwm_batch* b, int32_t n, int use_graph, int32_t* nsub) {
  if (!b) return fail(UMWM_ERR_ARGUMENT, "null batch");
  if (n < 1) return fail(UMWM_ERR_ARGUMENT, "batch_substeps: n must be >= 1");
  if (b->precision == 4)
    return batch_substeps_impl(*static_cast<BatchImpl<float>*>(b->impl), n, use_graph, nsub);
  return batch_substeps_impl(*static_cast<BatchImpl<double>*>(b->impl), n, use_graph, nsub);
}

int umwm_hip_batch_tile(const umwm_batch* b, int32_t out[5]) {
  if (!b || !out) return fail(UMWM_ERR_ARGUMENT, "null argument");
  const Tile t = b->precision == 4 ? static_cast<BatchImpl<float>*>(b->impl)->tile
                                   : static_cast<BatchImpl<double>*>(b->impl)->tile;
  tile_to(t, out);
  return 0;
}

int umwm_hip_batch_set_tile(umwm_batch* b, const int32_t tile[5]) {
  if (!b || !tile) return fail(UMWM_ERR_ARGUMENT, "null argument");
  auto apply = [&](auto* B) {
    int pmmax = 0;
    for (auto* D : B->doms) pmmax = std::max(pmmax, D->S.pm);
    Tile t = tile_from(tile, pmmax);
    if (!tile_ok(t, pmmax)) return fail(UMWM_ERR_ARGUMENT, "tile does not cover pm");
    B->tile = t;
    for (auto& g : B->groups) {
      Tile gt = t;  // same D and NB, waves for this class's pm
      gt.W1 = (g.pm + t.D1 - 1) / t.D1;
      gt.W2 = (g.pm + t.D2 - 1) / t.D2;
      gt.pmax = g.pm;
      g.tile = tile_ok(gt, g.pm) ? gt : choose_tile(g.pm);
    }
    if (B->exec) {  // the captured graph holds the old launch shape
      (void)hipGraphExecDestroy(B->exec);
      (void)hipGraphDestroy(B->graph);
      B->exec = nullptr;
      B->graph = nullptr;
    }
    return 0;
  };
  if (b->precision == 4) return apply(static_cast<BatchImpl<float>*>(b->impl));
  return apply(static_cast<BatchImpl<double>*>(b->impl));
}

}  // extern "C"

// ---------------------------------------------------------------------
// surrogate entry points that take a domain
// ---------------------------------------------------------------------
namespace {

__global__ void k_zero_f64(double* p, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0;
}

int attach_impl(DomainImpl<float>& D, SurrogateImpl* S, const umwm_surrogate_policy* pol) {
  UMWM_HIP_TRY(hipStreamSynchronize(D.stream));
  if (!S) {
    D.sur = nullptr;
    D.h.sur_r = nullptr;
    UMWM_HIP_TRY(hipMemcpy(D.d, &D.h, sizeof(DomainDev<float>), hipMemcpyHostToDevice));
    return 0;
  }
  if (S->om != D.h.om || S->pm != D.h.pm || D.S.P.fmin != S->fmin || D.S.P.fmax != S->fmax)
    return fail(UMWM_ERR_ARGUMENT,
                "surrogate_attach: the domain's spectral grid (om, pm, fmin, fmax) differs from "
                "the network's");
  if (!S->stats_ready) return fail(UMWM_ERR_STATE, "surrogate_attach: train the network first");
  if (!D.h.sur_r) {
    const size_t nb = (size_t)D.h.om * D.h.pm;
    uint16_t* r = D.arena.template alloc<uint16_t>(nb * D.h.ldi);
    int32_t* ok = D.arena.template alloc<int32_t>(D.h.ldi);
    int32_t* use = D.arena.template alloc<int32_t>(D.h.ldi);
    double* st = D.arena.template alloc<double>(SS_N);
    if (!r || !ok || !use || !st) return fail(UMWM_ERR_NOMEM, "surrogate_attach: allocation failed");
    D.h.sur_r = r;
    D.h.sur_ok = ok;
    D.h.sur_use = use;
    D.h.sur_stats = st;
    if (int rc = dfill<int32_t>(ok, D.h.ldi, 0, D.stream)) return rc;
    if (int rc = dfill<int32_t>(use, D.h.ldi, 0, D.stream)) return rc;
    hipLaunchKernelGGL(k_zero_f64, dim3(1), dim3(64), 0, D.stream, st, (int)SS_N);
    UMWM_HIP_TRY(hipGetLastError());
  }
  umwm_surrogate_policy p{};
  p.audit_stride = 16;
  p.resid_tol = 0.05;
  p.audit_tol = 0.25;
  p.range_margin = 0.1;
  p.zmax = 8.0;
  if (pol) p = *pol;
  if (p.audit_stride < 0 || !(p.resid_tol >= 0) || !(p.audit_tol >= 0) || !(p.range_margin >= 0) ||
      !(p.zmax > 0))
    return fail(UMWM_ERR_ARGUMENT, "surrogate policy: audit_stride >= 0, tolerances >= 0, zmax > 0");
  D.h.sur_audit_stride = p.audit_stride;
  D.h.sur_resid_tol = (float)p.resid_tol;
  D.h.sur_audit_tol = (float)p.audit_tol;
  D.sur_margin = p.range_margin;
  D.sur_zmax = p.zmax;
  D.sur = S;
  UMWM_HIP_TRY(hipMemcpyAsync(D.d, &D.h, sizeof(DomainDev<float>), hipMemcpyHostToDevice, D.stream));
  // the audit gate starts open, with empty sums
  struct {
    double anum, aden;
    int32_t closed, pad;
  } gate{0.0, 0.0, 0, 0};
  static_assert(sizeof(gate) == sizeof(Scalars) - offsetof(Scalars, sur_anum), "gate fields");
  UMWM_HIP_TRY(hipMemcpyAsync(reinterpret_cast<char*>(D.h.sc) + offsetof(Scalars, sur_anum), &gate,
                              sizeof(gate), hipMemcpyHostToDevice, D.stream));
  UMWM_HIP_TRY(hipStreamSynchronize(D.stream));
  return 0;
}

}  // namespace

extern "C" {

int umwm_hip_surrogate_collect(umwm_surrogate* s, umwm_domain* dom, int64_t* total) {
  if (!s || !s->impl) return fail(UMWM_ERR_ARGUMENT, "null surrogate");
  return dispatch(dom, [&](auto& D) {
    if (int rc = ensure_ops(D)) return rc;
    return surrogate_collect(*s->impl, D, total);
  });
}

int umwm_hip_surrogate_attach(umwm_domain* dom, umwm_surrogate* s, const umwm_surrogate_policy* pol) {
  if (!dom || !dom->base) return fail(UMWM_ERR_ARGUMENT, "null domain");
  if (dom->base->precision != 4)
    return fail(UMWM_ERR_ARGUMENT, "surrogate_attach: the surrogate serves FP32 domains");
  return attach_impl(*static_cast<DomainImpl<float>*>(dom->base), s ? s->impl : nullptr, pol);
}

int umwm_hip_surrogate_fields(umwm_domain* dom, double* dm_net, double* use) {
  if (!dom || !dom->base) return fail(UMWM_ERR_ARGUMENT, "null domain");
  if (dom->base->precision != 4) return fail(UMWM_ERR_STATE, "surrogate_fields: no surrogate");
  auto& D = *static_cast<DomainImpl<float>*>(dom->base);
  if (!D.h.sur_r) return fail(UMWM_ERR_STATE, "surrogate_fields: no surrogate attached");
  UMWM_HIP_TRY(hipStreamSynchronize(D.stream));
  const size_t ldi = D.h.ldi, im = D.h.im, om = D.h.om, pm = D.h.pm;
  if (dm_net) {
    std::vector<uint16_t> h(om * pm * ldi);
    UMWM_HIP_TRY(hipMemcpy(h.data(), D.h.sur_r, h.size() * 2, hipMemcpyDeviceToHost));
    for (size_t o = 0; o < om; ++o)
      for (size_t p = 0; p < pm; ++p)
        for (size_t j = 0; j < im; ++j) {
          const uint32_t bits = (uint32_t)h[(o * pm + p) * ldi + j] << 16;
          float v;
          std::memcpy(&v, &bits, 4);
          dm_net[j * om * pm + p * om + o] = 1.0 + (double)v;  // stored as dm^2 - 1
        }
  }
  if (use) {
    std::vector<int32_t> u(im);
    UMWM_HIP_TRY(hipMemcpy(u.data(), D.h.sur_use, im * 4, hipMemcpyDeviceToHost));
    for (size_t j = 0; j < im; ++j) use[j] = u[j];
  }
  return 0;
}

int umwm_hip_surrogate_stats(umwm_domain* dom, double out[7], int reset) {
  if (!dom || !dom->base || !out) return fail(UMWM_ERR_ARGUMENT, "null argument");
  if (dom->base->precision != 4) return fail(UMWM_ERR_STATE, "surrogate_stats: no surrogate");
  auto& D = *static_cast<DomainImpl<float>*>(dom->base);
  if (!D.h.sur_stats) return fail(UMWM_ERR_STATE, "surrogate_stats: no surrogate attached");
  double st[SS_N];
  UMWM_HIP_TRY(hipStreamSynchronize(D.stream));
  UMWM_HIP_TRY(hipMemcpy(st, D.h.sur_stats, sizeof(st), hipMemcpyDeviceToHost));
  // words SS_POINTS .. SS_AUDITED hold unsigned 64-bit counts
  for (int k = 0; k < SS_AUDIT_NUM; ++k) {
    unsigned long long c;
    std::memcpy(&c, st + k, sizeof(c));
    out[k] = (double)c;
  }
  out[6] = st[SS_AUDIT_DEN] > 0 ? st[SS_AUDIT_NUM] / st[SS_AUDIT_DEN] : 0.0;
  if (reset) {
    hipLaunchKernelGGL(k_zero_f64, dim3(1), dim3(64), 0, D.stream, D.h.sur_stats, (int)SS_N);
    UMWM_HIP_TRY(hipGetLastError());
    UMWM_HIP_TRY(hipStreamSynchronize(D.stream));
  }
  return 0;
}

}  // extern "C"
