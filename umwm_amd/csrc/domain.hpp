// Host-side owner of one domain's device state (the opaque umwm_domain).
#pragma once

#include <hip/hip_runtime.h>

#include <memory>
#include <string>
#include <vector>

#include "common.hpp"
#include "setup.hpp"

namespace umwm {

// Tile shape of the frequency-sweeping fused kernels (k_rates: D1, W1;
// k_integrate: D2, W2): W waves per workgroup, each thread owns D
// consecutive directions of one sea point; 64 sea points per workgroup.
struct Tile {
  int D1 = 9, W1 = 4;
  int D2 = 9, W2 = 4;
  int NB = 2;     // advect: directions per prefetch block
  // pm of the domains launched together (a batch launches each pm class
  // on its own); W D = pmax selects the sweeps' mask-free variants
  int pmax = 36;
};

bool tile_supported(int D);

struct SurrogateImpl;  // surrogate.hpp

class DeviceArena {
 public:
  ~DeviceArena() { release(); }
  template <typename U>
  U* alloc(size_t n) {
    void* p = nullptr;
    if (n == 0) n = 1;
    if (hipMalloc(&p, n * sizeof(U)) != hipSuccess) return nullptr;
    ptrs_.push_back(p);
    bytes_ += n * sizeof(U);
    return static_cast<U*>(p);
  }
  void release() {
    for (void* p : ptrs_) (void)hipFree(p);
    ptrs_.clear();
    bytes_ = 0;
  }
  size_t bytes() const { return bytes_; }

 private:
  std::vector<void*> ptrs_;
  size_t bytes_ = 0;
};

struct DomainBase {
  virtual ~DomainBase() = default;
  int precision = 8;
  HostSetup S;
  hipStream_t stream = nullptr;
  bool first = true;      // src/umwm_module.F90:47
  bool firstdtg = true;
  bool timing = false;
  // the fused substep does not store E_s: after it, ef equals e (as after
  // the reference's e = ef) and reads of "ef" are served from e
  bool ef_mirrors_e = false;
  // rates, integrate, advect, stress, surrogate inference
  double ms[5] = {0, 0, 0, 0, 0};
  int64_t launches[5] = {0, 0, 0, 0, 0};
};

template <typename T>
struct DomainImpl : DomainBase {
  DeviceArena arena;
  DomainDev<T> h{};           // host mirror of the device struct
  DomainDev<T>* d = nullptr;  // device copy (array of one)
  ForcingDev<T> fh{};
  // device copy of fh (array of one), re-uploaded when forcinginput swaps
  // the levels, and the list {0} that k_forcing's one-domain launch reads
  ForcingDev<T>* fd = nullptr;
  int32_t* list0 = nullptr;
  // umwm_hip_substeps_noalloc: the device's shared scratch array (from
  // hipGetSymbolAddress at creation) and a second device copy of the domain
  // record whose pfx and spart point into it, with its host image
  unsigned char* scr = nullptr;
  DomainDev<T>* d_scr = nullptr;
  DomainDev<T> h_scr{};
  // attached learned surrogate (borrowed; FP32 domains only)
  SurrogateImpl* sur = nullptr;
  double sur_margin = 0.1, sur_zmax = 8.0;
  // one steady-state substep captured as a hipGraph (run_dtg and
  // substeps replay it); `gsig` identifies what was captured
  hipGraph_t graph = nullptr;
  hipGraphExec_t gexec = nullptr;
  uint64_t gsig = 0;
  bool use_graph = true;
  DiagDev diag{};
  OcnDev ocn{};  // allocated with the operator arrays
  StokesDev stokes{};
  BlockTask* tasks = nullptr;
  int ntasks = 0;
  Tile tile;
  Scalars* sc_host = nullptr;  // pinned mirror
  bool ops_ready = false;      // operator-path arrays allocated
  bool rot_valid = false;      // rotation computed at least once
  // ev[0..4]: around the fused kernels; ev[5], ev[6]: around the inference
  hipEvent_t ev[7] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
  // forcing device buffers (level n and n+1)
  T *uwb = nullptr, *uwf = nullptr, *vwb = nullptr, *vwf = nullptr;
  T *ucb = nullptr, *ucf = nullptr, *vcb = nullptr, *vcf = nullptr;
  T *ficeb = nullptr, *ficef = nullptr;
  T *rhoab = nullptr, *rhoaf = nullptr, *rhowb = nullptr, *rhowf = nullptr;

  ~DomainImpl() override;
};

// Kernel launchers (ops.hip, fused.hip).
// gate: skip once sumt >= dtg (fused substeps, which may run past the dtg)
template <typename T> int launch_forcinginterpolate(DomainImpl<T>& D, bool gate);
template <typename T> int launch_op_sin(DomainImpl<T>& D);
template <typename T> int launch_op_sds(DomainImpl<T>& D);
template <typename T> int launch_op_snl(DomainImpl<T>& D);
template <typename T> int launch_op_sice(DomainImpl<T>& D);
template <typename T> int launch_op_source(DomainImpl<T>& D);
template <typename T> int launch_op_propagation(DomainImpl<T>& D);
template <typename T> int launch_op_refraction(DomainImpl<T>& D);
template <typename T> int launch_op_update(DomainImpl<T>& D);
template <typename T> int launch_op_stress_atm(DomainImpl<T>& D);
template <typename T> int launch_op_stress_ocn(DomainImpl<T>& D);
template <typename T> int launch_diag(DomainImpl<T>& D);
template <typename T> int launch_stokes(DomainImpl<T>& D);
template <typename T> int launch_hs_ice(DomainImpl<T>& D);
template <typename T> int launch_jonswap(DomainImpl<T>& D, const umwm_jonswap& js);

// Forcing interpolation of the n domains list[0..n) (indices into doms and
// fds) in one launch, plus the reset of their anycur flags when any has
// forced currents; maximm is the largest imm among them.
template <typename T>
int launch_forcing_list(const DomainDev<T>* doms, const ForcingDev<T>* fds, const int32_t* list,
                        int n, int maximm, bool currents, bool gate, hipStream_t s);

// Fused substep kernels over a task list that may span many domains; `sur`
// selects the variants that take the rates of an attached surrogate where
// a domain has one.
template <typename T>
int launch_fused(const DomainDev<T>* doms, const BlockTask* tasks, int ntasks, int omax,
                 int first, int rotation, Tile tile, bool keep, hipStream_t s, hipEvent_t* ev,
                 bool sur = false);

// The attached surrogate's spreading factors for the domain's current
// spectrum and forcing (dm^2 - 1 into DomainDev::sur_r), on stream s.
int launch_surrogate_inference(DomainImpl<float>& D, hipStream_t s);

}  // namespace umwm
