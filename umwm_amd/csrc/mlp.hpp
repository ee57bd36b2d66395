// Multilayer perceptron on gfx950 BF16 MFMA: the learned source-term
// surrogate's network.  Weights and activations are bfloat16, products
// accumulate in FP32 (v_mfma_f32_32x32x16_bf16), biases, master weights,
// gradients and Adam moments are FP32.
//
// Layer l maps K_l inputs to N_l outputs, z = W_l x + b_l with W_l stored
// [N_l][K_l] (row per output, inputs contiguous); hidden layers apply the
// tanh form of GELU, the last layer is linear.  Layer inputs are padded to
// multiples of 64, the network's outputs to 8 (the padding rows and
// columns of W are zero and stay zero: their gradients are zero).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

namespace umwm {

using bf16 = __bf16;

// Operand layouts of the GEMM C[m][n] = sum_k A(m,k) B(k,n):
//   A_MK: A(m,k) = A[m*lda + k]      A_KM: A(m,k) = A[k*lda + m]
//   B_NK: B(k,n) = B[n*ldb + k]      B_KN: B(k,n) = B[k*ldb + n]
enum GemmLayout { A_MK = 0, A_KM = 1, B_NK = 0, B_KN = 1 };

// Epilogues:
//   EP_GELU      z = acc + bias[n]; Zout[m][n] = z (bf16, when given);
//                C[m][n] = gelu(z) (bf16)
//   EP_F32       C[m][n] = acc + bias[n] (fp32; bias may be null)
//   EP_BF16_BM   C[m][n] = acc + bias[m] (bf16): the transposed last layer
//   EP_DGELU     C[m][n] = acc * gelu'(Zin[m][n]) (bf16)
//   EP_PART      C[z][m][n] = acc (fp32 split-K partial, z = blockIdx.z)
//   EP_SUR       C[m][n] = exp(min(osd[m] (acc + bias[m]) + omu[m], 30)) - 1
//                (bf16): the surrogate's transposed last layer, its
//                normalised output turned back into the spreading factor
//                dm^2 it stands for, stored as dm^2 - 1 (>= 0 in exact
//                arithmetic), whose bf16 rounding keeps 3 significant digits
//                where dm^2 itself, close to 1, would keep 2
enum GemmEpilogue { EP_GELU = 0, EP_F32 = 1, EP_BF16_BM = 2, EP_DGELU = 3, EP_PART = 4,
                    EP_SUR = 5 };

// The surrogate's first-layer operand taken straight from a domain's
// spectrum: A(m, k) = (ln max(e[k lde + m], 1e-30) - mu[k]) is[k] for the
// nb spectral inputs k < nb, ext[(k - nb) ldx + m] (already normalised)
// for the next three, 0 beyond.  |A(m, k)| > zmax for a spectral input
// clears ok[m].
struct SpecOperand {
  const float* e = nullptr;
  int64_t lde = 0;
  int cols = 0;  // columns of a row from e on (>= M, a multiple of 4)
  int nb = 0;
  const float *mu = nullptr, *is = nullptr;
  const float* ext = nullptr;
  int64_t ldx = 0;
  float zmax = 0.f;
  int32_t* ok = nullptr;
};

struct GemmArgs {
  const bf16* A;
  const bf16* B;
  void* C;
  const float* bias;
  const bf16* Zin;
  bf16* Zout;
  int64_t lda, ldb, ldc, ldz;
  int M, N, K;
  int kchunk;  // K range of one blockIdx.z (split-K)
  const float *omu, *osd;  // EP_SUR per-row output transform
  SpecOperand sa;          // A_SPEC operand (inference GEMM)
};

int gemm(int alay, int blay, int ep, const GemmArgs& g, int splits, hipStream_t s);

// Inference GEMM (forward only, B_NK, epilogue EP_GELU without Zout or
// EP_SUR / EP_BF16_BM): 256 x 128 blocks, the A operand bf16 A_MK or the
// spectrum (A_SPEC).
enum { A_SPEC = 2 };
int gemm_infer(int amode, int ep, const GemmArgs& g, hipStream_t s);

// GELU, tanh form: 0.5 z (1 + tanh(c (z + 0.044715 z^3))), c = sqrt(2/pi),
// and its derivative.
__host__ __device__ inline float gelu_f(float z) {
  const float c = 0.7978845608028654f;
  const float u = c * (z + 0.044715f * z * z * z);
  const float t = tanhf(u);
  return 0.5f * z * (1.f + t);
}
__host__ __device__ inline float gelu_grad_f(float z) {
  const float c = 0.7978845608028654f;
  const float u = c * (z + 0.044715f * z * z * z);
  const float t = tanhf(u);
  return 0.5f * (1.f + t) + 0.5f * z * (1.f - t * t) * c * (1.f + 3.f * 0.044715f * z * z);
}

// The same GELU for outputs rounded to bfloat16 (the inference GEMM): 0.5 z
// (1 + tanh u) = z / (1 + exp(-2u)), with the hardware exp2 and
// reciprocal (relative error about 1e-6, far below bf16's 4e-3).
__device__ inline float gelu_fast_f(float z) {
  const float u = 0.7978845608028654f * (z + 0.044715f * z * z * z);
  return z * __builtin_amdgcn_rcpf(1.f + __builtin_amdgcn_exp2f(-2.8853900817779268f * u));
}

struct MlpLayerDev {
  int K, N;            // padded input and output widths
  float *w, *b;        // master weights [N][K], bias [N]
  bf16* wb;            // bf16 copy of w
  float *gw, *gb;      // gradients
  float *mw, *vw, *mb, *vb;  // Adam moments
};

// Device network and the work buffers of one training or inference batch.
struct Mlp {
  std::vector<MlpLayerDev> L;
  int nin = 0, nout = 0;      // unpadded widths
  int cap_rows = 0;           // rows the activation buffers hold
  // activations: x (bf16 [rows][K0]); per hidden layer z and h (bf16
  // [rows][N_l]); y (fp32 [rows][N_last]); backward: dy, dh (bf16)
  bf16* x = nullptr;
  std::vector<bf16*> z, h;
  float* y = nullptr;
  bf16 *dy = nullptr, *da = nullptr, *db = nullptr;
  float* part = nullptr;      // split-K partials of the weight gradients
  size_t part_elems = 0;
  double* loss_acc = nullptr; // device loss accumulator
  int64_t step = 0;
  std::vector<void*> allocs;
  ~Mlp();
};

// Creates the network with widths dims[0..nl] (unpadded), He-uniform
// weights from `seed`, zero biases, activation buffers for `rows` rows.
int mlp_create(const std::vector<int>& dims, uint64_t seed, int rows, Mlp& m);
// Forward on m.x (rows x K0 bf16 already written); training keeps z and h.
int mlp_forward(Mlp& m, int rows, hipStream_t s);
// Loss = mean over rows and the nout real outputs of w_o (y - t)^2 with
// t fp32 [rows][N_last]; writes dy and accumulates the loss (device).
int mlp_loss(Mlp& m, const float* t, const float* wout, int rows, hipStream_t s);
// Backward from m.dy: gradients of every layer into gw, gb.
int mlp_backward(Mlp& m, int rows, hipStream_t s);
// Adam step on every layer (lr, betas, eps, decoupled weight decay), then
// the bf16 weight copies.
int mlp_adam(Mlp& m, float lr, float b1, float b2, float eps, float wd, hipStream_t s);
// Refreshes the bf16 copies from the master weights.
int mlp_sync_bf16(Mlp& m, hipStream_t s);
// Inference with the last layer transposed: out[o][ldo] (bf16) for `rows`
// input rows in m.x; hidden layers run as in mlp_forward (without keeping
// z).  With omu/osd the EP_SUR transform is applied, else the plain output
// plus bias.
int mlp_infer_transposed(Mlp& m, int rows, bf16* out, int64_t ldo, hipStream_t s,
                         const float* omu = nullptr, const float* osd = nullptr);
// The same with the first layer's input read from a spectrum (SpecOperand)
// instead of m.x.
int mlp_infer_spectrum(Mlp& m, const SpecOperand& in, int rows, bf16* out, int64_t ldo,
                       hipStream_t s, const float* omu, const float* osd);

}  // namespace umwm
