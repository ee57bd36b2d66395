// BF16 MFMA kernels of the surrogate network (mlp.hpp): one tiled GEMM
// with operand-layout and epilogue variants, the loss, the gradient
// reductions and the Adam step.
//
// GEMM tiling: a workgroup of 4 waves computes a 128x128 block of C; wave
// (wm, wn) owns the 64x64 quarter (wm, wn) as 2x2 tiles of the 32x32x16
// bf16 MFMA (v_mfma_f32_32x32x16_bf16, FP32 accumulators).  The K loop
// steps by 32: each thread stages two 16-byte chunks of A and of B in
// registers while the waves multiply the tile held in LDS, then writes
// them to the other LDS buffer (one barrier per step).  In LDS both tiles
// are k-contiguous rows ([m][k], [n][k], padded by 8 to spread banks), so
// every MFMA operand is one 16-byte read; operands stored m- or
// n-contiguous in memory (the weight-gradient products over the batch)
// are transposed on their way into LDS.
//
// MFMA operand layout (32x32x16): lane l supplies A(m = l%32, k = 8(l/32)
// + j) and B(k = 8(l/32) + j, n = l%32), j = 0..7; accumulator register i
// of lane l holds C(m = 8(i/4) + 4(l/32) + i%4, n = l%32).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <random>
#include <string>

#include "bufio.hpp"
#include "common.hpp"
#include "mlp.hpp"

namespace umwm {
namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

constexpr int BM = 128, BN = 128, BK = 32, LDK = BK + 8;

template <int AL, int BL, int EP>
__global__ __launch_bounds__(256) void k_gemm(GemmArgs g) {
  __shared__ __attribute__((aligned(16))) bf16 As[2][BM][LDK];
  __shared__ __attribute__((aligned(16))) bf16 Bs[2][BN][LDK];
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int wm = w >> 1, wn = w & 1;
  const int m0 = blockIdx.y * BM, n0 = blockIdx.x * BN;
  const int kb = blockIdx.z * g.kchunk;
  const int ke = min(g.K, kb + g.kchunk);
  const bf16x8 zero = {};
  bf16x8 ra[2], rb[2];
  // Every index below stays inside the operand: rows (m, n) are checked
  // against M and N, k against the chunk end; a 16-byte chunk never
  // straddles the end because K, and M or N on their contiguous axis, are
  // multiples of 8 (checked by gemm()).
  auto gload = [&](int k0) {
#pragma unroll
    for (int c = 0; c < 2; ++c) {
      const int q = tid + c * 256;
      if constexpr (AL == A_MK) {
        const int r = q >> 2, kc = (q & 3) * 8, m = m0 + r, k = k0 + kc;
        ra[c] = (m < g.M && k < ke) ? *(const bf16x8*)(g.A + (int64_t)m * g.lda + k) : zero;
      } else {
        const int kr = q >> 4, mc = (q & 15) * 8, m = m0 + mc, k = k0 + kr;
        ra[c] = (k < ke && m < g.M) ? *(const bf16x8*)(g.A + (int64_t)k * g.lda + m) : zero;
      }
      if constexpr (BL == B_NK) {
        const int r = q >> 2, kc = (q & 3) * 8, n = n0 + r, k = k0 + kc;
        rb[c] = (n < g.N && k < ke) ? *(const bf16x8*)(g.B + (int64_t)n * g.ldb + k) : zero;
      } else {
        const int kr = q >> 4, nc = (q & 15) * 8, n = n0 + nc, k = k0 + kr;
        rb[c] = (k < ke && n < g.N) ? *(const bf16x8*)(g.B + (int64_t)k * g.ldb + n) : zero;
      }
    }
  };
  auto sstore = [&](int buf) {
#pragma unroll
    for (int c = 0; c < 2; ++c) {
      const int q = tid + c * 256;
      if constexpr (AL == A_MK) {
        *(bf16x8*)&As[buf][q >> 2][(q & 3) * 8] = ra[c];
      } else {
        const int kr = q >> 4, mc = (q & 15) * 8;
#pragma unroll
        for (int i = 0; i < 8; ++i) As[buf][mc + i][kr] = ra[c][i];
      }
      if constexpr (BL == B_NK) {
        *(bf16x8*)&Bs[buf][q >> 2][(q & 3) * 8] = rb[c];
      } else {
        const int kr = q >> 4, nc = (q & 15) * 8;
#pragma unroll
        for (int i = 0; i < 8; ++i) Bs[buf][nc + i][kr] = rb[c][i];
      }
    }
  };
  f32x16 acc[2][2];
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b)
#pragma unroll
      for (int i = 0; i < 16; ++i) acc[a][b][i] = 0.f;
  int buf = 0;
  if (kb < ke) {
    gload(kb);
    sstore(0);
  }
  __syncthreads();
  for (int k0 = kb; k0 < ke; k0 += BK) {
    const bool more = k0 + BK < ke;
    if (more) gload(k0 + BK);
#pragma unroll
    for (int ks = 0; ks < 2; ++ks) {
      const int kk = ks * 16 + (lane >> 5) * 8;
      bf16x8 fa[2], fb[2];
#pragma unroll
      for (int t = 0; t < 2; ++t) {
        fa[t] = *(const bf16x8*)&As[buf][wm * 64 + t * 32 + (lane & 31)][kk];
        fb[t] = *(const bf16x8*)&Bs[buf][wn * 64 + t * 32 + (lane & 31)][kk];
      }
#pragma unroll
      for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b)
          acc[a][b] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(fa[a], fb[b], acc[a][b], 0, 0, 0);
    }
    if (more) sstore(buf ^ 1);
    __syncthreads();
    buf ^= 1;
  }
  // epilogue
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b)
#pragma unroll
      for (int i = 0; i < 16; ++i) {
        const int m = m0 + wm * 64 + a * 32 + 8 * (i >> 2) + 4 * (lane >> 5) + (i & 3);
        const int n = n0 + wn * 64 + b * 32 + (lane & 31);
        if (m >= g.M || n >= g.N) continue;
        const float v = acc[a][b][i];
        const int64_t o = (int64_t)m * g.ldc + n;
        if constexpr (EP == EP_GELU) {
          const float z = v + g.bias[n];
          if (g.Zout) g.Zout[o] = (bf16)z;
          ((bf16*)g.C)[o] = (bf16)gelu_f(z);
        } else if constexpr (EP == EP_F32) {
          ((float*)g.C)[o] = v + (g.bias ? g.bias[n] : 0.f);
        } else if constexpr (EP == EP_BF16_BM) {
          ((bf16*)g.C)[o] = (bf16)(v + g.bias[m]);
        } else if constexpr (EP == EP_DGELU) {
          ((bf16*)g.C)[o] = (bf16)(v * gelu_grad_f((float)g.Zin[(int64_t)m * g.ldz + n]));
        } else if constexpr (EP == EP_SUR) {
          ((bf16*)g.C)[o] = (bf16)expm1f(fminf(g.osd[m] * (v + g.bias[m]) + g.omu[m], 30.f));
        } else {
          ((float*)g.C)[(int64_t)blockIdx.z * g.M * g.ldc + o] = v;
        }
      }
}

// Inference GEMM.  A workgroup of 4 waves (2 x 2) computes a BMT x BNT
// block of C, each wave a (BMT/2) x (BNT/2) part as TA x TB tiles of the
// 32x32x16 MFMA (8 tiles: each 16-byte LDS read feeds 8/6 of an MFMA, the
// training GEMM's 2 x 2 wave tiles feed one).  bf16 A operands use 256 x
// 128 blocks; the spectrum operand uses 128 x 256, so that its ln /
// normalise / bf16 transform, done once per column block, is repeated
// over 2 column blocks of the 512 hidden outputs rather than 4.  The K
// loop steps by 32 with two tiles staged in registers ahead.  Blocks are
// numbered so that the blocks sharing one A block (or, for the transposed
// last layer, one B block) land on the same XCD back to back (workgroups
// go round-robin over the 8 XCDs): the shared operand then comes from that
// XCD's L2 after the first read.
//
// A_SPEC stages the first layer's input from the spectrum (fp32, points
// contiguous): each thread loads BMT/64 points x 8 inputs as 8 vector
// reads (a wave reads BMT x 4 bytes contiguous per input row), forms
// ln e, normalises, rounds to bf16 and writes each point's 8 inputs as
// one 16-byte LDS row segment.
constexpr int ILDK = BK + 8;

template <int AM, int BMT>
struct InfStage;
template <int BMT>
struct InfStage<A_MK, BMT> {
  bf16x8 a[BMT / 64];
};
template <int BMT>
struct InfStage<A_SPEC, BMT> {
  using V = float __attribute__((ext_vector_type(BMT / 64)));
  V a[8];
};

template <int AM, int EP, int BMT, int BNT>
__global__ __launch_bounds__(256, 2) void k_gemm_inf(GemmArgs g, int ntm, int ntn, int nmajor) {
  constexpr int TA = BMT / 64, TB = BNT / 64;  // MFMA tiles per wave
  constexpr int CA = BMT / 64, CB = BNT / 64;  // 16-byte bf16 chunks per thread per K tile
  constexpr int PPT = BMT / 64;                // A_SPEC: points per thread
  static_assert(TA * TB == 8 && (BMT == 256 || BMT == 128), "block shapes 256x128, 128x256");
  // one LDS array: the A and B tiles of the K loop, then the epilogue's
  // staging rows
  __shared__ __attribute__((aligned(16))) bf16 lds[2 * (BMT + BNT) * ILDK];
  bf16(*const As)[BMT][ILDK] = reinterpret_cast<bf16(*)[BMT][ILDK]>(lds);
  bf16(*const Bs)[BNT][ILDK] = reinterpret_cast<bf16(*)[BNT][ILDK]>(lds + 2 * BMT * ILDK);
  // XCD-aware tile order (the grid is padded to a multiple of 8): row
  // blocks outer (the blocks of one A block adjacent) or, with nmajor,
  // column blocks outer (the blocks of one B block adjacent)
  const int nt = ntm * ntn, per = (int)(gridDim.x / 8);
  const int t = (int)(blockIdx.x % 8) * per + (int)(blockIdx.x / 8);
  if (t >= nt) return;
  const int tm = nmajor ? t % ntm : t / ntn, tn = nmajor ? t / ntm : t % ntn;
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int wm = w >> 1, wn = w & 1;
  const int m0 = tm * BMT, n0 = tn * BNT;
  const int K = g.K;
  struct Stage {
    InfStage<AM, BMT> a;
    bf16x8 b[CB];
  };
  Stage st[2];  // two K tiles in flight
  // A_SPEC: this thread's PPT points and (wave-uniform) 8-input group
  [[maybe_unused]] const int smc = (tid & 63) * PPT;
  [[maybe_unused]] const int skg = __builtin_amdgcn_readfirstlane(tid >> 6) * 8;
  // A_SPEC: the largest |normalised input| of each of its points
  [[maybe_unused]] float zabs[PPT] = {};
  // Every load stays inside its operand and none is conditional (a
  // masked load would make the compiler wait for every load in flight
  // before the next LDS store): K is a multiple of 64 (checked by
  // gemm_infer()), so the K tiles hold no partial 16-byte chunk; rows past
  // M or N read row M - 1 or N - 1 again, and their results are never
  // stored; a spectrum vector lies wholly inside its row or wholly past it
  // (lde, the chunk start and the points per thread all multiples of 4 or
  // of PPT) and reads 0 there.  Points past M feed only their own rows of
  // C.
  // bf16 operands through buffer resources: each thread's 16-byte chunks
  // keep a fixed 32-bit byte offset (row, column within the K tile), and the
  // K tile's position is the scalar offset (gemm_infer() checks that the
  // operands span less than 2 GiB)
  [[maybe_unused]] const rsrc_t rA =
      make_rsrc(AM == A_MK ? (const void*)g.A : (const void*)g.B,
                AM == A_MK ? (uint32_t)((int64_t)g.M * g.lda * 2) : 0u);
  const rsrc_t rB = make_rsrc(g.B, (uint32_t)((int64_t)g.N * g.ldb * 2));
  [[maybe_unused]] uint32_t offA[CA];
  uint32_t offB[CB];
#pragma unroll
  for (int c = 0; c < CA; ++c) {
    const int q = tid + c * 256;
    if constexpr (AM == A_MK)
      offA[c] = (uint32_t)(((int64_t)min(m0 + (q >> 2), g.M - 1) * g.lda + (q & 3) * 8) * 2);
  }
#pragma unroll
  for (int c = 0; c < CB; ++c) {
    const int q = tid + c * 256;
    offB[c] = (uint32_t)(((int64_t)min(n0 + (q >> 2), g.N - 1) * g.ldb + (q & 3) * 8) * 2);
  }
  auto gload = [&](Stage& S, int k0) {
    if constexpr (AM == A_MK) {
#pragma unroll
      for (int c = 0; c < CA; ++c)
        S.a.a[c] = __builtin_bit_cast(bf16x8, bld16<float>(rA, offA[c], (uint32_t)k0 * 2u));
    } else {
      // one buffer resource per input row (scalar registers: the row is
      // wave-uniform) covering this block's columns of the row; a vector
      // past the row's end (the last row of the spectrum only) reads 0,
      // and rows past the inputs (K padding) get an empty resource
      const SpecOperand& sa = g.sa;
      const int kb = k0 + skg;
      const uint32_t voff = (uint32_t)smc * 4u;
#pragma unroll
      for (int i = 0; i < 8; ++i) {
        const int k = kb + i;
        const bool spec = k < sa.nb, in = k < sa.nb + 3;
        const int kk = in ? k : sa.nb;
        const float* row = spec ? sa.e + (int64_t)k * sa.lde : sa.ext + (int64_t)(kk - sa.nb) * sa.ldx;
        const int cols = spec ? sa.cols : (int)sa.ldx;
        const rsrc_t r = make_rsrc(row + m0, in ? (uint32_t)std::min(cols - m0, BMT) * 4u : 0u);
        if constexpr (PPT == 4) {
          S.a.a[i] = bld16<float>(r, voff);
        } else {
          S.a.a[i] = __builtin_bit_cast(typename InfStage<AM, BMT>::V,
                                        __builtin_amdgcn_raw_buffer_load_b64(r, (int)voff, 0, 0));
        }
      }
    }
#pragma unroll
    for (int c = 0; c < CB; ++c)
      S.b[c] = __builtin_bit_cast(bf16x8, bld16<float>(rB, offB[c], (uint32_t)k0 * 2u));
  };
  auto sstore = [&](const Stage& S, int buf, int k0) {
    if constexpr (AM == A_MK) {
#pragma unroll
      for (int c = 0; c < CA; ++c) {
        const int q = tid + c * 256;
        *(bf16x8*)&As[buf][q >> 2][(q & 3) * 8] = S.a.a[c];
      }
    } else {
      const SpecOperand& sa = g.sa;
      const int kb = k0 + skg;
      bf16x8 o[PPT];
#pragma unroll
      for (int i = 0; i < 8; ++i) {
        const int k = kb + i;
        float x[PPT];
#pragma unroll
        for (int c = 0; c < PPT; ++c) x[c] = S.a.a[i][c];
        if (k < sa.nb) {
          // (ln e - mu) is = log2(e) (ln 2 is) - mu is: the hardware log2
          // and one fma per input (wave-uniform constants per k)
          const float is = sa.is[k], a = is * 0.6931471805599453f, c0 = -sa.mu[k] * is;
#pragma unroll
          for (int c = 0; c < PPT; ++c) {
            x[c] = fmaf(__builtin_amdgcn_logf(fmaxf(x[c], 1e-30f)), a, c0);
            zabs[c] = fmaxf(zabs[c], fabsf(x[c]));
          }
        }
#pragma unroll
        for (int c = 0; c < PPT; ++c) o[c][i] = (bf16)x[c];
      }
#pragma unroll
      for (int c = 0; c < PPT; ++c) *(bf16x8*)&As[buf][smc + c][skg] = o[c];
    }
#pragma unroll
    for (int c = 0; c < CB; ++c) {
      const int q = tid + c * 256;
      *(bf16x8*)&Bs[buf][q >> 2][(q & 3) * 8] = S.b[c];
    }
  };
  f32x16 acc[TA][TB];
#pragma unroll
  for (int a = 0; a < TA; ++a)
#pragma unroll
    for (int b = 0; b < TB; ++b)
#pragma unroll
      for (int i = 0; i < 16; ++i) acc[a][b][i] = 0.f;
  auto mfma = [&](int buf) {
#pragma unroll
    for (int ks = 0; ks < 2; ++ks) {
      const int kk = ks * 16 + (lane >> 5) * 8;
      bf16x8 fa[TA], fb[TB];
#pragma unroll
      for (int a = 0; a < TA; ++a)
        fa[a] = *(const bf16x8*)&As[buf][wm * (BMT / 2) + a * 32 + (lane & 31)][kk];
#pragma unroll
      for (int b = 0; b < TB; ++b)
        fb[b] = *(const bf16x8*)&Bs[buf][wn * (BNT / 2) + b * 32 + (lane & 31)][kk];
#pragma unroll
      for (int a = 0; a < TA; ++a)
#pragma unroll
        for (int b = 0; b < TB; ++b)
          acc[a][b] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(fa[a], fb[b], acc[a][b], 0, 0, 0);
    }
  };
  // K tiles 0 .. nkt-1 of BK (nkt even: K a multiple of 64); tile u lives
  // in LDS buffer u % 2, odd tiles wait in st[0], even ones in st[1].  Each
  // step stores the next tile, refills its stage three tiles ahead and
  // multiplies the current one.  Loads and stores are unconditional (tile
  // indices past the end are clamped to the last tile, whose reload and
  // restore change nothing): a load under a branch leaves the compiler
  // unsure which stage's loads are the newest, and it then waits for all of
  // them before each LDS store.
  const int nkt = K / BK, last = nkt - 1;
  auto tile = [&](int u) { return min(u, last) * BK; };
  gload(st[0], 0);
  sstore(st[0], 0, 0);
  gload(st[0], tile(1));
  gload(st[1], tile(2));
  __syncthreads();
#pragma unroll 1
  for (int u = 0; u + 1 < nkt; u += 2) {
    sstore(st[0], 1, tile(u + 1));
    gload(st[0], tile(u + 3));
    mfma(0);
    __syncthreads();
    sstore(st[1], 0, tile(u + 2));
    gload(st[1], tile(u + 4));
    mfma(1);
    __syncthreads();
  }
  if constexpr (AM == A_SPEC) {
    // the first column of blocks reports each point's zmax verdict (inputs
    // are at least 1e-30 before the log, so no normalised input is NaN)
    if (tn == 0) {
#pragma unroll
      for (int c = 0; c < PPT; ++c)
        if (!(zabs[c] <= g.sa.zmax) && m0 + smc + c < g.M) atomicAnd(g.sa.ok + m0 + smc + c, 0);
    }
  }
  // Epilogue through LDS: each wave writes its part of C in 64 x 64 pieces
  // into its own LDS region (the MFMA layout puts one column per lane),
  // then reads rows back 16 bytes per lane, so that every store to C is a
  // whole 128-byte row segment.  Scattered 2-byte stores of the MFMA layout
  // ran at about 0.6 TB/s and made the output-heavy last layer store-bound.
  constexpr int SLD = 64 + 8;  // staging row, padded
  static_assert(4 * 64 * SLD <= 2 * (BMT + BNT) * ILDK, "staging fits in the LDS array");
  __syncthreads();  // every wave done with the K loop's tiles
  bf16(*stg)[SLD] = reinterpret_cast<bf16(*)[SLD]>(lds) + w * 64;
  bf16* const C = static_cast<bf16*>(g.C);
#pragma unroll
  for (int pr = 0; pr < TA / 2; ++pr)
#pragma unroll
    for (int pc = 0; pc < TB / 2; ++pc) {
#pragma unroll
      for (int aa = 0; aa < 2; ++aa)
#pragma unroll
        for (int bb = 0; bb < 2; ++bb)
#pragma unroll
          for (int i = 0; i < 16; ++i) {
            const int r = aa * 32 + 8 * (i >> 2) + 4 * (lane >> 5) + (i & 3), c = bb * 32 + (lane & 31);
            // indices of the epilogue's vectors, clamped: rows past M and
            // columns past N are staged but never stored
            const int m = min(m0 + wm * (BMT / 2) + pr * 64 + r, g.M - 1);
            const int n = min(n0 + wn * (BNT / 2) + pc * 64 + c, g.N - 1);
            const float v = acc[2 * pr + aa][2 * pc + bb][i];
            float y;
            if constexpr (EP == EP_GELU) {
              y = gelu_fast_f(v + g.bias[n]);
            } else if constexpr (EP == EP_BF16_BM) {
              y = v + g.bias[m];
            } else {
              static_assert(EP == EP_SUR, "inference epilogues: GELU, BF16_BM, SUR");
              // exp(x) - 1 with the hardware exp2: its cancellation near
              // x = 0 costs relative precision only where dm^2 - 1 is below
              // 1e-4, which moves dm^2 by less than bf16 resolves
              const float x = fminf(g.osd[m] * (v + g.bias[m]) + g.omu[m], 30.f);
              y = __builtin_amdgcn_exp2f(1.4426950408889634f * x) - 1.f;
            }
            stg[r][c] = (bf16)y;
          }
      // LDS accesses of one wave complete in order: the rows read below are
      // the ones written above, by any lane of the wave
#pragma unroll
      for (int it = 0; it < 8; ++it) {
        const int q = it * 64 + lane, r = q >> 3, cc = (q & 7) * 8;
        const int m = m0 + wm * (BMT / 2) + pr * 64 + r, n = n0 + wn * (BNT / 2) + pc * 64 + cc;
        const bf16x8 v = *(const bf16x8*)&stg[r][cc];
        if (m < g.M) {
          bf16* dst = C + (int64_t)m * g.ldc + n;
          if (n + 8 <= g.N) {
            *(bf16x8*)dst = v;
          } else {
#pragma unroll
            for (int e = 0; e < 8; ++e)
              if (n + e < g.N) dst[e] = v[e];
          }
        }
      }
    }
}

// gw[i] = sum over s of part[s][i] (deterministic split-K reduction).
__global__ void k_reduce_parts(const float* __restrict__ part, int splits, int64_t n,
                               float* __restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float s = 0.f;
  for (int k = 0; k < splits; ++k) s += part[(int64_t)k * n + i];
  out[i] = s;
}

// gb[n] = sum over rows of d[r][n] (bf16 [rows][ld]).  A block of 256
// threads takes 32 columns: 8 row groups sum every 8th row each, then the 8
// partial sums are added in a fixed order (the result does not depend on
// scheduling).
__global__ __launch_bounds__(256) void k_colsum(const bf16* __restrict__ d, int rows, int cols,
                                                int64_t ld, float* __restrict__ out) {
  __shared__ float part[8][33];
  const int c = threadIdx.x & 31, rg = threadIdx.x >> 5, n = blockIdx.x * 32 + c;
  float s = 0.f;
  if (n < cols)
    for (int r = rg; r < rows; r += 8) s += (float)d[(int64_t)r * ld + n];
  part[rg][c] = s;
  __syncthreads();
  if (rg == 0 && n < cols) {
    float t = 0.f;
#pragma unroll
    for (int q = 0; q < 8; ++q) t += part[q][c];
    out[n] = t;
  }
}

// Mean squared error over the real outputs with per-output weights: one
// block per row (no index division), one atomic per block.
__global__ __launch_bounds__(256) void k_loss(const float* __restrict__ y,
                                              const float* __restrict__ t,
                                              const float* __restrict__ wout, int rows, int nout,
                                              int ld, bf16* __restrict__ dy,
                                              double* __restrict__ loss) {
  __shared__ double wsum[4];
  const int64_t r0 = (int64_t)blockIdx.x * ld;
  const float inv = 1.f / ((float)rows * (float)nout);
  double l = 0;
  for (int c = threadIdx.x; c < ld; c += 256) {
    float g = 0.f;
    if (c < nout) {
      const float wo = wout ? wout[c] : 1.f;
      const float d = y[r0 + c] - t[r0 + c];
      g = 2.f * wo * d * inv;
      l += (double)wo * d * d * inv;
    }
    dy[r0 + c] = (bf16)g;
  }
  for (int off = 32; off > 0; off >>= 1) l += __shfl_down(l, off, 64);
  if ((threadIdx.x & 63) == 0) wsum[threadIdx.x >> 6] = l;
  __syncthreads();
  if (threadIdx.x == 0) {
    const double b = wsum[0] + wsum[1] + wsum[2] + wsum[3];
    if (b != 0) atomicAdd(loss, b);
  }
}

__global__ void k_adam(float* __restrict__ p, const float* __restrict__ gr, float* __restrict__ m,
                       float* __restrict__ v, int64_t n, float lr, float b1, float b2, float eps,
                       float wd, float c1, float c2) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float gi = gr[i];
  const float mi = b1 * m[i] + (1.f - b1) * gi;
  const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  p[i] = p[i] * (1.f - lr * wd) - lr * (mi * c1) / (sqrtf(vi * c2) + eps);
}

__global__ void k_to_bf16(const float* __restrict__ a, bf16* __restrict__ b, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) b[i] = (bf16)a[i];
}

inline dim3 grid_for(int64_t n, int t = 256) { return dim3((unsigned)((n + t - 1) / t)); }

template <int AL, int BL>
int gemm_ep(int ep, const GemmArgs& g, dim3 grid, hipStream_t s) {
  switch (ep) {
    case EP_GELU: hipLaunchKernelGGL((k_gemm<AL, BL, EP_GELU>), grid, dim3(256), 0, s, g); break;
    case EP_F32: hipLaunchKernelGGL((k_gemm<AL, BL, EP_F32>), grid, dim3(256), 0, s, g); break;
    case EP_BF16_BM: hipLaunchKernelGGL((k_gemm<AL, BL, EP_BF16_BM>), grid, dim3(256), 0, s, g); break;
    case EP_DGELU: hipLaunchKernelGGL((k_gemm<AL, BL, EP_DGELU>), grid, dim3(256), 0, s, g); break;
    case EP_PART: hipLaunchKernelGGL((k_gemm<AL, BL, EP_PART>), grid, dim3(256), 0, s, g); break;
    case EP_SUR: hipLaunchKernelGGL((k_gemm<AL, BL, EP_SUR>), grid, dim3(256), 0, s, g); break;
    default: return fail(UMWM_ERR_ARGUMENT, "gemm: unknown epilogue");
  }
  return 0;
}

}  // namespace

int gemm(int alay, int blay, int ep, const GemmArgs& g0, int splits, hipStream_t s) {
  GemmArgs g = g0;
  if (g.M <= 0 || g.N <= 0 || g.K <= 0) return 0;
  // 16-byte chunks: K and the contiguous axis of transposed operands must
  // be multiples of 8, and so must the leading dimensions
  if (g.K % 8 || g.lda % 8 || g.ldb % 8 || (alay == A_KM && g.M % 8) || (blay == B_KN && g.N % 8))
    return fail(UMWM_ERR_ARGUMENT, "gemm: K, lda, ldb (and M or N when transposed) must be "
                                   "multiples of 8");
  splits = std::max(1, splits);
  int kchunk = (g.K + splits - 1) / splits;
  kchunk = (kchunk + BK - 1) / BK * BK;
  splits = (g.K + kchunk - 1) / kchunk;
  g.kchunk = kchunk;
  if (ep != EP_PART && splits != 1) return fail(UMWM_ERR_ARGUMENT, "gemm: split-K needs EP_PART");
  const dim3 grid((g.N + BN - 1) / BN, (g.M + BM - 1) / BM, splits);
  int rc;
  if (alay == A_MK && blay == B_NK) rc = gemm_ep<A_MK, B_NK>(ep, g, grid, s);
  else if (alay == A_MK && blay == B_KN) rc = gemm_ep<A_MK, B_KN>(ep, g, grid, s);
  else if (alay == A_KM && blay == B_KN) rc = gemm_ep<A_KM, B_KN>(ep, g, grid, s);
  else rc = gemm_ep<A_KM, B_NK>(ep, g, grid, s);
  if (rc) return rc;
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail(UMWM_ERR_HIP, std::string("gemm: ") + hipGetErrorString(e));
  return 0;
}

int gemm_infer(int amode, int ep, const GemmArgs& g, hipStream_t s) {
  if (g.M <= 0 || g.N <= 0 || g.K <= 0) return 0;
  if (g.K % (2 * BK) || g.ldb % 8 || g.ldc % 8 || (amode == A_MK && g.lda % 8) ||
      reinterpret_cast<uintptr_t>(g.C) % 16)
    return fail(UMWM_ERR_ARGUMENT, "gemm_infer: K a multiple of 64, lda, ldb and ldc of 8, C "
                                   "16-byte aligned");
  if ((int64_t)g.N * g.ldb * 2 >= (1ll << 31) || (amode == A_MK && (int64_t)g.M * g.lda * 2 >= (1ll << 31)))
    return fail(UMWM_ERR_SIZE, "gemm_infer: an operand spans 2 GiB or more; use smaller chunks");
  if (amode == A_SPEC) {
    const SpecOperand& a = g.sa;
    if (!a.e || !a.mu || !a.is || !a.ext || !a.ok || a.nb < 1 || a.nb + 3 > g.K ||
        a.lde % 4 || a.ldx % 4 || a.cols < g.M || a.cols % 4 || a.ldx < g.M ||
        reinterpret_cast<uintptr_t>(a.e) % 16 || reinterpret_cast<uintptr_t>(a.ext) % 16)
      return fail(UMWM_ERR_ARGUMENT, "gemm_infer: bad spectrum operand");
  }
  // the spectrum operand takes 128 x 256 blocks, bf16 operands 256 x 128
  const int bm = amode == A_SPEC ? 128 : 256, bn = amode == A_SPEC ? 256 : 128;
  const int ntm = (g.M + bm - 1) / bm, ntn = (g.N + bn - 1) / bn;
  const int64_t nt = (int64_t)ntm * ntn;
  if (nt > (1ll << 30)) return fail(UMWM_ERR_SIZE, "gemm_infer: too many tiles");
  const dim3 grid((unsigned)((nt + 7) / 8 * 8));
  // the transposed last layer (EP_SUR, EP_BF16_BM) has few row blocks (the
  // outputs) and many column blocks (the points): column blocks outer, so
  // that the row blocks of one activation block share it in L2
  const int nmajor = ep == EP_GELU ? 0 : 1;
#define UMWM_INF(AM_, EP_, BM_, BN_) \
  hipLaunchKernelGGL((k_gemm_inf<AM_, EP_, BM_, BN_>), grid, dim3(256), 0, s, g, ntm, ntn, nmajor)
  if (amode == A_MK && ep == EP_GELU) UMWM_INF(A_MK, EP_GELU, 256, 128);
  else if (amode == A_SPEC && ep == EP_GELU) UMWM_INF(A_SPEC, EP_GELU, 128, 256);
  else if (amode == A_MK && ep == EP_SUR) UMWM_INF(A_MK, EP_SUR, 256, 128);
  else if (amode == A_MK && ep == EP_BF16_BM) UMWM_INF(A_MK, EP_BF16_BM, 256, 128);
  else return fail(UMWM_ERR_ARGUMENT, "gemm_infer: unsupported operand/epilogue pair");
#undef UMWM_INF
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail(UMWM_ERR_HIP, std::string("gemm_infer: ") + hipGetErrorString(e));
  return 0;
}

Mlp::~Mlp() {
  for (void* p : allocs) (void)hipFree(p);
}

namespace {
template <typename U>
U* dalloc(Mlp& m, size_t n) {
  void* p = nullptr;
  if (hipMalloc(&p, std::max<size_t>(n, 1) * sizeof(U)) != hipSuccess) return nullptr;
  m.allocs.push_back(p);
  return static_cast<U*>(p);
}
inline int pad8(int v) { return (v + 7) / 8 * 8; }
inline int pad64(int v) { return (v + 63) / 64 * 64; }
}  // namespace

int mlp_create(const std::vector<int>& dims, uint64_t seed, int rows, Mlp& m) {
  const int nl = (int)dims.size() - 1;
  if (nl < 1 || rows < 1) return fail(UMWM_ERR_ARGUMENT, "mlp: need >= 1 layer and >= 1 row");
  m.nin = dims.front();
  m.nout = dims.back();
  m.cap_rows = rows;
  std::mt19937_64 rng(seed);
  std::uniform_real_distribution<double> U(-1.0, 1.0);
  size_t part = 0;
  for (int l = 0; l < nl; ++l) {
    MlpLayerDev L{};
    // inputs padded to an even number of 32-wide K tiles of the inference
    // GEMM (a hidden layer's outputs are the next layer's inputs), the
    // network's outputs to 8
    L.K = pad64(dims[l]);
    L.N = l + 1 < nl ? pad64(dims[l + 1]) : pad8(dims[l + 1]);
    const size_t nw = (size_t)L.N * L.K;
    L.w = dalloc<float>(m, nw); L.b = dalloc<float>(m, L.N); L.wb = dalloc<bf16>(m, nw);
    L.gw = dalloc<float>(m, nw); L.gb = dalloc<float>(m, L.N);
    L.mw = dalloc<float>(m, nw); L.vw = dalloc<float>(m, nw);
    L.mb = dalloc<float>(m, L.N); L.vb = dalloc<float>(m, L.N);
    if (!L.vb) return fail(UMWM_ERR_NOMEM, "mlp: device allocation failed");
    // He-uniform for the GELU layers, Glorot-uniform for the linear output;
    // padding rows and columns zero
    const double fan_in = dims[l], fan_out = dims[l + 1];
    const double a = l + 1 < nl ? std::sqrt(6.0 / fan_in) : std::sqrt(6.0 / (fan_in + fan_out));
    std::vector<float> hw(nw, 0.f), zero(nw, 0.f), zb(L.N, 0.f);
    for (int n = 0; n < dims[l + 1]; ++n)
      for (int k = 0; k < dims[l]; ++k) hw[(size_t)n * L.K + k] = (float)(a * U(rng));
    UMWM_HIP_TRY(hipMemcpy(L.w, hw.data(), nw * 4, hipMemcpyHostToDevice));
    for (float* p : {L.gw, L.mw, L.vw}) UMWM_HIP_TRY(hipMemcpy(p, zero.data(), nw * 4, hipMemcpyHostToDevice));
    for (float* p : {L.b, L.gb, L.mb, L.vb}) UMWM_HIP_TRY(hipMemcpy(p, zb.data(), L.N * 4, hipMemcpyHostToDevice));
    m.L.push_back(L);
    part = std::max(part, nw);
  }
  // activation buffers
  m.x = dalloc<bf16>(m, (size_t)rows * m.L[0].K);
  int wmax = m.L[0].K;
  for (int l = 0; l + 1 < nl; ++l) {
    m.z.push_back(dalloc<bf16>(m, (size_t)rows * m.L[l].N));
    m.h.push_back(dalloc<bf16>(m, (size_t)rows * m.L[l].N));
    wmax = std::max(wmax, m.L[l].N);
  }
  wmax = std::max(wmax, m.L.back().N);
  m.y = dalloc<float>(m, (size_t)rows * m.L.back().N);
  m.dy = dalloc<bf16>(m, (size_t)rows * m.L.back().N);
  m.da = dalloc<bf16>(m, (size_t)rows * wmax);
  m.db = dalloc<bf16>(m, (size_t)rows * wmax);
  // split-K partials: at most 32 splits of the largest weight matrix
  m.part_elems = part * 32;
  m.part = dalloc<float>(m, m.part_elems);
  m.loss_acc = dalloc<double>(m, 1);
  if (!m.loss_acc || !m.part || !m.db) return fail(UMWM_ERR_NOMEM, "mlp: device allocation failed");
  if (int rc = mlp_sync_bf16(m, nullptr)) return rc;
  UMWM_HIP_TRY(hipDeviceSynchronize());
  return 0;
}

int mlp_sync_bf16(Mlp& m, hipStream_t s) {
  for (auto& L : m.L) {
    const int64_t nw = (int64_t)L.N * L.K;
    hipLaunchKernelGGL(k_to_bf16, grid_for(nw), dim3(256), 0, s, (const float*)L.w, L.wb, nw);
  }
  UMWM_HIP_TRY(hipGetLastError());
  return 0;
}

int mlp_forward(Mlp& m, int rows, hipStream_t s) {
  if (rows > m.cap_rows) return fail(UMWM_ERR_SIZE, "mlp: more rows than the buffers hold");
  const int nl = (int)m.L.size();
  const bf16* in = m.x;
  for (int l = 0; l < nl; ++l) {
    const MlpLayerDev& L = m.L[l];
    GemmArgs g{};
    g.A = in; g.lda = L.K;
    g.B = L.wb; g.ldb = L.K;
    g.M = rows; g.N = L.N; g.K = L.K;
    g.bias = L.b;
    g.ldc = L.N;
    int rc;
    if (l + 1 < nl) {
      g.C = m.h[l]; g.Zout = m.z[l];
      rc = gemm(A_MK, B_NK, EP_GELU, g, 1, s);
      in = m.h[l];
    } else {
      g.C = m.y;
      rc = gemm(A_MK, B_NK, EP_F32, g, 1, s);
    }
    if (rc) return rc;
  }
  return 0;
}

int mlp_loss(Mlp& m, const float* t, const float* wout, int rows, hipStream_t s) {
  const int ld = m.L.back().N;
  if (rows < 1) return 0;
  hipLaunchKernelGGL(k_loss, dim3(rows), dim3(256), 0, s, (const float*)m.y, t, wout, rows,
                     m.nout, ld, m.dy, m.loss_acc);
  UMWM_HIP_TRY(hipGetLastError());
  return 0;
}

int mlp_backward(Mlp& m, int rows, hipStream_t s) {
  const int nl = (int)m.L.size();
  const bf16* dz = m.dy;
  bf16* bufs[2] = {m.da, m.db};
  int pb = 0;
  for (int l = nl - 1; l >= 0; --l) {
    const MlpLayerDev& L = m.L[l];
    const bf16* a = l == 0 ? m.x : m.h[l - 1];
    // weight gradient over the batch, split over rows so that the small
    // output (N x K) still fills the GPU
    const int tiles = ((L.N + BM - 1) / BM) * ((L.K + BN - 1) / BN);
    int splits = std::max(1, std::min(rows / 256, (1024 + tiles - 1) / tiles));
    splits = std::min(splits, 32);
    GemmArgs g{};
    g.A = dz; g.lda = L.N;       // A(m = out, k = row) = dz[row][out]
    g.B = a; g.ldb = L.K;        // B(k = row, n = in) = a[row][in]
    g.M = L.N; g.N = L.K; g.K = rows;
    g.C = m.part; g.ldc = L.K;
    if ((size_t)splits * L.N * L.K > m.part_elems) splits = (int)(m.part_elems / ((size_t)L.N * L.K));
    // gemm() may lower the split count to whole chunks of 32 rows
    int kchunk = (rows + splits - 1) / splits;
    kchunk = (kchunk + BK - 1) / BK * BK;
    const int used = (rows + kchunk - 1) / kchunk;
    if (rows % 8) return fail(UMWM_ERR_ARGUMENT, "mlp: training rows must be a multiple of 8");
    if (int rc = gemm(A_KM, B_KN, EP_PART, g, splits, s)) return rc;
    const int64_t nw = (int64_t)L.N * L.K;
    hipLaunchKernelGGL(k_reduce_parts, grid_for(nw), dim3(256), 0, s, (const float*)m.part, used,
                       nw, L.gw);
    hipLaunchKernelGGL(k_colsum, dim3((L.N + 31) / 32), dim3(256), 0, s, dz, rows, L.N,
                       (int64_t)L.N, L.gb);
    UMWM_HIP_TRY(hipGetLastError());
    if (l > 0) {
      // d a_{l-1} = dz W_l, times gelu'(z_{l-1})
      GemmArgs d{};
      d.A = dz; d.lda = L.N;      // A(m = row, k = out)
      d.B = L.wb; d.ldb = L.K;    // B(k = out, n = in) = W[out][in]
      d.M = rows; d.N = L.K; d.K = L.N;
      d.C = bufs[pb]; d.ldc = L.K;
      d.Zin = m.z[l - 1]; d.ldz = L.K;
      if (int rc = gemm(A_MK, B_KN, EP_DGELU, d, 1, s)) return rc;
      dz = bufs[pb];
      pb ^= 1;
    }
  }
  return 0;
}

int mlp_adam(Mlp& m, float lr, float b1, float b2, float eps, float wd, hipStream_t s) {
  m.step += 1;
  const float c1 = 1.f / (1.f - std::pow(b1, (float)m.step));
  const float c2 = 1.f / (1.f - std::pow(b2, (float)m.step));
  for (auto& L : m.L) {
    const int64_t nw = (int64_t)L.N * L.K;
    hipLaunchKernelGGL(k_adam, grid_for(nw), dim3(256), 0, s, L.w, (const float*)L.gw, L.mw, L.vw,
                       nw, lr, b1, b2, eps, wd, c1, c2);
    hipLaunchKernelGGL(k_adam, grid_for(L.N), dim3(256), 0, s, L.b, (const float*)L.gb, L.mb, L.vb,
                       (int64_t)L.N, lr, b1, b2, eps, 0.f, c1, c2);
  }
  UMWM_HIP_TRY(hipGetLastError());
  return mlp_sync_bf16(m, s);
}

namespace {
// Hidden layers 1.. and the transposed last layer of an inference pass
// whose first hidden layer already ran into m.h[0].
int infer_rest(Mlp& m, int rows, bf16* out, int64_t ldo, hipStream_t s, const float* omu,
               const float* osd) {
  const int nl = (int)m.L.size();
  const bf16* in = m.h[0];
  for (int l = 1; l + 1 < nl; ++l) {
    const MlpLayerDev& L = m.L[l];
    GemmArgs g{};
    g.A = in; g.lda = L.K;
    g.B = L.wb; g.ldb = L.K;
    g.M = rows; g.N = L.N; g.K = L.K;
    g.bias = L.b;
    g.C = m.h[l]; g.ldc = L.N;
    if (int rc = gemm_infer(A_MK, EP_GELU, g, s)) return rc;
    in = m.h[l];
  }
  // out[o][j] = sum_k W[o][k] h[j][k] + b[o]: the network's outputs of point
  // j along a contiguous row per output (the spectra's [bin][point] layout)
  const MlpLayerDev& L = m.L.back();
  GemmArgs g{};
  g.A = L.wb; g.lda = L.K;
  g.B = in; g.ldb = L.K;
  g.M = m.nout; g.N = rows; g.K = L.K;
  g.bias = L.b;
  g.C = out; g.ldc = ldo;
  if (omu && osd) {
    g.omu = omu; g.osd = osd;
    return gemm_infer(A_MK, EP_SUR, g, s);
  }
  return gemm_infer(A_MK, EP_BF16_BM, g, s);
}
}  // namespace

int mlp_infer_transposed(Mlp& m, int rows, bf16* out, int64_t ldo, hipStream_t s,
                         const float* omu, const float* osd) {
  if (rows > m.cap_rows) return fail(UMWM_ERR_SIZE, "mlp: more rows than the buffers hold");
  if (m.L.size() < 2) return fail(UMWM_ERR_ARGUMENT, "mlp: inference needs a hidden layer");
  const MlpLayerDev& L = m.L[0];
  GemmArgs g{};
  g.A = m.x; g.lda = L.K;
  g.B = L.wb; g.ldb = L.K;
  g.M = rows; g.N = L.N; g.K = L.K;
  g.bias = L.b;
  g.C = m.h[0]; g.ldc = L.N;
  if (int rc = gemm_infer(A_MK, EP_GELU, g, s)) return rc;
  return infer_rest(m, rows, out, ldo, s, omu, osd);
}

int mlp_infer_spectrum(Mlp& m, const SpecOperand& in, int rows, bf16* out, int64_t ldo,
                       hipStream_t s, const float* omu, const float* osd) {
  if (rows > m.cap_rows) return fail(UMWM_ERR_SIZE, "mlp: more rows than the buffers hold");
  if (m.L.size() < 2) return fail(UMWM_ERR_ARGUMENT, "mlp: inference needs a hidden layer");
  const MlpLayerDev& L = m.L[0];
  if (in.nb + 3 != m.nin) return fail(UMWM_ERR_ARGUMENT, "mlp: spectrum operand width");
  GemmArgs g{};
  g.sa = in;
  g.B = L.wb; g.ldb = L.K;
  g.M = rows; g.N = L.N; g.K = L.K;
  g.bias = L.b;
  g.C = m.h[0]; g.ldc = L.N;
  if (int rc = gemm_infer(A_SPEC, EP_GELU, g, s)) return rc;
  return infer_rest(m, rows, out, ldo, s, omu, osd);
}

}  // namespace umwm
