// surrogate.hip -- coordinate MLP surrogate of activation maps and the warm
// start it feeds (no counterpart in the reference).
//
// Network: 16 inputs (vertex xyz, three pacing-site xyz, four regional speed
// factors), 6 hidden layers of 256 with SiLU, one output (time / t_scale).
// Hidden-layer products run on MFMA (v_mfma_f32_32x32x16_bf16): BF16 weights
// and activations, FP32 accumulation; FP32 master weights and Adam moments.
//
// Device layout of the FP32 master parameters (the first layer is padded to
// 64 inputs, the padding weights stay zero):
//   W0[256 x 64] b0[256]  (W_l[256 x 256] b_l[256]) x 5  w_out[256] b_out
// A training step is three kernels:
//   k_mlp_chain<true>: per tile of 64 rows, the six forward layers with the
//     activations kept in LDS, the output layer and squared error, and the
//     five backward-data layers dZ_{l-1} = (dZ_l W_l) * SiLU'(Z_{l-1});
//     Z_0..Z_4 and dZ_0..dZ_5 leave the tile for the weight gradients;
//   k_mlp_wgrad: dW_l = dZ_l^T H_{l-1} and db_l = column sums of dZ_l as
//     partial sums over spans of rows (H = SiLU(Z) recomputed from Z);
//   k_reduce_parts + k_adam: the sum of the partials, the Adam update and
//     the BF16 weight copies.
// Prediction is k_mlp_chain<false> (forward and output layer only).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "host.hpp"

#define HIPCHECK(x)                                                                   \
  do {                                                                                \
    hipError_t err_ = (x);                                                            \
    if (err_ != hipSuccess)                                                           \
      throw fim::Error(FIM_EDEVICE, std::string(#x) + ": " + hipGetErrorString(err_)); \
  } while (0)

namespace fim {
namespace nn {
namespace {

constexpr int NIN = FIM_NN_INPUTS;  // logical inputs
constexpr int K0 = 64;              // padded inputs
constexpr int HID = FIM_NN_HIDDEN;
constexpr int NL = FIM_NN_LAYERS;
constexpr size_t W0_N = (size_t)HID * K0;
constexpr size_t WL_N = (size_t)HID * HID;

// FP32 master offsets
constexpr size_t off_w(int l) { return l == 0 ? 0 : W0_N + HID + (size_t)(l - 1) * (WL_N + HID); }
constexpr size_t off_b(int l) { return off_w(l) + (l == 0 ? W0_N : WL_N); }
constexpr size_t OFF_WO = off_w(NL);  // = off_b(5) + HID
constexpr size_t OFF_BO = OFF_WO + HID;
constexpr size_t NPAD = OFF_BO + 1;
// BF16 weight copies in MFMA fragment order (below): wb holds W_l as the
// operand of the forward products (l = 0..5), wbt W_l^T as the operand of the
// backward-data products (l = 1..5)
constexpr size_t offb_w(int l) { return l == 0 ? 0 : W0_N + (size_t)(l - 1) * WL_N; }
constexpr size_t NWB = W0_N + (size_t)(NL - 1) * WL_N;
constexpr size_t offbt(int l) { return (size_t)(l - 1) * WL_N; }
constexpr size_t NWBT = (size_t)(NL - 1) * WL_N;
static_assert(off_w(NL) == off_b(NL - 1) + HID, "layout");
static_assert((size_t)HID * NIN + HID + (NL - 1) * (WL_N + HID) + HID + 1 == FIM_NN_PARAMS,
              "logical parameter count");

typedef unsigned short u16x8 __attribute__((ext_vector_type(8)));
typedef unsigned short u16x4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

// float -> bf16, round to nearest even (finite inputs): the hardware
// conversion on the device, the same rounding written out on the host
__host__ __device__ inline uint16_t f2bf(float f) {
#ifdef __HIP_DEVICE_COMPILE__
  return __builtin_bit_cast(uint16_t, (__bf16)f);
#else
  uint32_t u;
  std::memcpy(&u, &f, 4);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
#endif
}
__host__ __device__ inline float bf2f(uint16_t h) {
  const uint32_t u = (uint32_t)h << 16;
  float f;
  std::memcpy(&f, &u, 4);
  return f;
}
// logistic function with the hardware exponential and reciprocal (about 1
// ulp each; the results are rounded to BF16 afterwards)
__device__ inline float sigm(float z) { return __builtin_amdgcn_rcpf(1.0f + __expf(-z)); }
__device__ inline float silu(float z) { return z * sigm(z); }
__device__ inline float dsilu(float z) {
  const float s = sigm(z);
  return s * (1.0f + z * (1.0f - s));
}

// ---------------------------------------------------------------------------
// Layouts of the fused kernels.
//
// k-step layout of a matrix M[rows][C] (activations Z, dZ and the inputs X0):
// rows in groups of 16, each group stored column by column, so the 16 values
// of one column in one group are 32 contiguous bytes:
//   M(row, c) at ((row / 16) * C + c) * 16 + row % 16.
// An MFMA operand fragment whose contraction index is the row (weight
// gradients) is one 16-byte load in this layout.
//
// Fragment order of an operand A[M][K] of v_mfma_f32_32x32x16_bf16: lane l of
// a 32 x 16 fragment holds A(32 mb + l % 32, 16 ks + 8 (l / 32) + 0..7), and
// fragment (mb, ks) is the 1 KB block (mb * K / 16 + ks) -- one coalesced
// 16-byte load per lane.  wb holds the hidden-layer weights W_l[n][k] with
// m = n (forward products), wbt the layers l >= 1 with m = k (backward-data
// products).
//
// The accumulator element r of lane l of a 32 x 32 product is
// (row 8 (r / 4) + 4 (l / 32) + r % 4, col l % 32).
// ---------------------------------------------------------------------------
__host__ __device__ inline size_t ks_off(size_t row, uint32_t c, uint32_t C) {
  return (((row >> 4) * C + c) << 4) + (row & 15u);
}
__host__ __device__ constexpr size_t frag_off(uint32_t m, uint32_t k, uint32_t K) {
  return ((size_t)((m >> 5) * (K >> 4) + (k >> 4)) * 64u + (m & 31u) + 32u * ((k >> 3) & 1u)) * 8u + (k & 7u);
}

__device__ inline f32x16 mfma(const u16x8 a, const u16x8 b, const f32x16 c) {
  return __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(bf16x8, a), __builtin_bit_cast(bf16x8, b), c,
                                                 0, 0, 0);
}

constexpr int RG = 1;         // row groups of 64 per tile of the fused chain kernel
constexpr int TR = 64 * RG;   // rows per tile
constexpr int NT = 256 * RG;  // threads per workgroup: four waves per row group
constexpr int HP = HID + 8;   // LDS row pitch of the tile (BF16 values)
// LDS of the chain kernel: the tile, w_out, dy per row, loss and
// bias-gradient sums
constexpr size_t CHAIN_LDS = (size_t)TR * HP * 2 + (HID + TR + 2) * 4;

// One layer's product for the tile: acc[i][j] (rows 64 rg + 32 i, columns
// 64 fg + 32 j) = T(rows, 0..K) B, with T the tile in LDS (row-major, pitch
// HP) and B given by fragments (fragment (mb, ks) of B^T at bfr).  The next
// step's weight fragments are loaded before the current step's products.
template <int K>
__device__ inline void tile_product(const uint16_t *Ts, const uint16_t *__restrict__ bfr, uint32_t rg, uint32_t fg,
                                    uint32_t lane, f32x16 (&acc)[2][2]) {
  constexpr int NS = K / 16;
  const uint32_t l32 = lane & 31u, lh = lane >> 5;
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int r = 0; r < 16; ++r) acc[i][j][r] = 0.0f;
  const u16x8 *bp0 = reinterpret_cast<const u16x8 *>(bfr) + (size_t)(2 * fg) * NS * 64 + lane;
  const u16x8 *bp1 = bp0 + NS * 64;
  const uint16_t *ta0 = Ts + (64 * rg + l32) * HP + 8 * lh;
  const uint16_t *ta1 = ta0 + 32 * HP;
  // weight fragments two steps ahead, tile fragments one step ahead
  u16x8 b[2][2], t[2][2];
  b[0][0] = bp0[0];
  b[0][1] = bp1[0];
  b[1][0] = bp0[64];
  b[1][1] = bp1[64];
  t[0][0] = *reinterpret_cast<const u16x8 *>(ta0);
  t[0][1] = *reinterpret_cast<const u16x8 *>(ta1);
#pragma unroll
  for (int s = 0; s < NS; ++s) {
    const int c = s & 1;
    if (s + 1 < NS) {
      t[c ^ 1][0] = *reinterpret_cast<const u16x8 *>(ta0 + 16 * (s + 1));
      t[c ^ 1][1] = *reinterpret_cast<const u16x8 *>(ta1 + 16 * (s + 1));
    }
    acc[0][0] = mfma(t[c][0], b[c][0], acc[0][0]);
    acc[0][1] = mfma(t[c][0], b[c][1], acc[0][1]);
    acc[1][0] = mfma(t[c][1], b[c][0], acc[1][0]);
    acc[1][1] = mfma(t[c][1], b[c][1], acc[1][1]);
    if (s + 2 < NS) {
      b[c][0] = bp0[(s + 2) * 64];
      b[c][1] = bp1[(s + 2) * 64];
    }
    // keep each step's loads in place (no hoisting across steps): the
    // fragments stay two steps ahead without holding the whole layer
    __builtin_amdgcn_sched_barrier(0);
  }
}

struct ChainArgs {
  const uint16_t *X0;        // inputs, k-step layout, K0 columns
  const uint16_t *wb, *wbt;  // fragment-ordered BF16 weights
  const float *w;            // FP32 masters (biases, output layer)
  const float *tgt;          // training: targets / t_scale
  uint16_t *Z[NL - 1];       // training: Z_0..Z_4, k-step layout
  uint16_t *dZ[NL];          // training: dZ_0..dZ_5, k-step layout
  float *g, *loss;           // training: output-layer gradient (accumulated), loss sum
  float *out;                // inference: predictions
  uint32_t nreal;            // rows with data; the rest of the last tile is padding
  float inv_b, t_scale, delta;
};

// The whole network for one tile of TR = 64 RG rows, activations in LDS
// between layers.  Four waves per row group: rows 64 rg .. +63 (rg = wave /
// 4; RG = 1, so one group of four waves), columns 64 fg .. +63 (fg = wave % 4)
// of every product, as 2 x 2 accumulators of 32 x 32.
//   forward: Z_l = H_{l-1} W_l^T + b_l, H_l = SiLU(Z_l) (BF16 in LDS; Z_l to
//     HBM in the k-step layout when training);
//   output: y = H_5 . w_out + b_out (inference: t = y t_scale + delta);
//   training: e = y - t, dZ_5 = (2 e / B) w_out SiLU'(Z_5), the output-layer
//     gradient, then dZ_{l-1} = (dZ_l W_l) * SiLU'(Z_{l-1}) for l = 5..1 with
//     dZ_l in LDS; every dZ_l goes to HBM in the k-step layout.
// Rounding to BF16 where the separate-kernel formulation stores values:
// Z, H and dZ; SiLU'(Z) of the stored Z.
template <bool TRAIN>
__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(4, 4))) void k_mlp_chain(ChainArgs a) {
  extern __shared__ __align__(16) unsigned char chain_lds[];
  uint16_t *Ts = reinterpret_cast<uint16_t *>(chain_lds);
  float *wos = reinterpret_cast<float *>(chain_lds + TR * HP * 2);  // w_out
  float *dys = wos + HID;
  float *misc = dys + TR;
  const uint32_t tid = threadIdx.x, lane = tid & 63u, wave = tid >> 6;
  const uint32_t rg = wave >> 2, fg = wave & 3u, l32 = lane & 31u, lh = lane >> 5;
  const size_t r0 = (size_t)blockIdx.x * TR;
  // k-step offset, within the tile's rows, of the four accumulator values
  // (rows 64 rg + 32 i + 8 q + 4 lh + 0..3, column 64 fg + 32 j + l32) of
  // block (i, j), group q: ((row / 16) HID + column) 16 + row % 16
  const uint32_t ks_lane = ((4 * rg * HID + 64 * fg + l32) << 4) + 4 * lh;
  auto ks_acc = [&](int i, int j, int q) -> uint32_t {
    return ks_lane + i * 32 * HID + (q >> 1) * 16 * HID + j * 32 * 16 + (q & 1) * 8;
  };
  if (tid < 2) misc[tid] = 0.0f;
  for (uint32_t c = tid; c < HID; c += NT) wos[c] = a.w[OFF_WO + c];
  // inputs: 16-byte pieces of the k-step layout (8 rows of one column)
  for (uint32_t p = tid; p < TR * K0 / 8; p += NT) {
    const uint32_t kb = p / (2 * K0), c = (p >> 1) % K0, hf = p & 1u;
    const u16x8 v = *reinterpret_cast<const u16x8 *>(a.X0 + ((((r0 >> 4) + kb) * K0 + c) << 4) + 8 * hf);
#pragma unroll
    for (int j = 0; j < 8; ++j) Ts[(kb * 16 + 8 * hf + j) * HP + c] = v[j];
  }
  __syncthreads();
  f32x16 acc[2][2];
#pragma unroll 1
  for (int l = 0; l < NL; ++l) {
    if (l == 0)
      tile_product<K0>(Ts, a.wb + offb_w(0), rg, fg, lane, acc);
    else
      tile_product<HID>(Ts, a.wb + offb_w(l), rg, fg, lane, acc);
    __syncthreads();  // every wave has read the layer's input
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      const uint32_t f = 64 * fg + 32 * j + l32;
      const float bias = a.w[off_b(l) + f];
#pragma unroll
      for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int r = 0; r < 16; ++r) acc[i][j][r] += bias;
    }
    uint16_t *zl = TRAIN && l < NL - 1 ? a.Z[l] + r0 * HID : nullptr;
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
      for (int j = 0; j < 2; ++j) {
        const uint32_t f = 64 * fg + 32 * j + l32;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const uint32_t R = 64 * rg + 32 * i + 8 * q + 4 * lh;
          if (zl) {
            u16x4 zv;
#pragma unroll
            for (int c = 0; c < 4; ++c) zv[c] = f2bf(acc[i][j][4 * q + c]);
            *reinterpret_cast<u16x4 *>(zl + ks_acc(i, j, q)) = zv;
          }
#pragma unroll
          for (int c = 0; c < 4; ++c) Ts[(R + c) * HP + f] = f2bf(silu(acc[i][j][4 * q + c]));
        }
      }
    __syncthreads();
  }
  // output layer: H_5 in LDS, Z_5 (FP32) in acc.  Four threads per row, 64
  // columns each.
  {
    const uint32_t row = tid >> 2, c0 = (tid & 3u) * 64;
    float p = 0.0f;
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const u16x8 h = *reinterpret_cast<const u16x8 *>(Ts + row * HP + c0 + 8 * u);
#pragma unroll
      for (int k = 0; k < 8; ++k) p += bf2f(h[k]) * wos[c0 + 8 * u + k];
    }
    p += __shfl_xor(p, 1);
    p += __shfl_xor(p, 2);
    const float y = p + a.w[OFF_BO];
    const size_t grow = r0 + row;
    if (!TRAIN) {
      if ((tid & 3u) == 0 && grow < a.nreal) a.out[grow] = y * a.t_scale + a.delta;
      return;
    }
    const float e = grow < a.nreal ? y - a.tgt[grow] : 0.0f;
    const float dy = 2.0f * e * a.inv_b;
    if ((tid & 3u) == 0) dys[row] = dy;
    float ls = (tid & 3u) == 0 ? e * e : 0.0f, gb = (tid & 3u) == 0 ? dy : 0.0f;
    for (int off = 32; off > 0; off >>= 1) {
      ls += __shfl_xor(ls, off);
      gb += __shfl_xor(gb, off);
    }
    if (lane == 0) {
      atomicAdd(&misc[0], ls);
      atomicAdd(&misc[1], gb);
    }
  }
  __syncthreads();
  // output-layer gradient sum_rows dy H_5 (one column per thread)
  if (tid < HID) {
    float gw = 0.0f;
    for (uint32_t row = 0; row < TR; ++row) gw += dys[row] * bf2f(Ts[row * HP + tid]);
    atomicAdd(&a.g[OFF_WO + tid], gw);
  }
  if (tid == 0) {
    atomicAdd(a.loss, misc[0]);
    atomicAdd(&a.g[OFF_BO], misc[1]);
  }
  __syncthreads();  // H_5 is read; dZ_5 takes its place
  // dZ_5 = dy w_out SiLU'(Z_5)
#pragma unroll
  for (int j = 0; j < 2; ++j) {
    const uint32_t f = 64 * fg + 32 * j + l32;
    const float wo = wos[f];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const uint32_t R = 64 * rg + 32 * i + 8 * q + 4 * lh;
        u16x4 dv;
#pragma unroll
        for (int c = 0; c < 4; ++c) {
          dv[c] = f2bf(dys[R + c] * wo * dsilu(bf2f(f2bf(acc[i][j][4 * q + c]))));
          Ts[(R + c) * HP + f] = dv[c];
        }
        *reinterpret_cast<u16x4 *>(a.dZ[NL - 1] + r0 * HID + ks_acc(i, j, q)) = dv;
      }
  }
  __syncthreads();
  // backward through the hidden layers
#pragma unroll 1
  for (int l = NL - 1; l >= 1; --l) {
    tile_product<HID>(Ts, a.wbt + offbt(l), rg, fg, lane, acc);
    __syncthreads();  // every wave has read dZ_l
#pragma unroll
    for (int i = 0; i < 2; ++i) {
      u16x4 zp[2][4];  // Z_{l-1} of this row block, loaded before any is used
#pragma unroll
      for (int j = 0; j < 2; ++j)
#pragma unroll
        for (int q = 0; q < 4; ++q)
          zp[j][q] = *reinterpret_cast<const u16x4 *>(a.Z[l - 1] + r0 * HID + ks_acc(i, j, q));
#pragma unroll
      for (int j = 0; j < 2; ++j) {
        const uint32_t k = 64 * fg + 32 * j + l32;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const uint32_t R = 64 * rg + 32 * i + 8 * q + 4 * lh;
          u16x4 dv;
#pragma unroll
          for (int c = 0; c < 4; ++c) {
            dv[c] = f2bf(acc[i][j][4 * q + c] * dsilu(bf2f(zp[j][q][c])));
            Ts[(R + c) * HP + k] = dv[c];
          }
          *reinterpret_cast<u16x4 *>(a.dZ[l - 1] + r0 * HID + ks_acc(i, j, q)) = dv;
        }
      }
    }
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// Weight gradients dW_l = dZ_l^T H_{l-1} and db_l = column sums of dZ_l over
// one span of rows per workgroup (blockIdx.x = span, blockIdx.y = layer),
// written as that span's partial sum (P[span][off_w(l) ..], the master
// layout).  H_{l-1} = SiLU(Z_{l-1}) for l >= 1, X0 for l = 0.  Operands are
// staged 32 rows at a time into LDS in their k-step layout (SiLU applied
// once per element); every fragment is then one 16-byte LDS read.  Eight
// waves tile the 256 x K output: 4 x 2 waves of 64 x 128 (K = 256) or 8 x 1
// waves of 32 x 64 (K = 64).
// ---------------------------------------------------------------------------
struct WGradArgs {
  const uint16_t *dZ[NL];   // k-step layout, HID columns
  const uint16_t *Hin[NL];  // l = 0: X0 (K0 columns); l >= 1: Z_{l-1} (HID columns)
  float *P;                 // partial sums, OFF_WO floats per span
  uint32_t steps;           // 32-row steps of the batch; span s takes steps
                            // [s steps / S, (s + 1) steps / S), S = gridDim.x
};

// LDS of the weight-gradient kernel: two buffers of dZ and H, each two
// 16-row k-steps x 256 columns; a column's 16 rows take 24 values (48 B), so
// the 32 lanes of a fragment read hit distinct banks
constexpr int KSP = 24;
constexpr int WG_LDS = 2 * 2 * 2 * HID * KSP;  // BF16 values

template <int K>
__device__ inline void wgrad_body(const WGradArgs &a, uint32_t l, uint16_t *lds) {
  constexpr int WN = K == HID ? 4 : 8, WK = 8 / WN;
  constexpr int BN = HID / WN / 32, BK = K / WK / 32;
  constexpr uint32_t BUF = 2 * 2 * HID * KSP;  // one buffer: dZ (2 x HID x KSP) then H (2 x K x KSP)
  constexpr uint32_t HPC = 2 * K * 16 / 8;    // 16-byte pieces of H per step
  const uint32_t tid = threadIdx.x, lane = tid & 63u, wave = tid >> 6;
  const uint32_t wn = wave / WK, wk = wave % WK, l32 = lane & 31u, lh = lane >> 5;
  const uint32_t s0 = (uint32_t)((uint64_t)blockIdx.x * a.steps / gridDim.x);
  const uint32_t s1 = (uint32_t)((uint64_t)(blockIdx.x + 1) * a.steps / gridDim.x);
  const uint32_t n = s1 - s0;                   // >= 1: the host launches at most a.steps spans
  const uint32_t npad = (n + 5) / 6 * 6;        // steps run: the last ones stage zeros
  const uint16_t *dz = a.dZ[l], *hz = a.Hin[l];  // k-step blocks of consecutive rows are contiguous
  // Three register slots of global loads: relative step i + 4 is loaded while
  // step i is multiplied and step i + 1 is staged.  Every load is issued (the
  // step index clamped), so the number of loads in flight is the same on
  // every path and each wait leaves the two younger steps in flight.
  u16x8 rz[3][2], rh[3][2];
  float csum = 0.0f;  // column sum of dZ for column tid / 2
  auto gload = [&](int slot, uint32_t i) {
    const uint32_t st = s0 + (i < n ? i : n - 1);
    const u16x8 *sz = reinterpret_cast<const u16x8 *>(dz + (size_t)st * 32 * HID);
    const u16x8 *sh = reinterpret_cast<const u16x8 *>(hz + (size_t)st * 32 * K);
    rz[slot][0] = sz[tid];
    rz[slot][1] = sz[tid + 512];
#pragma unroll
    for (int u = 0; u < 2; ++u) rh[slot][u] = sh[(tid + 512u * u) % HPC];
  };
  // stages relative step i (zeros beyond the span) into LDS buffer buf
  auto lstore = [&](int slot, uint32_t buf, uint32_t i) {
    const uint32_t msk = i < n ? 0xffffffffu : 0u;  // zeros beyond the span (SiLU(0) = 0)
    auto masked = [&](u16x8 v) {
      uint4 w = __builtin_bit_cast(uint4, v);
      w.x &= msk;
      w.y &= msk;
      w.z &= msk;
      w.w &= msk;
      return __builtin_bit_cast(u16x8, w);
    };
    const u16x8 z0 = masked(rz[slot][0]), z1 = masked(rz[slot][1]);
    // piece p of a step (16 B: 8 rows of one column) goes to column p / 2 of
    // its k-step, half p % 2
    uint16_t *bz = lds + buf * BUF, *bh = lds + buf * BUF + 2 * HID * KSP;
    *reinterpret_cast<u16x8 *>(bz + (tid >> 1) * KSP + 8 * (tid & 1u)) = z0;
    *reinterpret_cast<u16x8 *>(bz + (HID + (tid >> 1)) * KSP + 8 * (tid & 1u)) = z1;
#pragma unroll
    for (int j = 0; j < 8; ++j) csum += bf2f(z0[j]) + bf2f(z1[j]);
#pragma unroll
    for (int u = 0; u < 2; ++u) {
      const uint32_t p = tid + 512u * u;
      if (p < HPC) {
        u16x8 v = masked(rh[slot][u]);
        if (K == HID)
#pragma unroll
          for (int j = 0; j < 8; ++j) v[j] = f2bf(silu(bf2f(v[j])));
        *reinterpret_cast<u16x8 *>(bh + (p >> 1) * KSP + 8 * (p & 1u)) = v;
      }
    }
  };
  f32x16 acc[BN][BK];
#pragma unroll
  for (int i = 0; i < BN; ++i)
#pragma unroll
    for (int j = 0; j < BK; ++j)
#pragma unroll
      for (int r = 0; r < 16; ++r) acc[i][j][r] = 0.0f;
  auto product = [&](uint32_t buf) {
    const uint16_t *bz = lds + buf * BUF, *bh = lds + buf * BUF + 2 * HID * KSP;
#pragma unroll
    for (int u = 0; u < 2; ++u) {  // two 16-row k-steps
      u16x8 fa[BN], fb[BK];
#pragma unroll
      for (int i = 0; i < BN; ++i)
        fa[i] = *reinterpret_cast<const u16x8 *>(bz + (u * HID + wn * BN * 32 + 32 * i + l32) * KSP + 8 * lh);
#pragma unroll
      for (int j = 0; j < BK; ++j)
        fb[j] = *reinterpret_cast<const u16x8 *>(bh + (u * K + wk * BK * 32 + 32 * j + l32) * KSP + 8 * lh);
#pragma unroll
      for (int i = 0; i < BN; ++i)
#pragma unroll
        for (int j = 0; j < BK; ++j) acc[i][j] = mfma(fa[i], fb[j], acc[i][j]);
    }
  };
  // relative steps 6 m + r use LDS buffer r % 2 and register slot r % 3
  gload(0, 0);
  lstore(0, 0, 0);
  gload(1, 1);
  gload(2, 2);
  gload(0, 3);
  __syncthreads();
  for (uint32_t i0 = 0; i0 < npad; i0 += 6) {
#pragma unroll
    for (int r = 0; r < 6; ++r) {
      product((uint32_t)(r & 1));
      lstore((r + 1) % 3, (uint32_t)((r + 1) & 1), i0 + r + 1);
      gload((r + 1) % 3, i0 + r + 4);
      __syncthreads();
    }
  }
  float *P = a.P + (size_t)blockIdx.x * OFF_WO;
#pragma unroll
  for (int i = 0; i < BN; ++i)
#pragma unroll
    for (int j = 0; j < BK; ++j)
#pragma unroll
      for (int r = 0; r < 16; ++r) {
        const uint32_t n = wn * BN * 32 + 32 * i + 8 * (r >> 2) + 4 * lh + (r & 3);
        const uint32_t k = wk * BK * 32 + 32 * j + l32;
        P[off_w(l) + (size_t)n * K + k] = acc[i][j][r];
      }
  csum += __shfl_xor(csum, 1);
  if ((tid & 1u) == 0) P[off_b(l) + (tid >> 1)] = csum;
}

__global__ __launch_bounds__(512) void k_mlp_wgrad(WGradArgs a) {
  extern __shared__ __align__(16) unsigned char wg_lds[];
  uint16_t *lds = reinterpret_cast<uint16_t *>(wg_lds);
  const uint32_t l = blockIdx.y;
  if (l == 0)
    wgrad_body<K0>(a, 0, lds);
  else
    wgrad_body<HID>(a, l, lds);
}

// ---------------------------------------------------------------------------
// small kernels
// ---------------------------------------------------------------------------
__device__ inline uint64_t mix64(uint64_t x) {
  x += 0x9E3779B97F4A7C15ull;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
  return x ^ (x >> 31);
}

// per map: 13 floats = three normalised site positions, four factors - 1
constexpr int MAPP = 16;

// the 16 inputs of one row into X0 (k-step layout; the padding columns
// NIN..K0-1 are zero from the start and never written)
__device__ inline void write_features(uint16_t *X0, size_t row, const float *vxyz, const float *mp) {
  for (int k = 0; k < 3; ++k) X0[ks_off(row, k, K0)] = f2bf(vxyz[k]);
  for (int k = 0; k < 13; ++k) X0[ks_off(row, 3 + k, K0)] = f2bf(mp[k]);
}

__device__ inline void zero_features(uint16_t *X0, size_t row) {
  for (int k = 0; k < NIN; ++k) X0[ks_off(row, k, K0)] = 0;
}

// training rows: random (map, vertex) samples; rows B..rows-1 (padding of
// the last tile) are zero with target 0
__global__ void k_sample(const float *coords, const float *mparams, const float *T, uint32_t nn,
                         uint32_t nmaps, uint64_t seed, uint32_t B, uint32_t rows, float inv_ts, uint16_t *X0,
                         float *tgt) {
  for (uint32_t m = blockIdx.x * blockDim.x + threadIdx.x; m < rows; m += gridDim.x * blockDim.x) {
    if (m >= B) {
      zero_features(X0, m);
      tgt[m] = 0.0f;
      continue;
    }
    const uint64_t h = mix64(seed ^ mix64((uint64_t)m + 1));
    const uint32_t j = (uint32_t)((h >> 32) % nmaps), v = (uint32_t)((h & 0xffffffffu) % nn);
    write_features(X0, m, coords + 3 * (size_t)v, mparams + MAPP * (size_t)j);
    tgt[m] = T[(size_t)j * nn + v] * inv_ts;
  }
}

// rows of one map: every vertex (padding rows repeat vertex 0)
__global__ void k_map_rows(const float *coords, const float *mp, uint32_t nn, uint32_t rows,
                           uint16_t *X0) {
  for (uint32_t m = blockIdx.x * blockDim.x + threadIdx.x; m < rows; m += gridDim.x * blockDim.x)
    write_features(X0, m, coords + 3 * (size_t)(m < nn ? m : 0), mp);
}

// rows of several solves over one vertex sample: row r is solve r / m at
// sample vertex r % m (padding rows repeat the last real row)
__global__ void k_sample_rows(const float *coords, const float *mp, const uint32_t *sample, uint32_t m,
                              uint32_t nreal, uint32_t rows, uint16_t *X0) {
  for (uint32_t r = blockIdx.x * blockDim.x + threadIdx.x; r < rows; r += gridDim.x * blockDim.x) {
    const uint32_t q = r < nreal ? r : nreal - 1;
    const uint32_t j = q / m, v = sample[q - j * m];
    write_features(X0, r, coords + 3 * (size_t)v, mp + MAPP * (size_t)j);
  }
}

// per solve: the largest of its m predictions (one wave per solve)
__global__ __launch_bounds__(256) void k_solve_max(const float *t, uint32_t m, uint32_t nsolve, float *out) {
  const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  for (uint32_t j = blockIdx.x * 4 + wave; j < nsolve; j += gridDim.x * 4) {
    float mx = -__builtin_huge_valf();
    for (uint32_t i = lane; i < m; i += 64) mx = fmaxf(mx, t[(size_t)j * m + i]);
    for (int off = 32; off > 0; off >>= 1) mx = fmaxf(mx, __shfl_xor(mx, off));
    if (lane == 0) out[j] = mx;
  }
}

// given inputs (n x 16 floats) -> BF16 rows of X0 (rows n..rows-1 zero)
__global__ void k_input_rows(const float *x, uint32_t n, uint32_t rows, uint16_t *X0) {
  for (uint32_t m = blockIdx.x * blockDim.x + threadIdx.x; m < rows; m += gridDim.x * blockDim.x)
    for (int k = 0; k < NIN; ++k) X0[ks_off(m, k, K0)] = m < n ? f2bf(x[(size_t)m * NIN + k]) : (uint16_t)0;
}

// sum of the batch-split partials
// (eight loads in flight per thread: partial k goes to accumulator k % 8,
// and the eight are added in a fixed order, so the sum is reproducible)
__global__ void k_reduce_parts(const float *P, uint32_t S, size_t n, float *g) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
    float a[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    uint32_t k = 0;
    for (; k + 8 <= S; k += 8) {
#pragma unroll
      for (int j = 0; j < 8; ++j) a[j] += P[(size_t)(k + j) * n + i];
    }
    for (; k < S; ++k) a[k & 7u] += P[(size_t)k * n + i];
    g[i] = ((a[0] + a[1]) + (a[2] + a[3])) + ((a[4] + a[5]) + (a[6] + a[7]));
  }
}

// Adam on the FP32 masters; BF16 copies of the hidden-layer weights
struct AdamArgs {
  float *w, *m, *v;
  const float *g;
  uint16_t *wb, *wbt;
  float lr, b1, b2, eps, c1, c2;  // c1 = 1 - b1^t, c2 = 1 - b2^t
};

__global__ void k_adam(AdamArgs a) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < NPAD; i += (size_t)gridDim.x * blockDim.x) {
    const float g = a.g[i];
    const float m = a.b1 * a.m[i] + (1.0f - a.b1) * g;
    const float v = a.b2 * a.v[i] + (1.0f - a.b2) * g * g;
    a.m[i] = m;
    a.v[i] = v;
    const float w = a.w[i] - a.lr * (m / a.c1) / (sqrtf(v / a.c2) + a.eps);
    a.w[i] = w;
    // weights of layer l live at off_w(l) .. off_b(l) in the masters
    for (int l = 0; l < NL; ++l)
      if (i >= off_w(l) && i < off_b(l)) {
        const size_t q = i - off_w(l);
        const uint32_t K = l == 0 ? K0 : HID, n = (uint32_t)(q / K), k = (uint32_t)(q % K);
        a.wb[offb_w(l) + frag_off(n, k, K)] = f2bf(w);
        if (l > 0) a.wbt[offbt(l) + frag_off(k, n, HID)] = f2bf(w);
      }
  }
}

__global__ void k_to_bf16(const float *w, uint16_t *wb, uint16_t *wbt) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < NPAD; i += (size_t)gridDim.x * blockDim.x)
    for (int l = 0; l < NL; ++l)
      if (i >= off_w(l) && i < off_b(l)) {
        const size_t q = i - off_w(l);
        const uint32_t K = l == 0 ? K0 : HID, n = (uint32_t)(q / K), k = (uint32_t)(q % K);
        wb[offb_w(l) + frag_off(n, k, K)] = f2bf(w[i]);
        if (l > 0) wbt[offbt(l) + frag_off(k, n, HID)] = f2bf(w[i]);
      }
}

__global__ void k_zero_sites(const uint32_t *v, uint32_t n, float *f) {
  if (threadIdx.x < n) f[v[threadIdx.x]] = 0.0f;
}

int grid_for(size_t n) { return (int)std::max<size_t>(1, std::min<size_t>((n + 255) / 256, 65536)); }

template <class T>
T *dalloc(size_t n) {
  void *p = nullptr;
  HIPCHECK(hipMalloc(&p, std::max<size_t>(n, 1) * sizeof(T)));
  return static_cast<T *>(p);
}

}  // namespace
}  // namespace nn
}  // namespace fim

using namespace fim;
using namespace fim::nn;

struct fim_nn {
  int dev = 0;
  hipStream_t stream = nullptr;
  fim_nn_opts o{};
  uint32_t nn = 0;
  std::vector<float> coords_h;  // normalised vertex coordinates
  double ctr[3] = {0, 0, 0}, half = 1.0;
  // parameters
  float *w = nullptr, *g = nullptr, *m = nullptr, *v = nullptr;
  uint16_t *wb = nullptr, *wbt = nullptr;
  int64_t t = 0;  // Adam steps taken
  // activations (rows = cap, a multiple of the chain kernel's tile), k-step layout
  uint32_t cap = 0;
  uint16_t *X0 = nullptr, *Z[NL - 1] = {}, *dZ[NL] = {};
  float *tgt = nullptr, *out = nullptr, *parts = nullptr, *loss = nullptr;
  uint32_t nsplit = 1;  // row spans of the weight gradients (partial sums)
  // dataset
  float *coords = nullptr, *T = nullptr, *mparams = nullptr;
  int32_t nmaps = 0, max_maps = 0;
  std::vector<void *> owned;
};

namespace {

template <class T>
T *nn_alloc(fim_nn *n, size_t count) {
  T *p = nn::dalloc<T>(count);
  n->owned.push_back(p);
  return p;
}

void nn_destroy(fim_nn *n) {
  if (!n) return;
  for (void *p : n->owned) (void)hipFree(p);
  if (n->stream) (void)hipStreamDestroy(n->stream);
  delete n;
}

// normalised per-map parameters (13 values, 16 stored)
std::vector<float> map_params(const fim_nn *n, const int64_t *sites, int32_t ns, const float *scales) {
  if (ns < 1 || ns > 3) throw Error(FIM_EINVAL, "a solve needs 1 to 3 pacing sites");
  std::vector<float> p(MAPP, 0.0f);
  for (int k = 0; k < 3; ++k) {
    const int64_t s = sites[k < ns ? k : 0];
    if (s < 0 || s >= (int64_t)n->nn) throw Error(FIM_EINVAL, "pacing site out of range");
    for (int d = 0; d < 3; ++d) p[3 * k + d] = n->coords_h[3 * s + d];
  }
  for (int k = 0; k < 4; ++k) {
    if (!(scales[k] > 0.0f)) throw Error(FIM_EINVAL, "speed factors must be positive");
    p[9 + k] = scales[k] - 1.0f;
  }
  return p;
}

// activation buffers for `rows` rows
void ensure_rows(fim_nn *n, uint32_t rows) {
  if (rows <= n->cap) return;
  throw Error(FIM_ELIMIT, "batch larger than the surrogate's row capacity");
}

// rows rounded up to whole tiles of the chain kernel
uint32_t tile_rows(uint32_t rows) { return (rows + TR - 1) / TR * TR; }

float t_scale(const fim_nn *n) { return (float)(n->o.t_scale > 0 ? n->o.t_scale : 1e5); }

// predictions for rows 0..nreal-1 of X0 (filled up to tile_rows(nreal)):
// out[r] = y_r t_scale + delta
void infer(fim_nn *n, uint32_t nreal, float delta, float *out) {
  ChainArgs a{};
  a.X0 = n->X0;
  a.wb = n->wb;
  a.wbt = n->wbt;
  a.w = n->w;
  a.out = out;
  a.nreal = nreal;
  a.t_scale = t_scale(n);
  a.delta = delta;
  k_mlp_chain<false><<<tile_rows(nreal) / TR, NT, CHAIN_LDS, n->stream>>>(a);
  HIPCHECK(hipGetLastError());
}

// forward and backward pass over rows 0..B-1 of X0 (targets in n->tgt; rows
// up to tile_rows(B) filled, as padding): gradient in n->g, loss sum in
// n->loss
void train_pass(fim_nn *n, uint32_t B) {
  const uint32_t rows = tile_rows(B);
  HIPCHECK(hipMemsetAsync(n->g + OFF_WO, 0, (NPAD - OFF_WO) * sizeof(float), n->stream));
  HIPCHECK(hipMemsetAsync(n->loss, 0, sizeof(float), n->stream));
  ChainArgs a{};
  a.X0 = n->X0;
  a.wb = n->wb;
  a.wbt = n->wbt;
  a.w = n->w;
  a.tgt = n->tgt;
  for (int l = 0; l < NL - 1; ++l) a.Z[l] = n->Z[l];
  for (int l = 0; l < NL; ++l) a.dZ[l] = n->dZ[l];
  a.g = n->g;
  a.loss = n->loss;
  a.nreal = B;
  a.inv_b = 1.0f / (float)B;
  k_mlp_chain<true><<<rows / TR, NT, CHAIN_LDS, n->stream>>>(a);
  // weight gradients: S spans of whole 32-row steps, one partial sum each
  const uint32_t S = std::min<uint32_t>(n->nsplit, rows / 32);
  WGradArgs wa;
  for (int l = 0; l < NL; ++l) {
    wa.dZ[l] = n->dZ[l];
    wa.Hin[l] = l == 0 ? n->X0 : n->Z[l - 1];
  }
  wa.P = n->parts;
  wa.steps = rows / 32;
  k_mlp_wgrad<<<dim3(S, NL), 512, WG_LDS * 2, n->stream>>>(wa);
  k_reduce_parts<<<grid_for(OFF_WO), 256, 0, n->stream>>>(n->parts, S, OFF_WO, n->g);
  HIPCHECK(hipGetLastError());
}

void adam(fim_nn *n) {
  n->t += 1;
  AdamArgs a;
  a.w = n->w;
  a.m = n->m;
  a.v = n->v;
  a.g = n->g;
  a.wb = n->wb;
  a.wbt = n->wbt;
  a.lr = (float)(n->o.lr > 0 ? n->o.lr : 1e-3);
  a.b1 = (float)(n->o.beta1 > 0 ? n->o.beta1 : 0.9);
  a.b2 = (float)(n->o.beta2 > 0 ? n->o.beta2 : 0.999);
  a.eps = (float)(n->o.adam_eps > 0 ? n->o.adam_eps : 1e-8);
  a.c1 = (float)(1.0 - std::pow((double)a.b1, (double)n->t));
  a.c2 = (float)(1.0 - std::pow((double)a.b2, (double)n->t));
  k_adam<<<grid_for(NPAD), 256, 0, n->stream>>>(a);
  HIPCHECK(hipGetLastError());
}

// logical <-> device parameter order
void to_device_order(const float *logical, std::vector<float> &dev) {
  dev.assign(NPAD, 0.0f);
  size_t q = 0;
  for (int r = 0; r < HID; ++r)
    for (int k = 0; k < NIN; ++k) dev[off_w(0) + (size_t)r * K0 + k] = logical[q++];
  for (int r = 0; r < HID; ++r) dev[off_b(0) + r] = logical[q++];
  for (int l = 1; l < NL; ++l) {
    for (size_t i = 0; i < WL_N; ++i) dev[off_w(l) + i] = logical[q++];
    for (int r = 0; r < HID; ++r) dev[off_b(l) + r] = logical[q++];
  }
  for (int r = 0; r < HID; ++r) dev[OFF_WO + r] = logical[q++];
  dev[OFF_BO] = logical[q++];
}

void to_logical_order(const std::vector<float> &dev, float *logical) {
  size_t q = 0;
  for (int r = 0; r < HID; ++r)
    for (int k = 0; k < NIN; ++k) logical[q++] = dev[off_w(0) + (size_t)r * K0 + k];
  for (int r = 0; r < HID; ++r) logical[q++] = dev[off_b(0) + r];
  for (int l = 1; l < NL; ++l) {
    for (size_t i = 0; i < WL_N; ++i) logical[q++] = dev[off_w(l) + i];
    for (int r = 0; r < HID; ++r) logical[q++] = dev[off_b(l) + r];
  }
  for (int r = 0; r < HID; ++r) logical[q++] = dev[OFF_WO + r];
  logical[q++] = dev[OFF_BO];
}

template <class F>
int guarded(F &&f) {
  try {
    f();
    return FIM_OK;
  } catch (const Error &e) {
    set_last_error(e.what());
    return e.code;
  } catch (const std::exception &e) {
    set_last_error(e.what());
    return FIM_EINVAL;
  } catch (...) {
    set_last_error("unknown error");
    return FIM_EINVAL;
  }
}

// prediction for all vertices of one solve into dst (device, nn floats)
void predict_map_device(fim_nn *n, const std::vector<float> &mp, float delta, float *dst) {
  const uint32_t rows = tile_rows(n->nn);
  ensure_rows(n, rows);
  HIPCHECK(hipMemcpyAsync(n->mparams + (size_t)n->max_maps * MAPP, mp.data(), MAPP * sizeof(float),
                          hipMemcpyHostToDevice, n->stream));
  k_map_rows<<<grid_for(rows), 256, 0, n->stream>>>(n->coords, n->mparams + (size_t)n->max_maps * MAPP,
                                                    n->nn, rows, n->X0);
  infer(n, n->nn, delta, dst);
}

}  // namespace

extern "C" {

int fim_nn_create(const fim_mesh *mesh, const fim_nn_opts *o, fim_nn **out) {
  return guarded([&] {
    if (!mesh || !o || !out) throw Error(FIM_EINVAL, "null argument");
    *out = nullptr;
    const Mesh &mm = mesh->m;
    if (mm.nn <= 0 || mm.nn >= (1ll << 31)) throw Error(FIM_EINVAL, "mesh size");
    fim_nn *n = new fim_nn();
    try {
      n->o = *o;
      n->dev = o->device;
      n->nn = (uint32_t)mm.nn;
      HIPCHECK(hipSetDevice(n->dev));
      HIPCHECK(hipStreamCreateWithFlags(&n->stream, hipStreamNonBlocking));
      // coordinates: centred on the bounding box, scaled by half its largest side
      double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
      for (int64_t v = 0; v < mm.nn; ++v)
        for (int d = 0; d < 3; ++d) {
          lo[d] = std::min(lo[d], mm.xyz[3 * v + d]);
          hi[d] = std::max(hi[d], mm.xyz[3 * v + d]);
        }
      double ext = 0.0;
      for (int d = 0; d < 3; ++d) {
        n->ctr[d] = 0.5 * (lo[d] + hi[d]);
        ext = std::max(ext, hi[d] - lo[d]);
      }
      n->half = ext > 0 ? 0.5 * ext : 1.0;
      n->coords_h.resize(3 * (size_t)mm.nn);
      for (int64_t v = 0; v < mm.nn; ++v)
        for (int d = 0; d < 3; ++d)
          n->coords_h[3 * v + d] = (float)((mm.xyz[3 * v + d] - n->ctr[d]) / n->half);
      n->max_maps = o->max_maps > 0 ? o->max_maps : 4096;
      const uint32_t mb = o->max_batch > 0 ? (uint32_t)o->max_batch : 65536u;
      if (mb % 64) throw Error(FIM_EINVAL, "max_batch must be a multiple of 64");
      n->cap = std::max<uint32_t>(tile_rows(mb), tile_rows(n->nn));
      // row spans of the weight gradients: 6 x 40 workgroups fill the CUs
      // once (32 / 48 / 64 spans measured 4-8 % slower per training step)
      n->nsplit = 40;
      // parameters: He-uniform hidden layers, small output layer, zero biases
      std::vector<float> w(NPAD, 0.0f);
      uint64_t st = o->seed ^ 0x5DEECE66Dull;
      auto uni = [&]() {
        st += 0x9E3779B97F4A7C15ull;
        uint64_t z = st;
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        z ^= z >> 31;
        return (double)(z >> 11) * (1.0 / 9007199254740992.0) * 2.0 - 1.0;
      };
      for (int l = 0; l < NL; ++l) {
        const int fan = l == 0 ? NIN : HID;
        const double a = std::sqrt(6.0 / fan);
        for (int r = 0; r < HID; ++r)
          for (int k = 0; k < fan; ++k) w[off_w(l) + (size_t)r * (l == 0 ? K0 : HID) + k] = (float)(a * uni());
      }
      for (int r = 0; r < HID; ++r) w[OFF_WO + r] = (float)(0.1 * std::sqrt(6.0 / HID) * uni());
      n->w = nn_alloc<float>(n, NPAD);
      n->g = nn_alloc<float>(n, NPAD);
      n->m = nn_alloc<float>(n, NPAD);
      n->v = nn_alloc<float>(n, NPAD);
      n->wb = nn_alloc<uint16_t>(n, NWB);
      n->wbt = nn_alloc<uint16_t>(n, NWBT);
      HIPCHECK(hipMemcpy(n->w, w.data(), NPAD * sizeof(float), hipMemcpyHostToDevice));
      HIPCHECK(hipMemset(n->m, 0, NPAD * sizeof(float)));
      HIPCHECK(hipMemset(n->v, 0, NPAD * sizeof(float)));
      HIPCHECK(hipMemset(n->g, 0, NPAD * sizeof(float)));
      k_to_bf16<<<grid_for(NPAD), 256, 0, n->stream>>>(n->w, n->wb, n->wbt);
      HIPCHECK(hipGetLastError());
      n->X0 = nn_alloc<uint16_t>(n, (size_t)n->cap * K0);
      HIPCHECK(hipMemset(n->X0, 0, (size_t)n->cap * K0 * sizeof(uint16_t)));  // padding columns stay 0
      for (int l = 0; l < NL - 1; ++l) n->Z[l] = nn_alloc<uint16_t>(n, (size_t)n->cap * HID);
      for (int l = 0; l < NL; ++l) n->dZ[l] = nn_alloc<uint16_t>(n, (size_t)n->cap * HID);
      n->tgt = nn_alloc<float>(n, n->cap);
      n->out = nn_alloc<float>(n, n->cap);
      n->parts = nn_alloc<float>(n, (size_t)n->nsplit * OFF_WO);
      HIPCHECK(hipFuncSetAttribute((const void *)k_mlp_chain<true>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                   (int)CHAIN_LDS));
      HIPCHECK(hipFuncSetAttribute((const void *)k_mlp_chain<false>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                   (int)CHAIN_LDS));
      HIPCHECK(hipFuncSetAttribute((const void *)k_mlp_wgrad, hipFuncAttributeMaxDynamicSharedMemorySize,
                                   WG_LDS * 2));
      n->loss = nn_alloc<float>(n, 1);
      n->coords = nn_alloc<float>(n, 3 * (size_t)n->nn);
      HIPCHECK(hipMemcpy(n->coords, n->coords_h.data(), 3 * (size_t)n->nn * sizeof(float), hipMemcpyHostToDevice));
      n->T = nn_alloc<float>(n, (size_t)n->max_maps * n->nn);
      n->mparams = nn_alloc<float>(n, (size_t)(n->max_maps + 1) * MAPP);
      HIPCHECK(hipStreamSynchronize(n->stream));
      *out = n;
    } catch (...) {
      nn_destroy(n);
      throw;
    }
  });
}

void fim_nn_destroy(fim_nn *n) { nn_destroy(n); }


int fim_nn_get_params(fim_nn *n, float *params) {
  return guarded([&] {
    if (!n || !params) throw Error(FIM_EINVAL, "null argument");
    HIPCHECK(hipSetDevice(n->dev));
    std::vector<float> dev(NPAD);
    HIPCHECK(hipMemcpyAsync(dev.data(), n->w, NPAD * sizeof(float), hipMemcpyDeviceToHost, n->stream));
    HIPCHECK(hipStreamSynchronize(n->stream));
    to_logical_order(dev, params);
  });
}

int fim_nn_set_params(fim_nn *n, const float *params) {
  return guarded([&] {
    if (!n || !params) throw Error(FIM_EINVAL, "null argument");
    HIPCHECK(hipSetDevice(n->dev));
    std::vector<float> dev;
    to_device_order(params, dev);
    HIPCHECK(hipMemcpyAsync(n->w, dev.data(), NPAD * sizeof(float), hipMemcpyHostToDevice, n->stream));
    k_to_bf16<<<grid_for(NPAD), 256, 0, n->stream>>>(n->w, n->wb, n->wbt);
    HIPCHECK(hipGetLastError());
    HIPCHECK(hipStreamSynchronize(n->stream));
  });
}

int fim_nn_adam_step(fim_nn *n, const float *grad) {
  return guarded([&] {
    if (!n || !grad) throw Error(FIM_EINVAL, "null argument");
    HIPCHECK(hipSetDevice(n->dev));
    std::vector<float> dev;
    to_device_order(grad, dev);  // padding entries stay 0
    HIPCHECK(hipMemcpyAsync(n->g, dev.data(), NPAD * sizeof(float), hipMemcpyHostToDevice, n->stream));
    adam(n);
    HIPCHECK(hipStreamSynchronize(n->stream));
  });
}

int fim_nn_get_moments(fim_nn *n, float *m, float *v, int64_t *steps) {
  return guarded([&] {
    if (!n) throw Error(FIM_EINVAL, "null argument");
    HIPCHECK(hipSetDevice(n->dev));
    std::vector<float> dev(NPAD);
    for (int k = 0; k < 2; ++k) {
      float *dst = k == 0 ? m : v;
      if (!dst) continue;
      HIPCHECK(hipMemcpyAsync(dev.data(), k == 0 ? n->m : n->v, NPAD * sizeof(float), hipMemcpyDeviceToHost,
                              n->stream));
      HIPCHECK(hipStreamSynchronize(n->stream));
      to_logical_order(dev, dst);
    }
    if (steps) *steps = n->t;
  });
}

int fim_nn_features(fim_nn *n, const int64_t *vtx, const int64_t *sites, int32_t n_sites,
                    const float *scales, int64_t count, float *x) {
  return guarded([&] {
    if (!n || (!vtx && count > 0) || !sites || !scales || (!x && count > 0) || count < 0)
      throw Error(FIM_EINVAL, "bad feature arguments");
    const std::vector<float> mp = map_params(n, sites, n_sites, scales);
    for (int64_t i = 0; i < count; ++i) {
      if (vtx[i] < 0 || vtx[i] >= (int64_t)n->nn) throw Error(FIM_EINVAL, "vertex out of range");
      for (int d = 0; d < 3; ++d) x[NIN * i + d] = n->coords_h[3 * vtx[i] + d];
      for (int k = 0; k < 13; ++k) x[NIN * i + 3 + k] = mp[k];
    }
  });
}

int fim_nn_add_maps(fim_nn *n, fim_ctx *c, int32_t inst0, int32_t count, const int64_t *sites,
                    const float *scales) {
  return guarded([&] {
    if (!n || !c || !sites || !scales || count < 0) throw Error(FIM_EINVAL, "null argument");
    if (!ctx_converge(c)) throw Error(FIM_ESTATE, "maps come from a converge-mode context");
    if (ctx_vertices(c) != n->nn) throw Error(FIM_EINVAL, "context mesh differs from the surrogate's");
    if (ctx_device(c) != n->dev) throw Error(FIM_EINVAL, "context on another device");
    if (n->nmaps + count > n->max_maps) throw Error(FIM_ELIMIT, "surrogate dataset full");
    HIPCHECK(hipSetDevice(n->dev));
    hipStream_t cs = (hipStream_t)ctx_stream(c);
    for (int32_t i = 0; i < count; ++i) {
      int32_t ns = 0;
      while (ns < 3 && sites[3 * i + ns] >= 0) ++ns;
      const std::vector<float> mp = map_params(n, sites + 3 * i, ns, scales + 4 * i);
      const float *src = ctx_gather(c, inst0 + i);
      HIPCHECK(hipMemcpyAsync(n->T + (size_t)n->nmaps * n->nn, src, (size_t)n->nn * sizeof(float),
                              hipMemcpyDeviceToDevice, cs));
      HIPCHECK(hipMemcpyAsync(n->mparams + (size_t)n->nmaps * MAPP, mp.data(), MAPP * sizeof(float),
                              hipMemcpyHostToDevice, cs));
      HIPCHECK(hipStreamSynchronize(cs));
      n->nmaps += 1;
    }
  });
}

int fim_nn_dataset_size(const fim_nn *n, int32_t *maps) {
  return guarded([&] {
    if (!n || !maps) throw Error(FIM_EINVAL, "null argument");
    *maps = n->nmaps;
  });
}

int fim_nn_gradient(fim_nn *n, const float *x, const float *t, int32_t b, double *loss, float *grad) {
  return guarded([&] {
    if (!n || !x || !t || !loss || !grad) throw Error(FIM_EINVAL, "null argument");
    if (b <= 0 || b % 64) throw Error(FIM_EINVAL, "batch must be a positive multiple of 64");
    ensure_rows(n, tile_rows((uint32_t)b));
    HIPCHECK(hipSetDevice(n->dev));
    std::vector<float> ts(b);
    for (int32_t i = 0; i < b; ++i) ts[i] = t[i] / t_scale(n);
    float *xd = nullptr;
    HIPCHECK(hipMalloc(&xd, (size_t)b * NIN * sizeof(float)));
    try {
      HIPCHECK(hipMemcpyAsync(xd, x, (size_t)b * NIN * sizeof(float), hipMemcpyHostToDevice, n->stream));
      HIPCHECK(hipMemcpyAsync(n->tgt, ts.data(), (size_t)b * sizeof(float), hipMemcpyHostToDevice, n->stream));
      k_input_rows<<<grid_for(tile_rows((uint32_t)b)), 256, 0, n->stream>>>(xd, (uint32_t)b,
                                                                            tile_rows((uint32_t)b), n->X0);
      train_pass(n, (uint32_t)b);
      std::vector<float> dev(NPAD);
      float l = 0.0f;
      HIPCHECK(hipMemcpyAsync(dev.data(), n->g, NPAD * sizeof(float), hipMemcpyDeviceToHost, n->stream));
      HIPCHECK(hipMemcpyAsync(&l, n->loss, sizeof(float), hipMemcpyDeviceToHost, n->stream));
      HIPCHECK(hipStreamSynchronize(n->stream));
      to_logical_order(dev, grad);
      *loss = (double)l / b;
    } catch (...) {
      (void)hipFree(xd);
      throw;
    }
    HIPCHECK(hipFree(xd));
  });
}

int fim_nn_train(fim_nn *n, int32_t steps, int32_t batch, uint64_t seed, double *loss) {
  return guarded([&] {
    if (!n) throw Error(FIM_EINVAL, "null argument");
    if (steps < 0 || batch <= 0 || batch % 64) throw Error(FIM_EINVAL, "bad training arguments");
    if (n->nmaps == 0) throw Error(FIM_ESTATE, "surrogate dataset is empty");
    const uint32_t rows = tile_rows((uint32_t)batch);
    ensure_rows(n, rows);
    HIPCHECK(hipSetDevice(n->dev));
    float l = 0.0f;
    for (int32_t s = 0; s < steps; ++s) {
      k_sample<<<grid_for(rows), 256, 0, n->stream>>>(n->coords, n->mparams, n->T, n->nn, (uint32_t)n->nmaps,
                                                      seed * 0x100000001B3ull + (uint64_t)s, (uint32_t)batch,
                                                      rows, 1.0f / t_scale(n), n->X0, n->tgt);
      train_pass(n, (uint32_t)batch);
      adam(n);
    }
    HIPCHECK(hipMemcpyAsync(&l, n->loss, sizeof(float), hipMemcpyDeviceToHost, n->stream));
    HIPCHECK(hipStreamSynchronize(n->stream));
    if (loss) *loss = (double)l / batch;
  });
}

int fim_nn_predict(fim_nn *n, const float *x, int64_t count, float *t) {
  return guarded([&] {
    if (!n || (!x && count > 0) || (!t && count > 0) || count < 0) throw Error(FIM_EINVAL, "null argument");
    if (count == 0) return;
    if (count > (int64_t)n->cap) throw Error(FIM_ELIMIT, "batch larger than the surrogate's row capacity");
    const uint32_t rows = tile_rows((uint32_t)count);
    ensure_rows(n, rows);
    HIPCHECK(hipSetDevice(n->dev));
    float *xd = nullptr;
    HIPCHECK(hipMalloc(&xd, (size_t)count * NIN * sizeof(float)));
    try {
      HIPCHECK(hipMemcpyAsync(xd, x, (size_t)count * NIN * sizeof(float), hipMemcpyHostToDevice, n->stream));
      k_input_rows<<<grid_for(rows), 256, 0, n->stream>>>(xd, (uint32_t)count, rows, n->X0);
      infer(n, (uint32_t)count, 0.0f, n->out);
      HIPCHECK(hipMemcpyAsync(t, n->out, (size_t)count * sizeof(float), hipMemcpyDeviceToHost, n->stream));
      HIPCHECK(hipStreamSynchronize(n->stream));
    } catch (...) {
      (void)hipFree(xd);
      throw;
    }
    HIPCHECK(hipFree(xd));
  });
}

int fim_nn_predict_max(fim_nn *n, int32_t count, const int64_t *sites, const float *scales,
                       const int64_t *sample, int32_t m, float *out) {
  return guarded([&] {
    if (!n || count < 0 || (count > 0 && (!sites || !scales || !out)) || m <= 0 || !sample)
      throw Error(FIM_EINVAL, "bad arguments");
    if (count == 0) return;
    if ((uint32_t)m > n->cap) throw Error(FIM_ELIMIT, "vertex sample larger than the row capacity");
    std::vector<uint32_t> sv(m);
    for (int32_t i = 0; i < m; ++i) {
      if (sample[i] < 0 || sample[i] >= (int64_t)n->nn) throw Error(FIM_EINVAL, "sample vertex out of range");
      sv[i] = (uint32_t)sample[i];
    }
    std::vector<float> mp((size_t)count * MAPP);
    for (int32_t j = 0; j < count; ++j) {
      int32_t ns = 0;
      while (ns < 3 && sites[3 * j + ns] >= 0) ++ns;
      const std::vector<float> p = map_params(n, sites + 3 * j, ns, scales + 4 * j);
      std::copy(p.begin(), p.end(), mp.begin() + (size_t)j * MAPP);
    }
    HIPCHECK(hipSetDevice(n->dev));
    // solves per forward pass: as many whole samples as the rows hold
    const uint32_t per = std::max<uint32_t>(1u, n->cap / (uint32_t)m);
    uint32_t *sd = nullptr;
    float *md = nullptr, *od = nullptr;
    try {
      HIPCHECK(hipMalloc(&sd, sv.size() * sizeof(uint32_t)));
      HIPCHECK(hipMalloc(&md, mp.size() * sizeof(float)));
      HIPCHECK(hipMalloc(&od, (size_t)count * sizeof(float)));
      HIPCHECK(hipMemcpyAsync(sd, sv.data(), sv.size() * sizeof(uint32_t), hipMemcpyHostToDevice, n->stream));
      HIPCHECK(hipMemcpyAsync(md, mp.data(), mp.size() * sizeof(float), hipMemcpyHostToDevice, n->stream));
      for (uint32_t j0 = 0; j0 < (uint32_t)count; j0 += per) {
        const uint32_t nj = std::min<uint32_t>(per, (uint32_t)count - j0);
        const uint32_t nreal = nj * (uint32_t)m;
        const uint32_t rows = tile_rows(nreal);  // <= cap: cap is a whole number of tiles
        k_sample_rows<<<grid_for(rows), 256, 0, n->stream>>>(n->coords, md + (size_t)j0 * MAPP, sd, (uint32_t)m,
                                                             nreal, rows, n->X0);
        infer(n, nreal, 0.0f, n->out);
        k_solve_max<<<std::min<uint32_t>((nj + 3) / 4, 4096), 256, 0, n->stream>>>(n->out, (uint32_t)m, nj,
                                                                                  od + j0);
        HIPCHECK(hipGetLastError());
      }
      HIPCHECK(hipMemcpyAsync(out, od, (size_t)count * sizeof(float), hipMemcpyDeviceToHost, n->stream));
      HIPCHECK(hipStreamSynchronize(n->stream));
    } catch (...) {
      (void)hipFree(sd);
      (void)hipFree(md);
      (void)hipFree(od);
      throw;
    }
    HIPCHECK(hipFree(sd));
    HIPCHECK(hipFree(md));
    HIPCHECK(hipFree(od));
  });
}

int fim_nn_predict_map(fim_nn *n, const int64_t *sites, int32_t n_sites, const float *scales, float *t) {
  return guarded([&] {
    if (!n || !sites || !scales || !t) throw Error(FIM_EINVAL, "null argument");
    HIPCHECK(hipSetDevice(n->dev));
    const std::vector<float> mp = map_params(n, sites, n_sites, scales);
    predict_map_device(n, mp, 0.0f, n->out);
    HIPCHECK(hipMemcpyAsync(t, n->out, (size_t)n->nn * sizeof(float), hipMemcpyDeviceToHost, n->stream));
    HIPCHECK(hipStreamSynchronize(n->stream));
  });
}

int fim_nn_warm_start(fim_nn *n, fim_ctx *c, int32_t inst, const int64_t *sites, int32_t n_sites,
                      const float *scales, double delta, int32_t all_active) {
  return guarded([&] {
    if (!n || !c || !sites || !scales) throw Error(FIM_EINVAL, "null argument");
    if (!ctx_converge(c)) throw Error(FIM_ESTATE, "warm start needs a converge-mode context");
    if (ctx_vertices(c) != n->nn) throw Error(FIM_EINVAL, "context mesh differs from the surrogate's");
    if (ctx_device(c) != n->dev) throw Error(FIM_EINVAL, "context on another device");
    if (inst < 0 || inst >= ctx_instances(c)) throw Error(FIM_EINVAL, "instance index out of range");
    if (!(delta >= 0.0)) throw Error(FIM_EINVAL, "delta must be >= 0");
    HIPCHECK(hipSetDevice(n->dev));
    const std::vector<float> mp = map_params(n, sites, n_sites, scales);
    predict_map_device(n, mp, (float)delta, n->out);
    uint32_t sv[3];
    for (int k = 0; k < n_sites; ++k) sv[k] = (uint32_t)sites[k];
    uint32_t *svd = reinterpret_cast<uint32_t *>(n->tgt);  // scratch of >= 3 words
    HIPCHECK(hipMemcpyAsync(svd, sv, n_sites * sizeof(uint32_t), hipMemcpyHostToDevice, n->stream));
    k_zero_sites<<<1, 64, 0, n->stream>>>(svd, (uint32_t)n_sites, n->out);
    // hand the field to the solver's stream
    float *dst = ctx_field_buffer(c);
    HIPCHECK(hipMemcpyAsync(dst, n->out, (size_t)n->nn * sizeof(float), hipMemcpyDeviceToDevice, n->stream));
    HIPCHECK(hipStreamSynchronize(n->stream));
    ctx_init_field(c, inst, all_active != 0);
    HIPCHECK(hipStreamSynchronize((hipStream_t)ctx_stream(c)));
  });
}

}  // extern "C"
