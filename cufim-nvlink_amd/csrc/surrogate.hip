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
// Activations of a batch of B rows (B a multiple of 64) are row-major BF16:
//   X0[B x 64], Z_l and H_l = SiLU(Z_l) [B x 256] for l = 0..5.
// Backward: dZ_5 from the head, then per layer dW_l = dZ_l^T H_{l-1} (split
// over the batch into partial sums, reduced), db_l = column sums of dZ_l,
// dZ_{l-1} = (dZ_l W_l) * SiLU'(Z_{l-1}).
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
// BF16 weight copies
constexpr size_t offb_w(int l) { return l == 0 ? 0 : W0_N + (size_t)(l - 1) * WL_N; }
constexpr size_t NWB = W0_N + (size_t)(NL - 1) * WL_N;
// BF16 copies of the hidden layers transposed (WT_l[k][n] = W_l[n][k], l >= 1):
// the backward pass reads them contiguous in its K dimension
constexpr size_t offbt(int l) { return (size_t)(l - 1) * WL_N; }
constexpr size_t NWBT = (size_t)(NL - 1) * WL_N;
static_assert(off_w(NL) == off_b(NL - 1) + HID, "layout");
static_assert((size_t)HID * NIN + HID + (NL - 1) * (WL_N + HID) + HID + 1 == FIM_NN_PARAMS,
              "logical parameter count");

typedef unsigned short u16x8 __attribute__((ext_vector_type(8)));
typedef unsigned short u16x4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

// float -> bf16, round to nearest even (finite inputs)
__host__ __device__ inline uint16_t f2bf(float f) {
  uint32_t u;
  std::memcpy(&u, &f, 4);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}
__host__ __device__ inline float bf2f(uint16_t h) {
  const uint32_t u = (uint32_t)h << 16;
  float f;
  std::memcpy(&f, &u, 4);
  return f;
}
__device__ inline float sigm(float z) { return 1.0f / (1.0f + __expf(-z)); }
__device__ inline float silu(float z) { return z * sigm(z); }
__device__ inline float dsilu(float z) {
  const float s = sigm(z);
  return s * (1.0f + z * (1.0f - s));
}

// ---------------------------------------------------------------------------
// C[R x Cc] = A(R x K) B(K x Cc) over the K range of blockIdx.z, BF16 in,
// FP32 accumulate.  A(r, k) = A[r*sar + k*sak], B(k, c) = B[k*sbk + c*sbc];
// AK: A contiguous in k (sak == 1), else in r (sar == 1); BK likewise for B.
// Workgroup tile 64 x 64, four waves of 32 x 32 (one MFMA accumulator),
// K staged through LDS 32 at a time.  R, Cc multiples of 64; the K span a
// multiple of 32.  MFMA 32x32x16 operands: lane l holds row (col) l % 32,
// k = 8 * (l / 32) + 0..7 of a 16-deep step; the accumulator element r of
// lane l is row 8 * (r / 4) + 4 * (l / 32) + r % 4, col l % 32.
// ---------------------------------------------------------------------------
template <bool AK, bool BK, class Epi>
__global__ __launch_bounds__(256) void k_gemm(const uint16_t *__restrict__ A, size_t sar, size_t sak,
                                              const uint16_t *__restrict__ B, size_t sbk, size_t sbc,
                                              uint32_t kspan, Epi epi) {
  __shared__ __align__(16) uint16_t As[64][40];  // [row][k], rows padded to 80 B
  __shared__ __align__(16) uint16_t Bs[64][40];  // [col][k]
  const uint32_t tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const uint32_t wr = w & 1u, wc = w >> 1;
  const size_t r0 = (size_t)blockIdx.y * 64, c0 = (size_t)blockIdx.x * 64;
  const size_t kb = (size_t)blockIdx.z * kspan, ke = kb + kspan;
  f32x16 acc;
  for (int i = 0; i < 16; ++i) acc[i] = 0.0f;
  for (size_t k0 = kb; k0 < ke; k0 += 32) {
    if (AK) {
      const uint32_t r = tid >> 2, kq = (tid & 3u) * 8u;
      *reinterpret_cast<u16x8 *>(&As[r][kq]) =
          *reinterpret_cast<const u16x8 *>(A + (r0 + r) * sar + k0 + kq);
    } else {
      const uint32_t k = tid >> 3, rq = (tid & 7u) * 8u;
      const u16x8 v = *reinterpret_cast<const u16x8 *>(A + (k0 + k) * sak + r0 + rq);
      for (int i = 0; i < 8; ++i) As[rq + i][k] = v[i];
    }
    if (BK) {
      const uint32_t c = tid >> 2, kq = (tid & 3u) * 8u;
      *reinterpret_cast<u16x8 *>(&Bs[c][kq]) =
          *reinterpret_cast<const u16x8 *>(B + (c0 + c) * sbc + k0 + kq);
    } else {
      const uint32_t k = tid >> 3, cq = (tid & 7u) * 8u;
      const u16x8 v = *reinterpret_cast<const u16x8 *>(B + (k0 + k) * sbk + c0 + cq);
      for (int i = 0; i < 8; ++i) Bs[cq + i][k] = v[i];
    }
    __syncthreads();
#pragma unroll
    for (uint32_t kk = 0; kk < 32; kk += 16) {
      const u16x8 a = *reinterpret_cast<const u16x8 *>(&As[wr * 32 + (lane & 31u)][kk + (lane >> 5) * 8]);
      const u16x8 b = *reinterpret_cast<const u16x8 *>(&Bs[wc * 32 + (lane & 31u)][kk + (lane >> 5) * 8]);
      acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(bf16x8, a),
                                                   __builtin_bit_cast(bf16x8, b), acc, 0, 0, 0);
    }
    __syncthreads();
  }
#pragma unroll
  for (int r = 0; r < 16; ++r) {
    const size_t row = r0 + wr * 32 + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
    const size_t col = c0 + wc * 32 + (lane & 31u);
    epi(row, col, acc[r]);
  }
}

// EpiLoads<E>::value: the epilogue E reads an input per output element
template <class E, class = void>
struct EpiLoads {
  static constexpr bool value = false;
};
template <class E>
struct EpiLoads<E, decltype(void(E::kLoads))> {
  static constexpr bool value = E::kLoads;
};

// ---------------------------------------------------------------------------
// The same product on 128 x 128 workgroup tiles: four waves of 64 x 64, each
// wave 2 x 2 MFMA accumulators (32 x 32 each), so every A and B fragment read
// from LDS feeds two MFMAs.  K advances 32 at a time through two LDS buffers:
// the next step's global loads are issued into registers before the current
// step's MFMAs and stored into the other buffer after them, one barrier per
// step.  R, Cc multiples of 128; the K span a multiple of 32.
// ---------------------------------------------------------------------------
template <bool AK, bool BK, class Epi>
__global__ __launch_bounds__(256) void k_gemm128(const uint16_t *__restrict__ A, size_t sar, size_t sak,
                                                 const uint16_t *__restrict__ B, size_t sbk, size_t sbc,
                                                 uint32_t kspan, Epi epi) {
  constexpr int LD = 40;  // LDS row pitch in BF16 values (80 B)
  __shared__ __align__(16) uint16_t As[2][128][LD];  // [buffer][row][k]
  __shared__ __align__(16) uint16_t Bs[2][128][LD];  // [buffer][col][k]
  const uint32_t tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const uint32_t wr = w & 1u, wc = w >> 1;
  const size_t r0 = (size_t)blockIdx.y * 128, c0 = (size_t)blockIdx.x * 128;
  const size_t kb = (size_t)blockIdx.z * kspan;
  const uint32_t nk = kspan / 32;
  u16x8 ra[2], rb[2];
  auto gload = [&](size_t k0) {
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      if (AK) {
        const uint32_t r = (tid >> 2) + 64u * j, kq = (tid & 3u) * 8u;
        ra[j] = *reinterpret_cast<const u16x8 *>(A + (r0 + r) * sar + k0 + kq);
      } else {  // rows k0 + 2 (tid >> 4) + j: two neighbouring k per thread
        const uint32_t k = 2u * (tid >> 4) + j, rq = (tid & 15u) * 8u;
        ra[j] = *reinterpret_cast<const u16x8 *>(A + (k0 + k) * sak + r0 + rq);
      }
      if (BK) {
        const uint32_t c = (tid >> 2) + 64u * j, kq = (tid & 3u) * 8u;
        rb[j] = *reinterpret_cast<const u16x8 *>(B + (c0 + c) * sbc + k0 + kq);
      } else {
        const uint32_t k = 2u * (tid >> 4) + j, cq = (tid & 15u) * 8u;
        rb[j] = *reinterpret_cast<const u16x8 *>(B + (k0 + k) * sbk + c0 + cq);
      }
    }
  };
  // transposed staging writes the two neighbouring k of a row as one
  // 32-bit word (k is even, rows are 80 B apart, so the word is aligned)
  auto lstore = [&](int buf) {
    if (AK) {
#pragma unroll
      for (int j = 0; j < 2; ++j) {
        const uint32_t r = (tid >> 2) + 64u * j, kq = (tid & 3u) * 8u;
        *reinterpret_cast<u16x8 *>(&As[buf][r][kq]) = ra[j];
      }
    } else {
      const uint32_t k = 2u * (tid >> 4), rq = (tid & 15u) * 8u;
#pragma unroll
      for (int i = 0; i < 8; ++i)
        *reinterpret_cast<uint32_t *>(&As[buf][rq + i][k]) = (uint32_t)ra[0][i] | ((uint32_t)ra[1][i] << 16);
    }
    if (BK) {
#pragma unroll
      for (int j = 0; j < 2; ++j) {
        const uint32_t c = (tid >> 2) + 64u * j, kq = (tid & 3u) * 8u;
        *reinterpret_cast<u16x8 *>(&Bs[buf][c][kq]) = rb[j];
      }
    } else {
      const uint32_t k = 2u * (tid >> 4), cq = (tid & 15u) * 8u;
#pragma unroll
      for (int i = 0; i < 8; ++i)
        *reinterpret_cast<uint32_t *>(&Bs[buf][cq + i][k]) = (uint32_t)rb[0][i] | ((uint32_t)rb[1][i] << 16);
    }
  };
  f32x16 acc[2][2];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int r = 0; r < 16; ++r) acc[i][j][r] = 0.0f;
  gload(kb);
  lstore(0);
  __syncthreads();
  for (uint32_t st = 0; st < nk; ++st) {
    const int buf = st & 1;
    if (st + 1 < nk) gload(kb + 32 * (size_t)(st + 1));
#pragma unroll
    for (uint32_t kk = 0; kk < 32; kk += 16) {
      u16x8 a[2], b[2];
#pragma unroll
      for (int i = 0; i < 2; ++i) {
        a[i] = *reinterpret_cast<const u16x8 *>(&As[buf][wr * 64 + i * 32 + (lane & 31u)][kk + (lane >> 5) * 8]);
        b[i] = *reinterpret_cast<const u16x8 *>(&Bs[buf][wc * 64 + i * 32 + (lane & 31u)][kk + (lane >> 5) * 8]);
      }
#pragma unroll
      for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
          acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(bf16x8, a[i]),
                                                             __builtin_bit_cast(bf16x8, b[j]), acc[i][j], 0, 0, 0);
    }
    if (st + 1 < nk) lstore(buf ^ 1);
    __syncthreads();
  }
  if constexpr (EpiLoads<Epi>::value) {
    // epilogues that read an input per element: all loads of one 32 x 32
    // block first, then the stores
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
      for (int j = 0; j < 2; ++j) {
        uint16_t in[16];
#pragma unroll
        for (int r = 0; r < 16; ++r)
          in[r] = epi.load(r0 + wr * 64 + i * 32 + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3),
                           c0 + wc * 64 + j * 32 + (lane & 31u));
#pragma unroll
        for (int r = 0; r < 16; ++r)
          epi.store(r0 + wr * 64 + i * 32 + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3),
                    c0 + wc * 64 + j * 32 + (lane & 31u), acc[i][j][r], in[r]);
      }
  } else {
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
      for (int j = 0; j < 2; ++j)
#pragma unroll
        for (int r = 0; r < 16; ++r) {
          const size_t row = r0 + wr * 64 + i * 32 + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
          const size_t col = c0 + wc * 64 + j * 32 + (lane & 31u);
          epi(row, col, acc[i][j][r]);
        }
  }
}

// hidden layer forward: Z = acc + b (BF16, kept for the backward pass when Z
// is given), H = SiLU(Z)
struct EpiFwd {
  const float *bias;
  uint16_t *Z, *H;
  __device__ void operator()(size_t row, size_t col, float v) const {
    const float z = v + bias[col];
    if (Z) Z[row * HID + col] = f2bf(z);
    H[row * HID + col] = f2bf(silu(z));
  }
};

// backward through a hidden layer: dZ_prev = (dZ W) * SiLU'(Z_prev).  The
// tile epilogue loads all its Z values (load) before it writes any dZ
// (store), so the loads are in flight together.
struct EpiBwd {
  const uint16_t *Zp;
  uint16_t *dZ;
  static constexpr bool kLoads = true;
  __device__ uint16_t load(size_t row, size_t col) const { return Zp[row * HID + col]; }
  __device__ void store(size_t row, size_t col, float v, uint16_t z) const {
    dZ[row * HID + col] = f2bf(v * dsilu(bf2f(z)));
  }
  __device__ void operator()(size_t row, size_t col, float v) const {
    store(row, col, v, load(row, col));
  }
};

// weight gradient: partial sum of one batch split
struct EpiPart {
  float *P;
  uint32_t ld;
  __device__ void operator()(size_t row, size_t col, float v) const {
    P[(size_t)blockIdx.z * HID * ld + row * ld + col] = v;
  }
};

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

__device__ inline void write_features(uint16_t *x, const float *vxyz, const float *mp) {
  for (int k = 0; k < 3; ++k) x[k] = f2bf(vxyz[k]);
  for (int k = 0; k < 13; ++k) x[3 + k] = f2bf(mp[k]);
  for (int k = NIN; k < K0; ++k) x[k] = 0;
}

// training rows: random (map, vertex) samples
__global__ void k_sample(const float *coords, const float *mparams, const float *T, uint32_t nn,
                         uint32_t nmaps, uint64_t seed, uint32_t B, float inv_ts, uint16_t *X0,
                         float *tgt) {
  for (uint32_t m = blockIdx.x * blockDim.x + threadIdx.x; m < B; m += gridDim.x * blockDim.x) {
    const uint64_t h = mix64(seed ^ mix64((uint64_t)m + 1));
    const uint32_t j = (uint32_t)((h >> 32) % nmaps), v = (uint32_t)((h & 0xffffffffu) % nn);
    write_features(X0 + (size_t)m * K0, coords + 3 * (size_t)v, mparams + MAPP * (size_t)j);
    tgt[m] = T[(size_t)j * nn + v] * inv_ts;
  }
}

// rows of one map: every vertex (padding rows repeat vertex 0)
__global__ void k_map_rows(const float *coords, const float *mp, uint32_t nn, uint32_t rows,
                           uint16_t *X0) {
  for (uint32_t m = blockIdx.x * blockDim.x + threadIdx.x; m < rows; m += gridDim.x * blockDim.x)
    write_features(X0 + (size_t)m * K0, coords + 3 * (size_t)(m < nn ? m : 0), mp);
}

// rows of several solves over one vertex sample: row r is solve r / m at
// sample vertex r % m (padding rows repeat the last real row)
__global__ void k_sample_rows(const float *coords, const float *mp, const uint32_t *sample, uint32_t m,
                              uint32_t nreal, uint32_t rows, uint16_t *X0) {
  for (uint32_t r = blockIdx.x * blockDim.x + threadIdx.x; r < rows; r += gridDim.x * blockDim.x) {
    const uint32_t q = r < nreal ? r : nreal - 1;
    const uint32_t j = q / m, v = sample[q - j * m];
    write_features(X0 + (size_t)r * K0, coords + 3 * (size_t)v, mp + MAPP * (size_t)j);
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

// given inputs (n x 16 floats) -> padded BF16 rows
__global__ void k_input_rows(const float *x, uint32_t n, uint32_t rows, uint16_t *X0) {
  for (uint32_t m = blockIdx.x * blockDim.x + threadIdx.x; m < rows; m += gridDim.x * blockDim.x)
    for (int k = 0; k < K0; ++k)
      X0[(size_t)m * K0 + k] = (m < n && k < NIN) ? f2bf(x[(size_t)m * NIN + k]) : (uint16_t)0;
}

// output layer and squared error, one wave per row: y = H5 . w_out + b_out;
// e = y - t; dZ5 = (2 e / B) w_out * SiLU'(Z5); gradients of w_out, b_out
__global__ __launch_bounds__(256) void k_head_train(const uint16_t *H5, const uint16_t *Z5,
                                                    const float *wo, const float *bo,
                                                    const float *tgt, uint32_t B, float inv_b,
                                                    uint16_t *dZ5, float *g_wo, float *g_bo,
                                                    float *loss) {
  __shared__ float s_gwo[HID];
  __shared__ float s_misc[2];
  const uint32_t tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  s_gwo[tid] = 0.0f;
  if (tid < 2) s_misc[tid] = 0.0f;
  __syncthreads();
  const uint32_t c = lane * 4;
  const float w4[4] = {wo[c], wo[c + 1], wo[c + 2], wo[c + 3]};
  float gw[4] = {0.0f, 0.0f, 0.0f, 0.0f}, gb = 0.0f, ls = 0.0f;
  for (uint32_t row = blockIdx.x * 4 + wave; row < B; row += gridDim.x * 4) {
    const u16x4 hv = *reinterpret_cast<const u16x4 *>(H5 + (size_t)row * HID + c);
    const float h[4] = {bf2f(hv[0]), bf2f(hv[1]), bf2f(hv[2]), bf2f(hv[3])};
    float p = h[0] * w4[0] + h[1] * w4[1] + h[2] * w4[2] + h[3] * w4[3];
    for (int off = 32; off > 0; off >>= 1) p += __shfl_xor(p, off);
    const float e = p + bo[0] - tgt[row];
    const float dy = 2.0f * e * inv_b;
    if (lane == 0) {
      ls += e * e;
      gb += dy;
    }
    const u16x4 zv = *reinterpret_cast<const u16x4 *>(Z5 + (size_t)row * HID + c);
    u16x4 o;
    for (int j = 0; j < 4; ++j) {
      gw[j] += dy * h[j];
      o[j] = f2bf(dy * w4[j] * dsilu(bf2f(zv[j])));
    }
    *reinterpret_cast<u16x4 *>(dZ5 + (size_t)row * HID + c) = o;
  }
  for (int j = 0; j < 4; ++j) atomicAdd(&s_gwo[c + j], gw[j]);
  if (lane == 0) {
    atomicAdd(&s_misc[0], gb);
    atomicAdd(&s_misc[1], ls);
  }
  __syncthreads();
  atomicAdd(&g_wo[tid], s_gwo[tid]);
  if (tid == 0) {
    atomicAdd(g_bo, s_misc[0]);
    atomicAdd(loss, s_misc[1]);
  }
}

// output layer for prediction: t = (H5 . w_out + b_out) * t_scale + delta
__global__ __launch_bounds__(256) void k_head_infer(const uint16_t *H5, const float *wo, const float *bo,
                                                    uint32_t n, float t_scale, float delta, float *out) {
  const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const uint32_t c = lane * 4;
  const float w4[4] = {wo[c], wo[c + 1], wo[c + 2], wo[c + 3]};
  for (uint32_t row = blockIdx.x * 4 + wave; row < n; row += gridDim.x * 4) {
    const u16x4 hv = *reinterpret_cast<const u16x4 *>(H5 + (size_t)row * HID + c);
    float p = bf2f(hv[0]) * w4[0] + bf2f(hv[1]) * w4[1] + bf2f(hv[2]) * w4[2] + bf2f(hv[3]) * w4[3];
    for (int off = 32; off > 0; off >>= 1) p += __shfl_xor(p, off);
    if (lane == 0) out[row] = (p + bo[0]) * t_scale + delta;
  }
}

// column sums of dZ (B x 256) into g_b (bias gradient)
// column sums of dZ (B x 256) into g_b: a thread sums 8 columns of every
// eighth row of the block's rows (16-byte loads), the block reduces in LDS
__global__ __launch_bounds__(256) void k_colsum(const uint16_t *dZ, uint32_t B, float *g_b) {
  __shared__ float part[8][HID];
  const uint32_t tid = threadIdx.x, cg = tid & 31u, rl = tid >> 5;  // column group, row lane
  float s[8] = {0, 0, 0, 0, 0, 0, 0, 0};
#pragma unroll 4
  for (uint32_t row = blockIdx.x * 8u + rl; row < B; row += gridDim.x * 8u) {
    const u16x8 v = *reinterpret_cast<const u16x8 *>(dZ + (size_t)row * HID + cg * 8u);
#pragma unroll
    for (int i = 0; i < 8; ++i) s[i] += bf2f(v[i]);
  }
#pragma unroll
  for (int i = 0; i < 8; ++i) part[rl][cg * 8u + i] = s[i];
  __syncthreads();
  float t = 0.0f;
#pragma unroll
  for (int r = 0; r < 8; ++r) t += part[r][tid];
  atomicAdd(&g_b[tid], t);
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
        a.wb[offb_w(l) + q] = f2bf(w);
        if (l > 0) a.wbt[offbt(l) + (q % HID) * HID + q / HID] = f2bf(w);
      }
  }
}

__global__ void k_to_bf16(const float *w, uint16_t *wb, uint16_t *wbt) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < NPAD; i += (size_t)gridDim.x * blockDim.x)
    for (int l = 0; l < NL; ++l)
      if (i >= off_w(l) && i < off_b(l)) {
        const size_t q = i - off_w(l);
        wb[offb_w(l) + q] = f2bf(w[i]);
        if (l > 0) wbt[offbt(l) + (q % HID) * HID + q / HID] = f2bf(w[i]);
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
  // activations (rows = cap)
  uint32_t cap = 0;
  uint16_t *X0 = nullptr, *Z[NL] = {}, *H[NL] = {}, *dZa = nullptr, *dZb = nullptr;
  float *tgt = nullptr, *out = nullptr, *parts = nullptr, *loss = nullptr;
  uint32_t nsplit = 1;
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

// forward pass over `rows` rows of X0 (store_z keeps Z for backward)
void forward(fim_nn *n, uint32_t rows, bool store_z) {
  const bool big = rows % 128 == 0;
  const dim3 g = big ? dim3(HID / 128, rows / 128, 1) : dim3(HID / 64, rows / 64, 1);
  for (int l = 0; l < NL; ++l) {
    const uint16_t *in = l == 0 ? n->X0 : n->H[l - 1];
    const size_t K = l == 0 ? K0 : HID;
    EpiFwd e{n->w + off_b(l), store_z ? n->Z[l] : nullptr, n->H[l]};
    if (big)
      k_gemm128<true, true, EpiFwd><<<g, 256, 0, n->stream>>>(in, K, 1, n->wb + offb_w(l), 1, K, (uint32_t)K, e);
    else
      k_gemm<true, true, EpiFwd><<<g, 256, 0, n->stream>>>(in, K, 1, n->wb + offb_w(l), 1, K, (uint32_t)K, e);
  }
  HIPCHECK(hipGetLastError());
}

// backward pass for `rows` rows (targets in n->tgt); gradients in n->g
void backward(fim_nn *n, uint32_t rows) {
  HIPCHECK(hipMemsetAsync(n->g, 0, NPAD * sizeof(float), n->stream));
  HIPCHECK(hipMemsetAsync(n->loss, 0, sizeof(float), n->stream));
  k_head_train<<<std::min<uint32_t>(rows / 4, 1024), 256, 0, n->stream>>>(
      n->H[NL - 1], n->Z[NL - 1], n->w + OFF_WO, n->w + OFF_BO, n->tgt, rows, 1.0f / rows, n->dZa,
      n->g + OFF_WO, n->g + OFF_BO, n->loss);
  uint16_t *dz = n->dZa, *dzp = n->dZb;
  // batch splits of the weight gradient: spans of >= 32 rows
  uint32_t S = std::min<uint32_t>(n->nsplit, rows / 32);
  while (rows % (S * 32)) --S;
  const uint32_t span = rows / S;
  for (int l = NL - 1; l >= 0; --l) {
    const uint16_t *hin = l == 0 ? n->X0 : n->H[l - 1];
    const uint32_t K = l == 0 ? K0 : HID;
    // dW_l[n][k] = sum_m dZ[m][n] hin[m][k]: A(r=n, k=m) = dZ (contiguous in r),
    // B(k=m, c=k) = hin (contiguous in c)
    EpiPart ep{n->parts, K};
    if (K % 128 == 0)
      k_gemm128<false, false, EpiPart><<<dim3(K / 128, HID / 128, S), 256, 0, n->stream>>>(dz, 1, HID, hin, K, 1,
                                                                                           span, ep);
    else
      k_gemm<false, false, EpiPart><<<dim3(K / 64, HID / 64, S), 256, 0, n->stream>>>(dz, 1, HID, hin, K, 1, span,
                                                                                       ep);
    k_reduce_parts<<<grid_for((size_t)HID * K), 256, 0, n->stream>>>(n->parts, S, (size_t)HID * K,
                                                                     n->g + off_w(l));
    k_colsum<<<std::min<uint32_t>(rows / 8, 512), 256, 0, n->stream>>>(dz, rows, n->g + off_b(l));
    if (l > 0) {
      // dZ_{l-1}[m][k] = sum_n dZ[m][n] W_l[n][k] * SiLU'(Z_{l-1})
      EpiBwd eb{n->Z[l - 1], dzp};
      // B(k = n, c = k') = W_l[n][k'] = WT_l[k'][n]: contiguous in k
      if (rows % 128 == 0)
        k_gemm128<true, true, EpiBwd><<<dim3(HID / 128, rows / 128, 1), 256, 0, n->stream>>>(
            dz, HID, 1, n->wbt + offbt(l), 1, HID, HID, eb);
      else
        k_gemm<true, true, EpiBwd><<<dim3(HID / 64, rows / 64, 1), 256, 0, n->stream>>>(
            dz, HID, 1, n->wbt + offbt(l), 1, HID, HID, eb);
      std::swap(dz, dzp);
    }
  }
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

float t_scale(const fim_nn *n) { return (float)(n->o.t_scale > 0 ? n->o.t_scale : 1e5); }

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
  const uint32_t rows = (n->nn + 127u) & ~127u;  // 128-row tiles
  ensure_rows(n, rows);
  HIPCHECK(hipMemcpyAsync(n->mparams + (size_t)n->max_maps * MAPP, mp.data(), MAPP * sizeof(float),
                          hipMemcpyHostToDevice, n->stream));
  k_map_rows<<<grid_for(rows), 256, 0, n->stream>>>(n->coords, n->mparams + (size_t)n->max_maps * MAPP,
                                                    n->nn, rows, n->X0);
  forward(n, rows, false);
  k_head_infer<<<std::min<uint32_t>(rows / 4, 4096), 256, 0, n->stream>>>(n->H[NL - 1], n->w + OFF_WO,
                                                                          n->w + OFF_BO, n->nn,
                                                                          t_scale(n), delta, dst);
  HIPCHECK(hipGetLastError());
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
      n->cap = std::max<uint32_t>(mb, (n->nn + 127u) & ~127u);
      n->nsplit = 128;  // batch splits of the weight gradients: 4 x 128 tiles of 128 x 128
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
      for (int l = 0; l < NL; ++l) {
        n->Z[l] = nn_alloc<uint16_t>(n, (size_t)n->cap * HID);
        n->H[l] = nn_alloc<uint16_t>(n, (size_t)n->cap * HID);
      }
      n->dZa = nn_alloc<uint16_t>(n, (size_t)n->cap * HID);
      n->dZb = nn_alloc<uint16_t>(n, (size_t)n->cap * HID);
      n->tgt = nn_alloc<float>(n, n->cap);
      n->out = nn_alloc<float>(n, n->cap);
      n->parts = nn_alloc<float>(n, (size_t)n->nsplit * HID * HID);
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
    ensure_rows(n, (uint32_t)b);
    HIPCHECK(hipSetDevice(n->dev));
    std::vector<float> ts(b);
    for (int32_t i = 0; i < b; ++i) ts[i] = t[i] / t_scale(n);
    float *xd = nullptr;
    HIPCHECK(hipMalloc(&xd, (size_t)b * NIN * sizeof(float)));
    try {
      HIPCHECK(hipMemcpyAsync(xd, x, (size_t)b * NIN * sizeof(float), hipMemcpyHostToDevice, n->stream));
      HIPCHECK(hipMemcpyAsync(n->tgt, ts.data(), (size_t)b * sizeof(float), hipMemcpyHostToDevice, n->stream));
      k_input_rows<<<grid_for((size_t)b), 256, 0, n->stream>>>(xd, (uint32_t)b, (uint32_t)b, n->X0);
      forward(n, (uint32_t)b, true);
      backward(n, (uint32_t)b);
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
    ensure_rows(n, (uint32_t)batch);
    HIPCHECK(hipSetDevice(n->dev));
    float l = 0.0f;
    for (int32_t s = 0; s < steps; ++s) {
      k_sample<<<grid_for((size_t)batch), 256, 0, n->stream>>>(
          n->coords, n->mparams, n->T, n->nn, (uint32_t)n->nmaps, seed * 0x100000001B3ull + (uint64_t)s,
          (uint32_t)batch, 1.0f / t_scale(n), n->X0, n->tgt);
      forward(n, (uint32_t)batch, true);
      backward(n, (uint32_t)batch);
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
    const uint32_t r128 = (uint32_t)((count + 127) & ~int64_t(127));
    const uint32_t rows = r128 <= n->cap ? r128 : (uint32_t)((count + 63) & ~int64_t(63));
    ensure_rows(n, rows);
    HIPCHECK(hipSetDevice(n->dev));
    float *xd = nullptr;
    HIPCHECK(hipMalloc(&xd, (size_t)count * NIN * sizeof(float)));
    try {
      HIPCHECK(hipMemcpyAsync(xd, x, (size_t)count * NIN * sizeof(float), hipMemcpyHostToDevice, n->stream));
      k_input_rows<<<grid_for(rows), 256, 0, n->stream>>>(xd, (uint32_t)count, rows, n->X0);
      forward(n, rows, false);
      k_head_infer<<<std::min<uint32_t>(rows / 4, 4096), 256, 0, n->stream>>>(
          n->H[NL - 1], n->w + OFF_WO, n->w + OFF_BO, (uint32_t)count, t_scale(n), 0.0f, n->out);
      HIPCHECK(hipGetLastError());
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
        uint32_t rows = (nreal + 127u) & ~127u;
        if (rows > n->cap) rows = (nreal + 63u) & ~63u;
        k_sample_rows<<<grid_for(rows), 256, 0, n->stream>>>(n->coords, md + (size_t)j0 * MAPP, sd, (uint32_t)m,
                                                             nreal, rows, n->X0);
        forward(n, rows, false);
        k_head_infer<<<std::min<uint32_t>(rows / 4, 4096), 256, 0, n->stream>>>(
            n->H[NL - 1], n->w + OFF_WO, n->w + OFF_BO, nreal, t_scale(n), 0.0f, n->out);
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
