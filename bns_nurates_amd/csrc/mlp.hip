// mlp.hip -- BF16-MFMA multilayer perceptron (ensemble) for the learned form
// of the gray M1 opacities: training kernels (GEMM, bias/SiLU forward and
// backward, MSE gradient, Adam), the native training loop, and the fused
// inference kernel that evaluates every ensemble member with activations in
// LDS.  No counterpart in the reference.
//
// Network of one member: 15 input features (padded to K = 32) -> 256 -> 256
// -> 256 -> 256 -> 15 outputs (padded to N = 16), SiLU on the hidden layers.
// Weights W_l [K_l][N_l] and biases b_l [N_l] are FP32 master copies; the
// matrix products take BF16 weights and activations and accumulate in FP32
// on the matrix cores (v_mfma_f32_16x16x32_bf16).
//
// MFMA 16x16x32 operand layout used here (64 lanes): A fragment of a 16 x 32
// tile, lane l holds row (l % 16), eight consecutive k starting at
// 8 (l / 16); B fragment of a 32 x 16 tile, lane l holds column (l % 16) at
// the same eight k; the 16 x 16 result, lane l holds rows 4 (l / 16) + v,
// v = 0..3, of column (l % 16).  (Any k order shared by A and B gives the
// same product.)
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <string>
#include <vector>

#include "../../include/bns_nurates_amd.h"
#include "context.h"
#include "gray_math.h"
#include "mlp.h"

namespace bns {
namespace mlp {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float floatx4 __attribute__((ext_vector_type(4)));

__host__ __device__ inline uint16_t f2bf(float f) {  // round to nearest even
#if defined(__HIP_DEVICE_COMPILE__)
  // v_cvt_pk_bf16_f32: the same rounding (NaN stays NaN)
  return __builtin_bit_cast(uint16_t, static_cast<__bf16>(f));
#else
  uint32_t u;
  std::memcpy(&u, &f, 4);
  if ((u & 0x7f800000u) == 0x7f800000u) return uint16_t((u >> 16) | ((u & 0xffff) ? 0x40 : 0));
  u += 0x7fffu + ((u >> 16) & 1u);
  return uint16_t(u >> 16);
#endif
}
__host__ __device__ inline float bf2f(uint16_t h) {
  uint32_t u = uint32_t(h) << 16;
  float f;
  std::memcpy(&f, &u, 4);
  return f;
}

__device__ inline bf16x8 ld8(const uint16_t* p) {  // 16-byte aligned LDS or global read
  uint4 v = *reinterpret_cast<const uint4*>(p);
  return __builtin_bit_cast(bf16x8, v);
}

// SiLU z / (1 + e^-z) with the hardware exponential and reciprocal (both
// within a few ulp of FP32; the activations are rounded to BF16 next)
__device__ inline float sigmoid_fast(float z) { return __builtin_amdgcn_rcpf(1.f + __expf(-z)); }
__device__ inline float silu(float z) { return z * sigmoid_fast(z); }
__device__ inline float silu_grad(float z) {
  float s = sigmoid_fast(z);
  return s * (1.f + z * (1.f - s));
}

// ---- training kernels ---------------------------------------------------------
// C[m][n] (+)= sum_k A(m,k) B(k,n), A(m,k) = A[m sam + k sak],
// B(k,n) = B[k sbk + n sbn] (BF16), C FP32 with leading dimension ldc.
// 64 x 64 tile per workgroup, 4 waves of 32 x 32, K in steps of 32; gridDim.z
// splits K into chunks of kchunk (SPLIT: chunk z writes its partial product
// to C + z csplit, summed in chunk order by k_sum_splits, so that the result
// does not depend on the order in which workgroups finish).
template <bool SPLIT>
__global__ __launch_bounds__(256) void k_gemm(int M, int N, int K, const uint16_t* __restrict__ A,
                                              int64_t sam, int64_t sak,
                                              const uint16_t* __restrict__ B, int64_t sbk,
                                              int64_t sbn, float* __restrict__ C, int64_t ldc,
                                              int kchunk, int64_t csplit) {
  __shared__ __attribute__((aligned(16))) uint16_t As[64][40];  // [m][k]
  __shared__ __attribute__((aligned(16))) uint16_t Bs[64][40];  // [n][k]
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  const int m0 = blockIdx.y * 64, n0 = blockIdx.x * 64;
  const int kbeg = blockIdx.z * kchunk;
  const int kend = min(K, kbeg + kchunk);
  const int wm = (wave >> 1) * 32, wn = (wave & 1) * 32;
  floatx4 acc[2][2];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) acc[i][j] = floatx4{0.f, 0.f, 0.f, 0.f};
  for (int k0 = kbeg; k0 < kend; k0 += 32) {
#pragma unroll
    for (int i = 0; i < 8; ++i) {
      int e = t + i * 256;
      int mm = sak == 1 ? e >> 5 : e & 63, kk = sak == 1 ? e & 31 : e >> 6;
      int gm = m0 + mm, gk = k0 + kk;
      As[mm][kk] = (gm < M && gk < kend) ? A[gm * sam + gk * sak] : uint16_t(0);
    }
#pragma unroll
    for (int i = 0; i < 8; ++i) {
      int e = t + i * 256;
      int nn = sbn == 1 ? e & 63 : e >> 5, kk = sbn == 1 ? e >> 6 : e & 31;
      int gn = n0 + nn, gk = k0 + kk;
      Bs[nn][kk] = (gn < N && gk < kend) ? B[gk * sbk + gn * sbn] : uint16_t(0);
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < 2; ++i) {
      bf16x8 a = ld8(&As[wm + i * 16 + (lane & 15)][8 * (lane >> 4)]);
#pragma unroll
      for (int j = 0; j < 2; ++j) {
        bf16x8 b = ld8(&Bs[wn + j * 16 + (lane & 15)][8 * (lane >> 4)]);
        acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, acc[i][j], 0, 0, 0);
      }
    }
    __syncthreads();
  }
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        int row = m0 + wm + i * 16 + 4 * (lane >> 4) + v, col = n0 + wn + j * 16 + (lane & 15);
        if (row < M && col < N) (SPLIT ? C + blockIdx.z * csplit : C)[row * ldc + col] = acc[i][j][v];
      }
}

// The same product for operands whose contiguous dimension has stride 1 and
// whose tiles can be moved in 16-byte pieces (the shapes of the training
// step): AK: A(m,k) = A[m lda + k] (else A[m + k lda]); BK: B(k,n) =
// B[k + n ldb] (else B[k ldb + n]).  Each thread moves one 16-byte piece of
// each operand per k step (transposed into the [m][k] / [n][k] LDS tiles
// where the contiguous dimension is m or n), and the next step's pieces are
// loaded while the current step multiplies.
template <bool SPLIT, bool AK, bool BK>
__global__ __launch_bounds__(256) void k_gemm_v(int M, int N, int K,
                                                const uint16_t* __restrict__ A, int64_t lda,
                                                const uint16_t* __restrict__ B, int64_t ldb,
                                                float* __restrict__ C, int64_t ldc, int kchunk,
                                                int64_t csplit, const float* __restrict__ bias,
                                                int act, uint16_t* __restrict__ H) {
  __shared__ __attribute__((aligned(16))) uint16_t As[64][40];  // [m][k]
  __shared__ __attribute__((aligned(16))) uint16_t Bs[64][40];  // [n][k]
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  const int m0 = blockIdx.y * 64, n0 = blockIdx.x * 64;
  const int kbeg = blockIdx.z * kchunk;
  const int kend = min(K, kbeg + kchunk);
  const int wm = (wave >> 1) * 32, wn = (wave & 1) * 32;
  floatx4 acc[2][2];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) acc[i][j] = floatx4{0.f, 0.f, 0.f, 0.f};
  // this thread's pieces: a row and 8 consecutive k, or a k and 8
  // consecutive rows
  const int ar = AK ? t >> 2 : (t & 7) * 8, ak = AK ? (t & 3) * 8 : t >> 3;
  const int br = BK ? t >> 2 : (t & 7) * 8, bk = BK ? (t & 3) * 8 : t >> 3;
  uint4 ra, rb;
  auto gload = [&](int k0) {
    const int gm = m0 + ar, gka = k0 + ak;
    const int gn = n0 + br, gkb = k0 + bk;
    ra = (gm < M && gka < kend)
             ? *reinterpret_cast<const uint4*>(A + (AK ? int64_t(gm) * lda + gka
                                                      : gm + int64_t(gka) * lda))
             : uint4{0u, 0u, 0u, 0u};
    rb = (gn < N && gkb < kend)
             ? *reinterpret_cast<const uint4*>(B + (BK ? gkb + int64_t(gn) * ldb
                                                      : int64_t(gkb) * ldb + gn))
             : uint4{0u, 0u, 0u, 0u};
  };
  auto sstore = [&]() {
    if (AK) {
      *reinterpret_cast<uint4*>(&As[ar][ak]) = ra;
    } else {
      const uint16_t* v = reinterpret_cast<const uint16_t*>(&ra);
#pragma unroll
      for (int j = 0; j < 8; ++j) As[ar + j][ak] = v[j];
    }
    if (BK) {
      *reinterpret_cast<uint4*>(&Bs[br][bk]) = rb;
    } else {
      const uint16_t* v = reinterpret_cast<const uint16_t*>(&rb);
#pragma unroll
      for (int j = 0; j < 8; ++j) Bs[br + j][bk] = v[j];
    }
  };
  gload(kbeg);
  for (int k0 = kbeg; k0 < kend; k0 += 32) {
    sstore();
    __syncthreads();
    if (k0 + 32 < kend) gload(k0 + 32);
#pragma unroll
    for (int i = 0; i < 2; ++i) {
      bf16x8 a = ld8(&As[wm + i * 16 + (lane & 15)][8 * (lane >> 4)]);
#pragma unroll
      for (int j = 0; j < 2; ++j) {
        bf16x8 b = ld8(&Bs[wn + j * 16 + (lane & 15)][8 * (lane >> 4)]);
        acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, acc[i][j], 0, 0, 0);
      }
    }
    __syncthreads();
  }
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        int row = m0 + wm + i * 16 + 4 * (lane >> 4) + v, col = n0 + wn + j * 16 + (lane & 15);
        if (row < M && col < N) {
          if (!SPLIT && bias) {
            // the forward pass's bias and activation (what k_bias_act does
            // to the stored product, the same operations)
            const float z = acc[i][j][v] + bias[col];
            C[int64_t(row) * ldc + col] = z;
            if (H) H[int64_t(row) * ldc + col] = f2bf(act ? silu(z) : z);
          } else {
            (SPLIT ? C + blockIdx.z * csplit : C)[int64_t(row) * ldc + col] = acc[i][j][v];
          }
        }
      }
}

// batch rows per split-K chunk of the weight-gradient products (256: 16
// chunks at batch 4096; 512 measured 12% slower per step, 128 no faster), at
// most kMaxSplits chunks (the partial products' workspace)
constexpr int kDwRows = 256;
constexpr int kMaxSplits = 16;
constexpr int kMaxSplitElems = kHidden * kHidden;  // largest M x N of a weight gradient

// C[i] += sum over z of part[z n + i], in z order
__global__ void k_sum_splits(int64_t n, int z, const float* __restrict__ part,
                             float* __restrict__ C) {
  int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float s = part[i];
  for (int k = 1; k < z; ++k) s += part[int64_t(k) * n + i];
  C[i] += s;
}

// db[c] += sum over b of part[b cols + c] for one column per workgroup: each
// thread sums a fixed stride of the partials, then a fixed tree
__global__ __launch_bounds__(256) void k_sum_columns(int nblk, int cols,
                                                     const float* __restrict__ part,
                                                     float* __restrict__ db) {
  __shared__ float red[256];
  const int c = blockIdx.x, t = threadIdx.x;
  float s = 0.f;
  for (int b = t; b < nblk; b += 256) s += part[int64_t(b) * cols + c];
  red[t] = s;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (t < w) red[t] += red[t + w];
    __syncthreads();
  }
  if (t == 0) db[c] += red[0];
}
// rows per workgroup of the element-wise training kernels (one thread per
// column): 512 workgroups at batch 4096
constexpr int kEwRows = 8;

// Z += b (kept as the pre-activation), H = BF16(act(Z)); one thread per
// column, kEwRows rows per workgroup
__global__ void k_bias_act(int rows, int cols, float* __restrict__ Z, const float* __restrict__ b,
                           int act, uint16_t* __restrict__ H) {
  int c = threadIdx.x, r0 = blockIdx.x * kEwRows;
  if (c >= cols) return;
  float bc = b[c];
  for (int r = r0; r < min(rows, r0 + kEwRows); ++r) {
    float z = Z[int64_t(r) * cols + c] + bc;
    Z[int64_t(r) * cols + c] = z;
    if (H) H[int64_t(r) * cols + c] = f2bf(act ? silu(z) : z);
  }
}

// dZ = dH act'(Z) as BF16, and this workgroup's column sums of dZ into
// part[blockIdx.x cols + c] (one thread per column, kEwRows rows per
// workgroup; k_sum_columns adds them to db)
__global__ void k_act_bwd(int rows, int cols, const float* __restrict__ dH,
                          const float* __restrict__ Z, uint16_t* __restrict__ dZ,
                          float* __restrict__ part) {
  int c = threadIdx.x, r0 = blockIdx.x * kEwRows;
  if (c >= cols) return;
  float s = 0.f;
  for (int r = r0; r < min(rows, r0 + kEwRows); ++r) {
    int64_t i = int64_t(r) * cols + c;
    float d = dH[i] * silu_grad(Z[i]);
    dZ[i] = f2bf(d);
    s += d;
  }
  part[int64_t(blockIdx.x) * cols + c] = s;
}

// last layer: y = Z (bias added), e = y - t on the first nout columns (0
// where the target is NaN: an output missing for that row); dZ = 2 e /
// (rows nout) as BF16; this workgroup's column sums of dZ into
// part[blockIdx.x cols + c] and its sum of e^2 into lpart[blockIdx.x]
// (summed in workgroup order by k_sum_columns and k_sum_loss)
__global__ void k_mse_grad(int rows, int cols, int nout, const float* __restrict__ Z,
                           const float* __restrict__ T, uint16_t* __restrict__ dZ,
                           float* __restrict__ part, float* __restrict__ lpart) {
  __shared__ float red[256], cs[256];
  int c = threadIdx.x % cols, rl = threadIdx.x / cols, rstep = blockDim.x / cols;
  int r0 = blockIdx.x * 256;
  float s = 0.f, l2 = 0.f;
  const float scale = 2.f / float(rows * nout);
  if (rl < rstep) {
    for (int r = r0 + rl; r < min(rows, r0 + 256); r += rstep) {
      int64_t i = int64_t(r) * cols + c;
      const float tv = T[i];
      float e = (c < nout && tv == tv) ? Z[i] - tv : 0.f;
      float d = e * scale;
      dZ[i] = f2bf(d);
      s += d;
      l2 += e * e;
    }
  }
  red[threadIdx.x] = l2;
  cs[threadIdx.x] = s;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) lpart[blockIdx.x] = red[0];
  if (threadIdx.x < cols) {
    float t = 0.f;
    for (int r = 0; r < rstep; ++r) t += cs[r * cols + threadIdx.x];
    part[int64_t(blockIdx.x) * cols + threadIdx.x] = t;
  }
}

// loss[0] += sum of the n workgroup sums, in order (one thread)
__global__ void k_sum_loss(int n, const float* __restrict__ lpart, float* __restrict__ loss) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  float s = 0.f;
  for (int i = 0; i < n; ++i) s += lpart[i];
  loss[0] += s;
}

// Adam step over the whole parameter blob; refreshes the BF16 copy
__global__ void k_adam(int64_t n, float* __restrict__ w, const float* __restrict__ g,
                       float* __restrict__ m, float* __restrict__ v, uint16_t* __restrict__ wb,
                       float lr, float b1, float b2, float eps, float c1, float c2) {
  int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float gi = g[i];
  float mi = b1 * m[i] + (1.f - b1) * gi;
  float vi = b2 * v[i] + (1.f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  float wi = w[i] - lr * (mi * c1) / (sqrtf(vi * c2) + eps);
  w[i] = wi;
  wb[i] = f2bf(wi);
}

// batch assembly: rows perm[start + r] of the normalised features (FP32,
// [n][15]) and targets ([n][15]) into the padded BF16 input [B][32] and FP32
// targets [B][16]
__global__ void k_gather_batch(int B, const int32_t* __restrict__ perm, int64_t start,
                               const float* __restrict__ X, const float* __restrict__ Y,
                               uint16_t* __restrict__ Xb, float* __restrict__ Tb) {
  int r = blockIdx.x * blockDim.y + threadIdx.y, c = threadIdx.x;  // 32 x 8
  if (r >= B) return;
  int64_t src = perm[start + r];
  uint16_t h = 0;
  if (c < kFeat) {
    h = f2bf(X[src * kFeat + c]);
  } else if (c < 2 * kFeat) {  // low part of feature c - kFeat
    const float x = X[src * kFeat + c - kFeat];
    h = f2bf(x - bf2f(f2bf(x)));
  }
  Xb[int64_t(r) * kK0 + c] = h;
  if (Y && c < kNL) Tb[int64_t(r) * kNL + c] = c < kOut ? Y[src * kOut + c] : 0.f;
}

__global__ void k_fill(int64_t n, float* __restrict__ p, float v) {
  int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

// tied first layer: the gradient of the low-part rows adds into the rows of
// their features (and is cleared), so that Adam moves the shared weights
__global__ void k_tie_grad_l0(float* __restrict__ g) {
  const int n = threadIdx.x;  // kHidden threads
#pragma unroll
  for (int i = 0; i < kFeat; ++i) {
    g[i * kHidden + n + kWOff[0]] += g[(kFeat + i) * kHidden + n + kWOff[0]];
    g[(kFeat + i) * kHidden + n + kWOff[0]] = 0.f;
  }
}
// the BF16 copy of W_0: low-part rows mirror the feature rows
__global__ void k_mirror_l0(uint16_t* __restrict__ wb) {
  const int n = threadIdx.x;  // kHidden threads
#pragma unroll
  for (int i = 0; i < kFeat; ++i)
    wb[(kFeat + i) * kHidden + n + kWOff[0]] = wb[i * kHidden + n + kWOff[0]];
}

// ---- inference ---------------------------------------------------------------------
// BF16 weights of one layer in MFMA fragment order: block (kb, nb) of 32 x 16
// holds, for lane l, W[32 kb + 8 (l / 16) + j][16 nb + l % 16], j = 0..7.
// mirror: the first layer, whose rows kFeat .. 2 kFeat - 1 take the weights
// of rows 0 .. kFeat - 1.
__global__ void k_pack(int K, int N, const float* __restrict__ Wf, uint16_t* __restrict__ Wp,
                       int mirror) {
  int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= int64_t(K) * N) return;
  int j = int(i & 7), lane = int((i >> 3) & 63);
  int64_t blk = i >> 9;
  int NB = N / 16, kb = int(blk / NB), nb = int(blk % NB);
  int k = 32 * kb + 8 * (lane >> 4) + j, n = 16 * nb + (lane & 15);
  if (mirror && k >= kFeat && k < 2 * kFeat) k -= kFeat;
  Wp[i] = f2bf(Wf[int64_t(k) * N + n]);
}

struct InferArgs {
  int64_t npts, ld;
  const double* pts;                 // [21][ld]
  const bnsamd_mlp_norm* norm;
  const uint16_t* wp[kMaxMembers];   // packed weights per member
  const float* w[kMaxMembers];       // FP32 blobs (biases)
  int members;
  float threshold;                   // predicted log error that triggers fallback
  // acceptance judge (bnsamd_mlp_set_judge): one network of the members'
  // shape whose outputs are the predicted log margins of the 15 outputs;
  // a point falls back when the largest exceeds jthreshold (jwp == nullptr:
  // no judge)
  const uint16_t* jwp;
  const float* jw;
  const bnsamd_mlp_norm* jnorm;
  double jthreshold;
  double dU, dm_eff;                 // training-domain values of these point fields
  double mu_np_shift;                // dQ + dU of the beta rates in the training domain
  double beta_guard;                 // |mu_n - mu_p - mu_np_shift| < beta_guard T: fall back
  double* out;                       // [20][ld]
  int32_t* fallback_idx;             // compacted indices of the points to recompute
  int32_t* fallback_count;
  // calibration form: raw predictions instead of outputs and fallback
  double* cal_mean;                  // [15][ld] mean log output
  float* cal_spread;                 // [15][ld] members' spread of the log output
  int32_t* cal_outside;              // [ld] outside the training box or domain
};

constexpr int kRows = 64, kLds = kHidden + 8;  // points per workgroup, padded row
constexpr int kXLd = kK0 + 8;                    // padded row of the input features


// One layer for this wave's 64 output features (16 nb0 .. 16 nb0 + 63) and
// the workgroup's 64 points, with the operands of the MFMA swapped: the
// packed weight fragment is the A operand (rows = output features) and the
// activation fragment (rows = points in LDS) the B operand, so the result is
// the transposed tile and lane l holds output features 4 (l / 16) + v,
// v = 0..3, of point l % 16 -- four consecutive features of one point,
// stored with one 8-byte LDS write.  K is fixed so that the k loop unrolls;
// the weights of step kb + 1 are loaded while step kb multiplies.
template <int K>
__device__ inline void mma_layer_t(const uint16_t* __restrict__ A, int stride,
                                   const uint16_t* __restrict__ Wp, int NB, int nb0, int lane,
                                   floatx4 (&acc)[4][4]) {
#pragma unroll
  for (int rt = 0; rt < 4; ++rt)
#pragma unroll
    for (int ct = 0; ct < 4; ++ct) acc[rt][ct] = floatx4{0.f, 0.f, 0.f, 0.f};
  bf16x8 w[2][4];
#pragma unroll
  for (int ct = 0; ct < 4; ++ct) w[0][ct] = ld8(Wp + ((nb0 + ct) * 64 + lane) * 8);
#pragma unroll
  for (int kb = 0; kb < K / 32; ++kb) {
    const int cur = kb & 1;
    if (kb + 1 < K / 32) {
#pragma unroll
      for (int ct = 0; ct < 4; ++ct)
        w[cur ^ 1][ct] = ld8(Wp + ((int64_t(kb + 1) * NB + nb0 + ct) * 64 + lane) * 8);
    }
#pragma unroll
    for (int rt = 0; rt < 4; ++rt) {
      bf16x8 x = ld8(A + (rt * 16 + (lane & 15)) * stride + 32 * kb + 8 * (lane >> 4));
#pragma unroll
      for (int ct = 0; ct < 4; ++ct)
        acc[rt][ct] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(w[cur][ct], x, acc[rt][ct], 0, 0, 0);
    }
  }
}

// Whether the FP64 path's density thresholds set output kind q (0 eta_0,
// 1 eta, 2 kappa_0_a, 3 kappa_a, 4 kappa_s) of species sp to zero
// (m1_opacities.hpp:1199-1292): kappa_0_a needs n > 1e-21, kappa_a and
// kappa_s need J > 1e-25, and kappa_a of nue is zero when n is.  Such
// outputs are not the network's (training masks them), so they neither
// enter the spread test nor take its value.
__device__ inline bool threshold_zero(int q, int sp, double ns, double Js) {
  if (q == 2) return !(ns > bns::c::ThresholdN);
  if (q >= 3) return !(Js > bns::c::ThresholdJ) || (q == 3 && sp == 0 && ns == 0.);
  return false;
}

// bias + SiLU of one transposed tile, four features of one point per lane,
// written as one 8-byte row segment of H
__device__ inline void store_act4(uint16_t* __restrict__ dst, floatx4 v, const float4 b) {
  typedef __bf16 bf16x4 __attribute__((ext_vector_type(4)));
  bf16x4 h;
  h[0] = static_cast<__bf16>(silu(v[0] + b.x));
  h[1] = static_cast<__bf16>(silu(v[1] + b.y));
  h[2] = static_cast<__bf16>(silu(v[2] + b.z));
  h[3] = static_cast<__bf16>(silu(v[3] + b.w));
  *reinterpret_cast<uint2*>(dst) = __builtin_bit_cast(uint2, h);
}

// One workgroup (4 waves) evaluates every ensemble member on 64 points.
// LDS holds the points' BF16 features and one activation buffer, updated in
// place: a layer's outputs stay in the MFMA accumulators until every wave
// has read its inputs.  212 VGPRs: 2 waves per SIMD (3 spill, and measured
// no faster).  CAL: write the raw predictions (mean log output, spread,
// outside flag) instead of outputs and fallback indices.
template <bool CAL>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2, 2))) void k_infer(InferArgs a) {
  __shared__ __attribute__((aligned(16))) uint16_t X[kRows][kXLd];
  __shared__ __attribute__((aligned(16))) uint16_t H[kRows][kLds];
  __shared__ float spread[kRows][kNL];
  __shared__ double jmargin[kRows][kNL];
  __shared__ float xs[kRows][kFeat];  // standardised features (linear output model)
  __shared__ int oob[kRows];
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  const int64_t p0 = int64_t(blockIdx.x) * kRows;
  const bnsamd_mlp_norm& nm = *a.norm;

  // features of this workgroup's points (thread p), normalised, checked
  // against the training box
  if (t < kRows) {
    int64_t p = p0 + t;
    float f[kFeat];
    bool in = p < a.npts;
    if (in) {
      const double* q = a.pts + p;
      const int64_t ld = a.ld;
      f[0] = float(log(q[0 * ld]));
      f[1] = float(log(q[1 * ld]));
      f[2] = float(q[2 * ld]);
#pragma unroll
      for (int s = 0; s < 3; ++s) {
        f[3 + 3 * s] = float(log(q[(9 + s) * ld]));
        f[4 + 3 * s] = float(log(q[(13 + s) * ld]));
        f[5 + 3 * s] = float(q[(17 + s) * ld]);
      }
      const double rT = 1. / q[1 * ld];
      f[12] = float(q[6 * ld] * rT);                // mu_e / T
      f[13] = float((q[5 * ld] - q[4 * ld]) * rT);  // (mu_n - mu_p) / T
      f[14] = float(rT);
      // the inputs that are not features must take their training-domain
      // values: anux moments equal to nux's, yn = 1 - yp, dU and dm_eff
      in = q[12 * ld] == q[11 * ld] && q[16 * ld] == q[15 * ld] && q[20 * ld] == q[19 * ld] &&
           q[3 * ld] == 1. - q[2 * ld] && q[7 * ld] == a.dU && q[8 * ld] == a.dm_eff;
      // near the branch surfaces of the beta rates' nucleon factors
      in = in && !(fabs(q[5 * ld] - q[4 * ld] - a.mu_np_shift) < a.beta_guard * q[1 * ld]);
    } else {
#pragma unroll
      for (int i = 0; i < kFeat; ++i) f[i] = nm.feat_mean[i];
    }
    bool out_of_box = !in;
#pragma unroll
    for (int i = 0; i < kFeat; ++i) {
      out_of_box = out_of_box || !(f[i] >= nm.box_lo[i] && f[i] <= nm.box_hi[i]);
      const float x = (f[i] - nm.feat_mean[i]) / nm.feat_std[i];
      xs[t][i] = out_of_box ? 0.f : x;
      const uint16_t hi = f2bf(x);
      X[t][i] = hi;
      X[t][kFeat + i] = f2bf(x - bf2f(hi));
    }
#pragma unroll
    for (int i = 2 * kFeat; i < kXLd; ++i) X[t][i] = 0;
    oob[t] = out_of_box;
  }
  __syncthreads();

  // running ensemble sums of this lane's four outputs: outputs
  // 4 (lane / 16) + v of point 16 wave + lane % 16
  float sum[4] = {0.f, 0.f, 0.f, 0.f}, sq[4] = {0.f, 0.f, 0.f, 0.f};
  float jy[4] = {0.f, 0.f, 0.f, 0.f};  // the judge's outputs (standardised)
  const int q4 = 4 * (lane >> 4);
  const int passes = a.members + (!CAL && a.jwp ? 1 : 0);
  for (int e = 0; e < passes; ++e) {
    const bool judge = e == a.members;
    const uint16_t* wp = judge ? a.jwp : a.wp[e];
    const float* wf = judge ? a.jw : a.w[e];
#pragma unroll 1
    for (int l = 0; l < kLayers - 1; ++l) {
      floatx4 acc[4][4];
      if (l == 0)
        mma_layer_t<kK0>(&X[0][0], kXLd, wp + kPOff[0], kHidden / 16, 4 * wave, lane, acc);
      else
        mma_layer_t<kHidden>(&H[0][0], kLds, wp + kPOff[l], kHidden / 16, 4 * wave, lane, acc);
      __syncthreads();  // every wave has read H: overwrite it in place
      const float* bias = wf + kBOff[l];
#pragma unroll
      for (int ct = 0; ct < 4; ++ct) {
        const int col = 64 * wave + 16 * ct + q4;
        const float4 bc = *reinterpret_cast<const float4*>(bias + col);
#pragma unroll
        for (int rt = 0; rt < 4; ++rt) store_act4(&H[rt * 16 + (lane & 15)][col], acc[rt][ct], bc);
      }
      __syncthreads();
    }
    // output layer (16 padded outputs): wave w takes points 16 w .. 16 w + 15
    floatx4 o = floatx4{0.f, 0.f, 0.f, 0.f};
    const uint16_t* wl = wp + kPOff[kLayers - 1];
#pragma unroll
    for (int kb = 0; kb < kHidden / 32; ++kb) {
      bf16x8 b = ld8(wl + (int64_t(kb) * 64 + lane) * 8);
      bf16x8 x = ld8(&H[16 * wave + (lane & 15)][32 * kb + 8 * (lane >> 4)]);
      o = __builtin_amdgcn_mfma_f32_16x16x32_bf16(b, x, o, 0, 0, 0);
    }
    const float4 bo = *reinterpret_cast<const float4*>(wf + kBOff[kLayers - 1] + q4);
    const float bov[4] = {bo.x, bo.y, bo.z, bo.w};
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      float y = o[v] + bov[v];
      if (judge) {
        jy[v] = y;
      } else {
        sum[v] += y;
        sq[v] += y * y;
      }
    }
    __syncthreads();  // the next member overwrites H
  }

  // ensemble mean -> outputs; spread (natural-log units) -> fallback flag
  const int r = 16 * wave + (lane & 15);
  const int64_t p = p0 + r;
  const bool live = p < a.npts;
  const float inv = 1.f / float(a.members);
  float mean[4];
#pragma unroll
  for (int v = 0; v < 4; ++v) {
    const int c = q4 + v;
    const float os = c < kOut ? nm.out_std[c] : 0.f;
    const float rlo = c < kOut ? nm.res_lo[c] : 0.f, rhi = c < kOut ? nm.res_hi[c] : 0.f;
    mean[v] = sum[v] * inv;
    float var = fmaxf(sq[v] * inv - mean[v] * mean[v], 0.f);
    bool zero = false;
    if (c < kOut && live) {
      const int sp = c / 5;
      zero = threshold_zero(c % 5, sp, a.pts[(9 + sp) * a.ld + p], a.pts[(13 + sp) * a.ld + p]);
    }
    // a residual outside the training range counts as unbounded spread
    bool in_range = mean[v] >= rlo && mean[v] <= rhi;
    const float sd = (c < kOut && !zero) ? (in_range ? sqrtf(var) * os : __builtin_huge_valf())
                                         : 0.f;
    if constexpr (CAL) {
      if (c < kOut && live) {
        float lin = nm.lin[0][c];
#pragma unroll
        for (int i = 0; i < kFeat; ++i) lin = fmaf(nm.lin[1 + i][c], xs[r][i], lin);
        a.cal_mean[c * a.ld + p] = double(lin) + double(mean[v] * nm.out_std[c] + nm.out_mean[c]);
        a.cal_spread[c * a.ld + p] = zero ? -1.f : sd;
      }
    } else {
      // predicted error of this log output (0 for a threshold zero; a
      // residual outside the training range rejects whatever err_k is)
      spread[r][c] = (c < kOut && !zero)
                         ? (sd < __builtin_huge_valf() ? nm.err_a[c] + nm.err_k[c] * sd : sd)
                         : 0.f;
      // the judge's predicted log margin, formed as the calibration form
      // forms a one-member ensemble's output (a residual outside the
      // judge's training range: unbounded); threshold zeros are exact
      double jm = -__builtin_huge_val();
      if (a.jwp && c < kOut && !zero) {
        const bnsamd_mlp_norm& jn = *a.jnorm;
        if (jy[v] >= jn.res_lo[c] && jy[v] <= jn.res_hi[c]) {
          float lin = jn.lin[0][c];
#pragma unroll
          for (int i = 0; i < kFeat; ++i) lin = fmaf(jn.lin[1 + i][c], xs[r][i], lin);
          jm = double(lin) + double(jy[v] * jn.out_std[c] + jn.out_mean[c]);
        } else {
          jm = __builtin_huge_val();
        }
      }
      if (c < kNL) jmargin[r][c] = jm;
    }
  }
  if constexpr (CAL) {
    if (t < kRows && p0 + t < a.npts) a.cal_outside[p0 + t] = oob[t];
    return;
  }
  __syncthreads();
  if (t < kRows) {
    float s = 0.f;
#pragma unroll
    for (int c = 0; c < kOut; ++c) s = fmaxf(s, spread[t][c]);
    double jmax = -__builtin_huge_val();
    if (a.jwp) {
#pragma unroll
      for (int c = 0; c < kOut; ++c) jmax = fmax(jmax, jmargin[t][c]);
    }
    int fb = oob[t] || !(s <= a.threshold) || (a.jwp && !(jmax <= a.jthreshold));
    oob[t] = fb;
    int64_t p = p0 + t;
    if (fb && p < a.npts) {
      int slot = atomicAdd(a.fallback_count, 1);
      a.fallback_idx[slot] = int32_t(p);
    }
  }
  __syncthreads();
  // outputs of the accepted points: column c = 5 s + q (q: eta_0, eta,
  // kappa_0_a, kappa_a, kappa_s) goes to field 4 q' + s, anux copying nux.
  // The opacities keep the zeros the FP64 path gives below the density
  // thresholds (for anux, its own n and J decide).
#pragma unroll
  for (int v = 0; v < 4; ++v) {
    const int c = q4 + v;
    if (c < kOut && live && !oob[r]) {
      const int s = c / 5, q = c % 5;
      const int base = q == 0 ? 0 : q == 1 ? 8 : q == 2 ? 4 : q == 3 ? 12 : 16;
      float lin = nm.lin[0][c];
#pragma unroll
      for (int i = 0; i < kFeat; ++i) lin = fmaf(nm.lin[1 + i][c], xs[r][i], lin);
      const double y = exp(double(lin) + double(mean[v] * nm.out_std[c] + nm.out_mean[c]));
      auto value = [&](int sp) {
        const double ns = a.pts[(9 + sp) * a.ld + p], Js = a.pts[(13 + sp) * a.ld + p];
        return threshold_zero(q, sp, ns, Js) ? 0. : y;
      };
      a.out[(base + s) * a.ld + p] = value(s);
      if (s == 2) a.out[(base + 3) * a.ld + p] = value(3);
    }
  }
}

// compact copies of the fallback points and back
__global__ void k_gather_pts(int32_t n, const int32_t* __restrict__ idx, const double* __restrict__ pts,
                             int64_t ld, double* __restrict__ cp) {
  int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  int64_t p = idx[i];
  for (int f = 0; f < BNSAMD_PT_NFIELDS; ++f) cp[int64_t(f) * n + i] = pts[f * ld + p];
}
__global__ void k_scatter_out(int32_t n, const int32_t* __restrict__ idx, const double* __restrict__ co,
                              double* __restrict__ out, int64_t ld) {
  int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  int64_t p = idx[i];
  for (int f = 0; f < BNSAMD_OUT_NFIELDS; ++f) out[f * ld + p] = co[int64_t(f) * n + i];
}

}  // namespace mlp
}  // namespace bns

namespace bns {
namespace mlp {
// training workspace of one stream: maxb rows of every layer's activations
struct Work {
  hipStream_t stream = nullptr;
  bool own_stream = false;
  uint16_t* Xb = nullptr;
  float* Tb = nullptr;
  float* Z[kLayers] = {};
  uint16_t* H[kLayers] = {};
  uint16_t* dZ[kLayers] = {};
  float* dH = nullptr;
  float* loss = nullptr;
  // per-workgroup partial sums, added in a fixed order so that training is
  // reproducible: split-K weight-gradient products, bias-gradient column
  // sums, loss sums
  float* part = nullptr;
  float* colpart = nullptr;
  float* lpart = nullptr;
};
}  // namespace mlp
}  // namespace bns

using namespace bns::mlp;

// ---- the MLP object -------------------------------------------------------------
struct bnsamd_mlp {
  int device = 0, members = 0, maxb = 0;
  hipStream_t stream = nullptr;
  std::string err;
  // per member: FP32 parameters, gradients, Adam moments, BF16 copy, packed
  // BF16 weights for the inference kernel
  std::vector<float*> w, g, m, v;
  std::vector<uint16_t*> wb, wp;
  std::vector<int64_t> steps;
  float* norm = nullptr;  // device copy of the normalisation block
  bnsamd_mlp_norm hnorm{};
  // training workspaces: ws[0] on the main stream (single-member training,
  // gradients, prediction); ws[e] for member e of bnsamd_mlp_train_ensemble,
  // on its own stream, allocated on first use
  std::vector<Work> ws;
  // fallback workspace (grown on demand)
  // training domain (bnsamd_mlp_set_domain)
  bool domain_set = false;
  bnsamd_mlp_domain domain{};
  int64_t fb_cap = 0;
  int32_t* fb_idx = nullptr;
  int32_t* fb_count = nullptr;
  double* fb_pts = nullptr;
  double* fb_out = nullptr;
  bool packed_valid = false;
  // acceptance judge: a copy of another instance's member 0 (FP32 blob,
  // packed BF16 weights) and of its normalisation block
  bool judge_set = false;
  float* jw = nullptr;
  uint16_t* jwp = nullptr;
  float* jnorm = nullptr;
  double jthreshold = 0.;
};

namespace {

int mlp_fail(bnsamd_mlp* m, int code, const std::string& msg) {
  if (m) m->err = msg;
  return code;
}

template <bool SPLIT>
void gemm_v(hipStream_t s, dim3 grid, bool ak, bool bk, int M, int N, int K, const uint16_t* A,
            int64_t lda, const uint16_t* B, int64_t ldb, float* C, int64_t ldc, int kchunk,
            int64_t csplit, const float* bias, int act, uint16_t* H) {
  const dim3 block(256);
  if (ak && bk)
    hipLaunchKernelGGL((k_gemm_v<SPLIT, true, true>), grid, block, 0, s, M, N, K, A, lda, B, ldb,
                       C, ldc, kchunk, csplit, bias, act, H);
  else if (ak)
    hipLaunchKernelGGL((k_gemm_v<SPLIT, true, false>), grid, block, 0, s, M, N, K, A, lda, B,
                       ldb, C, ldc, kchunk, csplit, bias, act, H);
  else if (bk)
    hipLaunchKernelGGL((k_gemm_v<SPLIT, false, true>), grid, block, 0, s, M, N, K, A, lda, B,
                       ldb, C, ldc, kchunk, csplit, bias, act, H);
  else
    hipLaunchKernelGGL((k_gemm_v<SPLIT, false, false>), grid, block, 0, s, M, N, K, A, lda, B,
                       ldb, C, ldc, kchunk, csplit, bias, act, H);
}

// C = A B, or with splitk > 1 C += A B as splitk partial products over
// chunks of K written to part (room for kMaxSplits x M x N floats; C then
// M x N contiguous, ldc = N) and summed in chunk order.  With bias: C = A B
// + bias per column and H = BF16(act ? SiLU(C) : C) (ldc = N), in the
// product's epilogue where it has one (else by k_bias_act)
void gemm(hipStream_t s, int M, int N, int K, const uint16_t* A, int64_t sam, int64_t sak,
          const uint16_t* B, int64_t sbk, int64_t sbn, float* C, int64_t ldc, int splitk,
          float* part = nullptr, const float* bias = nullptr, int act = 0,
          uint16_t* H = nullptr) {
  int kchunk = (K + splitk - 1) / splitk;
  kchunk = (kchunk + 31) / 32 * 32;
  int z = (K + kchunk - 1) / kchunk;
  dim3 grid((N + 63) / 64, (M + 63) / 64, z), block(256);
  const int64_t mn = int64_t(M) * N;
  float* const Cz = z > 1 ? part : C;  // where the kernel writes
  auto finish = [&](bool epilogue_done) {
    if (z > 1)
      hipLaunchKernelGGL(k_sum_splits, dim3(unsigned((mn + 255) / 256)), dim3(256), 0, s, mn, z,
                         part, C);
    if (bias && !epilogue_done)
      hipLaunchKernelGGL(k_bias_act, dim3((M + kEwRows - 1) / kEwRows), dim3(256), 0, s, M, N, C,
                         bias, act, H);
  };
  // 16-byte pieces: the contiguous dimension of each operand has stride 1,
  // its extent and the other stride are multiples of 8, the bases are
  // 16-byte aligned
  auto a16 = [](const void* p) { return (reinterpret_cast<uintptr_t>(p) & 15) == 0; };
  const bool ak = sak == 1, bk = sbk == 1;
  const int64_t lda = ak ? sam : sak, ldb = bk ? sbn : sbk;
  const bool vec = (ak || sam == 1) && (bk || sbn == 1) && a16(A) && a16(B) && lda % 8 == 0 &&
                   ldb % 8 == 0 && (ak ? K % 8 == 0 : M % 8 == 0) &&
                   (bk ? K % 8 == 0 : N % 8 == 0);
  if (vec) {
    if (z > 1) {
      gemm_v<true>(s, grid, ak, bk, M, N, K, A, lda, B, ldb, Cz, ldc, kchunk, mn, nullptr, 0,
                   nullptr);
      finish(false);
    } else {
      gemm_v<false>(s, grid, ak, bk, M, N, K, A, lda, B, ldb, Cz, ldc, kchunk, 0, bias, act, H);
      finish(true);
    }
    return;
  }
  if (z > 1)
    hipLaunchKernelGGL(k_gemm<true>, grid, block, 0, s, M, N, K, A, sam, sak, B, sbk, sbn, Cz, ldc,
                       kchunk, mn);
  else
    hipLaunchKernelGGL(k_gemm<false>, grid, block, 0, s, M, N, K, A, sam, sak, B, sbk, sbn, Cz,
                       ldc, kchunk, int64_t(0));
  finish(false);
}

void fill(hipStream_t s, int64_t n, float* p, float v) {
  hipLaunchKernelGGL(k_fill, dim3(unsigned((n + 255) / 256)), dim3(256), 0, s, n, p, v);
}

// forward pass of member e on B rows of W.Xb (BF16 [B][32]); W.Z[l] hold the
// pre-activations, W.H[l] the BF16 activations
void forward(bnsamd_mlp* m, Work& W, int e, int B) {
  hipStream_t s = W.stream;
  const uint16_t* in = W.Xb;
  for (int l = 0; l < kLayers; ++l) {
    const int K = kLayerK[l], N = kLayerN[l];
    gemm(s, B, N, K, in, K, 1, m->wb[e] + kWOff[l], N, 1, W.Z[l], N, 1, nullptr,
         m->w[e] + kBOff[l], l + 1 < kLayers ? 1 : 0,
         l + 1 < kLayers ? W.H[l] : static_cast<uint16_t*>(nullptr));
    in = W.H[l];
  }
}

// loss, gradients of member e on B rows (W.Xb, W.Tb) into g[e]
void backward(bnsamd_mlp* m, Work& W, int e, int B) {
  hipStream_t s = W.stream;
  fill(s, kParams, m->g[e], 0.f);
  const int L = kLayers - 1;
  const int nmse = (B + 255) / 256, nbwd = (B + kEwRows - 1) / kEwRows;
  hipLaunchKernelGGL(k_mse_grad, dim3(nmse), dim3(256), 0, s, B, kNL, kOut, W.Z[L], W.Tb,
                     W.dZ[L], W.colpart, W.lpart);
  hipLaunchKernelGGL(k_sum_columns, dim3(kNL), dim3(256), 0, s, nmse, kNL, W.colpart,
                     m->g[e] + kBOff[L]);
  hipLaunchKernelGGL(k_sum_loss, dim3(1), dim3(64), 0, s, nmse, W.lpart, W.loss);
  for (int l = L; l >= 0; --l) {
    const int K = kLayerK[l], N = kLayerN[l];
    const uint16_t* Hin = l ? W.H[l - 1] : W.Xb;
    // dW = H_in^T dZ (reduction over the batch, split into chunks of at
    // least kDwRows rows)
    gemm(s, K, N, B, Hin, 1, K, W.dZ[l], N, 1, m->g[e] + kWOff[l], N,
         std::min((B + kDwRows - 1) / kDwRows, kMaxSplits), W.part);
    if (l) {
      // dH_in = dZ W^T, then through the SiLU of layer l - 1
      gemm(s, B, K, N, W.dZ[l], N, 1, m->wb[e] + kWOff[l], 1, N, W.dH, K, 1);
      hipLaunchKernelGGL(k_act_bwd, dim3(nbwd), dim3(256), 0, s, B, K, W.dH, W.Z[l - 1],
                         W.dZ[l - 1], W.colpart);
      hipLaunchKernelGGL(k_sum_columns, dim3(K), dim3(256), 0, s, nbwd, K, W.colpart,
                         m->g[e] + kBOff[l - 1]);
    }
  }
  hipLaunchKernelGGL(k_tie_grad_l0, dim3(1), dim3(kHidden), 0, s, m->g[e]);
}

void adam(bnsamd_mlp* m, hipStream_t s, int e, float lr) {
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  int64_t t = ++m->steps[e];
  float c1 = 1.f / (1.f - std::pow(b1, float(t))), c2 = 1.f / (1.f - std::pow(b2, float(t)));
  hipLaunchKernelGGL(k_adam, dim3(unsigned((kParams + 255) / 256)), dim3(256), 0, s,
                     int64_t(kParams), m->w[e], m->g[e], m->m[e], m->v[e], m->wb[e], lr, b1, b2,
                     eps, c1, c2);
  hipLaunchKernelGGL(k_mirror_l0, dim3(1), dim3(kHidden), 0, s, m->wb[e]);
}

// buffers of one workspace (maxb rows); false when an allocation fails
bool work_alloc(Work& W, int maxb) {
  bool ok = true;
  auto alloc = [&](void** p, size_t bytes) {
    ok = ok && hipMalloc(p, bytes) == hipSuccess && hipMemset(*p, 0, bytes) == hipSuccess;
  };
  const size_t B = size_t(maxb);
  alloc(reinterpret_cast<void**>(&W.Xb), sizeof(uint16_t) * B * kK0);
  alloc(reinterpret_cast<void**>(&W.Tb), sizeof(float) * B * kNL);
  for (int l = 0; l < kLayers; ++l) {
    alloc(reinterpret_cast<void**>(&W.Z[l]), sizeof(float) * B * kLayerN[l]);
    alloc(reinterpret_cast<void**>(&W.H[l]), sizeof(uint16_t) * B * kLayerN[l]);
    alloc(reinterpret_cast<void**>(&W.dZ[l]), sizeof(uint16_t) * B * kLayerN[l]);
  }
  alloc(reinterpret_cast<void**>(&W.dH), sizeof(float) * B * kHidden);
  alloc(reinterpret_cast<void**>(&W.loss), sizeof(float));
  alloc(reinterpret_cast<void**>(&W.part), sizeof(float) * kMaxSplits * kMaxSplitElems);
  alloc(reinterpret_cast<void**>(&W.colpart),
        sizeof(float) * ((B + kEwRows - 1) / kEwRows) * kHidden);
  alloc(reinterpret_cast<void**>(&W.lpart), sizeof(float) * ((B + 255) / 256));
  return ok;
}

void work_free(Work& W) {
  auto fr = [](void* p) {
    if (p) (void)hipFree(p);
  };
  if (W.stream) (void)hipStreamSynchronize(W.stream);
  fr(W.Xb);
  fr(W.Tb);
  for (int l = 0; l < kLayers; ++l) {
    fr(W.Z[l]);
    fr(W.H[l]);
    fr(W.dZ[l]);
  }
  fr(W.dH);
  fr(W.loss);
  fr(W.part);
  fr(W.colpart);
  fr(W.lpart);
  if (W.own_stream && W.stream) (void)hipStreamDestroy(W.stream);
  W = Work{};
}

// every stream that training may still be using
hipError_t sync_all(bnsamd_mlp* m) {
  hipError_t e = hipSuccess;
  for (auto& W : m->ws)
    if (W.stream && e == hipSuccess) e = hipStreamSynchronize(W.stream);
  return e;
}

}  // namespace

extern "C" {

int bnsamd_mlp_create(int device, int members, int max_batch, bnsamd_mlp** out) {
  if (!out || members < 1 || members > kMaxMembers || max_batch < 64 || max_batch > (1 << 16))
    return BNSAMD_ERR_INVALID;
  *out = nullptr;
  if (hipSetDevice(device) != hipSuccess) return BNSAMD_ERR_NO_DEVICE;
  auto* m = new bnsamd_mlp();
  m->device = device;
  m->members = members;
  m->maxb = max_batch;
  for (int o = 0; o < kOut; ++o) m->hnorm.err_k[o] = 1.f;  // predicted error = spread
  bool ok = hipStreamCreateWithFlags(&m->stream, hipStreamNonBlocking) == hipSuccess;
  auto alloc = [&](void** p, size_t bytes) {
    ok = ok && hipMalloc(p, bytes) == hipSuccess && hipMemset(*p, 0, bytes) == hipSuccess;
  };
  for (int e = 0; e < members && ok; ++e) {
    float *w, *g, *mm, *v;
    uint16_t *wb, *wp;
    alloc(reinterpret_cast<void**>(&w), sizeof(float) * kParams);
    alloc(reinterpret_cast<void**>(&g), sizeof(float) * kParams);
    alloc(reinterpret_cast<void**>(&mm), sizeof(float) * kParams);
    alloc(reinterpret_cast<void**>(&v), sizeof(float) * kParams);
    alloc(reinterpret_cast<void**>(&wb), sizeof(uint16_t) * kParams);
    alloc(reinterpret_cast<void**>(&wp), sizeof(uint16_t) * kPacked);
    m->w.push_back(w);
    m->g.push_back(g);
    m->m.push_back(mm);
    m->v.push_back(v);
    m->wb.push_back(wb);
    m->wp.push_back(wp);
    m->steps.push_back(0);
  }
  alloc(reinterpret_cast<void**>(&m->norm), sizeof(bnsamd_mlp_norm));
  m->ws.resize(1);
  m->ws[0].stream = m->stream;
  ok = ok && work_alloc(m->ws[0], max_batch);
  if (!ok) {
    bnsamd_mlp_destroy(m);
    return BNSAMD_ERR_NOMEM;
  }
  *out = m;
  return BNSAMD_OK;
}

void bnsamd_mlp_destroy(bnsamd_mlp* m) {
  if (!m) return;
  (void)hipSetDevice(m->device);
  (void)sync_all(m);
  auto fr = [](void* p) {
    if (p) (void)hipFree(p);
  };
  for (auto* p : m->w) fr(p);
  for (auto* p : m->g) fr(p);
  for (auto* p : m->m) fr(p);
  for (auto* p : m->v) fr(p);
  for (auto* p : m->wb) fr(p);
  for (auto* p : m->wp) fr(p);
  fr(m->norm);
  for (auto& W : m->ws) work_free(W);
  fr(m->fb_idx);
  fr(m->fb_count);
  fr(m->fb_pts);
  fr(m->fb_out);
  fr(m->jw);
  fr(m->jwp);
  fr(m->jnorm);
  if (m->stream) (void)hipStreamDestroy(m->stream);
  delete m;
}

const char* bnsamd_mlp_last_error(const bnsamd_mlp* m) { return m ? m->err.c_str() : ""; }

int bnsamd_mlp_param_count(void) { return kParams; }

int bnsamd_mlp_set_params(bnsamd_mlp* m, int member, const float* h_params) {
  if (!m || member < 0 || member >= m->members || !h_params) return BNSAMD_ERR_INVALID;
  (void)hipSetDevice(m->device);
  // the FP32 master of the first layer's low-part rows stays zero
  std::vector<float> hp(h_params, h_params + kParams);
  for (int i = kFeat * kHidden; i < 2 * kFeat * kHidden; ++i) hp[kWOff[0] + i] = 0.f;
  std::vector<uint16_t> hb(kParams);
  for (int i = 0; i < kParams; ++i) hb[i] = f2bf(hp[i]);
  // the first layer's low-part rows: the feature rows' weights
  for (int i = 0; i < kFeat; ++i)
    for (int n = 0; n < kHidden; ++n)
      hb[kWOff[0] + (kFeat + i) * kHidden + n] = hb[kWOff[0] + i * kHidden + n];
  if (hipMemcpy(m->w[member], hp.data(), sizeof(float) * kParams, hipMemcpyHostToDevice) !=
          hipSuccess ||
      hipMemcpy(m->wb[member], hb.data(), sizeof(uint16_t) * kParams, hipMemcpyHostToDevice) !=
          hipSuccess ||
      hipMemset(m->m[member], 0, sizeof(float) * kParams) != hipSuccess ||
      hipMemset(m->v[member], 0, sizeof(float) * kParams) != hipSuccess)
    return mlp_fail(m, BNSAMD_ERR_HIP, "bnsamd_mlp_set_params: copy");
  m->steps[member] = 0;
  m->packed_valid = false;
  return BNSAMD_OK;
}

int bnsamd_mlp_get_params(bnsamd_mlp* m, int member, float* h_params) {
  if (!m || member < 0 || member >= m->members || !h_params) return BNSAMD_ERR_INVALID;
  (void)hipSetDevice(m->device);
  if (sync_all(m) != hipSuccess ||
      hipMemcpy(h_params, m->w[member], sizeof(float) * kParams, hipMemcpyDeviceToHost) !=
          hipSuccess)
    return mlp_fail(m, BNSAMD_ERR_HIP, "bnsamd_mlp_get_params: copy");
  return BNSAMD_OK;
}

int bnsamd_mlp_set_norm(bnsamd_mlp* m, const bnsamd_mlp_norm* norm) {
  if (!m || !norm) return BNSAMD_ERR_INVALID;
  (void)hipSetDevice(m->device);
  // the acceptance coefficients are bnsamd_mlp_set_acceptance's
  bnsamd_mlp_norm h = *norm;
  std::memcpy(h.err_a, m->hnorm.err_a, sizeof(h.err_a));
  std::memcpy(h.err_k, m->hnorm.err_k, sizeof(h.err_k));
  m->hnorm = h;
  if (hipMemcpy(m->norm, &m->hnorm, sizeof(m->hnorm), hipMemcpyHostToDevice) != hipSuccess)
    return mlp_fail(m, BNSAMD_ERR_HIP, "bnsamd_mlp_set_norm: copy");
  return BNSAMD_OK;
}

int bnsamd_mlp_gradients(bnsamd_mlp* m, int member, int rows, const float* d_x, const float* d_y,
                         float* h_grad, float* h_loss) {
  if (!m || member < 0 || member >= m->members || rows < 1 || rows > m->maxb || !d_x || !d_y ||
      !h_grad)
    return BNSAMD_ERR_INVALID;
  (void)hipSetDevice(m->device);
  std::vector<int32_t> ident(rows);
  for (int i = 0; i < rows; ++i) ident[i] = i;
  int32_t* perm = nullptr;
  if (hipMalloc(&perm, sizeof(int32_t) * rows) != hipSuccess)
    return mlp_fail(m, BNSAMD_ERR_NOMEM, "bnsamd_mlp_gradients: hipMalloc");
  hipError_t e = hipMemcpy(perm, ident.data(), sizeof(int32_t) * rows, hipMemcpyHostToDevice);
  Work& W = m->ws[0];
  if (e == hipSuccess) e = sync_all(m);  // no member is still training
  fill(W.stream, 1, W.loss, 0.f);
  hipLaunchKernelGGL(k_gather_batch, dim3((rows + 7) / 8), dim3(32, 8), 0, W.stream, rows, perm,
                     int64_t(0), d_x, d_y, W.Xb, W.Tb);
  forward(m, W, member, rows);
  backward(m, W, member, rows);
  float loss = 0.f;
  if (e == hipSuccess) e = hipStreamSynchronize(W.stream);
  if (e == hipSuccess)
    e = hipMemcpy(h_grad, m->g[member], sizeof(float) * kParams, hipMemcpyDeviceToHost);
  if (e == hipSuccess) e = hipMemcpy(&loss, W.loss, sizeof(float), hipMemcpyDeviceToHost);
  (void)hipFree(perm);
  if (e != hipSuccess) return mlp_fail(m, BNSAMD_ERR_HIP, "bnsamd_mlp_gradients");
  if (h_loss) *h_loss = loss / float(rows * kOut);
  return BNSAMD_OK;
}

namespace {
// one Adam step of member e on workspace W: batch rows of d_x / d_y in the
// order d_perm from `start`
void train_step(bnsamd_mlp* m, Work& W, int e, int batch, int64_t start, const float* d_x,
                const float* d_y, const int32_t* d_perm, float lr) {
  hipLaunchKernelGGL(k_gather_batch, dim3((batch + 7) / 8), dim3(32, 8), 0, W.stream, batch,
                     d_perm, start, d_x, d_y, W.Xb, W.Tb);
  forward(m, W, e, batch);
  backward(m, W, e, batch);
  adam(m, W.stream, e, lr);
}

// cosine decay from lr0 to lr1 over `total` steps
float cosine_lr(float lr0, float lr1, int64_t k, int64_t total) {
  return lr1 + 0.5f * (lr0 - lr1) * (1.f + std::cos(3.14159265f * float(k) / float(total)));
}

// row offset of step st of epoch ep: each epoch reads the permutation from a
// rotating offset
int64_t batch_start(int64_t st, int ep, int batch, int64_t n) {
  return (st * batch + int64_t(ep) * 7919 * batch) % (n - batch + 1);
}
}  // namespace

int bnsamd_mlp_train(bnsamd_mlp* m, int member, int64_t n, const float* d_x, const float* d_y,
                     const int32_t* d_perm, int epochs, int batch, float lr0, float lr1,
                     float* h_epoch_loss) {
  if (!m || member < 0 || member >= m->members || n < batch || batch < 64 || batch > m->maxb ||
      !d_x || !d_y || !d_perm || epochs < 1)
    return BNSAMD_ERR_INVALID;
  (void)hipSetDevice(m->device);
  if (sync_all(m) != hipSuccess) return mlp_fail(m, BNSAMD_ERR_HIP, "bnsamd_mlp_train: sync");
  m->packed_valid = false;
  Work& W = m->ws[0];
  const int64_t steps_per_epoch = n / batch, total = steps_per_epoch * epochs;
  int64_t k = 0;
  for (int ep = 0; ep < epochs; ++ep) {
    fill(W.stream, 1, W.loss, 0.f);
    for (int64_t st = 0; st < steps_per_epoch; ++st, ++k)
      train_step(m, W, member, batch, batch_start(st, ep, batch, n), d_x, d_y, d_perm,
                 cosine_lr(lr0, lr1, k, total));
    float loss = 0.f;
    if (hipMemcpyAsync(&loss, W.loss, sizeof(float), hipMemcpyDeviceToHost, W.stream) !=
            hipSuccess ||
        hipStreamSynchronize(W.stream) != hipSuccess)
      return mlp_fail(m, BNSAMD_ERR_HIP, "bnsamd_mlp_train: loss readback");
    if (h_epoch_loss) h_epoch_loss[ep] = loss / float(steps_per_epoch * batch * kOut);
  }
  hipError_t e = hipStreamSynchronize(W.stream);
  return e == hipSuccess ? BNSAMD_OK : mlp_fail(m, BNSAMD_ERR_HIP, "bnsamd_mlp_train");
}

int bnsamd_mlp_train_ensemble(bnsamd_mlp* m, int64_t n, const float* d_x, const float* d_y,
                              const int32_t* d_perm, int epochs, int batch, float lr0, float lr1,
                              float* h_epoch_loss) {
  if (!m || n < batch || batch < 64 || batch > m->maxb || !d_x || !d_y || !d_perm || epochs < 1)
    return BNSAMD_ERR_INVALID;
  (void)hipSetDevice(m->device);
  if (sync_all(m) != hipSuccess)
    return mlp_fail(m, BNSAMD_ERR_HIP, "bnsamd_mlp_train_ensemble: sync");
  // one workspace and stream per member (default priority)
  while (int(m->ws.size()) < m->members) {
    Work W;
    bool ok = hipStreamCreateWithFlags(&W.stream, hipStreamNonBlocking) == hipSuccess;
    W.own_stream = ok;
    ok = ok && work_alloc(W, m->maxb);
    if (!ok) {
      work_free(W);
      return mlp_fail(m, BNSAMD_ERR_NOMEM, "bnsamd_mlp_train_ensemble: workspace");
    }
    m->ws.push_back(W);
  }
  m->packed_valid = false;
  const int M = m->members;
  const int64_t steps_per_epoch = n / batch, total = steps_per_epoch * epochs;
  int64_t k = 0;
  for (int ep = 0; ep < epochs; ++ep) {
    for (int e = 0; e < M; ++e) fill(m->ws[e].stream, 1, m->ws[e].loss, 0.f);
    // the members' steps interleaved, each on its own stream: the small
    // GEMMs of one member leave most of the chip idle
    for (int64_t st = 0; st < steps_per_epoch; ++st, ++k) {
      const float lr = cosine_lr(lr0, lr1, k, total);
      const int64_t start = batch_start(st, ep, batch, n);
      for (int e = 0; e < M; ++e)
        train_step(m, m->ws[e], e, batch, start, d_x, d_y, d_perm + int64_t(e) * n, lr);
    }
    for (int e = 0; e < M; ++e) {
      float loss = 0.f;
      if (hipMemcpyAsync(&loss, m->ws[e].loss, sizeof(float), hipMemcpyDeviceToHost,
                         m->ws[e].stream) != hipSuccess ||
          hipStreamSynchronize(m->ws[e].stream) != hipSuccess)
        return mlp_fail(m, BNSAMD_ERR_HIP, "bnsamd_mlp_train_ensemble: loss readback");
      if (h_epoch_loss)
        h_epoch_loss[int64_t(e) * epochs + ep] = loss / float(steps_per_epoch * batch * kOut);
    }
  }
  hipError_t e = sync_all(m);
  return e == hipSuccess ? BNSAMD_OK : mlp_fail(m, BNSAMD_ERR_HIP, "bnsamd_mlp_train_ensemble");
}

int bnsamd_mlp_predict(bnsamd_mlp* m, int member, int rows, const float* d_x, float* d_y) {
  if (!m || member < 0 || member >= m->members || rows < 1 || !d_x || !d_y)
    return BNSAMD_ERR_INVALID;
  (void)hipSetDevice(m->device);
  std::vector<int32_t> ident(rows);
  for (int i = 0; i < rows; ++i) ident[i] = i;
  int32_t* perm = nullptr;
  if (hipMalloc(&perm, sizeof(int32_t) * rows) != hipSuccess)
    return mlp_fail(m, BNSAMD_ERR_NOMEM, "bnsamd_mlp_predict: hipMalloc");
  hipError_t e = hipMemcpy(perm, ident.data(), sizeof(int32_t) * rows, hipMemcpyHostToDevice);
  Work& W = m->ws[0];
  if (e == hipSuccess) e = sync_all(m);  // no member is still training
  for (int r0 = 0; r0 < rows && e == hipSuccess; r0 += m->maxb) {  // batches of maxb rows
    const int b = rows - r0 < m->maxb ? rows - r0 : m->maxb;
    hipLaunchKernelGGL(k_gather_batch, dim3((b + 7) / 8), dim3(32, 8), 0, W.stream, b, perm,
                       int64_t(r0), d_x, static_cast<const float*>(nullptr), W.Xb, W.Tb);
    forward(m, W, member, b);
    e = hipMemcpy2DAsync(d_y + int64_t(r0) * kOut, sizeof(float) * kOut, W.Z[kLayers - 1],
                         sizeof(float) * kNL, sizeof(float) * kOut, b, hipMemcpyDeviceToDevice,
                         W.stream);
  }
  if (e == hipSuccess) e = hipStreamSynchronize(W.stream);
  (void)hipFree(perm);
  return e == hipSuccess ? BNSAMD_OK : mlp_fail(m, BNSAMD_ERR_HIP, "bnsamd_mlp_predict");
}

int bnsamd_mlp_set_domain(bnsamd_mlp* m, const bnsamd_mlp_domain* d) {
  if (!m || !d) return BNSAMD_ERR_INVALID;
  if (d->nquad < 1 || d->nquad > 32 || (d->stim_abs != 0 && d->stim_abs != 1) ||
      !(d->beta_guard >= 0.))
    return mlp_fail(m, BNSAMD_ERR_INVALID, "bnsamd_mlp_set_domain: invalid domain");
  m->domain = *d;
  m->domain_set = true;
  return BNSAMD_OK;
}

int bnsamd_mlp_set_acceptance(bnsamd_mlp* m, const float* err_a, const float* err_k) {
  if (!m || !err_a || !err_k) return BNSAMD_ERR_INVALID;
  for (int o = 0; o < kOut; ++o)
    if (!(err_a[o] >= 0.f) || !(err_k[o] >= 0.f))
      return mlp_fail(m, BNSAMD_ERR_INVALID, "bnsamd_mlp_set_acceptance: negative coefficient");
  (void)hipSetDevice(m->device);
  std::memcpy(m->hnorm.err_a, err_a, sizeof(float) * kOut);
  std::memcpy(m->hnorm.err_k, err_k, sizeof(float) * kOut);
  // the norm block may be read by an inference still queued on a stream
  if (hipDeviceSynchronize() != hipSuccess ||
      hipMemcpy(m->norm, &m->hnorm, sizeof(m->hnorm), hipMemcpyHostToDevice) != hipSuccess)
    return mlp_fail(m, BNSAMD_ERR_HIP, "bnsamd_mlp_set_acceptance: copy");
  return BNSAMD_OK;
}

int bnsamd_mlp_set_judge(bnsamd_mlp* m, bnsamd_mlp* judge, double threshold) {
  if (!m || judge == m) return BNSAMD_ERR_INVALID;
  (void)hipSetDevice(m->device);
  // an inference still queued may read the judge block
  if (hipDeviceSynchronize() != hipSuccess)
    return mlp_fail(m, BNSAMD_ERR_HIP, "bnsamd_mlp_set_judge: synchronize");
  if (!judge) {
    m->judge_set = false;
    return BNSAMD_OK;
  }
  if (judge->device != m->device)
    return mlp_fail(m, BNSAMD_ERR_INVALID, "bnsamd_mlp_set_judge: judge on another device");
  if (std::isnan(threshold))
    return mlp_fail(m, BNSAMD_ERR_INVALID, "bnsamd_mlp_set_judge: NaN threshold");
  // the judge reads the features the ensemble standardises: same feature
  // statistics and training box
  const bnsamd_mlp_norm &a = m->hnorm, &b = judge->hnorm;
  if (std::memcmp(a.feat_mean, b.feat_mean, sizeof(a.feat_mean)) != 0 ||
      std::memcmp(a.feat_std, b.feat_std, sizeof(a.feat_std)) != 0 ||
      std::memcmp(a.box_lo, b.box_lo, sizeof(a.box_lo)) != 0 ||
      std::memcmp(a.box_hi, b.box_hi, sizeof(a.box_hi)) != 0)
    return mlp_fail(m, BNSAMD_ERR_INVALID,
                    "bnsamd_mlp_set_judge: feature standardisation differs from the ensemble's");
  if (!m->jw) {
    // first judge of this ensemble: on failure nothing is kept (no judge)
    if (hipMalloc(&m->jw, sizeof(float) * kParams) != hipSuccess ||
        hipMalloc(&m->jwp, sizeof(uint16_t) * kPacked) != hipSuccess ||
        hipMalloc(&m->jnorm, sizeof(bnsamd_mlp_norm)) != hipSuccess) {
      (void)hipFree(m->jw);
      (void)hipFree(m->jwp);
      (void)hipFree(m->jnorm);
      m->jw = nullptr;
      m->jwp = nullptr;
      m->jnorm = nullptr;
      m->judge_set = false;
      return mlp_fail(m, BNSAMD_ERR_NOMEM, "bnsamd_mlp_set_judge: hipMalloc");
    }
  }
  // while the blocks are rewritten the judge rejects every point: a copy
  // that fails leaves that judge, never a partly copied one
  m->jthreshold = -std::numeric_limits<double>::infinity();
  m->judge_set = true;
  // the judge's training has finished; copy its member 0 and pack it
  hipError_t e = sync_all(judge);
  if (e == hipSuccess)
    e = hipMemcpy(m->jw, judge->w[0], sizeof(float) * kParams, hipMemcpyDeviceToDevice);
  if (e == hipSuccess)
    e = hipMemcpy(m->jnorm, &judge->hnorm, sizeof(bnsamd_mlp_norm), hipMemcpyHostToDevice);
  for (int l = 0; l < kLayers && e == hipSuccess; ++l) {
    int64_t n = int64_t(kLayerK[l]) * kLayerN[l];
    hipLaunchKernelGGL(k_pack, dim3(unsigned((n + 255) / 256)), dim3(256), 0, nullptr, kLayerK[l],
                       kLayerN[l], m->jw + kWOff[l], m->jwp + kPOff[l], l == 0 ? 1 : 0);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipDeviceSynchronize();
  if (e != hipSuccess) return mlp_fail(m, BNSAMD_ERR_HIP, "bnsamd_mlp_set_judge: copy");
  m->jthreshold = threshold;
  m->judge_set = true;
  return BNSAMD_OK;
}

namespace {
// inference arguments shared by the learned form and the raw predictions;
// packs the weights when training changed them
int infer_setup(bnsamd_mlp* m, hipStream_t s, int64_t npts, int64_t ld, const double* d_pts,
                InferArgs& a, const char* who) {
  // training runs on the MLP's own streams: parameters are final before use
  if (sync_all(m) != hipSuccess) return mlp_fail(m, BNSAMD_ERR_HIP, std::string(who) + ": training stream");
  if (!m->packed_valid) {
    for (int e = 0; e < m->members; ++e)
      for (int l = 0; l < kLayers; ++l) {
        int64_t n = int64_t(kLayerK[l]) * kLayerN[l];
        hipLaunchKernelGGL(k_pack, dim3(unsigned((n + 255) / 256)), dim3(256), 0, s, kLayerK[l],
                           kLayerN[l], m->w[e] + kWOff[l], m->wp[e] + kPOff[l], l == 0 ? 1 : 0);
      }
    m->packed_valid = true;
  }
  a = InferArgs{};
  a.npts = npts;
  a.ld = ld;
  a.pts = d_pts;
  a.norm = reinterpret_cast<const bnsamd_mlp_norm*>(m->norm);
  for (int e = 0; e < m->members; ++e) {
    a.wp[e] = m->wp[e];
    a.w[e] = m->w[e];
  }
  a.members = m->members;
  if (m->judge_set) {
    a.jwp = m->jwp;
    a.jw = m->jw;
    a.jnorm = reinterpret_cast<const bnsamd_mlp_norm*>(m->jnorm);
    a.jthreshold = m->jthreshold;
  }
  a.dU = m->domain.dU;
  a.dm_eff = m->domain.dm_eff;
  const bnsamd_params& pr = m->domain.params;
  a.mu_np_shift = (pr.use_dm_eff ? m->domain.dm_eff : bns::c::Q) + (pr.use_dU ? m->domain.dU : 0.);
  a.beta_guard = m->domain.beta_guard;
  return BNSAMD_OK;
}

bool same_flags(const bnsamd_flags& x, const bnsamd_flags& y) {
  return x.use_abs_em == y.use_abs_em && x.use_pair == y.use_pair && x.use_brem == y.use_brem &&
         x.use_inelastic_scatt == y.use_inelastic_scatt && x.use_iso == y.use_iso;
}
bool same_params(const bnsamd_params& x, const bnsamd_params& y) {
  return x.use_dU == y.use_dU && x.use_dm_eff == y.use_dm_eff && x.use_WM_ab == y.use_WM_ab &&
         x.use_WM_sc == y.use_WM_sc && x.use_decay == y.use_decay &&
         x.brem_implementation == y.brem_implementation &&
         x.use_NN_medium_corr == y.use_NN_medium_corr && x.neglect_blocking == y.neglect_blocking;
}
}  // namespace

int bnsamd_mlp_infer(bnsamd_context* ctx, bnsamd_mlp* m, int64_t npts, int64_t ld,
                     const double* d_pts, double* d_logmean, float* d_spread, int32_t* d_outside) {
  if (!ctx || !m || npts < 0 || ld < npts || !d_pts || !d_logmean || !d_spread || !d_outside ||
      npts > int64_t(0x7fffffff))
    return m ? mlp_fail(m, BNSAMD_ERR_INVALID, "bnsamd_mlp_infer: invalid argument")
             : BNSAMD_ERR_INVALID;
  if (!m->domain_set)
    return mlp_fail(m, BNSAMD_ERR_INVALID, "bnsamd_mlp_infer: training domain not set");
  if (npts == 0) return BNSAMD_OK;
  (void)hipSetDevice(m->device);
  hipStream_t s = static_cast<hipStream_t>(bnsamd_get_stream(ctx));
  InferArgs a;
  if (int rc = infer_setup(m, s, npts, ld, d_pts, a, "bnsamd_mlp_infer")) return rc;
  a.threshold = 0.f;
  a.cal_mean = d_logmean;
  a.cal_spread = d_spread;
  a.cal_outside = d_outside;
  hipLaunchKernelGGL(k_infer<true>, dim3(unsigned((npts + kRows - 1) / kRows)), dim3(256), 0, s, a);
  hipError_t e = hipGetLastError();
  return e == hipSuccess ? BNSAMD_OK : mlp_fail(m, BNSAMD_ERR_HIP, "bnsamd_mlp_infer: launch");
}

int bnsamd_surrogate_m1_opacities(bnsamd_context* ctx, bnsamd_mlp* m, int64_t npts, int64_t ld,
                                  const double* d_pts, float error_bound,
                                  const bnsamd_flags* flags, const bnsamd_params* params,
                                  int stim_abs, double* d_out, int64_t* n_fallback) {
  if (!ctx || !m || npts < 0 || ld < npts || !d_pts || !d_out || !flags || !params ||
      npts > int64_t(0x7fffffff))
    return m ? mlp_fail(m, BNSAMD_ERR_INVALID, "bnsamd_surrogate_m1_opacities: invalid argument")
             : BNSAMD_ERR_INVALID;
  if (!m->domain_set)
    return mlp_fail(m, BNSAMD_ERR_INVALID,
                    "bnsamd_surrogate_m1_opacities: training domain not set");
  if (npts == 0) {
    if (n_fallback) *n_fallback = 0;
    return BNSAMD_OK;
  }
  (void)hipSetDevice(m->device);
  hipStream_t s = static_cast<hipStream_t>(bnsamd_get_stream(ctx));
  // a call outside the configuration the targets were made with: every
  // point through the FP64 path
  int n1 = 0, n2 = 0;
  if (!bns::context_quadrature_orders(ctx, &n1, &n2))
    return mlp_fail(m, BNSAMD_ERR_NO_QUADRATURE, "bnsamd_set_quadrature not called");
  const bnsamd_mlp_domain& dm = m->domain;
  if (!same_flags(*flags, dm.flags) || !same_params(*params, dm.params) ||
      stim_abs != dm.stim_abs || n1 != dm.nquad || n2 != dm.nquad) {
    int rc = bnsamd_m1_opacities(ctx, npts, ld, d_pts, nullptr, flags, params, stim_abs,
                                 BNSAMD_FP64, d_out);
    if (rc != BNSAMD_OK) return mlp_fail(m, rc, std::string("fp64 path: ") + bnsamd_last_error(ctx));
    if (hipStreamSynchronize(s) != hipSuccess)
      return mlp_fail(m, BNSAMD_ERR_HIP, "bnsamd_surrogate_m1_opacities: fp64 path");
    if (n_fallback) *n_fallback = npts;
    return BNSAMD_OK;
  }
  if (npts > m->fb_cap) {
    auto fr = [](void* p) {
      if (p) (void)hipFree(p);
    };
    // the previous call's gather / scatter have finished (the call is synchronous)
    fr(m->fb_idx);
    fr(m->fb_count);
    fr(m->fb_pts);
    fr(m->fb_out);
    m->fb_idx = nullptr;
    m->fb_count = nullptr;
    m->fb_pts = m->fb_out = nullptr;
    m->fb_cap = 0;
    if (hipMalloc(&m->fb_idx, sizeof(int32_t) * npts) != hipSuccess ||
        hipMalloc(&m->fb_count, sizeof(int32_t)) != hipSuccess ||
        hipMalloc(&m->fb_pts, sizeof(double) * BNSAMD_PT_NFIELDS * npts) != hipSuccess ||
        hipMalloc(&m->fb_out, sizeof(double) * BNSAMD_OUT_NFIELDS * npts) != hipSuccess)
      return mlp_fail(m, BNSAMD_ERR_NOMEM, "bnsamd_surrogate_m1_opacities: workspace");
    m->fb_cap = npts;
  }
  InferArgs a;
  if (int rc = infer_setup(m, s, npts, ld, d_pts, a, "bnsamd_surrogate_m1_opacities")) return rc;
  a.threshold = error_bound;
  a.out = d_out;
  a.fallback_idx = m->fb_idx;
  a.fallback_count = m->fb_count;
  hipError_t e = hipMemsetAsync(m->fb_count, 0, sizeof(int32_t), s);
  hipLaunchKernelGGL(k_infer<false>, dim3(unsigned((npts + kRows - 1) / kRows)), dim3(256), 0, s,
                     a);
  int32_t nfb = 0;
  if (e == hipSuccess) e = hipGetLastError();
  if (e == hipSuccess)
    e = hipMemcpyAsync(&nfb, m->fb_count, sizeof(int32_t), hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  if (e != hipSuccess) return mlp_fail(m, BNSAMD_ERR_HIP, "bnsamd_surrogate_m1_opacities: infer");
  if (nfb > 0) {
    unsigned g = unsigned((nfb + 255) / 256);
    hipLaunchKernelGGL(k_gather_pts, dim3(g), dim3(256), 0, s, nfb, m->fb_idx, d_pts, ld,
                       m->fb_pts);
    int rc = bnsamd_m1_opacities(ctx, nfb, nfb, m->fb_pts, nullptr, flags, params, stim_abs,
                                 BNSAMD_FP64, m->fb_out);
    if (rc != BNSAMD_OK)
      return mlp_fail(m, rc, std::string("fallback: ") + bnsamd_last_error(ctx));
    hipLaunchKernelGGL(k_scatter_out, dim3(g), dim3(256), 0, s, nfb, m->fb_idx, m->fb_out, d_out,
                       ld);
    e = hipGetLastError();
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess)
      return mlp_fail(m, BNSAMD_ERR_HIP, "bnsamd_surrogate_m1_opacities: fallback");
  }
  if (n_fallback) *n_fallback = nfb;
  return BNSAMD_OK;
}

}  // extern "C"
