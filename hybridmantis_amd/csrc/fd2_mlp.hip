// fd2_mlp.hip -- the surrogate network: 5 -> 256 -> 256 -> 256 -> 65 MLP
// with GELU, BF16 weights/activations on the matrix cores (FP32 accumulate,
// FP32 master weights), trained with Adam by the kernels below.
//
// Layout (row-major everywhere, feature dims padded to multiples of 16):
//   X0 [B x 32] bf16 (5 features + zeros)
//   W0 [32 x 256], W1, W2 [256 x 256], W3 [256 x 80] (65 outputs + zeros)
//   Zl = A(l-1) Wl + bl (FP32, kept for backward), Al = bf16(gelu(Zl))
//   output logits Z4 [B x 80]: column 0 detection logit, 1..64 radial logits
// Loss per sample: BCE(sigmoid(z0), y0) + w * CE(softmax(z1..64), q) with
// y0 the detected fraction, q the radial distribution and w = 1 when the
// sample detected photons (else 0); the batch loss is the mean.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <chrono>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <algorithm>
#include <new>
#include <thread>
#include <vector>

#include "../../include/fd2.h"
#include "fd2_internal.h"
#include "fd2_mlp_internal.h"

namespace fd2mlp {

typedef short short4_ __attribute__((ext_vector_type(4)));
typedef float float4_ __attribute__((ext_vector_type(4)));

__device__ __forceinline__ uint16_t f2bf(float f) {
  uint32_t u = __float_as_uint(f);
  u += 0x7FFFu + ((u >> 16) & 1u);  // round to nearest even
  return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float bf2f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

// GELU, tanh form: 0.5 x (1 + tanh(c (x + 0.044715 x^3))), c = sqrt(2/pi)
__device__ __forceinline__ float gelu_f(float x) {
  const float c = 0.7978845608f;
  const float t = tanhf(c * (x + 0.044715f * x * x * x));
  return 0.5f * x * (1.0f + t);
}
__device__ __forceinline__ float gelu_grad(float x) {
  const float c = 0.7978845608f;
  const float u = c * (x + 0.044715f * x * x * x);
  const float t = tanhf(u);
  const float du = c * (1.0f + 3.0f * 0.044715f * x * x);
  return 0.5f * (1.0f + t) + 0.5f * x * (1.0f - t * t) * du;
}

// ------------------------------------------------------------- GEMM
//
// C[M x N] = op(A)[M x K] * op(B)[K x N], K a multiple of 16.
// TA: A is stored transposed ([K x M]); TB: B stored transposed ([N x K]).
// 256 threads = 4 waves compute a 64 x 64 tile; each wave a 32 x 32 quarter
// as 2 x 2 blocks of v_mfma_f32_16x16x16bf16_1k.  Operand tiles of 64 x 16
// go through LDS so every layout is read with coalesced 8-byte loads.
// Split-K: blockIdx.z takes K range [z*kchunk, (z+1)*kchunk).
enum Epilogue : int {
  kStoreF32 = 0,      // C (+bias) -> outf; gelu(C) as bf16 -> outb if set
  kAtomicF32 = 1,     // outf += C (split-K weight gradients)
  kGeluGradBf16 = 2,  // outb = bf16(C * gelu'(Z)) and bias grad += colsum
};

struct GemmArgs {
  const uint16_t* A;
  const uint16_t* B;
  int M, N, K;
  int lda, ldb;        // leading dims of the stored arrays
  int kchunk;
  const float* bias;   // kStoreF32
  float* outf;         // kStoreF32 / kAtomicF32 (ld = N)
  uint16_t* outb;      // kStoreF32 (gelu) / kGeluGradBf16 (ld = ldo)
  const float* Z;      // kGeluGradBf16: pre-activations (ld = N)
  float* dbias;        // kGeluGradBf16: column sums of the product
  int ldo;
};

template <bool TA, bool TB, int EPI>
__global__ void __launch_bounds__(256) k_gemm(GemmArgs g) {
  __shared__ uint16_t As[64][16 + 4];
  __shared__ uint16_t Bs[16][64 + 4];
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  const int m0 = blockIdx.y * 64, n0 = blockIdx.x * 64;
  const int wm = (w >> 1) * 32, wn = (w & 1) * 32;
  const int kbeg = blockIdx.z * g.kchunk;
  const int kend = min(g.K, kbeg + g.kchunk);
  float4_ acc[2][2];
  for (int i = 0; i < 2; ++i)
    for (int j = 0; j < 2; ++j) acc[i][j] = float4_{0.f, 0.f, 0.f, 0.f};

  for (int k0 = kbeg; k0 < kend; k0 += 16) {
    // A tile: 64 rows x 16 k
    if (!TA) {
      const int r = t >> 2, kk = (t & 3) * 4;
      const int gm = m0 + r;
      uint16_t v[4] = {0, 0, 0, 0};
      if (gm < g.M) {
        const uint2 q = *reinterpret_cast<const uint2*>(g.A + (size_t)gm * g.lda + k0 + kk);
        v[0] = q.x & 0xFFFF; v[1] = q.x >> 16; v[2] = q.y & 0xFFFF; v[3] = q.y >> 16;
      }
      for (int j = 0; j < 4; ++j) As[r][kk + j] = v[j];
    } else {
      const int kk = t >> 4, r = (t & 15) * 4;
      const int gm = m0 + r;
      const uint16_t* src = g.A + (size_t)(k0 + kk) * g.lda + gm;
      for (int j = 0; j < 4; ++j) As[r + j][kk] = (gm + j < g.M) ? src[j] : (uint16_t)0;
    }
    // B tile: 16 k x 64 cols
    if (!TB) {
      const int kk = t >> 4, c = (t & 15) * 4;
      const int gn = n0 + c;
      const uint16_t* src = g.B + (size_t)(k0 + kk) * g.ldb + gn;
      for (int j = 0; j < 4; ++j) Bs[kk][c + j] = (gn + j < g.N) ? src[j] : (uint16_t)0;
    } else {
      const int c = t >> 2, kk = (t & 3) * 4;
      const int gn = n0 + c;
      uint16_t v[4] = {0, 0, 0, 0};
      if (gn < g.N) {
        const uint2 q = *reinterpret_cast<const uint2*>(g.B + (size_t)gn * g.ldb + k0 + kk);
        v[0] = q.x & 0xFFFF; v[1] = q.x >> 16; v[2] = q.y & 0xFFFF; v[3] = q.y >> 16;
      }
      for (int j = 0; j < 4; ++j) Bs[kk + j][c] = v[j];
    }
    __syncthreads();
    // fragments: A lane l -> row l%16, k 4*(l/16)..+3; B lane l -> col l%16
    short4_ af[2], bfr[2];
    const int lr = lane & 15, lk = (lane >> 4) * 4;
    for (int i = 0; i < 2; ++i) {
      const uint16_t* ap = &As[wm + i * 16 + lr][lk];
      af[i] = short4_{(short)ap[0], (short)ap[1], (short)ap[2], (short)ap[3]};
      bfr[i] = short4_{(short)Bs[lk][wn + i * 16 + lr], (short)Bs[lk + 1][wn + i * 16 + lr],
                       (short)Bs[lk + 2][wn + i * 16 + lr], (short)Bs[lk + 3][wn + i * 16 + lr]};
    }
    for (int i = 0; i < 2; ++i)
      for (int j = 0; j < 2; ++j)
        acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(af[i], bfr[j], acc[i][j], 0, 0, 0);
    __syncthreads();
  }
  // D lane l: col l%16, rows 4*(l/16)+r
  const int lc = lane & 15, lr4 = (lane >> 4) * 4;
  for (int i = 0; i < 2; ++i) {
    for (int j = 0; j < 2; ++j) {
      const int col = n0 + wn + j * 16 + lc;
      float colsum = 0.0f;
      for (int r = 0; r < 4; ++r) {
        const int row = m0 + wm + i * 16 + lr4 + r;
        const float v = acc[i][j][r];
        if (row >= g.M || col >= g.N) continue;
        if (EPI == kStoreF32) {
          const float z = v + (g.bias ? g.bias[col] : 0.0f);
          g.outf[(size_t)row * g.N + col] = z;
          if (g.outb) g.outb[(size_t)row * g.ldo + col] = f2bf(gelu_f(z));
        } else if (EPI == kAtomicF32) {
          atomicAdd(&g.outf[(size_t)row * g.N + col], v);
        } else {
          const float d = v * gelu_grad(g.Z[(size_t)row * g.N + col]);
          const uint16_t h = f2bf(d);
          g.outb[(size_t)row * g.ldo + col] = h;
          colsum += bf2f(h);
        }
      }
      if (EPI == kGeluGradBf16 && g.dbias) {
        // lanes l, l+16, l+32, l+48 share the column
        colsum += __shfl_xor(colsum, 16, 64);
        colsum += __shfl_xor(colsum, 32, 64);
        if ((lane >> 4) == 0 && col < g.N) atomicAdd(&g.dbias[col], colsum);
      }
    }
  }
}

template <bool TA, bool TB, int EPI>
static hipError_t gemm(const GemmArgs& a, int splits, hipStream_t s) {
  dim3 grid((a.N + 63) / 64, (a.M + 63) / 64, splits);
  hipLaunchKernelGGL((k_gemm<TA, TB, EPI>), grid, dim3(256), 0, s, a);
  return hipGetLastError();
}

// ----------------------------------------------------- elementwise kernels

// features (FP32 [n x 5], host-normalised) -> X0 bf16 [n x 32]
__global__ void k_pack_inputs(const float* x, int n, uint16_t* x0) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n * kIn) return;
  const int row = i / kIn, c = i % kIn;
  x0[(size_t)row * kInPad + c] = c < kFeat ? f2bf(x[(size_t)row * kFeat + c]) : (uint16_t)0;
}

// gather a minibatch: rows idx[b] of the resident data set, one wave per
// row (lanes over the row's columns, so the copies are coalesced)
__global__ void k_gather(const float* xs, const float* ys, const uint32_t* idx, int B,
                         uint16_t* x0, float* y) {
  const int i = (blockIdx.x * blockDim.x + threadIdx.x) >> 6, lane = threadIdx.x & 63;
  if (i >= B) return;
  const uint32_t r = idx[i];
  if (lane < kInPad)
    x0[(size_t)i * kInPad + lane] = lane < kFeat ? f2bf(xs[(size_t)r * kFeat + lane]) : (uint16_t)0;
  for (int c = lane; c < kTarget; c += 64) y[(size_t)i * kTarget + c] = ys[(size_t)r * kTarget + c];
}

// loss and dL/dZ4, one sample at a time per wave: lane k handles radial
// logit k (z[1+k]); lane 0 also the detection logit.  dz -> bf16 [B x 80].
// The column 0 and 1..64 bias gradients and the loss are summed over the
// wave's samples in registers, then over the block's waves in LDS, and
// added with one atomic per column and block: one atomic per sample and
// column made 4096 x 65 float atomics meet on 65 addresses (142 us per
// batch of 4096); 32 blocks of long per-wave loops took 69 us.
constexpr int kHeadLossBlocks = 256;
__global__ void __launch_bounds__(256) k_head_loss(const float* z4, const float* y, int B,
                                                   uint16_t* dz, float* db3, float* loss) {
  __shared__ float s_db[4][kOut];
  __shared__ float s_loss[4];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int nw = gridDim.x * (blockDim.x >> 6);
  const float invB = 1.0f / (float)B;
  float db_r = 0.0f, db_0 = 0.0f, lsum = 0.0f;
  for (int wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6; wave < B; wave += nw) {
    const float* zr = z4 + (size_t)wave * kOutPad;
    const float* yr = y + (size_t)wave * kTarget;
    // radial softmax over 64 lanes
    const float zl = zr[1 + lane];
    float mx = zl;
    for (int o = 32; o > 0; o >>= 1) mx = fmaxf(mx, __shfl_xor(mx, o, 64));
    const float e = expf(zl - mx);
    float se = e;
    for (int o = 32; o > 0; o >>= 1) se += __shfl_xor(se, o, 64);
    const float prob = e / se;
    const float q = yr[1 + lane];
    float qs = q;
    for (int o = 32; o > 0; o >>= 1) qs += __shfl_xor(qs, o, 64);
    const float wrad = qs > 0.0f ? 1.0f : 0.0f;
    float ce = q > 0.0f ? -q * (zl - mx - logf(se)) : 0.0f;
    for (int o = 32; o > 0; o >>= 1) ce += __shfl_xor(ce, o, 64);
    const float gr = wrad * (prob - q) * invB;
    uint16_t* dr = dz + (size_t)wave * kOutPad;
    const uint16_t grb = f2bf(gr);
    dr[1 + lane] = grb;
    db_r += bf2f(grb);
    if (lane < kOutPad - 1 - 64) dr[65 + lane] = 0;
    if (lane == 0) {
      const float z0 = zr[0], y0 = yr[0];
      const float p = 1.0f / (1.0f + expf(-z0));
      // BCE with logits: max(z,0) - z y + log(1 + e^-|z|)
      const float bce = fmaxf(z0, 0.0f) - z0 * y0 + log1pf(expf(-fabsf(z0)));
      const uint16_t g0 = f2bf((p - y0) * invB);
      dr[0] = g0;
      db_0 += bf2f(g0);
      lsum += (bce + wrad * ce) * invB;
    }
  }
  s_db[w][1 + lane] = db_r;
  if (lane == 0) {
    s_db[w][0] = db_0;
    s_loss[w] = lsum;
  }
  __syncthreads();
  const int t = threadIdx.x;
  if (t < kOut) atomicAdd(&db3[t], (s_db[0][t] + s_db[1][t]) + (s_db[2][t] + s_db[3][t]));
  if (t == 0) atomicAdd(loss, (s_loss[0] + s_loss[1]) + (s_loss[2] + s_loss[3]));
}

// logits -> [p_det, radial probabilities] (FP32 [n x 65])
__global__ void __launch_bounds__(256) k_head_probs(const float* z4, int n, float* out) {
  const int wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const int lane = threadIdx.x & 63;
  if (wave >= n) return;
  const float* zr = z4 + (size_t)wave * kOutPad;
  const float zl = zr[1 + lane];
  float mx = zl;
  for (int o = 32; o > 0; o >>= 1) mx = fmaxf(mx, __shfl_xor(mx, o, 64));
  const float e = expf(zl - mx);
  float se = e;
  for (int o = 32; o > 0; o >>= 1) se += __shfl_xor(se, o, 64);
  float* orow = out + (size_t)wave * kTarget;
  orow[1 + lane] = e / se;
  if (lane == 0) orow[0] = 1.0f / (1.0f + expf(-zr[0]));
}

// Adam (Kingma & Ba 2015) on FP32 master weights; refreshes the bf16 copy
// of weight matrices and clears the gradient.
__global__ void k_adam(float* w, float* g, float* m, float* v, uint16_t* wb, int n,
                       float lr, float b1, float b2, float eps, float c1, float c2) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i];
  const float mi = b1 * m[i] + (1.0f - b1) * gi;
  const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  const float wi = w[i] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
  w[i] = wi;
  g[i] = 0.0f;
  if (wb) wb[i] = f2bf(wi);
}

__global__ void k_to_bf16(const float* w, uint16_t* wb, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) wb[i] = f2bf(w[i]);
}

__global__ void k_zero(float* p, size_t n) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x)
    p[i] = 0.0f;
}

// ------------------------------------------------- fused inference forward

// Weight matrices in MFMA fragment order for the fused forward: layer l's
// W [K x N] (row-major at kWOff[l] in the bf16 copy) is stored at the same
// offset as the B fragments of v_mfma_f32_16x16x16bf16_1k, column block cb
// (16 columns) by k step ks (16 k) by lane by element: lane 16 q + c holds
// W[16 ks + 4 q + e][16 cb + c], e = 0..3.  One fragment is then 512
// contiguous bytes, a lane's part one 8-byte load.  Bias slots are unused.
__global__ void k_fragment_weights(const uint16_t* __restrict__ wb, uint16_t* __restrict__ wt) {
  const int ks[4] = {kInPad, kHidden, kHidden, kHidden};
  const int ns[4] = {kHidden, kHidden, kHidden, kOutPad};
  const int l = blockIdx.y;
  const int K = ks[l], N = ns[l];
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < K * N; i += gridDim.x * blockDim.x) {
    const int k = i / N, n = i - k * N;
    const int lane = 16 * ((k & 15) >> 2) + (n & 15);
    const size_t frag = (size_t)(n >> 4) * (K >> 4) + (k >> 4);
    wt[kWOff[l] + (frag * 64 + lane) * 4 + (k & 3)] = wb[kWOff[l] + i];
  }
}

// The inference forward with all four layers in one kernel: a workgroup
// of 4 waves takes kFuseRows = 64 rows, keeps their activations in LDS
// (bf16, 64 x 256) from layer to layer and writes only the output logits
// (FP32, B x kOutPad), where the layered forward writes every layer's FP32
// pre-activations and bf16 activations to HBM and reads them back.
// Hidden layers: wave w computes the 64 rows x columns [64w, 64w + 64) as
// 4 x 4 tiles of v_mfma_f32_16x16x16bf16_1k, A fragments from LDS and B
// fragments straight from the fragment-ordered weights (312 KB in all, read by
// every workgroup, so they stay in L2); after a barrier the epilogue
// (bias, GELU, bf16) overwrites the activations in place.  Output layer:
// wave w takes rows [16w, 16w + 16) and the 5 column tiles.  Fragments,
// the k order of the accumulation and the epilogue arithmetic are those of
// k_gemm, so the logits equal the layered forward's bit for bit.
constexpr int kFuseRows = 64;
constexpr int kActLd = kHidden + 8;  // LDS row stride (bf16): 528 B

__global__ void __launch_bounds__(256) k_forward_fused(const uint16_t* __restrict__ x0, int n,
                                                       const uint16_t* __restrict__ wt,
                                                       const float* __restrict__ w,
                                                       float* __restrict__ out) {
  __shared__ __attribute__((aligned(16))) uint16_t act[kFuseRows][kActLd];
  const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
  const int row0 = blockIdx.x * kFuseRows;
  // input rows (32 bf16 each) -> act[r][0 .. 31]; rows past n are zero
  for (int i = t; i < kFuseRows * (kInPad / 4); i += 256) {
    const int r = i / (kInPad / 4), c = (i % (kInPad / 4)) * 4;
    uint2 v = make_uint2(0u, 0u);
    if (row0 + r < n) v = *reinterpret_cast<const uint2*>(x0 + (size_t)(row0 + r) * kInPad + c);
    *reinterpret_cast<uint2*>(&act[r][c]) = v;
  }
  __syncthreads();
  const int lr = lane & 15, lq = lane >> 4;
  for (int l = 0; l < 3; ++l) {
    const int K = l == 0 ? kInPad : kHidden;
    const int n0 = wv * 64;
    const int steps = K / 16;  // 2 or 16
    // fragment (column block 4 wv + j, k step s) at wl + (j steps + s) 256
    const uint16_t* wl = wt + kWOff[l] + (size_t)(4 * wv * steps) * 256 + 4 * lane;
    float4_ acc[4][4];
    for (int i = 0; i < 4; ++i)
      for (int j = 0; j < 4; ++j) acc[i][j] = float4_{0.f, 0.f, 0.f, 0.f};
    // Weight fragments come from L2 and load kPre k steps ahead: a ring of
    // kPre x 4 fragments, the k loop unrolled by kPre so that the ring
    // index is a constant.  With the fragment-ordered weights (one 512-byte
    // read per fragment) depths 2, 4 and 8 time the same on MI355X (0.078
    // to 0.080 ms per 65,536 rows); with one column per lane from the
    // row-major transpose (16 partial cache lines per fragment) the forward
    // took 0.17, 0.16 and 0.14 ms.
    constexpr int kPre = 2;
    short4_ bq[kPre][4];
    for (int d = 0; d < kPre; ++d)
      for (int j = 0; j < 4; ++j)
        bq[d][j] = d < steps ? *reinterpret_cast<const short4_*>(wl + (size_t)(j * steps + d) * 256)
                             : short4_{0, 0, 0, 0};
    for (int s0 = 0; s0 < steps; s0 += kPre) {
#pragma unroll
      for (int d = 0; d < kPre; ++d) {
        const int s = s0 + d;
        if (s < steps) {
          short4_ af[4];
          for (int i = 0; i < 4; ++i)
            af[i] = *reinterpret_cast<const short4_*>(&act[i * 16 + lr][16 * s + 4 * lq]);
          for (int i = 0; i < 4; ++i)
            for (int j = 0; j < 4; ++j)
              acc[i][j] =
                  __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(af[i], bq[d][j], acc[i][j], 0, 0, 0);
          if (s + kPre < steps)
            for (int j = 0; j < 4; ++j)
              bq[d][j] = *reinterpret_cast<const short4_*>(wl + (size_t)(j * steps + s + kPre) * 256);
        }
      }
    }
    __syncthreads();  // every wave has read this layer's inputs
    const float* b = w + kBOff[l];
    for (int j = 0; j < 4; ++j) {
      const int col = n0 + j * 16 + lr;
      const float bj = b[col];
      for (int i = 0; i < 4; ++i)
        for (int r = 0; r < 4; ++r) act[i * 16 + 4 * lq + r][col] = f2bf(gelu_f(acc[i][j][r] + bj));
    }
    __syncthreads();
  }
  const uint16_t* wl = wt + kWOff[3] + 4 * lane;  // fragment (j, s) at wl + (16 j + s) 256
  float4_ acc[5];
  for (int j = 0; j < 5; ++j) acc[j] = float4_{0.f, 0.f, 0.f, 0.f};
  constexpr int kPre = 4, kSteps = kHidden / 16;
  short4_ bq[kPre][5];
  for (int d = 0; d < kPre; ++d)
    for (int j = 0; j < 5; ++j)
      bq[d][j] = *reinterpret_cast<const short4_*>(wl + (size_t)(j * kSteps + d) * 256);
  for (int s0 = 0; s0 < kSteps; s0 += kPre) {
#pragma unroll
    for (int d = 0; d < kPre; ++d) {
      const int s = s0 + d;
      const short4_ af = *reinterpret_cast<const short4_*>(&act[wv * 16 + lr][16 * s + 4 * lq]);
      for (int j = 0; j < 5; ++j)
        acc[j] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(af, bq[d][j], acc[j], 0, 0, 0);
      if (s + kPre < kSteps)
        for (int j = 0; j < 5; ++j)
          bq[d][j] = *reinterpret_cast<const short4_*>(wl + (size_t)(j * kSteps + s + kPre) * 256);
    }
  }
  const float* b = w + kBOff[3];
  for (int j = 0; j < 5; ++j) {
    const int col = j * 16 + lr;
    const float bj = b[col];
    for (int r = 0; r < 4; ++r) {
      const int row = row0 + wv * 16 + 4 * lq + r;
      if (row < n) out[(size_t)row * kOutPad + col] = acc[j][r] + bj;
    }
  }
}

}  // namespace fd2mlp

using namespace fd2mlp;

// -------------------------------------------------------------- host side

struct fd2_mlp {
  int device = 0;
  hipStream_t stream = nullptr;
  int cap = 0;  // batch capacity
  float* w = nullptr;   // FP32 master parameters (kParams)
  float* g = nullptr;   // gradients
  float* m = nullptr;   // Adam moments
  float* v = nullptr;
  uint16_t* wb = nullptr;  // bf16 copy (same layout as w)
  uint16_t* wt = nullptr;  // bf16 weight matrices transposed (fused forward)
  bool wt_stale = true;    // wb changed since wt was built
  uint16_t* x0 = nullptr;
  float* z[4] = {};
  uint16_t* a[3] = {};
  uint16_t* dz = nullptr;    // bf16 gradient of the current layer's output
  uint16_t* dz2 = nullptr;
  float* y = nullptr;
  float* loss = nullptr;
  float* data_x = nullptr;
  float* data_y = nullptr;
  int64_t n_data = 0;
  uint32_t* idx = nullptr;
  int64_t step = 0;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  // fd2_surrogate_submit's buffers, kept between calls and grown on demand
  // (a call of the same size allocates nothing): device buffers by SurrBuf
  // index, and the pinned copy of the audit predictions
  void* sbuf[9] = {};
  size_t sbuf_cap[9] = {};
  float* h_pred = nullptr;
  size_t h_pred_cap = 0;
};

namespace {

#define MLP_HIP(call)                          \
  do {                                         \
    if ((call) != hipSuccess) return FD2_E_DEVICE; \
  } while (0)

void free_batch(fd2_mlp* p) {
  void* bufs[] = {p->x0, p->z[0], p->z[1], p->z[2], p->z[3], p->a[0], p->a[1], p->a[2],
                  p->dz, p->dz2, p->y, p->idx};
  for (void* b : bufs)
    if (b) (void)hipFree(b);
  p->x0 = nullptr;
  for (auto& q : p->z) q = nullptr;
  for (auto& q : p->a) q = nullptr;
  p->dz = p->dz2 = nullptr;
  p->y = nullptr;
  p->idx = nullptr;
  p->cap = 0;
}

// fd2_surrogate_submit's device buffers (fd2_mlp::sbuf)
enum SurrBuf : int {
  kBufAll, kBufEv, kBufIdx, kBufIds, kBufRad, kBufLogits, kBufSpread, kBufFeat, kBufPred, kBufCount
};
static_assert(kBufCount == sizeof(fd2_mlp::sbuf) / sizeof(void*), "one slot per buffer");

// Buffer `which` with room for `bytes`: the kept one, or a larger one (the
// old one freed; the previous call has synchronised its streams).
template <class T>
bool reserve(fd2_mlp* p, int which, size_t bytes, T** out) {
  void*& b = p->sbuf[which];
  size_t& cap = p->sbuf_cap[which];
  if (!b || cap < bytes) {
    if (b) (void)hipFree(b);
    b = nullptr;
    cap = 0;
    if (hipMalloc(&b, bytes ? bytes : 1) != hipSuccess) return false;
    cap = bytes;
  }
  *out = static_cast<T*>(b);
  return true;
}

int ensure_batch(fd2_mlp* p, int B) {
  if (B <= p->cap) return FD2_OK;
  free_batch(p);
  const size_t b = (size_t)B;
  MLP_HIP(hipMalloc(&p->x0, b * kInPad * 2));
  for (int l = 0; l < 3; ++l) {
    MLP_HIP(hipMalloc(&p->z[l], b * kHidden * 4));
    MLP_HIP(hipMalloc(&p->a[l], b * kHidden * 2));
  }
  MLP_HIP(hipMalloc(&p->z[3], b * kOutPad * 4));
  MLP_HIP(hipMalloc(&p->dz, b * kHidden * 2));
  MLP_HIP(hipMalloc(&p->dz2, b * kHidden * 2));
  MLP_HIP(hipMalloc(&p->y, b * kTarget * 4));
  MLP_HIP(hipMalloc(&p->idx, b * 4));
  p->cap = B;
  return FD2_OK;
}

// forward pass on p->x0 (B rows) -> z[3] logits
// forward pass of B rows of p->x0; the output layer goes to `out` (B x
// kOutPad f32) when given, else to p->z[3]
int forward(fd2_mlp* p, int B, hipStream_t stream, float* out = nullptr) {
  const int ins[4] = {kInPad, kHidden, kHidden, kHidden};
  const int outs[4] = {kHidden, kHidden, kHidden, kOutPad};
  const uint16_t* in = p->x0;
  for (int l = 0; l < 4; ++l) {
    GemmArgs a{};
    a.A = in;
    a.B = p->wb + kWOff[l];
    a.M = B;
    a.N = outs[l];
    a.K = ins[l];
    a.lda = ins[l];
    a.ldb = outs[l];
    a.kchunk = ins[l];
    a.bias = p->w + kBOff[l];
    a.outf = l == 3 && out ? out : p->z[l];
    a.outb = l < 3 ? p->a[l] : nullptr;
    a.ldo = outs[l];
    MLP_HIP((gemm<false, false, kStoreF32>(a, 1, stream)));
    in = l < 3 ? p->a[l] : nullptr;
  }
  return FD2_OK;
}

// inference forward of B rows of p->x0 in one kernel (k_forward_fused):
// the output logits go to `out` (B x kOutPad f32) when given, else to
// p->z[3]; the other layers' buffers are not written.  Rebuilds the
// transposed weights first when the parameters changed.
int forward_fused(fd2_mlp* p, int B, hipStream_t stream, float* out = nullptr) {
  if (p->wt_stale) {
    hipLaunchKernelGGL(k_fragment_weights, dim3(64, 4), dim3(256), 0, stream, p->wb, p->wt);
    MLP_HIP(hipGetLastError());
    p->wt_stale = false;
  }
  if (B <= 0) return FD2_OK;
  hipLaunchKernelGGL(k_forward_fused, dim3((unsigned)((B + kFuseRows - 1) / kFuseRows)),
                     dim3(256), 0, stream, p->x0, B, p->wt, p->w, out ? out : p->z[3]);
  MLP_HIP(hipGetLastError());
  return FD2_OK;
}

// backward from the bf16 output gradient in p->dz (B x 80) into p->g
int backward(fd2_mlp* p, int B) {
  const int ins[4] = {kInPad, kHidden, kHidden, kHidden};
  const int outs[4] = {kHidden, kHidden, kHidden, kOutPad};
  if (B % 16) return FD2_E_INVAL;
  const int target = B >= 1024 ? 16 : 1;
  const int kchunk = ((B + target - 1) / target + 15) / 16 * 16;
  const int splits = (B + kchunk - 1) / kchunk;
  uint16_t* dcur = p->dz;   // gradient w.r.t. layer l's output Z_l
  uint16_t* dnext = p->dz2;
  for (int l = 3; l >= 0; --l) {
    const uint16_t* in = l == 0 ? p->x0 : p->a[l - 1];
    // dW_l = in^T dZ_l  ([ins x B] x [B x outs]), split over B
    GemmArgs gw{};
    gw.A = in;
    gw.B = dcur;
    gw.M = ins[l];
    gw.N = outs[l];
    gw.K = B;
    gw.lda = ins[l];
    gw.ldb = outs[l];
    gw.kchunk = kchunk;
    gw.outf = p->g + kWOff[l];
    MLP_HIP((gemm<true, false, kAtomicF32>(gw, splits, p->stream)));
    if (l == 0) break;
    // dZ_{l-1} = (dZ_l W_l^T) * gelu'(Z_{l-1}); bias grad of layer l-1
    GemmArgs gx{};
    gx.A = dcur;
    gx.B = p->wb + kWOff[l];
    gx.M = B;
    gx.N = ins[l];
    gx.K = outs[l];
    gx.lda = outs[l];
    gx.ldb = outs[l];
    gx.kchunk = outs[l];
    gx.Z = p->z[l - 1];
    gx.outb = dnext;
    gx.ldo = ins[l];
    gx.dbias = p->g + kBOff[l - 1];
    MLP_HIP((gemm<false, true, kGeluGradBf16>(gx, 1, p->stream)));
    uint16_t* t = dcur;
    dcur = dnext;
    dnext = t;
  }
  return FD2_OK;
}

int loss_grad(fd2_mlp* p, int B, float* d_loss) {
  hipLaunchKernelGGL(k_zero, dim3(1), dim3(64), 0, p->stream, d_loss, (size_t)1);
  hipLaunchKernelGGL(k_head_loss, dim3(std::min((B * 64 + 255) / 256, kHeadLossBlocks)), dim3(256),
                     0, p->stream,
                     p->z[3], p->y, B, p->dz, p->g + kBOff[3], d_loss);
  MLP_HIP(hipGetLastError());
  return FD2_OK;
}

void init_params(std::vector<float>& w, uint32_t seed) {
  // Glorot-uniform weights, zero biases (padded rows/cols stay zero)
  uint64_t s = 0x9E3779B97F4A7C15ull ^ seed;
  auto rnd = [&]() {
    s ^= s >> 12; s ^= s << 25; s ^= s >> 27;
    return (double)((s * 2685821657736338717ull) >> 11) * (1.0 / 9007199254740992.0);
  };
  const int ins[4] = {kFeat, kHidden, kHidden, kHidden};
  const int insp[4] = {kInPad, kHidden, kHidden, kHidden};
  const int outs[4] = {kHidden, kHidden, kHidden, kOut};
  const int outsp[4] = {kHidden, kHidden, kHidden, kOutPad};
  std::fill(w.begin(), w.end(), 0.0f);
  for (int l = 0; l < 4; ++l) {
    const double lim = std::sqrt(6.0 / (ins[l] + outs[l]));
    for (int i = 0; i < ins[l]; ++i)
      for (int j = 0; j < outs[l]; ++j)
        w[kWOff[l] + (size_t)i * outsp[l] + j] = (float)((2.0 * rnd() - 1.0) * lim);
    (void)insp;
  }
}

}  // namespace

extern "C" {

size_t fd2_mlp_num_params(void) { return kParams; }

int fd2_mlp_create(int device, uint32_t seed, fd2_mlp** out) {
  if (!out) return FD2_E_INVAL;
  *out = nullptr;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return FD2_E_NO_DEVICE;
  if (device < 0 || device >= ndev) return FD2_E_INVAL;
  fd2_mlp* p = new (std::nothrow) fd2_mlp();
  if (!p) return FD2_E_NOMEM;
  p->device = device;
  int rc = FD2_OK;
  do {
    if (hipSetDevice(device) != hipSuccess ||
        hipStreamCreateWithFlags(&p->stream, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreate(&p->e0) != hipSuccess || hipEventCreate(&p->e1) != hipSuccess) {
      rc = FD2_E_DEVICE;
      break;
    }
    const size_t n = kParams;
    if (hipMalloc(&p->w, n * 4) != hipSuccess || hipMalloc(&p->g, n * 4) != hipSuccess ||
        hipMalloc(&p->m, n * 4) != hipSuccess || hipMalloc(&p->v, n * 4) != hipSuccess ||
        hipMalloc(&p->wb, n * 2) != hipSuccess || hipMalloc(&p->wt, n * 2) != hipSuccess ||
        hipMalloc(&p->loss, 4) != hipSuccess) {
      rc = FD2_E_NOMEM;
      break;
    }
    std::vector<float> h(n);
    init_params(h, seed);
    rc = fd2_mlp_set_params(p, h.data(), n);
  } while (0);
  if (rc) {
    fd2_mlp_destroy(p);
    return rc;
  }
  *out = p;
  return FD2_OK;
}

void fd2_mlp_destroy(fd2_mlp* p) {
  if (!p) return;
  (void)hipSetDevice(p->device);
  if (p->stream) (void)hipStreamSynchronize(p->stream);
  free_batch(p);
  void* bufs[] = {p->w, p->g, p->m, p->v, p->wb, p->wt, p->loss, p->data_x, p->data_y};
  for (void* b : bufs)
    if (b) (void)hipFree(b);
  for (void* b : p->sbuf)
    if (b) (void)hipFree(b);
  if (p->h_pred) (void)hipHostFree(p->h_pred);
  if (p->e0) (void)hipEventDestroy(p->e0);
  if (p->e1) (void)hipEventDestroy(p->e1);
  if (p->stream) (void)hipStreamDestroy(p->stream);
  delete p;
}

int fd2_mlp_set_params(fd2_mlp* p, const float* host, size_t n) {
  if (!p || !host || n != kParams) return FD2_E_INVAL;
  MLP_HIP(hipSetDevice(p->device));
  MLP_HIP(hipMemcpyAsync(p->w, host, n * 4, hipMemcpyHostToDevice, p->stream));
  hipLaunchKernelGGL(k_to_bf16, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, p->stream,
                     p->w, p->wb, (int)n);
  p->wt_stale = true;
  const unsigned nb = (unsigned)((n + 255) / 256);
  hipLaunchKernelGGL(k_zero, dim3(nb), dim3(256), 0, p->stream, p->g, n);
  hipLaunchKernelGGL(k_zero, dim3(nb), dim3(256), 0, p->stream, p->m, n);
  hipLaunchKernelGGL(k_zero, dim3(nb), dim3(256), 0, p->stream, p->v, n);
  MLP_HIP(hipGetLastError());
  MLP_HIP(hipStreamSynchronize(p->stream));
  p->step = 0;
  return FD2_OK;
}

int fd2_mlp_get_params(fd2_mlp* p, float* host, size_t n) {
  if (!p || !host || n != kParams) return FD2_E_INVAL;
  MLP_HIP(hipSetDevice(p->device));
  MLP_HIP(hipStreamSynchronize(p->stream));
  MLP_HIP(hipMemcpy(host, p->w, n * 4, hipMemcpyDeviceToHost));
  return FD2_OK;
}

int fd2_mlp_forward(fd2_mlp* p, const float* x, int64_t n, float* out) {
  if (!p || n < 0 || (n > 0 && (!x || !out))) return FD2_E_INVAL;
  MLP_HIP(hipSetDevice(p->device));
  const int64_t chunk = 65536;
  float* dx = nullptr;
  float* dout = nullptr;
  int rc = ensure_batch(p, (int)chunk);
  if (rc) return rc;
  MLP_HIP(hipMalloc(&dx, (size_t)chunk * kFeat * 4));
  if (hipMalloc(&dout, (size_t)chunk * kTarget * 4) != hipSuccess) {
    (void)hipFree(dx);
    return FD2_E_NOMEM;
  }
  for (int64_t s = 0; s < n && rc == FD2_OK; s += chunk) {
    const int b = (int)(n - s < chunk ? n - s : chunk);
    if (hipMemcpyAsync(dx, x + s * kFeat, (size_t)b * kFeat * 4, hipMemcpyHostToDevice,
                       p->stream) != hipSuccess) {
      rc = FD2_E_DEVICE;
      break;
    }
    hipLaunchKernelGGL(k_pack_inputs, dim3((b * kIn + 255) / 256), dim3(256), 0, p->stream,
                       dx, b, p->x0);
    rc = forward_fused(p, b, p->stream);
    if (rc) break;
    hipLaunchKernelGGL(k_head_probs, dim3((b * 64 + 255) / 256), dim3(256), 0, p->stream,
                       p->z[3], b, dout);
    if (hipMemcpyAsync(out + s * kTarget, dout, (size_t)b * kTarget * 4,
                       hipMemcpyDeviceToHost, p->stream) != hipSuccess ||
        hipStreamSynchronize(p->stream) != hipSuccess)
      rc = FD2_E_DEVICE;
  }
  (void)hipFree(dx);
  (void)hipFree(dout);
  return rc;
}

int fd2_mlp_logits(fd2_mlp* p, const float* x, int64_t n, float* logits, int layered) {
  if (!p || n < 0 || (n > 0 && (!x || !logits))) return FD2_E_INVAL;
  MLP_HIP(hipSetDevice(p->device));
  const int64_t chunk = 65536;
  float* dx = nullptr;
  int rc = ensure_batch(p, (int)chunk);
  if (rc) return rc;
  MLP_HIP(hipMalloc(&dx, (size_t)chunk * kFeat * 4));
  for (int64_t s = 0; s < n && rc == FD2_OK; s += chunk) {
    const int b = (int)(n - s < chunk ? n - s : chunk);
    if (hipMemcpyAsync(dx, x + s * kFeat, (size_t)b * kFeat * 4, hipMemcpyHostToDevice,
                       p->stream) != hipSuccess) {
      rc = FD2_E_DEVICE;
      break;
    }
    hipLaunchKernelGGL(k_pack_inputs, dim3((b * kIn + 255) / 256), dim3(256), 0, p->stream,
                       dx, b, p->x0);
    rc = layered ? forward(p, b, p->stream) : forward_fused(p, b, p->stream);
    if (rc) break;
    // the kOut logits of each row (of kOutPad computed)
    if (hipMemcpy2DAsync(logits + s * kOut, kOut * 4, p->z[3], kOutPad * 4, kOut * 4, b,
                         hipMemcpyDeviceToHost, p->stream) != hipSuccess ||
        hipStreamSynchronize(p->stream) != hipSuccess)
      rc = FD2_E_DEVICE;
  }
  (void)hipFree(dx);
  return rc;
}

int fd2_mlp_time_forward(fd2_mlp* p, int B, int reps, float* ms) {
  if (!p || !ms || B <= 0 || B % 16 || reps <= 0) return FD2_E_INVAL;
  MLP_HIP(hipSetDevice(p->device));
  int rc = ensure_batch(p, B);
  if (rc) return rc;
  MLP_HIP(hipMemsetAsync(p->x0, 0, (size_t)B * kInPad * 2, p->stream));  // defined inputs
  rc = forward_fused(p, B, p->stream);  // warm-up
  if (rc) return rc;
  MLP_HIP(hipEventRecord(p->e0, p->stream));
  for (int r = 0; r < reps && rc == FD2_OK; ++r) rc = forward_fused(p, B, p->stream);
  if (rc) return rc;
  MLP_HIP(hipEventRecord(p->e1, p->stream));
  MLP_HIP(hipStreamSynchronize(p->stream));
  float t = 0.0f;
  MLP_HIP(hipEventElapsedTime(&t, p->e0, p->e1));
  *ms = t / (float)reps;
  return FD2_OK;
}

// gradients of the mean batch loss (no update): grads (kParams) and loss
int fd2_mlp_gradients(fd2_mlp* p, const float* x, const float* y, int B, float* grads,
                      float* loss) {
  if (!p || !x || !y || !grads || B <= 0 || B % 16) return FD2_E_INVAL;
  MLP_HIP(hipSetDevice(p->device));
  int rc = ensure_batch(p, B);
  if (rc) return rc;
  float* dx = nullptr;
  MLP_HIP(hipMalloc(&dx, (size_t)B * kFeat * 4));
  rc = FD2_OK;
  if (hipMemcpyAsync(dx, x, (size_t)B * kFeat * 4, hipMemcpyHostToDevice, p->stream) != hipSuccess ||
      hipMemcpyAsync(p->y, y, (size_t)B * kTarget * 4, hipMemcpyHostToDevice, p->stream) != hipSuccess)
    rc = FD2_E_DEVICE;
  if (!rc) {
    hipLaunchKernelGGL(k_zero, dim3((unsigned)((kParams + 255) / 256)), dim3(256), 0, p->stream,
                       p->g, (size_t)kParams);
    hipLaunchKernelGGL(k_pack_inputs, dim3((B * kIn + 255) / 256), dim3(256), 0, p->stream,
                       dx, B, p->x0);
    rc = forward(p, B, p->stream);
  }
  if (!rc) rc = loss_grad(p, B, p->loss);
  if (!rc) rc = backward(p, B);
  if (!rc && (hipMemcpyAsync(grads, p->g, kParams * 4, hipMemcpyDeviceToHost, p->stream) != hipSuccess ||
              (loss && hipMemcpyAsync(loss, p->loss, 4, hipMemcpyDeviceToHost, p->stream) != hipSuccess) ||
              hipStreamSynchronize(p->stream) != hipSuccess))
    rc = FD2_E_DEVICE;
  (void)hipFree(dx);
  return rc;
}

int fd2_mlp_set_data(fd2_mlp* p, const float* x, const float* y, int64_t n) {
  if (!p || !x || !y || n <= 0) return FD2_E_INVAL;
  MLP_HIP(hipSetDevice(p->device));
  if (p->data_x) (void)hipFree(p->data_x);
  if (p->data_y) (void)hipFree(p->data_y);
  p->data_x = p->data_y = nullptr;
  MLP_HIP(hipMalloc(&p->data_x, (size_t)n * kFeat * 4));
  MLP_HIP(hipMalloc(&p->data_y, (size_t)n * kTarget * 4));
  MLP_HIP(hipMemcpy(p->data_x, x, (size_t)n * kFeat * 4, hipMemcpyHostToDevice));
  MLP_HIP(hipMemcpy(p->data_y, y, (size_t)n * kTarget * 4, hipMemcpyHostToDevice));
  p->n_data = n;
  return FD2_OK;
}

// `steps` Adam steps on minibatches of B rows drawn (host RNG, seeded) from
// the resident data set; losses[steps] receives each step's batch loss
// (may be NULL); *ms the GPU time of the steps (may be NULL).
int fd2_mlp_train(fd2_mlp* p, int steps, int B, float lr, uint32_t seed, float* losses,
                  float* ms) {
  if (!p || steps < 0 || B <= 0 || B % 16 || p->n_data <= 0) return FD2_E_INVAL;
  MLP_HIP(hipSetDevice(p->device));
  int rc = ensure_batch(p, B);
  if (rc) return rc;
  // all minibatch indices drawn up front (xorshift64*, seeded)
  std::vector<uint32_t> idx((size_t)B * steps);
  uint64_t s = 0xD1B54A32D192ED03ull ^ ((uint64_t)seed << 17) ^ (uint64_t)p->step;
  for (auto& v : idx) {
    s ^= s >> 12; s ^= s << 25; s ^= s >> 27;
    v = (uint32_t)(((s * 2685821657736338717ull) >> 32) % (uint64_t)p->n_data);
  }
  uint32_t* d_idx = nullptr;
  float* d_losses = nullptr;
  MLP_HIP(hipMalloc(&d_idx, idx.size() * 4 + 4));
  if (hipMalloc(&d_losses, (size_t)(steps + 1) * 4) != hipSuccess) {
    (void)hipFree(d_idx);
    return FD2_E_NOMEM;
  }
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  rc = hipMemcpyAsync(d_idx, idx.data(), idx.size() * 4, hipMemcpyHostToDevice, p->stream) ==
               hipSuccess && hipEventRecord(p->e0, p->stream) == hipSuccess
           ? FD2_OK
           : FD2_E_DEVICE;
  for (int st = 0; st < steps && rc == FD2_OK; ++st) {
    hipLaunchKernelGGL(k_gather, dim3((B * 64 + 255) / 256), dim3(256), 0, p->stream, p->data_x,
                       p->data_y, d_idx + (size_t)st * B, B, p->x0, p->y);
    rc = forward(p, B, p->stream);
    if (!rc) rc = loss_grad(p, B, d_losses + st);
    if (!rc) rc = backward(p, B);
    if (rc) break;
    p->step++;
    const float c1 = 1.0f - std::pow(b1, (float)p->step), c2 = 1.0f - std::pow(b2, (float)p->step);
    hipLaunchKernelGGL(k_adam, dim3((unsigned)((kParams + 255) / 256)), dim3(256), 0, p->stream,
                       p->w, p->g, p->m, p->v, p->wb, (int)kParams, lr, b1, b2, eps, c1, c2);
    p->wt_stale = true;
    if (hipGetLastError() != hipSuccess) rc = FD2_E_DEVICE;
  }
  if (!rc && (hipEventRecord(p->e1, p->stream) != hipSuccess ||
              (losses && hipMemcpyAsync(losses, d_losses, (size_t)steps * 4,
                                        hipMemcpyDeviceToHost, p->stream) != hipSuccess) ||
              hipStreamSynchronize(p->stream) != hipSuccess))
    rc = FD2_E_DEVICE;
  if (!rc && ms && hipEventElapsedTime(ms, p->e0, p->e1) != hipSuccess) rc = FD2_E_DEVICE;
  (void)hipStreamSynchronize(p->stream);
  (void)hipFree(d_idx);
  (void)hipFree(d_losses);
  return rc;
}

}  // extern "C"

// ===================================================== surrogate transport

namespace fd2mlp {

// Sampler shapes: the side (pixels) of the LDS tile of geometry 0's PRF,
// centred on the detector, and the workgroup size.  The sampler runs after
// the Monte Carlo launch has drained, so the tile is sized for the
// sampler's own occupancy, not the transport's.  Large launches take a
// 176-pixel tile (124 KB, one 1024-thread workgroup per CU); each
// workgroup flushes its tile once, which small launches would not repay,
// so they take a 128-pixel tile (64 KB, two 256-thread workgroups per CU).
constexpr int kSampleTile = 128, kSampleThreads = 256;
constexpr int kSampleTileLarge = 176, kSampleThreadsLarge = 1024;
constexpr int kSampleLargeEventsPerWave = 64;  // events per wave for the large shape

// network features of an event (see fd2.h): depth below the top / H,
// col_dx / R, H / 500 um, R / 5 um, top reflectivity
__device__ __forceinline__ float feature(const fd2_event& e, const fd2::GeomDev& g, int c) {
  switch (c) {
    case 0: return (g.half_h - e.z) / g.H;
    case 1: return e.col_dx / g.R;
    case 2: return g.H * (1.0f / 500.0f);
    case 3: return g.R * (1.0f / 5.0f);
    case 4: return 1.0f - g.top_absfrac;
    default: return 0.0f;
  }
}

__global__ void k_surr_features(const fd2_event* ev, int n, const fd2::GeomDev* geoms,
                                uint16_t* x0) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n * kInPad) return;
  const int row = i / kInPad, c = i % kInPad;
  const fd2_event e = ev[row];
  x0[i] = c < kFeat ? f2bf(feature(e, geoms[e.geom], c)) : (uint16_t)0;
}

// Network events gathered from the whole submitted array by index; their
// random-stream keys are the global event numbers base + index.
__global__ void k_surr_gather(const fd2_event* __restrict__ all, const uint32_t* __restrict__ idx,
                              int64_t n, uint64_t base, fd2_event* __restrict__ out,
                              uint64_t* __restrict__ ids) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint32_t k = idx[i];
  out[i] = all[k];
  ids[i] = base + k;
}

// edges of the radial bins: 0, then 0.5 * 1600^((j-1)/62) um for j = 1..63
// (0.5 to 800 um), then 1000 um: bin 0 = [0, 0.5), 62 log bins, [800, 1000)
__device__ __forceinline__ float radial_edge(int j) {
  if (j <= 0) return 0.0f;
  if (j >= FD2_RADIAL_BINS) return 1000.0f;
  return 0.5f * expf((float)(j - 1) * 0.118996111f);  // ln(1600)/62
}

struct SampleArgs {
  const fd2_event* events;
  const uint64_t* ids;       // global event numbers (random-stream keys)
  const float* logits;       // n x kOutPad
  const fd2::GeomDev* geoms;
  unsigned long long* counters;
  int n;
  int n_geoms;
  uint32_t seed;
  int tile;                  // LDS PRF tile side (geometry 0), 0 = none
  int tile_i0, tile_j0;      // pixel origin of the tile
  uint32_t* spread;          // kSpreadCopies u32 copies of geometry 0's PRF, or null
};

// Geometry 0's detections outside the LDS tile go to one of several u32
// copies of its PRF, picked by workgroup number, so that the atomics of
// different workgroups (and XCDs, which take workgroups in turn) meet on
// the same address less often; k_spread_flush adds the copies to the u64
// PRF.  Each copy holds fewer counts than the photons of one launch, which
// the host keeps below 2^32.
constexpr int kSpreadCopies = 8;

// Area-weighted mean radius of radial bin j, (2/3)(r1^3 - r0^3)/(r1^2 - r0^2).
__device__ __forceinline__ float radial_centroid(int j) {
  const float r0 = radial_edge(j), r1 = radial_edge(j + 1);
  return (2.0f / 3.0f) * (r1 * r1 + r1 * r0 + r0 * r0) / (r1 + r0);
}

// One wave per event: lanes are the 64 radial bins for the softmax and its
// CDF; the binomial detected count is a sum of N Bernoulli(p_det) draws
// spread over the lanes; each detected photon gets a bin by inverse CDF, a
// radius within the bin and a uniform azimuth, and is tallied like a Monte
// Carlo detection.  Within a bin the density per unit area is linear in r,
// rho_j (1 + s_j (r - c_j)) with c_j the bin's area centroid (so the bin
// keeps its probability for any slope), and s_j the slope of the bins'
// mean densities across the neighbouring bins, bounded so that the density
// stays >= 0 in the bin: the radial density is then continuous across bin
// edges instead of a staircase whose steps (bins 12.6% wide in r, ~25 um
// at r = 200 um) print rings into the PRF.  The radius is drawn uniform in
// area and accepted with probability (1 + s_j (r - c_j)) / max over the
// bin.  A workgroup loops over many events and
// keeps the central part of geometry 0's PRF in LDS.  The read-only tables
// come again as __restrict__ parameters, so that the geometry constants are
// read with scalar loads rather than reloaded around every tally atomic.
template <int THREADS>
__global__ void __launch_bounds__(THREADS)
k_surr_sample(SampleArgs a, const fd2_event* __restrict__ events,
              const uint64_t* __restrict__ ids, const float* __restrict__ logits,
              const fd2::GeomDev* __restrict__ geoms) {
  extern __shared__ uint32_t s_tile[];
  __shared__ uint32_t s_cnt[FD2_MAX_GEOMS * 3];
  // the radial bins' edges and area centroids (constants), and per wave its
  // event's CDF, slopes and guide table (s_guide[k]: the first bin whose CDF
  // reaches k/64), so that a detection finds its bin, slope and edges with
  // a few independent LDS reads instead of six dependent shuffles and four
  // exponentials
  __shared__ float s_edge[FD2_RADIAL_BINS + 1];
  __shared__ float s_cent[FD2_RADIAL_BINS];
  __shared__ float s_cdf[THREADS / 64][64];
  __shared__ float s_slope[THREADS / 64][64];
  __shared__ int s_guide[THREADS / 64][64];
  const int T = a.tile, T2 = a.tile * a.tile;
  for (int i = threadIdx.x; i < T2; i += blockDim.x) s_tile[i] = 0u;
  for (int i = threadIdx.x; i < a.n_geoms * 3; i += blockDim.x) s_cnt[i] = 0u;
  if (threadIdx.x <= FD2_RADIAL_BINS) s_edge[threadIdx.x] = radial_edge(threadIdx.x);
  if (threadIdx.x < FD2_RADIAL_BINS) s_cent[threadIdx.x] = radial_centroid(threadIdx.x);
  __syncthreads();
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int waves = gridDim.x * (THREADS / 64);
  for (int e = blockIdx.x * (THREADS / 64) + w; e < a.n; e += waves) {
    const fd2_event ev = events[e];
    const fd2::GeomDev& g = geoms[ev.geom];
    const float* z = logits + (size_t)e * kOutPad;
    const float zl = z[1 + lane];
    float mx = zl;
    for (int o = 32; o > 0; o >>= 1) mx = fmaxf(mx, __shfl_xor(mx, o, 64));
    const float ex = expf(zl - mx);
    float cdf = ex;  // inclusive scan over lanes
    for (int o = 1; o < 64; o <<= 1) {
      const float t = __shfl_up(cdf, o, 64);
      if (lane >= o) cdf += t;
    }
    const float total = __shfl(cdf, 63, 64);
    cdf = lane == 63 ? 1.0f : cdf / total;
    // this lane's bin: mean density and relative slope within the bin
    float slope = 0.0f;
    {
      const float e0 = s_edge[lane], e1 = s_edge[lane + 1];
      const float rho = (ex / total) / (e1 * e1 - e0 * e0);  // x pi, cancels
      const float rho_dn = __shfl_up(rho, 1, 64), rho_up = __shfl_down(rho, 1, 64);
      const float c = s_cent[lane];
      const float c_dn = lane > 0 ? s_cent[lane - 1] : 0.0f;
      const float c_up = lane < 63 ? s_cent[lane + 1] : 0.0f;
      if (lane > 0 && lane < 63 && rho > 0.0f) {
        slope = (rho_up - rho_dn) / (rho * (c_up - c_dn));
        // keep 1 + slope (r - c) >= 0 on [e0, e1]
        const float lim_lo = 1.0f / (c - e0), lim_hi = 1.0f / (e1 - c);
        slope = fminf(fmaxf(slope, -lim_hi), lim_lo);
      }
    }
    // the CDF is non-decreasing (an inclusive scan of non-negative terms
    // over one divisor, the last set to 1), so the first bin reaching t is
    // the number of bins below t
    s_cdf[w][lane] = cdf;
    s_slope[w][lane] = slope;
    {
      int gk = 0;
      for (int k = 0; k < 64; ++k) {
        const unsigned long long below = __ballot(cdf < (float)k * (1.0f / 64.0f));
        if (lane == k) gk = (int)__popcll(below);
      }
      s_guide[w][lane] = gk;
    }
    // the wave's LDS writes complete in order before its reads below; the
    // barrier keeps the compiler from moving the reads above them
    __builtin_amdgcn_wave_barrier();
    const float pdet = 1.0f / (1.0f + expf(-z[0]));
    fd2::CounterRng rng;
    rng.seed(a.seed ^ 0x53555252u, ids[e], (uint32_t)lane);
    int cnt = 0;
    for (int j = lane; j < ev.N; j += 64) cnt += rng.u() < pdet ? 1 : 0;
    for (int o = 32; o > 0; o >>= 1) cnt += __shfl_xor(cnt, o, 64);
    const int K = cnt;
    for (int j0 = 0; j0 < K; j0 += 64) {
      if (j0 + lane >= K) continue;
      const float u = rng.u();  // in (0, 1]
      // first bin whose CDF reaches u: from the guide entry of floor(64 u)
      // (u >= k/64, so the bin is at or after it) forward; the last bin's
      // CDF is 1, so the walk ends by bin 63
      int bin = s_guide[w][min((int)(u * 64.0f), 63)];
      while (s_cdf[w][bin] < u) ++bin;
      const float sb = s_slope[w][bin];
      const float r0 = s_edge[bin], r1 = s_edge[bin + 1];
      const float cb = s_cent[bin];
      const float wmax = fmaxf(1.0f + sb * (r0 - cb), 1.0f + sb * (r1 - cb));
      float r;
      for (int tries = 0;; ++tries) {  // acceptance >= 1/2 per try
        r = sqrtf(r0 * r0 + rng.u() * (r1 * r1 - r0 * r0));
        if (tries >= 63 || rng.u() * wmax <= 1.0f + sb * (r - cb)) break;
      }
      float sn, cs;
      fd2_sincos(rng.u() * 6.283185307f, &sn, &cs);
      const float x = ev.x + r * cs, y = ev.y + r * sn;
      // x / pitch, as a product when the pitch is a power of two (exact)
      const float qx = g.inv_pixel != 0.0f ? (x - g.lbx) * g.inv_pixel : (x - g.lbx) / (float)g.pixel;
      const float qy = g.inv_pixel != 0.0f ? (y - g.lby) * g.inv_pixel : (y - g.lby) / (float)g.pixel;
      const int pi = (int)floorf(qx);
      const int pj = (int)floorf(qy);
      if (x <= g.ubx && y <= g.uby && x >= g.lbx && y >= g.lby && pi >= 0 && pi < g.xdim &&
          pj >= 0 && pj < g.ydim) {
        const int ti = pi - a.tile_i0, tj = pj - a.tile_j0;
        if (ev.geom == 0 && ti >= 0 && ti < T && tj >= 0 && tj < T)
          atomicAdd(&s_tile[ti * T + tj], 1u);
        else
          if (ev.geom == 0 && a.spread)
            fd2::tally_add(&a.spread[(size_t)(blockIdx.x % kSpreadCopies) * g.xdim * g.ydim +
                                     (size_t)pi * g.ydim + pj], 1u);
          else
            fd2::tally_add(&g.prf[(size_t)pi * g.ydim + pj], 1ull);
      }
    }
    if (lane == 0) {
      if (ev.hist >= 1 && ev.hist <= g.num_hist && K > 0) fd2::tally_add(&g.hist[ev.hist - 1], (uint32_t)K);
      atomicAdd(&s_cnt[ev.geom * 3 + 0], (uint32_t)ev.N);
      atomicAdd(&s_cnt[ev.geom * 3 + 1], (uint32_t)K);
      atomicAdd(&s_cnt[ev.geom * 3 + 2], (uint32_t)(ev.N - K));
    }
  }
  __syncthreads();
  if (T > 0) {
    const fd2::GeomDev& g0 = geoms[0];
    for (int i = threadIdx.x; i < T2; i += blockDim.x) {
      const uint32_t v = s_tile[i];
      if (!v) continue;
      const int pi = a.tile_i0 + i / T, pj = a.tile_j0 + i % T;
      fd2::tally_add(&g0.prf[(size_t)pi * g0.ydim + pj], (unsigned long long)v);
    }
  }
  for (int gi = threadIdx.x; gi < a.n_geoms; gi += blockDim.x) {
    unsigned long long* c = a.counters + (size_t)gi * FD2_NCOUNT;
    if (s_cnt[gi * 3 + 0]) atomicAdd(&c[FD2_C_GENERATED], (unsigned long long)s_cnt[gi * 3 + 0]);
    if (s_cnt[gi * 3 + 1]) atomicAdd(&c[FD2_C_DETECT], (unsigned long long)s_cnt[gi * 3 + 1]);
    if (s_cnt[gi * 3 + 2])
      atomicAdd(&c[FD2_C_SURR_UNDETECTED], (unsigned long long)s_cnt[gi * 3 + 2]);
  }
}

// Adds the u32 PRF copies to the u64 PRF and clears them.  The PRF takes
// atomics: the Monte Carlo launch may be tallying into it at the same time.
__global__ void k_spread_flush(uint32_t* __restrict__ spread, size_t npix,
                               unsigned long long* __restrict__ prf) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= npix) return;
  unsigned long long sum = 0;
  for (int c = 0; c < kSpreadCopies; ++c) {
    sum += spread[(size_t)c * npix + i];
    spread[(size_t)c * npix + i] = 0u;
  }
  if (sum) atomicAdd(&prf[i], sum);
}

uint32_t mix32(uint32_t x) {  // integer hash (murmur3 finaliser)
  x ^= x >> 16;
  x *= 0x85EBCA6Bu;
  x ^= x >> 13;
  x *= 0xC2B2AE35u;
  x ^= x >> 16;
  return x;
}

}  // namespace fd2mlp

static int surrogate_submit(fd2_ctx* c, fd2_mlp* p, const fd2_event* events, int64_t n,
                            const fd2_surrogate_cfg* cfg_in, fd2_surrogate_stats* st) {
  if (!c || !p || n < 0 || (n > 0 && !events)) return FD2_E_INVAL;
  if (c->device != p->device) return FD2_E_INVAL;
  if (n > (int64_t)UINT32_MAX) return FD2_E_INVAL;  // network events are routed by u32 index
  fd2_surrogate_cfg cfg;
  if (cfg_in) {
    cfg = *cfg_in;
  } else {
    std::memset(&cfg, 0, sizeof cfg);
    cfg.surface_band = 5.0f;
    cfg.audit_fraction = 0.01f;
    cfg.h_min = 100.0f; cfg.h_max = 500.0f;
    cfg.r_min = 2.55f; cfg.r_max = 5.0f;
    cfg.refl_min = 0.5f; cfg.refl_max = 0.95f;
    cfg.center_band = 25.0f;
    cfg.det_extent = 1001.0f;
  }
  fd2_surrogate_stats stats;
  std::memset(&stats, 0, sizeof stats);
  const auto t0 = std::chrono::steady_clock::now();
  MLP_HIP(hipSetDevice(c->device));
  bool in_range[FD2_MAX_GEOMS];
  for (int g = 0; g < c->n_geoms; ++g) {
    const fd2_params& q = c->params[g];
    const double refl = 1.0 - q.dettop;
    // parameters the network does not take as input must be the training
    // set's (hybridmantis_amd/events.py BASE_PARAMS)
    const bool fixed = std::fabs(q.detnC - 1.8) < 1e-9 && std::fabs(q.detnIC - 1.0) < 1e-9 &&
                       std::fabs(q.detbulk - 1e-4) < 1e-12 && std::fabs(q.detbeta - 0.2) < 1e-9 &&
                       std::fabs(q.detdmin - 1.0) < 1e-9 && std::fabs(q.detdmax - 280.0) < 1e-9 &&
                       q.detsensorRefl == 0.0 &&
                       std::fabs(q.detx - cfg.det_extent) < 1e-3 &&
                       std::fabs(q.dety - cfg.det_extent) < 1e-3;
    in_range[g] = fixed && q.detheight >= cfg.h_min && q.detheight <= cfg.h_max &&
                  q.detradius >= cfg.r_min && q.detradius <= cfg.r_max &&
                  refl >= cfg.refl_min && refl <= cfg.refl_max;
  }
  // Validation and routing over contiguous ranges of `events`, on up to 8
  // host threads.  Monte Carlo and audit events (a few per cent) are
  // copied; network events are kept as indices into `events` and gathered
  // on the device.  The ranges' lists joined in order are what one pass in
  // order would give, and the first range holding an invalid event reports
  // the first invalid event.
  const int64_t base = c->next_event;
  const uint64_t audit_cut = (uint64_t)((double)cfg.audit_fraction * 4294967296.0);
  const uint32_t seed_mix = fd2mlp::mix32(cfg.seed);
  struct Routed {
    std::vector<fd2_event> mc, audit;
    std::vector<uint32_t> idx;
    int64_t photons_mc = 0, photons_surrogate = 0;
    int status = FD2_OK;
  };
  auto route_range = [&](int64_t lo, int64_t hi, Routed* r) {
    r->idx.reserve((size_t)(hi - lo));
    for (int64_t i = lo; i < hi; ++i) {
      const fd2_event& e = events[i];
      if (e.N < 0 || e.N > FD2_MAX_PHOTONS_PER_EVENT) {
        r->status = FD2_E_TOO_MANY;
        return;
      }
      if (e.geom >= (uint32_t)c->n_geoms) {
        r->status = FD2_E_INVAL;
        return;
      }
      const fd2::GeomDev& g = c->hgeom[e.geom];
      // audit hash of the 64-bit event number (its low word below 2^32)
      const uint64_t gid64 = (uint64_t)(base + i);
      const uint32_t gid = (uint32_t)gid64 ^ ((uint32_t)(gid64 >> 32) * 0x9E3779B9u);
      const bool near = e.z > g.half_h - cfg.surface_band || e.z < -g.half_h + cfg.surface_band;
      const float ox = e.x - 0.5f * g.xdet, oy = e.y - 0.5f * g.ydet;
      const bool off_centre = ox * ox + oy * oy > cfg.center_band * cfg.center_band;
      const bool off_axis = !(e.col_dx >= 0.0f && e.col_dx < g.R);  // trained on [0, R)
      if (near || off_centre || off_axis || !in_range[e.geom]) {
        r->mc.push_back(e);
        r->photons_mc += e.N;
      } else if ((uint64_t)fd2mlp::mix32(gid ^ seed_mix) < audit_cut) {
        r->audit.push_back(e);
        r->photons_mc += e.N;
      } else {
        r->idx.push_back((uint32_t)i);
        r->photons_surrogate += e.N;
      }
    }
  };
  // an allocation failure ends the range with a status, also on a worker thread
  auto route = [&](int64_t lo, int64_t hi, Routed* r) {
    try {
      route_range(lo, hi, r);
    } catch (...) {
      r->status = FD2_E_NOMEM;
    }
  };
  const int64_t per_part = (int64_t)1 << 18;
  int n_parts = (int)std::min<int64_t>(8, (n + per_part - 1) / per_part);
  const unsigned hw = std::thread::hardware_concurrency();
  if (hw > 0 && n_parts > (int)hw) n_parts = (int)hw;
  if (n_parts < 1) n_parts = 1;
  std::vector<Routed> parts((size_t)n_parts);
  {
    std::vector<std::thread> workers;
    try {
      workers.reserve((size_t)n_parts);
    } catch (...) {
    }
    for (int t = 1; t < n_parts; ++t) {
      const int64_t lo = n * t / n_parts, hi = n * (t + 1) / n_parts;
      try {
        workers.emplace_back(route, lo, hi, &parts[t]);
      } catch (...) {
        route(lo, hi, &parts[t]);
      }
    }
    route(0, n / n_parts, &parts[0]);
    for (std::thread& w : workers) w.join();
  }
  std::vector<fd2_event> mc, audit;
  int64_t ns = 0;
  for (const Routed& r : parts) {
    if (r.status) return r.status;
    mc.insert(mc.end(), r.mc.begin(), r.mc.end());
    audit.insert(audit.end(), r.audit.begin(), r.audit.end());
    ns += (int64_t)r.idx.size();
    stats.photons_mc += r.photons_mc;
    stats.photons_surrogate += r.photons_surrogate;
  }
  // network events per sampling launch: fewer than 2^32 photons (the u32
  // PRF tallies of a launch cannot wrap) and at most 2^30 events (int
  // indexing in the kernel); slice_start holds each launch's first index
  std::vector<int64_t> slice_start;
  if (ns > 0 && (uint64_t)stats.photons_surrogate <= 0xFFFFFFFFull &&
      ns <= ((int64_t)1 << 30)) {
    slice_start.push_back(0);
  } else if (ns > 0) {
    uint64_t slice_photons = 0;
    int64_t k = 0;
    for (const Routed& r : parts)
      for (const uint32_t i : r.idx) {
        const uint64_t N = (uint64_t)events[i].N;
        if (slice_start.empty() || slice_photons + N > 0xFFFFFFFFull ||
            k - slice_start.back() >= ((int64_t)1 << 30)) {
          slice_start.push_back(k);
          slice_photons = 0;
        }
        slice_photons += N;
        ++k;
      }
  }
  const int64_t n_mc = (int64_t)mc.size(), n_audit = (int64_t)audit.size();
  stats.n_events = n;
  stats.n_mc = n_mc;
  stats.n_audit = n_audit;
  stats.n_surrogate = ns;

  fd2_event* d_all = nullptr;
  fd2_event* d_ev = nullptr;
  uint32_t* d_idx = nullptr;
  uint64_t* d_ids = nullptr;
  uint32_t* d_rad = nullptr;
  float* d_logits = nullptr;
  uint32_t* d_spread = nullptr;
  float* d_feat = nullptr;  // audit features and predictions
  float* d_pred = nullptr;
  float* h_pred = nullptr;  // pinned
  std::vector<float> feat;
  // waits for both streams before returning; the buffers stay with the
  // network for its next call
  auto release = [&](int status) {
    if (hipStreamSynchronize(c->stream) != hipSuccess && !status) status = FD2_E_DEVICE;
    if (hipStreamSynchronize(p->stream) != hipSuccess && !status) status = FD2_E_DEVICE;
    return status;
  };

  // Monte Carlo part: the fallback events, then the audit events, in one
  // launch with per-event radial tallies (only the audit slice is read),
  // on the context's stream.  The events keep the numbers that separate
  // submits of the two sets would give them.
  int rc = FD2_OK;
  mc.insert(mc.end(), audit.begin(), audit.end());
  if (!mc.empty()) {
    if (!reserve(p, kBufRad, mc.size() * FD2_RADIAL_BINS * sizeof(uint32_t), &d_rad))
      return release(FD2_E_NOMEM);
    rc = fd2::submit_radial_async(c, mc.data(), (int64_t)mc.size(), d_rad);
    if (rc) return release(rc);
  }

  // The network's prediction for the audit events needs no Monte Carlo
  // result: it goes first on the network's stream, with its copy to pinned
  // host memory, so that the end of the call only reads it.
  if (n_audit > 0) {
    const int chunk = 65536;
    rc = ensure_batch(p, chunk);
    if (rc) return release(rc);
    feat.resize((size_t)n_audit * kFeat);
    for (int64_t i = 0; i < n_audit; ++i) {
      const fd2_event& e = audit[i];
      const fd2::GeomDev& g = c->hgeom[e.geom];
      feat[i * kFeat + 0] = (g.half_h - e.z) / g.H;
      feat[i * kFeat + 1] = e.col_dx / g.R;
      feat[i * kFeat + 2] = g.H * (1.0f / 500.0f);
      feat[i * kFeat + 3] = g.R * (1.0f / 5.0f);
      feat[i * kFeat + 4] = 1.0f - g.top_absfrac;
    }
    const size_t pred_bytes = (size_t)n_audit * kTarget * sizeof(float);
    if (!reserve(p, kBufFeat, feat.size() * sizeof(float), &d_feat) ||
        !reserve(p, kBufPred, pred_bytes, &d_pred))
      return release(FD2_E_NOMEM);
    if (p->h_pred_cap < pred_bytes) {
      if (p->h_pred) (void)hipHostFree(p->h_pred);
      p->h_pred = nullptr;
      p->h_pred_cap = 0;
      if (hipHostMalloc(&p->h_pred, pred_bytes, hipHostMallocDefault) != hipSuccess)
        return release(FD2_E_NOMEM);
      p->h_pred_cap = pred_bytes;
    }
    h_pred = p->h_pred;
    if (hipMemcpyAsync(d_feat, feat.data(), feat.size() * sizeof(float), hipMemcpyHostToDevice,
                       p->stream) != hipSuccess)
      return release(FD2_E_DEVICE);
    for (int64_t s0 = 0; s0 < n_audit && rc == FD2_OK; s0 += chunk) {
      const int b = (int)(n_audit - s0 < chunk ? n_audit - s0 : chunk);
      hipLaunchKernelGGL(k_pack_inputs, dim3((b * kIn + 255) / 256), dim3(256), 0, p->stream,
                         d_feat + (size_t)s0 * kFeat, b, p->x0);
      rc = forward_fused(p, b, p->stream);
      if (rc) break;
      hipLaunchKernelGGL(k_head_probs, dim3((b * 64 + 255) / 256), dim3(256), 0, p->stream,
                         p->z[3], b, d_pred + (size_t)s0 * kTarget);
    }
    if (rc) return release(rc);
    if (hipMemcpyAsync(h_pred, d_pred, pred_bytes, hipMemcpyDeviceToHost, p->stream) !=
        hipSuccess)
      return release(FD2_E_DEVICE);
  }

  // Network part, on the network's own stream: its kernels take the compute
  // units that the Monte Carlo launch frees as it drains.  The forward runs
  // in batches into one logits array; one sampling launch then covers all
  // network events (one per 2^32 photons), so that each workgroup flushes
  // its LDS PRF tile once.
  bool timed_net = false;
  if (ns > 0) {
    const int chunk = 65536;
    rc = ensure_batch(p, chunk);
    if (rc) return release(rc);
    if (!reserve(p, kBufAll, (size_t)n * sizeof(fd2_event), &d_all) ||
        !reserve(p, kBufEv, (size_t)ns * sizeof(fd2_event), &d_ev) ||
        !reserve(p, kBufIdx, (size_t)ns * 4, &d_idx) ||
        !reserve(p, kBufIds, (size_t)ns * sizeof(uint64_t), &d_ids) ||
        !reserve(p, kBufLogits, (size_t)ns * kOutPad * sizeof(float), &d_logits))
      return release(FD2_E_NOMEM);
    const size_t npix0 = (size_t)c->hgeom[0].xdim * c->hgeom[0].ydim;
    if (!reserve(p, kBufSpread, npix0 * fd2mlp::kSpreadCopies * sizeof(uint32_t), &d_spread))
      return release(FD2_E_NOMEM);
    const fd2::GeomDev& g0 = c->hgeom[0];
    const int dev_cu = c->n_cu > 0 ? c->n_cu : 1;
    // a sampler shape: tile side, dynamic LDS, workgroups per CU
    struct Shape {
      int tile = 0, per_cu = 0;
      size_t smem = 0;
    };
    auto shape = [&](int side, int threads, const void* kernel, Shape* sh) {
      int t = side < g0.xdim ? side : g0.xdim;
      if (t > g0.ydim) t = g0.ydim;
      sh->tile = t;
      sh->smem = (size_t)t * t * 4;
      if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&sh->per_cu, kernel, threads, sh->smem) !=
              hipSuccess ||
          sh->per_cu < 1)
        return false;
      if (sh->per_cu > 8) sh->per_cu = 8;
      return true;
    };
    Shape small, large;
    if (!shape(fd2mlp::kSampleTile, fd2mlp::kSampleThreads,
               (const void*)fd2mlp::k_surr_sample<fd2mlp::kSampleThreads>, &small) ||
        !shape(fd2mlp::kSampleTileLarge, fd2mlp::kSampleThreadsLarge,
               (const void*)fd2mlp::k_surr_sample<fd2mlp::kSampleThreadsLarge>, &large))
      return release(FD2_E_DEVICE);
    if (hipEventRecord(p->e0, p->stream) != hipSuccess ||
        hipMemcpyAsync(d_all, events, (size_t)n * sizeof(fd2_event), hipMemcpyHostToDevice,
                       p->stream) != hipSuccess)
      return release(FD2_E_DEVICE);
    int64_t off = 0;
    for (const Routed& r : parts) {
      if (!r.idx.empty() && hipMemcpyAsync(d_idx + off, r.idx.data(), r.idx.size() * 4,
                                           hipMemcpyHostToDevice, p->stream) != hipSuccess)
        return release(FD2_E_DEVICE);
      off += (int64_t)r.idx.size();
    }
    if (fd2::launch_fill_u32(d_spread, npix0 * fd2mlp::kSpreadCopies, p->stream) != hipSuccess)
      return release(FD2_E_DEVICE);
    hipLaunchKernelGGL(fd2mlp::k_surr_gather, dim3((unsigned)((ns + 255) / 256)), dim3(256), 0,
                       p->stream, d_all, d_idx, ns, (uint64_t)base, d_ev, d_ids);
    for (int64_t s0 = 0; s0 < ns && rc == FD2_OK; s0 += chunk) {
      const int b = (int)(ns - s0 < chunk ? ns - s0 : chunk);
      hipLaunchKernelGGL(fd2mlp::k_surr_features, dim3((b * kInPad + 255) / 256), dim3(256), 0,
                         p->stream, d_ev + s0, b, c->d_geoms, p->x0);
      rc = forward_fused(p, b, p->stream, d_logits + (size_t)s0 * kOutPad);
    }
    for (size_t k = 0; k < slice_start.size() && rc == FD2_OK; ++k) {
      const int64_t s0 = slice_start[k];
      const int64_t s1 = k + 1 < slice_start.size() ? slice_start[k + 1] : ns;
      const int b = (int)(s1 - s0);
      fd2mlp::SampleArgs a;
      a.events = d_ev + s0;
      a.ids = d_ids + s0;
      a.logits = d_logits + (size_t)s0 * kOutPad;
      a.geoms = c->d_geoms;
      a.counters = c->d_counters;
      a.n = b;
      a.n_geoms = c->n_geoms;
      a.seed = cfg.seed;
      const int large_waves = dev_cu * large.per_cu * (fd2mlp::kSampleThreadsLarge / 64);
      const bool big = (int64_t)b >= (int64_t)large_waves * fd2mlp::kSampleLargeEventsPerWave;
      const Shape& sh = big ? large : small;
      const int threads = big ? fd2mlp::kSampleThreadsLarge : fd2mlp::kSampleThreads;
      a.tile = sh.tile;
      a.tile_i0 = g0.xdim / 2 - sh.tile / 2;
      a.tile_j0 = g0.ydim / 2 - sh.tile / 2;
      a.spread = d_spread;
      const int waves_per_block = threads / 64;
      int64_t blocks = (int64_t)dev_cu * sh.per_cu;
      if (blocks * waves_per_block > b) blocks = (b + waves_per_block - 1) / waves_per_block;
      if (big)
        hipLaunchKernelGGL(fd2mlp::k_surr_sample<fd2mlp::kSampleThreadsLarge>,
                           dim3((unsigned)blocks), dim3(threads), sh.smem, p->stream, a, a.events,
                           a.ids, a.logits, (const fd2::GeomDev*)a.geoms);
      else
        hipLaunchKernelGGL(fd2mlp::k_surr_sample<fd2mlp::kSampleThreads>, dim3((unsigned)blocks),
                           dim3(threads), sh.smem, p->stream, a, a.events, a.ids, a.logits,
                           (const fd2::GeomDev*)a.geoms);
      hipLaunchKernelGGL(fd2mlp::k_spread_flush, dim3((unsigned)((npix0 + 255) / 256)), dim3(256),
                         0, p->stream, d_spread, npix0, c->d_prf[0]);
      if (hipGetLastError() != hipSuccess) rc = FD2_E_DEVICE;
    }
    if (rc) return release(rc);
    if (hipEventRecord(p->e1, p->stream) != hipSuccess) return release(FD2_E_DEVICE);
    timed_net = true;
  }

  // the Monte Carlo part's end; the audit slice of its radial tallies
  std::vector<uint32_t> rad((size_t)n_audit * FD2_RADIAL_BINS);
  if (!mc.empty()) {
    rc = fd2_sync(c);
    if (!rc && n_audit > 0 &&
        (hipMemcpyAsync(rad.data(), d_rad + (size_t)n_mc * FD2_RADIAL_BINS,
                        rad.size() * sizeof(uint32_t), hipMemcpyDeviceToHost,
                        c->stream) != hipSuccess ||
         hipStreamSynchronize(c->stream) != hipSuccess))
      rc = FD2_E_DEVICE;
    if (rc) return release(rc);
  }
  // the audit comparison, once the network's stream is done
  if (n_audit > 0) {
    if (hipStreamSynchronize(p->stream) != hipSuccess) return release(FD2_E_DEVICE);
    const float* pred = h_pred;
    std::vector<double> pm(FD2_RADIAL_BINS, 0.0), pp(FD2_RADIAL_BINS, 0.0);
    for (int64_t i = 0; i < n_audit; ++i) {
      double det = 0.0;
      for (int b = 0; b < FD2_RADIAL_BINS; ++b) {
        det += rad[i * FD2_RADIAL_BINS + b];
        pm[b] += rad[i * FD2_RADIAL_BINS + b];
      }
      const double nd = (double)audit[i].N * pred[i * kTarget];
      stats.audit_detected_mc += det;
      stats.audit_detected_pred += nd;
      for (int b = 0; b < FD2_RADIAL_BINS; ++b) pp[b] += nd * pred[i * kTarget + 1 + b];
    }
    double sm = 0, sp = 0, l1 = 0;
    for (int b = 0; b < FD2_RADIAL_BINS; ++b) {
      sm += pm[b];
      sp += pp[b];
    }
    for (int b = 0; b < FD2_RADIAL_BINS; ++b)
      l1 += std::fabs((sm > 0 ? pm[b] / sm : 0) - (sp > 0 ? pp[b] / sp : 0));
    stats.audit_radial_l1 = l1;
  }
  rc = release(FD2_OK);
  if (rc) return rc;
  const auto t1 = std::chrono::steady_clock::now();
  c->next_event = base + n > c->next_event ? base + n : c->next_event;
  float ms = 0.0f;
  if (!mc.empty() && hipEventElapsedTime(&ms, c->ev_start, c->ev_stop) == hipSuccess)
    stats.ms_mc = ms;
  if (timed_net && hipEventElapsedTime(&ms, p->e0, p->e1) == hipSuccess) stats.ms_surrogate = ms;
  stats.ms_total = std::chrono::duration<double, std::milli>(t1 - t0).count();
  if (st) *st = stats;
  return FD2_OK;
}

// host allocation failures come back as a status, not an exception
extern "C" int fd2_surrogate_submit(fd2_ctx* c, fd2_mlp* p, const fd2_event* events,
                                    int64_t n, const fd2_surrogate_cfg* cfg_in,
                                    fd2_surrogate_stats* st) {
  try {
    return surrogate_submit(c, p, events, n, cfg_in, st);
  } catch (...) {
    return FD2_E_NOMEM;
  }
}
