// guide.hip — the guiding network on MFMA: inference, fused forward + backward
// with MFMA weight gradients, Adam, and the C ABI of the guide object.
//
// Layout on the matrix cores (v_mfma_f32_16x16x32_bf16 of gfx950, 64 lanes;
// lane l, s = l % 16, g = l / 16):
//   A fragment (16 x 32): lane holds A[s][k] for the 8 k of lane group g
//   B fragment (32 x 16): lane holds B[k][s] for the same 8 k
//   D (16 x 16, FP32):    lane holds D[4g .. 4g+3][s]
// The forward pass computes Z^T = W X^T (weights as A, a 16-sample tile of
// activations as B).  A lane group's 8 k are rows 4g .. 4g+3 of two 16-row D
// tiles (the fragment k order of guide.h's frag_elem), so the D tiles of layer
// l, two at a time, are the B fragments of layer l+1: activations stay in
// registers between layers.  The backward pass does the same with W^T
// fragments.  Weight gradients dW = dZ H^T take the batch as the contraction
// dimension; dZ and H are staged in LDS as [feature][sample] BF16 and read back
// as A / B fragments of 8 consecutive samples per lane group.
#include <hip/hip_runtime.h>

#include <chrono>
#include <cmath>
#include <cstring>
#include <memory>
#include <random>
#include <vector>

#include "guide_host.h"

using namespace mupsi;
using namespace mupsi::guide;

namespace {

typedef __bf16 vbf16x8 __attribute__((ext_vector_type(8)));
__device__ __forceinline__ f32x4 mfma(bf16x8 a, bf16x8 b, f32x4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(vbf16x8, a),
                                                 __builtin_bit_cast(vbf16x8, b), c, 0, 0, 0);
}
// Two 16-row tiles as one 32-wide k operand (guide.h's fragment k order).
__device__ __forceinline__ bf16x8 cat(bf16x4 lo, bf16x4 hi) {
  return bf16x8{lo[0], lo[1], lo[2], lo[3], hi[0], hi[1], hi[2], hi[3]};
}

__device__ __forceinline__ uint32_t mix32(uint32_t x) {
  x ^= x >> 16;
  x *= 0x7feb352du;
  x ^= x >> 15;
  x *= 0x846ca68bu;
  x ^= x >> 16;
  return x;
}

// One hidden layer of the forward pass: 64 outputs from NK 16-row input tiles
// (NK / 2 k-tiles of 32).
template <int NK>
__device__ __forceinline__ void hidden_layer(const bf16x8* __restrict__ wf, int frag0,
                                             const float* __restrict__ bias, const bf16x4* in,
                                             bf16x4* out, int lane) {
  const int g = lane >> 4;
  constexpr int KT = NK / 2;
#pragma unroll
  for (int mt = 0; mt < 4; ++mt) {
    f32x4 z = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int kt = 0; kt < KT; ++kt)
      z = mfma(wf[(frag0 + mt * KT + kt) * 64 + lane], cat(in[2 * kt], in[2 * kt + 1]), z);
    const int r = 16 * mt + 4 * g;
    out[mt] = pack4(fmaxf(z[0] + bias[r], 0.f), fmaxf(z[1] + bias[r + 1], 0.f),
                    fmaxf(z[2] + bias[r + 2], 0.f), fmaxf(z[3] + bias[r + 3], 0.f));
  }
}

// Output layer: 3 tiles of 16 rows (rows 41..47 are padding and read as 0).
__device__ __forceinline__ void output_layer(const bf16x8* __restrict__ wf,
                                             const float* __restrict__ b4, const bf16x4* h3,
                                             f32x4* o, int lane) {
  const int g = lane >> 4;
#pragma unroll
  for (int mt = 0; mt < 3; ++mt) {
    f32x4 z = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int kt = 0; kt < 2; ++kt)
      z = mfma(wf[(kFragW4 + mt * 2 + kt) * 64 + lane], cat(h3[2 * kt], h3[2 * kt + 1]), z);
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int r = 16 * mt + 4 * g + i;
      o[mt][i] = r < kOut ? z[i] + b4[r] : 0.0f;
    }
  }
}

// The network as a block keeps it in LDS: the forward fragments, the
// transposed fragments (training only) and the biases b1 | b2 | b3 | b4.  A
// layer then waits on LDS, not on a global load, for its weights.
constexpr int kBiasLds = 3 * kHid + kOutPad;
struct NetLds {
  bf16x8 wf[kNumFragF * 64];
  float bias[kBiasLds];
};
struct NetLdsT {
  bf16x8 wf[kNumFragF * 64];
  bf16x8 wt[kNumFragT * 64];
  float bias[kBiasLds];
};
__device__ __forceinline__ void stage_bias(const NetDev& net, float* bias) {
  for (int i = threadIdx.x; i < kBiasLds; i += blockDim.x) {
    const int l = i / kHid, r = i % kHid;
    bias[i] = l < 3 ? net.theta[(l == 0 ? kOffB1 : l == 1 ? kOffB2 : kOffB3) + r]
                    : (r < kOut ? net.theta[kOffB4 + r] : 0.0f);
  }
}
__device__ __forceinline__ void stage_frags(const bf16x8* src, bf16x8* dst, int n) {
  for (int i = threadIdx.x; i < n; i += blockDim.x) dst[i] = src[i];
}

__device__ __forceinline__ void forward_tile(const bf16x8* wf, const float* bias, bf16x4 x0,
                                             bf16x4 x1, bf16x4* h1, bf16x4* h2, bf16x4* h3,
                                             f32x4* o, int lane) {
  const bf16x4 x[2] = {x0, x1};
  hidden_layer<2>(wf, kFragW1, bias, x, h1, lane);
  hidden_layer<4>(wf, kFragW2, bias + kHid, h1, h2, lane);
  hidden_layer<4>(wf, kFragW3, bias + 2 * kHid, h2, h3, lane);
  output_layer(wf, bias + 3 * kHid, h3, o, lane);
}

__device__ __forceinline__ float4 ld4(const float* p) { return *reinterpret_cast<const float4*>(p); }

// --------------------------------------------------------------- inference
struct InferArgs {
  const float* feats;
  const int* count_dev;
  int count_host;
  const int* q_pid;
  float* raw;   // n x kOut, or null
  float* mix;   // kMix floats per output row, or null
  float alpha_max;
};

__global__ __launch_bounds__(256) void k_mlp_infer(NetDev net, InferArgs a) {
  __shared__ float Ow[4][kOutPad][17];
  __shared__ NetLds N;
  stage_frags(net.wf, N.wf, kNumFragF * 64);
  stage_bias(net, N.bias);
  __syncthreads();
  const int n = a.count_dev ? *a.count_dev : a.count_host;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63, s = lane & 15, g = lane >> 4;
  // every wave of a block runs the same number of iterations (block tiles of 64
  // rows), so the block barriers below are uniform
  for (int bt = blockIdx.x; bt * 64 < n; bt += gridDim.x) {
    const int row = bt * 64 + wave * 16 + s;
    const bool valid = row < n;
    float4 f0 = {0, 0, 0, 0}, f1 = {0, 0, 0, 0};
    if (valid) {
      f0 = ld4(a.feats + (size_t)row * kIn + 4 * g);
      f1 = ld4(a.feats + (size_t)row * kIn + 16 + 4 * g);
    }
    bf16x4 h1[4], h2[4], h3[4];
    f32x4 o[3];
    forward_tile(N.wf, N.bias, pack4(f0.x, f0.y, f0.z, f0.w), pack4(f1.x, f1.y, f1.z, f1.w), h1,
                 h2, h3, o, lane);
#pragma unroll
    for (int mt = 0; mt < 3; ++mt)
#pragma unroll
      for (int i = 0; i < 4; ++i) Ow[wave][16 * mt + 4 * g + i][s] = o[mt][i];
    __syncthreads();
    if (g == 0 && valid) {
      float ov[kOut];
#pragma unroll
      for (int r = 0; r < kOut; ++r) ov[r] = Ow[wave][r][s];
      if (a.raw) {
        float* dst = a.raw + (size_t)row * kOut;
#pragma unroll
        for (int r = 0; r < kOut; ++r) dst[r] = ov[r];
      }
      if (a.mix) {
        float m[kMix];
        raw_to_mixture(ov, a.alpha_max, m);
        const int pid = a.q_pid ? a.q_pid[row] : row;
        float4* dst = reinterpret_cast<float4*>(a.mix + (size_t)pid * kMix);
#pragma unroll
        for (int r = 0; r < kMix / 4; ++r) dst[r] = make_float4(m[4 * r], m[4 * r + 1], m[4 * r + 2], m[4 * r + 3]);
      }
    }
    __syncthreads();
  }
}

// --------------------------------------------------------------- training
// LDS rows ([feature][sample], 64 samples + 8 padding per row, BF16).
constexpr int kLdsStride = 72;
constexpr int kRowX = 0, kRowH1 = 32, kRowH2 = 96, kRowH3 = 160;
constexpr int kRowD1 = 224, kRowD2 = 288, kRowD3 = 352, kRowD4 = 416, kRowsLds = 464;
constexpr int kMaxTrainBlocks = 256;

struct TrainArgs {
  const float* feats;
  const float* dirs;
  const float* w;
  const int* n_dev;
  int n_host;
  int batch;      // samples in this step, multiple of 64
  int random;     // 1: draw records uniformly with the hash; 0: sample i is record i
  uint32_t seed;
  float* grad_part;  // gridDim.x x kNumParams
  float* loss_part;  // gridDim.x
  float alpha_max;
};

// The 52 weight-gradient output tiles: (A rows in LDS, B rows in LDS, param
// offset, in_dim, out_dim, mo, mi).
struct GradTile {
  int arow, brow, off, in_dim, out_dim;
};
__device__ __forceinline__ GradTile grad_tile(int t) {
  GradTile r;
  if (t < 8) {  // dW1 = dZ1 X^T, 64 x 32
    r.arow = kRowD1 + 16 * (t >> 1);
    r.brow = kRowX + 16 * (t & 1);
    r.off = kOffW1 + 16 * (t >> 1) * kIn + 16 * (t & 1);
    r.in_dim = kIn;
    r.out_dim = 16;
  } else if (t < 40) {  // dW2 = dZ2 H1^T, dW3 = dZ3 H2^T
    const int l = t < 24 ? 0 : 1;
    const int u = t - (l ? 24 : 8);
    r.arow = (l ? kRowD3 : kRowD2) + 16 * (u >> 2);
    r.brow = (l ? kRowH2 : kRowH1) + 16 * (u & 3);
    r.off = (l ? kOffW3 : kOffW2) + 16 * (u >> 2) * kHid + 16 * (u & 3);
    r.in_dim = kHid;
    r.out_dim = 16;
  } else {  // dW4 = dO H3^T, 41 x 64
    const int u = t - 40;
    r.arow = kRowD4 + 16 * (u >> 2);
    r.brow = kRowH3 + 16 * (u & 3);
    r.off = kOffW4 + 16 * (u >> 2) * kHid + 16 * (u & 3);
    r.in_dim = kHid;
    r.out_dim = (u >> 2) == 2 ? kOut - 32 : 16;  // rows 41..47 do not exist
  }
  return r;
}

// 8 consecutive BF16 of an LDS row (16-byte aligned: col is a multiple of 8 and
// a row is 144 bytes).
__device__ __forceinline__ bf16x8 lds_frag(const uint16_t (*L)[kLdsStride], int row, int col) {
  return *reinterpret_cast<const bf16x8*>(&L[row][col]);
}
__device__ __forceinline__ void lds_put4(uint16_t (*L)[kLdsStride], int row0, int col, bf16x4 v) {
#pragma unroll
  for (int i = 0; i < 4; ++i) L[row0 + i][col] = (uint16_t)v[i];
}

// Backward through one hidden layer: dZ_prev = (W^T dZ) masked by h_prev != 0;
// NK 16-row tiles of dZ (the last k-tile of 32 is padded with zeros for odd NK).
template <int NK>
__device__ __forceinline__ void backward_layer(const bf16x8* __restrict__ wt, int frag0,
                                               const bf16x4* dz, const bf16x4* hprev,
                                               bf16x4* dzprev, int lane) {
  constexpr int KT = (NK + 1) / 2;
  const bf16x4 zero4 = {0, 0, 0, 0};
#pragma unroll
  for (int mt = 0; mt < 4; ++mt) {
    f32x4 z = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int kt = 0; kt < KT; ++kt)
      z = mfma(wt[(frag0 + mt * KT + kt) * 64 + lane],
               cat(dz[2 * kt], 2 * kt + 1 < NK ? dz[2 * kt + 1] : zero4), z);
    dzprev[mt] = pack4(hprev[mt][0] != 0 ? z[0] : 0.f, hprev[mt][1] != 0 ? z[1] : 0.f,
                       hprev[mt][2] != 0 ? z[2] : 0.f, hprev[mt][3] != 0 ? z[3] : 0.f);
  }
}

__global__ __launch_bounds__(256) void k_mlp_train(NetDev net, TrainArgs a) {
  __shared__ __align__(16) uint16_t L[kRowsLds][kLdsStride];
  __shared__ float Ow[4][kOutPad][17];
  __shared__ float lsum[4];
  __shared__ NetLdsT N;
  stage_frags(net.wf, N.wf, kNumFragF * 64);
  stage_frags(net.wt, N.wt, kNumFragT * 64);
  stage_bias(net, N.bias);
  __syncthreads();
  const int n = a.n_dev ? *a.n_dev : a.n_host;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63, s = lane & 15, g = lane >> 4;
  const int col = wave * 16 + s;
  f32x4 acc[13];
#pragma unroll
  for (int t = 0; t < 13; ++t) acc[t] = f32x4{0.f, 0.f, 0.f, 0.f};
  float bacc = 0.f, lacc = 0.f;
  const int iters = a.batch >> 6;
  for (int it = blockIdx.x; it < iters; it += gridDim.x) {
    const int si = it * 64 + col;
    int rec;
    bool valid;
    if (a.random) {
      valid = n > 0;
      rec = valid ? (int)(mix32(a.seed ^ mix32((uint32_t)si * 0x9e3779b9u + 0x632be5abu)) % (uint32_t)n) : 0;
    } else {
      rec = si;
      valid = si < n;
    }
    float4 f0 = {0, 0, 0, 0}, f1 = {0, 0, 0, 0};
    if (valid) {
      f0 = ld4(a.feats + (size_t)rec * kIn + 4 * g);
      f1 = ld4(a.feats + (size_t)rec * kIn + 16 + 4 * g);
    }
    const bf16x4 x0 = pack4(f0.x, f0.y, f0.z, f0.w), x1 = pack4(f1.x, f1.y, f1.z, f1.w);
    lds_put4(L, kRowX + 4 * g, col, x0);
    lds_put4(L, kRowX + 16 + 4 * g, col, x1);
    bf16x4 h1[4], h2[4], h3[4];
    f32x4 o[3];
    forward_tile(N.wf, N.bias, x0, x1, h1, h2, h3, o, lane);
#pragma unroll
    for (int mt = 0; mt < 4; ++mt) {
      lds_put4(L, kRowH1 + 16 * mt + 4 * g, col, h1[mt]);
      lds_put4(L, kRowH2 + 16 * mt + 4 * g, col, h2[mt]);
      lds_put4(L, kRowH3 + 16 * mt + 4 * g, col, h3[mt]);
    }
#pragma unroll
    for (int mt = 0; mt < 3; ++mt)
#pragma unroll
      for (int i = 0; i < 4; ++i) Ow[wave][16 * mt + 4 * g + i][s] = o[mt][i];
    __syncthreads();
    if (g == 0) {  // per-sample mixture loss and its gradient w.r.t. the 40 outputs
      float ov[kOut], d[kOut];
#pragma unroll
      for (int r = 0; r < kOut; ++r) ov[r] = Ow[wave][r][s];
      float l = 0.f;
      if (valid) {
        const float* dp = a.dirs + (size_t)rec * 3;
        const float* np = a.feats + (size_t)rec * kIn + 24;  // the vertex normal
        const float cosn = fmaf(np[0], dp[0], fmaf(np[1], dp[1], np[2] * dp[2]));
        l = mixture_loss_grad(ov, dp[0], dp[1], dp[2], cosn, a.w[rec], a.alpha_max, d);
      } else {
#pragma unroll
        for (int r = 0; r < kOut; ++r) d[r] = 0.f;
      }
      lacc += l;
#pragma unroll
      for (int r = 0; r < kOut; ++r) Ow[wave][r][s] = d[r];
#pragma unroll
      for (int r = kOut; r < kOutPad; ++r) Ow[wave][r][s] = 0.f;
    }
    __syncthreads();
    bf16x4 dO[3];
#pragma unroll
    for (int kt = 0; kt < 3; ++kt) {
      const int r = 16 * kt + 4 * g;
      dO[kt] = pack4(Ow[wave][r][s], Ow[wave][r + 1][s], Ow[wave][r + 2][s], Ow[wave][r + 3][s]);
      lds_put4(L, kRowD4 + r, col, dO[kt]);
    }
    bf16x4 d3[4], d2[4], d1[4];
    backward_layer<3>(N.wt, kFragT4, dO, h3, d3, lane);
    backward_layer<4>(N.wt, kFragT3, d3, h2, d2, lane);
    backward_layer<4>(N.wt, kFragT2, d2, h1, d1, lane);
#pragma unroll
    for (int mt = 0; mt < 4; ++mt) {
      lds_put4(L, kRowD3 + 16 * mt + 4 * g, col, d3[mt]);
      lds_put4(L, kRowD2 + 16 * mt + 4 * g, col, d2[mt]);
      lds_put4(L, kRowD1 + 16 * mt + 4 * g, col, d1[mt]);
    }
    __syncthreads();
    // weight gradients over the 64 samples of this tile: this wave's 13 tiles
#pragma unroll
    for (int tt = 0; tt < 13; ++tt) {
      const GradTile gt = grad_tile(wave + 4 * tt);
#pragma unroll
      for (int sg = 0; sg < 2; ++sg) {  // samples 32 sg + 8 g .. + 7 of the tile
        const bf16x8 A = lds_frag(L, gt.arow + s, 32 * sg + 8 * g);
        const bf16x8 B = lds_frag(L, gt.brow + s, 32 * sg + 8 * g);
        acc[tt] = mfma(A, B, acc[tt]);
      }
    }
    // bias gradients: rows D1 (64), D2 (64), D3 (64), D4 (41)
    if (threadIdx.x < 3 * kHid + kOut) {  // 8 reads of 8 BF16 (in column order)
      const int t = threadIdx.x;
      const int row = t < 192 ? kRowD1 + t : kRowD4 + (t - 192);
      float sacc = 0.f;
#pragma unroll
      for (int c = 0; c < 64; c += 8) {
        const bf16x8 v = lds_frag(L, row, c);
#pragma unroll
        for (int e = 0; e < 8; ++e) sacc += bf2f((uint16_t)v[e]);
      }
      bacc += sacc;
    }
    __syncthreads();
  }
  float* gp = a.grad_part + (size_t)blockIdx.x * kNumParams;
#pragma unroll
  for (int tt = 0; tt < 13; ++tt) {
    const GradTile gt = grad_tile(wave + 4 * tt);
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int r = 4 * g + i;
      if (r < gt.out_dim) gp[gt.off + r * gt.in_dim + s] = acc[tt][i];
    }
  }
  if (threadIdx.x < 3 * kHid + kOut) {
    const int t = threadIdx.x;
    const int off = t < 64 ? kOffB1 + t : t < 128 ? kOffB2 + t - 64 : t < 192 ? kOffB3 + t - 128
                                                                             : kOffB4 + t - 192;
    gp[off] = bacc;
  }
  for (int o = 32; o > 0; o >>= 1) lacc += __shfl_xor(lacc, o);
  if (lane == 0) lsum[wave] = lacc;
  __syncthreads();
  if (threadIdx.x == 0) a.loss_part[blockIdx.x] = lsum[0] + lsum[1] + lsum[2] + lsum[3];
}

// --------------------------------------------------------------- parameters
__device__ __forceinline__ void store_frags(int p, float th, uint16_t* wf, uint16_t* wt) {
  const uint16_t b = f2bf(th);
  if (p < kOffB1) {
    wf[frag_elem(kFragW1, 1, p / kIn, p % kIn)] = b;
  } else if (p >= kOffW2 && p < kOffB2) {
    const int q = p - kOffW2, r = q / kHid, c = q % kHid;
    wf[frag_elem(kFragW2, 2, r, c)] = b;
    wt[frag_elem(kFragT2, 2, c, r)] = b;
  } else if (p >= kOffW3 && p < kOffB3) {
    const int q = p - kOffW3, r = q / kHid, c = q % kHid;
    wf[frag_elem(kFragW3, 2, r, c)] = b;
    wt[frag_elem(kFragT3, 2, c, r)] = b;
  } else if (p >= kOffW4 && p < kOffB4) {
    const int q = p - kOffW4, r = q / kHid, c = q % kHid;
    wf[frag_elem(kFragW4, 2, r, c)] = b;
    wt[frag_elem(kFragT4, 2, c, r)] = b;
  }
}

__global__ void k_pack(const float* theta, uint16_t* wf, uint16_t* wt) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p < kNumParams) store_frags(p, theta[p], wf, wt);
}

// Sum of the per-block partial gradients, times inv_n, by kRedLanes lanes per
// parameter: lane j adds the partials of blocks j, j + kRedLanes, ... in block
// order (loads issued 8 at a time ahead of the additions), then an xor
// butterfly over the lanes gives every lane the same total.  The order is
// fixed, so a gradient's bits do not depend on the run.  Every lane of a
// kRedLanes group must call it (p >= kNumParams reads nothing).
constexpr int kRedLanes = 8;
__device__ __forceinline__ float reduce_grad(const float* part, int nparts, int p, int j,
                                             float inv_n) {
  constexpr int kBatch = 8;
  float gsum = 0.f;
  if (p < kNumParams) {
    int b = j;
    for (; b + (kBatch - 1) * kRedLanes < nparts; b += kBatch * kRedLanes) {
      float v[kBatch];
#pragma unroll
      for (int i = 0; i < kBatch; ++i) v[i] = part[(size_t)(b + i * kRedLanes) * kNumParams + p];
#pragma unroll
      for (int i = 0; i < kBatch; ++i) gsum += v[i];
    }
    for (; b < nparts; b += kRedLanes) gsum += part[(size_t)b * kNumParams + p];
  }
  static_assert(kRedLanes == 8, "three butterfly steps");
  gsum += __shfl_xor(gsum, 1);
  gsum += __shfl_xor(gsum, 2);
  gsum += __shfl_xor(gsum, 4);
  return gsum * inv_n;
}

__global__ void k_grad_reduce(const float* part, int nparts, float inv_n, float* out) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const int p = t / kRedLanes, j = t % kRedLanes;
  const float gr = reduce_grad(part, nparts, p, j, inv_n);
  if (j == 0 && p < kNumParams) out[p] = gr;
}

// Adam (beta1 0.9, beta2 0.999, eps 1e-8) on the reduced minibatch gradient, then
// the BF16 fragments of the updated parameter (one parameter per kRedLanes lanes).
__global__ void k_adam(float* theta, float* m, float* v, uint16_t* wf, uint16_t* wt,
                       const float* part, int nparts, float inv_n, float lr, float bc1, float bc2) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const int p = t / kRedLanes, j = t % kRedLanes;
  const float gr = reduce_grad(part, nparts, p, j, inv_n);
  if (j != 0 || p >= kNumParams) return;
  const float mm = 0.9f * m[p] + 0.1f * gr;
  const float vv = 0.999f * v[p] + 0.001f * gr * gr;
  m[p] = mm;
  v[p] = vv;
  const float th = theta[p] - lr * (mm / bc1) / (sqrtf(vv / bc2) + 1e-8f);
  theta[p] = th;
  store_frags(p, th, wf, wt);
}

__global__ void k_encode(EncodeBox b, const float* p, const float* nrm, const float* wo, int64_t n,
                         float* feats) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float f[kIn];
  encode_features(b, p[3 * i], p[3 * i + 1], p[3 * i + 2], nrm[3 * i], nrm[3 * i + 1],
                  nrm[3 * i + 2], wo[3 * i], wo[3 * i + 1], wo[3 * i + 2], f);
  float4* dst = reinterpret_cast<float4*>(feats + i * kIn);
#pragma unroll
  for (int r = 0; r < kIn / 4; ++r) dst[r] = make_float4(f[4 * r], f[4 * r + 1], f[4 * r + 2], f[4 * r + 3]);
}

// Host: one gradient evaluation (k_mlp_train) over `batch` samples into grad_part.
int train_block_count(int batch) { return std::max(1, std::min(batch / 64, kMaxTrainBlocks)); }

void launch_train(const mupsi_guide* g, const float* feats, const float* dirs, const float* w,
                  const int* n_dev, int n_host, int batch, int random, uint32_t seed,
                  hipStream_t st) {
  TrainArgs ta{feats, dirs, w, n_dev, n_host, batch, random, seed, g->grad_part, g->loss_part,
               g->cfg.alpha};
  hipLaunchKernelGGL(k_mlp_train, dim3(train_block_count(batch)), dim3(256), 0, st, g->net(), ta);
}

void launch_adam(mupsi_guide* g, int batch, hipStream_t st) {
  g->step += 1;
  const float bc1 = (float)(1.0 - std::pow(0.9, (double)g->step));
  const float bc2 = (float)(1.0 - std::pow(0.999, (double)g->step));
  hipLaunchKernelGGL(k_adam, dim3(grid_for((int64_t)kNumParams * kRedLanes)), dim3(256), 0, st, g->theta, g->adam_m,
                     g->adam_v, reinterpret_cast<uint16_t*>(g->wf), reinterpret_cast<uint16_t*>(g->wt),
                     g->grad_part, train_block_count(batch), 1.0f / (float)batch, g->cfg.learning_rate,
                     bc1, bc2);
}

// He-uniform initialisation (biases zero); the output layer is scaled down so the
// initial mixture is broad and nearly uniform in weight.
std::vector<float> init_params(uint64_t seed) {
  std::vector<float> th(kNumParams, 0.0f);
  std::mt19937_64 rng(seed);
  auto fill = [&](int off, int out, int in, float scale) {
    std::uniform_real_distribution<float> u(-1.0f, 1.0f);
    const float lim = scale * std::sqrt(6.0f / (float)in);
    for (int i = 0; i < out * in; ++i) th[off + i] = lim * u(rng);
  };
  fill(kOffW1, kHid, kIn, 1.0f);
  fill(kOffW2, kHid, kHid, 1.0f);
  fill(kOffW3, kHid, kHid, 1.0f);
  fill(kOffW4, kOut, kHid, 0.1f);
  // log kappa = -3 + 11 sigmoid(s): s = -0.56 starts every lobe near kappa = e;
  // the selection output starts low (alpha = alpha_max sigmoid(-4) = 0.018 alpha_max),
  // so the first passes sample close to the cosine lobe until the guide has learnt
  for (int k = 0; k < kLobes; ++k) th[kOffB4 + 5 * k + 3] = -0.56f;
  th[kOffB4 + kAlphaIdx] = -4.0f;
  return th;
}

}  // namespace

namespace mupsi {
namespace guide {

int train_blocks(int batch) { return train_block_count(batch); }

void launch_infer(const mupsi_guide* g, const float* feats, const int* count_dev, int count_host,
                  const int* q_pid, float* raw_out, float* mix_out, hipStream_t st) {
  // blocks that are resident at once (each keeps the network in 40 KiB of LDS)
  const int cap_blocks = count_dev ? g->ctx->cus * 4 : std::max(1, (count_host + 63) / 64);
  const int blocks = std::max(1, std::min(cap_blocks, g->ctx->cus * 4));
  InferArgs ia{feats, count_dev, count_host, q_pid, raw_out, mix_out, g->cfg.alpha};
  hipLaunchKernelGGL(k_mlp_infer, dim3(blocks), dim3(256), 0, st, g->net(), ia);
}

mupsi_status train_random(mupsi_guide* g, const float* feats, const float* dirs, const float* w,
                          int n, int steps, int batch, hipStream_t st) {
  for (int i = 0; i < steps; ++i) {
    const uint32_t seed = (uint32_t)(g->cfg.seed * 0x9E3779B97F4A7C15ull >> 32) ^
                          (uint32_t)(g->step * 0x85ebca6bull + 0x1b873593ull);
    launch_train(g, feats, dirs, w, nullptr, n, batch, 1, seed, st);
    launch_adam(g, batch, st);
    HIP_TRY(hipGetLastError());
  }
  g->stats.train_steps += steps;
  return MUPSI_OK;
}

}  // namespace guide
}  // namespace mupsi

// ====================================================================== C ABI

static mupsi_status upload_params(mupsi_guide* g, const float* host) {
  HIP_TRY(hipMemcpy(g->theta, host, kNumParams * sizeof(float), hipMemcpyHostToDevice));
  HIP_TRY(hipMemset(g->adam_m, 0, kNumParams * sizeof(float)));
  HIP_TRY(hipMemset(g->adam_v, 0, kNumParams * sizeof(float)));
  hipLaunchKernelGGL(k_pack, dim3(grid_for(kNumParams)), dim3(256), 0, g->ctx->stream, g->theta,
                     reinterpret_cast<uint16_t*>(g->wf), reinterpret_cast<uint16_t*>(g->wt));
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipStreamSynchronize(g->ctx->stream));
  g->step = 0;
  return MUPSI_OK;
}

extern "C" mupsi_status mupsi_guide_create(mupsi_ctx* ctx, const mupsi_guide_config* cfg,
                                           mupsi_guide** out) {
  if (!ctx || !cfg || !out) return fail(MUPSI_ERR_INVALID, "mupsi_guide_create: null argument");
  *out = nullptr;
  if (!(cfg->learning_rate > 0.0f) || !(cfg->alpha >= 0.0f && cfg->alpha <= 1.0f) ||
      cfg->train_passes < 0 || cfg->batch_size < 64 || cfg->batch_size % 64 != 0 ||
      cfg->min_steps_per_pass < 0 || cfg->max_steps_per_pass < cfg->min_steps_per_pass ||
      cfg->max_records < 1 || cfg->max_records > 4 ||
      !(cfg->box_radius > 0.0f) || !(cfg->indirect_clamp >= 0.0f) || !(cfg->rr_boost >= 1.0f) ||
      cfg->passes_per_update < 1)
    return fail(MUPSI_ERR_INVALID, "mupsi_guide_create: bad configuration (lr > 0, alpha in [0,1], "
                                   "batch a positive multiple of 64, max_records 1..4, radius > 0, rr_boost >= 1, "
                                   "passes_per_update >= 1)");
  HIP_TRY(hipSetDevice(ctx->device));
  auto g = std::make_unique<mupsi_guide>();
  g->ctx = ctx;
  g->cfg = *cfg;
  HIP_TRY(g->mem.alloc(kNumParams, &g->theta));
  HIP_TRY(g->mem.alloc(kNumParams, &g->adam_m));
  HIP_TRY(g->mem.alloc(kNumParams, &g->adam_v));
  HIP_TRY(g->mem.alloc(kNumParams, &g->grad_out));
  HIP_TRY(g->mem.alloc((size_t)kNumFragF * 64, &g->wf));
  HIP_TRY(g->mem.alloc((size_t)kNumFragT * 64, &g->wt));
  HIP_TRY(hipMemset(g->wf, 0, (size_t)kNumFragF * 64 * sizeof(bf16x8)));
  HIP_TRY(hipMemset(g->wt, 0, (size_t)kNumFragT * 64 * sizeof(bf16x8)));
  g->part_blocks = kMaxTrainBlocks;
  HIP_TRY(g->mem.alloc((size_t)kMaxTrainBlocks * kNumParams, &g->grad_part));
  HIP_TRY(g->mem.alloc(kMaxTrainBlocks, &g->loss_part));
  const std::vector<float> th = init_params(cfg->seed);
  const mupsi_status st = upload_params(g.get(), th.data());
  if (st != MUPSI_OK) return st;
  ctx->guides.push_back(g.get());
  *out = g.release();
  return MUPSI_OK;
}

void destroy_guide(mupsi_guide* g) {
  (void)hipStreamSynchronize(g->ctx->stream);
  delete g;
}

extern "C" void mupsi_guide_free(mupsi_guide* g) {
  if (!g) return;
  mupsi_ctx* ctx = g->ctx;
  auto it = std::find(ctx->guides.begin(), ctx->guides.end(), g);
  if (it != ctx->guides.end()) ctx->guides.erase(it);
  for (mupsi_scene* s : ctx->scenes) detach_guide(s, g);
  (void)hipSetDevice(ctx->device);
  destroy_guide(g);
}

extern "C" mupsi_status mupsi_guide_get_params(mupsi_guide* g, float* host) {
  if (!g || !host) return fail(MUPSI_ERR_INVALID, "mupsi_guide_get_params: null argument");
  HIP_TRY(hipSetDevice(g->ctx->device));
  HIP_TRY(hipStreamSynchronize(g->ctx->stream));
  HIP_TRY(hipMemcpy(host, g->theta, kNumParams * sizeof(float), hipMemcpyDeviceToHost));
  return MUPSI_OK;
}

extern "C" mupsi_status mupsi_guide_set_params(mupsi_guide* g, const float* host) {
  if (!g || !host) return fail(MUPSI_ERR_INVALID, "mupsi_guide_set_params: null argument");
  for (int i = 0; i < kNumParams; ++i)
    if (!std::isfinite(host[i])) return fail(MUPSI_ERR_INVALID, "mupsi_guide_set_params: non-finite parameter");
  HIP_TRY(hipSetDevice(g->ctx->device));
  return upload_params(g, host);
}

extern "C" mupsi_status mupsi_guide_stats_get(mupsi_guide* g, mupsi_guide_stats* out) {
  if (!g || !out) return fail(MUPSI_ERR_INVALID, "mupsi_guide_stats_get: null argument");
  *out = g->stats;
  return MUPSI_OK;
}

extern "C" mupsi_status mupsi_guide_forward(mupsi_guide* g, const float* feats, int64_t n,
                                            float* raw, float* mixture) {
  if (!g) return fail(MUPSI_ERR_INVALID, "mupsi_guide_forward: null guide");
  if (n < 0 || n > INT32_MAX) return fail(MUPSI_ERR_INVALID, "mupsi_guide_forward: bad n");
  if (n == 0) return MUPSI_OK;
  if (!feats || (!raw && !mixture)) return fail(MUPSI_ERR_INVALID, "mupsi_guide_forward: null array");
  HIP_TRY(hipSetDevice(g->ctx->device));
  launch_infer(g, feats, nullptr, (int)n, nullptr, raw, mixture, g->ctx->stream);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipStreamSynchronize(g->ctx->stream));
  return MUPSI_OK;
}

extern "C" mupsi_status mupsi_guide_gradient(mupsi_guide* g, const float* feats, const float* dirs,
                                             const float* weights, int64_t n, float* grad_host,
                                             double* loss_host) {
  if (!g || !feats || !dirs || !weights) return fail(MUPSI_ERR_INVALID, "mupsi_guide_gradient: null argument");
  if (n <= 0 || n > (1 << 20)) return fail(MUPSI_ERR_INVALID, "mupsi_guide_gradient: n must be in 1..2^20");
  HIP_TRY(hipSetDevice(g->ctx->device));
  const int batch = (int)((n + 63) / 64 * 64);
  hipStream_t st = g->ctx->stream;
  launch_train(g, feats, dirs, weights, nullptr, (int)n, batch, 0, 0u, st);
  hipLaunchKernelGGL(k_grad_reduce, dim3(grid_for((int64_t)kNumParams * kRedLanes)), dim3(256), 0, st, g->grad_part,
                     train_block_count(batch), 1.0f / (float)n, g->grad_out);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipStreamSynchronize(st));
  if (grad_host)
    HIP_TRY(hipMemcpy(grad_host, g->grad_out, kNumParams * sizeof(float), hipMemcpyDeviceToHost));
  if (loss_host) {
    std::vector<float> lp(train_block_count(batch));
    HIP_TRY(hipMemcpy(lp.data(), g->loss_part, lp.size() * sizeof(float), hipMemcpyDeviceToHost));
    double s = 0;
    for (float v : lp) s += v;
    *loss_host = s / (double)n;
  }
  return MUPSI_OK;
}

extern "C" mupsi_status mupsi_guide_train_step(mupsi_guide* g, const float* feats, const float* dirs,
                                               const float* weights, int64_t n) {
  if (!g || !feats || !dirs || !weights) return fail(MUPSI_ERR_INVALID, "mupsi_guide_train_step: null argument");
  if (n <= 0 || n > 65536) return fail(MUPSI_ERR_INVALID, "mupsi_guide_train_step: n must be in 1..65536");
  HIP_TRY(hipSetDevice(g->ctx->device));
  const int batch = (int)((n + 63) / 64 * 64);
  hipStream_t st = g->ctx->stream;
  launch_train(g, feats, dirs, weights, nullptr, (int)n, batch, 0, 0u, st);
  // the minibatch mean is over the n real samples
  g->step += 1;
  const float bc1 = (float)(1.0 - std::pow(0.9, (double)g->step));
  const float bc2 = (float)(1.0 - std::pow(0.999, (double)g->step));
  hipLaunchKernelGGL(k_adam, dim3(grid_for((int64_t)kNumParams * kRedLanes)), dim3(256), 0, st, g->theta, g->adam_m,
                     g->adam_v, reinterpret_cast<uint16_t*>(g->wf), reinterpret_cast<uint16_t*>(g->wt),
                     g->grad_part, train_block_count(batch), 1.0f / (float)n, g->cfg.learning_rate, bc1, bc2);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipStreamSynchronize(st));
  g->stats.train_steps += 1;
  return MUPSI_OK;
}

extern "C" mupsi_status mupsi_guide_encode(mupsi_guide* g, const float* p, const float* nrm,
                                           const float* wo, int64_t n, float* feats) {
  if (!g) return fail(MUPSI_ERR_INVALID, "mupsi_guide_encode: null guide");
  if (n <= 0) return MUPSI_OK;
  if (!p || !nrm || !wo || !feats) return fail(MUPSI_ERR_INVALID, "mupsi_guide_encode: null array");
  HIP_TRY(hipSetDevice(g->ctx->device));
  hipLaunchKernelGGL(k_encode, dim3(grid_for(n)), dim3(256), 0, g->ctx->stream, g->box(), p, nrm, wo,
                     n, feats);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipStreamSynchronize(g->ctx->stream));
  return MUPSI_OK;
}
