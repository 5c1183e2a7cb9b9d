// guide.h — learned path guiding: network layout, feature encoding and the
// directional mixture, shared by the MLP kernels (guide.hip) and the tracer
// (kernels.hip).
//
// The reference has no guiding; this is an addition on top of its trace loop
// (the plain tracer stays available and unchanged).  At a guided vertex (a GP-medium hit, whose BSDF is the fixed
// Lambertian 0.8 of gpmedium.h:46-48) the next direction is drawn by one-sample
// MIS from
//     q(w) = alpha * p_guide(w | x) + (1 - alpha) * cos(w, n) / pi,
// and the throughput is multiplied by f(w) cos / q(w) = albedo (cos/pi) / q(w),
// so the estimator keeps the expectation of the reference's Lambertian step
// (albedo, with p = cos/pi) exactly.
//
// Network: 32 -> 64 -> 64 -> 64 -> 41, ReLU on the hidden layers, BF16 operands
// on MFMA with FP32 accumulation, FP32 master parameters.  Outputs 0..39 are 8
// lobes x (v.x, v.y, v.z, s, a): mean direction v/|v|, concentration
// kappa = exp(lo + (hi - lo) sigmoid(s)) (a smooth bound: a hard clamp would
// stop the gradient of a lobe that overshoots), mixture weight softmax(a).  A
// lobe is a von Mises-Fisher density on the sphere,
//     vMF(w; mu, kappa) = kappa / (2 pi (1 - e^{-2 kappa})) exp(kappa (mu.w - 1)).
// Output 40 is the selection probability of the guide at this vertex,
// alpha = alpha_max sigmoid(t): the loss -w log q is taken on the combined
// density q, so the network also learns where the guide should not be used.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace mupsi {
namespace guide {

constexpr int kIn = 32, kHid = 64, kOut = 41, kOutPad = 48, kLobes = 8;
constexpr int kMix = 44;     // floats per vertex: 8 x (mu, kappa, weight), alpha, 3 unused
constexpr int kAlphaIdx = 40;

// Flat FP32 parameter vector: W1 [64][32], b1, W2 [64][64], b2, W3 [64][64], b3,
// W4 [41][64], b4 (row-major [out][in]).
constexpr int kOffW1 = 0;
constexpr int kOffB1 = kOffW1 + kHid * kIn;    // 2048
constexpr int kOffW2 = kOffB1 + kHid;          // 2112
constexpr int kOffB2 = kOffW2 + kHid * kHid;   // 6208
constexpr int kOffW3 = kOffB2 + kHid;          // 6272
constexpr int kOffB3 = kOffW3 + kHid * kHid;   // 10368
constexpr int kOffW4 = kOffB3 + kHid;          // 10432
constexpr int kOffB4 = kOffW4 + kOut * kHid;   // 13056
constexpr int kNumParams = kOffB4 + kOut;      // 13097

// MFMA 16x16x32 BF16 fragments (64 lanes x 8 bf16 each).  Forward: A = W_l tiles
// (out-tile mt of 16 rows, in-tile kt of 32 columns); transposed: A = W_l^T tiles
// (in-tile mt, out-tile kt) for the backward pass of layers 2-4.
constexpr int kFragW1 = 0;                 // 4 x 1
constexpr int kFragW2 = kFragW1 + 4;       // 4 x 2
constexpr int kFragW3 = kFragW2 + 8;       // 4 x 2
constexpr int kFragW4 = kFragW3 + 8;       // 3 x 2 (rows 41..47 are zero)
constexpr int kNumFragF = kFragW4 + 6;     // 26
constexpr int kFragT2 = 0;                 // W2^T: 4 x 2
constexpr int kFragT3 = kFragT2 + 8;       // W3^T: 4 x 2
constexpr int kFragT4 = kFragT3 + 8;       // W4^T: 4 x 2 (columns 41..63 are zero)
constexpr int kNumFragT = kFragT4 + 8;     // 24

constexpr float kLogKappaMin = -3.0f;  // kappa in (0.0498, 2981)
constexpr float kLogKappaMax = 8.0f;

__device__ __forceinline__ float log_kappa(float s) {
  return kLogKappaMin + (kLogKappaMax - kLogKappaMin) / (1.0f + __expf(-s));
}

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef short bf16x8 __attribute__((ext_vector_type(8)));  // 8 BF16 bit patterns
typedef float f32x4 __attribute__((ext_vector_type(4)));

// FP32 -> BF16, round to nearest even (inputs are finite).
__host__ __device__ __forceinline__ uint16_t f2bf(float f) {
  uint32_t u;
  __builtin_memcpy(&u, &f, 4);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}
__host__ __device__ __forceinline__ float bf2f(uint16_t h) {
  const uint32_t u = (uint32_t)h << 16;
  float f;
  __builtin_memcpy(&f, &u, 4);
  return f;
}
__device__ __forceinline__ bf16x4 pack4(float a, float b, float c, float d) {
  bf16x4 r;
  r[0] = (short)f2bf(a);
  r[1] = (short)f2bf(b);
  r[2] = (short)f2bf(c);
  r[3] = (short)f2bf(d);
  return r;
}

// Where W_l[o][i] lives in the fragment arrays (element index in units of bf16).
// In a 32-wide k-tile, lane group g (lanes 16g .. 16g+15) holds k = 4g .. 4g+3
// in elements 0-3 and k = 16+4g .. 16+4g+3 in elements 4-7: the rows that the
// 16x16 D tiles 2kt and 2kt+1 of a layer leave in lane group g (a D tile puts
// rows 4g .. 4g+3 of column s in lane s + 16g), so a layer's output tiles are
// the next layer's B operand as they stand, with no data moved between lanes.
// The B operand uses the same k order, so the products pair up as the sum
// needs them.
__host__ __device__ __forceinline__ int frag_elem(int frag_base, int nkt, int row, int col) {
  const int mt = row >> 4, r = row & 15, kt = col >> 5, c = col & 31;
  const int lane = r + 16 * ((c & 15) >> 2);
  return ((frag_base + mt * nkt + kt) * 64 + lane) * 8 + (c & 3) + ((c >> 4) << 2);
}

// Device view of one network.
struct NetDev {
  const float* theta;   // FP32 master parameters (biases are read from here)
  const bf16x8* wf;     // kNumFragF x 64 forward fragments
  const bf16x8* wt;     // kNumFragT x 64 transposed fragments
};

// Position normalisation for the encoding: p_hat = clamp((p - c) / R, -1, 1).
struct EncodeBox {
  float c[3];
  float inv_r;
};

// Feature vector (32): sin/cos(2^i pi p_hat_a) for a = x, y, z and i = 0..3
// (entries 8a + 2i, 8a + 2i + 1), then n (24..26), wo (27..29), 1, n.wo.
__device__ __forceinline__ void encode_features(const EncodeBox& b, float px, float py, float pz,
                                                float nx, float ny, float nz, float wx, float wy,
                                                float wz, float* f) {
  const float ph[3] = {fminf(fmaxf((px - b.c[0]) * b.inv_r, -1.0f), 1.0f),
                       fminf(fmaxf((py - b.c[1]) * b.inv_r, -1.0f), 1.0f),
                       fminf(fmaxf((pz - b.c[2]) * b.inv_r, -1.0f), 1.0f)};
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    float fr = ph[a];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      float s, c;
      sincospif(fr, &s, &c);  // sin(pi * fr), cos(pi * fr)
      f[8 * a + 2 * i] = s;
      f[8 * a + 2 * i + 1] = c;
      fr *= 2.0f;
    }
  }
  f[24] = nx;
  f[25] = ny;
  f[26] = nz;
  f[27] = wx;
  f[28] = wy;
  f[29] = wz;
  f[30] = 1.0f;
  f[31] = fmaf(nx, wx, fmaf(ny, wy, nz * wz));
}

// ------------------------------------------------------------------ mixture

__device__ __forceinline__ float sigmoidf_(float x) { return 1.0f / (1.0f + __expf(-x)); }

// Raw network outputs -> mixture parameters, 8 x (mu.x, mu.y, mu.z, kappa, pi),
// then alpha (m[40]) and three zeros.
__device__ __forceinline__ void raw_to_mixture(const float* o, float alpha_max, float* m) {
  float amax = -INFINITY;
#pragma unroll
  for (int k = 0; k < kLobes; ++k) amax = fmaxf(amax, o[5 * k + 4]);
  float e[kLobes], esum = 0.0f;
#pragma unroll
  for (int k = 0; k < kLobes; ++k) {
    e[k] = __expf(o[5 * k + 4] - amax);
    esum += e[k];
  }
#pragma unroll
  for (int k = 0; k < kLobes; ++k) {
    const float vx = o[5 * k], vy = o[5 * k + 1], vz = o[5 * k + 2];
    const float len = sqrtf(fmaf(vx, vx, fmaf(vy, vy, vz * vz)));
    const bool ok = len > 1e-12f;
    m[5 * k + 0] = ok ? vx / len : 0.0f;
    m[5 * k + 1] = ok ? vy / len : 0.0f;
    m[5 * k + 2] = ok ? vz / len : 1.0f;
    m[5 * k + 3] = __expf(log_kappa(o[5 * k + 3]));
    m[5 * k + 4] = e[k] / esum;
  }
  m[kAlphaIdx] = alpha_max * sigmoidf_(o[kAlphaIdx]);
  m[41] = m[42] = m[43] = 0.0f;
}

// log of the vMF normaliser kappa / (2 pi (1 - e^{-2 kappa})).
__device__ __forceinline__ float vmf_log_norm(float kappa) {
  return logf(kappa) - 1.8378770664093453f - log1pf(-__expf(-2.0f * kappa));
}

__device__ __forceinline__ float mixture_pdf(const float* m, float wx, float wy, float wz) {
  float p = 0.0f;
#pragma unroll
  for (int k = 0; k < kLobes; ++k) {
    const float kap = m[5 * k + 3];
    const float c = fmaf(m[5 * k], wx, fmaf(m[5 * k + 1], wy, m[5 * k + 2] * wz));
    p = fmaf(m[5 * k + 4], __expf(vmf_log_norm(kap) + kap * (c - 1.0f)), p);
  }
  return p;
}

// Draw a direction from the mixture with three uniforms: u1 picks the lobe by
// its weight, u2 the cosine to the lobe mean by inverting the vMF marginal
// (w = 1 + log(u2 + (1 - u2) e^{-2 kappa}) / kappa), u3 the azimuth.
__device__ __forceinline__ void mixture_sample(const float* m, float u1, float u2, float u3,
                                               float& wx, float& wy, float& wz) {
  int k = kLobes - 1;
  float acc = 0.0f;
#pragma unroll
  for (int j = 0; j < kLobes - 1; ++j) {
    acc += m[5 * j + 4];
    if (u1 < acc && k == kLobes - 1) k = j;
  }
  const float mx = m[5 * k], my = m[5 * k + 1], mz = m[5 * k + 2], kap = m[5 * k + 3];
  float w = 1.0f + logf(fmaf(1.0f - u2, __expf(-2.0f * kap), u2)) / kap;
  w = fminf(fmaxf(w, -1.0f), 1.0f);
  const float r = sqrtf(fmaxf(0.0f, fmaf(-w, w, 1.0f)));
  float sp, cp;
  sincospif(2.0f * u3, &sp, &cp);
  // orthonormal basis around mu (Duff et al. 2017)
  const float sg = copysignf(1.0f, mz);
  const float a = -1.0f / (sg + mz);
  const float b = mx * my * a;
  const float tx = fmaf(sg * mx * mx, a, 1.0f), ty = sg * b, tz = -sg * mx;
  const float bx = b, by = fmaf(my * my, a, sg), bz = -my;
  const float lx = r * cp, ly = r * sp;
  wx = fmaf(mx, w, fmaf(bx, ly, tx * lx));
  wy = fmaf(my, w, fmaf(by, ly, ty * lx));
  wz = fmaf(mz, w, fmaf(bz, ly, tz * lx));
  const float len = sqrtf(fmaf(wx, wx, fmaf(wy, wy, wz * wz)));
  wx /= len;
  wy /= len;
  wz /= len;
}

// Per-sample loss -wt log q(w) of the combined sampling density
//     q = alpha p_guide(w) + (1 - alpha) max(cos, 0) / pi
// given raw outputs o (cosn = cos(w, n) of the vertex), and its gradient d[41]
// with respect to o.  With A = alpha p_guide / q, the lobe terms are the
// gradient of -wt A log p_guide; the alpha output gets -wt (p_guide - cos/pi)/q
// times d alpha / d t.  Returns the loss.
__device__ __forceinline__ float mixture_loss_grad(const float* o, float wx, float wy, float wz,
                                                   float cosn, float wt, float alpha_max,
                                                   float* d) {
  float amax = -INFINITY;
#pragma unroll
  for (int k = 0; k < kLobes; ++k) amax = fmaxf(amax, o[5 * k + 4]);
  float lse_a = 0.0f;
#pragma unroll
  for (int k = 0; k < kLobes; ++k) lse_a += __expf(o[5 * k + 4] - amax);
  lse_a = amax + logf(lse_a);
  float ell[kLobes], kap[kLobes], c[kLobes], len[kLobes], mu[kLobes][3];
  float lmax = -INFINITY;
#pragma unroll
  for (int k = 0; k < kLobes; ++k) {
    const float vx = o[5 * k], vy = o[5 * k + 1], vz = o[5 * k + 2];
    len[k] = sqrtf(fmaf(vx, vx, fmaf(vy, vy, vz * vz)));
    const bool ok = len[k] > 1e-12f;
    mu[k][0] = ok ? vx / len[k] : 0.0f;
    mu[k][1] = ok ? vy / len[k] : 0.0f;
    mu[k][2] = ok ? vz / len[k] : 1.0f;
    kap[k] = __expf(log_kappa(o[5 * k + 3]));
    c[k] = fmaf(mu[k][0], wx, fmaf(mu[k][1], wy, mu[k][2] * wz));
    ell[k] = (o[5 * k + 4] - lse_a) + vmf_log_norm(kap[k]) + kap[k] * (c[k] - 1.0f);
    lmax = fmaxf(lmax, ell[k]);
  }
  float sum = 0.0f;
#pragma unroll
  for (int k = 0; k < kLobes; ++k) sum += __expf(ell[k] - lmax);
  const float logpg = lmax + logf(sum);
  const float pg = __expf(logpg);
  const float sa = sigmoidf_(o[kAlphaIdx]);
  const float alpha = alpha_max * sa;
  const float pb = fmaxf(cosn, 0.0f) * 0.318309886183790671538f;
  const float q = alpha * pg + (1.0f - alpha) * pb;
  if (!(q > 0.0f) || !(wt > 0.0f)) {
#pragma unroll
    for (int r = 0; r < kOut; ++r) d[r] = 0.0f;
    return 0.0f;
  }
  const float wA = wt * (alpha * pg / q);
#pragma unroll
  for (int k = 0; k < kLobes; ++k) {
    const float r = __expf(ell[k] - logpg);       // responsibility within the guide
    const float pi = __expf(o[5 * k + 4] - lse_a);
    const float g = -wA * r;
    // d log vMF / d kappa = 1/kappa - 2 e^{-2k} / (1 - e^{-2k}) + (c - 1)
    const float e2 = __expf(-2.0f * kap[k]);
    const float dlk = 1.0f / kap[k] - 2.0f * e2 / (-expm1f(-2.0f * kap[k])) + (c[k] - 1.0f);
    // d log kappa / d s = (hi - lo) sigmoid(s) (1 - sigmoid(s))
    const float sg = sigmoidf_(o[5 * k + 3]);
    d[5 * k + 3] = g * dlk * kap[k] * (kLogKappaMax - kLogKappaMin) * sg * (1.0f - sg);
    d[5 * k + 4] = -wA * (r - pi);
    if (len[k] > 1e-12f) {
      const float f = g * kap[k] / len[k];
      d[5 * k + 0] = f * (wx - c[k] * mu[k][0]);
      d[5 * k + 1] = f * (wy - c[k] * mu[k][1]);
      d[5 * k + 2] = f * (wz - c[k] * mu[k][2]);
    } else {
      d[5 * k + 0] = d[5 * k + 1] = d[5 * k + 2] = 0.0f;
    }
  }
  d[kAlphaIdx] = -wt * (pg - pb) / q * alpha_max * sa * (1.0f - sa);
  return -wt * logf(q);
}

}  // namespace guide
}  // namespace mupsi
