// march_cached.h — GP-medium marcher with an LDS impulse window (gfx950).
//
// Cost model that motivates it: regenerating one impulse (make_seed, a 64-bit
// LCG jump and four PCG-XSH-RR draws) costs ~110 VALU slots, mostly quarter-rate
// 32-bit integer multiplies, while evaluating it costs ~12.  A ray advancing by
// dt = 0.5 through cells of size kr keeps 2/3 or more of its 27-cell neighbourhood
// from one step to the next, so the impulses are generated once per cell visit
// and evaluated from LDS at every step.
//
// Mapping: a group of 8 or 16 lanes marches one ray; a 64-lane workgroup (one
// wave) holds 8 or 4 groups.  Each group owns a window of 27 cells x ppc impulses
// in LDS, slot (n_x mod 3, n_y mod 3, n_z mod 3) x ppc + i, so moving the window
// centre by one cell along an axis overwrites exactly the slab of cells that left
// it.  The impulses of a cell are written as float4 (position, weight);
// evaluation reads them with one ds_read_b128 per slot, 27*ppc slots over the
// group's lanes, and reduces the group's partial sums with cross-lane swizzles.
//
// Results (hit flag, hit t bits) are the reference's sampleDistance/transmittance
// decisions; only the summation order of psi differs from the reference's
// sequential loop (within the FP32 error bound the parity tests use).
#pragma once
#include "device.h"

namespace mupsi {
namespace dev {

constexpr int kGroup = 32;        // default lanes per ray of the templates (the launches pick below)
// Lanes per ray of the per-scene instantiations.  Fewer lanes per ray put more
// rays in a wave (less per-ray control work, more BVH loads in flight for mesh
// means) but more window slots on each lane; measured best: sphere means 16
// lanes at ppc 3 (8 lanes 10% slower), mesh means 8 lanes at ppc 4 and 16 lanes
// for the larger windows of the generated scenes.
constexpr int sphere_group(int ppc) { return ppc <= 2 ? 8 : 16; }
constexpr int mesh_group(int ppc) { return ppc <= 4 ? 8 : 16; }
constexpr int kMultiGroup = 16;   // batched launches (ppc per ray, LDS for 8)
constexpr int kMaxPpcCached = 8;  // window capacity: 27 * 8 impulses

struct LcgJump {
  uint64_t a[kMaxPpcCached];
  uint64_t c[kMaxPpcCached];
};
// Jump of 4*i LCG steps: s_{4i} = A^{4i} s_0 + (A^{4i-1} + ... + 1).
constexpr LcgJump make_jumps() {
  LcgJump j{};
  uint64_t a = 1, c = 0;
  for (int i = 0; i < kMaxPpcCached; ++i) {
    j.a[i] = a;
    j.c[i] = c;
    for (int k = 0; k < 4; ++k) {
      c = c * 6364136223846793005ULL + 1ULL;
      a = a * 6364136223846793005ULL;
    }
  }
  return j;
}
__constant__ constexpr LcgJump kJump = make_jumps();

// State k of a cell's draw sequence from its first state s0 (k LCG steps):
// s_k = A^k s0 + (A^{k-1} + ... + 1), as (A^k, that sum) for k < 4 * kMaxPpcCached.
// With s0 a 32-bit value each state costs one 32 x 64-bit product, and the four
// draws of an impulse do not wait on one another.
struct LcgStates {
  uint64_t a[4 * kMaxPpcCached];
  uint64_t c[4 * kMaxPpcCached];
};
constexpr LcgStates make_states() {
  LcgStates j{};
  uint64_t a = 1, c = 0;
  for (int k = 0; k < 4 * kMaxPpcCached; ++k) {
    j.a[k] = a;
    j.c[k] = c;
    c = c * 6364136223846793005ULL + 1ULL;
    a = a * 6364136223846793005ULL;
  }
  return j;
}
__constant__ constexpr LcgStates kStates = make_states();

// The draw made from state A s0 + c of a stream whose first state s0 is 32-bit.
__device__ __forceinline__ uint32_t pcg_draw_from(uint32_t s0, uint64_t a, uint64_t c) {
  return pcg_output((uint64_t)s0 * a + c);
}

// New cells of the window when its centre moves by delta (each |delta_a| <= 1):
// {d in {-1,0,1}^3 : some axis a has delta_a != 0 and d_a == delta_a}.
// Entry [delta index][k] is the neighbour index 9(dx+1) + 3(dy+1) + (dz+1);
// off[delta index][k] the same neighbour's offsets packed as (dx+1) | (dy+1) << 2 |
// (dz+1) << 4.
struct NewCells {
  uint8_t n[28];
  uint8_t idx[28][27];
  uint8_t off[28][27];
};
constexpr NewCells make_new_cells() {
  NewCells t{};
  for (int di = 0; di < 28; ++di) {
    int cnt = 0;
    for (int ni = 0; ni < 27; ++ni) {
      bool add = di == 27;  // 27: refill the whole window
      if (di < 27) {
        const int dd[3] = {di / 9 - 1, (di / 3) % 3 - 1, di % 3 - 1};
        const int nn[3] = {ni / 9 - 1, (ni / 3) % 3 - 1, ni % 3 - 1};
        for (int a = 0; a < 3; ++a) add = add || (dd[a] != 0 && nn[a] == dd[a]);
      }
      if (add) {
        t.idx[di][cnt] = (uint8_t)ni;
        t.off[di][cnt++] = (uint8_t)((ni / 9) | ((ni / 3) % 3) << 2 | (ni % 3) << 4);
      }
    }
    t.n[di] = (uint8_t)cnt;
  }
  return t;
}
__constant__ constexpr NewCells kNewCells = make_new_cells();

// Number of cells entered for centre move delta: 27 minus the cells that stay,
// 2 per axis that moves and 3 per axis that does not (27 for a refill).
constexpr int new_cell_count(int ex, int ey, int ez) {
  return 27 - (ex ? 2 : 3) * (ey ? 2 : 3) * (ez ? 2 : 3);
}
constexpr bool new_cell_counts_match() {
  const NewCells t = make_new_cells();
  for (int di = 0; di < 27; ++di)
    if (new_cell_count(di / 9 - 1, (di / 3) % 3 - 1, di % 3 - 1) != t.n[di]) return false;
  return t.n[27] == 27;
}
static_assert(new_cell_counts_match(), "entered-cell count formula");

// floor(j / D) for 0 <= j < 64 D as one 24-bit multiply and a shift (full-rate
// VALU, where a general division takes a quarter-rate 32-bit multiply-high):
// M = ceil(2^s / D) for the smallest s exact over that whole range, found and
// checked at compile time.
struct DivSmall {
  uint32_t m, s;
};
constexpr DivSmall make_div_small(int d) {
  for (uint32_t s = 1; s < 24; ++s) {
    const uint32_t m = ((1u << s) + (uint32_t)d - 1u) / (uint32_t)d;
    bool ok = (uint64_t)(64 * d - 1) * m < (1ull << 24);
    for (int j = 0; ok && j < 64 * d; ++j) ok = ((uint32_t)j * m) >> s == (uint32_t)(j / d);
    if (ok) return DivSmall{m, s};
  }
  return DivSmall{0, 0};
}
template <int D>
__device__ __forceinline__ int div_small(int j) {
  constexpr DivSmall k = make_div_small(D);
  static_assert(k.m != 0, "no exact multiply-shift for this divisor");
  return (int)(__umul24((uint32_t)j, k.m) >> k.s);
}
__device__ __forceinline__ int mod3(int n) {
  const int r = n % 3;
  return r < 0 ? r + 3 : r;
}

// trunc(q / kr) with the reference's correctly rounded division.  A product with
// the rounded reciprocal is within 2 ulp of the quotient; it is used unless an
// integer lies that close to it (or kr is a power of two, where it is exact).
__device__ __forceinline__ int cell_coord(float q, const DevGP& g, float inv_kr, bool pow2) {
  const float a = q * inv_kr;
  if (pow2) return (int)a;
  const float r = rintf(a);
  if (fabsf(a - r) > 4.0f * 1.1920929e-7f * fmaxf(1.0f, fabsf(a))) return (int)a;
  return (int)(q / g.kr);
}

// cell_coord of the three coordinates of q, the power-of-two test taken once.
__device__ __forceinline__ void cell_coords(F3 q, const DevGP& g, float inv_kr, bool pow2, int& cx,
                                            int& cy, int& cz) {
  const float ax = q.x * inv_kr, ay = q.y * inv_kr, az = q.z * inv_kr;
  cx = (int)ax;
  cy = (int)ay;
  cz = (int)az;
  if (!pow2) {
    const auto near_int = [](float a) {
      return !(fabsf(a - rintf(a)) > 4.0f * 1.1920929e-7f * fmaxf(1.0f, fabsf(a)));
    };
    if (near_int(ax)) cx = (int)(q.x / g.kr);
    if (near_int(ay)) cy = (int)(q.y / g.kr);
    if (near_int(az)) cz = (int)(q.z / g.kr);
  }
}

// Impulse of cell (nx, ny, nz) of realization seed as (position, weight); kr is
// the cell size, (ja, jc) the jump to the impulse's first draw (kJump.a[i],
// kJump.c[i] for impulse i).
__device__ __forceinline__ float4 make_impulse(float kr, int nx, int ny, int nz, uint32_t seed,
                                               uint64_t ja, uint64_t jc) {
  const uint32_t cs = make_seed4((uint32_t)nx, (uint32_t)ny, (uint32_t)nz, seed);
  uint64_t s = (uint32_t)(cs + 1u);
  s = s * ja + jc;
  const uint32_t r0 = pcg_next(s);
  const uint32_t r1 = pcg_next(s);
  const uint32_t r2 = pcg_next(s);
  const uint32_t r3 = pcg_next(s);
  return make_float4(fmaf(u01(r2), kr, (float)nx * kr), fmaf(u01(r1), kr, (float)ny * kr),
                     fmaf(u01(r0), kr, (float)nz * kr), (r3 >> 31) ? 1.0f : -1.0f);
}

}  // namespace dev
}  // namespace mupsi
