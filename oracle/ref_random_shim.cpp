// ref_random_shim.cpp — TEST INFRASTRUCTURE ONLY.
//
// Thin extern "C" wrappers around the reference's own, dependency-free
// src/math/random.h, compiled from where it lies in the reference tree by
// oracle/Makefile into oracle/_ref/libref_random.so.  Used only by the golden
// vector generator (tests/golden/make_rng_golden.py) and by the not-gpu tests
// that pin the oracle's RNG restatement; nothing here is product code.
#include <cstdint>

#include "math/random.h"

extern "C" {

uint32_t ref_make_seed4(int32_t a, int32_t b, int32_t c, uint32_t d) {
  return mupsi::make_seed(a, b, c, d);
}
uint32_t ref_make_seed5(int32_t a, int32_t b, int32_t c, uint32_t d, int32_t e) {
  return mupsi::make_seed(a, b, c, d, e);
}
// Random(seed) followed by n nextI() draws.
void ref_random_ctor_draws(uint64_t seed, int32_t n, uint32_t* out) {
  mupsi::Random r(seed);
  for (int32_t i = 0; i < n; ++i) out[i] = r.nextI();
}
// setState(s) followed by n nextI() draws.
void ref_random_state_draws(uint64_t s, int32_t n, uint32_t* out) {
  mupsi::Random r(1u);
  r.setState(s);
  for (int32_t i = 0; i < n; ++i) out[i] = r.nextI();
}
float ref_normalized_uint(uint32_t i) { return mupsi::Random::normalizedUint(i); }
float ref_bernoulli(float rv) { return mupsi::Random::Bernoulli(rv, -1.0f, 1.0f, 0.5f); }

}  // extern "C"
