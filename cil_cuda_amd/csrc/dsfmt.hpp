// dsfmt.hpp — the reference's sequential random stream, for host and device.
//
// The reference draws every random number from one global dSFMT-19937 generator
// (Random::dsfmt, Random.hpp:8-16; vendored mt/dSFMT.c, 2.2.x), through Random::uniform0x,
// Random::uniform, Random::normal (Leva's ratio-of-uniforms method) and Random::onDSphere.
// This restates the published generator (Saito & Matsumoto, "SIMD-oriented Fast Mersenne
// Twister", MCQMC 2008; parameter set MEXP 19937 of dSFMT-params19937.h) and those samplers so
// that a colony can be advanced on the very stream the reference consumes.  The state is 192
// 128-bit words (the last one is the "lung") and an index into the 382 doubles of a block.
#pragma once
#include <stdint.h>

#include <hip/hip_runtime.h>

namespace cil {

constexpr int kDsN = 191;            // 128-bit words of state (MEXP 19937)
constexpr int kDsN64 = 2 * kDsN;     // 64-bit outputs per block
constexpr int kDsPos1 = 117;
constexpr int kDsSl1 = 19;
constexpr int kDsSr = 12;
constexpr uint64_t kDsMsk1 = 0x000ffafffffffb3fULL;
constexpr uint64_t kDsMsk2 = 0x000ffdfffc90fffdULL;
constexpr uint64_t kDsFix1 = 0x90014964b32f4329ULL;
constexpr uint64_t kDsFix2 = 0x3b8d12ac548a7c7aULL;
constexpr uint64_t kDsPcv1 = 0x3d84e1ac0dc82880ULL;
constexpr uint64_t kDsPcv2 = 0x0000000000000001ULL;
constexpr uint64_t kDsLowMask = 0x000FFFFFFFFFFFFFULL;
constexpr uint64_t kDsHighConst = 0x3FF0000000000000ULL;

struct Dsfmt {
  uint64_t s[kDsN + 1][2];
  int idx;        // next 64-bit output of the current block (kDsN64: a new block is due)
  uint64_t used;  // 64-bit outputs consumed since seeding
};

// One step of the recursion on a 128-bit word (two 64-bit halves).
__host__ __device__ inline void dsfmt_recursion(uint64_t* r, const uint64_t* a, const uint64_t* b, uint64_t* lung) {
  const uint64_t t0 = a[0], t1 = a[1], l0 = lung[0], l1 = lung[1];
  lung[0] = (t0 << kDsSl1) ^ (l1 >> 32) ^ (l1 << 32) ^ b[0];
  lung[1] = (t1 << kDsSl1) ^ (l0 >> 32) ^ (l0 << 32) ^ b[1];
  r[0] = (lung[0] >> kDsSr) ^ (lung[0] & kDsMsk1) ^ t0;
  r[1] = (lung[1] >> kDsSr) ^ (lung[1] & kDsMsk2) ^ t1;
}

// A new block of kDsN64 outputs in place.
__host__ __device__ inline void dsfmt_next_block(Dsfmt& g) {
  uint64_t lung[2] = {g.s[kDsN][0], g.s[kDsN][1]};
  int i = 0;
  for (; i < kDsN - kDsPos1; ++i) dsfmt_recursion(g.s[i], g.s[i], g.s[i + kDsPos1], lung);
  for (; i < kDsN; ++i) dsfmt_recursion(g.s[i], g.s[i], g.s[i + kDsPos1 - kDsN], lung);
  g.s[kDsN][0] = lung[0];
  g.s[kDsN][1] = lung[1];
}

// Seeding by a 32-bit integer: a linear-congruential fill of the 32-bit words, the mask that
// puts every output word in [1, 2), and the period certification of the lung.
__host__ __device__ inline void dsfmt_seed(Dsfmt& g, uint32_t seed) {
  uint32_t prev = seed;
  for (int k = 0; k < (kDsN + 1) * 4; ++k) {
    const uint32_t v = k == 0 ? seed : 1812433253u * (prev ^ (prev >> 30)) + (uint32_t)k;
    prev = v;
    const int w = k >> 2, h = (k >> 1) & 1;
    if ((k & 1) == 0) g.s[w][h] = (uint64_t)v;
    else g.s[w][h] |= (uint64_t)v << 32;
  }
  for (int w = 0; w < kDsN; ++w)
    for (int h = 0; h < 2; ++h) g.s[w][h] = (g.s[w][h] & kDsLowMask) | kDsHighConst;
  uint64_t inner = ((g.s[kDsN][0] ^ kDsFix1) & kDsPcv1) ^ ((g.s[kDsN][1] ^ kDsFix2) & kDsPcv2);
  for (int k = 32; k > 0; k >>= 1) inner ^= inner >> k;
  if ((inner & 1) != 1) g.s[kDsN][1] ^= 1;  // kDsPcv2's lowest set bit
  g.idx = kDsN64;
  g.used = 0;
}

__host__ __device__ inline uint64_t dsfmt_bits(Dsfmt& g) {
  if (g.idx >= kDsN64) {
    dsfmt_next_block(g);
    g.idx = 0;
  }
  const uint64_t v = g.s[g.idx >> 1][g.idx & 1];
  ++g.idx;
  ++g.used;
  return v;
}

__host__ __device__ inline double dsfmt_as_double(uint64_t b) {
  union {
    uint64_t u;
    double d;
  } c;
  c.u = b;
  return c.d;
}

// Skip n outputs (the draws another party already consumed).
__host__ __device__ inline void dsfmt_skip(Dsfmt& g, uint64_t n) {
  for (uint64_t k = 0; k < n; ++k) (void)dsfmt_bits(g);
}

// dsfmt_genrand_close_open: [0, 1); Random::uniform0x(x) = x * this (Random.hpp:13-15).
__host__ __device__ inline double dsfmt_close_open(Dsfmt& g) { return dsfmt_as_double(dsfmt_bits(g)) - 1.0; }
// dsfmt_genrand_open_open: (0, 1); Random::uniform() = 2 (this - 1/2) (Random.hpp:10-12).
__host__ __device__ inline double dsfmt_open_open(Dsfmt& g) { return dsfmt_as_double(dsfmt_bits(g) | 1) - 1.0; }

// Random::normal (Random.hpp:24-52): Leva's ratio-of-uniforms sampler (ACM TOMS 18, 1992) with
// its quadratic inner and outer bounds.
__host__ __device__ inline double dsfmt_normal(Dsfmt& g) {
  const double s = 0.449871, t = -0.386595, a = 0.19600, b = 0.25472, r1 = 0.27597, r2 = 0.27846;
  double u, v, Q;
  do {
    u = 1.0 - dsfmt_close_open(g);
    v = 1.7156 * (dsfmt_close_open(g) - 0.5);
    const double x = u - s, y = fabs(v) - t;
    Q = x * x + y * (a * y - b * x);
  } while (Q >= r1 && (Q > r2 || v * v > -4.0 * u * u * log(u)));
  return v / u;
}

// Random::onDSphere (Random.hpp:82-114): a point uniform in the unit D-ball by rejection with
// Random::uniform(), scaled to the sphere.
__host__ __device__ inline void dsfmt_on_sphere(Dsfmt& g, int D, double* n) {
  double r2;
  do {
    r2 = 0.0;
    for (int d = 0; d < D; ++d) {
      n[d] = (dsfmt_open_open(g) - 0.5) * 2.0;
      r2 += n[d] * n[d];
    }
  } while (!(r2 <= 1.0));
  const double norm = 1.0 / sqrt(r2);
  for (int d = 0; d < D; ++d) n[d] *= norm;
}

}  // namespace cil
