// device.hpp — device-side building blocks of the gfx950 colony step: counter RNG,
// min-image geometry, pair laws and intra-cell forces.  FP64 paths reproduce the
// reference's operation order (the library is compiled with -ffp-contract=off, so no
// multiply-add is fused unless written as fma()).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/cil_hip.h"

namespace cil {

constexpr double kPi = 3.14159265358979323846;
constexpr double kTwoPi = 2.0 * kPi;
constexpr double kBigTolMp = 1.0e3 * 10.0 * 2.220446049250313e-16;  // Utils.hpp:47-50
constexpr double kTolMp = 10.0 * 2.220446049250313e-16;
constexpr double kWallCutoff = 0.82;  // Wall.cpp:3

enum Stream : uint32_t { ST_WIENER = 1, ST_FDT = 2, ST_ACTIVE = 3, ST_DIV = 4, ST_ZERO = 5, ST_PAIR = 6 };
// The stochastic Heun step evaluates forces twice per step: the corrector's evaluation (at the
// half-step positions, before the step index moves on) adds ST_PHASE to the streams of its
// noise and zero-distance draws, so it never repeats the previous step's final evaluation.
constexpr uint32_t ST_PHASE = 8;
enum CellStateV : int { CS_DIVIDING = 0, CS_CRAWLING = 1, CS_UNDEFINED = 2 };  // Cell.cuh:26-30

// Error bits raised by kernels and collected by the host at synchronisation points.
enum ErrBit : unsigned { ERR_BIN_OVERFLOW = 1u, ERR_CAPACITY = 2u, ERR_BOND = 4u, ERR_NONFINITE = 8u,
                         ERR_CUTOFF = 16u, ERR_NOISE = 32u, ERR_GUARD = 64u,
                         ERR_KNN = 128u, ERR_LIST = 256u, ERR_SYNC = 512u };

// ---------------------------------------------------------------------------------------
// Philox4x32-10 (Salmon et al., SC'11).  Counter = (id, ts, stream, k), key = seed.
// ---------------------------------------------------------------------------------------
struct u4 {
  uint32_t x, y, z, w;
};

__host__ __device__ inline u4 philox(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3, uint64_t seed) {
  uint32_t k0 = (uint32_t)seed, k1 = (uint32_t)(seed >> 32);
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    uint64_t p0 = (uint64_t)0xD2511F53u * c0;
    uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
    uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0;
    uint32_t n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
    c1 = (uint32_t)p1;
    c3 = (uint32_t)p0;
    c0 = n0;
    c2 = n2;
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
  return u4{c0, c1, c2, c3};
}

__device__ inline void philox_uniform2(uint64_t seed, uint32_t id, uint32_t ts, uint32_t st, uint32_t k,
                                       double u[2]) {
  u4 o = philox(id, ts, st, k, seed);
  u[0] = (double)(((((uint64_t)o.y) << 32) | o.x) >> 11) * 0x1.0p-53;
  u[1] = (double)(((((uint64_t)o.w) << 32) | o.z) >> 11) * 0x1.0p-53;
}

__device__ inline void philox_normal2(uint64_t seed, uint32_t id, uint32_t ts, uint32_t st, uint32_t k,
                                      double z[2]) {
  u4 o = philox(id, ts, st, k, seed);
  double u1 = (double)((((((uint64_t)o.y) << 32) | o.x) >> 11) + 1) * 0x1.0p-53;
  double u2 = (double)(((((uint64_t)o.w) << 32) | o.z) >> 11) * 0x1.0p-53;
  double r = sqrt(-2.0 * log(u1));
  double a = kTwoPi * u2;
  double sa, ca;
  sincos(a, &sa, &ca);
  z[0] = r * ca;
  z[1] = r * sa;
}

// The same pair of normals for the FP32 force mode: the same Philox draws and 53-bit
// uniforms, the Box-Muller transform in FP32 (u1 converted to FP32 keeps its relative
// precision down to the smallest uniform, so the tails match).  The normals agree with
// philox_normal2 to about 1e-6 relative.
__device__ inline void philox_normal2_f32(uint64_t seed, uint32_t id, uint32_t ts, uint32_t st, uint32_t k,
                                          double z[2]) {
  u4 o = philox(id, ts, st, k, seed);
  const float u1 = (float)((double)((((((uint64_t)o.y) << 32) | o.x) >> 11) + 1) * 0x1.0p-53);
  const float u2 = (float)((double)(((((uint64_t)o.w) << 32) | o.z) >> 11) * 0x1.0p-53);
  const float r = sqrtf(-2.0f * __logf(u1));
  float sa, ca;
  __sincosf((float)kTwoPi * u2, &sa, &ca);
  z[0] = (double)(r * ca);
  z[1] = (double)(r * sa);
}

__host__ __device__ inline uint32_t cell_key(int s, int c) { return ((uint32_t)s << 28) | (uint32_t)c; }

// ---------------------------------------------------------------------------------------
// Geometry (Utils.hpp:171-255, 447-602)
// ---------------------------------------------------------------------------------------
__device__ inline int nint(double a) { return a >= 0.0 ? (int)(a + 0.5) : (int)(a - 0.5); }
__device__ inline double sq(double a) { return a * a; }

// Exact wrap of fmod(s, L) for s in [0, 2L) (Sterbenz: s - L is exact there); fmod otherwise.
__device__ inline double wrap_fmod(double s, double L) {
  if (s >= 0.0 && s < L) return s;
  if (s >= L && s < 2.0 * L) return s - L;
  return fmod(s, L);
}

struct Box {
  int dim;
  double L[3];
  double invL[3];
  int pbc[3];
};

// BoundingBox::distance: r12 = r2 - r1 with minimum image.  Nint(x/L) is taken as
// Nint(x*invL): the two can only differ when |x| is near L/2, where no pair interacts.
__device__ inline double box_dist(const Box& b, double* r12, const double* r1, const double* r2) {
  double ans = 0.0;
  for (int d = 0; d < b.dim; ++d) {
    double x = r2[d] - r1[d];
    if (b.pbc[d]) x -= (double)nint(x / b.L[d]) * b.L[d];
    r12[d] = x;
    ans += sq(x);
  }
  return ans;
}

// BoundingBox::updatePos(r, v, dt) (Utils.hpp:561-566)
__device__ inline double box_update1(const Box& b, int d, double r, double v, double dt) {
  return b.pbc[d] ? wrap_fmod(r + dt * v + b.L[d], b.L[d]) : r + v * dt;
}
// BoundingBox::updatePos(r, dr) (Utils.hpp:556-559)
__device__ inline double box_update_dr1(const Box& b, int d, double r, double dr) {
  return b.pbc[d] ? wrap_fmod(r + dr + b.L[d], b.L[d]) : r + dr;
}

__device__ inline double sigma2area(double a) { return kPi * sq(a / 2); }
__device__ inline double area2sigma(double a) { return 2 * sqrt(a / kPi); }
__device__ inline double sigma2volume(double a) { return 4 * kPi * ((a / 2) * (a / 2) * (a / 2)) / 3; }
__device__ inline double volume2sigma(double a) { return 2 * cbrt(3 * a / (4 * kPi)); }

__device__ inline double contact_area(double r, double a, double b) {  // Utils.hpp:236-241
  double ra = a / 2, rb = b / 2;
  double t = sqrt(4 * sq(r) * sq(ra) - sq(sq(r) - sq(rb) + sq(ra))) / (2 * r);
  return t > 0 ? kPi * sq(t) : 0;
}
__device__ inline double contact_length(double a, double b, double h) {  // Utils.hpp:243-250
  double ra = a / 2, rb = b / 2;
  double t = 2 * sqrt(sq(ra) - sq(ra - (2 * rb * h - sq(h)) / (2 * (ra + rb - h))));
  double small_r = (a <= b) ? a : b;
  return isnan(t) ? small_r : t;
}

// ---------------------------------------------------------------------------------------
// Pair-law table (per element-type pair), built on the host from cil_params.
// ---------------------------------------------------------------------------------------
struct LawEntry {
  int law;
  double rcut2;
  double eps;      // LJ: 24 sqrt(e_i e_j)
  double xi;       // SC
  double e0, ew;   // SC
  double hzk;      // HZ: 0.75*((1-v_i^2)/E_i + (1-v_j^2)/E_j)
  double ewell_i;  // HZ/EA adhesion of the lower-UID element (canonical i)
  double ea_ei;    // EA: E_i
};

struct StepParams {
  Box box;
  int nE, nspec, ntypes;
  int base[CIL_MAX_SPECIES];  // first cell slot of each species region
  int cap[CIL_MAX_SPECIES];
  LawEntry law[CIL_MAX_TYPES][CIL_MAX_TYPES];
  double cross_adh[CIL_MAX_SPECIES];
  int force_shifted;
  double dt;
  int noise, active_noise, rng, division, nwalls, drain, precision, pressure;
  uint64_t seed;
  double osmotic_pi;
  double max_disp;
  // neighbour-list mode (FP32): the list holds every pair within its cut-off + skin at the
  // last build; it is rebuilt once any particle has moved more than skin / 2 since then
  int list_mode;
  int list_max;     // list capacity per particle
  float skin;
  float skin_in;    // two-level lists: the inner list's skin (0: one level)
  float list_rca_max;  // largest rca of any type pair with a law: a pair cut-off grows by at most
                       // list_rca_max per unit of diameter growth of either particle
  cil_wall walls[CIL_MAX_WALLS];
  cil_species sp[CIL_MAX_SPECIES];
  // bins: spatial grid ns[0] x ns[1] x ns[2]; particles are counting-sorted into buckets
  // (element type, spatial bin), bucket = type * nbins + bin
  int ns[3];
  double w[3];      // bin widths (L/ns)
  double invw[3];
  int nbins;
  int nbuckets;
  int reach[3];     // largest stencil half-width per dim (0 when ns==1)
  int reach_tt[CIL_MAX_TYPES][CIL_MAX_TYPES][3];  // per (target, source) type; -1: no law
  double rca_tt[CIL_MAX_TYPES][CIL_MAX_TYPES];    // force cut-off rc = rca*(s_p+s_q) + rcb
  double rcb_tt[CIL_MAX_TYPES][CIL_MAX_TYPES];
  double sigma_bound[CIL_MAX_TYPES];              // diameters the bins were sized for
};

__device__ inline int species_of_cell(const StepParams& P, int cslot) {
  int s = 0;
  for (int k = 1; k < P.nspec; ++k)
    if (cslot >= P.base[k]) s = k;
  return s;
}

// LJ bare force modulus (MD.cuh:50-56)
template <typename T>
__device__ inline T lj_f0(T sij2, T r2) {
  T ir2 = T(1) / r2;
  T sr2n = sij2 * ir2;
  T srn = sr2n * sr2n * sr2n;
  sr2n = srn * srn;
  return (srn - T(2) * sr2n) * ir2;
}
template <typename T>
__device__ inline T sc_tprime(T y, T xi) {
  return T(6) * y * (-y + xi) / (xi * xi * xi);
}
template <typename T>
__device__ inline T sc_f0(T sigCore, T r2, T xi, T e0, T ew) {  // MD.cuh:145-164
  T r = sqrt(r2);
  T y = r - sigCore;
  if (y <= -xi) return T(0);
  if (-xi <= y && y < T(0)) return -(e0 + ew) * sc_tprime(-y, xi) / r;
  if (T(0) <= y && y < xi) return ew * sc_tprime(y, xi) / r;
  return T(0);
}

}  // namespace cil
