/*
 * cil_oracle.c — CPU oracle (FP64, strict FP) for the overdamped cell-colony Euler step.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP path and the timed CPU baseline.
 * Build with -ffp-contract=off (parity build) or -O3 -ffast-math (timing build mirroring the
 * reference's -Ofast, include.mk:5-9).  Each function cites the reference text it restates.
 *
 * Parity status per part (pins live in tests/):
 *  - dSFMT / Leva normal: pinned against the reference's own mt/dSFMT.c compiled into
 *    oracle/_ref, and by the InitConf known answers (tswitch 47049 / 139509).
 *  - Philox4x32-10: pinned by the published Random123 known-answer vectors.
 *  - Min-image, wrap, link list, pair predicates: pinned by brute force and closed forms.
 *  - Pair laws: pinned by closed forms (LJ minimum, SC polynomial, Hertz h^1.5 law).
 *  - Host cellCellForcePressure: the reference ships no source for it (only MD.o, closed);
 *    restated from the device body MD.cu:302-366 and the host sibling MD.cu:757-790.
 *    Its exact host semantics are "parity unpinned" beyond those texts.
 *  - Extensions (harmonic spring, FDT noise, linear growth, law matrix): parity unpinned
 *    against the reference (no reference text); pinned by closed forms / invariants only.
 */
#include "cil_oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define OR_PI 3.14159265358979323846
#define OR_TWOPI (2.0 * OR_PI)
#define OR_BIG_TOL_MP (1.0e3 * 10.0 * 2.220446049250313e-16) /* Utils.hpp:47-50 */
#define OR_TOL_MP (10.0 * 2.220446049250313e-16)
#define WALL_CUTOFF 0.82 /* Wall.cpp:3 */

enum { ST_WIENER = 1, ST_FDT = 2, ST_ACTIVE = 3, ST_DIV = 4, ST_ZERO = 5, ST_PAIR = 6 };
/* The stochastic Heun step evaluates forces twice per step: the corrector's evaluation (at the
 * half-step positions, before ++ts) draws its noise and its zero-distance directions from the
 * streams above with ST_PHASE added, so that it never repeats the draws of the previous step's
 * final evaluation at the same step index. */
enum { ST_PHASE = 8 };
enum { CS_DIVIDING = 0, CS_CRAWLING = 1, CS_UNDEFINED = 2 }; /* Cell.cuh:26-30 */

/* ===================================================================================== */
/* dSFMT-19937 restatement (mt/dSFMT.c: do_recursion, gen_rand_all, dsfmt_chk_init_gen_rand, */
/* initial_mask, period_certification; mt/dSFMT.h: dsfmt_genrand_close1_open2 etc.)       */
/* ===================================================================================== */
#define DS_N 191
#define DS_N64 (DS_N * 2)
#define DS_POS1 117
#define DS_SL1 19
#define DS_SR 12
#define DS_MSK1 0x000ffafffffffb3fULL
#define DS_MSK2 0x000ffdfffc90fffdULL
#define DS_FIX1 0x90014964b32f4329ULL
#define DS_FIX2 0x3b8d12ac548a7c7aULL
#define DS_PCV1 0x3d84e1ac0dc82880ULL
#define DS_PCV2 0x0000000000000001ULL
#define DS_LOW_MASK 0x000FFFFFFFFFFFFFULL
#define DS_HIGH_CONST 0x3FF0000000000000ULL

static void ds_recursion(uint64_t r[2], const uint64_t a[2], const uint64_t b[2], uint64_t lung[2]) {
  uint64_t t0 = a[0], t1 = a[1], L0 = lung[0], L1 = lung[1];
  lung[0] = (t0 << DS_SL1) ^ (L1 >> 32) ^ (L1 << 32) ^ b[0];
  lung[1] = (t1 << DS_SL1) ^ (L0 >> 32) ^ (L0 << 32) ^ b[1];
  r[0] = (lung[0] >> DS_SR) ^ (lung[0] & DS_MSK1) ^ t0;
  r[1] = (lung[1] >> DS_SR) ^ (lung[1] & DS_MSK2) ^ t1;
}

static void ds_gen_all(or_dsfmt* g) {
  uint64_t lung[2] = {g->s[DS_N][0], g->s[DS_N][1]};
  int i;
  ds_recursion(g->s[0], g->s[0], g->s[DS_POS1], lung);
  for (i = 1; i < DS_N - DS_POS1; i++) ds_recursion(g->s[i], g->s[i], g->s[i + DS_POS1], lung);
  for (; i < DS_N; i++) ds_recursion(g->s[i], g->s[i], g->s[i + DS_POS1 - DS_N], lung);
  g->s[DS_N][0] = lung[0];
  g->s[DS_N][1] = lung[1];
}

void or_dsfmt_init(or_dsfmt* g, uint32_t seed) {
  uint32_t w[(DS_N + 1) * 4];
  int i;
  w[0] = seed;
  for (i = 1; i < (DS_N + 1) * 4; i++) w[i] = 1812433253UL * (w[i - 1] ^ (w[i - 1] >> 30)) + (uint32_t)i;
  for (i = 0; i < DS_N + 1; i++) {
    g->s[i][0] = (uint64_t)w[4 * i] | ((uint64_t)w[4 * i + 1] << 32);
    g->s[i][1] = (uint64_t)w[4 * i + 2] | ((uint64_t)w[4 * i + 3] << 32);
  }
  /* initial_mask over the first DS_N words */
  for (i = 0; i < DS_N; i++) {
    g->s[i][0] = (g->s[i][0] & DS_LOW_MASK) | DS_HIGH_CONST;
    g->s[i][1] = (g->s[i][1] & DS_LOW_MASK) | DS_HIGH_CONST;
  }
  /* period_certification */
  {
    uint64_t t0 = g->s[DS_N][0] ^ DS_FIX1, t1 = g->s[DS_N][1] ^ DS_FIX2;
    uint64_t inner = (t0 & DS_PCV1) ^ (t1 & DS_PCV2);
    int k;
    for (k = 32; k > 0; k >>= 1) inner ^= inner >> k;
    inner &= 1;
    if (inner != 1) g->s[DS_N][1] ^= 1; /* DS_PCV2 has bit 0 set */
  }
  g->idx = DS_N64;
}

static uint64_t ds_next_bits(or_dsfmt* g) {
  uint64_t v;
  if (g->idx >= DS_N64) {
    ds_gen_all(g);
    g->idx = 0;
  }
  v = g->s[g->idx >> 1][g->idx & 1];
  g->idx++;
  return v;
}

static double bits_to_double(uint64_t b) {
  double d;
  memcpy(&d, &b, sizeof d);
  return d;
}

double or_dsfmt_close1_open2(or_dsfmt* g) { return bits_to_double(ds_next_bits(g)); }
double or_dsfmt_close_open(or_dsfmt* g) { return or_dsfmt_close1_open2(g) - 1.0; }
double or_dsfmt_open_open(or_dsfmt* g) { return bits_to_double(ds_next_bits(g) | 1) - 1.0; }

/* Random::normal, Leva (1992) ratio of uniforms, Random.hpp:24-52. */
double or_dsfmt_normal(or_dsfmt* g) {
  const double s = 0.449871, t = -0.386595, a = 0.19600, b = 0.25472, r1 = 0.27597, r2 = 0.27846;
  double u, v, x, y, Q;
  do {
    u = 1.0 - or_dsfmt_close_open(g);
    v = 1.7156 * (or_dsfmt_close_open(g) - 0.5);
    x = u - s;
    y = fabs(v) - t;
    Q = x * x + y * (a * y - b * x);
  } while (Q >= r1 && (Q > r2 || v * v > -4.0 * u * u * log(u)));
  return v / u;
}

void or_dsfmt_fill(uint32_t seed, int kind, double* out, size_t n) {
  or_dsfmt g;
  size_t i;
  or_dsfmt_init(&g, seed);
  for (i = 0; i < n; i++)
    out[i] = kind == 0 ? or_dsfmt_close_open(&g) : kind == 1 ? or_dsfmt_normal(&g) : or_dsfmt_open_open(&g);
}

/* ===================================================================================== */
/* Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11) — same recipe as the HIP kernels.     */
/* ===================================================================================== */
void or_philox4x32(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
  uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3], k0 = key[0], k1 = key[1];
  int r;
  for (r = 0; r < 10; r++) {
    uint64_t p0 = (uint64_t)0xD2511F53u * c0, p1 = (uint64_t)0xCD9E8D57u * c2;
    uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0, hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
    uint32_t n0 = hi1 ^ c1 ^ k0, n2 = hi0 ^ c3 ^ k1;
    c0 = n0;
    c1 = lo1;
    c2 = n2;
    c3 = lo0;
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
  out[0] = c0;
  out[1] = c1;
  out[2] = c2;
  out[3] = c3;
}

static void philox_words(uint64_t seed, uint32_t id, uint32_t ts, uint32_t stream, uint32_t k, uint32_t o[4]) {
  uint32_t ctr[4] = {id, ts, stream, k};
  uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
  or_philox4x32(ctr, key, o);
}

void or_philox_uniform2(uint64_t seed, uint32_t id, uint32_t ts, uint32_t stream, uint32_t k, double u[2]) {
  uint32_t o[4];
  philox_words(seed, id, ts, stream, k, o);
  /* [0,1) with 53 random bits */
  u[0] = (double)(((((uint64_t)o[1]) << 32) | o[0]) >> 11) * 0x1.0p-53;
  u[1] = (double)(((((uint64_t)o[3]) << 32) | o[2]) >> 11) * 0x1.0p-53;
}

void or_philox_normal2(uint64_t seed, uint32_t id, uint32_t ts, uint32_t stream, uint32_t k, double z[2]) {
  uint32_t o[4];
  double u1, u2, r, a;
  philox_words(seed, id, ts, stream, k, o);
  u1 = (double)((((((uint64_t)o[1]) << 32) | o[0]) >> 11) + 1) * 0x1.0p-53; /* (0,1] */
  u2 = (double)(((((uint64_t)o[3]) << 32) | o[2]) >> 11) * 0x1.0p-53;        /* [0,1) */
  r = sqrt(-2.0 * log(u1));
  a = OR_TWOPI * u2;
  z[0] = r * cos(a);
  z[1] = r * sin(a);
}

/* ===================================================================================== */
/* Geometry (Utils.hpp:171-255, 447-602)                                                  */
/* ===================================================================================== */
static inline int or_nint(double a) { return a >= 0.0 ? (int)(a + 0.5) : (int)(a - 0.5); } /* Utils.hpp:173-175 */
static inline double SQ(double a) { return a * a; }

typedef struct {
  int dim;
  double L[3];
  int pbc[3];
} box_t;

/* BoundingBox::distance, Utils.hpp:529-537: r12 = r2 - r1 (min image), returns |r12|^2 */
static inline double box_distance(const box_t* b, double* r12, const double* r1, const double* r2) {
  double ans = 0.0;
  int d;
  for (d = 0; d < b->dim; d++) {
    r12[d] = r2[d] - r1[d];
    if (b->pbc[d]) r12[d] -= (double)(or_nint(r12[d] / b->L[d])) * b->L[d];
    ans += SQ(r12[d]);
  }
  return ans;
}

/* BoundingBox::updatePos(r, dr), Utils.hpp:556-559 */
static inline void box_update_pos_dr(const box_t* b, double* r, const double* dr) {
  int d;
  for (d = 0; d < b->dim; d++) r[d] = b->pbc[d] ? fmod(r[d] + dr[d] + b->L[d], b->L[d]) : r[d] + dr[d];
}

/* BoundingBox::updatePos(r, v, dt), Utils.hpp:561-566 */
static inline void box_update_pos(const box_t* b, double* r, const double* v, double dt) {
  int d;
  for (d = 0; d < b->dim; d++) r[d] = b->pbc[d] ? fmod(r[d] + dt * v[d] + b->L[d], b->L[d]) : r[d] + v[d] * dt;
}

/* BoundingBox::isBoundedPos, Utils.hpp:523-527 */
static inline int box_bounded(const box_t* b, const double* r) {
  int d, ans = 1;
  for (d = 0; d < b->dim; d++) ans = ans && (r[d] >= 0.0 && r[d] < b->L[d]);
  return ans;
}

static inline double sigma2area(double a) { return OR_PI * SQ(a / 2); }                 /* Utils.hpp:227 */
static inline double area2sigma(double a) { return 2 * sqrt(a / OR_PI); }               /* Utils.hpp:252 */
static inline double sigma2volume(double a) { return 4 * OR_PI * ((a / 2) * (a / 2) * (a / 2)) / 3; } /* :218 */
static inline double volume2sigma(double a) { return 2 * cbrt(3 * a / (4 * OR_PI)); }  /* :222 */

/* twoSigma2ContactArea, Utils.hpp:236-241 */
static inline double contact_area(double r, double a, double b) {
  double ra = a / 2, rb = b / 2;
  double tempA = sqrt(4 * SQ(r) * SQ(ra) - SQ(SQ(r) - SQ(rb) + SQ(ra))) / (2 * r);
  return tempA > 0 ? OR_PI * SQ(tempA) : 0;
}
/* twoSigma2ContactLength, Utils.hpp:243-250 */
static inline double contact_length(double a, double b, double h) {
  double ra = a / 2, rb = b / 2;
  double tempA = 2 * sqrt(SQ(ra) - SQ(ra - (2 * rb * h - SQ(h)) / (2 * (ra + rb - h))));
  double small_r = (a <= b) ? a : b;
  return isnan(tempA) ? small_r : tempA;
}

/* ===================================================================================== */
/* Workspace                                                                              */
/* ===================================================================================== */
struct or_ws {
  cil_params p;
  box_t box;
  int nE;
  int rng_mode;
  or_dsfmt g;
  double* ext;
  size_t ext_per_step;
  int ext_nsteps, ext_cursor;
  int traversal;
  int nthreads;
  double* wiener[CIL_MAX_SPECIES];  /* (ncells, dim) */
  double* awiener[CIL_MAX_SPECIES]; /* (ncells, 2) */
  int wcap[CIL_MAX_SPECIES];
  int* head;
  int* list;
  int* ihead;
  int nbins_cap, list_cap;
  int ns[3];
  double width;
  double* com;
  int com_cap;
  int phase; /* 1 while the Heun corrector's forces are evaluated (stream offset ST_PHASE) */
  int sample; /* threaded gather: pair forces only for cells whose UID is a multiple of sample (0/1: all) */
  /* division records (posCellsDiv / cellT0Div, Colony.cpp:319-326; CIL.cu:396-399, 542-545) */
  int log_on;
  int log_n[CIL_MAX_SPECIES], log_cap[CIL_MAX_SPECIES];
  double* log_pos[CIL_MAX_SPECIES];
  int* log_ts[CIL_MAX_SPECIES];
  char err[256];
};

static int fail(or_ws* w, int code, const char* msg) {
  snprintf(w->err, sizeof w->err, "%s", msg);
  return code;
}
const char* or_last_error(or_ws* w) { return w->err; }
size_t or_sizeof(int which) {
  return which == 0 ? sizeof(cil_params) : which == 1 ? sizeof(cil_species) : which == 2 ? sizeof(cil_species_view)
                                                                                      : sizeof(cil_box);
}

or_ws* or_create(const cil_params* p, const cil_box* box) {
  or_ws* w = (or_ws*)calloc(1, sizeof(or_ws));
  int d;
  if (!w) return NULL;
  w->p = *p;
  w->box.dim = p->dim;
  for (d = 0; d < 3; d++) {
    w->box.L[d] = box->length[d];
    w->box.pbc[d] = box->periodic[d];
  }
  w->nE = p->species[0].nelem;
  w->rng_mode = p->rng == CIL_RNG_EXTERNAL ? OR_RNG_EXTERNAL : OR_RNG_COUNTER;
  if (p->rng == CIL_RNG_DSFMT) { /* the reference's stream from `seed`, past rng_offset outputs */
    uint64_t k;
    w->rng_mode = OR_RNG_DSFMT;
    or_dsfmt_init(&w->g, (uint32_t)p->seed);
    for (k = 0; k < p->rng_offset; k++) (void)ds_next_bits(&w->g);
  }
  w->traversal = 0;
  w->nthreads = 1;
  return w;
}

void or_destroy(or_ws* w) {
  int s;
  if (!w) return;
  for (s = 0; s < CIL_MAX_SPECIES; s++) {
    free(w->wiener[s]);
    free(w->awiener[s]);
  }
  free(w->ext);
  free(w->head);
  free(w->list);
  free(w->ihead);
  free(w->com);
  {
    int s;
    for (s = 0; s < CIL_MAX_SPECIES; s++) {
      free(w->log_pos[s]);
      free(w->log_ts[s]);
    }
  }
  free(w);
}

void or_set_rng(or_ws* w, int mode, uint32_t seed) {
  w->rng_mode = mode;
  if (mode == OR_RNG_DSFMT) or_dsfmt_init(&w->g, seed);
}
double or_rng_draw(or_ws* w, int kind) { return kind == 1 ? or_dsfmt_normal(&w->g) : or_dsfmt_close_open(&w->g); }
void or_set_traversal(or_ws* w, int mode) { w->traversal = mode; }
void or_set_sample(or_ws* w, int every) { w->sample = every; }
void or_set_division_log(or_ws* w, int on) { w->log_on = on; }

/* Records the new cell inew of species s right after its division, as the reference appends
 * posCellsDiv[n] = pos[inew] and cellT0Div[n] = ts. */
static void log_division(or_ws* w, int s, const cil_species_view* V, int inew, int ts) {
  const int per = w->nE * w->box.dim;
  int k;
  if (!w->log_on) return;
  if (w->log_n[s] == w->log_cap[s]) {
    int cap = w->log_cap[s] ? 2 * w->log_cap[s] : 64;
    w->log_pos[s] = (double*)realloc(w->log_pos[s], sizeof(double) * (size_t)cap * per);
    w->log_ts[s] = (int*)realloc(w->log_ts[s], sizeof(int) * (size_t)cap);
    w->log_cap[s] = cap;
  }
  for (k = 0; k < per; k++) w->log_pos[s][(size_t)w->log_n[s] * per + k] = V->pos[(size_t)inew * per + k];
  w->log_ts[s][w->log_n[s]] = ts;
  w->log_n[s]++;
}

int or_division_log(or_ws* w, int s, double* pos, int* ts, int cap, int* n, int clear) {
  const int per = w->nE * w->box.dim;
  int m;
  if (s < 0 || s >= CIL_MAX_SPECIES || !n) return CIL_E_INVALID;
  *n = w->log_n[s];
  m = w->log_n[s] < cap ? w->log_n[s] : cap;
  if (m > 0 && pos) memcpy(pos, w->log_pos[s], sizeof(double) * (size_t)m * per);
  if (m > 0 && ts) memcpy(ts, w->log_ts[s], sizeof(int) * (size_t)m);
  if (clear) w->log_n[s] = 0;
  return 0;
}
int or_set_threads(or_ws* w, int n) {
  w->nthreads = n < 1 ? 1 : n;
  return w->nthreads;
}

int or_set_external_noise(or_ws* w, const double* normals, size_t per_step, int nsteps) {
  free(w->ext);
  w->ext = (double*)malloc(sizeof(double) * per_step * (size_t)(nsteps > 0 ? nsteps : 1));
  if (!w->ext) return fail(w, CIL_E_INVALID, "oom");
  memcpy(w->ext, normals, sizeof(double) * per_step * (size_t)nsteps);
  w->ext_per_step = per_step;
  w->ext_nsteps = nsteps;
  w->ext_cursor = 0;
  return 0;
}

static int ensure_wiener(or_ws* w, const cil_species_view* v, int nspec) {
  int s;
  for (s = 0; s < nspec; s++) {
    if (v[s].capacity > w->wcap[s]) {
      free(w->wiener[s]);
      free(w->awiener[s]);
      w->wiener[s] = (double*)calloc((size_t)v[s].capacity * 3, sizeof(double));
      w->awiener[s] = (double*)calloc((size_t)v[s].capacity * 2, sizeof(double));
      if (!w->wiener[s] || !w->awiener[s]) return fail(w, CIL_E_INVALID, "oom");
      w->wcap[s] = v[s].capacity;
    }
  }
  return 0;
}

/* uniform draw in [0, x) for the current RNG source (Random::uniform0x, Random.hpp:13-15) */
static double rng_uniform0x(or_ws* w, double x, uint32_t id, uint32_t ts, uint32_t stream, uint32_t k, int half) {
  if (w->rng_mode == OR_RNG_DSFMT) return or_dsfmt_close_open(&w->g) * x;
  {
    double u[2];
    or_philox_uniform2(w->p.seed, id, ts, stream, k, u);
    return u[half] * x;
  }
}
static double rng_normal(or_ws* w, uint32_t id, uint32_t ts, uint32_t stream, uint32_t k, int half) {
  if (w->rng_mode == OR_RNG_DSFMT) return or_dsfmt_normal(&w->g);
  {
    double z[2];
    or_philox_normal2(w->p.seed, id, ts, stream, k, z);
    return z[half];
  }
}

/* Random direction on the D-sphere.  dSFMT source: Random::onDSphere (Random.hpp:82-114,
 * rejection in the unit ball with Random::uniform()).  Counter source: normalised Gaussian. */
static void rng_on_sphere_k(or_ws* w, double* n, uint32_t id, uint32_t ts, uint32_t stream, uint32_t k0);
static void rng_on_sphere(or_ws* w, double* n, uint32_t id, uint32_t ts, uint32_t stream) {
  rng_on_sphere_k(w, n, id, ts, stream, 0);
}
/* counter source: Philox blocks k0 and k0 + 1 of (id, ts, stream) */
static void rng_on_sphere_k(or_ws* w, double* n, uint32_t id, uint32_t ts, uint32_t stream, uint32_t k0) {
  int d, D = w->box.dim;
  double r2 = 0.0, norm;
  if (w->rng_mode == OR_RNG_DSFMT) {
    do {
      r2 = 0.0;
      for (d = 0; d < D; d++) {
        n[d] = (or_dsfmt_open_open(&w->g) - 0.5) * 2.0;
        r2 += n[d] * n[d];
      }
    } while (!(r2 <= 1.0));
  } else {
    double z[4];
    or_philox_normal2(w->p.seed, id, ts, stream, k0, z);
    or_philox_normal2(w->p.seed, id, ts, stream, k0 + 1, z + 2);
    for (d = 0; d < D; d++) {
      n[d] = z[d];
      r2 += n[d] * n[d];
    }
  }
  norm = 1.0 / sqrt(r2);
  for (d = 0; d < D; d++) n[d] *= norm;
}

static inline uint32_t cell_key(int s, int c) { return ((uint32_t)s << 28) | (uint32_t)c; }

/* ===================================================================================== */
/* Pair laws (MD.cuh:47-344) and combination tables (MD.cu:7-245)                         */
/* ===================================================================================== */
static inline double lj_f0(double sij2, double r2) { /* LJParams::elemForce0, MD.cuh:50-56 */
  double ir2 = 1.0 / r2;
  double sr2n = sij2 * ir2;
  double srn = sr2n * sr2n * sr2n;
  sr2n = srn * srn;
  return (srn - 2.0 * sr2n) * ir2;
}
static inline double sc_tprime(double y, double xi) { return 6.0 * y * (-y + xi) / (xi * xi * xi); } /* MD.cuh:137-139 */
static inline double sc_f0(double sigCore, double r2, double xi, double e0, double ew) { /* MD.cuh:145-164 */
  double r = sqrt(r2);
  double y = r - sigCore;
  if (y <= -xi) return 0.0;
  if (-xi <= y && y < 0.0) return -(e0 + ew) * sc_tprime(-y, xi) / r;
  if (0 <= y && y < xi) return ew * sc_tprime(y, xi) / r;
  return 0.0;
}

typedef struct {
  int law;
  double sij2;    /* LJ: sigma_ij^2 */
  double eps;     /* LJ: 24 sqrt(e_i e_j); SC: E0_ij */
  double ew;      /* SC: Ew_ij */
  double xi;      /* SC: xi_ij */
  double rcut2;
} law_tab;

/* Element parameters of type t (species s = t / nE, element e = t % nE). */
static inline const cil_species* spec_of(const cil_params* p, int t, int nE) { return &p->species[t / nE]; }

/* LJParams/SCParams::elemForce (MD.cuh:89-110, 190-207) and the contact laws of the device
 * body MD.cu:302-366 (HZ MD.cuh:240-252, EA MD.cuh:296-304).  For HZ/EA, r2 is the squared
 * distance here and the contact overlap h is formed inside, as MD.cu:330-334 does. */
int or_pair_f0(const cil_params* p, int dim, int ti, int tj, double si, double sj, double r2, double* f0) {
  int nE = p->species[0].nelem;
  int law = p->law[ti][tj];
  const cil_species* Si = spec_of(p, ti, nE);
  const cil_species* Sj = spec_of(p, tj, nE);
  int ei = ti % nE, ej = tj % nE;
  double rc2;
  (void)dim;
  *f0 = 0.0;
  if (law == CIL_LAW_LJ) {
    double sij2 = SQ(0.5 * (si + sj));
    double eps = 24.0 * sqrt(Si->eps_well[ei] * Sj->eps_well[ej]);
    rc2 = sij2 * p->rcut2[ti][tj];
    if (r2 <= rc2) {
      double f = lj_f0(sij2, r2) - (p->force_shifted ? lj_f0(sij2, rc2) * sqrt(rc2 / r2) : 0.0);
      *f0 = f * eps;
      return 1;
    }
    return 0;
  }
  if (law == CIL_LAW_SC) {
    double sigCore = 0.5 * (si + sj);
    double xi = 0.5 * (Si->sig_well[ei] + Sj->sig_well[ej]);
    double e0 = sqrt(Si->eps_core[ei] * Sj->eps_core[ej]);
    double ew = sqrt(Si->eps_well[ei] * Sj->eps_well[ej]);
    rc2 = p->rcut2[ti][tj] * SQ(sigCore + xi);
    if (r2 > p->rcut2[ti][tj] * SQ(sigCore - xi) && r2 <= rc2) {
      *f0 = sc_f0(sigCore, r2, xi, e0, ew) - (p->force_shifted ? sc_f0(sigCore, rc2, xi, e0, ew) * sqrt(rc2 / r2) : 0.0);
      return 1;
    }
    return 0;
  }
  if (law == CIL_LAW_HZ || law == CIL_LAW_EA) {
    double r = sqrt(r2);
    if (r <= (si + sj) / 2) {
      double h = (si + sj) / 2 - r;
      if (h < 0.0) h = 0.0;
      if (law == CIL_LAW_HZ) {
        double Ri = si / 2, Rj = sj / 2, vi = Si->sig_well[ei], vj = Sj->sig_well[ej];
        double Ei = Si->eps_core[ei], Ej = Sj->eps_core[ej];
        double fEl = pow(h, 1.5) / ((0.75) * ((1 - SQ(vi)) / Ei + (1 - SQ(vj)) / Ej) * (sqrt((1 / Ri) + (1 / Rj))));
        *f0 = -fEl;
      } else {
        *f0 = -(h * Si->eps_core[ei]);
      }
      return 1;
    }
    return 0;
  }
  return 0;
}

/* One element pair in canonical orientation (i = lower UID), accumulating into
 * fi/fj and pri/prj.  LJ/SC: host cellCellForce semantics (MD.cu:757-790).  HZ/EA: device
 * cellCellForcePressureDevice semantics (MD.cu:317-362) with DIM in place of MAX_DIM.  */
static void elem_pair(or_ws* w, int ti, int tj, int spi, int spj, const double* ri, const double* rj, double si,
                      double sj, double* fi, double* fj, double* pri, double* prj, double* sci, double* scj,
                      uint32_t pid_i, uint32_t pid_j, int ts) {
  const cil_params* p = &w->p;
  int D = w->box.dim, d;
  int law = p->law[ti][tj];
  double rij[3] = {0, 0, 0};
  double r2 = box_distance(&w->box, rij, ri, rj);
  double f0;
  double cai = p->species[spi].cross_adhesion, caj = p->species[spj].cross_adhesion;
  if (law == CIL_LAW_NONE) return;
  if (law == CIL_LAW_LJ || law == CIL_LAW_SC) {
    if (!or_pair_f0(p, D, ti, tj, si, sj, r2, &f0)) return;
    if (f0 < 0.0 || spi == spj) {
      for (d = 0; d < D; d++) {
        fi[d] += f0 * rij[d];
        fj[d] -= f0 * rij[d];
      }
      if (sci) {
        double m = fabs(f0) * sqrt(r2);
        *sci += m;
        *scj += m;
      }
    } else {
      for (d = 0; d < D; d++) {
        fi[d] += cai * f0 * rij[d];
        fj[d] -= caj * f0 * rij[d];
      }
      if (sci) {
        *sci += fabs(cai * f0) * sqrt(r2);
        *scj += fabs(caj * f0) * sqrt(r2);
      }
    }
    return;
  }
  {
    double r = sqrt(r2);
    int nE = w->nE;
    const cil_species* Si = spec_of(p, ti, nE);
    double A, temp;
    if (!(r <= (si + sj) / 2)) return;
    if (D == 2 && rij[0] == 0 && rij[1] == 0) { /* MD.cu:321-331 */
      /* counter draw in every RNG mode: the reference's host body that would draw it from the
       * sequential stream is absent (its device body uses curand, MD.cu:321-331) */
      double u2[2], alpha;
      or_philox_uniform2(w->p.seed, pid_i, (uint32_t)ts, ST_PAIR + ST_PHASE * w->phase, pid_j, u2);
      alpha = u2[0] * OR_TWOPI;
      rij[0] = sin(alpha);
      rij[1] = cos(alpha);
    }
    or_pair_f0(p, D, ti, tj, si, sj, r2, &f0);
    {
      double h = (si + sj) / 2 - r;
      if (h < 0.0) h = 0.0;
      A = D == 3 ? contact_area(r, si, sj) : contact_length(si, sj, h);
    }
    f0 += Si->eps_well[ti % nE] * A;
    temp = (A == 0 ? 0 : (4 * OR_PI * SQ(si / 2) / 3));
    temp = (D == 3 ? temp : A);
    if (f0 < 0.0 || spi == spj) {
      if (temp != 0.0 && p->pressure) {
        *pri -= f0 / temp;
        *prj -= f0 / temp;
      }
      for (d = 0; d < D; d++) {
        fi[d] += f0 * rij[d];
        fj[d] -= f0 * rij[d];
      }
      if (sci) {
        double m = fabs(f0) * r;
        *sci += m;
        *scj += m;
      }
    } else {
      if (temp != 0.0 && p->pressure) {
        *pri -= cai * f0 / temp;
        *prj -= caj * f0 / temp;
      }
      for (d = 0; d < D; d++) {
        fi[d] += cai * f0 * rij[d];
        fj[d] -= caj * f0 * rij[d];
      }
      if (sci) {
        *sci += fabs(cai * f0) * r;
        *scj += fabs(caj * f0) * r;
      }
    }
  }
}

/* ===================================================================================== */
/* UID helpers (Colony.cpp:22-67)                                                         */
/* ===================================================================================== */
typedef struct {
  int nspec;
  int cumul[CIL_MAX_SPECIES + 1];
} uidmap;

static void make_uidmap(uidmap* m, const cil_species_view* v, int nspec) {
  int s;
  m->nspec = nspec;
  m->cumul[0] = 0;
  for (s = 0; s < nspec; s++) m->cumul[s + 1] = m->cumul[s] + v[s].ncells;
}
static inline void uid2cell(const uidmap* m, int uid, int* s, int* c) {
  int k = 0;
  while (k + 1 < m->nspec && uid >= m->cumul[k + 1]) k++;
  *s = k;
  *c = uid - m->cumul[k];
}

/* cellCellForcePressure over element pairs of two cells (canonical orientation). */
static void cell_cell(or_ws* w, cil_species_view* v, int si_, int ci, int sj_, int cj, int ts) {
  int nE = w->nE, D = w->box.dim, ei, ej;
  for (ei = 0; ei < nE; ei++) {
    for (ej = 0; ej < nE; ej++) {
      size_t pi = (size_t)ci * nE + ei, pj = (size_t)cj * nE + ej;
      elem_pair(w, si_ * nE + ei, sj_ * nE + ej, si_, sj_, &v[si_].pos[pi * D], &v[sj_].pos[pj * D],
                v[si_].sigma[pi], v[sj_].sigma[pj], &v[si_].frc[pi * D], &v[sj_].frc[pj * D], &v[si_].pressure[ci],
                &v[sj_].pressure[cj], NULL, NULL, (uint32_t)((cell_key(si_, ci) << 1) | ei),
                (uint32_t)((cell_key(sj_, cj) << 1) | ej), ts);
    }
  }
}

/* ===================================================================================== */
/* Link list over cell centres of mass (Container.cpp:81-103, Utils.hpp:872-1069)         */
/* ===================================================================================== */
static const int DNS2D[10] = {0, 0, 1, 0, 1, 1, 0, 1, -1, 1}; /* Utils.hpp:700 */
static const int DNS3D[42] = {0, 0, 0, 1, 0, 0, 1, 1, 0, 0, 1, 0, -1, 1, 0, 0, 0, 1, 1, 0, 1,
                              1, 1, 1, 0, 1, 1, -1, 1, 1, -1, 0, 1, -1, -1, 1, 0, -1, 1, 1, -1, 1}; /* :701-704 */

/* Largest pair cut-off over the actual per-disc diameters (the reference sizes the list from
 * species defaults, CIL.cuh:73-80; per-disc polydispersity needs the actual maximum). */
static double max_cutoff(or_ws* w, const cil_species_view* v, int nspec) {
  const cil_params* p = &w->p;
  int nE = w->nE, ti, tj, s;
  double smax[CIL_MAX_TYPES] = {0}, rc = 0.0;
  for (s = 0; s < nspec; s++) {
    int c, e;
    for (c = 0; c < v[s].ncells; c++)
      for (e = 0; e < nE; e++) {
        double sg = v[s].sigma[(size_t)c * nE + e];
        if (sg > smax[s * nE + e]) smax[s * nE + e] = sg;
      }
  }
  for (ti = 0; ti < nspec * nE; ti++)
    for (tj = 0; tj < nspec * nE; tj++) {
      int law = p->law[ti][tj];
      double sbar = 0.5 * (smax[ti] + smax[tj]), r = 0.0;
      if (law == CIL_LAW_LJ) r = sqrt(p->rcut2[ti][tj]) * sbar;
      else if (law == CIL_LAW_SC) {
        const cil_species* Si = spec_of(p, ti, nE);
        const cil_species* Sj = spec_of(p, tj, nE);
        double xi = 0.5 * (Si->sig_well[ti % nE] + Sj->sig_well[tj % nE]);
        r = sqrt(p->rcut2[ti][tj]) * (sbar + xi);
      } else if (law == CIL_LAW_HZ || law == CIL_LAW_EA)
        r = sbar;
      if (r > rc) rc = r;
    }
  return rc;
}

/* Largest element-to-element extent within a cell (bond length); the reference uses
 * sqrt(max r2max) (CIL.cuh:125-136). */
static double max_extent(or_ws* w, const cil_species_view* v, int nspec) {
  int s, c, D = w->box.dim;
  double m = 0.0;
  if (w->nE < 2) return 0.0;
  for (s = 0; s < nspec; s++)
    for (c = 0; c < v[s].ncells; c++) {
      double rij[3];
      double r2 = box_distance(&w->box, rij, &v[s].pos[(size_t)c * 2 * D], &v[s].pos[((size_t)c * 2 + 1) * D]);
      if (r2 > m) m = r2;
    }
  return sqrt(m);
}

static int build_linklist(or_ws* w, cil_species_view* v, int nspec, const uidmap* um, int* nboxes_out) {
  int D = w->box.dim, d, ncellsTot = um->cumul[nspec], i, nb = 1, overkill = 0;
  double width = max_cutoff(w, v, nspec) + max_extent(w, v, nspec);
  double lmin = w->box.L[0];
  for (d = 1; d < D; d++)
    if (w->box.L[d] < lmin) lmin = w->box.L[d];
  {
    double delta = lmin * 1.0e-8, ir = 1.0 / (width + delta); /* get_opt_size, Utils.hpp:872-876 */
    for (d = 0; d < D; d++) {
      w->ns[d] = (int)(w->box.L[d] * ir);
      if (w->ns[d] < 3) overkill = 1;
    }
  }
  w->width = width;
  if (overkill) { /* Utils.hpp:897-901 and CIL.cuh:158-162: list disabled */
    for (d = 0; d < D; d++) w->ns[d] = 1;
    *nboxes_out = 0;
    return 1;
  }
  for (d = 0; d < D; d++) nb *= w->ns[d];
  if (nb > w->nbins_cap) {
    free(w->head);
    free(w->ihead);
    w->head = (int*)malloc(sizeof(int) * nb);
    w->ihead = (int*)malloc(sizeof(int) * nb);
    w->nbins_cap = nb;
  }
  if (ncellsTot > w->list_cap) {
    free(w->list);
    w->list = (int*)malloc(sizeof(int) * (ncellsTot > 0 ? ncellsTot : 1));
    w->list_cap = ncellsTot;
  }
  if (ncellsTot * D > w->com_cap) {
    free(w->com);
    w->com = (double*)malloc(sizeof(double) * (size_t)(ncellsTot > 0 ? ncellsTot : 1) * D);
    w->com_cap = ncellsTot * D;
  }
  or_coms(w, v, nspec, w->com);
  for (i = 0; i < nb; i++) w->head[i] = -1;
  for (i = 0; i < ncellsTot; i++) w->list[i] = -1;
  for (i = 0; i < ncellsTot; i++) { /* populate_list, Utils.hpp:982-1001 */
    int ls[3] = {0, 0, 0}, k;
    for (d = 0; d < D; d++) {
      double iH = 1.0 / w->box.L[d];
      int b = (int)((w->com[(size_t)i * D + d] * iH) * w->ns[d]);
      ls[d] = b < w->ns[d] - 1 ? b : w->ns[d] - 1;
    }
    k = D == 2 ? ls[0] * w->ns[1] + ls[1] : (ls[0] * w->ns[1] + ls[1]) * w->ns[2] + ls[2];
    if (w->head[k] == -1) w->head[k] = i;
    else {
      w->list[i] = w->head[k];
      w->head[k] = i;
    }
  }
  /* LINKLIST_SORTED: boxes ordered by head descending (Container.cpp:93-100) */
  {
    int endCell = 0, b;
    /* heads are unique among non-empty boxes: counting by value gives the same order as std::sort */
    int* byhead = (int*)malloc(sizeof(int) * (ncellsTot > 0 ? ncellsTot : 1));
    for (i = 0; i < ncellsTot; i++) byhead[i] = -1;
    for (b = 0; b < nb; b++)
      if (w->head[b] >= 0) byhead[w->head[b]] = b;
    for (i = ncellsTot - 1; i >= 0; i--)
      if (byhead[i] >= 0) w->ihead[endCell++] = byhead[i];
    free(byhead);
    *nboxes_out = endCell;
  }
  return 0;
}

static void cell_coord(const or_ws* w, int ic, int* ixyz) {
  if (w->box.dim == 2) {
    ixyz[0] = ic / w->ns[1];
    ixyz[1] = ic - w->ns[1] * ixyz[0];
  } else {
    int ii;
    ixyz[0] = ic / (w->ns[1] * w->ns[2]);
    ii = ic - ixyz[0] * w->ns[1] * w->ns[2];
    ixyz[1] = ii / w->ns[2];
    ixyz[2] = ii - ixyz[1] * w->ns[2];
  }
}
static int neighbor_id(const or_ws* w, const int* ixyz, const int* dxyz) { /* get_cellID, Utils.hpp:918-936 */
  int ls[3] = {0, 0, 0}, d, D = w->box.dim, in = 1;
  for (d = 0; d < D; d++) {
    ls[d] = ixyz[d] + dxyz[d];
    if (w->box.pbc[d]) ls[d] = (ls[d] + w->ns[d]) % w->ns[d];
    else if (ls[d] < 0 || ls[d] >= w->ns[d]) in = 0;
  }
  if (!in) return -1;
  return D == 2 ? ls[0] * w->ns[1] + ls[1] : (ls[0] * w->ns[1] + ls[1]) * w->ns[2] + ls[2];
}

/* Visit candidate cell pairs as calculInterForces does (CIL.cu:177-263).  The callback gets
 * the pair in traversal orientation (ii first). */
typedef void (*pair_cb)(or_ws* w, cil_species_view* v, const uidmap* um, int ii, int jj, void* ctx);

static void traverse(or_ws* w, cil_species_view* v, int nspec, const uidmap* um, pair_cb cb, void* ctx) {
  int ncellsTot = um->cumul[nspec];
  int nboxes = 0, disabled = 1;
  if (w->traversal == 0) disabled = build_linklist(w, v, nspec, um, &nboxes);
  if (!disabled) {
    int D = w->box.dim, nnb = D == 2 ? 5 : 14, iic, dc;
    const int* dns = D == 2 ? DNS2D : DNS3D;
    for (iic = 0; iic < nboxes; iic++) {
      int ic = w->ihead[iic], icell[3];
      cell_coord(w, ic, icell);
      for (dc = 0; dc < nnb; dc++) {
        int jc = neighbor_id(w, icell, &dns[dc * D]), ii;
        if (jc == -1) continue;
        for (ii = w->head[ic]; ii >= 0; ii = w->list[ii]) {
          int jj;
          for (jj = w->head[jc]; jj >= 0; jj = w->list[jj])
            if ((ic != jc || jj > ii)) cb(w, v, um, ii, jj, ctx);
        }
      }
    }
  } else {
    int ii, jj;
    for (ii = 0; ii < ncellsTot; ii++)
      for (jj = ii + 1; jj < ncellsTot; jj++) cb(w, v, um, ii, jj, ctx);
  }
}

static void force_cb(or_ws* w, cil_species_view* v, const uidmap* um, int ii, int jj, void* ctx) {
  int si, ci, sj, cj;
  int ts = *(int*)ctx;
  uid2cell(um, ii, &si, &ci);
  uid2cell(um, jj, &sj, &cj);
  if (si == sj && ci == cj) return;
  cell_cell(w, v, si, ci, sj, cj, ts);
}

/* ===================================================================================== */
/* Intra-cell forces (CIL.cu:30-100; MD.cu:525-644; Wall.hpp:68-79)                       */
/* ===================================================================================== */
static void binary_cell_force(or_ws* w, cil_species_view* v, int s, int c, int ts) {
  const cil_species* P = &w->p.species[s];
  int D = w->box.dim, d;
  double* p0 = &v[s].pos[((size_t)c * 2) * D];
  double* p1 = &v[s].pos[((size_t)c * 2 + 1) * D];
  double* f0 = &v[s].frc[((size_t)c * 2) * D];
  double* f1 = &v[s].frc[((size_t)c * 2 + 1) * D];
  double* v0 = &v[s].vel[((size_t)c * 2) * D];
  double* v1 = &v[s].vel[((size_t)c * 2 + 1) * D];
  const double* sig = &v[s].sigma[(size_t)c * 2];
  double rij[3] = {0, 0, 0}, nij[3], rnd[3] = {0, 0, 0};
  double r2 = box_distance(&w->box, rij, p0, p1); /* rij = r1 - r0 */
  double r = sqrt(r2), spring, sc;
  double act0 = P->motility[0], act1 = P->motility[1];
  if (r2 == 0.0) { /* MD.cu:534-540 */
    r = OR_BIG_TOL_MP * 0.5 * (P->sigma[0] + P->sigma[1]);
    rng_on_sphere(w, rij, cell_key(s, c), (uint32_t)ts, ST_ZERO + ST_PHASE * w->phase);
    for (d = 0; d < D; d++) rij[d] *= r;
    r2 = SQ(r);
    box_update_pos_dr(&w->box, p1, rij);
  }
  for (d = 0; d < 3; d++) nij[d] = rij[d];
  if (P->spring == CIL_SPRING_HARMONIC) {
    double l = v[s].rest_len ? v[s].rest_len[c] : 0.0;
    spring = -P->k_harm * (r - l) / r;
    sc = 1.0;
    for (d = 0; d < D; d++) nij[d] = rij[d] / r;
  } else {
    /* the Dividing state (MD.cu:547,560-568): the bond's r2max becomes divisionR2max, both
     * discs' activity becomes -growthk0 fene(divisionR2min) + growthk1 + growthk2 r2^2 along
     * the unscaled axis, and the contraction factor drops from 2 to 1 */
    const int dividing = v[s].state && v[s].state[c] == CS_DIVIDING;
    double r2max0 = dividing ? P->div_r2max : P->r2max, smax = sig[0] >= sig[1] ? sig[0] : sig[1];
    double r2max = r2max0 > smax ? r2max0 : smax;
    double ratio = P->spring_ratio;
    if (!(r2max - r2 > OR_TOL_MP)) snprintf(w->err, sizeof w->err, "FENE bond over-stretched");
    spring = r2 <= (r2max * ratio) ? P->kappa_push * fabs(r2 - (r2max * ratio))
                                   : -P->kappa / (1.0 - SQ((r2 - r2max * ratio) / (r2max * (1 - ratio))));
    if (dividing) {
      double fene_div_max = -P->kappa / (1.0 - SQ(P->division_r2min / r2max)); /* FENE, MD.cu:255-257 */
      act0 = (-P->growthk0[0] * fene_div_max + P->growthk1[0] + P->growthk2[0] * SQ(r2));
      act1 = (-P->growthk0[1] * fene_div_max + P->growthk1[1] + P->growthk2[1] * SQ(r2));
      sc = 1.0;
    } else {
      sc = 2.0;
      if (!P->cil) {
        double scaling = P->rss / r;
        for (d = 0; d < D; d++) nij[d] *= scaling;
      }
    }
  }
  if (w->p.active_noise && D == 2) { /* MD.cu:573-578 */
    const double* aw = &w->awiener[s][(size_t)c * 2];
    double Wr = aw[0], Wt = aw[1];
    rnd[0] = 0.5 * (Wr * rij[0] / r - Wt * rij[1]);
    rnd[1] = 0.5 * (Wr * rij[1] / r + Wt * rij[0]);
  }
  for (d = 0; d < D; d++) {
    double tc0 = -(act0 * nij[d] + spring * rij[d] * sc + rnd[d] * rij[d]);
    double tc1 = (act1 * nij[d] + spring * rij[d] * sc + rnd[d] * rij[d]);
    if (P->quad_drag) {
      double tf0 = P->zeta[0] * SQ(v0[d]) * (v0[d] / fabs(v0[d]));
      double tf1 = P->zeta[1] * SQ(v1[d]) * (v1[d] / fabs(v1[d]));
      f0[d] = (isinf(tf0) || isnan(tf0)) ? tc0 : (tc0 - tf0);
      f1[d] = (isinf(tf1) || isnan(tf1)) ? tc1 : (tc1 - tf1);
    } else {
      f0[d] = tc0;
      f1[d] = tc1;
    }
  }
}

static void single_cell_force(or_ws* w, cil_species_view* v, int s, int c, int ts) { /* MD.cu:600-644 */
  const cil_species* P = &w->p.species[s];
  int D = w->box.dim, d;
  double* v0 = &v[s].vel[(size_t)c * D];
  double* f0 = &v[s].frc[(size_t)c * D];
  double vl = 0.0;
  for (d = 0; d < D; d++) vl += v0[d] * v0[d];
  vl = sqrt(vl);
  if (vl == 0.0) {
    if (w->rng_mode == OR_RNG_DSFMT || P->motility[0] != 0.0) {
      rng_on_sphere(w, v0, cell_key(s, c) << 1, (uint32_t)ts, ST_ZERO + ST_PHASE * w->phase);
      for (d = 0; d < D; d++) v0[d] *= OR_BIG_TOL_MP;
    }
    vl = 0.0;
    for (d = 0; d < D; d++) vl += v0[d] * v0[d];
    vl = sqrt(vl);
  }
  for (d = 0; d < D; d++) {
    double n = vl > 0.0 ? v0[d] / vl : 0.0;
    f0[d] = P->motility[0] * n;
  }
}

static void flat_wall(const cil_wall* wl, double sigma, const double* r, double* f) { /* Wall.hpp:68-79 */
  double dr = (wl->positive ? (r[wl->axis] - wl->x0) : (wl->x0 - r[wl->axis])) / sigma;
  if (dr <= wl->rcut) {
    double ir2, f0;
    dr = dr >= WALL_CUTOFF ? dr : WALL_CUTOFF;
    ir2 = 1.0 / (dr * dr);
    f0 = ir2 * ir2 * ir2;
    f0 = wl->epsilon * (2.0 * f0 - 1.0) * f0 * ir2 / SQ(sigma);
    f0 = wl->positive ? f0 : -f0;
    f[wl->axis] += f0; /* the wall's pressure term is reset by calculInterForces (CIL.cu:113-118) */
  }
}

static void intra_forces(or_ws* w, cil_species_view* v, int nspec, int ts) { /* CIL.cu:30-100 */
  int s, c, e, d, k, D = w->box.dim, nE = w->nE;
  for (s = 0; s < nspec; s++) {
    const cil_species* P = &w->p.species[s];
#pragma omp parallel for num_threads(w->nthreads) if (w->nthreads > 1 && w->rng_mode != OR_RNG_DSFMT)
    for (c = 0; c < v[s].ncells; c++) {
      if (nE == 2) binary_cell_force(w, v, s, c, ts);
      else single_cell_force(w, v, s, c, ts);
    }
    if (w->p.nwalls > 0)
      for (c = 0; c < v[s].ncells; c++)
        for (e = 0; e < nE; e++)
          for (k = 0; k < w->p.nwalls; k++)
            flat_wall(&w->p.walls[k], v[s].sigma[(size_t)c * nE + e], &v[s].pos[((size_t)c * nE + e) * D],
                      &v[s].frc[((size_t)c * nE + e) * D]);
    if (w->p.noise == CIL_NOISE_CELL_WIENER)
      for (c = 0; c < v[s].ncells; c++)
        for (e = 0; e < nE; e++)
          for (d = 0; d < D; d++) v[s].frc[((size_t)c * nE + e) * D + d] += P->zeta[e] * w->wiener[s][(size_t)c * D + d];
  }
}

/* updateWienerProcess, CIL.cu:793-818 */
static int draw_noise(or_ws* w, cil_species_view* v, int nspec, int ts) {
  int s, c, d, D = w->box.dim;
  const double* ext = NULL;
  size_t pos = 0;
  if (w->p.noise != CIL_NOISE_CELL_WIENER && !w->p.active_noise) return 0;
  if (w->rng_mode == OR_RNG_EXTERNAL) {
    if (w->ext_cursor >= w->ext_nsteps) return fail(w, CIL_E_INVALID, "external noise exhausted");
    ext = w->ext + (size_t)w->ext_cursor * w->ext_per_step;
    w->ext_cursor++;
  }
  if (w->p.noise == CIL_NOISE_CELL_WIENER)
    for (s = 0; s < nspec; s++) {
      const cil_species* P = &w->p.species[s];
      if (!(P->d_noise > 0.0)) continue;
      for (c = 0; c < v[s].ncells; c++)
        for (d = 0; d < D; d++) {
          double z = ext ? ext[pos++] : rng_normal(w, cell_key(s, c), (uint32_t)ts, ST_WIENER + ST_PHASE * w->phase, (uint32_t)(d >> 1), d & 1);
          w->wiener[s][(size_t)c * D + d] = P->d_noise * z;
        }
    }
  if (w->p.active_noise)
    for (s = 0; s < nspec; s++) {
      const cil_species* P = &w->p.species[s];
      if (w->nE < 2) continue;
      if (P->d_noise_radial > 0.0)
        for (c = 0; c < v[s].ncells; c++)
          w->awiener[s][(size_t)c * 2] =
              P->d_noise_radial * (ext ? ext[pos++] : rng_normal(w, cell_key(s, c), (uint32_t)ts, ST_ACTIVE + ST_PHASE * w->phase, 0, 0));
      if (P->d_noise_angular > 0.0)
        for (c = 0; c < v[s].ncells; c++)
          w->awiener[s][(size_t)c * 2 + 1] =
              P->d_noise_angular * (ext ? ext[pos++] : rng_normal(w, cell_key(s, c), (uint32_t)ts, ST_ACTIVE + ST_PHASE * w->phase, 0, 1));
    }
  if (ext && pos != w->ext_per_step) return fail(w, CIL_E_INVALID, "external noise block size mismatch");
  return 0;
}

/* Race-free multi-threaded variant of the link-list traversal (CIL.cu:177-231 has OpenMP
 * pragmas but writes both cells of a pair; this gathers instead): every cell visits the full
 * 3^D box shell around its own box and accumulates only its own force and pressure.  Pair
 * terms are evaluated in canonical orientation (lower UID first).  Counter RNG only. */
static void inter_forces_gather(or_ws* w, cil_species_view* v, int nspec, const uidmap* um, int ts, double* scale) {
  int D = w->box.dim, nE = w->nE, ncellsTot = um->cumul[nspec], nboxes = 0;
  int disabled = build_linklist(w, v, nspec, um, &nboxes);
  int ii;
#pragma omp parallel for schedule(dynamic, 64) num_threads(w->nthreads)
  for (ii = 0; ii < ncellsTot; ii++) {
    int si, ci, icell[3] = {0, 0, 0}, ox, oy, oz, d, ib = 0;
    double dumf[3 * CIL_MAX_ELEMS], dump = 0.0;
    if (w->sample > 1 && ii % w->sample != 0) continue;
    uid2cell(um, ii, &si, &ci);
    if (!disabled) {
      const double* cm = &w->com[(size_t)ii * D];
      for (d = 0; d < D; d++) {
        int b = (int)((cm[d] * (1.0 / w->box.L[d])) * w->ns[d]);
        icell[d] = b < w->ns[d] - 1 ? b : w->ns[d] - 1;
      }
      ib = D == 2 ? icell[0] * w->ns[1] + icell[1] : (icell[0] * w->ns[1] + icell[1]) * w->ns[2] + icell[2];
    }
    for (ox = -1; ox <= 1; ox++)
      for (oy = -1; oy <= 1; oy++)
        for (oz = (D == 3 ? -1 : 0); oz <= (D == 3 ? 1 : 0); oz++) {
          int jj, jstart;
          if (!disabled) {
            int dxyz[3] = {ox, oy, oz}, jc;
            /* skip offsets that wrap onto an already visited box (ns == 3 is fine) */
            jc = neighbor_id(w, icell, dxyz);
            if (jc == -1) continue;
            jstart = w->head[jc];
          } else {
            if (ox != 0 || oy != 0 || oz != 0) continue;
            jstart = 0;
          }
          for (jj = jstart; jj >= 0 && jj < ncellsTot; jj = disabled ? jj + 1 : w->list[jj]) {
            int sj, cj, ei, ej;
            if (jj == ii) continue;
            uid2cell(um, jj, &sj, &cj);
            for (ei = 0; ei < nE; ei++)
              for (ej = 0; ej < nE; ej++) {
                size_t pi = (size_t)ci * nE + ei, pj = (size_t)cj * nE + ej;
                uint32_t ki = (uint32_t)((cell_key(si, ci) << 1) | ei), kj = (uint32_t)((cell_key(sj, cj) << 1) | ej);
                if (scale) { /* the force scale of or_force_scale, gathered */
                  double rij[3], f0, r2 = box_distance(&w->box, rij, &v[si].pos[pi * D], &v[sj].pos[pj * D]);
                  int ti = si * nE + ei, tj = sj * nE + ej;
                  if (w->p.law[ti][tj] != CIL_LAW_NONE &&
                      or_pair_f0(&w->p, D, ti, tj, v[si].sigma[pi], v[sj].sigma[pj], r2, &f0))
                    scale[(size_t)ii * nE + ei] += fabs(f0) * sqrt(r2) + 1e-300;
                  continue;
                }
                for (d = 0; d < 3 * CIL_MAX_ELEMS; d++) dumf[d] = 0.0;
                if (ii < jj)
                  elem_pair(w, si * nE + ei, sj * nE + ej, si, sj, &v[si].pos[pi * D], &v[sj].pos[pj * D],
                            v[si].sigma[pi], v[sj].sigma[pj], &v[si].frc[pi * D], dumf, &v[si].pressure[ci], &dump, NULL,
                            NULL, ki, kj, ts);
                else
                  elem_pair(w, sj * nE + ej, si * nE + ei, sj, si, &v[sj].pos[pj * D], &v[si].pos[pi * D],
                            v[sj].sigma[pj], v[si].sigma[pi], dumf, &v[si].frc[pi * D], &dump, &v[si].pressure[ci], NULL,
                            NULL, kj, ki, ts);
              }
          }
          (void)ib;
        }
  }
}

/* calculInterForces, CIL.cu:108-264 */
static void inter_forces(or_ws* w, cil_species_view* v, int nspec, int ts) {
  uidmap um;
  int s, c;
  make_uidmap(&um, v, nspec);
  for (s = 0; s < nspec; s++)
    for (c = 0; c < v[s].ncells; c++) v[s].pressure[c] = 0;
  if (w->nthreads > 1 && w->traversal == 0 && w->rng_mode == OR_RNG_COUNTER) {
    inter_forces_gather(w, v, nspec, &um, ts, NULL);
    return;
  }
  traverse(w, v, nspec, &um, force_cb, &ts);
}

int or_forces(or_ws* w, cil_species_view* v, int nspec, int ts) {
  int rc;
  w->err[0] = 0;
  if ((rc = ensure_wiener(w, v, nspec))) return rc;
  if ((rc = draw_noise(w, v, nspec, ts))) return rc;
  intra_forces(w, v, nspec, ts);
  inter_forces(w, v, nspec, ts);
  if (w->err[0]) return CIL_E_BOND;
  return 0;
}

/* ===================================================================================== */
/* Update (CIL.cu:902-939) and centres of mass (Colony.cpp:409-432)                       */
/* ===================================================================================== */
/* Relaxation extension: scale a displacement down to length `cap` (cap <= 0: no-op). */
static void cap_disp(double cap, double* dx, int D) {
  double n2 = 0.0;
  int d;
  if (!(cap > 0.0)) return;
  for (d = 0; d < D; d++) n2 += dx[d] * dx[d];
  if (n2 > cap * cap) {
    double f = cap / sqrt(n2);
    for (d = 0; d < D; d++) dx[d] *= f;
  }
}

static void update_positions(or_ws* w, cil_species_view* v, int nspec, int ts) {
  int s, c, e, d, D = w->box.dim, nE = w->nE;
  double dt = w->p.dt;
  for (s = 0; s < nspec; s++) {
    const cil_species* P = &w->p.species[s];
    double p0 = P->cycle_p0;
#pragma omp parallel for private(e, d) num_threads(w->nthreads) if (w->nthreads > 1)
    for (c = 0; c < v[s].ncells; c++) {
      double pi = v[s].pressure[c];
      for (e = 0; e < nE; e++) {
        size_t pidx = (size_t)c * nE + e;
        double area = sigma2area(v[s].sigma[(size_t)c * nE + 0]); /* element 0's area (CIL.cu:921) */
        double pND = pi / p0;
        double pterm = P->zeta_p[e] == 0.0 ? 0.0 : P->zeta_p[e] * pow(pND, P->zeta_p_beta[e]);
        double temp = dt * 1 / ((P->zeta[e] + pterm) * area);
        double* R = &v[s].pos[pidx * D];
        double* F = &v[s].frc[pidx * D];
        double* V = &v[s].vel[pidx * D];
        double mu = temp / dt;
        if (w->p.noise == CIL_NOISE_FDT && P->kT > 0.0) {
          double z[4], amp = sqrt(2.0 * P->kT * temp), dx[3];
          or_philox_normal2(w->p.seed, (uint32_t)((cell_key(s, c) << 1) | e), (uint32_t)ts, ST_FDT, 0, z);
          or_philox_normal2(w->p.seed, (uint32_t)((cell_key(s, c) << 1) | e), (uint32_t)ts, ST_FDT, 1, z + 2);
          for (d = 0; d < D; d++) dx[d] = temp * F[d] + amp * z[d];
          cap_disp(w->p.max_disp, dx, D);
          box_update_pos_dr(&w->box, R, dx);
        } else if (w->p.max_disp > 0.0) {
          double dx[3];
          for (d = 0; d < D; d++) dx[d] = temp * F[d];
          cap_disp(w->p.max_disp, dx, D);
          box_update_pos_dr(&w->box, R, dx);
        } else {
          box_update_pos(&w->box, R, F, temp);
        }
        for (d = 0; d < D; d++) V[d] = F[d] * mu;
      }
    }
  }
}

int or_update(or_ws* w, cil_species_view* v, int nspec, int ts) {
  update_positions(w, v, nspec, ts);
  return 0;
}

int or_coms(or_ws* w, cil_species_view* v, int nspec, double* com) {
  int s, c, e, d, D = w->box.dim, nE = w->nE;
  size_t k = 0;
  for (s = 0; s < nspec; s++)
    for (c = 0; c < v[s].ncells; c++, k++) {
      const double* r0 = &v[s].pos[((size_t)c * nE) * D];
      double cm[3] = {0, 0, 0}, rij[3];
      for (e = 0; e < nE; e++) {
        box_distance(&w->box, rij, r0, &v[s].pos[((size_t)c * nE + e) * D]);
        for (d = 0; d < D; d++) cm[d] += rij[d];
      }
      for (d = 0; d < D; d++) cm[d] /= nE;
      box_update_pos_dr(&w->box, cm, r0);
      for (d = 0; d < D; d++) com[k * D + d] = cm[d];
    }
  return 0;
}

/* ===================================================================================== */
/* Divisions                                                                               */
/* ===================================================================================== */
static int add_cell(cil_species_view* vs) {
  if (vs->ncells >= vs->capacity) return -1;
  return vs->ncells++;
}

/* randomTimeSpan / randomTimeSpanLogNormal, MD.cuh:433-440; Random::logNormal Random.hpp:55-60 */
static int time_span(double dt, double mu, double sigma, double z) {
  int n = or_nint((mu + sigma * z) / dt);
  return n >= 1 ? n : 1;
}
static int time_span_lognormal(double dt, double mu, double sigma, double Z) {
  double mu2 = log((mu * mu) / (sqrt(mu * mu + sigma * sigma)));
  double sigma2 = sqrt(log(1 + (sigma * sigma) / (mu * mu)));
  int n = or_nint(exp(mu2 + sigma2 * Z) / dt);
  return n >= 1 ? n : 1;
}

/* calculDivisionsSingle, CIL.cu:469-619, for single-disc cells. */
static int divisions_theta(or_ws* w, cil_species_view* v, int nspec, int ts) {
  int s, i, d, D = w->box.dim;
  double dt = w->p.dt;
  for (s = 0; s < nspec; s++) {
    const cil_species* P = &w->p.species[s];
    cil_species_view* V = &v[s];
    const int nCells = V->ncells;
    double pShift = P->cycle_p_shift * exp(-ts * P->cycle_p_shift_tau_inv);
    for (i = 0; i < nCells; i++) {
      if ((V->state[i] != CS_DIVIDING) || (V->pressure[i] < P->apop_pressure)) V->tpressure[i] = 0;
      else V->tpressure[i] += 1;
      if (V->state[i] == CS_DIVIDING && !P->no_division) { /* func != NULL, CIL.cu:496 */
        double pn = V->pressure[i] + pShift;
        double pNonDim = (0. >= pn ? 0. : pn) / P->cycle_p0;
        double rr = P->cycle_rmaxed_inv * V->rho[i];
        double theta_prime = V->ang_vel[i] * (rr <= 1.0 ? rr : 1.0);
        double vt, vv, c0, vt_prime, v_prime, r_prime_np, tau_r_pre, r_prime;
        V->theta[i] += theta_prime;
        vt = D == 3 ? sigma2volume(V->sigma_t[i]) : sigma2area(V->sigma_t[i]);
        vv = D == 3 ? sigma2volume(V->sigma[i]) : sigma2area(V->sigma[i]);
        c0 = V->ang_vel[i] * log(2.0) + (P->cycle_v_1 - 1) / P->cycle_tau_v;
        vt_prime = c0 * V->rho[i] - (((vt - P->cycle_v_s) / P->cycle_v_s) / P->cycle_tau_v);
        vt += vt_prime * vt;
        v_prime = -(vv - (vt / (1 + P->cycle_beta_vs * (w->p.osmotic_pi + pNonDim)))) / P->cycle_tau_p;
        vv += v_prime;
        vv = P->cycle_tau_p == 0 ? vt : vv;
        V->sigma_t[i] = D == 3 ? volume2sigma(vt) : area2sigma(vt);
        V->sigma[i] = D == 3 ? volume2sigma(vv) : area2sigma(vv);
        r_prime_np = (1 - (V->rho[i]) - pNonDim);
        tau_r_pre = r_prime_np > 0 ? 1 : P->cycle_rho_ratio;
        r_prime = r_prime_np / (P->cycle_tau_r * tau_r_pre);
        V->rho[i] += r_prime * V->rho[i];
        if (V->theta[i] >= 1) {
          int inew = add_cell(V);
          double lMom = V->sigma[i], newSig, dr[3];
          uint32_t key = cell_key(s, i);
          double z[4];
          if (inew < 0) return fail(w, CIL_E_CAPACITY, "division beyond capacity");
          /* singleCellDivision, MD.cu:668-719 */
          if (D == 2) {
            double lD = lMom / sqrt(2.0);
            double r = (lMom - lD) / 2;
            double alpha = rng_uniform0x(w, OR_TWOPI, key, (uint32_t)ts, ST_DIV, 0, 0);
            dr[0] = r * sin(alpha);
            dr[1] = r * cos(alpha);
          } else {
            double rMom = lMom / 2;
            double volumeMom = 4 * OR_PI * rMom * rMom * rMom / 3;
            double a = 9 * lMom / 4, b = -(3 * (lMom * lMom) / 2);
            double c = lMom * lMom * lMom / 4 - 3 * volumeMom / (OR_PI);
            double lD = (-b + sqrt(b * b - 4 * a * c)) / (2 * a);
            double r = (lMom - lD) / 2;
            double alpha = rng_uniform0x(w, OR_TWOPI, key, (uint32_t)ts, ST_DIV, 0, 0);
            double beta = rng_uniform0x(w, OR_TWOPI, key, (uint32_t)ts, ST_DIV, 0, 1);
            dr[0] = r * sin(alpha) * cos(beta);
            dr[1] = r * sin(alpha) * sin(beta);
            dr[2] = r * cos(alpha);
          }
          for (d = 0; d < D; d++) V->pos[(size_t)inew * D + d] = V->pos[(size_t)i * D + d];
          box_update_pos_dr(&w->box, &V->pos[(size_t)inew * D], dr);
          for (d = 0; d < D; d++) dr[d] *= -1;
          box_update_pos_dr(&w->box, &V->pos[(size_t)i * D], dr);
          log_division(w, s, V, inew, ts);
          /* sigma halving, CIL.cu:546-570 */
          if (D == 2) newSig = lMom / sqrt(2.0);
          else {
            double rMom = (V->sigma[i]) / 2;
            double volumeMom = 4 * OR_PI * rMom * rMom * rMom / 3;
            double a = 9 * lMom / 4, b = -(3 * (lMom * lMom) / 2);
            double c = lMom * lMom * lMom / 4 - 3 * volumeMom / (OR_PI);
            newSig = (-b + sqrt(b * b - 4 * a * c)) / (2 * a);
          }
          V->sigma[inew] = newSig;
          V->sigma[i] = newSig;
          V->sigma_t[inew] = newSig;
          V->sigma_t[i] = newSig;
          V->t0[inew] = ts;
          V->t0[i] = ts;
          if (w->rng_mode == OR_RNG_DSFMT) {
            z[0] = or_dsfmt_normal(&w->g);
            z[1] = or_dsfmt_normal(&w->g);
          } else
            or_philox_normal2(w->p.seed, key, (uint32_t)ts, ST_DIV, 1, z);
          V->tswitch[inew] = ts + time_span(dt, P->migration_tau.mu, P->migration_tau.sigma, z[0]);
          V->tswitch[i] = ts + time_span(dt, P->migration_tau.mu, P->migration_tau.sigma, z[1]);
          V->state[i] = CS_DIVIDING;
          V->state[inew] = CS_DIVIDING;
          V->theta[i] = 0;
          V->theta[inew] = 0;
          V->pressure[inew] = V->pressure[i];
          V->rho[inew] = V->rho[i];
          if (w->rng_mode == OR_RNG_DSFMT) {
            z[2] = or_dsfmt_normal(&w->g);
            z[3] = or_dsfmt_normal(&w->g);
          } else
            or_philox_normal2(w->p.seed, key, (uint32_t)ts, ST_DIV, 2, z + 2);
          V->ang_vel[i] = 1.0 / time_span_lognormal(dt, P->division_tau.mu, P->division_tau.sigma, z[2]);
          V->ang_vel[inew] = 1.0 / time_span_lognormal(dt, P->division_tau.mu, P->division_tau.sigma, z[3]);
          V->tpressure[inew] = 0;
          for (d = 0; d < D; d++) {
            V->vel[(size_t)inew * D + d] = 0;
            V->vel[(size_t)i * D + d] = 0;
            V->frc[(size_t)inew * D + d] = 0;
            V->frc[(size_t)i * D + d] = 0;
          }
          if (V->rest_len) V->rest_len[inew] = V->rest_len[i];
        }
      }
    }
  }
  return 0;
}

/* Linear-growth extension for two-disc cells (no reference text; parity unpinned against
 * the reference).  Rest length grows by growth_rate*dt per step; at l_div the cell splits
 * into a back daughter (keeps the slot) and a front daughter (appended), both with rest
 * length l_birth; daughter diameters get sigma*(1 + jitter*N(0,1)). */
static int divisions_linear(or_ws* w, cil_species_view* v, int nspec, int ts) {
  int s, i, d, e, D = w->box.dim;
  for (s = 0; s < nspec; s++) {
    const cil_species* P = &w->p.species[s];
    cil_species_view* V = &v[s];
    const int nCells = V->ncells;
    if (!(P->growth_rate > 0.0) || P->no_division) continue;
    for (i = 0; i < nCells; i++) {
      V->rest_len[i] += P->growth_rate * w->p.dt;
      if (V->rest_len[i] >= P->l_div) {
        int inew = add_cell(V);
        double *b = &V->pos[((size_t)i * 2) * D], *f = &V->pos[((size_t)i * 2 + 1) * D];
        double u[3] = {0, 0, 0}, uh[3] = {1, 0, 0}, r2, r, dx[3];
        double z[4];
        uint32_t key = cell_key(s, i);
        if (inew < 0) return fail(w, CIL_E_CAPACITY, "division beyond capacity");
        r2 = box_distance(&w->box, u, b, f);
        r = sqrt(r2);
        if (r > 0.0)
          for (d = 0; d < D; d++) uh[d] = u[d] / r;
        /* daughter B (appended): back = f - l_birth*uh, front = f */
        for (d = 0; d < D; d++) {
          V->pos[((size_t)inew * 2 + 1) * D + d] = f[d];
          V->pos[((size_t)inew * 2) * D + d] = f[d];
          dx[d] = -P->l_birth * uh[d];
        }
        box_update_pos_dr(&w->box, &V->pos[((size_t)inew * 2) * D], dx);
        log_division(w, s, V, inew, ts);
        /* daughter A (parent slot): back = b, front = b + l_birth*uh */
        for (d = 0; d < D; d++) {
          f[d] = b[d];
          dx[d] = P->l_birth * uh[d];
        }
        box_update_pos_dr(&w->box, f, dx);
        V->rest_len[i] = P->l_birth;
        V->rest_len[inew] = P->l_birth;
        or_philox_normal2(w->p.seed, key, (uint32_t)ts, ST_DIV, 0, z);
        or_philox_normal2(w->p.seed, key, (uint32_t)ts, ST_DIV, 1, z + 2);
        for (e = 0; e < 2; e++) {
          V->sigma[(size_t)i * 2 + e] = P->sigma[e] * (1.0 + P->sigma_jitter * z[e]);
          V->sigma[(size_t)inew * 2 + e] = P->sigma[e] * (1.0 + P->sigma_jitter * z[2 + e]);
        }
        V->state[inew] = V->state[i];
        V->pressure[inew] = V->pressure[i];
        V->t0[inew] = ts;
        V->t0[i] = ts;
        for (e = 0; e < 2; e++)
          for (d = 0; d < D; d++) {
            V->vel[((size_t)inew * 2 + e) * D + d] = 0;
            V->vel[((size_t)i * 2 + e) * D + d] = 0;
            V->frc[((size_t)inew * 2 + e) * D + d] = 0;
            V->frc[((size_t)i * 2 + e) * D + d] = 0;
          }
      }
    }
  }
  return 0;
}

/* binaryRandomReorientation (MD.cu:653-659): slave <- master, then master += radius*dr and
 * slave -= radius*dr for a random unit direction dr (both wrapped into the box). */
static void binary_reorient(or_ws* w, double* slave, double* master, double radius, uint32_t key, int ts,
                            uint32_t k0) {
  int d, D = w->box.dim;
  double dr[3] = {0, 0, 0}, m[3];
  for (d = 0; d < D; d++) slave[d] = master[d];
  rng_on_sphere_k(w, dr, key, (uint32_t)ts, ST_DIV, k0);
  for (d = 0; d < D; d++) m[d] = -dr[d];
  box_update_pos(&w->box, master, dr, radius);
  box_update_pos(&w->box, slave, m, radius);
}

/* resetVel / resetFrc of both discs of cell i (Utils.hpp:591-599) */
static void reset_cell_motion(or_ws* w, cil_species_view* V, int i) {
  int k, D = w->box.dim;
  for (k = 0; k < 2 * D; k++) {
    V->vel[(size_t)i * 2 * D + k] = 0;
    V->frc[(size_t)i * 2 * D + k] = 0;
  }
}

/* randomTimeSpan (MD.cuh:433-435): max(Nint(N(mu, sigma) / dt), 1) with the normal drawn from
 * the dSFMT stream (Leva) or from Philox block k of (key, ts, ST_DIV). */
static int random_span(or_ws* w, const cil_gaussian* g, uint32_t key, int ts, uint32_t k, int half) {
  double z = rng_normal(w, key, (uint32_t)ts, ST_DIV, k, half);
  return time_span(w->p.dt, g->mu, g->sigma, z);
}

/* calculDivisions, CIL.cu:369-455, for two-disc cells.  Per dividing species, cells i in
 * ascending order with the count frozen at entry:
 *   Dividing, and binaryCellDivisionCriteria (MD.cuh:426-432: |r_0 - r_1|^2 >= divisionR2min,
 *   or >= divisionR2minRatio * sigma_0 when that ratio is set): the cell divides.  The daughter
 *   is appended; binaryCellDivision (MD.cu:721-731) gives the daughter the mother's back disc
 *   position and re-places both cells' discs +-divisionR0 around the mother's front and the
 *   old back position along fresh random directions (mother first).  Both get t0 = ts and
 *   state Crawling, vel = frc = 0 and the species' diameters; the daughter's switch time is
 *   ts + span(tau_migration), the mother's grows by another such span.
 *   Dividing, check failed, and tswitch == ts: binaryCellDivisionFail (MD.cu:736-752) moves
 *   the back disc to the midpoint (unwrapped, as the reference does) and re-places both discs
 *   +-divisionR0 around it; vel = frc = 0, tswitch += span(tau_migration), Crawling, species
 *   diameters.
 *   Crawling and tswitch == ts: state Dividing, tswitch += span(tau_division), vel = frc = 0.
 * The reference draws from one sequential stream in that order (onDSphere mother, onDSphere
 * daughter, normal daughter, normal mother; failure: onDSphere, normal; start: normal); with
 * the counter source the draws of cell i at step ts are Philox blocks of (cell key, ts, ST_DIV):
 * directions k = 0,1 and 2,3, the success and failure spans block 4 (daughter half 0, mother
 * half 1), the start span block 5.  The fields the reference leaves stale in the daughter's
 * slot (addCell, Colony.cpp:355-359: pressure, theta, rho, ang_vel, sigma_t, rest_len) take
 * the mother's values and tpressure 0, so that runs are reproducible. */
static int divisions_binary(or_ws* w, cil_species_view* v, int nspec, int ts) {
  int s, i, d, e, D = w->box.dim;
  for (s = 0; s < nspec; s++) {
    const cil_species* P = &w->p.species[s];
    cil_species_view* V = &v[s];
    const int nCells = V->ncells;
    if (P->no_division) continue;
    for (i = 0; i < nCells; i++) {
      const uint32_t key = cell_key(s, i);
      double* b = &V->pos[((size_t)i * 2) * D];
      double* f = &V->pos[((size_t)i * 2 + 1) * D];
      if (V->state[i] == CS_DIVIDING) {
        double rij[3];
        double r2 = box_distance(&w->box, rij, f, b);
        int ok = P->division_r2min_ratio != 0.0 ? r2 >= P->division_r2min_ratio * V->sigma[(size_t)i * 2]
                                                : r2 >= P->division_r2min;
        if (ok) {
          int inew = add_cell(V);
          double *nb, *nf;
          if (inew < 0) return fail(w, CIL_E_CAPACITY, "division beyond capacity");
          nb = &V->pos[((size_t)inew * 2) * D];
          nf = &V->pos[((size_t)inew * 2 + 1) * D];
          for (d = 0; d < D; d++) nb[d] = b[d];
          binary_reorient(w, b, f, P->division_r0, key, ts, 0);   /* mother keeps the front */
          binary_reorient(w, nf, nb, P->division_r0, key, ts, 2); /* daughter the old back */
          log_division(w, s, V, inew, ts);
          V->t0[inew] = ts;
          V->t0[i] = ts;
          V->tswitch[inew] = ts + random_span(w, &P->migration_tau, key, ts, 4, 0);
          V->state[inew] = CS_CRAWLING;
          for (e = 0; e < 2; e++) V->sigma[(size_t)inew * 2 + e] = P->sigma[e];
          reset_cell_motion(w, V, inew);
          reset_cell_motion(w, V, i);
          V->pressure[inew] = V->pressure[i];
          V->tpressure[inew] = 0;
          if (V->theta) V->theta[inew] = V->theta[i];
          if (V->rho) V->rho[inew] = V->rho[i];
          if (V->ang_vel) V->ang_vel[inew] = V->ang_vel[i];
          if (V->sigma_t) V->sigma_t[inew] = V->sigma_t[i];
          if (V->rest_len) V->rest_len[inew] = V->rest_len[i];
          V->tswitch[i] += random_span(w, &P->migration_tau, key, ts, 4, 1);
          V->state[i] = CS_CRAWLING;
          for (e = 0; e < 2; e++) V->sigma[(size_t)i * 2 + e] = P->sigma[e];
        } else if (V->tswitch[i] == ts) {
          double dr[3];
          box_distance(&w->box, dr, b, f); /* dr = r_1 - r_0 */
          for (d = 0; d < D; d++) b[d] += dr[d] * 0.5;
          binary_reorient(w, f, b, P->division_r0, key, ts, 0);
          reset_cell_motion(w, V, i);
          V->tswitch[i] += random_span(w, &P->migration_tau, key, ts, 4, 1);
          V->state[i] = CS_CRAWLING;
          for (e = 0; e < 2; e++) V->sigma[(size_t)i * 2 + e] = P->sigma[e];
        }
      } else if (V->tswitch[i] == ts) {
        V->state[i] = CS_DIVIDING;
        V->tswitch[i] += random_span(w, &P->division_tau, key, ts, 5, 0);
        reset_cell_motion(w, V, i);
      }
    }
  }
  return 0;
}

int or_divisions(or_ws* w, cil_species_view* v, int nspec, int ts) {
  if (w->p.division == CIL_DIV_THETA_CYCLE) return divisions_theta(w, v, nspec, ts);
  if (w->p.division == CIL_DIV_LINEAR_GROWTH) return divisions_linear(w, v, nspec, ts);
  if (w->p.division == CIL_DIV_BINARY) return divisions_binary(w, v, nspec, ts);
  return 0;
}

/* updateDiameters, CIL.cu:970-990: in a dividing species with DIVISION.SWELLING, every disc of
 * a Dividing cell smaller than the species' largest default diameter grows by
 * (max - sigma) / (tswitch - ts), so that it reaches the maximum at the switch time. */
int or_update_diameters(or_ws* w, cil_species_view* v, int nspec, int ts) {
  int s, i, e, nE = w->nE;
  if (w->p.division == CIL_DIV_NONE) return 0;
  for (s = 0; s < nspec; s++) {
    const cil_species* P = &w->p.species[s];
    cil_species_view* V = &v[s];
    double smax = P->sigma[0];
    if (P->no_division || !P->division_swelling) continue;
    for (e = 1; e < nE; e++) smax = P->sigma[e] > smax ? P->sigma[e] : smax;
    for (i = 0; i < V->ncells; i++)
      if (V->state[i] == CS_DIVIDING)
        for (e = 0; e < nE; e++) {
          double* sg = &V->sigma[(size_t)i * nE + e];
          if (*sg < smax) *sg += (smax - *sg) / (V->tswitch[i] - ts);
        }
  }
  return 0;
}

/* Colony removeCell, Colony.cpp:365-399: copy the last cell into the slot. */
static void remove_cell(or_ws* w, cil_species_view* V, int i) {
  int last = V->ncells - 1, nE = w->nE, D = w->box.dim, k;
  for (k = 0; k < nE * D; k++) {
    V->pos[(size_t)i * nE * D + k] = V->pos[(size_t)last * nE * D + k];
    V->frc[(size_t)i * nE * D + k] = V->frc[(size_t)last * nE * D + k];
    V->vel[(size_t)i * nE * D + k] = V->vel[(size_t)last * nE * D + k];
  }
  for (k = 0; k < nE; k++) V->sigma[(size_t)i * nE + k] = V->sigma[(size_t)last * nE + k];
#define CPY(f)      \
  if (V->f) V->f[i] = V->f[last]
  CPY(sigma_t);
  CPY(pressure);
  CPY(state);
  CPY(t0);
  CPY(tswitch);
  CPY(tpressure);
  CPY(theta);
  CPY(rho);
  CPY(ang_vel);
  CPY(rest_len);
#undef CPY
  V->ncells--;
}

/* checkPositions, CIL.cu:945-966 (including the skip-after-swap quirk at :957-963). */
int or_check_positions(or_ws* w, cil_species_view* v, int nspec, int* removed) {
  int s, nE = w->nE, D = w->box.dim;
  *removed = 0;
  for (s = 0; s < nspec; s++) {
    cil_species_view* V = &v[s];
    int i = 0, nCells = V->ncells;
    while (i < nCells) {
      int j;
      for (j = 0; j < nE; j++) {
        const double* R = &V->pos[((size_t)i * nE + j) * D];
        if (!w->p.drain) {
          if (!box_bounded(&w->box, R)) return fail(w, CIL_E_OUTSIDE, "Particle outside domain!");
        } else if (!box_bounded(&w->box, R)) {
          remove_cell(w, V, i);
          (*removed)++;
          nCells = V->ncells;
          break;
        }
      }
      i += 1;
    }
  }
  return 0;
}

/* ===================================================================================== */
/* Euler, CIL.cu:1013-1020                                                                */
/* ===================================================================================== */
/* updatePositions (CIL.cu:849-866): x += deltaT * zetaInv * F (zetaInv = 1/zeta, Input.cpp:93),
 * wrapped; velocities untouched. */
static void heun_half_update(or_ws* w, cil_species_view* v, int nspec) {
  int s, c, e, D = w->box.dim, nE = w->nE;
  const double hdt = 0.5 * w->p.dt; /* paramsMD.hdt, CIL.cuh:208 */
  for (s = 0; s < nspec; s++) {
    const cil_species* P = &w->p.species[s];
    for (c = 0; c < v[s].ncells; c++)
      for (e = 0; e < nE; e++) {
        size_t pidx = (size_t)c * nE + e;
        double temp = hdt * (1.0 / P->zeta[e]);
        box_update_pos(&w->box, &v[s].pos[pidx * D], &v[s].frc[pidx * D], temp);
      }
  }
}

int or_step(or_ws* w, cil_species_view* v, int nspec, int nsteps, int* ts) {
  int k, rc;
  for (k = 0; k < nsteps; k++) {
    if (w->p.integrator == CIL_INT_HEUN) {
      /* StochasticHeun, CIL.cu:1028-1049.  The predictor's full step into pos0 is never read
       * (pos0 is overwritten next step); the half step with the old forces, the forces at the
       * half-step positions and a second half step with them give the new positions. */
      heun_half_update(w, v, nspec);
      w->phase = 1;
      rc = or_forces(w, v, nspec, *ts);
      w->phase = 0;
      if (rc) return rc;
      heun_half_update(w, v, nspec);
    } else {
      update_positions(w, v, nspec, *ts);
    }
    /* calculComs runs here in the reference; the link list recomputes them when it is built. */
    ++(*ts);
    if ((rc = or_divisions(w, v, nspec, *ts))) return rc;
    if ((rc = or_update_diameters(w, v, nspec, *ts))) return rc;
    if ((rc = or_forces(w, v, nspec, *ts))) return rc;
  }
  return 0;
}

/* ===================================================================================== */
/* Pair enumeration                                                                       */
/* ===================================================================================== */
typedef struct {
  int32_t* ij;
  long cap, n;
} pair_acc;

static void pair_cb_fn(or_ws* w, cil_species_view* v, const uidmap* um, int ii, int jj, void* ctx) {
  pair_acc* acc = (pair_acc*)ctx;
  int si, ci, sj, cj, ei, ej, nE = w->nE, D = w->box.dim;
  uid2cell(um, ii, &si, &ci);
  uid2cell(um, jj, &sj, &cj);
  if (si == sj && ci == cj) return;
  for (ei = 0; ei < nE; ei++)
    for (ej = 0; ej < nE; ej++) {
      size_t pi = (size_t)ci * nE + ei, pj = (size_t)cj * nE + ej;
      double rij[3], f0;
      double r2 = box_distance(&w->box, rij, &v[si].pos[pi * D], &v[sj].pos[pj * D]);
      int ti = si * nE + ei, tj = sj * nE + ej;
      if (w->p.law[ti][tj] == CIL_LAW_NONE) continue;
      if (or_pair_f0(&w->p, D, ti, tj, v[si].sigma[pi], v[sj].sigma[pj], r2, &f0)) {
        int32_t a = (int32_t)((size_t)um->cumul[si] * nE + pi), b = (int32_t)((size_t)um->cumul[sj] * nE + pj);
        if (a > b) {
          int32_t t = a;
          a = b;
          b = t;
        }
        if (acc->n < acc->cap) {
          acc->ij[2 * acc->n] = a;
          acc->ij[2 * acc->n + 1] = b;
        }
        acc->n++;
      }
    }
}

static int cmp_pair(const void* x, const void* y) {
  const int32_t *a = (const int32_t*)x, *b = (const int32_t*)y;
  if (a[0] != b[0]) return a[0] < b[0] ? -1 : 1;
  if (a[1] != b[1]) return a[1] < b[1] ? -1 : 1;
  return 0;
}

long or_pairs(or_ws* w, cil_species_view* v, int nspec, int mode, int32_t* ij, long cap) {
  uidmap um;
  pair_acc acc = {ij, cap, 0};
  int saved = w->traversal;
  make_uidmap(&um, v, nspec);
  w->traversal = mode == 0 ? 0 : 1;
  traverse(w, v, nspec, &um, pair_cb_fn, &acc);
  w->traversal = saved;
  if (acc.n <= cap) qsort(ij, (size_t)acc.n, 2 * sizeof(int32_t), cmp_pair);
  return acc.n;
}

void or_linklist_info(or_ws* w, int* ns, double* width) {
  int d;
  for (d = 0; d < 3; d++) ns[d] = w->ns[d];
  *width = w->width;
}

/* Per-particle sum of |pair force contributions| (all-pairs), for relative tolerances. */
int or_force_scale_gather(or_ws* w, cil_species_view* v, int nspec, double* scale) {
  uidmap um;
  w->err[0] = 0;
  if (w->traversal != 0) return fail(w, CIL_E_INVALID, "force_scale_gather needs the link-list traversal");
  make_uidmap(&um, v, nspec);
  memset(scale, 0, sizeof(double) * (size_t)um.cumul[nspec] * w->nE);
  inter_forces_gather(w, v, nspec, &um, 0, scale);
  return 0;
}

int or_force_scale(or_ws* w, cil_species_view* v, int nspec, double* scale) {
  uidmap um;
  int D = w->box.dim, nE = w->nE, ii, jj, ntot;
  make_uidmap(&um, v, nspec);
  ntot = um.cumul[nspec];
  memset(scale, 0, sizeof(double) * (size_t)ntot * nE);
  for (ii = 0; ii < ntot; ii++)
    for (jj = ii + 1; jj < ntot; jj++) {
      int si, ci, sj, cj, ei, ej;
      uid2cell(&um, ii, &si, &ci);
      uid2cell(&um, jj, &sj, &cj);
      for (ei = 0; ei < nE; ei++)
        for (ej = 0; ej < nE; ej++) {
          size_t pi = (size_t)ci * nE + ei, pj = (size_t)cj * nE + ej;
          double rij[3], f0;
          double r2 = box_distance(&w->box, rij, &v[si].pos[pi * D], &v[sj].pos[pj * D]);
          int ti = si * nE + ei, tj = sj * nE + ej;
          if (w->p.law[ti][tj] == CIL_LAW_NONE) continue;
          if (or_pair_f0(&w->p, D, ti, tj, v[si].sigma[pi], v[sj].sigma[pj], r2, &f0)) {
            double m = fabs(f0) * sqrt(r2) + 1e-300;
            scale[(size_t)ii * nE + ei] += m;
            scale[(size_t)jj * nE + ej] += m;
          }
        }
    }
  return 0;
}
