/*
 * cil_oracle.h — CPU oracle for the overdamped cell-colony Euler step.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C, FP64, strict-FP restatement of the
 * reference's algorithm (KokoroSinx/cil_cuda: CIL.cu, MD.cu, MD.cuh, Utils.hpp, Colony.cpp,
 * Container.cpp, Random.hpp, mt/dSFMT.c).  Only tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg may load it, and only as the checker or the timed CPU
 * baseline; the HIP product path never calls it.
 *
 * Every function cites the reference file:line it restates.  Parity pins per function are
 * listed in the test-suite (tests/test_oracle_*.py); functions without an independent pin
 * are marked "parity unpinned" below.
 */
#ifndef CIL_ORACLE_H
#define CIL_ORACLE_H

#include <stddef.h>
#include <stdint.h>
#include "../include/cil_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

/* ---- RNG sources ------------------------------------------------------------------ */
enum { OR_RNG_COUNTER = 0, OR_RNG_EXTERNAL = 1, OR_RNG_DSFMT = 2 };

/* dSFMT-19937 restatement (mt/dSFMT.c, mt/dSFMT.h, mt/dSFMT-params19937.h). */
typedef struct {
  uint64_t s[192][2]; /* DSFMT_N + 1 128-bit words; s[191] is the "lung" */
  int idx;
} or_dsfmt;
void or_dsfmt_init(or_dsfmt* g, uint32_t seed);
double or_dsfmt_close1_open2(or_dsfmt* g);
double or_dsfmt_close_open(or_dsfmt* g);
double or_dsfmt_open_open(or_dsfmt* g);
double or_dsfmt_normal(or_dsfmt* g); /* Random::normal (Leva 1992), Random.hpp:24-52 */
/* fill `n` values of the Random:: call named by `kind` (0 close_open, 1 normal, 2 open_open) */
void or_dsfmt_fill(uint32_t seed, int kind, double* out, size_t n);

/* Philox4x32-10 (Salmon et al., SC'11), shared bit-for-bit with the HIP kernels. */
void or_philox4x32(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
/* Two N(0,1) deviates for counter (id, ts, stream, k); Box-Muller on 53-bit uniforms. */
void or_philox_normal2(uint64_t seed, uint32_t id, uint32_t ts, uint32_t stream, uint32_t k, double z[2]);
void or_philox_uniform2(uint64_t seed, uint32_t id, uint32_t ts, uint32_t stream, uint32_t k, double u[2]);

/* ---- colony workspace ----------------------------------------------------------------- */
typedef struct or_ws or_ws;

or_ws* or_create(const cil_params* p, const cil_box* box);
void or_destroy(or_ws* w);
void or_set_rng(or_ws* w, int mode, uint32_t dsfmt_seed);
/* One draw from the workspace's sequential dSFMT stream (OR_RNG_DSFMT mode): kind 0 =
 * Random::uniform0x(1), kind 1 = Random::normal() (Random.hpp:13-52).  For host-side setup
 * draws the reference makes on the same stream, such as initializeCellStatus. */
double or_rng_draw(or_ws* w, int kind);
int or_set_external_noise(or_ws* w, const double* normals, size_t per_step, int nsteps);
/* Pair-traversal back end for inter forces: 0 = COM link list (reference structure,
 * CIL.cu:177-231), 1 = all pairs (CIL.cu:232-263). */
void or_set_traversal(or_ws* w, int mode);
/* threaded gather (threads > 1, counter RNG): pair forces of the cells whose UID is a multiple of
 * `every` only, the other cells keep their intra forces (a sample of a large colony) */
void or_set_sample(or_ws* w, int every);
/* division records: on = 1 appends, per division, the new cell's element positions right after
 * the division and the step; or_division_log copies species s's records (pos [n][nelem][dim])
 * and clears them when clear != 0 */
void or_set_division_log(or_ws* w, int on);
int or_division_log(or_ws* w, int s, double* pos, int* ts, int cap, int* n, int clear);
int or_set_threads(or_ws* w, int nthreads); /* OpenMP gather variant when > 1 */

/* calculForcesPressure, CIL.cu:999-1007 (draws noise when enabled). */
int or_forces(or_ws* w, cil_species_view* v, int nspec, int ts);
/* Euler(ts) x nsteps, CIL.cu:1013-1020. */
int or_step(or_ws* w, cil_species_view* v, int nspec, int nsteps, int* ts);
/* Pieces, exposed for unit tests. */
int or_update(or_ws* w, cil_species_view* v, int nspec, int ts);         /* CIL.cu:902-939 */
int or_coms(or_ws* w, cil_species_view* v, int nspec, double* com);       /* Colony.cpp:409-432 */
int or_divisions(or_ws* w, cil_species_view* v, int nspec, int ts);        /* CIL.cu:369-455, 469-619 */
int or_update_diameters(or_ws* w, cil_species_view* v, int nspec, int ts); /* CIL.cu:970-990 */
int or_check_positions(or_ws* w, cil_species_view* v, int nspec, int* removed); /* CIL.cu:945-966 */

/* Interacting element pairs (UID_i < UID_j), sorted ascending.  mode 0: via the COM link
 * list, mode 1: brute force O(N^2).  Returns the count; writes at most cap pairs. */
long or_pairs(or_ws* w, cil_species_view* v, int nspec, int mode, int32_t* ij, long cap);

/* Per-particle force scale sum_j |f_ij| (norm of each pair term), for relative tolerances. */
/* or_force_scale over the link list with the threaded gather (the sample of or_set_sample, the
 * other entries stay 0); traversal 0 only */
int or_force_scale_gather(or_ws* w, cil_species_view* v, int nspec, double* scale);
int or_force_scale(or_ws* w, cil_species_view* v, int nspec, double* scale);

/* Scalar pair law f0 for element types ti,tj with diameters si,sj at squared distance r2
 * (and the contact-law h path); returns 1 when the law's predicate holds. */
int or_pair_f0(const cil_params* p, int dim, int ti, int tj, double si, double sj, double r2, double* f0);

/* Link-list geometry the oracle used last (bins per dim). */
void or_linklist_info(or_ws* w, int* ns, double* width);

const char* or_last_error(or_ws* w);
size_t or_sizeof(int which);

#ifdef __cplusplus
}
#endif
#endif
