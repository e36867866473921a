// kernels.hpp — gfx950 kernels of the colony step.
//
// Data layout in HBM (all canonical arrays in cell-slot order, slot = species base + cell id,
// particle p = slot*nE + elem, matching the reference's pos(cell, elem, d) ordering):
//   pos/vel/frc : FP64 SoA [D][P]      sigma : FP64 [P]
//   per-cell    : FP64 / int32 [C]
// Per-step scratch (rebuilt every step):
//   key/rank    : int32 [P]  bin index and arrival rank (atomic counting sort)
//   count/start : int32 [nbins(+1)]
//   idx         : int32 [P]  sorted slot -> particle, ascending particle index inside a bin
//                            (sorting network per lane, wave bitonic for large bins), so
//                            every run is bit-reproducible
//   spos        : float4 (bin-local x,y,z, sigma) in FP32 mode, double4 (global) in FP64
//   smeta       : int32 [P]  particle*8 + element type
//   fpair/ppair : [D][P] / [P] pair forces and pressure partials, in sorted order
#pragma once
#include "device.hpp"
#include "dsfmt.hpp"

namespace cil {

struct Dev {
  int P, C, D;
  double* pos;
  double* vel;
  double* frc;
  double* sigma;
  double* sigma_t;
  double* pressure;
  int* state;
  int* t0;
  int* tswitch;
  int* tpressure;
  double* theta;
  double* rho;
  double* ang_vel;
  double* rest_len;
  double* wiener;   // [C][3]
  double* awiener;  // [C][2]
  int* ncells;      // [CIL_MAX_SPECIES]
  int* ts;          // [1] step index of the last force evaluation
  int ts_off;       // added to *ts by every reader: step offset inside a chunk of launched steps
  unsigned* err;    // [1]
  int* key;
  int* rank;
  int* count;
  int* start;       // [nbins+1]
  int* blocksum;
  int* idx;
  int* inv;
  void* spos;
  int* smeta;
  int* sbin;        // bin of each sorted particle, packed coordinates (pack_bin)
  int* gslot;       // [P][4] FP32 LJ pairs inside the cut-off guard band (sorted source index),
                    // resolved in FP64 by the pair kernel itself after its pair loop
  void* fpair;
  void* ppair;
  const double* ext;
  long ext_per_step;
  int ext_nsteps;
  int ext_ts0;
  // division / removal scratch
  int* dflag;
  int* dscan;
  int* dblock;
  int* nnew;        // [CIL_MAX_SPECIES]
  // neighbour-list mode
  unsigned* rebuild;  // [0..1] by the parity of a force pass's step index: nonzero = that pass
                      // re-sorts and rebuilds the list (set by the skin test, division and
                      // host-initiated sorts; the update of step u clears word u & 1, whose
                      // pass has run); [2] list epoch, counts the builds
  int* nl;            // [list_max / 4][P][4] groups of (sorted index << 10 | bin offset code);
                      // each source type starts on a group boundary
  int* nlt;           // [ntypes][P] entries per source type
  float* xref;        // [4][P] positions (x, y, z) and diameter at the last build (particle order)
  float* spos0;       // [P][4] or null: the sorted bin-local positions and diameter at the last
                      // build (sorted order), kept for the surrogate's kNN candidate lists
  // two-level lists (null: one level): the inner list holds the (outer) list's entries within
  // cut-off + inner skin at its last refresh, in the same order; a pass refreshes it, while
  // evaluating the outer list, when the outer list was rebuilt or a particle has moved more
  // than inner skin / 2 since the refresh (flags rebuild[4..5], by parity like [0..1])
  int* nl_in;         // [list_max / 4][P][4], the layout of nl
  int* nlt_in;        // [ntypes][P]
  float* xref_in;     // [3][P] positions at the last refresh (particle order)
  // integrator: 0 Euler (one force pass per step, at step index f); 1 stochastic Heun (two
  // passes per step: the corrector's at the step's own index with phase 1, the final one at the
  // next index with phase 0).  `phase` is set per launch.
  int heun;
  int phase;
  // CIL_RNG_DSFMT: the reference's sequential stream (one state, drawn by one device thread)
  // and each cell's draws of the current step's division event, [C][8]
  Dsfmt* dsfmt;
  double* ddraw;
  // division records (the reference's posCellsDiv / cellT0Div, Colony.cpp:319-326), on when
  // requested: per species region of cap[s] entries from base[s], the new cell's element
  // positions [C][CIL_MAX_ELEMS][3] and the step, dlog_cnt[s] entries since the last read
  double* dlog_pos = nullptr;
  int* dlog_ts = nullptr;
  int* dlog_cnt = nullptr;
};

// Records a division of species s (its rank-th this step) whose new cell is slot inew, in
// the order of the reference's loop (parent order); entries beyond the region are dropped
// (the host reads and clears the records once per frame).
__device__ inline void log_division(const StepParams& P, const Dev& S, int s, int rank, int inew, int ts) {
  if (!S.dlog_pos) return;
  const int k = S.dlog_cnt[s] + rank;
  if (k >= P.cap[s]) return;
  const size_t e = (size_t)P.base[s] + k;
  for (int q = 0; q < P.nE; ++q)
    for (int d = 0; d < 3; ++d)
      S.dlog_pos[(e * CIL_MAX_ELEMS + q) * 3 + d] =
          d < P.box.dim ? S.pos[(size_t)d * S.P + (size_t)inew * P.nE + q] : 0.0;
  S.dlog_ts[e] = ts;
}

// ---------------------------------------------------------------------------------------
// Small helpers
// ---------------------------------------------------------------------------------------
// Step index seen by a kernel: the device counter plus the launch's offset within its chunk
// (a chunk of steps advances the counter once, at its end).
__device__ inline int step_ts(const Dev& S) { return *S.ts + S.ts_off; }
// Rebuild flag of the force pass at step index f (list mode).  Consecutive force passes use
// alternate words: by the step index's parity under Euler, by the phase under Heun (corrector
// passes word 1, final passes word 0).  A launch's own pass, and the update's previous pass,
// is the one of its (step index, phase).
__device__ inline unsigned* rebuild_flag(const Dev& S, int f) { return S.rebuild + (S.heun ? S.phase : (f & 1)); }
// Inner-list refresh flag of the force pass at step index f (two-level lists, Euler only).
__device__ inline unsigned* refresh_flag(const Dev& S, int f) { return S.rebuild + 4 + (f & 1); }
// Flag of the force pass that follows an update launched at step index f.
__device__ inline unsigned* next_pass_flag(const Dev& S, int f) {
  return S.rebuild + (S.heun ? 1 - S.phase : ((f + 1) & 1));
}
// Streams of the draws made while forces are evaluated, offset in the Heun corrector phase.
__device__ inline uint32_t force_stream(const Dev& S, uint32_t st) { return st + ST_PHASE * (uint32_t)S.phase; }

__device__ inline bool cell_active(const StepParams& P, const int* ncells, int cslot, int* s_out, int* c_out) {
  int s = species_of_cell(P, cslot);
  int c = cslot - P.base[s];
  *s_out = s;
  *c_out = c;
  return c < ncells[s];
}

__device__ inline int bin_coord(const StepParams& P, int d, double x) {
  int b = (int)floor(x * P.invw[d]);
  b = b < 0 ? 0 : b;
  b = b > P.ns[d] - 1 ? P.ns[d] - 1 : b;
  return b;
}

__device__ inline int bin_of(const StepParams& P, const double* x) {
  int b = bin_coord(P, 0, x[0]) * P.ns[1] + bin_coord(P, 1, x[1]);
  if (P.box.dim == 3) b = b * P.ns[2] + bin_coord(P, 2, x[2]);
  return b;
}

// Counting-sort key of particle p (slot index) at position x: element type major.
__device__ inline int bucket_key(const StepParams& P, int p, int s, const double* x) {
  int t = s * P.nE + (p % P.nE);
  return t * P.nbins + bin_of(P, x);
}

// ---------------------------------------------------------------------------------------
// Intra-cell forces (CIL.cu:30-100): migration (MD.cu:525-644), walls (Wall.hpp:68-79),
// passive noise.  Writes the full intra force of each element of the cell into F.
// `vel` may be modified (singleCellForce's zero-velocity kick), pos0/pos1 too (zero bond).
// ---------------------------------------------------------------------------------------
// The axis is a run-time value: the component is picked with selects, not an index, so the
// caller's position and force arrays stay in registers.
__device__ inline void flat_wall(const cil_wall& wl, double sigma, const double* r, double* f) {
  const double ra = wl.axis == 0 ? r[0] : (wl.axis == 1 ? r[1] : r[2]);
  double dr = (wl.positive ? (ra - wl.x0) : (wl.x0 - ra)) / sigma;
  if (dr <= wl.rcut) {
    dr = dr >= kWallCutoff ? dr : kWallCutoff;
    double ir2 = 1.0 / (dr * dr);
    double f0 = ir2 * ir2 * ir2;
    f0 = wl.epsilon * (2.0 * f0 - 1.0) * f0 * ir2 / sq(sigma);
    f0 = wl.positive ? f0 : -f0;
    if (wl.axis == 0) f[0] += f0;
    else if (wl.axis == 1) f[1] += f0;
    else f[2] += f0;
  }
}

// Random unit vector from Philox blocks k0, k0 + 1 of (id, ts, stream): normalised Gaussians.
__device__ inline void on_sphere_counter(uint64_t seed, uint32_t id, uint32_t ts, int D, double* n,
                                         uint32_t stream = ST_ZERO, uint32_t k0 = 0) {
  double z[4];
  philox_normal2(seed, id, ts, stream, k0, z);
  philox_normal2(seed, id, ts, stream, k0 + 1, z + 2);
  double r2 = 0.0;
  for (int d = 0; d < D; ++d) {
    n[d] = z[d];
    r2 += n[d] * n[d];
  }
  double norm = 1.0 / sqrt(r2);
  for (int d = 0; d < D; ++d) n[d] *= norm;
}

// DIM, NE: box dimension and discs per cell (compile-time, so the per-disc arrays stay in
// registers); they must equal P.box.dim and P.nE.
template <int DIM, int NE>
__device__ inline void intra_cell(const StepParams& P, const Dev& S, int s, int c, int cslot, int ts,
                                  double pos[2][3], double vel[2][3], const double* sig, double F[2][3],
                                  bool* pos1_changed) {
  const cil_species& sp = P.sp[s];
  constexpr int D = DIM, nE = NE;
  *pos1_changed = false;
  if (nE == 2) {
    double rij[3] = {0, 0, 0}, nij[3] = {0, 0, 0}, rnd[3] = {0, 0, 0};
    double r2 = box_dist(P.box, rij, pos[0], pos[1]);
    double r = sqrt(r2);
    if (r2 == 0.0) {  // MD.cu:534-540
      r = kBigTolMp * 0.5 * (sp.sigma[0] + sp.sigma[1]);
      on_sphere_counter(P.seed, cell_key(s, c), (uint32_t)ts, D, rij, force_stream(S, ST_ZERO));
      for (int d = 0; d < D; ++d) rij[d] *= r;
      r2 = sq(r);
      for (int d = 0; d < D; ++d) pos[1][d] = box_update_dr1(P.box, d, pos[1][d], rij[d]);
      *pos1_changed = true;
    }
    for (int d = 0; d < 3; ++d) nij[d] = rij[d];
    double spring, scf, a_div[2] = {0.0, 0.0};
    bool use_div = false;
    if (sp.spring == CIL_SPRING_HARMONIC) {
      double l = S.rest_len[cslot];
      spring = -sp.k_harm * (r - l) / r;
      scf = 1.0;
      for (int d = 0; d < D; ++d) nij[d] = rij[d] / r;
    } else {
      // Dividing state (MD.cu:547,560-568): divisionR2max replaces r2max, the activity of both
      // discs becomes -growthk0 fene(divisionR2min) + growthk1 + growthk2 r2^2 along the
      // unscaled axis, and the contraction factor 2 becomes 1
      const bool dividing = S.state[cslot] == CS_DIVIDING;
      double smax = sig[0] >= sig[1] ? sig[0] : sig[1];
      double r2max0 = dividing ? sp.div_r2max : sp.r2max;
      double r2max = r2max0 > smax ? r2max0 : smax;
      double ratio = sp.spring_ratio;
      if (!(r2max - r2 > kTolMp)) atomicOr(S.err, (unsigned)ERR_BOND);
      spring = r2 <= (r2max * ratio) ? sp.kappa_push * fabs(r2 - (r2max * ratio))
                                     : -sp.kappa / (1.0 - sq((r2 - r2max * ratio) / (r2max * (1 - ratio))));
      if (dividing) {
        const double fene_div_max = -sp.kappa / (1.0 - sq(sp.division_r2min / r2max));  // FENE, MD.cu:255-257
        a_div[0] = (-sp.growthk0[0] * fene_div_max + sp.growthk1[0] + sp.growthk2[0] * sq(r2));
        a_div[1] = (-sp.growthk0[1] * fene_div_max + sp.growthk1[1] + sp.growthk2[1] * sq(r2));
        use_div = true;
        scf = 1.0;
      } else {
        scf = 2.0;
        if (!sp.cil) {
          double scaling = sp.rss / r;
          for (int d = 0; d < D; ++d) nij[d] *= scaling;
        }
      }
    }
    if (P.active_noise && D == 2) {
      double Wr = S.awiener[(size_t)cslot * 2], Wt = S.awiener[(size_t)cslot * 2 + 1];
      rnd[0] = 0.5 * (Wr * rij[0] / r - Wt * rij[1]);
      rnd[1] = 0.5 * (Wr * rij[1] / r + Wt * rij[0]);
    }
    const double a0 = use_div ? a_div[0] : sp.motility[0], a1 = use_div ? a_div[1] : sp.motility[1];
    for (int d = 0; d < D; ++d) {
      double tc0 = -(a0 * nij[d] + spring * rij[d] * scf + rnd[d] * rij[d]);
      double tc1 = (a1 * nij[d] + spring * rij[d] * scf + rnd[d] * rij[d]);
      if (sp.quad_drag) {
        double v0 = vel[0][d], v1 = vel[1][d];
        double tf0 = sp.zeta[0] * sq(v0) * (v0 / fabs(v0));
        double tf1 = sp.zeta[1] * sq(v1) * (v1 / fabs(v1));
        F[0][d] = (isinf(tf0) || isnan(tf0)) ? tc0 : (tc0 - tf0);
        F[1][d] = (isinf(tf1) || isnan(tf1)) ? tc1 : (tc1 - tf1);
      } else {
        F[0][d] = tc0;
        F[1][d] = tc1;
      }
    }
  } else {
    double vl = 0.0;
    for (int d = 0; d < D; ++d) vl += vel[0][d] * vel[0][d];
    vl = sqrt(vl);
    if (vl == 0.0 && sp.motility[0] != 0.0) {  // MD.cu:609-614
      on_sphere_counter(P.seed, cell_key(s, c) << 1, (uint32_t)ts, D, vel[0], force_stream(S, ST_ZERO));
      for (int d = 0; d < D; ++d) vel[0][d] *= kBigTolMp;
      vl = 0.0;
      for (int d = 0; d < D; ++d) vl += vel[0][d] * vel[0][d];
      vl = sqrt(vl);
    }
    for (int d = 0; d < D; ++d) F[0][d] = sp.motility[0] * (vl > 0.0 ? vel[0][d] / vl : 0.0);
  }
  for (int e = 0; e < nE; ++e)
    for (int k = 0; k < P.nwalls; ++k) flat_wall(P.walls[k], sig[e], pos[e], F[e]);
  if (P.noise == CIL_NOISE_CELL_WIENER)
    for (int e = 0; e < nE; ++e)
      for (int d = 0; d < D; ++d) F[e][d] += sp.zeta[e] * S.wiener[(size_t)cslot * 3 + d];
}

}  // namespace cil
