/*
 * cil_hip.h — C ABI of the MI355X (gfx950) overdamped cell-colony MD step.
 *
 * The library replaces the reference's Euler step (`void Euler(int& ts)`, CIL.cu:1013-1020),
 * which the reference calls through the `MDIntegrator mdStep` function pointer
 * (MD.cuh:450-452, invoked GTS times per frame at CIL.cu:1106).  A host program keeps
 * its own per-species cell arrays (the reference's `CellColony`, Colony.hpp:81-113) and
 * hands plain pointers to this ABI; device state stays resident between calls.
 *
 * Conventions
 *  - Every function returns 0 (CIL_OK) or a negative cil_status; nothing throws.
 *  - Host pointers are never cached across calls (the reference re-points its arrays on
 *    x1.5 growth, Colony.cpp:193-273).
 *  - Arrays use the reference's row-major layout: pos(cell, elem, d) at
 *    (cell*nelem + elem)*dim + d (Array.h:557-600).
 *  - One context = one colony (or one batch of colonies), one HIP stream, one host thread.
 *  - Only plain C types cross the boundary.
 */
#ifndef CIL_HIP_H
#define CIL_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define CIL_ABI_VERSION 1
#define CIL_MAX_SPECIES 4
#define CIL_MAX_ELEMS 2
#define CIL_MAX_TYPES (CIL_MAX_SPECIES * CIL_MAX_ELEMS)
#define CIL_MAX_WALLS 4

typedef enum {
  CIL_OK = 0,
  CIL_E_INVALID = -1,     /* bad argument or parameter combination */
  CIL_E_NO_DEVICE = -2,   /* no HIP device / runtime failure */
  CIL_E_CAPACITY = -3,    /* colony grew beyond device capacity or host views */
  CIL_E_OUTSIDE = -4,     /* particle outside the box and drain disabled (CIL.cu:955) */
  CIL_E_NONFINITE = -5,   /* NaN/Inf found in positions or forces */
  CIL_E_HIP = -6,         /* a HIP call failed; see cil_hip_last_error */
  CIL_E_BOND = -7,        /* FENE bond over-stretched (MD.cu:549 assert) */
  CIL_E_IO = -8           /* a file could not be created or written */
} cil_status;

/* Pair-force law per element-type pair (reference: one global PairForceParams, MD.cuh:23-344;
 * a per-type-pair matrix is an extension so that LJ and soft-core can be mixed). */
typedef enum {
  CIL_LAW_NONE = 0,
  CIL_LAW_LJ = 1, /* LJParams  MD.cuh:47-129 */
  CIL_LAW_SC = 2, /* SCParams  MD.cuh:134-232 */
  CIL_LAW_HZ = 3, /* HZParams  MD.cuh:237-290, contact semantics of MD.cu:302-366 */
  CIL_LAW_EA = 4  /* EAParams  MD.cuh:295-344, contact semantics of MD.cu:302-366 */
} cil_law;

typedef enum {
  CIL_SPRING_FENE = 0,     /* reference FENE + push spring in r^2 (MD.cu:547-559) */
  CIL_SPRING_HARMONIC = 1  /* extension: F = -k (r - l) r_hat, rest length l per cell */
} cil_spring;

typedef enum {
  CIL_NOISE_NONE = 0,
  CIL_NOISE_CELL_WIENER = 1, /* reference: one Wiener vector per cell, F_e += zeta_e W (CIL.cu:84-98) */
  CIL_NOISE_FDT = 2          /* extension: per-disc dx += sqrt(2 kT mu dt) xi */
} cil_noise;

typedef enum {
  CIL_RNG_COUNTER = 0,  /* Philox4x32-10 keyed by (seed, step, id, stream) */
  CIL_RNG_EXTERNAL = 1, /* caller supplies the N(0,1) deviates (e.g. the reference's dSFMT stream) */
  CIL_RNG_DSFMT = 2     /* the reference's own sequential stream (Random::dsfmt, Random.hpp:8-52):
                           dSFMT-19937 seeded with `seed`, advanced by `rng_offset` outputs, drawn
                           on the device in the reference's order (Wiener noise, zero-velocity and
                           zero-bond directions, division geometry and timers).  A replay mode: the
                           draws are sequential, one device thread */
} cil_rng;

typedef enum {
  CIL_DIV_NONE = 0,
  CIL_DIV_THETA_CYCLE = 1,   /* calculDivisionsSingle, CIL.cu:469-619 (single-disc cells) */
  CIL_DIV_LINEAR_GROWTH = 2, /* extension: two-disc rest length grows at rate 1/tau, divide at l_div */
  CIL_DIV_BINARY = 3         /* calculDivisions, CIL.cu:369-455 (two-disc cells: Crawling <-> Dividing,
                                division check binaryCellDivisionCriteria MD.cuh:426-432, geometry
                                binaryCellDivision / binaryCellDivisionFail MD.cu:721-752) */
} cil_division;

typedef enum {
  CIL_INT_EULER = 0, /* Euler(int&), CIL.cu:1013-1020 */
  CIL_INT_HEUN = 1   /* StochasticHeun(int&), CIL.cu:1028-1049: half step with the old forces, forces at the
                        half-step positions, second half step, divisions, forces */
} cil_integrator;

typedef enum {
  CIL_PREC_FP64 = 0, /* forces in FP64 (parity mode) */
  CIL_PREC_FP32 = 1  /* forces in FP32, positions accumulated in FP64 (fast mode) */
} cil_precision;

/* Normal-distribution parameters (reference `Gaussian`, Cell.cuh). */
typedef struct {
  double mu;
  double sigma;
} cil_gaussian;

/* Per-species parameters; field names follow CellParams/ElemParams (Cell.cuh:32-266) after
 * the unit scaling and derivations of readSpecs (Input.cpp:6-293). */
typedef struct {
  int nelem;                        /* 1 or 2; all species of a context share it */
  double sigma[CIL_MAX_ELEMS];      /* SIGMA_CORE, default element diameters */
  double zeta[CIL_MAX_ELEMS];       /* FRICTION */
  double zeta_p[CIL_MAX_ELEMS];     /* P_FRICTION */
  double zeta_p_beta[CIL_MAX_ELEMS];/* P_FRICTION_BETA */
  double motility[CIL_MAX_ELEMS];   /* MOTILITY */
  /* pair-law element parameters (meaning depends on the law, Input.cpp:48-82):
   *   SC: eps_core = EPSILON_CORE, eps_well = EPSILON_RATIO*EPSILON_CORE, sig_well = SIGMA_RATIO*SIGMA_CORE
   *   HZ/EA: eps_core = ELASTIC_MODULUS, eps_well = ADHESION_CONSTANT, sig_well = POISSON_RATIO
   *   LJ: eps_well = EPSILON */
  double eps_core[CIL_MAX_ELEMS];
  double eps_well[CIL_MAX_ELEMS];
  double sig_well[CIL_MAX_ELEMS];
  double cross_adhesion;            /* CellParams::crossAdhesion */
  /* intra-cell bond of two-disc cells */
  int spring;                       /* cil_spring */
  double kappa, kappa_push, spring_ratio, r2max, div_r2max; /* FENE block (Input.cpp:117-135) */
  double rss;                       /* steady-state distance (Input.cpp:138-143) */
  int cil;                          /* contact inhibition of locomotion flag */
  int quad_drag;                    /* binaryCellForce's -zeta v^2 sign(v) term (MD.cu:588-595) */
  double k_harm;                    /* harmonic extension: spring constant */
  /* noise */
  double d_noise;                   /* passive Wiener amplitude sqrt(2D)/sqrt(dt) (Input.cpp:108-114, CIL.cuh:210-215) */
  double d_noise_radial, d_noise_angular; /* active noise amplitudes (two-disc, 2D) */
  double kT;                        /* FDT extension */
  /* theta cell cycle (calculDivisionsSingle, CIL.cu:469-619); tau_* already divided by dt */
  double cycle_p0, cycle_tau_r, cycle_tau_v, cycle_tau_p, cycle_v_s, cycle_v_1, cycle_beta_vs;
  double cycle_rmaxed_inv, cycle_rho_ratio, cycle_p_shift, cycle_p_shift_tau_inv, apop_pressure;
  cil_gaussian division_tau;        /* log-normal division time (CIL.cu:591-592) */
  cil_gaussian migration_tau;       /* normal switch time (CIL.cu:580-581) */
  /* linear-growth extension (two-disc) */
  double growth_rate;               /* d l / d t = growth_rate (= 1/tau) */
  double l_div;                     /* divide when the rest length reaches l_div */
  double l_birth;                   /* daughters' rest length */
  double sigma_jitter;              /* daughter diameters sigma*(1 + jitter*N(0,1)) */
  /* division switches and the two-disc division cycle (Input.cpp:148-245) */
  int no_division;                  /* CellParams::DIV == false: the species never divides (its divisionFunc
                                       is NULL, CIL.cuh:18-47); the default 0 divides under `division` */
  int division_swelling;            /* DIVISION.SWELLING: updateDiameters (CIL.cu:970-990) */
  double division_r0;               /* divisionR0: daughter half-separation (MD.cu:721-752) */
  double division_r2min;            /* divisionR2min = Rmin^2: division check and feneDivMax (MD.cu:560) */
  double division_r2min_ratio;      /* divisionR2minRatio: when nonzero the check is r2 >= ratio*sigma0 */
  double growthk0[CIL_MAX_ELEMS];   /* GROWTHK0..2: Dividing-state activity a_e = -k0 fene(Rmin^2) + k1 + k2 r^4 */
  double growthk1[CIL_MAX_ELEMS];   /*   (binaryCellForce, MD.cu:561-566) */
  double growthk2[CIL_MAX_ELEMS];
} cil_species;

typedef struct {
  int kind;        /* 0 = FlatWall (Wall.hpp:62-101) */
  int axis;        /* 0,1,2 */
  double x0;
  double rcut;     /* in units of sigma (default 2^(1/6), Input.cpp:381) */
  double epsilon;
  int positive;    /* wall pushes toward +axis */
} cil_wall;

typedef struct {
  int abi_version;                  /* must be CIL_ABI_VERSION */
  int dim;                          /* 2 or 3 */
  int nspecies;
  cil_species species[CIL_MAX_SPECIES];
  /* element type t = species*nelem + elem; law and cut-off factor per type pair */
  int law[CIL_MAX_TYPES][CIL_MAX_TYPES];
  double rcut2[CIL_MAX_TYPES][CIL_MAX_TYPES]; /* PairForceParams::rcut2 (RCUT^2; default (2^(1/6))^2 LJ, 1 otherwise) */
  int force_shifted;                /* OPT::FORCE_SHIFTED */
  double dt;
  int noise;                        /* cil_noise */
  int active_noise;                 /* OPT::NOISE_ACTIVE */
  int rng;                          /* cil_rng */
  uint64_t seed;
  int division;                     /* cil_division */
  double osmotic_pi;                /* MDParams::pi */
  int nwalls;
  cil_wall walls[CIL_MAX_WALLS];
  int drain;                        /* MDParams::drain (CIL.cu:945-966) */
  int precision;                    /* cil_precision */
  int pressure;                     /* accumulate pair pressure (needed when zeta_p != 0 or theta cycle) */
  double max_disp;                  /* extension for relaxing random initial states: per-step displacement
                                       cap (0 = off, the reference behaviour) */
  int integrator;                   /* cil_integrator (MDParams::INTEGRATOR, Input.cpp:520-540) */
  uint64_t rng_offset;              /* CIL_RNG_DSFMT: outputs of the stream consumed before the first
                                       draw of this context (e.g. by initializeCellStatus) */
} cil_params;

typedef struct {
  double length[3];
  int periodic[3];
} cil_box;

/* View of one species' arrays in the host colony (CellColony, Colony.hpp:81-113).
 * Optional arrays may be NULL.  `capacity` = number of cells the arrays can hold. */
typedef struct {
  int ncells;
  int capacity;
  double* pos;       /* ncells*nelem*dim */
  double* frc;       /* ncells*nelem*dim   (download only) */
  double* vel;       /* ncells*nelem*dim */
  double* sigma;     /* ncells*nelem      sigmaCore */
  double* sigma_t;   /* ncells            sigmaCoreT (theta cycle) */
  double* pressure;  /* ncells */
  int* state;        /* ncells CellState */
  int* t0;           /* ncells */
  int* tswitch;      /* ncells */
  int* tpressure;    /* ncells */
  double* theta;     /* ncells */
  double* rho;       /* ncells */
  double* ang_vel;   /* ncells */
  double* rest_len;  /* ncells  (linear-growth / harmonic extension) */
} cil_species_view;

/* A reference run's settings beyond cil_params (cil_hip_config_read). */
typedef struct {
  int nspecies;
  int gts;                               /* MD.GTS: steps per frame */
  int frames;                            /* MD.FRAMES */
  int ncells[CIL_MAX_SPECIES];           /* cells the INIT block places per species */
  char species[CIL_MAX_SPECIES][32];     /* SPECIES.DEFINE names */
  char elements[CIL_MAX_SPECIES][CIL_MAX_ELEMS][32]; /* SPECIES.<name>.NAMES */
  double kbt;                            /* MD.KBT (read; only the DPD integrator uses it) */
  long seed;                             /* RUNTIME.RNG.SEED as used (<= 0: time(NULL), Input.cpp:577-583) */
  /* parameters the step does not use, kept for the run's HDF5 parameters */
  double sigma_max[CIL_MAX_SPECIES][CIL_MAX_ELEMS]; /* PAIRFORCES.SC.SIGMA_MAX (0 for other laws) */
  double f_ext[CIL_MAX_SPECIES][3];      /* EXTFORCE (the reference adds 0 for it, CIL.cu:55-66) */
  int div_app_record;                    /* OUTPUT.DIV_APP_RECORD: the HDF5 file keeps /div_app */
} cil_run;

typedef struct cil_ctx cil_ctx;

/* Create a context: device buffers sized for `capacity[s]` cells per species.
 * Replaces the reference's initialize* chain for the step (CIL.cuh:332-393). */
int cil_hip_create(const cil_params* params, const cil_box* box, const int* capacity, cil_ctx** out);
int cil_hip_destroy(cil_ctx* ctx);

/* Copy the host colony to the device (all species). */
int cil_hip_upload(cil_ctx* ctx, const cil_species_view* views, int nspec);

/* Run the Euler phase order (CIL.cu:1013-1020) `nsteps` times; *ts is incremented per step
 * exactly like Euler(int& ts).  Requires forces to be initialised (cil_hip_forces or a
 * previous step), as the reference does at CIL.cu:1096. */
int cil_hip_step(cil_ctx* ctx, int nsteps, int* ts);

/* Copy the device colony back.  If a species holds more cells than views[s].capacity,
 * returns CIL_E_CAPACITY and writes the needed counts to required_cells[s]. */
int cil_hip_download(cil_ctx* ctx, cil_species_view* views, int nspec, int* required_cells);

/* calculForcesPressure at the current state (CIL.cu:999-1007; init call CIL.cu:1096).
 * Draws noise when the noise mode is on.  If frc/pressure are non-NULL they receive the
 * forces (UID order, FP64) and pressures. */
int cil_hip_forces(cil_ctx* ctx, double* frc, double* pressure);

/* Enumerate interacting element pairs (UID_i < UID_j, different cells, law predicate true)
 * in ascending (i, j) order.  Writes up to `cap` pairs as ij[2k], ij[2k+1]; *n = total. */
int cil_hip_pairs(cil_ctx* ctx, int32_t* ij, size_t cap, size_t* n);

/* cil_hip_pairs without device allocation (no counterpart in the reference): for callers that
 * cannot allocate device memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  Same result as cil_hip_pairs; the pairs and their counter are staged in a
 * fixed device array of the library (one per GPU, held until the process ends) instead of a
 * buffer allocated to `cap`.  cap <= 1048576 pairs, else CIL_E_INVALID.  Calls on one GPU share
 * that array: calls that overlap on one GPU (contexts used from different threads) race. */
int cil_hip_pairs_noalloc(cil_ctx* ctx, int32_t* ij, size_t cap, size_t* n);

/* checkPositions (CIL.cu:945-966): out-of-box cells raise CIL_E_OUTSIDE unless drain,
 * in which case they are removed by swap-with-last with the reference's skip quirk. */
int cil_hip_check_positions(cil_ctx* ctx, int* removed);

/* Supply external N(0,1) deviates for CIL_RNG_EXTERNAL: `nsteps` blocks of `per_step`
 * doubles in the reference's draw order (CIL.cu:793-818: all species' passive vectors,
 * species-major, cell-major, dim-minor; then active radial and angular blocks).  Block k
 * feeds the noise draw at step index ts + k, where ts is the context's current step index
 * (the draw of cil_hip_forces is at ts, the draw of the j-th following step at ts + j). */
int cil_hip_set_external_noise(cil_ctx* ctx, const double* normals, size_t per_step, int nsteps);

/* Current number of cells per species (device truth). */
int cil_hip_num_cells(cil_ctx* ctx, int* ncells);

/* Time spent per kernel family over the last cil_hip_step call, in microseconds
 * (measured with HIP events on the context stream); `names` is a ';'-separated list. */
int cil_hip_timings(cil_ctx* ctx, double* us, int cap, int* n, char* names, size_t names_cap);

/* Bin geometry in use: number of bins per dimension and bin width. */
int cil_hip_bin_info(cil_ctx* ctx, int* nbins, double* width);

/* Autotuner knobs (project addition): bins per cut-off (1..4) and threads per block of the
 * pair kernel (64, 128 or 256; 0 keeps the current value). */
int cil_hip_set_tuning(cil_ctx* ctx, int bins_per_cutoff, int block);

/* FP32 pair-kernel variant (project addition, autotuner knob): 0 = type-specialised lean
 * kernel (default), 1 = generic kernel. */
int cil_hip_set_kernel(cil_ctx* ctx, int variant);

/* Division records (project addition; the reference's posCellsDiv / cellT0Div, written to the
 * HDF5 group /div_app, Colony.cpp:319-344, OutputWriter.cpp:439-470).  on = 1 records, for each
 * division, the new cell's element positions right after the division and its step, per
 * species in the order the reference's loop divides (parent order), up to the species'
 * capacity between reads; on = 0 stops and frees.  Enable before creating a batch that holds
 * the context.  cil_hip_division_log copies species s's records (pos [n][nelem][dim], ts [n];
 * either may be NULL) and clears them when clear != 0; *n = records held (only cap copied). */
int cil_hip_set_division_log(cil_ctx* ctx, int on);
int cil_hip_division_log(cil_ctx* ctx, int s, double* pos, int* ts, int cap, int* n, int clear);

/* Autotuner (project addition): times `reps` force passes for every bins-per-cut-off /
 * pair-block / kernel-variant candidate, keeps the fastest and reports it.  In neighbour-list
 * mode it chooses the bins per cut-off (1-3) that minimise a pass without a rebuild plus a
 * rebuild pass spread over 45 x (skin / 0.3)^2 steps, plus with two-level lists an inner-list
 * refresh spread over 45 x (skin_in / 0.3)^2 steps (block and variant are reported unchanged).  Needs forces
 * computed by cil_hip_forces or a step (the state is not advanced). */
int cil_hip_autotune(cil_ctx* ctx, int reps, int* best_bins, int* best_block, int* best_variant, double* best_us);

/* Per-kernel-family HIP-event timing (project addition; off by default).  on = 1 times every
 * family; any other nonzero value is a mask with bit f+1 set for family f, in the order of
 * the names cil_hip_timings returns (update, division, sort, pair, noise, guard, build, knn,
 * mlp, macro).  Timed steps are launched directly (no graph): each timed family adds two
 * event records per step.  The FP32 guard-band fix and the neighbour-list build run inside
 * the pair kernels, so "guard" and "build" stay at zero and "pair" includes them. */
int cil_hip_set_timing(cil_ctx* ctx, int on);

/* Neighbour-list mode (project addition; FP32 force precision, fewer than 2^22 particle slots,
 * max_neighbours * particle slots * 4 bytes of list below 4 GiB, 8 <= max_neighbours <= 4096).
 * max_neighbours is the list slots per particle, rounded up to a multiple of 4; the entries of
 * each source element type start on a 4-slot boundary, so a particle holds at least
 * max_neighbours - 3 * (element types - 1) entries before the step fails with CIL_E_CAPACITY.
 * Pairs within their cut-off + skin are listed at a build; each step evaluates only the list
 * and re-sorts / rebuilds when a particle has moved more than skin / 2 since the build, when
 * cells divide or are removed, or after an upload.  The pair set equals the stencil kernel's
 * (FP32 summation order differs).  on = 0 returns to the stencil kernels.  Bins are sized for
 * cut-off + skin while it is on. */
int cil_hip_set_neighbour_list(cil_ctx* ctx, int on, double skin, int max_neighbours);

/* Two-level neighbour lists (project addition; call after cil_hip_set_neighbour_list, which
 * returns to one level).  The list above (cut-off + skin) becomes the outer list; an inner list
 * holds its entries within cut-off + skin_in (0 < skin_in < skin), in the same order, and the
 * passes evaluate the inner list.  A pass refreshes the inner list from the outer one, while
 * evaluating the outer list, when the outer list was rebuilt or a particle has moved more than
 * skin_in / 2 since the last refresh.  The pairs within their cut-off, and the order in which
 * they are summed, are those of a one-level list, so forces are bit-identical to it; a large
 * skin then makes re-sorts and list builds rare while the per-step list stays short.  Needs the
 * Euler integrator and no division (CIL_E_INVALID otherwise); skin_in = 0 returns to one level.
 * The drift surrogate needs one-level lists. */
int cil_hip_set_neighbour_list_inner(cil_ctx* ctx, double skin_in);

/* ---------------------------------------------------------------------------------------
 * Reference run configurations (project addition: the reference's own JSON input files).
 *
 * cil_hip_config_read parses a configuration in the reference's format (readMD, readSpecs,
 * readInteractions, readRunTime: Input.cpp:6-590; initializeTimeStep: CIL.cuh:167-250) into
 * cil_params / cil_box and the run settings.  The parameters come out FP64, with the reference's
 * noise (Euler: none; STOCHASTIC_EULER: per-cell Wiener; STOCHASTIC_HEUN: Wiener and active
 * noise), its division model (Euler: theta cell cycle; Heun: two-disc cycle) and its random
 * stream (CIL_RNG_DSFMT seeded by RUNTIME.RNG.SEED).  Supported: UNITS all 1, DT MANUAL (or
 * AUTOSCALE with LJ / SC), the EULER / STOCHASTIC_EULER / STOCHASTIC_HEUN integrators, the LJ /
 * SC / HZ / EA laws, SLAB walls, one- or two-element species with FENE bonds, division and no
 * apoptosis (EXTFORCE adds nothing, as in the reference, CIL.cu:55-66), INIT JSON or RANDOM.  Anything else returns CIL_E_INVALID with
 * the reason in err.
 * cil_hip_config_init builds the initial colony into the views (capacity >= run.ncells):
 * readInit, initializeCellStatus (CIL.cuh:253-330) on the reference's stream, readInit again
 * (CIL.cuh:372-376), and stores the stream outputs consumed in params->rng_offset, so that a
 * context created from params continues the reference's stream where the initialisation left it.
 * ------------------------------------------------------------------------------------- */
int cil_hip_config_read(const char* json, cil_params* params, cil_box* box, cil_run* run, char* err, size_t errcap);
int cil_hip_config_init(const char* json, cil_params* params, cil_species_view* views, int nspec, char* err,
                        size_t errcap);
/* n draws of the reference's stream (dSFMT-19937 from seed, after `skip` outputs): kind 0
 * Random::uniform0x(1), 1 Random::normal, 2 dsfmt_genrand_open_open (project addition). */
int cil_hip_dsfmt_fill(uint32_t seed, uint64_t skip, int kind, double* out, size_t n);

/* ---------------------------------------------------------------------------------------
 * Frames in the reference's GSD format (project addition: the reference writes them with the
 * vendored gsd.c through outputGSDFrame, OutputWriter.cpp:505-657).
 *
 * cil_hip_gsd_create / _write_chunk / _end_frame / _close replace gsd_create + gsd_open
 * (GSD_OPEN_APPEND), gsd_write_chunk, gsd_end_frame and gsd_close for writing (gsd.h:113-143):
 * file format 1.0, the same layout and growth of the index and name list, so a file written
 * chunk for chunk as the reference writes it is identical byte for byte.  Chunk types are
 * gsd.h's gsd_type values (CIL_GSD_*).
 *
 * cil_hip_gsd_write_frame writes one frame of a host colony as outputGSDFrame does:
 * configuration/step, /dimensions (3), /box (lx, ly, lz; lz in 2D = the largest species
 * SIGMA_CORE, OutputWriter.cpp:715); particles/N; particles/types at ts == 0 ("species" or
 * "species_element", 63 bytes, names from `run`, which may be NULL: C1, C2, ... and E0, E1);
 * particles/Z, /diameter (sigma) and /typeid (species * nelem + element) when a species divides
 * or ts == 0; particles/charge (rho) when a species divides; particles/position (centred on the
 * box; in 2D z = -0.01 * element); particles/velocity; particles/mass (the cell pressure).
 * cil_hip_gsd_write_frame_ctx does the same from a context's device colony. */
typedef enum {
  CIL_GSD_UINT8 = 1, CIL_GSD_UINT16 = 2, CIL_GSD_UINT32 = 3, CIL_GSD_UINT64 = 4,
  CIL_GSD_INT8 = 5, CIL_GSD_INT16 = 6, CIL_GSD_INT32 = 7, CIL_GSD_INT64 = 8,
  CIL_GSD_FLOAT = 9, CIL_GSD_DOUBLE = 10
} cil_gsd_type;
typedef struct cil_gsd cil_gsd;
int cil_hip_gsd_create(const char* path, const char* application, const char* schema, uint32_t schema_version,
                       cil_gsd** out);
int cil_hip_gsd_write_chunk(cil_gsd* g, const char* name, int type, uint64_t N, uint32_t M, const void* data);
int cil_hip_gsd_end_frame(cil_gsd* g);
int cil_hip_gsd_close(cil_gsd* g);
int cil_hip_gsd_write_frame(cil_gsd* g, const cil_params* params, const cil_box* box, const cil_run* run,
                            const cil_species_view* views, int nspec, int ts);
int cil_hip_gsd_write_frame_ctx(cil_ctx* ctx, cil_gsd* g, const cil_run* run, int ts);

/* HDF5 snapshots in the reference's layout (project addition: initHDF5Output and
 * outputHDF5Frame, OutputWriter.cpp:240-436, through the reference's SimIO wrapper).
 * cil_hip_h5_create writes /params (dt, L, Tdump = run->gts, Force), /params/units (all 1) and
 * /params/cell/<species> (flags, division and migration times, noise, element parameters,
 * bond matrices; `run` may be NULL) and creates /snapshots; with params NULL it creates a bare
 * file (the reference's `.Last.h5`).  cil_hip_h5_write_frame adds /snapshots/t_<id> (attribute
 * ts), or for id < 0 writes at the root, R, Vel, SigmaCore, F, FD, FR, RHO, Theta, PR, State,
 * T0 and CpuTime per species (a subgroup per species when there are several), reals as 32-bit
 * floats.  FD and FR (the DPD forces) are zero.  With run->div_app_record the file also holds
 * /div_app/DIV and /div_app/APP, and cil_hip_h5_write_frame_ctx writes each frame's division
 * records (d_<id>: R, T) from the context's division log (cil_hip_set_division_log must be on;
 * cil_hip_h5_write_frame, from host views, writes none).  There is no apoptosis: APP stays
 * empty.
 * The HDF5 library is loaded at the first create ($CIL_HDF5_LIB, libhdf5.so, or
 * /opt/conda/lib/libhdf5.so); without it the calls return CIL_E_IO. */
typedef struct cil_h5 cil_h5;
int cil_hip_h5_create(const char* path, const cil_params* params, const cil_box* box, const cil_run* run, cil_h5** out);
int cil_hip_h5_write_frame(cil_h5* f, const cil_params* params, const cil_species_view* views, int nspec,
                           const cil_run* run, int id, int ts, double cpu_time);
int cil_hip_h5_write_frame_ctx(cil_ctx* ctx, cil_h5* f, const cil_run* run, int id, int ts, double cpu_time);
int cil_hip_h5_close(cil_h5* f);

/* sizeof the ABI structs (0: cil_params, 1: cil_species, 2: cil_species_view, 3: cil_box,
 * 4: cil_surrogate_stats, 5: cil_run),
 * so an FFI binding can verify its layout (project addition). */
size_t cil_hip_sizeof(int which);

const char* cil_hip_last_error(cil_ctx* ctx);
const char* cil_hip_version(void);

/* ---------------------------------------------------------------------------------------
 * Batches of independent colonies (project addition; the reference runs one colony per
 * process, CIL.cu:1072-1120).
 *
 * A batch advances many contexts (one colony each: own parameters, box, seed, bins and
 * buffers) with one launch per kernel for all of them and one HIP graph per step loop.
 * Every colony evolves exactly as it would under cil_hip_step alone.  The contexts must share
 * the device, dimension, force precision, division model, noise kernel (cell Wiener / active
 * noise or not), LJ presence, pair kernel variant and block size, and use counter-based RNG.
 * They stay owned by the caller and must outlive the batch; cil_hip_upload / download /
 * check_positions on a member context remain valid between batch steps.
 * ------------------------------------------------------------------------------------- */
typedef struct cil_batch cil_batch;

int cil_hip_batch_create(cil_ctx* const* colonies, int n, cil_batch** out);
int cil_hip_batch_destroy(cil_batch* batch);
int cil_hip_batch_size(cil_batch* batch);
/* nsteps Euler steps of every colony; ts[c] is colony c's step counter (in/out).  Every
 * colony needs initialised forces (cil_hip_forces). */
int cil_hip_batch_step(cil_batch* batch, int nsteps, int* ts);
/* Times the batch force pass for bins per cut-off 1..3 x pair block 64/128/256 (x FP32
 * kernel variant) and applies the fastest to every colony. */
int cil_hip_batch_autotune(cil_batch* batch, int reps, int* best_bins, int* best_block, int* best_variant,
                           double* best_us);
const char* cil_hip_batch_last_error(cil_batch* batch);

/* ---------------------------------------------------------------------------------------
 * Drift surrogate (project addition; no counterpart in the reference).
 *
 * A per-particle MLP (72 -> 128 -> 128 -> 3, SiLU, BF16 weights and activations on MFMA,
 * FP32 accumulation) predicts the correction to the one-pass drift b = 8 dt mu F(x0) of a
 * macro-step of 8 Euler steps.  Features: the 16 nearest neighbours, each given as (u m, 20 g):
 * u the unit vector towards it, m = sigma_j (species 0) or sigma_j + 4 (species 1),
 * g = r / sigma_bar - 1; own sigma and species; b and the sum z of the 8 steps' FDT noise
 * increments, both / 0.01 (72 features).  A macro-step moves x0 + b + 0.002 y + z, the
 * correction 0.002 y capped at 0.05 sigma per particle (a trust bound).  Supported: 3D
 * periodic single-disc colonies in FP32 force mode, no
 * division, walls, motility, pressure friction, drag or active noise; other set-ups return
 * CIL_E_INVALID.  The kNN search visits shells of bins around each particle until the 16th
 * neighbour is covered, up to 3 shells (fewer in boxes of < 7 bins per dimension); a 16th
 * neighbour beyond that returns CIL_E_INVALID.  In neighbour-list mode the 16 nearest come from
 * per-particle candidate lists rebuilt with the neighbour list (exact; particles whose list
 * cannot guarantee it take the bin walk), and macro-steps keep the list's skin bookkeeping.
 * ------------------------------------------------------------------------------------- */
typedef struct cil_surrogate_stats {
  long macro_steps;     /* macro-steps advanced by the network */
  long plain_steps;     /* Euler steps run by the plain integrator (judging and fallback) */
  long judged;          /* judged macro-steps */
  long failed;          /* judged macro-steps whose RMS error exceeded the tolerance */
  double last_err_rms;  /* RMS over particles of |x_network - x_plain| at the last judgement */
  double last_err_max;
  double worst_err_rms;
  long knn_list_builds; /* neighbour-list mode: kNN candidate-list builds */
  long knn_fallbacks;   /* neighbour-list mode: particles whose kNN came from the bin walk */
} cil_surrogate_stats;

/* Number of network parameters (W1[128][72], b1, W2[128][128], b2, W3[3][128], b3; row-major,
 * out x in, in this order) = 26243. */
int cil_hip_mlp_nparams(void);
int cil_hip_mlp_set_params(cil_ctx* ctx, const float* params, size_t n);
int cil_hip_mlp_get_params(cil_ctx* ctx, float* params, size_t n);
/* Weights and biases uniform in (-1/sqrt(fan_in), 1/sqrt(fan_in)), from Philox keyed by seed;
 * resets the Adam state. */
int cil_hip_mlp_init(cil_ctx* ctx, uint64_t seed);
/* Y (n, 3) = 0.002 x network output (the drift correction, length units) for features
 * X (n, 72), rounded to BF16 on entry. */
int cil_hip_mlp_forward(cil_ctx* ctx, const float* X, size_t n, float* Y);
/* Gradient of the loss mean((y - T / 0.002)^2) / 2 over the n x 3 outputs (y = network output
 * before the 0.002 output scale; T in length units), no update.  n <= 65536. */
int cil_hip_mlp_grad(cil_ctx* ctx, const float* X, const float* T, size_t n, float* grad, float* loss);
/* Features of the current state at the context's step counter (the last ts an entry point
 * returned), rows in UID order (species-major); computes the forces if none are held. */
int cil_hip_knn_features(cil_ctx* ctx, float* F, size_t rows);
/* Runs nmacro x 8 plain Euler steps from the current state (ts advances by 8 per macro-step)
 * and appends, for every stride-th sorted particle of each macro-step start (rotating phase),
 * its features and its 8-step displacement minus the noise sum z and the one-pass drift b
 * (the correction the network learns) to the device training set (capacity cap samples). */
int cil_hip_surrogate_collect(cil_ctx* ctx, int nmacro, int stride, long cap, int* ts, long* nsamples);
/* Replaces the training set with n host samples: X (n, 72) rounded to BF16, T (n, 3) drift
 * correction in length units. */
int cil_hip_surrogate_set_samples(cil_ctx* ctx, const float* X, const float* T, long n);
/* Copies up to cap training samples back (X as float, T in length units); *n = samples held. */
int cil_hip_surrogate_get_samples(cil_ctx* ctx, float* X, float* T, long cap, long* n);
int cil_hip_surrogate_clear_samples(cil_ctx* ctx);
/* nsteps Adam steps (beta 0.9/0.999, eps 1e-8) on minibatches of `batch` samples drawn with
 * replacement (Philox keyed by seed); losses[it] receives each minibatch loss (may be NULL). */
int cil_hip_surrogate_train(cil_ctx* ctx, int nsteps, int batch, float lr, uint64_t seed, float* losses);
/* One judged macro-step from the current state: the network, zero drift and one Euler step of
 * 8 dt are each compared with 8 plain steps (whose result is kept; ts += 8).
 * err_rms[3], err_max[3] in that order. */
int cil_hip_surrogate_eval(cil_ctx* ctx, int* ts, double* err_rms, double* err_max);
/* nmacro surrogate-assisted macro-steps (ts += 8 each).  Every judge_every-th macro-step is
 * judged: the plain integrator runs it (its result is kept) and the network's prediction is
 * compared; an RMS error above tol switches to the plain integrator for the next judge_every
 * macro-steps, after which the next macro-step is judged again.  judge_every 0: no judging.
 * st (may be NULL) receives the cumulative statistics. */
int cil_hip_surrogate_step(cil_ctx* ctx, int nmacro, int judge_every, double tol, int* ts,
                           cil_surrogate_stats* st);
int cil_hip_surrogate_stats_reset(cil_ctx* ctx);

#ifdef __cplusplus
}
#endif
#endif /* CIL_HIP_H */
