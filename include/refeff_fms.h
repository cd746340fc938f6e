/*
 * refeff_fms.h -- C ABI of the MI355X-native FMS (full multiple scattering)
 * stage that stands in for refeff's real-space FMS solve.
 *
 * The entry points mirror refeff's own seams (paths relative to the refeff
 * repository root):
 *   - cluster preparation   crates/refeff-core/src/fms/cluster.rs:10-241
 *   - state kets            crates/refeff-core/src/state.rs:97-148
 *   - plan                  crates/refeff-core/src/fms/driver.rs:105-140
 *                           (FmsRealSpacePlanInput, fms/types.rs:110-157)
 *   - energy sweep          crates/refeff-core/src/fms/driver.rs:231-236
 *                           (fms_real_space_spectrum)
 *   - MKGTR trace           crates/refeff-core/src/fms/mkgtr.rs:790-953
 *   - module stage          crates/refeff-engine/src/fms.rs:462-539 (run_fms_in_dir)
 *                           crates/refeff-engine/src/fms.rs:559-631 (run_mkgtr_in_dir)
 *
 * Conventions: plain pointers and sizes only.  Complex numbers are two
 * consecutive doubles (re, im).  Every function returns a refeff_fms_status;
 * none aborts or prints.  The caller owns every array; the library owns the
 * plan and its device memory.  A plan may be used from any thread, one batch
 * at a time (calls on one plan are serialised internally).
 */
#ifndef REFEFF_FMS_H
#define REFEFF_FMS_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes; one per FmsError family (crates/refeff-core/src/fms/types.rs:830-981). */
typedef enum refeff_fms_status {
    REFEFF_FMS_OK = 0,
    REFEFF_FMS_INVALID_ARGUMENT = 1,       /* InvalidAngularLimit, TableIndexOutOfRange, InvalidClusterRadius, ... */
    REFEFF_FMS_MISSING_CENTRAL_ATOM = 2,   /* MissingCentralAtom */
    REFEFF_FMS_DUPLICATE_ABSORBER = 3,     /* DuplicateAbsorber */
    REFEFF_FMS_NON_FINITE = 4,             /* NonFiniteCoordinate / WaveNumber / PhaseShift / MeanSquareDisplacement */
    REFEFF_FMS_ZERO_RHO = 5,               /* ZeroRho */
    REFEFF_FMS_POTENTIAL_OUT_OF_RANGE = 6, /* PotentialOutOfRange, StateKetPotentialOutOfRange */
    REFEFF_FMS_MISSING_REPRESENTATIVE = 7, /* MissingRepresentativePotential */
    REFEFF_FMS_SINGULAR = 8,               /* LinearAlgebra(SingularMatrix) */
    REFEFF_FMS_BREAKDOWN = 9,              /* IterativeSolverBreakdown */
    REFEFF_FMS_CENTRAL_MISMATCH = 10,      /* CentralAtomMismatch */
    REFEFF_FMS_UNSUPPORTED = 11,           /* a solver branch this library does not provide */
    REFEFF_FMS_NO_MEMORY = 12,
    REFEFF_FMS_NO_CONVERGENCE = 13,        /* IterativeSolverNoConvergence (refinement / Krylov) */
    REFEFF_FMS_DEVICE_ERROR = 14,          /* HIP runtime failure or no device */
    REFEFF_FMS_INVALID_SPIN_COUNT = 15,    /* InvalidSpinChannelCount */
    REFEFF_FMS_EMPTY_CLUSTER = 16,         /* EmptyCluster */
    REFEFF_FMS_IO_ERROR = 17               /* file handoff could not be read / written / parsed */
} refeff_fms_status;

/* Solver selection inside a plan.  All of them return the LU answer of
 * fms_lu_scattering (crates/refeff-core/src/fms/solvers.rs:366-484) except
 * REFEFF_FMS_SOLVER_GGBI, which reproduces refeff's one-shot minv=1 branch
 * (crates/refeff-core/src/fms/internals/iterative.rs:181-254). */
typedef enum refeff_fms_solver {
    REFEFF_FMS_SOLVER_AUTO = 0,    /* fms_scattering_method_selection (fms/driver.rs:72-95): minv 0 -> MIXED,
                                      1 -> GGBI, 2 -> GGRM, 3 -> GGGM, any other -> GGTF; a full scattering
                                      request -> MIXED */
    REFEFF_FMS_SOLVER_MIXED = 1,   /* complex FP32 MFMA LU + FP64 iterative refinement */
    REFEFF_FMS_SOLVER_FP64 = 2,    /* plain complex FP64 MFMA LU */
    REFEFF_FMS_SOLVER_GGBI = 3,    /* refeff minv=1 one-shot BiCGStab, in FP64 (refeff runs it in c32) */
    REFEFF_FMS_SOLVER_GGRM = 4,    /* minv=2 recursion (fms_recursion_scattering), in FP64 */
    REFEFF_FMS_SOLVER_GGGM = 5,    /* minv=3 Graves-Morris/Salam (fms_graves_morris_scattering), in FP64 */
    REFEFF_FMS_SOLVER_GGTF = 6,    /* other minv: TFQMR (fms_tfqmr_scattering), in FP64 */
    REFEFF_FMS_SOLVER_BICGSTAB = 7 /* converged BiCGSTAB per energy: FP32 MFMA matvecs, FP64 x and reliable
                                      updates from the FP64 residual (relative residual 1e-10 per column);
                                      energies it breaks down on or does not finish within the iteration
                                      cap, and every lower energy after the first such one, are solved by
                                      MIXED.  Same results as MIXED to the tolerance; a full scattering
                                      request -> MIXED.  Replaces the LU of fms_lu_scattering
                                      (fms/solvers.rs:366-484) where it converges */
} refeff_fms_solver;

/* ---------------------------------------------------------------------- */
/* host-side cluster preparation                                            */
/* ---------------------------------------------------------------------- */

/* FEFF athep: sort atoms in place by r^2 + (index+1)*1e-6.
 * Replaces refeff sort_atoms_by_radius, crates/refeff-core/src/fms/cluster.rs:91-107.
 * positions: [n][3] float (Angstrom); potentials: [n]; keys: [n] or NULL. */
int refeff_fms_sort_atoms_by_radius(int32_t n, float* positions, int32_t* potentials, double* keys);

/* FEFF yprep cluster prefix: centre on the first atom of central_potential,
 * sort (athep) and keep the prefix inside cluster_radius, truncated to
 * cluster_capacity.  Replaces refeff fms_yprep_cluster,
 * crates/refeff-core/src/fms/cluster.rs:10-84.  out_positions / out_potentials
 * must hold n entries; *out_count receives the kept prefix length. */
int refeff_fms_yprep_cluster(int32_t n, const float* positions, const int32_t* potentials,
                             int32_t central_potential, float cluster_radius, int32_t cluster_capacity,
                             float* out_positions, int32_t* out_potentials, int32_t* out_count,
                             int32_t* central_atom, int32_t* untruncated_count);

/* FEFF sortat: move one representative atom per potential into the prefix.
 * Replaces refeff sort_representative_atoms, crates/refeff-core/src/fms/cluster.rs:115-195.
 * representatives: [max_potential+1], -1 where a potential is absent. */
int refeff_fms_sort_representative_atoms(int32_t central_potential, int32_t max_potential, int32_t n,
                                         float* positions, int32_t* potentials, int32_t* representatives);

/* FEFF getkts state kets |atom,l,m,spin>.  Replaces refeff
 * construct_state_kets_with_limit, crates/refeff-core/src/state.rs:97-148.
 * Pass NULL arrays to query *state_count only.  atoms are zero-based. */
int refeff_fms_state_kets(int32_t spin_channels, int32_t atom_count, const int32_t* atom_potentials,
                          int32_t max_potential, const int32_t* raw_potential_lmax, int32_t global_lmax,
                          int32_t* state_count, int32_t* state_atom, int32_t* state_l, int32_t* state_m,
                          int32_t* state_spin, int32_t* representative_offsets);

/* FMS pair mean-square displacements of the Debye models, computed on the
 * current HIP device (one FP64 thread per ordered atom pair).  model is the
 * FEFF idwopt: 0 = correlated (quantum) Debye, FEFF sigms/corrfn; 3 = classical
 * Debye, FEFF sigcl/corrfn2; any other value returns REFEFF_FMS_UNSUPPORTED.
 * Replaces the idwopt 0 / 3 branches of fms_mean_square_displacements_with_metadata,
 * crates/refeff-engine/src/fms.rs:4345-4433, with fms_debye_pair_sigma
 * (:4849-4877) and quantum/classical_debye_waller_factor,
 * crates/refeff-core/src/debye/models.rs:138-436.
 *   temperature, debye_temperature: tk, thetad (K); average_norman_radius: rnrmav (bohr)
 *   positions: [atom_count][3] f32 cluster coordinates (Angstrom)
 *   atomic_numbers: [atom_count], 1..139
 *   sigsqr: [atom_count*atom_count] out; sigsqr(atom2, atom1) at atom2 + atom_count*atom1
 *     = f32(sig2g + sigma^2 of the path atom1 -> atom2 -> atom1), zero diagonal
 *   first_failure: [2] or NULL; on a pair failure, (atom1, atom2) of the first failing
 *     pair in refeff's order (atom2 outer, atom1 inner), else (-1, -1).
 * A Romberg integral that does not converge returns REFEFF_FMS_NO_CONVERGENCE;
 * coincident atoms (a zero-length leg) REFEFF_FMS_INVALID_ARGUMENT. */
int refeff_fms_debye_sigsqr(int32_t model, double temperature, double debye_temperature,
                            double average_norman_radius, int32_t atom_count, const float* positions,
                            const int32_t* atomic_numbers, float sig2g, float* sigsqr, int32_t* first_failure);

/* ---------------------------------------------------------------------- */
/* plan + batched energy sweep (the hot path)                               */
/* ---------------------------------------------------------------------- */

typedef struct refeff_fms_plan refeff_fms_plan;

/* Energy-independent inputs.  Mirrors FmsRealSpacePlanInput
 * (crates/refeff-core/src/fms/types.rs:110-157); the rotation (drix), xnlm and
 * spin-orbit tables are rebuilt inside the library from the geometry. */
typedef struct refeff_fms_plan_desc {
    int32_t lfms;                       /* 0: absorber potential only */
    int32_t minv;                       /* raw FEFF solver selector */
    int32_t spin_channels;              /* nsp: 1 or 2 */
    int32_t spin_selector;              /* ispin */
    int32_t atom_count;
    const float* positions;             /* [atom_count][3], absorber at index 0 (prepared cluster) */
    const int32_t* potentials;          /* [atom_count] */
    int32_t max_potential;              /* inclusive npot */
    int32_t global_lmax;                /* lx */
    const int32_t* raw_potential_lmax;  /* [max_potential+1] lmaxph before clamping */
    float direct_cutoff;                /* rdirec, Angstrom */
    const float* mean_square_displacements; /* [atom_count*atom_count], sigsqr(atom2,atom1) at atom2+atom_count*atom1; NULL = 0 */
    float convergence_tolerance;        /* toler1 */
    float zero_tolerance;               /* toler2 */
    int32_t full_scattering_requested;  /* save_gg_slice: also solve A^-1 G0 for every G0 column (forces LU) */
    int32_t solver;                     /* refeff_fms_solver */
    int32_t max_batch;                  /* energies per device batch; 0 = sized from free device memory */
    int32_t state_capacity;             /* istatx-style state-ket capacity (FmsRealSpacePlanInput::state_capacity,
                                           types.rs:126); 0 = no limit.  A larger state count fails with
                                           REFEFF_FMS_INVALID_ARGUMENT, variant REFEFF_FMS_E_STATE_CAPACITY_EXCEEDED */
    const uint8_t* calculated_l;        /* FEFF lcalc(l), [global_lmax+1] (types.rs:140); NULL = every l.  The
                                           iterative branches (GGBI/GGRM/GGGM/GGTF) leave the source columns of
                                           masked l at zero (internals/iterative.rs:138-143); the LU ignores it */
} refeff_fms_plan_desc;

typedef struct refeff_fms_plan_info {
    int32_t state_count;      /* N */
    int32_t padded_order;     /* N rounded up for the device tiles */
    int32_t channel_count;    /* (global_lmax+1)^2 * nsp */
    int32_t potential_count;  /* max_potential+1 */
    int32_t potential_start;  /* solved potential range (lfms) */
    int32_t potential_end;
    int32_t rhs_count;        /* right-hand sides per energy */
    int32_t solver;           /* effective refeff_fms_solver */
    int32_t batch;            /* energies per device batch */
    int32_t lu_block;         /* LU outer block = K of the trailing update (LU solvers) */
} refeff_fms_plan_info;

typedef struct refeff_fms_stats {
    double seconds_total;        /* wall time of the call */
    double seconds_assemble;     /* t-matrix + G0 + (1 - G0 t) build: device time on the first energy
                                    sub-batch's stream (all of it with one stream, the default);
                                    seconds_factor excludes it */
    double seconds_factor;       /* LU */
    double seconds_solve;        /* triangular solves + refinement */
    int32_t refinement_sweeps;   /* largest number of FP64 refinement sweeps over the energies */
    int32_t batches;
    double max_correction;       /* largest estimated relative error left in x after refinement:
                                    d_s * min(1, d_s / d_{s-1}) from the last two corrections
                                    |dx|/|x| (mixed solver) */
    int32_t krylov_energies;     /* BICGSTAB: energies BiCGSTAB solved */
    int32_t krylov_iterations;   /* BICGSTAB: their iterations, summed */
    int32_t lu_energies;         /* BICGSTAB: energies the mixed LU solved */
    int32_t fp64_fallback_energies; /* MIXED: energies the FP32 factors + FP64 refinement did not settle within
                                       8 sweeps, solved again by the plain FP64 LU (their status is that solve's) */
} refeff_fms_stats;

/* Typed failure detail.  The status codes above are the FmsError families; this
 * record carries the variant and its payload (crates/refeff-core/src/fms/types.rs:830-981,
 * crates/refeff-linalg/src/error.rs:1-42), so a Rust wrapper can rebuild the
 * exact FmsError value. */
typedef enum refeff_fms_error_variant {
    REFEFF_FMS_E_NONE = 0,
    REFEFF_FMS_E_INVALID_ANGULAR_LIMIT = 1,      /* InvalidAngularLimit{name, value, lx}: values = value, lx */
    REFEFF_FMS_E_EMPTY_CLUSTER = 2,              /* EmptyCluster */
    REFEFF_FMS_E_NON_FINITE_COORDINATE = 3,      /* NonFiniteCoordinate{atom, axis}: values = atom, axis */
    REFEFF_FMS_E_POTENTIAL_OUT_OF_RANGE = 4,     /* PotentialOutOfRange{potential, max_potential} */
    REFEFF_FMS_E_NON_FINITE_MSD = 5,             /* NonFiniteMeanSquareDisplacement */
    REFEFF_FMS_E_INVALID_DIRECT_CUTOFF = 6,      /* InvalidDirectCutoff */
    REFEFF_FMS_E_INVALID_TOLERANCE = 7,          /* InvalidTolerance{name, value}: name, real_value */
    REFEFF_FMS_E_INVALID_SPIN_CHANNEL_COUNT = 8, /* InvalidSpinChannelCount{value} */
    REFEFF_FMS_E_STATE_CAPACITY_EXCEEDED = 9,    /* StateCapacityExceeded{capacity} */
    REFEFF_FMS_E_MISSING_REPRESENTATIVE = 10,    /* MissingRepresentativePotential{potential} */
    REFEFF_FMS_E_FULL_SCATTERING_REQUIRES_LU = 11, /* FullScatteringRequiresLu{method}: values = minv */
    REFEFF_FMS_E_NON_FINITE_WAVE_NUMBER = 12,    /* NonFiniteWaveNumber */
    REFEFF_FMS_E_NON_FINITE_PHASE_SHIFT = 13,    /* NonFinitePhaseShift{spin, angular_momentum, potential} */
    REFEFF_FMS_E_SINGULAR_MATRIX = 14,           /* LinearAlgebra(SingularMatrix{pivot}): values = pivot (0-based) */
    REFEFF_FMS_E_ITERATIVE_BREAKDOWN = 15,       /* IterativeSolverBreakdown{solver, step}: name = solver */
    REFEFF_FMS_E_ITERATIVE_NO_CONVERGENCE = 16,  /* IterativeSolverNoConvergence{solver, restarts} */
    REFEFF_FMS_E_REFINEMENT_NO_CONVERGENCE = 17, /* no refeff counterpart: FP64 refinement left an estimated
                                                    relative error above 1e-9; values = sweeps, real_value = error */
    REFEFF_FMS_E_UNSUPPORTED = 18,               /* a branch this library does not provide */
    REFEFF_FMS_E_DEVICE = 19,                    /* HIP runtime failure, no device, or no memory */
    REFEFF_FMS_E_INVALID_ARGUMENT = 20           /* a malformed call (NULL pointer, negative count, ...) */
} refeff_fms_error_variant;

typedef struct refeff_fms_error {
    int32_t status;       /* refeff_fms_status family */
    int32_t variant;      /* refeff_fms_error_variant */
    int64_t values[3];    /* integer payload fields, in the order the variant lists them */
    double real_value;    /* floating-point payload (InvalidTolerance value, refinement error) */
    char name[24];        /* &'static str payload (tolerance or solver name, "lfms" ...), NUL-terminated */
    char message[160];    /* refeff's Display text for the variant */
} refeff_fms_error;

/* Detail of the last failing refeff_fms_plan_create on the calling thread
 * (REFEFF_FMS_E_NONE after a success). */
int refeff_fms_last_error(refeff_fms_error* out);

/* Detail of energy `energy` (index into the last solve call's energies) for
 * the last refeff_fms_solve_batch* call on this plan: the per-energy
 * Vec<Result<FmsRealSpaceEnergyResult, FmsError>> of fms_real_space_spectrum
 * (driver.rs:231-236).  REFEFF_FMS_E_NONE for an energy that succeeded. */
int refeff_fms_plan_energy_error(const refeff_fms_plan* plan, int32_t energy, refeff_fms_error* out);

/* Build a plan (uploads geometry to the current HIP device).  Mirrors
 * fms_real_space_plan, crates/refeff-core/src/fms/driver.rs:105-140. */
int refeff_fms_plan_create(const refeff_fms_plan_desc* desc, refeff_fms_plan** plan);
void refeff_fms_plan_destroy(refeff_fms_plan* plan);
int refeff_fms_plan_query(const refeff_fms_plan* plan, refeff_fms_plan_info* info);
/* No counterpart in refeff (its plans hold no device memory): give back the
 * plan's batch workspace (the next solve allocates it again), for callers that
 * keep many plans and solve them one after another. */
int refeff_fms_plan_release_workspace(refeff_fms_plan* plan);

/* Solve every energy point.  Replaces fms_real_space_spectrum,
 * crates/refeff-core/src/fms/driver.rs:231-236, and packs the scattering table
 * of fms_lu_scattering (solvers.rs:425-471): for each energy and each potential
 * ip in [potential_start, potential_end] the (ch x ch) block
 * scattering[e][ip] holds gg(row, col) at row + ch*col (zero elsewhere).
 *   wave_numbers: [ne][nsp] complex, Angstrom^-1 (before refeff narrows them, engine fms.rs:4972-4985)
 *   phase_shifts: [ne][nsp][2*global_lmax+1][max_potential+1] complex (engine fms.rs:5002-5054 layout)
 *   scattering:   [ne][max_potential+1][ch][ch] complex (output)
 *   status:       [ne] refeff_fms_status per energy (output; may be NULL)
 *   stats:        may be NULL
 * Host pointers. */
int refeff_fms_solve_batch(refeff_fms_plan* plan, int32_t energy_count, const double* wave_numbers,
                           const double* phase_shifts, double* scattering, int32_t* status,
                           refeff_fms_stats* stats);

/* Same as refeff_fms_solve_batch with device pointers (inputs already in HBM,
 * outputs left in HBM); status stays a host array.  stream: hipStream_t or NULL. */
int refeff_fms_solve_batch_device(refeff_fms_plan* plan, int32_t energy_count, const double* d_wave_numbers,
                                  const double* d_phase_shifts, double* d_scattering, int32_t* status,
                                  refeff_fms_stats* stats, void* stream);

/* No counterpart in refeff: for callers that cannot allocate device memory
 * (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  Same as refeff_fms_solve_batch_device, but the batch workspace
 * comes from a fixed per-device scratch of 256 MiB inside the library instead
 * of hipMalloc: the call allocates nothing.  Energies run in batches of as many
 * as the scratch holds (at most the plan's batch); a cluster whose single-energy
 * workspace exceeds 256 MiB (about N > 3300 for the mixed solver, N > 4000 for
 * FP64) fails with REFEFF_FMS_NO_MEMORY and refeff_fms_last_error describes it.
 * Mixed and FP64 LU plans only (REFEFF_FMS_UNSUPPORTED otherwise).  All calls on
 * one device share that scratch: calls within a process are serialised by a
 * per-device lock; the scratch keeps the last call's contents (every buffer is
 * rewritten before it is read). */
int refeff_fms_solve_batch_device_noalloc(refeff_fms_plan* plan, int32_t energy_count, const double* d_wave_numbers,
                                          const double* d_phase_shifts, double* d_scattering, int32_t* status,
                                          refeff_fms_stats* stats, void* stream);

/* Same as refeff_fms_solve_batch, and also returns full_scattering = A^-1 G0
 * (fms_lu_scattering with calculate_full_scattering, solvers.rs:473-477, as
 * engine fms.rs:3659-3665 collects it for save_gg_slice).  The plan must have
 * been created with full_scattering_requested.  Host pointers:
 *   full_scattering: [ne][N][N] complex, column-major per energy (row + N*col),
 *                    N = state_count. */
int refeff_fms_solve_batch_full(refeff_fms_plan* plan, int32_t energy_count, const double* wave_numbers,
                                const double* phase_shifts, double* scattering, double* full_scattering,
                                int32_t* status, refeff_fms_stats* stats);

/* ---------------------------------------------------------------------- */
/* autotuner and plan sets (many clusters in one graph launch)              */
/* ---------------------------------------------------------------------- */

/* No counterpart in refeff: faer picks its own blocking
 * (crates/refeff-linalg/src/lu.rs:385-397).  The LU variant of a plan:
 * outer block width (K of the trailing MFMA update), the row count from which
 * FP32 panels use tournament pivoting, the trailing-update kernel of the
 * mixed solver (REFEFF_FMS_UPDATE_*), and the most refinement sweeps a plan
 * set's graph launches (mixed solver; energies that have settled skip the
 * rest).  Mixed plans of padded order >= 1536 start with the split-fp16
 * update, smaller ones with the FP32 3M update; from padded order 4096 with
 * outer blocks of 512 (1024 from 8192) factored by recursive halving. */
enum {
    REFEFF_FMS_UPDATE_FP32 = 0,      /* FP32 MFMA, four real products per complex product */
    REFEFF_FMS_UPDATE_FP32_3M = 1,   /* FP32 MFMA, three real products (3M) */
    REFEFF_FMS_UPDATE_FP16X2 = 4     /* f16 MFMA on two-part fp16 splits with row / column scales (FP32-like) */
};
typedef struct refeff_fms_tuning {
    int32_t lu_block;
    int32_t calu_min_rows;
    int32_t trailing_update;
    int32_t graph_sweeps;
    double seconds_per_energy;  /* measured by refeff_fms_plan_tune (0 otherwise) */
    int32_t lu_split;           /* 1: each outer block factored by recursive halving (halves joined by
                                   K = half updates, down to 32-column panels), its far U12 rows
                                   solved the same way; 0: in 32-column steps */
} refeff_fms_tuning;

int refeff_fms_plan_get_tuning(const refeff_fms_plan* plan, refeff_fms_tuning* tuning);
int refeff_fms_plan_set_tuning(refeff_fms_plan* plan, const refeff_fms_tuning* tuning);
/* Time the LU variants on up to 8 of the given energies (host arrays, layout of
 * refeff_fms_solve_batch), keep the fastest on the plan and return it (graph
 * sweeps 8, the eager refinement's limit). */
int refeff_fms_plan_tune(refeff_fms_plan* plan, int32_t energy_count, const double* wave_numbers,
                         const double* phase_shifts, refeff_fms_tuning* best);

/* A plan set: one plan per cluster (descriptors as for refeff_fms_plan_create;
 * mixed or FP64 LU solvers), solved together.  Replaces the per-caller loop of
 * fms_real_space_plan / fms_real_space_spectrum over clusters
 * (crates/refeff-engine/src/fms.rs:1396-1458, 2154-2206, 2896-2931) with one
 * graph launch.  tune_cache: path of the autotuner's text cache (read at
 * creation, written by refeff_fms_plan_set_tune), or NULL. */
typedef struct refeff_fms_plan_set refeff_fms_plan_set;

typedef struct refeff_fms_set_stats {
    double seconds_total;
    double seconds_capture;     /* graph capture + instantiation (0 when the kept graph was replayed) */
    double seconds_graph;       /* the graph launch, to completion */
    int64_t graph_nodes;
    int32_t captured;           /* 1: this call captured a new graph */
    int32_t resolved_energies;  /* energies the refinement left unsettled, solved again by the FP64 LU */
    int32_t energies;
} refeff_fms_set_stats;

int refeff_fms_plan_set_create(const refeff_fms_plan_desc* descs, int32_t cluster_count, const char* tune_cache,
                               refeff_fms_plan_set** set);
void refeff_fms_plan_set_destroy(refeff_fms_plan_set* set);
/* The plan of one cluster (owned by the set; query, tune or solve it alone). */
int refeff_fms_plan_set_plan(refeff_fms_plan_set* set, int32_t cluster, refeff_fms_plan** plan);
/* Tune every cluster whose size the cache does not cover yet, on its own energies
 * (host arrays per cluster), once per size bucket; writes the cache. */
int refeff_fms_plan_set_tune(refeff_fms_plan_set* set, const int32_t* energy_counts, const double* const* wave_numbers,
                             const double* const* phase_shifts, int32_t* tuned);
/* Solve every cluster's energies: per cluster c, energy_counts[c] energies with
 * device arrays d_wave_numbers[c], d_phase_shifts[c] (layouts of
 * refeff_fms_solve_batch_device) and output d_scattering[c]; status: host
 * [sum of energy_counts], clusters in order.  The first call (and any call
 * with other pointers or counts) captures and instantiates the graph; later
 * calls replay it. */
int refeff_fms_plan_set_solve(refeff_fms_plan_set* set, const int32_t* energy_counts, const double* const* d_wave_numbers,
                              const double* const* d_phase_shifts, double* const* d_scattering, int32_t* status,
                              refeff_fms_set_stats* stats);
/* How refeff_fms_plan_set_solve runs the lanes: GRAPH (default) replays the
 * captured graph; EAGER launches directly, one host thread per lane, in the
 * same arenas, with the refinement's stopping rule on the host.  Results are
 * the same to the refinement tolerance; which is faster depends on the mix of
 * cluster sizes (launch overhead against idle refinement sweeps). */
enum { REFEFF_FMS_SET_GRAPH = 0, REFEFF_FMS_SET_EAGER = 1 };
int refeff_fms_plan_set_mode(refeff_fms_plan_set* set, int32_t mode);

/* ---------------------------------------------------------------------- */
/* BiCGSTAB preconditioners (REFEFF_FMS_SOLVER_BICGSTAB plans)              */
/* ---------------------------------------------------------------------- */

/* No counterpart in refeff: refeff's iterative branches are FEFF's one-shot
 * and restarted solvers without a preconditioner.  These change only the
 * iteration count of the converged BiCGSTAB path; convergence is judged by the
 * FP64 residual and the mixed LU stays the fallback, so results are the LU's
 * (crates/refeff-core/src/fms/solvers.rs:366-484) to the solver tolerance. */
typedef enum refeff_fms_precond {
    REFEFF_FMS_PRECOND_NONE = 0,
    REFEFF_FMS_PRECOND_BLOCK_JACOBI = 1, /* inverses of A on neighbour-shell atom groups (<= 64 states), FP32 */
    REFEFF_FMS_PRECOND_LEARNED = 2       /* M = (I + N_1)(I + N_2): N_h applies per-channel weights, predicted by a
                                            graph network over the atom blocks, to the neighbour blocks of G0 t */
} refeff_fms_precond;

/* The learned preconditioner: a small graph network over the atoms of one
 * cluster (edges: atom pairs closer than neighbor_cutoff), trained on the GPU
 * by this library's own forward / backward kernels on exact solves of the
 * mixed LU.  Global lmax <= 3 and one spin channel (REFEFF_FMS_UNSUPPORTED
 * otherwise).  Parameters start from a fixed function of seed. */
typedef struct refeff_fms_gnn refeff_fms_gnn;
int refeff_fms_gnn_create(const refeff_fms_plan* plan, float neighbor_cutoff, uint64_t seed, refeff_fms_gnn** gnn);
void refeff_fms_gnn_destroy(refeff_fms_gnn* gnn);
/* param_count: network parameters (float); edge_count: graph edges; samples: training energies held */
int refeff_fms_gnn_info(const refeff_fms_gnn* gnn, int32_t* param_count, int32_t* edge_count, int32_t* samples);
int refeff_fms_gnn_get_params(const refeff_fms_gnn* gnn, float* params);
int refeff_fms_gnn_set_params(refeff_fms_gnn* gnn, const float* params);
/* Solve energy_count energies (host arrays, layout of refeff_fms_solve_batch) with the
 * plan's mixed LU and keep each energy's neighbour blocks of G0 t, right-hand sides and
 * solutions as training samples (energies with a non-OK status are held but get no weight).
 * The plan must be built on the gnn's cluster. */
int refeff_fms_gnn_collect(refeff_fms_gnn* gnn, refeff_fms_plan* plan, int32_t energy_count, const double* wave_numbers,
                           const double* phase_shifts, int32_t* status);
/* Loss (mean over held samples and columns of log(1 + |M b - x|^2 / |x|^2)) and its gradient
 * [param_count] at the current parameters. */
int refeff_fms_gnn_loss_grad(refeff_fms_gnn* gnn, double* loss, float* grad);
/* epochs full-batch Adam steps (learning_rate); loss_history: [epochs] losses before each step, or NULL. */
int refeff_fms_gnn_train(refeff_fms_gnn* gnn, int32_t epochs, float learning_rate, double* loss_history);
/* Network output w of held sample `sample`: [edge_count][2 heads][16 channels] complex (float pairs),
 * edges in (source atom, destination atom) order. */
int refeff_fms_gnn_weights(refeff_fms_gnn* gnn, int32_t sample, float* weights);
/* y = M v for held sample `sample` (its blocks); v, y: [npad][rhs_count] complex doubles. */
int refeff_fms_gnn_apply(refeff_fms_gnn* gnn, int32_t sample, const double* v, double* y);
/* Select the BiCGSTAB preconditioner of a plan (gnn: required for LEARNED, built on the
 * plan's cluster; the plan keeps a reference, so destroy the plan first). */
int refeff_fms_plan_set_preconditioner(refeff_fms_plan* plan, int32_t kind, refeff_fms_gnn* gnn);

/* ---------------------------------------------------------------------- */
/* measurement hooks                                                        */
/* ---------------------------------------------------------------------- */

/* No counterpart in refeff: instrumentation for benchmarks.  When enabled,
 * every kernel launch of the plan is bracketed by HIP events on the plan's
 * stream; refeff_fms_plan_profile_read sums them per kernel class. */
enum refeff_fms_kernel_class {
    REFEFF_FMS_K_TMATRIX = 0,
    REFEFF_FMS_K_BUILD = 1,     /* G0 + (1 - G0 t) assembly */
    REFEFF_FMS_K_PANEL = 2,     /* LU panel factorisation */
    REFEFF_FMS_K_SWAP = 3,      /* row interchanges + U12 triangular solve */
    REFEFF_FMS_K_GEMM = 4,      /* LU trailing update (K = outer block width) and every update on the
                                   split-fp16 path (MFMA) */
    REFEFF_FMS_K_SOLVE = 5,     /* forward/backward substitution + diagonal-block inverses */
    REFEFF_FMS_K_RESIDUAL = 6,  /* FP64 residual */
    REFEFF_FMS_K_OTHER = 7,
    REFEFF_FMS_K_GEMM_SMALL = 8, /* narrower LU updates inside an outer block (FP32 MFMA) */
    REFEFF_FMS_K_COUNT = 9
};

typedef struct refeff_fms_profile {
    double milliseconds[REFEFF_FMS_K_COUNT];
    double flops[REFEFF_FMS_K_COUNT];       /* algorithmic real flops */
    double bytes[REFEFF_FMS_K_COUNT];       /* algorithmic HBM bytes */
    int64_t launches[REFEFF_FMS_K_COUNT];
} refeff_fms_profile;

/* enable != 0 starts (and clears) the per-launch event record: 1 brackets
 * every launch, any other value is a bitmask (1 << class) of the kernel
 * classes to bracket (the others run without events); 0 stops recording. */
int refeff_fms_plan_profile(refeff_fms_plan* plan, int32_t enable);
int refeff_fms_plan_profile_read(refeff_fms_plan* plan, refeff_fms_profile* out);

/* ---------------------------------------------------------------------- */
/* MKGTR absorber trace (FP64 on the device)                               */
/* ---------------------------------------------------------------------- */

/* Orbital momentum per transition slot and the orientation-averaged bcoef
 * matrix with the ispin fold.  Replaces transition_b_matrix for ipol=0
 * (crates/refeff-core/src/angular/transition.rs:9-167, 431-482).
 * bmat: (2lmax+1) x 2 x 8 x (2lmax+1) x 2 x 8 complex, column-major. */
int refeff_fms_transition_bmat_averaged(int32_t ihole, int32_t multipole, int32_t lmax, int32_t spin,
                                        int32_t spin_channels, int32_t* lind, double* bmat);

/* FEFF bcoef for any polarization: transition_b_matrix,
 * crates/refeff-core/src/angular/transition.rs:9-482.  ipol = 0 is the
 * averaged matrix; otherwise ptz holds the polarization tensor as 3 x 3
 * complex (re, im pairs) with ptz[p+1][p'+1], angks is the spin-axis angle.
 * trace_orbital folds the m-diagonal onto m = 0 within equal l and spin; spin
 * and spin_channels select the ispin fold.  Layout of lind and bmat as above. */
int refeff_fms_transition_bmat(int32_t ihole, int32_t multipole, int32_t lmax, int32_t ipol, const double* ptz,
                               int32_t trace_orbital, int32_t spin, int32_t spin_channels, double angks,
                               int32_t* lind, double* bmat);

/* gtr(ie) = sum gg bmat rkk rkk.  Replaces mkgtr_green_trace,
 * crates/refeff-core/src/fms/mkgtr.rs:790-953, keeping gg in FP64.
 *   gg:  [ne][nch][nch] complex, (row, col) at row + nch*col
 *   rkk: [ne][8][nsp] complex;  gtr: [ne] complex (output). Host pointers. */
int refeff_fms_mkgtr_trace(int32_t energy_count, int32_t channel_count, int32_t spin_channels, int32_t lmax,
                           const int32_t* lind, const double* bmat, const double* gg, const double* rkk,
                           double* gtr);

/* ---------------------------------------------------------------------- */
/* module stages over a work directory                                      */
/* ---------------------------------------------------------------------- */

/* Reads fms.inp, phase.bin, geom.dat, global.inp (optionally .dimensions.dat)
 * and writes gg.bin, gg.dat and log3.dat, plus gg_slice.bin and gg_diag.bin
 * when fms.inp sets save_gg_slice.  Replaces run_fms_in_dir,
 * crates/refeff-engine/src/fms.rs:462-539, on its ordinary source path
 * (build_fms_source_outputs, :3549-3695) with the global SIG2 damping
 * (idwopt < 0), the correlated and classical Debye models (0, 3) or sig2.dat
 * (4); the other Debye-Waller models, reciprocal-space FMS (reciprocal.inp
 * ispace 0) and Hubbard sources return REFEFF_FMS_UNSUPPORTED.  Without source files a cached
 * gg.bin / gg.dat (and gg_slice.bin / gg_diag.bin) is re-emitted
 * canonically.  solver: a refeff_fms_solver.  *files_written receives the
 * file count; message receives a context string on failure. */
int refeff_fms_run_in_dir(const char* work_dir, int32_t solver, int32_t* files_written, char* message,
                          size_t message_len);

/* Reads gg.bin, fms.inp, global.inp, phase.bin and writes fms.bin and gtr.dat
 * (the ordinary trace: orientation averaged for ipol = 0, else with
 * global.inp's polarization tensor and spin angle) and appends the MKGTR
 * lines to log3.dat.  Replaces run_mkgtr_in_dir,
 * crates/refeff-engine/src/fms.rs:559-631 with build_mkgtr_ordinary_outputs
 * (:5528-5608).  With eels.inp asking for ELNES, fms.bin holds one spectrum
 * per Cartesian polarization selector (:5610-5677) and gtr.dat the first.
 * NRIXS traces return REFEFF_FMS_UNSUPPORTED. */
int refeff_mkgtr_run_in_dir(const char* work_dir, int32_t* files_written, char* message, size_t message_len);

/* LDOS gtrNN.bin on the RHORRP energy grid that the LDOS stage hands over.
 * Replaces write_ldos_gtr_bin_source_grid_handoff and
 * has_supported_ldos_gtr_bin_source_grid_handoff,
 * crates/refeff-engine/src/fms.rs:1632-1687 (with :1759-1906, 2049-2093,
 * 2254-2434).  Reads ldos.inp, fms.inp, phase.bin, geom.dat, global.inp and an
 * optional .dimensions.dat from work_dir; for every potential c solves the
 * cluster around c's first atom under the ldos.inp controls with
 * k = wave_numbers_bohr[e][c] / a0 and writes gtrNN.bin (FEFF fmsdos trace).
 * ldos.inp's lmaxph is capped by phase_l_count - 1 and .dimensions.dat's lx.
 *   wave_numbers_bohr: [energy_count][wave_number_columns] complex, 1/bohr
 *   phase_shifts: [energy_count][phase_l_count][phase_potential_count] complex, by l
 * *supported: 1 where refeff produces the files (0: another LDOS path, nothing
 * written); dry_run != 0 stops there.  *files_written: gtrNN.bin files written.
 * rfms2 <= 0 writes zero traces without a solve. */
int refeff_ldos_gtr_source_grid_run_in_dir(const char* work_dir, int32_t solver, int32_t energy_count,
                                           const double* wave_numbers_bohr, int32_t wave_number_columns,
                                           int32_t phase_l_count, int32_t phase_potential_count,
                                           const double* phase_shifts, int32_t dry_run, int32_t* supported,
                                           int32_t* files_written, char* message, size_t message_len);

/* POT self-consistency FMS traces gtr(l, iph) on the POT energy grid.  Replaces
 * the real-space branch of build_pot_scf_fms_source_grid_handoff,
 * crates/refeff-engine/src/fms.rs:2650-2965 (with :3247-3275, 5118-5175 and
 * refeff-io screen_dat.rs:645-690).  geom.dat is read from work_dir; the pot.inp
 * controls come in as lfms1, rfms1 (cluster radius; direct cutoff 2 rfms1),
 * nph and lmaxsc[nph+1].  lfms1 != 0: one cluster around potential 0 (sortat),
 * every potential's block; otherwise one cluster per potential around its first
 * atom.  No damping.  rfms1 <= 0 gives zero traces without a solve; otherwise a
 * reciprocal.inp with ispace 0 (k-space FMS) returns REFEFF_FMS_UNSUPPORTED.
 *   energies: [energy_count] complex Hartree
 *   reference_energies: [energy_count][reference_count] complex; column nph sets k
 *   phase_shifts: [energy_count][phase_l_count][nph+1] complex, indexed by l
 *   traces: [energy_count][angular_count][nph+1] complex out,
 *     sum_m gg(l^2 + m, l^2 + m) exp(2 i delta_l) / (2 l + 1) */
int refeff_fms_pot_scf_traces(const char* work_dir, int32_t solver, int32_t lfms1, double rfms1, int32_t nph,
                              const int32_t* lmaxsc, int32_t energy_count, const double* energies,
                              int32_t reference_count, const double* reference_energies, int32_t phase_l_count,
                              const double* phase_shifts, int32_t angular_count, double* traces, char* message,
                              size_t message_len);

/* SCREEN's FMS cluster Green's function traces for the absorbing potential.
 * Replaces build_screen_fms_source_grid_handoff_with_potential_count,
 * crates/refeff-engine/src/fms.rs:2464-2648 (both the phase.bin and the
 * pot.bin variants, :2436-2462).  fms.inp and geom.dat are read from work_dir;
 * when either is missing *produced is 0 and nothing else is written (refeff's
 * None).  cluster_radius and direct_cutoff replace fms.inp's rfms2 and rdirec.
 * The cluster is the sortat cluster around potential 0, damped as fms.inp's
 * idwopt says (Debye models 0 and 3 take atomic_numbers[potential_count] and
 * average_norman_radius from phase.bin or pot.bin; other models ignore them, and
 * atomic_numbers may be NULL), solved with lfms 0 and fms.inp's tolerances.
 *   wave_numbers_bohr: [energy_count] complex, 1/bohr
 *   phase_shifts: [energy_count][phase_l_count][potential_count] complex, by l
 *   traces: [energy_count][angular_count] complex out, the absorber's
 *     sum_m gg(l^2 + m, l^2 + m) exp(2 i delta_l) / (2 l + 1)
 *   *potential_index: the absorber's potential; *produced: 1 */
int refeff_fms_screen_traces(const char* work_dir, int32_t solver, int32_t potential_count,
                             const int32_t* atomic_numbers, double average_norman_radius, int32_t energy_count,
                             const double* wave_numbers_bohr, int32_t phase_l_count, const double* phase_shifts,
                             int32_t angular_count, double cluster_radius, double direct_cutoff, double* traces,
                             int32_t* potential_index, int32_t* produced, char* message, size_t message_len);

/* LDOS gtrNN.bin from the FMS source handoffs.  Replaces
 * write_ldos_gtr_bin_source_handoffs, crates/refeff-engine/src/fms.rs:1226-1247
 * (with :1689-1760, 2095-2245): reads ldos.inp, fms.inp, phase.bin, geom.dat and
 * global.inp; for every potential c solves the cluster centred on c's first atom
 * with the ldos.inp FMS controls and writes gtrNN.bin with FEFF's fmsdos trace of
 * potential c.  Writes nothing (status OK, *files_written = 0) where refeff
 * leaves the work to its other LDOS paths; keeps a gtrNN.bin that already
 * matches.  solver: as refeff_fms_run_in_dir. */
int refeff_ldos_gtr_run_in_dir(const char* work_dir, int32_t solver, int32_t* files_written, char* message,
                               size_t message_len);

/* ---------------------------------------------------------------------- */
/* text codecs of the FEFF handoff files (crates/refeff-io)                 */
/* ---------------------------------------------------------------------- */

/* Text outputs: out receives a NUL-terminated string when out_len allows;
 * *needed (when given) receives strlen + 1; out = NULL only queries the size.
 * Codec failures return REFEFF_FMS_IO_ERROR (malformed text) or
 * REFEFF_FMS_INVALID_ARGUMENT (values that cannot be written);
 * refeff_io_last_error() then describes the failure on the calling thread. */
const char* refeff_io_last_error(void);

#define REFEFF_IO_MAX_POTENTIALS 32

/* PAD field of npack characters.  Replaces encode_f64 / decode_f64,
 * crates/refeff-io/src/pad.rs:33-120 (FEFF padlib, byte-identical output).
 * out must hold npack + 1 bytes. */
int refeff_io_pad_encode(double value, int32_t npack, char* out, size_t out_len);
int refeff_io_pad_decode(const char* field, int32_t npack, double* value);

/* PAD lines ('!' reals, '$' complex pairs).  Replaces encode_reals /
 * encode_complex / decode_reals / decode_complex, crates/refeff-io/src/pad.rs:122-220.
 * values: count reals, or count complex (re, im) pairs. */
int refeff_io_pad_encode_array(const double* values, size_t count, int32_t complex_values, int32_t npack,
                               char* out, size_t out_len, size_t* needed);
int refeff_io_pad_decode_array(const char* text, int32_t complex_values, int32_t npack, size_t expected,
                               double* values, size_t* decoded);

/* Fortran Ew.d with a [0.1, 1) mantissa (gg.dat's E20.10).  Replaces
 * write_fortran_zero_scaled_exp, crates/refeff-io/src/format.rs:164-209. */
int refeff_io_fortran_exp(double value, int32_t width, int32_t precision, char* out, size_t out_len);

/* fms.inp.  Mirrors FmsInput (crates/refeff-io/src/fms_input.rs:19-60);
 * parse replaces FmsInput::parse_str (:112, :274-410), string replaces
 * fms_input_string (:164-189). */
typedef struct refeff_io_fms_inp {
    int32_t mfms, idwopt, minv;
    double rfms2, rdirec, toler1, toler2;
    double tk, thetad, sig2g;
    int32_t lmaxph_count;
    int32_t lmaxph[REFEFF_IO_MAX_POTENTIALS];
    int32_t decomposition_channels;
    int32_t save_gg_slice;
    int32_t do_fms;
} refeff_io_fms_inp;
int refeff_io_fms_inp_parse(const char* text, refeff_io_fms_inp* out);
int refeff_io_fms_inp_string(const refeff_io_fms_inp* in, char* out, size_t out_len, size_t* needed);

/* phase.bin.  Mirrors PhaseBinData (crates/refeff-io/src/phase_bin/types.rs:85-127);
 * parse replaces parse_phase_bin (parse.rs:18-123), string replaces
 * phase_bin_string (render.rs:16-95).  Arrays (complex, C order):
 *   energy_grid [ne], reference_energy [ne][nsp],
 *   phase_shifts: per potential in order, [ne][2 lmax + 1][nsp], concatenated,
 *   transition_moments [ne][nq][ntr][nsp].
 * Parse with NULL arrays to read only the header. */
typedef struct refeff_io_phase_bin_info {
    int32_t spin_count, energy_count, main_energy_count, auxiliary_energy_count;
    int32_t potential_count, ihole, fermi_index, pad_width;
    int32_t final_state_count, transition_count, q_count;
    int32_t header_fields;  /* 8, 10 or 11 integers on the header line (parse only) */
    double scalars[3];      /* rnrmav, xmu, edge */
    int32_t lmax[REFEFF_IO_MAX_POTENTIALS];
    int32_t atomic_number[REFEFF_IO_MAX_POTENTIALS];
    char label[REFEFF_IO_MAX_POTENTIALS][8];
} refeff_io_phase_bin_info;
int refeff_io_phase_bin_parse(const char* text, refeff_io_phase_bin_info* info, double* energy_grid,
                              double* reference_energy, double* phase_shifts, double* transition_moments);
int refeff_io_phase_bin_string(const refeff_io_phase_bin_info* info, const double* energy_grid,
                               const double* reference_energy, const double* phase_shifts,
                               const double* transition_moments, char* out, size_t out_len, size_t* needed);

/* geom.dat.  Replaces GeomDat::parse_str / geom_dat_string,
 * crates/refeff-io/src/structure_output.rs:204-222, 470-530.  Parse with NULL
 * arrays to read nat / nph; model_atoms holds nph + 1 entries, xyz 3 nat. */
int refeff_io_geom_dat_parse(const char* text, int32_t* nat, int32_t* nph, int32_t* model_atoms, int32_t* index,
                             double* xyz, int32_t* iph, int32_t* boundary);
int refeff_io_geom_dat_string(int32_t nat, int32_t nph, const int32_t* model_atoms, const int32_t* index,
                              const double* xyz, const int32_t* iph, const int32_t* boundary, char* out,
                              size_t out_len, size_t* needed);

/* global.inp (the controls the FMS and MKGTR stages read).  Replaces
 * GlobalInput::parse_str, crates/refeff-io/src/global_input.rs:376-470. */
typedef struct refeff_io_global_inp {
    int32_t nabs, iphabs;
    double rclabs;
    int32_t ipol, ispin, le2;
    double elpty, angks;
    int32_t l2lp, do_nrixs, ldecmx, lj;
    double evec[3], xivec[3], spvec[3];
    double polarization_tensor[3][6];
    double evnorm, xivnorm, spvnorm;
    int32_t nq, imdff, qaverage, mixdff;
} refeff_io_global_inp;
int refeff_io_global_inp_parse(const char* text, refeff_io_global_inp* out);

/* gg.dat / gg.bin sections (E20.10 complex matrices).  Replaces parse_gg_dat /
 * gg_dat_string, crates/refeff-io/src/gg_dat.rs:89-170, 410-640.  values:
 * each section's (row, col) matrix row-major, sections concatenated.  Parse
 * with NULL arrays to read the counts. */
int refeff_io_gg_dat_parse(const char* text, int32_t* section_count, size_t* value_count, int32_t* section_numbers,
                           int32_t* rows, int32_t* cols, double* values);
int refeff_io_gg_dat_string(int32_t section_count, const int32_t* section_numbers, const int32_t* rows,
                            const int32_t* cols, const double* values, char* out, size_t out_len, size_t* needed);

/* fms.bin.  Replaces fms_bin_string / parse_fms_bin,
 * crates/refeff-io/src/fms_bin.rs:77-196.  declared_spectrum_count = -1
 * writes the five-integer legacy header; spectra [spectrum][energy] complex.
 * counts: ne, ne1, ne3, nph, npadx. */
int refeff_io_fms_bin_string(double cluster_radius, int32_t energy_count, int32_t main_energy_count,
                             int32_t auxiliary_energy_count, int32_t highest_potential_index, int32_t pad_width,
                             int32_t declared_spectrum_count, int32_t spectrum_count, const double* spectra,
                             char* out, size_t out_len, size_t* needed);
int refeff_io_fms_bin_parse(const char* text, double* cluster_radius, int32_t* counts,
                            int32_t* declared_spectrum_count, int32_t* spectrum_count, double* spectra);

/* gtr.dat (energy, trace) rows.  Replaces gtr_dat_string / parse_gtr_dat,
 * crates/refeff-io/src/gtr_dat.rs. */
int refeff_io_gtr_dat_string(int32_t count, const double* energy, const double* trace, char* out, size_t out_len,
                             size_t* needed);
int refeff_io_gtr_dat_parse(const char* text, int32_t* count, double* energy, double* trace);

/* gg_slice.bin / gg_diag.bin (binary: *needed = byte count, no terminator).
 * Replace rhorrp_gg_slice_bin_bytes / parse_rhorrp_gg_slice_bin and
 * rhorrp_gg_diag_bin_bytes / parse_rhorrp_gg_diag_bin,
 * crates/refeff-io/src/rhorrp_gg_bin.rs:98-260.  Values are complex float:
 * slice [ne][row][col] (dims = ne, ldim, istate); diag [ne][atom][row][col]
 * (dims = ne, inclus, ldim).  Parsing accepts either byte order; pass
 * values = NULL to read dims only. */
int refeff_io_gg_slice_bin_bytes(int32_t energy_count, int32_t row_count, int32_t column_count, const float* values,
                                 char* out, size_t out_len, size_t* needed);
int refeff_io_gg_slice_bin_parse(const char* bytes, size_t len, int32_t* dims, float* values);
int refeff_io_gg_diag_bin_bytes(int32_t energy_count, int32_t atom_count, int32_t dim, const float* values, char* out,
                                size_t out_len, size_t* needed);
int refeff_io_gg_diag_bin_parse(const char* bytes, size_t len, int32_t* dims, float* values);

/* gtrNN.bin, the FMS -> LDOS trace handoff (binary: *needed = byte count).
 * Replace gtr_bin_bytes / parse_gtr_bin, crates/refeff-io/src/gtr_bin.rs:86-190.
 * header: [6] = ne, ne1, ne3, nph, ifms, angular channel count; values:
 * [ne][nph+1][angular] complex, written as f32.  Parsing accepts either byte
 * order; pass values = NULL to read the header only. */
int refeff_io_gtr_bin_bytes(const int32_t* header, const double* values, char* out, size_t out_len, size_t* needed);
int refeff_io_gtr_bin_parse(const char* bytes, size_t len, int32_t* header, double* values);

/* ldos.inp.  Replaces LdosInput::parse_str / ldos_input_string,
 * crates/refeff-io/src/ldos_input.rs:59-300. */
typedef struct refeff_io_ldos_inp {
    int32_t mldos, lfms2, ixc, ispin, minv, neldos, iscfxc;
    double rfms2, emin, emax, eimag, rgrd;
    double rdirec, toler1, toler2;
    int32_t lmaxph_count;
    int32_t lmaxph[64];
    int32_t ldostype;
} refeff_io_ldos_inp;
int refeff_io_ldos_inp_parse(const char* text, refeff_io_ldos_inp* out);
int refeff_io_ldos_inp_string(const refeff_io_ldos_inp* in, char* out, size_t out_len, size_t* needed);

/* sig2.dat pair damping values (idwopt = 4): the third field of the first
 * pair_count data rows (pairs atom1 < atom2, atom1 outer), each finite and >= 0.
 * Replaces fms_sig2_dat_pair_values, crates/refeff-engine/src/fms.rs:4778-4847.
 * values: [pair_count]. */
int refeff_io_sig2_dat_parse(const char* text, size_t pair_count, double* values);

/* ---------------------------------------------------------------------- */
/* ff2x: XANES rows of xmu.dat from the FMS trace                           */
/* ---------------------------------------------------------------------- */

/* The array core of ff2x's regular XANES output (ispec +-1 and +-2): the
 * combined trace (FMS trace times S02 plus an optional path sum,
 * ff2x.rs:2569-2603, 2677-2702), FEFF's xscorr contour correction at zero
 * electronic temperature (ff2x_xscorr, ff2x.rs:4437-4641, with its vrcorr
 * shift of the contour), the corrected total and atomic components
 * (ff2x_xanes_corrected_components, :3186-3263), the output energy grid
 * (ff2x_output_energy_grid, :4333-4400) and the xmu.dat columns with the
 * xsedge (or absolute) normalization and the sig_gk damping
 * (ff2x_xanes_xmu_dat_from_components, :3309-3419).  Inputs are xsect.dat's
 * arrays in Hartree (XsectFf2xHandoff, crates/refeff-io/src/xsect_dat.rs:73-110),
 * feff.bin's edge energy and the output momentum grid (bohr^-1), and ff2x.inp's
 * controls.  Branches not provided (REFEFF_FMS_UNSUPPORTED): a positive
 * electronic temperature (ff2x_thermal_xscorr), mbconv > 0, and the Lorentzian
 * vicorr / gamma-channel convolution of a run with no listed paths
 * (ff2x_xanes_apply_vicorr_convolution). */
typedef struct refeff_ff2x_xanes_input {
    int32_t energy_count;             /* ne: the horizontal grid, then the contour */
    int32_t main_energy_count;        /* ne1 */
    int32_t fermi_index;              /* ik0, 0-based, < ne1 */
    const double* energy;             /* [ne] complex (re, im) energy grid, Hartree (emxs) */
    const double* omega;              /* [ne] photon energy, Hartree (omega) */
    const double* cross_section;      /* [ne] complex atomic cross section (xsec) */
    const double* background;         /* [ne] normalized atomic background (xsnorm) */
    double amplitude_reduction;       /* S02 */
    const double* fms_trace;          /* [ne] complex FMS trace (an fms.bin spectrum), NULL = none */
    const double* path_total;         /* [ne] complex path sum, NULL = no paths */
    int32_t listed_paths;             /* list.dat's path count */
    int32_t ispec, absolu, mbconv;
    int32_t atomic_background;        /* 1: the polarization carries the atomic background */
    int32_t gamma_channel;            /* ff2x.inp i_gamma_ch */
    double core_hole_width;           /* Hartree (gamach) */
    double vrcorr_ev, vicorr_ev;      /* ff2x.inp corrections, eV */
    double sig_gk;                    /* ff2x.inp sig_gk, Angstrom */
    double electronic_temperature_ev;
    double edge_energy;               /* feff.bin's edge energy, Hartree */
    int32_t output_count;             /* = ne1 */
    const double* output_momentum;    /* [ne1] bohr^-1 */
} refeff_ff2x_xanes_input;

/* omega_ev, e_ev, k, mu, mu0, chi: [output_count] (the xmu.dat columns, k in
 * Angstrom^-1); normalization: the divisor applied (1 with absolu). */
int refeff_ff2x_xanes_xmu(const refeff_ff2x_xanes_input* in, double* omega_ev, double* e_ev, double* k, double* mu,
                          double* mu0, double* chi, double* normalization, char* message, size_t message_len);

/* Library version string (static storage). */
const char* refeff_fms_version(void);

#ifdef __cplusplus
}
#endif

#endif /* REFEFF_FMS_H */
