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
    REFEFF_FMS_SOLVER_AUTO = 0,    /* minv=0 -> MIXED, minv=1 -> GGBI, other minv -> UNSUPPORTED */
    REFEFF_FMS_SOLVER_MIXED = 1,   /* complex FP32 MFMA LU + FP64 iterative refinement */
    REFEFF_FMS_SOLVER_FP64 = 2,    /* plain complex FP64 MFMA LU */
    REFEFF_FMS_SOLVER_GGBI = 3     /* refeff minv=1 one-shot BiCGStab (FP32, as refeff) */
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
    int32_t full_scattering_requested;  /* save_gg_slice: not provided (REFEFF_FMS_UNSUPPORTED) */
    int32_t solver;                     /* refeff_fms_solver */
    int32_t max_batch;                  /* energies per device batch; 0 = sized from free device memory */
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
} refeff_fms_plan_info;

typedef struct refeff_fms_stats {
    double seconds_total;        /* wall time of the call */
    double seconds_assemble;     /* t-matrix + G0 + (1 - G0 t) build */
    double seconds_factor;       /* LU */
    double seconds_solve;        /* triangular solves + refinement */
    int32_t refinement_sweeps;   /* largest number of FP64 refinement sweeps over the energies */
    int32_t batches;
    double max_correction;       /* largest estimated relative error left in x after refinement:
                                    d_s * min(1, d_s / d_{s-1}) from the last two corrections
                                    |dx|/|x| (mixed solver) */
} refeff_fms_stats;

/* Build a plan (uploads geometry to the current HIP device).  Mirrors
 * fms_real_space_plan, crates/refeff-core/src/fms/driver.rs:105-140. */
int refeff_fms_plan_create(const refeff_fms_plan_desc* desc, refeff_fms_plan** plan);
void refeff_fms_plan_destroy(refeff_fms_plan* plan);
int refeff_fms_plan_query(const refeff_fms_plan* plan, refeff_fms_plan_info* info);

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
    REFEFF_FMS_K_GEMM = 4,      /* LU trailing update, K = outer block width (MFMA) */
    REFEFF_FMS_K_SOLVE = 5,     /* forward/backward substitution + diagonal-block inverses */
    REFEFF_FMS_K_RESIDUAL = 6,  /* FP64 residual */
    REFEFF_FMS_K_OTHER = 7,
    REFEFF_FMS_K_GEMM_SMALL = 8, /* panel-width (K = 32) LU updates inside an outer block */
    REFEFF_FMS_K_COUNT = 9
};

typedef struct refeff_fms_profile {
    double milliseconds[REFEFF_FMS_K_COUNT];
    double flops[REFEFF_FMS_K_COUNT];       /* algorithmic real flops */
    double bytes[REFEFF_FMS_K_COUNT];       /* algorithmic HBM bytes */
    int64_t launches[REFEFF_FMS_K_COUNT];
} refeff_fms_profile;

/* enable != 0 starts (and clears) the per-launch event record. */
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
 * and writes gg.bin, gg.dat and log3.dat.  Replaces run_fms_in_dir,
 * crates/refeff-engine/src/fms.rs:462-539.  *files_written receives the file
 * count; message receives a context string on failure. */
int refeff_fms_run_in_dir(const char* work_dir, int32_t solver, int32_t* files_written, char* message,
                          size_t message_len);

/* Reads gg.bin, fms.inp, global.inp, phase.bin and writes fms.bin and gtr.dat.
 * Replaces run_mkgtr_in_dir, crates/refeff-engine/src/fms.rs:559-631. */
int refeff_mkgtr_run_in_dir(const char* work_dir, int32_t* files_written, char* message, size_t message_len);

/* Library version string (static storage). */
const char* refeff_fms_version(void);

#ifdef __cplusplus
}
#endif

#endif /* REFEFF_FMS_H */
