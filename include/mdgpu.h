/* mdgpu.h - C ABI of the MI355X (gfx950) energy-and-force evaluator.
 *
 * This is the thin boundary between the C++ host (md::System, which keeps the
 * reference's class API) and the hand-written HIP kernels.  Plain C types and
 * pointers only; no HIP or torch types appear in a signature.
 *
 * The reference has no FFI of its own: its evaluation is the private member
 * chain System::init() -> calculate_energies()/calculate_forces()
 * (src/System/init.cpp:10-81, src/System/System.cpp:9-39).  Each entry point
 * below names the reference routine(s) whose work it replaces.
 *
 * Conventions
 *   - every function returns an int status (MDGPU_OK == 0); no exception or
 *     abort crosses the boundary; mdgpu_last_error() gives the message;
 *   - host buffers belong to the caller, device buffers to the context;
 *   - one context = one device and its own HIP streams (a main stream, and a
 *     side stream for the PME chain); calls are synchronous on return unless
 *     stated; a context is not thread-safe (use one per thread);
 *   - atom indices are 0-based; coordinates and forces are 3N interleaved
 *     x,y,z doubles (the reference's layout, AmberTopology.h:252, System.h:175);
 *   - energies in kcal/mol, forces in kcal/mol/A, lengths in A.
 */
#ifndef MDGPU_H
#define MDGPU_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    MDGPU_OK = 0,
    MDGPU_ERR_ARG = 1,         /* invalid argument or inconsistent system   */
    MDGPU_ERR_HIP = 2,         /* a HIP runtime call failed                 */
    MDGPU_ERR_STATE = 3,       /* call out of order (e.g. eval before upload) */
    MDGPU_ERR_OVERFLOW = 4,    /* fixed-point force accumulator overflowed  */
    MDGPU_ERR_UNSUPPORTED = 5  /* a parameter combination is not supported  */
};

/* Semantics.  REF reproduces the reference's printed energies and final
 * forces, quirks included (empty exclusion lists, plain-Coulomb forces,
 * O(N^2) "EE" energy; src/System/System.cpp:9-39).  AMBER evaluates the
 * physics the reference's code is written for: exclusions and scaled 1-4
 * pairs honoured, smooth PME reciprocal space, forces = -grad E. */
enum { MDGPU_SEM_REF = 0, MDGPU_SEM_AMBER = 1 };

/* Precision forms.  FP64: all arithmetic and accumulation in double.
 * MIXED: pair, PME and bonded arithmetic in FP32, forces accumulated in
 * 64-bit fixed point (bit-reproducible), FP64 for the excluded/1-4 Ewald
 * corrections, self energy, energy/virial reductions and coordinate wrapping. */
enum { MDGPU_PREC_FP64 = 0, MDGPU_PREC_MIXED = 1 };

typedef struct mdgpu_ctx mdgpu_ctx;

/* One system.  Arrays are read during mdgpu_upload_system only. */
typedef struct {
    int32_t n_atoms;
    int32_t n_types;              /* LJ atom types T                          */
    const double* charge;         /* [n_atoms] prmtop CHARGE / 18.2206899283247 */
    const int32_t* type;          /* [n_atoms] 0-based LJ type                */
    const int32_t* nb_index;      /* [T*T] 1-based LJ table index, 0 = none   */
    int32_t n_lj;                 /* length of the LJ tables                  */
    const double* lj_a;           /* [n_lj] A (r^-12) coefficients            */
    const double* lj_b;           /* [n_lj] B (r^-6)                          */
    const double* lj14_a;         /* [n_lj] 1-4 A                             */
    const double* lj14_b;         /* [n_lj] 1-4 B                             */

    int32_t n_bonds;              /* harmonic bonds E = k (r - r0)^2          */
    const int32_t* bond_atoms;    /* [2*n]                                    */
    const int32_t* bond_type;     /* [n] 0-based into bond_k/bond_r0          */
    int32_t n_bond_types;
    const double* bond_k;
    const double* bond_r0;

    int32_t n_ubs;                /* Urey-Bradley E = k (r13 - r0)^2          */
    const int32_t* ub_atoms;      /* [2*n]                                    */
    const int32_t* ub_type;
    int32_t n_ub_types;
    const double* ub_k;
    const double* ub_r0;

    int32_t n_angles;             /* harmonic angles E = k (theta - theta0)^2 */
    const int32_t* angle_atoms;   /* [3*n]                                    */
    const int32_t* angle_type;
    int32_t n_angle_types;
    const double* angle_k;
    const double* angle_theta0;

    int32_t n_dihedrals;          /* cosine E = k (1 + cos(n phi - phase))    */
    const int32_t* dihedral_atoms;/* [4*n]                                    */
    const int32_t* dihedral_type;
    int32_t n_dihedral_types;
    const double* dihedral_k;
    const double* dihedral_n;     /* periodicity (double, as in the prmtop)   */
    const double* dihedral_phase;

    int32_t n_impropers;          /* CHARMM harmonic E = k (psi - psi0)^2     */
    const int32_t* improper_atoms;/* [4*n]                                    */
    const int32_t* improper_type;
    int32_t n_improper_types;
    const double* improper_k;
    const double* improper_psi0;

    int32_t n_cmaps;              /* CMAP bicubic patches on (phi, psi)       */
    const int32_t* cmap_atoms;    /* [5*n]                                    */
    const int32_t* cmap_set;      /* [n] 1-based parameter set                */
    int32_t cmap_n_sets;          /* number of sets in cmap_coeff / cmap_grid */
    int32_t cmap_resolution;      /* grid points per angle (same for all sets)*/
    const double* cmap_coeff;     /* [n_sets*res*res*16] patch coefficients   */
    const double* cmap_grid;      /* [n_sets*res*res] grid energies (phi-major)*/

    /* Nonbonded exclusions (AMBER semantics; REF passes none).  excl_* is the
     * full symmetric per-atom CSR list (both directions, 1-4 partners
     * included) used to mask pair tiles.  The explicit pair lists drive the
     * FP64 correction kernel. */
    const int32_t* excl_start;    /* [n_atoms+1] or NULL                      */
    const int32_t* excl_list;     /* [excl_start[n_atoms]] partner indices    */
    int32_t n_excl_pairs;         /* pure exclusions (1-2, 1-3), i < j        */
    const int32_t* excl_pairs;    /* [2*n]                                    */
    int32_t n_pairs14;            /* 1-4 pairs from proper dihedrals          */
    const int32_t* pairs14;       /* [2*n]                                    */
    const double* scee14;         /* [n] Coulomb 1-4 divisor                  */
} mdgpu_system_desc;

/* Evaluation parameters (the reference hard-codes these in System::init(),
 * src/System/init.cpp:18-23,53; its setters live in System.h:54-62). */
typedef struct {
    int32_t semantics;            /* MDGPU_SEM_*                              */
    int32_t precision;            /* MDGPU_PREC_*                             */
    double lj_cutoff;             /* A                                        */
    double ee_cutoff;             /* A                                        */
    double lj_switch;             /* switch radius (A); <= 0: force-shifted LJ */
    double ewald_alpha;           /* 1/A                                      */
    int32_t pme_grid[3];          /* 0,0,0 -> no reciprocal space             */
    int32_t pme_order;            /* B-spline order (4 supported)             */
} mdgpu_params;

/* Per-term energies (kcal/mol).  Names follow the reference's members
 * (System.h:176-188) plus the AMBER terms it does not compute. */
typedef struct {
    double bond;                  /* bond_energies_                           */
    double angle;                 /* angle_energy_                            */
    double dihedral;              /* dihedral_energy_                         */
    double urey_bradley;          /* urey_bradley_energy_                     */
    double improper;              /* improper_energy_                         */
    double cmap;                  /* CMAP_energy_                             */
    double vdw;                   /* LJ_energy_pairlist_ (switched / shifted) */
    double ee_all;                /* EE_energy_: plain Coulomb, all i<j < rc  */
    double ee_pairlist;           /* EE_energy_pairlist_                      */
    double ee_rf;                 /* EE_energy_pairlist_cutoff_               */
    double ewald_direct;          /* EE_energy_pairlist_ewald_direct_         */
    double ewald_self;            /* EE_energy_pairlist_ewald_self_           */
    double ewald_recip;           /* AMBER: smooth PME reciprocal energy      */
    double ewald_excl;            /* AMBER: -k qq erf(a r)/r, pure exclusions */
    double vdw14;                 /* AMBER: 1-4 LJ                            */
    double ee14;                  /* AMBER: 1-4 Coulomb (scaled) minus erf part */
    double total;                 /* energy whose -gradient the forces are    */
} mdgpu_energies;

/* Kernel ids for mdgpu_get_timing(). */
enum {
    MDGPU_K_SORT = 0,       /* spatial sort + cluster build                 */
    MDGPU_K_PAIRLIST = 1,   /* cluster-pair search                          */
    MDGPU_K_NONBONDED = 2,  /* LJ + Coulomb/Ewald pair kernel (+ bonded)    */
    MDGPU_K_BONDED = 3,     /* bonded terms: run as blocks of the pair
                               launch and timed there (stays 0)             */
    MDGPU_K_CORRECT = 4,    /* FP64 excluded / 1-4 corrections: likewise    */
    MDGPU_K_SPREAD = 5,     /* PME charge spreading                         */
    MDGPU_K_FFT = 6,        /* PME r2c + c2c passes (forward + inverse)     */
    MDGPU_K_CONV = 7,       /* PME influence function + energy + virial     */
    MDGPU_K_GATHER = 8,     /* PME force gather                             */
    MDGPU_K_REDUCE = 9,     /* energy reductions + force conversion         */
    MDGPU_K_COUNT = 10
};

/* Create a context on HIP device `device`. */
int mdgpu_create(int device, mdgpu_ctx** out);
/* Destroy a context and free its device memory (NULL is accepted). */
int mdgpu_destroy(mdgpu_ctx* ctx);
/* Message for the last non-zero status returned on this context. */
const char* mdgpu_last_error(const mdgpu_ctx* ctx);

/* Upload the topology (replaces System::build_nonbonded_cache,
 * NonBondedUtils.cpp:134-160, and the term-list reads of the energy/force
 * passes).  May be called again to replace the system. */
int mdgpu_upload_system(mdgpu_ctx* ctx, const mdgpu_system_desc* sys);

/* Set the evaluation parameters (replaces the setters and hard-coded values
 * of System::init, init.cpp:15-23,53-55, System.h:54-62). */
int mdgpu_set_params(mdgpu_ctx* ctx, const mdgpu_params* p);

/* One energy + force evaluation from host buffers.  Replaces the evaluation
 * part of System::init(): build_lj_pairlist (NonBondedUtils.cpp:185-282),
 * calculate_energies (System.cpp:9-23) and calculate_forces (System.cpp:25-39).
 *   coords     [3N] host, raw (unwrapped) coordinates;
 *   box        [3]  orthorhombic box lengths (all > 0);
 *   out_e      energies (may be NULL);
 *   out_forces [3N] host (may be NULL);
 *   out_virial [9]  host, reciprocal-space virial (may be NULL). */
int mdgpu_eval(mdgpu_ctx* ctx, const double* coords, const double box[3],
               mdgpu_energies* out_e, double* out_forces, double* out_virial);

/* Same evaluation with device-resident coordinates and forces (3N doubles in
 * device memory), for callers that keep state on the GPU.  The kernels run
 * on the context's stream; the call returns after they finish. */
int mdgpu_eval_device(mdgpu_ctx* ctx, const double* d_coords, const double box[3],
                      mdgpu_energies* out_e, double* d_forces, double* out_virial);

/* Parameter gradients of a weighted energy sum (see mdgpu_adjoint).  Host
 * arrays owned by the caller; a NULL pointer skips that block, every other
 * array is overwritten.  Lengths follow the uploaded mdgpu_system_desc:
 * charge [n_atoms]; lj_* and lj14_* [n_lj] (per LJ table entry, i.e. per
 * type pair); bond_* [n_bond_types]; ub_* [n_ub_types]; angle_*
 * [n_angle_types]; dihedral_* [n_dihedral_types]; improper_*
 * [n_improper_types]; cmap_coeff [cmap_n_sets * res * res * 16], the
 * gradient with respect to the patch coefficients (mdhost_cmap_grid_grad maps
 * it to the grid values through the transpose of the patch table). */
typedef struct {
    double* charge;
    double* lj_a;
    double* lj_b;
    double* lj14_a;
    double* lj14_b;
    double* bond_k;
    double* bond_r0;
    double* ub_k;
    double* ub_r0;
    double* angle_k;
    double* angle_theta0;
    double* dihedral_k;
    double* dihedral_phase;
    double* improper_k;
    double* improper_psi0;
    double* cmap_coeff;
} mdgpu_param_grads;

/* Adjoint: sum_t w_t dE_t/dtheta at one configuration, for the parameters
 * of mdgpu_param_grads, in AMBER semantics (the reference has no gradient
 * with respect to parameters; this is new).  weights holds one weight per
 * energy term (fields of mdgpu_energies); weights->total is added to every
 * term of the total.  The pair list and every pair contribution are
 * recomputed (nothing is kept from an evaluation).  Arithmetic follows the
 * context's precision form; sums are accumulated in 64-bit fixed point
 * (deterministic), except the LJ sums of the FP64 form (FP64 atomics).
 *   coords [3N] host coordinates; box [3]. */
int mdgpu_adjoint(mdgpu_ctx* ctx, const double* coords, const double box[3], const mdgpu_energies* weights,
                  mdgpu_param_grads* out);

/* ---- Batches: many independent systems in one graph launch ------------
 * The reference evaluates one system per process; this is new.  A batch owns
 * one context per system (own topology, parameters and fixed box) and
 * evaluates all of them with ONE hipGraph launch: every system's kernel
 * sequence is captured once, spread over a few concurrent streams, and
 * replayed.  Coordinates and forces of all systems are concatenated in the
 * order the systems were added (3 * sum N doubles).  The first graph
 * evaluation runs every system once on its own (sizing the pair lists), then
 * captures the graph, once per batch.  A later call whose buffer addresses
 * differ from the captured ones, or in which a system's pair list outgrows its
 * capacity (that system is redone on its own), leaves the graph stale, and the
 * batch then issues the same launches on the same streams directly, without a
 * graph (see mdgpu_batch_graph_state). */
typedef struct mdgpu_batch mdgpu_batch;

int mdgpu_batch_create(int device, mdgpu_batch** out);
int mdgpu_batch_destroy(mdgpu_batch* b);
const char* mdgpu_batch_last_error(const mdgpu_batch* b);
/* Add a system (copied to the device now); *index receives its position. */
int mdgpu_batch_add(mdgpu_batch* b, const mdgpu_system_desc* sys, const mdgpu_params* p, const double box[3],
                    int* index);
/* Evaluate every system.  d_coords / d_forces: device buffers of 3 * sum N
 * doubles (d_forces may be NULL); out_e: [n_systems] host energies.  The
 * graph is (re)captured when a buffer address changes. */
int mdgpu_batch_eval_device(mdgpu_batch* b, const double* d_coords, mdgpu_energies* out_e, double* d_forces);
/* Same from host buffers (copied in and out around the graph launch). */
int mdgpu_batch_eval(mdgpu_batch* b, const double* coords, mdgpu_energies* out_e, double* out_forces);
/* Execution settings: mode 0 = one graph launch (default), 1 = one
 * synchronous evaluation per system in turn (the serial baseline), 2 = the
 * graph's launches on the graph's streams issued directly, without a graph;
 * streams = concurrent streams (1..8); split = waves per i-cluster list in
 * the pair kernel (1..8); fft_lines = cap on FFT lines per block (0 =
 * automatic).  A changed split or fft_lines applies to every system (it
 * replaces the tuner's per-class choice).  streams, split and fft_lines are
 * fixed once the graph exists (MDGPU_ERR_STATE); mode can always change. */
int mdgpu_batch_configure(mdgpu_batch* b, int mode, int streams, int split, int fft_lines);
/* Autotuner, before the first graph evaluation.  Streams become
 * min(8, systems).  The systems are ordered by atom count and cut into up to
 * four size classes of equal count; for each class in turn (largest first)
 * every pair-tile split (1, 2, 4, 8) and FFT block shape (16, 8, 4 lines) is
 * timed over `reps` passes of the whole batch with the graph's launches and
 * streams issued directly (mode 2), the other classes at their current
 * choice, and the fastest is kept for that class.  Unless the batch is in
 * the serial mode, it then captures the graph (once, on these buffers), times
 * `reps` graph launches against `reps` direct passes and keeps the faster form
 * as the batch's mode (0 or 2; mdgpu_batch_mode).  Returns the streams, the
 * largest class's settings and the best pass time (ms);
 * mdgpu_batch_system_config gives every system's. */
int mdgpu_batch_tune(mdgpu_batch* b, const double* d_coords, double* d_forces, int reps, int* streams, int* split,
                     int* fft_lines, double* ms);
/* The batch's execution mode (0 graph, 1 serial, 2 direct launches). */
int mdgpu_batch_mode(const mdgpu_batch* b, int* mode);
/* Pair-tile split and FFT lines of system `index` (no counterpart in the
 * reference). */
int mdgpu_batch_system_config(const mdgpu_batch* b, int index, int* split, int* fft_lines);
/* 0: no graph yet; 1: the captured graph is in use; 2: the graph went stale
 * and the batch issues its launches directly (mode 2's passes). */
int mdgpu_batch_graph_state(const mdgpu_batch* b, int* state);
/* Number of systems and total atoms. */
int mdgpu_batch_size(const mdgpu_batch* b, int* n_systems, int64_t* n_atoms);

/* Per-kernel timing with HIP events on the context stream (for roofline
 * reporting).  mdgpu_set_timing(ctx, mode) resets the totals; mode 1 makes
 * each later evaluation add the event-measured duration of every kernel
 * group, mode 2 only that of the pair kernel (two events per evaluation),
 * mode 3 that of the pair kernel in one evaluation of every 16 (events in a
 * replayed graph cost time, so the others replay a graph without them),
 * mode 0 turns timing off. */
int mdgpu_set_timing(mdgpu_ctx* ctx, int mode);
int mdgpu_get_timing(mdgpu_ctx* ctx, int kernel_id, double* total_ms, int64_t* launches);
/* Work counts of the last evaluation: [0] atom pairs within cutoff (set when
 * counting is enabled), [1] cluster-pair tiles, [2] clusters, [3] atoms. */
int mdgpu_get_counts(mdgpu_ctx* ctx, int64_t counts[4]);
/* Replay of a captured evaluation (default on; no counterpart in the
 * reference).  When two consecutive successful evaluations of a context have
 * the same coordinate and force pointers, box, parameters and buffer sizes,
 * the next one is captured into a hipGraph and replayed from then on, which
 * removes the per-launch host work (timing events, if on, are captured as
 * event nodes).  Any change of those or of the timing mode, or a redo (list
 * over capacity), makes the context evaluate with direct launches again, for
 * good.  Contexts of a batch never capture their own
 * graph. */
int mdgpu_set_graph(mdgpu_ctx* ctx, int enable);
/* 0: no graph captured yet, 1: evaluations replay a captured graph,
 * 2: direct launches for good (-1 for a null context). */
int mdgpu_graph_state(const mdgpu_ctx* ctx);
/* Autotuner of one context (no counterpart in the reference, whose loops
 * have no launch shapes to choose).  Before the first graph evaluation
 * (graph state 0, else MDGPU_ERR_STATE), every pair-kernel split (1, 2, 4, 8
 * waves per i-cluster list) and FFT block shape (16, 8, 4 lines per block)
 * is timed over `reps` evaluations of d_coords in `box` with direct launches
 * (host clock around each synchronous evaluation; two rounds over the
 * candidates, each candidate's time the lower of its two medians), and the
 * fastest pair is kept for the context's later evaluations.  Mixed-form
 * forces are the same bits for every choice (fixed-point sums); energies and
 * FP64 forces are summed in a fixed order for a given choice and differ
 * between choices by the rounding of their partial sums (the split decides
 * which wave sums which pairs; FP32 partial energies in the mixed form, about
 * 1e-9 of the total).  The graph setting and timing mode and totals are as before
 * afterwards; the list-build counters include the tuning evaluations.
 * Returns the chosen split and FFT lines (0 = 16) and the median time (ms). */
int mdgpu_tune(mdgpu_ctx* ctx, const double* d_coords, const double box[3], double* d_forces, int reps, int* split,
               int* fft_lines, double* ms);
/* Launch shapes of one context, set by hand or read back (no counterpart in
 * the reference): split = waves per i-cluster list in the pair kernel
 * (1..8, default 4), fft_lines = cap on FFT lines per block (0 = 16, else a
 * power of two <= 16; default 8).  A change is refused (MDGPU_ERR_STATE)
 * while the context replays a captured graph (graph state 1). */
int mdgpu_set_shapes(mdgpu_ctx* ctx, int split, int fft_lines);
int mdgpu_get_shapes(const mdgpu_ctx* ctx, int* split, int* fft_lines);
/* Enable (1) or disable (0) counting of pairs within the cutoff. */
int mdgpu_set_counting(mdgpu_ctx* ctx, int enable);

/* Full virial and pressure (no counterpart in the reference, which forms no
 * virial).  With mdgpu_set_virial(ctx, 1)
 * each evaluation also forms the virial of the pair, bonded and correction
 * forces, W_ab = sum over interactions of (relative position)_a (force)_b,
 * and mdgpu_get_virial returns it plus the reciprocal-space virial of the
 * PME pass: W = -dE/d(strain), so the pressure tensor is (sum m v v^T + W) / V.
 * The evaluation then runs the pair kernel's statistics instance (which also
 * counts pairs).  mdgpu_get_virial: MDGPU_ERR_STATE when the last evaluation
 * did not form it. */
int mdgpu_set_virial(mdgpu_ctx* ctx, int enable);
int mdgpu_get_virial(const mdgpu_ctx* ctx, double virial[9]);

/* Verlet list with a GPU-resident rebuild criterion.  Replaces the
 * reference's skin handling: System::set_lj_skin_ / set_ee_skin
 * (src/System/NonBondedUtils.cpp:11-47), setup_neighbor_rebuild_threshold_lj
 * (:65-75), copy_ref_coords_lj (:77-85) and ljpairs_needs_rebuild (:87-100),
 * which the reference declares but never calls.
 * skin > 0: the cluster-pair list is built with radius max(cutoffs) + skin
 * and kept across evaluations; each evaluation first measures, on the GPU,
 * every atom's minimum-image displacement since the build and rebuilds the
 * list only when some atom has moved more than skin / 2 (the reference's
 * test dr^2 > (skin/2)^2).  The pair kernel reads the kept list pruned to the
 * tiles with a pair within cutoff + 0.2 A, pruned again once an atom has
 * moved 0.1 A since the last prune.  Kept clusters follow their atoms; energies and
 * forces are those of a fresh list (every pair within the cutoff is in the
 * list).  A new box, system, parameter set or skin, a redo and an adjoint
 * call force the next evaluation to build.  The decision is made on the
 * device, so captured evaluations replay it.  skin = 0 (default): the list
 * is built at every evaluation.  Negative or non-finite: MDGPU_ERR_ARG. */
int mdgpu_set_verlet(mdgpu_ctx* ctx, double skin);
/* Triclinic cells (no counterpart in the reference, whose periodic boundary
 * is orthorhombic only, PBC.cpp:72-81).  angles = alpha, beta, gamma in
 * degrees, as on the box line of an rst7 file; they hold for every later
 * evaluation, minimisation and dynamics run of the context, whose box
 * arguments then give the edge lengths |a|, |b|, |c|.  The cell is
 * a = (|a|, 0, 0), b = |b| (cos gamma, sin gamma, 0) and c in the half-space
 * z > 0 with the given angles; coordinates are taken in that frame.  All
 * three angles exactly 90 (the default) is the orthorhombic cell.  The
 * cell is evaluated in its reduced basis: b and c lose whole multiples of a
 * and b until |c_y| <= b_y / 2 and |b_x|, |c_x| <= a_x / 2 (the same lattice,
 * so the same periodic system; an already reduced cell is kept as given), and
 * the PME grid runs along the reduced edges.  A triclinic cell needs AMBER
 * semantics, every distance between opposite faces of the reduced cell at
 * least twice the cutoff and at least the cutoff plus the Verlet skin
 * (checked at each evaluation: MDGPU_ERR_UNSUPPORTED), and is not available
 * to mdgpu_adjoint.  Angles outside (0, 180) or that form no cell:
 * MDGPU_ERR_ARG.  A change of angles forces the next list build. */
int mdgpu_set_box_angles(mdgpu_ctx* ctx, const double angles[3]);
int mdgpu_get_box_angles(const mdgpu_ctx* ctx, double angles[3]);
/* Which cluster search the context's last list build ran (no counterpart in
 * the reference): 0 the minimum-image search, 1 the general per-shift search,
 * which runs when a cluster could be within the cutoff of a second image of
 * another (the box or the cell's face distances are small for cutoff + skin),
 * and for good once a cluster was too wide for the minimum-image search; -1
 * for a null context. */
int mdgpu_get_search_mode(const mdgpu_ctx* ctx);
/* List builds and successful evaluations since the system upload. */
int mdgpu_get_list_builds(const mdgpu_ctx* ctx, int64_t* builds, int64_t* evals);
/* Prunes of the kept list since the system upload, Verlet mode only (every
 * build is followed by a prune; between builds the list is pruned again once
 * some atom has moved 0.1 A since the last prune).  No counterpart in the
 * reference, which keeps no list. */
int mdgpu_get_list_prunes(const mdgpu_ctx* ctx, int64_t* prunes);
/* ljpairs_needs_rebuild (NonBondedUtils.cpp:87-100) on the GPU for device
 * coordinates d_coords (3N doubles): *needs = 1 when some atom's squared
 * minimum-image displacement from the positions of the last list build, in
 * that build's box, exceeds (skin/2)^2; also 1 when no list is kept (skin 0,
 * none built yet, or a build pending).  Changes nothing. */
int mdgpu_pairs_need_rebuild(mdgpu_ctx* ctx, const double* d_coords, int* needs);

/* Allocation-free form of mdgpu_pairs_need_rebuild.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code under
 * HIP graph capture, serving loops that keep their own reference positions),
 * and it allocates nothing and takes no context.  On `device`, it enqueues on
 * `stream` (a hipStream_t; NULL: the null stream) the test of
 * ljpairs_needs_rebuild (NonBondedUtils.cpp:87-100), with the arithmetic of
 * the context's query, so the decision is the same bit for bit:
 *   d_coords, d_coords_ref  device, 3 * n_atoms doubles: the current and the
 *                           reference positions (e.g. those of the last build);
 *   box                     host, orthorhombic edge lengths;
 *   d_needs                 device int32: 1 when some atom's squared
 *                           minimum-image displacement exceeds (skin / 2)^2;
 *   d_max_d2                device double (or NULL): that largest squared
 *                           displacement.
 * The outputs are written when the stream reaches the work; nothing is
 * synchronised.  Scratch: a fixed device array per GPU (one partial maximum
 * per 256 atoms and a counter, 32 KiB), held until the process ends, so
 * n_atoms is at most MDGPU_NOALLOC_MAX_ATOMS; larger, negative or NULL
 * arguments, a non-positive box edge, a negative skin or a bad device give
 * MDGPU_ERR_ARG and enqueue nothing.  Calls that overlap on one GPU (other
 * streams or threads) share that array and race. */
#define MDGPU_NOALLOC_MAX_ATOMS 1048576
int mdgpu_pairs_need_rebuild_noalloc(int device, const double* d_coords, const double* d_coords_ref, int32_t n_atoms,
                                     const double box[3], double skin, void* stream, int32_t* d_needs,
                                     double* d_max_d2);

/* Molecular dynamics: velocity Verlet, NVE, every step on the GPU.  No
 * counterpart in the reference, whose README describes integrating Newton's
 * equations but whose driver stops after one evaluation (apps/main.cpp:23-28).
 * Units: A, A/ps, ps, amu, kcal/mol (a = 418.4 F/m A/ps^2).
 * mdgpu_md_set_masses: masses [n_atoms] (host, amu, positive).
 * mdgpu_md_run: nsteps steps from the device state (d_x, d_v, d_f: 3N
 * doubles each; on entry d_f must hold the forces at d_x, e.g. from
 * mdgpu_eval_device or a previous run; on return all three hold the final
 * state).  The evaluation settings of the context apply (precision, Verlet
 * list, virial).  trace (host, [nsteps] or NULL) receives each step's
 * potential energy (the evaluation's total), kinetic energy after the step
 * and, with the virial on, the pressure; last
 * (or NULL) the final step's energies.  A step that needs a larger list
 * restarts the run from the entry state; a step that fails (e.g. a force
 * beyond the fixed-point range) leaves the entry state and returns the error. */
typedef struct {
    double potential;
    double kinetic;
    double pressure;   /* bar, (2 KE + trace W) / 3V; NaN unless mdgpu_set_virial is on */
} mdgpu_md_energy;
int mdgpu_md_set_masses(mdgpu_ctx* ctx, const double* masses);

/* Energy minimisation by L-BFGS on the GPU.  The reference's counterpart is
 * the sys.minimize() left commented out in apps/main.cpp:28, whose solver its
 * comments name as L-BFGS (include/System/System.h:84,113); it has no body.
 * d_x (device, 3N doubles) is minimised in place with the context's
 * evaluation settings.  Iterations stop when the RMS force per coordinate is
 * <= rms_force_tol (kcal/mol/A), after max_iter iterations, or when the
 * backtracking line search finds no decrease.  No atom moves more than
 * max_step A per iteration; history = L-BFGS pairs kept (1..16).  opts NULL:
 * {500, 1e-2, 0.2, 8}. */
typedef struct {
    int32_t max_iter;
    double rms_force_tol;
    double max_step;
    int32_t history;
} mdgpu_min_opts;
typedef struct {
    int32_t iterations;
    int32_t evaluations;
    int32_t converged;     /* 1: the RMS force reached the tolerance        */
    double energy;         /* total energy at the returned coordinates     */
    double rms_force;      /* RMS force per coordinate there (kcal/mol/A)  */
    double max_force;      /* largest atom force there                     */
} mdgpu_min_result;
int mdgpu_minimize(mdgpu_ctx* ctx, double* d_x, const double box[3], const mdgpu_min_opts* opts,
                   mdgpu_min_result* result);
/* The same for host coordinates x (3N doubles, updated in place). */
int mdgpu_minimize_host(mdgpu_ctx* ctx, double* x, const double box[3], const mdgpu_min_opts* opts,
                        mdgpu_min_result* result);
int mdgpu_md_run(mdgpu_ctx* ctx, double* d_x, double* d_v, double* d_f, const double box[3], double dt,
                 int32_t nsteps, mdgpu_md_energy* trace, mdgpu_energies* last);

#ifdef __cplusplus
}
#endif

#endif /* MDGPU_H */
