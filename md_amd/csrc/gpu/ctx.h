// Device-side state of one mdgpu context and the kernel launchers.
#ifndef MD_AMD_GPU_CTX_H
#define MD_AMD_GPU_CTX_H

#include "common.h"
#include "mdgpu.h"

#include <algorithm>
#include <string>
#include <vector>

namespace mdg {

constexpr int kFlags = 8;

template <typename T>
struct DevBuf {
    T* p = nullptr;
    size_t n = 0;
};

// Static topology arrays on the device (uploaded once per system).
struct Topo {
    int n_atoms = 0, n_types = 0;          // n_types excludes the padding type
    double* charge = nullptr;              // [N]
    int* type = nullptr;                   // [N]
    // pair table over (T+1)^2 types, the last type being the padding type:
    // x = A, y = B, z = U(rc), w = G(rc) (the force-shift constants)
    double4* lj_d = nullptr;
    float4* lj_f = nullptr;
    unsigned char* lj_listed = nullptr;    // (T+1)^2
    int* nb_index = nullptr;               // [T*T] 1-based LJ table index (0 = none)
    double2* lj14 = nullptr;               // [n_lj] (A14, B14)
    // bonded terms: per-term atoms and per-type parameters
    int n_bonds = 0, n_ubs = 0, n_angles = 0, n_dihedrals = 0, n_impropers = 0, n_cmaps = 0;
    int2* bond_atoms = nullptr;  int* bond_type = nullptr;  double2* bond_par = nullptr;   // (k, r0)
    int2* ub_atoms = nullptr;    int* ub_type = nullptr;    double2* ub_par = nullptr;
    int4* angle_atoms = nullptr; int* angle_type = nullptr; double2* angle_par = nullptr; // (k, theta0)
    int4* dih_atoms = nullptr;   int* dih_type = nullptr;   double4* dih_par = nullptr;   // (k, n, phase)
    int4* imp_atoms = nullptr;   int* imp_type = nullptr;   double2* imp_par = nullptr;   // (k, psi0)
    int* cmap_atoms = nullptr;   int* cmap_set = nullptr;                                // [5n], [n]
    double* cmap_coeff = nullptr;
    int cmap_res = 0, cmap_n_sets = 0;
    // exclusions
    int* excl_start = nullptr;  int* excl_list = nullptr; int has_excl = 0;
    int n_excl_pairs = 0; int2* excl_pairs = nullptr;
    int n_pairs14 = 0; int2* pairs14 = nullptr; double* scee14 = nullptr;
    double sum_q2 = 0.0;                   // for the self energy
    // parameter counts and the pair-table index map, for the adjoint
    int n_lj = 0, n_bond_types = 0, n_ub_types = 0, n_angle_types = 0, n_dih_types = 0, n_imp_types = 0;
    int* lj_pidx = nullptr;                // (T+1)^2: LJ table entry of a type pair, -1 if none
};

// Fixed-point accumulators of the adjoint (one allocation; offsets in
// elements).  Each block is the gradient of one energy term before its
// weight is applied.
struct AdjointBuf {
    size_t cap = 0;                        // elements allocated
    unsigned long long* acc = nullptr;
    double* pot_rec = nullptr;             // [N] reciprocal potential (FP64, plain stores)
    size_t cap_pot = 0;
    // FP64 form: LJ derivative sums in FP64 (a pair at a fraction of an
    // angstrom, possible for atoms without LJ, exceeds any fixed-point range
    // fine enough for the FP64 bar); [2 n_lj]: dE/dA then dE/dB
    double* lj_d = nullptr;
    size_t cap_lj_d = 0;
    size_t q_dir = 0, q_excl = 0, q_14 = 0;          // [N] each
    size_t lj_a = 0, lj_b = 0;                       // [2 n_lj] each: (high, remainder) word pairs (mixed form)
    size_t lj14_a = 0, lj14_b = 0;                   // [n_lj] each
    size_t bonded = 0;                               // [2 x (bond, UB, angle, dihedral, improper types)]
    size_t cmap = 0;                                 // [sets x res^2 x 16]
    size_t flag = 0;                                 // overflow flag (as an element)
    size_t total = 0;
};
constexpr double kAdjScaleQ = 68719476736.0;        // 2^36: potentials
constexpr double kAdjScaleLJ = 17592186044416.0;    // 2^44: 1-4 LJ and CMAP coefficient sums
constexpr double kAdjScaleLJhi = 65536.0;           // 2^16: LJ sums of the mixed form, high word
constexpr double kAdjScaleLJlo = 281474976710656.0; // 2^48: their remainder word
constexpr double kAdjScaleB = 1099511627776.0;      // 2^40: bonded constants and equilibria

// Verlet-list state of a context on the device.  k_disp folds every atom's
// squared minimum-image displacement since the last list build into max_d2
// (the bits of a non-negative double order like the double); its last block
// decides, publishes the decision in `build` for the sort and search kernels
// of the same evaluation, and clears max_d2, ticket and force.
struct VerletState {
    unsigned long long max_d2;
    unsigned int ticket;   // k_disp blocks finished
    unsigned int build;    // 1: this evaluation builds the list
    unsigned int force;    // set by the host: the next evaluation builds
    unsigned int query;    // decision of the last query (mdgpu_pairs_need_rebuild)
    unsigned long long max_p2;  // largest squared displacement since the last prune
    unsigned int prune;    // 1: this evaluation prunes the kept list again
    unsigned int pad;
};

// The pruned list keeps the tiles with a pair within cutoff + kPruneBuffer and
// is pruned again once some atom has moved more than kPruneBuffer / 2 since
// the last prune (two atoms then come at most kPruneBuffer closer).
constexpr double kPruneBuffer = 0.2;

// atoms per column at most (the in-LDS column sort's capacity, and the
// stride of the per-column key buckets)
constexpr int kColBucket = 2048;

// Work arrays that depend on the atom count / box.
struct Work {
    size_t cap_atoms = 0;
    double* coords = nullptr;        // [3N] input (owned copy for host-pointer evals)
    double* wrapped = nullptr;       // [3N] wrapped into the box
    int* col_of = nullptr;           // [N] column of each atom
    int* col_count = nullptr;        // [ncol] atoms per column (zero between evaluations)
    float* block_emax = nullptr;     // [sort blocks] largest excluded-pair distance per sort block
    int* col_cstart = nullptr;       // [ncol + 1] cluster offsets
    // atoms bucketed by column as (z key << 32 | atom): column c holds its
    // atoms at [c * kColBucket, c * kColBucket + col_count[c])
    unsigned long long* col_key = nullptr;
    size_t cap_colkey = 0;           // entries of col_key
    unsigned int* col_ticket = nullptr;  // [1] blocks of the column sort done (zero between evaluations)
    size_t cap_cols = 0;
    size_t cap_slots = 0;            // clusters * 8
    int* slot_atom = nullptr;        // [slots] original atom or -1
    int2* slot_meta = nullptr;       // [slots] (LJ type, original atom or -1)
    double4* slot_pos_d = nullptr;   // [slots] position relative to the cluster origin + charge
    float4* slot_pos_f = nullptr;
    // PME per slot (AMBER with a grid): base grid index floor(K x / L) mod K
    // per dimension packed 10 bits each, and the fractions w (x, y, z) + charge
    int* slot_pme_fl = nullptr;
    double4* slot_pme_wd = nullptr;
    float4* slot_pme_wf = nullptr;
    int* cluster_col = nullptr;      // [clusters]
    double4* cl_origin = nullptr;    // [clusters] bbox centre (w unused)
    float4* cl_half = nullptr;       // [clusters] bbox half extents
    // pair list (capacity per i-cluster)
    int list_cap = 0;
    size_t cap_list = 0;
    int* list_count = nullptr;       // [clusters]
    int* list_j = nullptr;           // [clusters * cap]
    unsigned long long* list_mask = nullptr;
    double4* list_delta_d = nullptr;
    float4* list_delta_f = nullptr;
    // Verlet mode: the kept list (built at cutoff + skin) pruned every
    // evaluation to the tiles with a pair inside the cutoff, same layout
    size_t cap_inner = 0;
    size_t cap_inner_cl = 0;
    int* inner_count = nullptr;      // [clusters]
    int* inner_j = nullptr;          // [clusters * cap]
    unsigned long long* inner_mask = nullptr;
    double4* inner_delta_d = nullptr;
    float4* inner_delta_f = nullptr;
    // forces
    double* force_d = nullptr;                 // [3N]
    unsigned long long* force_fx = nullptr;    // [3N] fixed point
    // pair-kernel and PME-gather forces in slot order (3 per slot), so the
    // atomics of one tile's cluster fall on consecutive addresses; the final
    // reduction adds them to the atom-ordered accumulators through atom_slot
    double* force_d_s = nullptr;               // [3 slots]
    unsigned long long* force_fx_s = nullptr;  // [3 slots] fixed point
    int* atom_slot = nullptr;                  // [N] slot of each atom (the last list build), -1 before one
    // energies: partial rows, one per block of each producing kernel
    size_t cap_partial = 0;
    double* partial = nullptr;       // [rows * E_NSLOTS]
    double* energies = nullptr;      // [E_NSLOTS + kFlags] final energies, then the flags as doubles
    double* force_out = nullptr;     // [3N] FP64 forces for host-buffer evaluations
    unsigned int* flags = nullptr;   // [kFlags]: 0 fixed overflow, 1 list overflow, 2 column overflow,
                                     // 3 near-tile overflow, 4 cluster too wide for the fast search,
                                     // 5 the list was built (Verlet mode; not an error)
    // Verlet mode: raw coordinates at the last list build, and the state
    double* x_ref = nullptr;         // [3N]
    double* x_prune = nullptr;       // [3N] positions of the last prune of the kept list
    VerletState* verlet = nullptr;
    float* excl_rmax = nullptr;      // [1] largest excluded-pair distance (as float)
    // PME
    int pme_k[3] = {0, 0, 0};
    size_t cap_grid = 0;
    float* grid_f = nullptr;                    // [K^3] potential (mixed form)
    double* grid_phi_d = nullptr;               // [K^3] potential (FP64 form)
    void* cgrid = nullptr;                      // [Kx*Ky*(Kz/2+1)] complex (double2 / float2)
    double* bsp_mod = nullptr;                  // [Kx + Ky + Kz] B-spline moduli 1/|b|^2
    float* bsp_mod_f = nullptr;
    double* twiddle_d = nullptr;                // [Kx + Ky + Kz] complex, e^{-2 pi i m/K}
    float* twiddle_f = nullptr;
};

// Host-side per-evaluation geometry of the column grid.
struct Grid {
    int ncx = 1, ncy = 1;
    double sx = 0, sy = 0;
    double box[3] = {0, 0, 0};
    int dcx = 0, dcy = 0;            // column search reach
    double r_list = 0;               // list radius: max cutoff + Verlet skin
    // Triclinic cells bin in fractional coordinates: column (cx, cy) holds
    // s_x in [cx, cx + 1) / ncx and s_y in [cy, cy + 1) / ncy, and sx, sy are
    // the column widths in units of the edge lengths (|a| / ncx, |b| / ncy).
    // wx, wy: the distance between a column's opposite faces (sx, sy when
    // orthorhombic); frac_scale: the largest edge length over face distance
    // ratio, which converts a distance into a bound on s_x |a| and s_y |b|.
    // the cell of the evaluation: the caller's cell with its off-diagonal edge
    // components reduced (the same lattice, so the minimum images and the
    // pairs within the cutoff are the same, but the columns of a strongly
    // sheared cell stay narrow); the PME grid runs along its edges
    Cell cell{};
    double face[3] = {0, 0, 0};  // distances between opposite faces of cell (the lengths when orthorhombic)
    double wx = 0, wy = 0;
    double frac_scale = 1.0;
};

constexpr int kTimingSample = 16;  // timing mode 3: one evaluation in 16 carries the events

struct Timing {
    int mode = 0;         // 0 off, 1 every kernel group, 2 the pair kernel, 3 the pair kernel sampled
    bool record = false;  // this evaluation records its events
    long long tick = 0;   // evaluations since the mode was set (mode 3 sampling)
    double total_ms[MDGPU_K_COUNT] = {};
    long long launches[MDGPU_K_COUNT] = {};
};

}  // namespace mdg

namespace mdg {
// What a captured single-context evaluation depends on besides buffer contents.
struct GraphKey {
    const double* x;
    const double* out;
    double box[3];
    double angles[3];
    int pme_k[3];
    unsigned long long gen;
    int list_cap, search_general, split, fft_lines, counting, timing, n_clusters, n_cols;
    size_t caps[7];
};
}  // namespace mdg

struct mdgpu_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string error;
    bool have_system = false;
    bool have_params = false;
    mdgpu_params params{};
    mdg::Topo topo;
    mdg::Work work;
    mdg::Grid grid;
    mdg::Timing timing;
    mdg::AdjointBuf adj;
    std::vector<double> host_charge;
    bool counting = false;
    bool virial = false;            // statistics runs also form the full virial
    bool have_virial = false;       // the last evaluation formed it
    double last_virial[6] = {};     // its xx, yy, zz, xy, xz, yz (recip + pair + bonded + corrections)
    long long counts[4] = {0, 0, 0, 0};
    int n_clusters = 0;
    int n_cols = 0;
    int split = 4;                  // waves per i-cluster list in the pair kernel
    int fft_lines = 8;              // cap on FFT lines per block (power of two; 0 = 16)
    const double* cur_x = nullptr;  // raw coordinates of the evaluation in progress (device)
    bool search_general = false;    // use the general (per-shift) cluster search
    bool search_fast = true;        // the last list build ran the minimum-image search
    std::vector<void*> topo_allocs;
    // host copies of the per-type-pair LJ parameters (the device table also
    // holds cutoff-dependent shift constants, rebuilt by mdgpu_set_params)
    std::vector<double> host_lj_a, host_lj_b;
    std::vector<unsigned char> host_listed;
    // event pairs per kernel group; the extra pair times the inverse FFT passes
    hipEvent_t ev[2 * (MDGPU_K_COUNT + 1)] = {};
    // energies and flags of the last evaluation in pinned, host-mapped memory
    // (written by the final reduction kernel; d_ is its device address)
    double* h_energies = nullptr;
    double* d_host_energies = nullptr;
    // captured evaluation (single contexts): 0 none yet, 1 captured, 2 off for good
    bool use_graph = true;
    int graph_state = 0;
    hipGraph_t graph = nullptr;
    hipGraphExec_t graph_exec = nullptr;        // without timing events
    hipGraphExec_t graph_exec_timed = nullptr;  // with the pair kernel's events (timing modes 2, 3)
    mdg::GraphKey graph_key{}, last_key{};
    bool have_last_key = false;
    bool capturing = false;         // an evaluation is being captured
    unsigned long long gen = 0;     // bumped by system upload and set_params
    // Verlet list (skin > 0): kept across evaluations until an atom has moved
    // more than skin / 2 since the build; verlet_force makes the next
    // evaluation build it whatever the displacements
    double skin = 0.0;
    // cell angles alpha, beta, gamma in degrees (all 90: orthorhombic)
    double angles[3] = {90.0, 90.0, 90.0};
    bool verlet_force = true;
    long long list_builds = 0, list_evals = 0, list_prunes = 0;
    void* md = nullptr;             // molecular-dynamics buffers (md.hip), after mdgpu_md_set_masses
    void* minbuf = nullptr;         // minimiser buffers (minimize.hip)
    // concurrent PME chain (single-context evaluations)
    bool fork = true;
    hipStream_t side[1] = {nullptr};
    hipEvent_t ev_fork = nullptr;
    hipEvent_t ev_join[1] = {nullptr};
};

namespace mdg {

// Timing events on c->stream, for evaluations that record them.  While an
// evaluation is captured nothing is recorded: the capture adds the pair
// kernel's two events as graph nodes itself (timing modes 2 and 3; mode 1 is
// never captured).
inline hipError_t record_timing_event(mdgpu_ctx* c, hipEvent_t ev) {
    return c->capturing || !c->timing.record ? hipSuccess : hipEventRecord(ev, c->stream);
}
// the pair-kernel instance a launch with the context's settings uses
const void* pair_kernel_fn(const mdgpu_ctx* c);

// ---- launchers (each returns hipSuccess or the first HIP error) ---------
// excl: also bound the excluded and 1-4 pair distances (AMBER with exclusions)
// verlet: the context's Verlet state (nullptr: build the list unconditionally).
// With it, launch_sort first measures the displacements (k_disp); the sort and
// search kernels then return at once unless the list is rebuilt, and a
// refresh kernel moves the kept clusters' slot records to the new positions.
hipError_t launch_sort(mdgpu_ctx* c, const double* coords, bool mixed, bool excl, VerletState* verlet);
hipError_t launch_pairsearch(mdgpu_ctx* c, bool mixed, bool amber, const VerletState* verlet);
// displacement criterion alone: the decision goes to verlet->query
hipError_t launch_disp_query(mdgpu_ctx* c, const double* coords);
// Verlet mode: when the displacement check asks for it, prunes the kept list
// to the tiles that hold a pair within cutoff + kPruneBuffer at the current
// positions (the pair kernel reads that list) and records those positions.
hipError_t launch_prune(mdgpu_ctx* c, bool mixed, const double* coords, const VerletState* verlet);
// pair kernel, with the bonded terms and (amber) the FP64 excluded / 1-4
// corrections as further blocks of the same launch; energy rows partial_row0
// .. + nonbonded_rows + bonded_rows + corrections_rows
hipError_t launch_nonbonded(mdgpu_ctx* c, bool mixed, bool amber, int partial_row0, const double* coords);
hipError_t launch_pme(mdgpu_ctx* c, bool mixed, int partial_row0, int* rows_used);
hipError_t launch_reduce(mdgpu_ctx* c, bool mixed, int n_rows, double* out_forces);
// adjoint kernels (AMBER semantics)
hipError_t launch_pair_adjoint(mdgpu_ctx* c, bool mixed);
hipError_t launch_bonded_adjoint(mdgpu_ctx* c, const double* coords, bool mixed);
hipError_t launch_corrections_adjoint(mdgpu_ctx* c);
hipError_t launch_pme_potential(mdgpu_ctx* c, bool mixed, int partial_row0, double* pot);

// one evaluation of one context, in steps (api.hip): prepare buffers for a
// box; enqueue the launches (no synchronisation or allocation); act on the
// flags of a finished evaluation (MDGPU_OK, kEvalRedo after adjusting the
// context, or an error status); fill the outputs; or all of it, synchronously
constexpr int kEvalRedo = -1;
int ctx_prepare(mdgpu_ctx* c, const double box[3]);
int ctx_enqueue(mdgpu_ctx* c, const double* d_coords, double* d_forces);
int ctx_check(mdgpu_ctx* c, const double* energies_and_flags);
void ctx_outputs(mdgpu_ctx* c, const double* energies_and_flags, mdgpu_energies* out_e, double* out_virial);
int ctx_eval_device(mdgpu_ctx* c, const double* d_coords, const double box[3], mdgpu_energies* out_e,
                    double* d_forces);
int ctx_error(mdgpu_ctx* c, int status, const std::string& msg);
// several evaluations in a row without synchronisation (molecular dynamics):
// request a forced Verlet-list build if one is pending (before the first
// launch); enqueue one evaluation with the context's stream layout; account
// for nsteps finished evaluations whose energy rows are rows[k * stride]
int ctx_verlet_prologue(mdgpu_ctx* c);
int ctx_enqueue_step(mdgpu_ctx* c, const double* d_coords, double* d_forces);
void ctx_steps_done(mdgpu_ctx* c, const double* rows, int stride, int nsteps);
void md_free(mdgpu_ctx* c);
void min_free(mdgpu_ctx* c);

// rows of the partial-energy buffer each kernel needs for a given launch
int pair_block_threads(const mdgpu_ctx* c);  // threads per block of the pair launch
int nonbonded_rows(const mdgpu_ctx* c);
int bonded_rows(const mdgpu_ctx* c);
int corrections_rows(const mdgpu_ctx* c);
int pme_rows(const mdgpu_ctx* c);
int sort_blocks(int n);

// fixed-point scale exponent of the mixed form for a semantics
inline double fixed_scale(int semantics) { return semantics == MDGPU_SEM_REF ? 1048576.0 : 4294967296.0; }


}  // namespace mdg

#endif
