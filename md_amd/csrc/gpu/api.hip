// mdgpu C ABI: context, upload, parameters, evaluation and the final
// reduction.  One evaluation runs, on the context's stream:
//   [AMBER] excluded-pair extent -> sort + clusters -> cluster-pair search ->
//   nonbonded -> bonded -> [AMBER] FP64 corrections -> [AMBER] PME ->
//   reduction (energy rows in fixed order, fixed-point -> FP64 forces)
// and returns after one device -> host copy of the energies and flags.
#include "ctx.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <new>
#include <vector>

using mdg::Topo;
using mdg::Work;

namespace {

int set_error(mdgpu_ctx* c, int code, const std::string& msg) {
    if (c) c->error = msg;
    return code;
}

int hip_error(mdgpu_ctx* c, hipError_t e, const char* where) {
    return set_error(c, MDGPU_ERR_HIP, std::string(where) + ": " + hipGetErrorString(e));
}

#define HIP_TRY(expr, where)                          \
    do {                                              \
        hipError_t _e = (expr);                       \
        if (_e != hipSuccess) return hip_error(c, _e, where); \
    } while (0)

template <typename T>
hipError_t upload(mdgpu_ctx* c, T** dst, const T* src, size_t n) {
    *dst = nullptr;
    if (n == 0) return hipSuccess;
    hipError_t e = hipMalloc(reinterpret_cast<void**>(dst), n * sizeof(T));
    if (e != hipSuccess) return e;
    c->topo_allocs.push_back(*dst);
    return hipMemcpy(*dst, src, n * sizeof(T), hipMemcpyHostToDevice);
}

void free_topo(mdgpu_ctx* c) {
    for (void* p : c->topo_allocs) (void)hipFree(p);
    c->topo_allocs.clear();
    c->topo = Topo{};
}

template <typename T>
hipError_t ensure(T** p, size_t* cap, size_t n) {
    if (*p && *cap >= n) return hipSuccess;
    if (*p) (void)hipFree(*p);
    *p = nullptr;
    *cap = 0;
    hipError_t e = hipMalloc(reinterpret_cast<void**>(p), std::max<size_t>(n, 1) * sizeof(T));
    if (e == hipSuccess) *cap = n;
    return e;
}

void free_work(mdgpu_ctx* c) {
    Work& w = c->work;
    void* ptrs[] = {w.coords, w.wrapped, w.col_of, w.col_count, w.block_emax, w.col_ticket, w.col_cstart, w.col_key,
                    w.slot_atom, w.slot_meta, w.slot_pos_d, w.slot_pos_f, w.slot_pme_fl, w.slot_pme_wd,
                    w.slot_pme_wf, w.cluster_col, w.cl_origin, w.cl_half,
                    w.list_count, w.list_j, w.list_mask, w.list_delta_d, w.list_delta_f, w.inner_count, w.inner_j,
                    w.inner_mask, w.inner_delta_d, w.inner_delta_f, w.force_d, w.force_fx,
                    w.force_d_s, w.force_fx_s, w.atom_slot,
                    w.partial, w.energies, w.force_out, w.flags, w.excl_rmax, w.grid_f, w.grid_phi_d,
                    w.cgrid, w.bsp_mod, w.bsp_mod_f, w.twiddle_d, w.twiddle_f, w.x_ref, w.x_prune, w.verlet};
    for (void* p : ptrs)
        if (p) (void)hipFree(p);
    c->work = Work{};
}

bool is_2a3b(int n) {
    if (n < 2) return false;
    while (n % 2 == 0) n /= 2;
    while (n % 3 == 0) n /= 3;
    return n == 1;
}

// LJ pair table over (T+1)^2 types with the force-shift constants at rc.
int build_lj_table(mdgpu_ctx* c) {
    const std::vector<double>& A = c->host_lj_a;
    const std::vector<double>& B = c->host_lj_b;
    const int T = c->topo.n_types, tp = T + 1;
    const double rc2 = c->params.lj_cutoff * c->params.lj_cutoff;
    std::vector<double4> td(static_cast<size_t>(tp) * tp);
    std::vector<float4> tf(td.size());
    for (int i = 0; i < tp; ++i)
        for (int j = 0; j < tp; ++j) {
            const size_t k = static_cast<size_t>(i) * tp + j;
            double a = 0, b = 0;
            if (i < T && j < T) {
                a = A[static_cast<size_t>(i) * T + j];
                b = B[static_cast<size_t>(i) * T + j];
            }
            // U(rc) and G(rc) = -U'(rc)/rc (NonBondedUtils.cpp:162-183)
            const double s2 = 1.0 / rc2, s6 = s2 * s2 * s2;
            const double u = a * s6 * s6 - b * s6;
            const double g = 12.0 * a * s6 * s6 * s2 - 6.0 * b * s6 * s2;
            td[k] = make_double4(a, b, u, g);
            tf[k] = make_float4(static_cast<float>(a), static_cast<float>(b), static_cast<float>(u),
                                static_cast<float>(g));
        }
    hipError_t e = hipMemcpy(c->topo.lj_d, td.data(), td.size() * sizeof(double4), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(c->topo.lj_f, tf.data(), tf.size() * sizeof(float4), hipMemcpyHostToDevice);
    if (e != hipSuccess) return hip_error(c, e, "lj table upload");
    return MDGPU_OK;
}

}  // namespace

// ------------------------------------------------------------------ API --

extern "C" {

int mdgpu_create(int device, mdgpu_ctx** out) {
    if (!out) return MDGPU_ERR_ARG;
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || device < 0 || device >= ndev) return MDGPU_ERR_ARG;
    auto* c = new (std::nothrow) mdgpu_ctx();
    if (!c) return MDGPU_ERR_HIP;
    c->device = device;
    hipError_t e = hipSetDevice(device);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    for (auto& ev : c->ev)
        if (e == hipSuccess) e = hipEventCreate(&ev);
    if (e == hipSuccess) e = hipMalloc(reinterpret_cast<void**>(&c->work.flags), mdg::kFlags * sizeof(unsigned int));
    if (e == hipSuccess) e = hipMalloc(reinterpret_cast<void**>(&c->work.energies), (mdg::E_NSLOTS + mdg::kFlags) * sizeof(double));
    if (e == hipSuccess) e = hipMalloc(reinterpret_cast<void**>(&c->work.excl_rmax), sizeof(float));
    if (e == hipSuccess) e = hipMalloc(reinterpret_cast<void**>(&c->work.verlet), sizeof(mdg::VerletState));
    // energies and flags of every evaluation, written by the final reduction
    if (e == hipSuccess)
        e = hipHostMalloc(reinterpret_cast<void**>(&c->h_energies), (mdg::E_NSLOTS + mdg::kFlags) * sizeof(double),
                          hipHostMallocCoherent | hipHostMallocMapped);
    if (e == hipSuccess) e = hipHostGetDevicePointer(reinterpret_cast<void**>(&c->d_host_energies), c->h_energies, 0);
    // zero once; afterwards the final reduction clears them for the next evaluation
    if (e == hipSuccess) e = hipMemset(c->work.flags, 0, mdg::kFlags * sizeof(unsigned int));
    if (e == hipSuccess) e = hipMemset(c->work.excl_rmax, 0, sizeof(float));
    if (e == hipSuccess) e = hipMemset(c->work.verlet, 0, sizeof(mdg::VerletState));
    if (e != hipSuccess) {
        mdgpu_destroy(c);
        return MDGPU_ERR_HIP;
    }
    *out = c;
    return MDGPU_OK;
}

int mdgpu_destroy(mdgpu_ctx* c) {
    if (!c) return MDGPU_OK;
    (void)hipSetDevice(c->device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    free_topo(c);
    free_work(c);
    mdg::md_free(c);
    mdg::min_free(c);
    if (c->adj.acc) (void)hipFree(c->adj.acc);
    if (c->adj.pot_rec) (void)hipFree(c->adj.pot_rec);
    if (c->adj.lj_d) (void)hipFree(c->adj.lj_d);
    for (auto& ev : c->ev)
        if (ev) (void)hipEventDestroy(ev);
    if (c->graph_exec) (void)hipGraphExecDestroy(c->graph_exec);
    if (c->graph_exec_timed) (void)hipGraphExecDestroy(c->graph_exec_timed);
    if (c->graph) (void)hipGraphDestroy(c->graph);
    if (c->h_energies) (void)hipHostFree(c->h_energies);
    for (auto& s : c->side)
        if (s) (void)hipStreamDestroy(s);
    if (c->ev_fork) (void)hipEventDestroy(c->ev_fork);
    for (auto& ev : c->ev_join)
        if (ev) (void)hipEventDestroy(ev);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
    return MDGPU_OK;
}

const char* mdgpu_last_error(const mdgpu_ctx* c) { return c ? c->error.c_str() : "null context"; }

}  // extern "C"

namespace {

int validate(mdgpu_ctx* c, const mdgpu_system_desc* s) {
    const int n = s->n_atoms;
    if (n <= 0) return set_error(c, MDGPU_ERR_ARG, "n_atoms must be positive");
    if (s->n_types <= 0 || !s->charge || !s->type || !s->nb_index)
        return set_error(c, MDGPU_ERR_ARG, "charges, types and nb_index are required");
    for (int i = 0; i < n; ++i)
        if (s->type[i] < 0 || s->type[i] >= s->n_types) return set_error(c, MDGPU_ERR_ARG, "atom type out of range");
    for (long k = 0; k < static_cast<long>(s->n_types) * s->n_types; ++k)
        if (s->nb_index[k] < 0 || s->nb_index[k] > s->n_lj)
            return set_error(c, MDGPU_ERR_ARG, "nb_index out of range");
    auto atoms_ok = [&](const int32_t* a, long cnt) {
        for (long k = 0; k < cnt; ++k)
            if (a[k] < 0 || a[k] >= n) return false;
        return true;
    };
    auto types_ok = [&](const int32_t* t, long cnt, int ntypes) {
        for (long k = 0; k < cnt; ++k)
            if (t[k] < 0 || t[k] >= ntypes) return false;
        return true;
    };
    if (!atoms_ok(s->bond_atoms, 2L * s->n_bonds) || !types_ok(s->bond_type, s->n_bonds, s->n_bond_types) ||
        !atoms_ok(s->ub_atoms, 2L * s->n_ubs) || !types_ok(s->ub_type, s->n_ubs, s->n_ub_types) ||
        !atoms_ok(s->angle_atoms, 3L * s->n_angles) || !types_ok(s->angle_type, s->n_angles, s->n_angle_types) ||
        !atoms_ok(s->dihedral_atoms, 4L * s->n_dihedrals) ||
        !types_ok(s->dihedral_type, s->n_dihedrals, s->n_dihedral_types) ||
        !atoms_ok(s->improper_atoms, 4L * s->n_impropers) ||
        !types_ok(s->improper_type, s->n_impropers, s->n_improper_types) ||
        !atoms_ok(s->cmap_atoms, 5L * s->n_cmaps) || !atoms_ok(s->excl_pairs, 2L * s->n_excl_pairs) ||
        !atoms_ok(s->pairs14, 2L * s->n_pairs14))
        return set_error(c, MDGPU_ERR_ARG, "term list index out of range");
    if (s->n_cmaps > 0) {
        if (s->cmap_resolution < 1 || s->cmap_n_sets < 1 || !s->cmap_coeff)
            return set_error(c, MDGPU_ERR_ARG, "CMAP terms need coefficient tables");
        for (int k = 0; k < s->n_cmaps; ++k)
            if (s->cmap_set[k] < 1 || s->cmap_set[k] > s->cmap_n_sets)
                return set_error(c, MDGPU_ERR_ARG, "CMAP parameter set out of range");
    }
    if (s->excl_start) {
        if (s->excl_start[0] != 0) return set_error(c, MDGPU_ERR_ARG, "excl_start[0] must be 0");
        for (int i = 0; i < n; ++i) {
            if (s->excl_start[i + 1] < s->excl_start[i]) return set_error(c, MDGPU_ERR_ARG, "excl_start not sorted");
            for (int k = s->excl_start[i]; k < s->excl_start[i + 1]; ++k) {
                if (s->excl_list[k] < 0 || s->excl_list[k] >= n)
                    return set_error(c, MDGPU_ERR_ARG, "exclusion partner out of range");
                if (k > s->excl_start[i] && s->excl_list[k] <= s->excl_list[k - 1])
                    return set_error(c, MDGPU_ERR_ARG, "exclusion lists must be sorted and unique");
            }
        }
    }
    return MDGPU_OK;
}

}  // namespace

extern "C" int mdgpu_upload_system(mdgpu_ctx* c, const mdgpu_system_desc* s) {
    if (!c || !s) return MDGPU_ERR_ARG;
    int st = validate(c, s);
    if (st != MDGPU_OK) return st;
    HIP_TRY(hipSetDevice(c->device), "hipSetDevice");
    HIP_TRY(hipStreamSynchronize(c->stream), "sync");
    free_topo(c);
    c->have_system = false;
    Topo& t = c->topo;
    const int n = s->n_atoms, T = s->n_types;
    t.n_atoms = n;
    t.n_types = T;
    HIP_TRY(upload(c, &t.charge, s->charge, n), "upload charge");
    HIP_TRY(upload(c, &t.type, reinterpret_cast<const int*>(s->type), n), "upload type");
    HIP_TRY(upload(c, &t.nb_index, reinterpret_cast<const int*>(s->nb_index), static_cast<size_t>(T) * T),
            "upload nb_index");
    // LJ tables per type pair; listed = the pair has an LJ entry
    c->host_lj_a.assign(static_cast<size_t>(T) * T, 0.0);
    c->host_lj_b.assign(static_cast<size_t>(T) * T, 0.0);
    c->host_listed.assign(static_cast<size_t>(T + 1) * (T + 1), 0);
    for (int i = 0; i < T; ++i)
        for (int j = 0; j < T; ++j) {
            const int nb = s->nb_index[i * T + j];
            if (nb > 0) {
                c->host_lj_a[static_cast<size_t>(i) * T + j] = s->lj_a[nb - 1];
                c->host_lj_b[static_cast<size_t>(i) * T + j] = s->lj_b[nb - 1];
                c->host_listed[static_cast<size_t>(i) * (T + 1) + j] = 1;
            }
        }
    const size_t tp2 = static_cast<size_t>(T + 1) * (T + 1);
    HIP_TRY(hipMalloc(reinterpret_cast<void**>(&t.lj_d), tp2 * sizeof(double4)), "alloc lj");
    c->topo_allocs.push_back(t.lj_d);
    HIP_TRY(hipMalloc(reinterpret_cast<void**>(&t.lj_f), tp2 * sizeof(float4)), "alloc lj");
    c->topo_allocs.push_back(t.lj_f);
    HIP_TRY(upload(c, &t.lj_listed, c->host_listed.data(), tp2), "upload listed");
    {
        std::vector<double2> l14(std::max(s->n_lj, 1));
        for (int k = 0; k < s->n_lj; ++k) l14[k] = make_double2(s->lj14_a ? s->lj14_a[k] : 0.0, s->lj14_b ? s->lj14_b[k] : 0.0);
        HIP_TRY(upload(c, &t.lj14, l14.data(), l14.size()), "upload lj14");
    }
    // bonded terms
    auto pack2 = [](const int32_t* a, int n2) {
        std::vector<int2> v(n2);
        for (int k = 0; k < n2; ++k) v[k] = make_int2(a[2 * k], a[2 * k + 1]);
        return v;
    };
    auto par2 = [](const double* a, const double* b, int nt) {
        std::vector<double2> v(nt);
        for (int k = 0; k < nt; ++k) v[k] = make_double2(a[k], b[k]);
        return v;
    };
    t.n_bonds = s->n_bonds;
    {
        auto a = pack2(s->bond_atoms, s->n_bonds);
        auto p = par2(s->bond_k, s->bond_r0, s->n_bond_types);
        HIP_TRY(upload(c, &t.bond_atoms, a.data(), a.size()), "upload bonds");
        HIP_TRY(upload(c, &t.bond_type, reinterpret_cast<const int*>(s->bond_type), s->n_bonds), "upload bonds");
        HIP_TRY(upload(c, &t.bond_par, p.data(), p.size()), "upload bonds");
    }
    t.n_ubs = s->n_ubs;
    {
        auto a = pack2(s->ub_atoms, s->n_ubs);
        auto p = par2(s->ub_k, s->ub_r0, s->n_ub_types);
        HIP_TRY(upload(c, &t.ub_atoms, a.data(), a.size()), "upload ub");
        HIP_TRY(upload(c, &t.ub_type, reinterpret_cast<const int*>(s->ub_type), s->n_ubs), "upload ub");
        HIP_TRY(upload(c, &t.ub_par, p.data(), p.size()), "upload ub");
    }
    t.n_angles = s->n_angles;
    {
        std::vector<int4> a(s->n_angles);
        for (int k = 0; k < s->n_angles; ++k)
            a[k] = make_int4(s->angle_atoms[3 * k], s->angle_atoms[3 * k + 1], s->angle_atoms[3 * k + 2], 0);
        auto p = par2(s->angle_k, s->angle_theta0, s->n_angle_types);
        HIP_TRY(upload(c, &t.angle_atoms, a.data(), a.size()), "upload angles");
        HIP_TRY(upload(c, &t.angle_type, reinterpret_cast<const int*>(s->angle_type), s->n_angles), "upload angles");
        HIP_TRY(upload(c, &t.angle_par, p.data(), p.size()), "upload angles");
    }
    t.n_dihedrals = s->n_dihedrals;
    {
        std::vector<int4> a(s->n_dihedrals);
        for (int k = 0; k < s->n_dihedrals; ++k)
            a[k] = make_int4(s->dihedral_atoms[4 * k], s->dihedral_atoms[4 * k + 1], s->dihedral_atoms[4 * k + 2],
                             s->dihedral_atoms[4 * k + 3]);
        std::vector<double4> p(s->n_dihedral_types);
        for (int k = 0; k < s->n_dihedral_types; ++k)
            p[k] = make_double4(s->dihedral_k[k], s->dihedral_n[k], s->dihedral_phase[k], 0.0);
        HIP_TRY(upload(c, &t.dih_atoms, a.data(), a.size()), "upload dihedrals");
        HIP_TRY(upload(c, &t.dih_type, reinterpret_cast<const int*>(s->dihedral_type), s->n_dihedrals),
                "upload dihedrals");
        HIP_TRY(upload(c, &t.dih_par, p.data(), p.size()), "upload dihedrals");
    }
    t.n_impropers = s->n_impropers;
    {
        std::vector<int4> a(s->n_impropers);
        for (int k = 0; k < s->n_impropers; ++k)
            a[k] = make_int4(s->improper_atoms[4 * k], s->improper_atoms[4 * k + 1], s->improper_atoms[4 * k + 2],
                             s->improper_atoms[4 * k + 3]);
        auto p = par2(s->improper_k, s->improper_psi0, s->n_improper_types);
        HIP_TRY(upload(c, &t.imp_atoms, a.data(), a.size()), "upload impropers");
        HIP_TRY(upload(c, &t.imp_type, reinterpret_cast<const int*>(s->improper_type), s->n_impropers),
                "upload impropers");
        HIP_TRY(upload(c, &t.imp_par, p.data(), p.size()), "upload impropers");
    }
    t.n_cmaps = s->n_cmaps;
    t.cmap_res = s->cmap_resolution;
    t.cmap_n_sets = s->cmap_n_sets;
    if (s->n_cmaps > 0) {
        const size_t nc = static_cast<size_t>(s->cmap_n_sets) * s->cmap_resolution * s->cmap_resolution * 16;
        HIP_TRY(upload(c, &t.cmap_atoms, reinterpret_cast<const int*>(s->cmap_atoms), 5 * static_cast<size_t>(s->n_cmaps)),
                "upload cmap");
        HIP_TRY(upload(c, &t.cmap_set, reinterpret_cast<const int*>(s->cmap_set), s->n_cmaps), "upload cmap");
        HIP_TRY(upload(c, &t.cmap_coeff, s->cmap_coeff, nc), "upload cmap");
    }
    // exclusions
    t.has_excl = s->excl_start != nullptr && s->excl_start[n] > 0;
    if (t.has_excl) {
        HIP_TRY(upload(c, &t.excl_start, reinterpret_cast<const int*>(s->excl_start), static_cast<size_t>(n) + 1),
                "upload exclusions");
        HIP_TRY(upload(c, &t.excl_list, reinterpret_cast<const int*>(s->excl_list), static_cast<size_t>(s->excl_start[n])),
                "upload exclusions");
    }
    t.n_excl_pairs = s->n_excl_pairs;
    {
        auto a = pack2(s->excl_pairs, s->n_excl_pairs);
        HIP_TRY(upload(c, &t.excl_pairs, a.data(), a.size()), "upload excluded pairs");
    }
    t.n_pairs14 = s->n_pairs14;
    {
        auto a = pack2(s->pairs14, s->n_pairs14);
        HIP_TRY(upload(c, &t.pairs14, a.data(), a.size()), "upload 1-4 pairs");
        HIP_TRY(upload(c, &t.scee14, s->scee14, s->n_pairs14), "upload 1-4 pairs");
    }
    double q2 = 0.0;
    for (int i = 0; i < n; ++i) q2 += s->charge[i] * s->charge[i];
    t.sum_q2 = q2;
    c->host_charge.assign(s->charge, s->charge + n);
    t.n_lj = s->n_lj;
    t.n_bond_types = s->n_bond_types;
    t.n_ub_types = s->n_ub_types;
    t.n_angle_types = s->n_angle_types;
    t.n_dih_types = s->n_dihedral_types;
    t.n_imp_types = s->n_improper_types;
    {
        std::vector<int> pidx(tp2, -1);
        for (int i = 0; i < T; ++i)
            for (int j = 0; j < T; ++j) pidx[static_cast<size_t>(i) * (T + 1) + j] = s->nb_index[i * T + j] - 1;
        HIP_TRY(upload(c, &t.lj_pidx, pidx.data(), pidx.size()), "upload LJ index map");
    }
    c->have_system = true;
    ++c->gen;  // a captured evaluation graph no longer applies
    c->verlet_force = true;
    c->list_builds = c->list_evals = c->list_prunes = 0;
    if (c->have_params) return build_lj_table(c);
    return MDGPU_OK;
}

extern "C" int mdgpu_set_params(mdgpu_ctx* c, const mdgpu_params* p) {
    if (!c || !p) return MDGPU_ERR_ARG;
    if (p->semantics != MDGPU_SEM_REF && p->semantics != MDGPU_SEM_AMBER)
        return set_error(c, MDGPU_ERR_ARG, "unknown semantics");
    if (p->precision != MDGPU_PREC_FP64 && p->precision != MDGPU_PREC_MIXED)
        return set_error(c, MDGPU_ERR_ARG, "unknown precision");
    if (!(p->lj_cutoff > 0) || !(p->ee_cutoff > 0)) return set_error(c, MDGPU_ERR_ARG, "cutoffs must be positive");
    // the reference throws for a switch radius >= cutoff (NonBondedUtils.cpp:30-37)
    if (p->lj_switch > 0 && !(p->lj_switch < p->lj_cutoff))
        return set_error(c, MDGPU_ERR_ARG, "lj_switch must be < lj_cutoff");
    if (!(p->ewald_alpha >= 0)) return set_error(c, MDGPU_ERR_ARG, "ewald_alpha must be >= 0");
    const bool pme = p->pme_grid[0] != 0 || p->pme_grid[1] != 0 || p->pme_grid[2] != 0;
    if (pme) {
        if (p->semantics != MDGPU_SEM_AMBER)
            return set_error(c, MDGPU_ERR_UNSUPPORTED, "PME is evaluated in AMBER semantics only");
        if (p->pme_order != 4) return set_error(c, MDGPU_ERR_UNSUPPORTED, "only PME order 4 is implemented");
        for (int d = 0; d < 3; ++d)
            if (!is_2a3b(p->pme_grid[d]) || p->pme_grid[d] < 8 || p->pme_grid[d] > 512)
                return set_error(c, MDGPU_ERR_ARG, "PME grid sizes must be 2^a 3^b in [8, 512]");
        if (!(p->ewald_alpha > 0)) return set_error(c, MDGPU_ERR_ARG, "PME needs ewald_alpha > 0");
    }
    HIP_TRY(hipSetDevice(c->device), "hipSetDevice");
    c->params = *p;
    c->have_params = true;
    ++c->gen;
    c->verlet_force = true;
    if (c->have_system) return build_lj_table(c);
    return MDGPU_OK;
}

namespace {

int ensure_pme(mdgpu_ctx* c, const double box[3]) {
    Work& w = c->work;
    const int* K = c->params.pme_grid;
    const bool same = w.pme_k[0] == K[0] && w.pme_k[1] == K[1] && w.pme_k[2] == K[2];
    if (K[0] == 0) {
        w.pme_k[0] = w.pme_k[1] = w.pme_k[2] = 0;
        return MDGPU_OK;
    }
    if (!same) {
        const size_t G = static_cast<size_t>(K[0]) * K[1] * K[2];
        const size_t Gc = static_cast<size_t>(K[0]) * K[1] * (K[2] / 2 + 1);
        void* ptrs[] = {w.grid_f, w.grid_phi_d, w.cgrid, w.bsp_mod, w.twiddle_d, w.twiddle_f};
        for (void* p : ptrs)
            if (p) (void)hipFree(p);
        HIP_TRY(hipMalloc(reinterpret_cast<void**>(&w.grid_f), G * sizeof(float)), "alloc grid");
        HIP_TRY(hipMalloc(reinterpret_cast<void**>(&w.grid_phi_d), G * sizeof(double)), "alloc grid");
        HIP_TRY(hipMalloc(&w.cgrid, Gc * sizeof(double2)), "alloc grid");
        const int nk = K[0] + K[1] + K[2];
        HIP_TRY(hipMalloc(reinterpret_cast<void**>(&w.bsp_mod), nk * sizeof(double)), "alloc moduli");
        HIP_TRY(hipMalloc(reinterpret_cast<void**>(&w.twiddle_d), 2 * nk * sizeof(double)), "alloc twiddles");
        HIP_TRY(hipMalloc(reinterpret_cast<void**>(&w.twiddle_f), 2 * nk * sizeof(float)), "alloc twiddles");
        std::vector<double> mod(nk), twd(2 * nk);
        std::vector<float> twf(2 * nk);
        int off = 0;
        for (int d = 0; d < 3; ++d) {
            const int n = K[d];
            // 1/|b(m)|^2 = |sum_{k=0}^{2} M4(k+1) e^{2 pi i m k/n}|^2 with M4(1..3) = 1/6, 4/6, 1/6
            const double m4[3] = {1.0 / 6.0, 4.0 / 6.0, 1.0 / 6.0};
            for (int m = 0; m < n; ++m) {
                long double sr = 0, si = 0;
                for (int k = 0; k < 3; ++k) {
                    const long double ang = 2.0L * 3.141592653589793238462643383279502884L * m * k / n;
                    sr += m4[k] * cosl(ang);
                    si += m4[k] * sinl(ang);
                }
                mod[off + m] = static_cast<double>(1.0L / (sr * sr + si * si));
                const long double ang = -2.0L * 3.141592653589793238462643383279502884L * m / n;
                twd[2 * (off + m)] = static_cast<double>(cosl(ang));
                twd[2 * (off + m) + 1] = static_cast<double>(sinl(ang));
                twf[2 * (off + m)] = static_cast<float>(cosl(ang));
                twf[2 * (off + m) + 1] = static_cast<float>(sinl(ang));
            }
            off += n;
        }
        HIP_TRY(hipMemcpy(w.bsp_mod, mod.data(), nk * sizeof(double), hipMemcpyHostToDevice), "upload moduli");
        HIP_TRY(hipMemcpy(w.twiddle_d, twd.data(), 2 * nk * sizeof(double), hipMemcpyHostToDevice), "upload tw");
        HIP_TRY(hipMemcpy(w.twiddle_f, twf.data(), 2 * nk * sizeof(float), hipMemcpyHostToDevice), "upload tw");
        w.pme_k[0] = K[0];
        w.pme_k[1] = K[1];
        w.pme_k[2] = K[2];
    }
    (void)box;
    return MDGPU_OK;
}

// The cell of edge lengths box and angles (degrees): a along x, b in the
// x-y plane, c completing a right-handed cell (the rst7 convention).  An
// angle of exactly 90 degrees has cosine 0, so all-90 angles give the
// orthorhombic cell.  face: the distances between opposite faces.  False
// when the angles do not form a cell.
bool make_cell(const double box[3], const double ang[3], mdg::Cell& cell, double face[3]) {
    cell = mdg::Cell{};
    for (int d = 0; d < 3; ++d) cell.l[d] = box[d];
    cell.tri = !(ang[0] == 90.0 && ang[1] == 90.0 && ang[2] == 90.0);
    if (!cell.tri) {
        for (int d = 0; d < 3; ++d) face[d] = box[d];
        return true;
    }
    const double rad = mdg::kPi / 180.0;
    auto cosd = [&](double a) { return a == 90.0 ? 0.0 : std::cos(a * rad); };
    const double ca = cosd(ang[0]), cb = cosd(ang[1]), cg = cosd(ang[2]);
    const double sg = ang[2] == 90.0 ? 1.0 : std::sin(ang[2] * rad);
    if (!(sg > 1e-6)) return false;
    cell.ax = box[0];
    cell.bx = box[1] * cg;
    cell.by = box[1] * sg;
    cell.cx = box[2] * cb;
    cell.cy = box[2] * (ca - cb * cg) / sg;
    const double cz2 = box[2] * box[2] - cell.cx * cell.cx - cell.cy * cell.cy;
    if (!(cz2 > 1e-12 * box[2] * box[2])) return false;
    cell.cz = std::sqrt(cz2);
    cell.r00 = 1.0 / cell.ax;
    cell.r11 = 1.0 / cell.by;
    cell.r22 = 1.0 / cell.cz;
    cell.r01 = -cell.bx / (cell.ax * cell.by);
    cell.r12 = -cell.cy / (cell.by * cell.cz);
    cell.r02 = (cell.bx * cell.cy - cell.by * cell.cx) / (cell.ax * cell.by * cell.cz);
    face[0] = 1.0 / std::sqrt(cell.r00 * cell.r00 + cell.r01 * cell.r01 + cell.r02 * cell.r02);
    face[1] = 1.0 / std::sqrt(cell.r11 * cell.r11 + cell.r12 * cell.r12);
    face[2] = cell.cz;
    return true;
}

// The same lattice in a reduced basis: c loses whole multiples of b until
// |c_y| <= b_y / 2, then c and b lose whole multiples of a until |c_x|,
// |b_x| <= a_x / 2 (a, the diagonal a_x, b_y, c_z and the volume are kept).
// A strongly sheared cell (say edges a, a + b, a + b + c of a cube) becomes
// nearly rectangular, so its fractional columns stay narrow and the
// minimum-image cluster search applies; the PME grid then runs along the
// reduced edges.  An already reduced cell is returned unchanged, bit for bit.
// face: the reduced cell's face distances.
mdg::Cell reduce_cell(const mdg::Cell& in, double face[3]) {
    mdg::Cell r = in;
    if (!in.tri) return r;
    double k = std::nearbyint(r.cy / r.by);
    if (k != 0.0) {
        r.cx -= k * r.bx;
        r.cy -= k * r.by;
    }
    k = std::nearbyint(r.cx / r.ax);
    if (k != 0.0) r.cx -= k * r.ax;
    k = std::nearbyint(r.bx / r.ax);
    if (k != 0.0) r.bx -= k * r.ax;
    if (r.bx == in.bx && r.cx == in.cx && r.cy == in.cy) {
        face[0] = 1.0 / std::sqrt(in.r00 * in.r00 + in.r01 * in.r01 + in.r02 * in.r02);
        face[1] = 1.0 / std::sqrt(in.r11 * in.r11 + in.r12 * in.r12);
        face[2] = in.cz;
        return r;
    }
    r.l[1] = std::sqrt(r.bx * r.bx + r.by * r.by);
    r.l[2] = std::sqrt(r.cx * r.cx + r.cy * r.cy + r.cz * r.cz);
    r.r01 = -r.bx / (r.ax * r.by);
    r.r12 = -r.cy / (r.by * r.cz);
    r.r02 = (r.bx * r.cy - r.by * r.cx) / (r.ax * r.by * r.cz);
    face[0] = 1.0 / std::sqrt(r.r00 * r.r00 + r.r01 * r.r01 + r.r02 * r.r02);
    face[1] = 1.0 / std::sqrt(r.r11 * r.r11 + r.r12 * r.r12);
    face[2] = r.cz;
    return r;
}

int prepare(mdgpu_ctx* c, const double box[3]) {
    const int n = c->topo.n_atoms;
    const mdgpu_params& p = c->params;
    for (int d = 0; d < 3; ++d)
        if (!(box[d] > 0)) return set_error(c, MDGPU_ERR_ARG, "box lengths must be positive");
    // a Verlet list holds shift vectors of the box it was built in
    for (int d = 0; d < 3; ++d)
        if (box[d] != c->grid.box[d]) c->verlet_force = true;
    mdg::Cell given{};
    double face[3];
    if (!make_cell(box, c->angles, given, face))
        return set_error(c, MDGPU_ERR_ARG, "the cell angles do not form a cell");
    // bin and search in the reduced basis of the same lattice
    const mdg::Cell cell = reduce_cell(given, face);
    if (cell.tri && p.semantics != MDGPU_SEM_AMBER)
        return set_error(c, MDGPU_ERR_UNSUPPORTED, "triclinic cells are evaluated in AMBER semantics only");
    const double rc = std::max(p.lj_cutoff, p.ee_cutoff);
    // list radius: the cutoff, plus the skin of a Verlet list
    const double rl = rc + (c->skin > 0 ? c->skin : 0.0);
    if (!cell.tri && std::min({box[0], box[1], box[2]}) < 2.0 * rc)
        return set_error(c, MDGPU_ERR_UNSUPPORTED, "every box length must be at least twice the cutoff");
    if (cell.tri && (std::min({face[0], face[1], face[2]}) < 2.0 * rc || std::min({face[0], face[1], face[2]}) < rl))
        return set_error(c, MDGPU_ERR_UNSUPPORTED,
                         "every distance between opposite cell faces must be at least twice the cutoff and at least "
                         "the cutoff plus the skin");
    mdg::Grid& g = c->grid;
    g.r_list = rl;
    g.box[0] = box[0];
    g.box[1] = box[1];
    g.box[2] = box[2];
    g.cell = cell;
    for (int d = 0; d < 3; ++d) g.face[d] = face[d];
    const double dens = n / cell_volume(cell);
    const double s = std::cbrt(mdg::kCluster / std::max(dens, 1e-6));
    g.ncx = std::max(1, static_cast<int>(face[0] / s));
    g.ncy = std::max(1, static_cast<int>(face[1] / s));
    g.sx = cell.l[0] / g.ncx;
    g.sy = cell.l[1] / g.ncy;
    g.wx = cell.tri ? face[0] / g.ncx : g.sx;
    g.wy = cell.tri ? face[1] / g.ncy : g.sy;
    g.frac_scale = cell.tri ? std::max(cell.l[0] / face[0], cell.l[1] / face[1]) : 1.0;
    g.dcx = static_cast<int>(std::floor(rl / g.wx)) + 1;
    g.dcy = static_cast<int>(std::floor(rl / g.wy)) + 1;
    const int ncol = g.ncx * g.ncy;
    c->n_cols = ncol;
    c->n_clusters = n / mdg::kCluster + ncol + 1;
    Work& w = c->work;
    size_t cap;
    cap = w.cap_atoms;
    if (w.cap_atoms < static_cast<size_t>(n)) {
        void* ptrs[] = {w.coords, w.wrapped, w.col_of, w.force_d, w.force_fx, w.force_out, w.x_ref,
                        w.x_prune, w.atom_slot};
        for (void* q : ptrs)
            if (q) (void)hipFree(q);
        w.coords = w.wrapped = w.force_d = w.force_out = w.x_ref = w.x_prune = nullptr;
        w.atom_slot = nullptr;
        w.col_of = nullptr;
        w.force_fx = nullptr;
        HIP_TRY(hipMalloc(reinterpret_cast<void**>(&w.coords), 3 * n * sizeof(double)), "alloc");
        HIP_TRY(hipMalloc(reinterpret_cast<void**>(&w.wrapped), 3 * n * sizeof(double)), "alloc");
        HIP_TRY(hipMalloc(reinterpret_cast<void**>(&w.col_of), n * sizeof(int)), "alloc");
        HIP_TRY(hipMalloc(reinterpret_cast<void**>(&w.force_d), 3 * n * sizeof(double)), "alloc");
        HIP_TRY(hipMalloc(reinterpret_cast<void**>(&w.force_fx), 3 * n * sizeof(unsigned long long)), "alloc");
        HIP_TRY(hipMalloc(reinterpret_cast<void**>(&w.force_out), 3 * n * sizeof(double)), "alloc");
        HIP_TRY(hipMalloc(reinterpret_cast<void**>(&w.x_ref), 3 * n * sizeof(double)), "alloc");
        HIP_TRY(hipMalloc(reinterpret_cast<void**>(&w.x_prune), 3 * n * sizeof(double)), "alloc");
        HIP_TRY(hipMalloc(reinterpret_cast<void**>(&w.atom_slot), n * sizeof(int)), "alloc");
        HIP_TRY(hipMemset(w.atom_slot, 0xff, n * sizeof(int)), "clear");  // -1: no slot yet
        c->verlet_force = true;  // no list positions to compare with
        // the accumulators start at zero; every evaluation's reduction re-zeroes them
        HIP_TRY(hipMemset(w.force_d, 0, 3 * n * sizeof(double)), "clear");
        HIP_TRY(hipMemset(w.force_fx, 0, 3 * n * sizeof(unsigned long long)), "clear");
        w.cap_atoms = n;
    }
    (void)cap;
    const size_t nbc = static_cast<size_t>(mdg::sort_blocks(n)) * (ncol + 1);
    if (w.cap_cols < nbc) {
        void* ptrs[] = {w.col_count, w.block_emax, w.col_ticket, w.col_cstart};
        for (void* q : ptrs)
            if (q) (void)hipFree(q);
        HIP_TRY(hipMalloc(reinterpret_cast<void**>(&w.col_count), nbc * sizeof(int)), "alloc");
        HIP_TRY(hipMalloc(reinterpret_cast<void**>(&w.block_emax), nbc * sizeof(float)), "alloc");
        // column totals and the ticket start at zero; the last block of the
        // column sort re-zeroes them every evaluation
        HIP_TRY(hipMemset(w.col_count, 0, nbc * sizeof(int)), "clear");
        HIP_TRY(hipMalloc(reinterpret_cast<void**>(&w.col_ticket), sizeof(unsigned int)), "alloc");
        HIP_TRY(hipMemset(w.col_ticket, 0, sizeof(unsigned int)), "clear");
        HIP_TRY(hipMalloc(reinterpret_cast<void**>(&w.col_cstart), nbc * sizeof(int)), "alloc");
        w.cap_cols = nbc;
    }
    const size_t nkey = static_cast<size_t>(ncol) * mdg::kColBucket;
    if (w.cap_colkey < nkey) {
        if (w.col_key) (void)hipFree(w.col_key);
        w.col_key = nullptr;
        HIP_TRY(hipMalloc(reinterpret_cast<void**>(&w.col_key), nkey * sizeof(unsigned long long)), "alloc");
        w.cap_colkey = nkey;
    }
    const size_t ncl = c->n_clusters, nslots = ncl * mdg::kCluster;
    if (w.cap_slots < nslots) {
        void* ptrs[] = {w.slot_atom, w.slot_meta, w.slot_pos_d, w.slot_pos_f, w.slot_pme_fl, w.slot_pme_wd,
                        w.slot_pme_wf, w.cluster_col, w.cl_origin, w.cl_half, w.list_count, w.force_d_s,
                        w.force_fx_s};
        for (void* q : ptrs)
            if (q) (void)hipFree(q);
        HIP_TRY(hipMalloc(reinterpret_cast<void**>(&w.slot_atom), nslots * sizeof(int)), "alloc");
        HIP_TRY(hipMalloc(reinterpret_cast<void**>(&w.slot_meta), nslots * sizeof(int2)), "alloc");
        HIP_TRY(hipMalloc(reinterpret_cast<void**>(&w.slot_pos_d), nslots * sizeof(double4)), "alloc");
        HIP_TRY(hipMalloc(reinterpret_cast<void**>(&w.slot_pos_f), nslots * sizeof(float4)), "alloc");
        HIP_TRY(hipMalloc(reinterpret_cast<void**>(&w.slot_pme_fl), nslots * sizeof(int)), "alloc");
        HIP_TRY(hipMalloc(reinterpret_cast<void**>(&w.slot_pme_wd), nslots * sizeof(double4)), "alloc");
        HIP_TRY(hipMalloc(reinterpret_cast<void**>(&w.slot_pme_wf), nslots * sizeof(float4)), "alloc");
        HIP_TRY(hipMalloc(reinterpret_cast<void**>(&w.cluster_col), ncl * sizeof(int)), "alloc");
        HIP_TRY(hipMalloc(reinterpret_cast<void**>(&w.cl_origin), ncl * sizeof(double4)), "alloc");
        HIP_TRY(hipMalloc(reinterpret_cast<void**>(&w.cl_half), ncl * sizeof(float4)), "alloc");
        HIP_TRY(hipMalloc(reinterpret_cast<void**>(&w.list_count), ncl * sizeof(int)), "alloc");
        HIP_TRY(hipMalloc(reinterpret_cast<void**>(&w.force_d_s), 3 * nslots * sizeof(double)), "alloc");
        HIP_TRY(hipMalloc(reinterpret_cast<void**>(&w.force_fx_s), 3 * nslots * sizeof(unsigned long long)), "alloc");
        // start at zero; every evaluation's reduction re-zeroes what it read
        HIP_TRY(hipMemset(w.force_d_s, 0, 3 * nslots * sizeof(double)), "clear");
        HIP_TRY(hipMemset(w.force_fx_s, 0, 3 * nslots * sizeof(unsigned long long)), "clear");
        c->verlet_force = true;  // new slot arrays: the slot map must be rebuilt
        w.cap_slots = nslots;
        w.cap_list = 0;  // list arrays are sized by clusters x capacity
    }
    if (w.list_cap == 0) {
        // half-list j-clusters per i-cluster: the clusters inside the cutoff
        // sphere grown by one cluster width, halved, with headroom
        const double width = std::cbrt(mdg::kCluster / std::max(dens, 1e-6));
        const double r = rl + 2.0 * width;
        const double est = 0.5 * (4.0 / 3.0) * 3.14159265 * r * r * r * dens / mdg::kCluster;
        w.list_cap = std::max(64, static_cast<int>(1.5 * est) + 32);
    }
    const size_t nl = ncl * static_cast<size_t>(w.list_cap);
    if (w.cap_list < nl) {
        void* ptrs[] = {w.list_j, w.list_mask, w.list_delta_d, w.list_delta_f};
        for (void* q : ptrs)
            if (q) (void)hipFree(q);
        HIP_TRY(hipMalloc(reinterpret_cast<void**>(&w.list_j), nl * sizeof(int)), "alloc list");
        HIP_TRY(hipMalloc(reinterpret_cast<void**>(&w.list_mask), nl * sizeof(unsigned long long)), "alloc list");
        HIP_TRY(hipMalloc(reinterpret_cast<void**>(&w.list_delta_d), nl * sizeof(double4)), "alloc list");
        HIP_TRY(hipMalloc(reinterpret_cast<void**>(&w.list_delta_f), nl * sizeof(float4)), "alloc list");
        w.cap_list = nl;
    }
    if (c->skin > 0 && (w.cap_inner < nl || !w.inner_count || w.cap_inner_cl < ncl)) {
        // the pruned copy of a kept Verlet list
        void* ptrs[] = {w.inner_count, w.inner_j, w.inner_mask, w.inner_delta_d, w.inner_delta_f};
        for (void* q : ptrs)
            if (q) (void)hipFree(q);
        HIP_TRY(hipMalloc(reinterpret_cast<void**>(&w.inner_count), ncl * sizeof(int)), "alloc list");
        HIP_TRY(hipMalloc(reinterpret_cast<void**>(&w.inner_j), nl * sizeof(int)), "alloc list");
        HIP_TRY(hipMalloc(reinterpret_cast<void**>(&w.inner_mask), nl * sizeof(unsigned long long)), "alloc list");
        HIP_TRY(hipMalloc(reinterpret_cast<void**>(&w.inner_delta_d), nl * sizeof(double4)), "alloc list");
        HIP_TRY(hipMalloc(reinterpret_cast<void**>(&w.inner_delta_f), nl * sizeof(float4)), "alloc list");
        w.cap_inner = nl;
        w.cap_inner_cl = ncl;
    }
    int st = ensure_pme(c, box);
    if (st != MDGPU_OK) return st;
    const size_t rows = static_cast<size_t>(mdg::nonbonded_rows(c) + mdg::bonded_rows(c) +
                                            mdg::corrections_rows(c) + mdg::pme_rows(c) + 1);
    if (w.cap_partial < rows) {
        if (w.partial) (void)hipFree(w.partial);
        w.partial = nullptr;
        HIP_TRY(hipMalloc(reinterpret_cast<void**>(&w.partial), rows * mdg::E_NSLOTS * sizeof(double)), "alloc");
        w.cap_partial = rows;
    }
    return MDGPU_OK;
}

// Final reduction.  Blocks [0, E_NSLOTS) each sum one energy slot over the
// partial rows (fixed per-thread order, four independent partial sums per
// thread, then a fixed wave and block tree, so the result is deterministic);
// block E_NSLOTS copies the flags to the output and clears them; the
// remaining blocks convert the forces (fixed point -> FP64 in the mixed form)
// into `out` and clear the accumulators for the next evaluation.  Energies
// and flags go to out_buf and, when out_host is set, also to that
// host-mapped copy, which the host reads after the stream synchronisation.
constexpr int kReduceThreads = 256;
__global__ __launch_bounds__(kReduceThreads) void k_reduce(const double* __restrict__ partial, int ld, int rows,
                                                           double* __restrict__ out_buf, double* out_host,
                                                           unsigned int* flags,
                                                           float* excl_rmax, int mixed, unsigned long long* fx,
                                                           double* fd, double inv_scale, int n3,
                                                           double* __restrict__ out,
                                                           const int* __restrict__ list_count,
                                                           const int* __restrict__ ncl_dev,
                                                           unsigned long long* fx_s, double* fd_s,
                                                           const int* __restrict__ atom_slot, int max_slots) {
    __shared__ double red[kReduceThreads / 64];
    const int b = blockIdx.x, t = threadIdx.x;
    if (b < mdg::E_NSLOTS) {
        double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
        if (b == mdg::E_TILES) {
            const int ncl = *ncl_dev;
            for (int k = t; k < ncl; k += kReduceThreads) s0 += list_count[k];
        } else {
            constexpr int kStep = 4 * kReduceThreads;
            const double* col = partial + static_cast<size_t>(b) * ld;  // this slot's rows, contiguous
            int r = t;
            for (; r + 3 * kReduceThreads < rows; r += kStep) {
                s0 += col[r];
                s1 += col[r + kReduceThreads];
                s2 += col[r + 2 * kReduceThreads];
                s3 += col[r + 3 * kReduceThreads];
            }
            for (; r < rows; r += kReduceThreads) s0 += col[r];
        }
        double s = mdg::wave_sum((s0 + s1) + (s2 + s3));
        if ((t & 63) == 0) red[t >> 6] = s;
        __syncthreads();
        if (t == 0) {
            s = (red[0] + red[1]) + (red[2] + red[3]);
            out_buf[b] = s;
            if (out_host) out_host[b] = s;
        }
        return;
    }
    if (b == mdg::E_NSLOTS) {
        if (t < mdg::kFlags) {
            const double v = static_cast<double>(flags[t]);
            out_buf[mdg::E_NSLOTS + t] = v;
            if (out_host) out_host[mdg::E_NSLOTS + t] = v;
            flags[t] = 0u;
        }
        if (t == 0) *excl_rmax = 0.f;
        return;
    }
    const int nb = gridDim.x - mdg::E_NSLOTS - 1;
    for (int i = (b - mdg::E_NSLOTS - 1) * kReduceThreads + t; i < n3; i += nb * kReduceThreads) {
        // the slot-ordered part of this coordinate (pair kernel and PME gather)
        const int atom = i / 3, slot = atom_slot[atom];
        const bool has = slot >= 0 && slot < max_slots;
        const int j = 3 * slot + (i - 3 * atom);
        if (mixed) {
            // integer sum first: the result does not depend on the split
            const unsigned long long v = fx[i] + (has ? fx_s[j] : 0ull);
            out[i] = static_cast<double>(static_cast<long long>(v)) * inv_scale;
            fx[i] = 0ull;
            if (has) fx_s[j] = 0ull;
        } else {
            out[i] = fd[i] + (has ? fd_s[j] : 0.0);
            fd[i] = 0.0;
            if (has) fd_s[j] = 0.0;
        }
    }
}

// Side stream for the PME chain, created on first use.
hipError_t ensure_side_streams(mdgpu_ctx* c) {
    if (c->side[0]) return hipSuccess;
    hipError_t e = hipStreamCreateWithFlags(&c->side[0], hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&c->ev_fork, hipEventDisableTiming);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&c->ev_join[0], hipEventDisableTiming);
    return e;
}

// Points c->stream at another stream for the launches of one scope.
struct StreamSwap {
    mdgpu_ctx* c;
    hipStream_t saved;
    StreamSwap(mdgpu_ctx* ctx, hipStream_t s) : c(ctx), saved(ctx->stream) { c->stream = s; }
    ~StreamSwap() { c->stream = saved; }
};

// Launches one evaluation on c->stream after prepare(): no host
// synchronisation and no allocation, so a caller can capture it into a graph
// (with fork = false).  With fork, the PME chain runs on a side stream after
// the sort, concurrently with the pair search and the pair launch (which also
// carries the bonded terms and corrections), and the final reduction waits
// for it.  Energies and flags
// land in w.energies, forces in d_out.
int enqueue_eval(mdgpu_ctx* c, const double* x, double* d_out, bool fork) {
    const mdgpu_params& p = c->params;
    const bool mixed = p.precision == MDGPU_PREC_MIXED;
    const bool amber = p.semantics == MDGPU_SEM_AMBER;
    const int n = c->topo.n_atoms;
    Work& w = c->work;
    c->cur_x = x;
    auto& T = c->timing;
    auto timed = [&](int k) { return T.mode == 1 || (T.mode >= 2 && k == MDGPU_K_NONBONDED); };
    auto mark = [&](int k, bool start) {
        if (timed(k)) (void)mdg::record_timing_event(c, c->ev[2 * k + (start ? 0 : 1)]);
    };
    if (fork) HIP_TRY(ensure_side_streams(c), "side streams");
    mdg::VerletState* vs = c->skin > 0 ? w.verlet : nullptr;
    mark(MDGPU_K_SORT, true);
    HIP_TRY(mdg::launch_sort(c, x, mixed, amber && c->topo.has_excl, vs), "sort");
    mark(MDGPU_K_SORT, false);
    // energy rows: pair kernel + bonded + corrections (one launch), PME
    const int row_nb = 0;
    const int row_pme = row_nb + mdg::nonbonded_rows(c) + mdg::bonded_rows(c) + (amber ? mdg::corrections_rows(c) : 0);
    int rows = row_pme;
    const hipStream_t s_pme = fork ? c->side[0] : c->stream;
    if (fork) HIP_TRY(hipEventRecord(c->ev_fork, c->stream), "fork");
    // the critical path first: the side stream's launches are issued while
    // the search and the pair kernel already run
    mark(MDGPU_K_PAIRLIST, true);
    HIP_TRY(mdg::launch_pairsearch(c, mixed, amber, vs), "pair search");
    if (vs) HIP_TRY(mdg::launch_prune(c, mixed, x, vs), "prune");
    mark(MDGPU_K_PAIRLIST, false);
    mark(MDGPU_K_NONBONDED, true);
    HIP_TRY(mdg::launch_nonbonded(c, mixed, amber, row_nb, x), "nonbonded");
    mark(MDGPU_K_NONBONDED, false);
    if (amber) {
        if (fork) HIP_TRY(hipStreamWaitEvent(s_pme, c->ev_fork, 0), "fork");
        StreamSwap sw(c, s_pme);
        int used = 0;
        HIP_TRY(mdg::launch_pme(c, mixed, row_pme, &used), "pme");
        rows += used;
        if (fork) HIP_TRY(hipEventRecord(c->ev_join[0], s_pme), "join");
    }
    if (fork && amber) HIP_TRY(hipStreamWaitEvent(c->stream, c->ev_join[0], 0), "join");
    mark(MDGPU_K_REDUCE, true);
    const int fblocks = std::max(1, std::min(512, (3 * n + kReduceThreads - 1) / kReduceThreads));
    hipLaunchKernelGGL(k_reduce, dim3(mdg::E_NSLOTS + 1 + fblocks), dim3(kReduceThreads), 0, c->stream, w.partial,
                       static_cast<int>(w.cap_partial), rows, w.energies, c->d_host_energies, w.flags, w.excl_rmax,
                       mixed ? 1 : 0, w.force_fx, w.force_d, 1.0 / mdg::fixed_scale(p.semantics), 3 * n, d_out,
                       vs ? w.inner_count : w.list_count, w.col_cstart + c->n_cols, w.force_fx_s, w.force_d_s,
                       w.atom_slot,
                       static_cast<int>(w.cap_slots));
    HIP_TRY(hipGetLastError(), "reduce");
    mark(MDGPU_K_REDUCE, false);
    return MDGPU_OK;
}

constexpr int kRedo = mdg::kEvalRedo;

// A forced Verlet-list build is requested through device memory, outside
// any capture, so that captured evaluations see it as well.
int verlet_prologue(mdgpu_ctx* c) {
    if (c->skin > 0 && c->verlet_force)
        HIP_TRY(hipMemsetD32Async(reinterpret_cast<hipDeviceptr_t>(&c->work.verlet->force), 1, 1, c->stream),
                "verlet force");
    return MDGPU_OK;
}

// Acts on the flags of a finished evaluation (e: host copy of w.energies).
// Returns MDGPU_OK, kRedo after adjusting the context so that a repeat
// succeeds (larger list, general search), or an error status.
int check_flags(mdgpu_ctx* c, const double* e) {
    unsigned int flags[mdg::kFlags];
    for (int k = 0; k < mdg::kFlags; ++k) flags[k] = static_cast<unsigned int>(e[mdg::E_NSLOTS + k]);
    if (flags[4]) {
        // a cluster is too wide for the minimum-image search: redo with the general one
        c->search_general = true;
        return kRedo;
    }
    if (flags[2]) return set_error(c, MDGPU_ERR_UNSUPPORTED, "a sort column holds more than 2048 atoms");
    if (flags[3]) return set_error(c, MDGPU_ERR_UNSUPPORTED, "too many tiles near excluded pairs in one cluster");
    if (flags[1]) {
        // list capacity exceeded: grow and evaluate again
        c->work.list_cap *= 2;
        return kRedo;
    }
    if (flags[0]) return set_error(c, MDGPU_ERR_OVERFLOW, "fixed-point force accumulator overflow");
    return MDGPU_OK;
}

// Energies, virial and work counts of a finished evaluation.
void fill_outputs(mdgpu_ctx* c, const double* e, mdgpu_energies* out_e, double* out_virial) {
    const mdgpu_params& p = c->params;
    const bool amber = p.semantics == MDGPU_SEM_AMBER;
    c->counts[0] = static_cast<long long>(llround(e[mdg::E_PAIRS]));
    c->counts[1] = static_cast<long long>(llround(e[mdg::E_TILES]));
    c->counts[2] = c->n_clusters;
    c->counts[3] = c->topo.n_atoms;
    if (out_e) {
        mdgpu_energies& o = *out_e;
        o.bond = e[mdg::E_BOND];
        o.angle = e[mdg::E_ANGLE];
        o.dihedral = e[mdg::E_DIHEDRAL];
        o.urey_bradley = e[mdg::E_UB];
        o.improper = e[mdg::E_IMPROPER];
        o.cmap = e[mdg::E_CMAP];
        o.vdw = e[mdg::E_VDW];
        o.ee_all = amber ? 0.0 : e[mdg::E_EE_ALL];
        o.ee_pairlist = amber ? 0.0 : e[mdg::E_EE_LIST];
        o.ee_rf = amber ? 0.0 : e[mdg::E_EE_RF];
        o.ewald_direct = e[mdg::E_EW_DIRECT];
        o.ewald_self = -mdg::kCoulomb * (p.ewald_alpha / std::sqrt(mdg::kPi)) * c->topo.sum_q2;
        o.ewald_recip = amber ? e[mdg::E_EW_RECIP] : 0.0;
        o.ewald_excl = amber ? e[mdg::E_EW_EXCL] : 0.0;
        o.vdw14 = amber ? e[mdg::E_VDW14] : 0.0;
        o.ee14 = amber ? e[mdg::E_EE14] : 0.0;
        const double bonded = o.bond + o.angle + o.dihedral + o.urey_bradley + o.improper + o.cmap;
        o.total = amber ? bonded + o.vdw + o.vdw14 + o.ewald_direct + o.ewald_recip + o.ewald_self + o.ewald_excl +
                              o.ee14
                        : bonded + o.vdw + o.ee_pairlist;
    }
    c->have_virial = c->counting || c->virial;
    for (int k = 0; k < 6; ++k) c->last_virial[k] = e[mdg::E_VIR_XX + k] + e[mdg::E_VSR_XX + k];
    if (out_virial) {
        const double xx = e[mdg::E_VIR_XX], yy = e[mdg::E_VIR_YY], zz = e[mdg::E_VIR_ZZ];
        const double xy = e[mdg::E_VIR_XY], xz = e[mdg::E_VIR_XZ], yz = e[mdg::E_VIR_YZ];
        const double v[9] = {xx, xy, xz, xy, yy, yz, xz, yz, zz};
        std::memcpy(out_virial, v, sizeof v);
    }
}

// Everything a captured evaluation bakes into its kernel arguments.
mdg::GraphKey graph_key(const mdgpu_ctx* c, const double* x, const double* out) {
    const Work& w = c->work;
    mdg::GraphKey k{};
    k.x = x;
    k.out = out;
    for (int d = 0; d < 3; ++d) {
        k.box[d] = c->grid.box[d];
        k.angles[d] = c->angles[d];
        k.pme_k[d] = w.pme_k[d];
    }
    k.gen = c->gen;
    k.list_cap = w.list_cap;
    k.search_general = c->search_general ? 1 : 0;
    k.split = c->split;
    k.fft_lines = c->fft_lines;
    k.counting = (c->counting ? 1 : 0) | (c->virial ? 2 : 0);
    k.timing = c->timing.mode;
    k.n_clusters = c->n_clusters;
    k.n_cols = c->n_cols;
    k.caps[0] = w.cap_atoms;
    k.caps[1] = w.cap_cols;
    k.caps[2] = w.cap_slots;
    k.caps[3] = w.cap_list + w.cap_inner;
    k.caps[4] = w.cap_partial;
    k.caps[5] = w.cap_grid;
    k.caps[6] = w.cap_colkey;
    return k;
}

bool same_key(const mdg::GraphKey& a, const mdg::GraphKey& b) {
    bool eq = a.x == b.x && a.out == b.out && a.gen == b.gen && a.list_cap == b.list_cap &&
              a.search_general == b.search_general && a.split == b.split && a.fft_lines == b.fft_lines &&
              a.counting == b.counting && a.timing == b.timing && a.n_clusters == b.n_clusters &&
              a.n_cols == b.n_cols;
    for (int d = 0; d < 3; ++d)
        eq = eq && a.box[d] == b.box[d] && a.angles[d] == b.angles[d] && a.pme_k[d] == b.pme_k[d];
    for (int k = 0; k < 7; ++k) eq = eq && a.caps[k] == b.caps[k];
    return eq;
}

// Timing mode 2 in a captured evaluation: event-record nodes for the pair
// kernel's two events, between the kernel node and its neighbours.
hipError_t add_pair_event_nodes(mdgpu_ctx* c, hipGraph_t g) {
    size_t n = 0;
    hipError_t e = hipGraphGetNodes(g, nullptr, &n);
    if (e != hipSuccess || n == 0) return e != hipSuccess ? e : hipErrorInvalidValue;
    std::vector<hipGraphNode_t> nodes(n);
    if ((e = hipGraphGetNodes(g, nodes.data(), &n)) != hipSuccess) return e;
    const void* fn = mdg::pair_kernel_fn(c);
    hipGraphNode_t pk = nullptr;
    for (hipGraphNode_t nd : nodes) {
        hipGraphNodeType ty;
        if (hipGraphNodeGetType(nd, &ty) != hipSuccess || ty != hipGraphNodeTypeKernel) continue;
        hipKernelNodeParams kp{};
        if (hipGraphKernelNodeGetParams(nd, &kp) == hipSuccess && kp.func == fn) {
            pk = nd;
            break;
        }
    }
    if (!pk) return hipErrorInvalidValue;
    size_t nd_in = 0, nd_out = 0;
    if ((e = hipGraphNodeGetDependencies(pk, nullptr, &nd_in)) != hipSuccess) return e;
    if ((e = hipGraphNodeGetDependentNodes(pk, nullptr, &nd_out)) != hipSuccess) return e;
    std::vector<hipGraphNode_t> in(nd_in), out(nd_out);
    if (nd_in && (e = hipGraphNodeGetDependencies(pk, in.data(), &nd_in)) != hipSuccess) return e;
    if (nd_out && (e = hipGraphNodeGetDependentNodes(pk, out.data(), &nd_out)) != hipSuccess) return e;
    const int k = MDGPU_K_NONBONDED;
    hipGraphNode_t e0 = nullptr, e1 = nullptr;
    if ((e = hipGraphAddEventRecordNode(&e0, g, in.data(), nd_in, c->ev[2 * k])) != hipSuccess) return e;
    for (hipGraphNode_t p : in)
        if ((e = hipGraphRemoveDependencies(g, &p, &pk, 1)) != hipSuccess) return e;
    if ((e = hipGraphAddDependencies(g, &e0, &pk, 1)) != hipSuccess) return e;
    if ((e = hipGraphAddEventRecordNode(&e1, g, &pk, 1, c->ev[2 * k + 1])) != hipSuccess) return e;
    for (hipGraphNode_t d : out) {
        if ((e = hipGraphRemoveDependencies(g, &pk, &d, 1)) != hipSuccess) return e;
        if ((e = hipGraphAddDependencies(g, &e1, &d, 1)) != hipSuccess) return e;
    }
    return hipSuccess;
}

// Captures one evaluation on c->stream into c->graph_exec.  A failure leaves
// the graph path switched off for this context.
int capture_eval(mdgpu_ctx* c, const double* x, double* out) {
    c->graph_state = 2;  // unless the capture below succeeds
    HIP_TRY(ensure_side_streams(c), "side streams");
    if (hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal) != hipSuccess) {
        (void)hipGetLastError();
        return MDGPU_OK;
    }
    c->capturing = true;
    const int st = enqueue_eval(c, x, out, true);
    c->capturing = false;
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(c->stream, &g);
    if (st != MDGPU_OK || e != hipSuccess || !g) {
        if (g) (void)hipGraphDestroy(g);
        (void)hipGetLastError();
        return MDGPU_OK;
    }
    hipGraphExec_t ge = nullptr, get = nullptr;
    if (hipGraphInstantiate(&ge, g, nullptr, nullptr, 0) != hipSuccess) {
        (void)hipGraphDestroy(g);
        (void)hipGetLastError();
        return MDGPU_OK;
    }
    // the timed instance: the same graph with the pair kernel's event nodes
    if (c->timing.mode >= 2 &&
        (add_pair_event_nodes(c, g) != hipSuccess || hipGraphInstantiate(&get, g, nullptr, nullptr, 0) != hipSuccess)) {
        (void)hipGraphExecDestroy(ge);
        (void)hipGraphDestroy(g);
        (void)hipGetLastError();
        return MDGPU_OK;
    }
    c->graph = g;
    c->graph_exec = ge;
    c->graph_exec_timed = get;
    c->graph_state = 1;
    return MDGPU_OK;
}

int eval_impl(mdgpu_ctx* c, const double* d_coords, const double box[3], mdgpu_energies* out_e, double* d_out,
              double* out_virial, bool coords_are_work) {
    if (!c->have_system || !c->have_params) return set_error(c, MDGPU_ERR_STATE, "upload a system and set params first");
    HIP_TRY(hipSetDevice(c->device), "hipSetDevice");
    const bool amber = c->params.semantics == MDGPU_SEM_AMBER;
    for (int attempt = 0; attempt < 4; ++attempt) {
        int st = prepare(c, box);
        if (st != MDGPU_OK) return st;
        Work& w = c->work;
        const double* x = coords_are_work ? w.coords : d_coords;
        double* out = d_out ? d_out : w.force_out;
        st = verlet_prologue(c);
        if (st != MDGPU_OK) return st;
        // Graph path: once an evaluation with the same key has succeeded
        // (capacities settled), the next one is captured and replayed from
        // then on; a different key or a redo switches the path off for good
        // (the context is never captured twice).
        bool graph = false;
        const mdg::GraphKey key = graph_key(c, x, out);
        auto& T = c->timing;
        T.record = T.mode == 1 || T.mode == 2 || (T.mode == 3 && T.tick % mdg::kTimingSample == 0);
        if (c->use_graph && c->fork && c->timing.mode != 1 && c->graph_state != 2) {
            if (c->graph_state == 1) {
                if (same_key(key, c->graph_key)) graph = true;
                else c->graph_state = 2;
            } else if (c->have_last_key && same_key(key, c->last_key)) {
                st = capture_eval(c, x, out);
                if (st != MDGPU_OK) return st;
                if (c->graph_state == 1) {
                    c->graph_key = key;
                    graph = true;
                }
            }
        }
        if (graph) {
            HIP_TRY(hipGraphLaunch(T.record ? c->graph_exec_timed : c->graph_exec, c->stream), "graph launch");
        } else {
            st = enqueue_eval(c, x, out, c->fork);
            if (st != MDGPU_OK) return st;
        }
        // the reduction wrote energies and flags into host-mapped memory
        HIP_TRY(hipStreamSynchronize(c->stream), "evaluation");
        double e[mdg::E_NSLOTS + mdg::kFlags];
        std::memcpy(e, c->h_energies, sizeof e);
        st = check_flags(c, e);
        c->have_last_key = false;
        if (st == kRedo) {
            if (graph) c->graph_state = 2;
            c->verlet_force = true;
            continue;
        }
        if (st != MDGPU_OK) return st;
        c->verlet_force = false;
        ++c->list_evals;
        if (c->skin <= 0 || e[mdg::E_NSLOTS + 5] != 0.0) ++c->list_builds;
        if (c->skin > 0 && e[mdg::E_NSLOTS + 6] != 0.0) ++c->list_prunes;
        c->last_key = key;
        c->have_last_key = true;
        ++T.tick;
        if (T.record) {
            auto timed = [&](int k) { return T.mode == 1 || (T.mode >= 2 && k == MDGPU_K_NONBONDED); };
            for (int k = 0; k <= MDGPU_K_COUNT; ++k) {
                if (!timed(k == MDGPU_K_COUNT ? MDGPU_K_FFT : k)) continue;
                if (k == MDGPU_K_SPREAD || k == MDGPU_K_FFT || k == MDGPU_K_CONV || k == MDGPU_K_GATHER ||
                    k == MDGPU_K_COUNT) {
                    if (!(amber && w.pme_k[0] > 0)) continue;
                }
                if (k == MDGPU_K_CORRECT || k == MDGPU_K_BONDED) continue;  // timed within the pair launch
                float ms = 0.f;
                if (hipEventElapsedTime(&ms, c->ev[2 * k], c->ev[2 * k + 1]) == hipSuccess) {
                    const int slot = k == MDGPU_K_COUNT ? MDGPU_K_FFT : k;
                    T.total_ms[slot] += ms;
                    if (k != MDGPU_K_COUNT) T.launches[slot] += 1;
                } else {
                    (void)hipGetLastError();  // a timing failure must not surface in a later launch check
                }
            }
        }
        fill_outputs(c, e, out_e, out_virial);
        return MDGPU_OK;
    }
    return set_error(c, MDGPU_ERR_HIP, "pair list did not fit after regrowing");
}

}  // namespace

namespace mdg {
int ctx_prepare(mdgpu_ctx* c, const double box[3]) { return prepare(c, box); }
int ctx_enqueue(mdgpu_ctx* c, const double* d_coords, double* d_forces) {
    return enqueue_eval(c, d_coords, d_forces ? d_forces : c->work.force_out, false);
}
int ctx_check(mdgpu_ctx* c, const double* e) { return check_flags(c, e); }
void ctx_outputs(mdgpu_ctx* c, const double* e, mdgpu_energies* out_e, double* out_virial) {
    fill_outputs(c, e, out_e, out_virial);
}
int ctx_eval_device(mdgpu_ctx* c, const double* d_coords, const double box[3], mdgpu_energies* out_e,
                    double* d_forces) {
    return eval_impl(c, d_coords, box, out_e, d_forces, nullptr, false);
}
int ctx_error(mdgpu_ctx* c, int status, const std::string& msg) { return set_error(c, status, msg.c_str()); }
int ctx_verlet_prologue(mdgpu_ctx* c) { return verlet_prologue(c); }
int ctx_enqueue_step(mdgpu_ctx* c, const double* d_coords, double* d_forces) {
    return enqueue_eval(c, d_coords, d_forces, c->fork);
}
void ctx_steps_done(mdgpu_ctx* c, const double* rows, int stride, int nsteps) {
    c->verlet_force = false;
    c->list_evals += nsteps;
    for (int k = 0; k < nsteps; ++k) {
        if (c->skin <= 0 || rows[static_cast<size_t>(k) * stride + mdg::E_NSLOTS + 5] != 0.0) ++c->list_builds;
        if (c->skin > 0 && rows[static_cast<size_t>(k) * stride + mdg::E_NSLOTS + 6] != 0.0) ++c->list_prunes;
    }
}
}  // namespace mdg

extern "C" {

int mdgpu_eval(mdgpu_ctx* c, const double* coords, const double box[3], mdgpu_energies* out_e, double* out_forces,
               double* out_virial) {
    if (!c || !coords || !box) return set_error(c, MDGPU_ERR_ARG, "null argument");
    if (!c->have_system) return set_error(c, MDGPU_ERR_STATE, "upload a system first");
    HIP_TRY(hipSetDevice(c->device), "hipSetDevice");
    const int n = c->topo.n_atoms;
    int st = prepare(c, box);
    if (st != MDGPU_OK) return st;
    HIP_TRY(hipMemcpyAsync(c->work.coords, coords, 3 * n * sizeof(double), hipMemcpyHostToDevice, c->stream),
            "upload coords");
    st = eval_impl(c, nullptr, box, out_e, nullptr, out_virial, true);
    if (st != MDGPU_OK) return st;
    if (out_forces) {
        HIP_TRY(hipMemcpy(out_forces, c->work.force_out, 3 * n * sizeof(double), hipMemcpyDeviceToHost), "read forces");
    }
    return MDGPU_OK;
}

int mdgpu_eval_device(mdgpu_ctx* c, const double* d_coords, const double box[3], mdgpu_energies* out_e,
                      double* d_forces, double* out_virial) {
    if (!c || !d_coords || !box) return set_error(c, MDGPU_ERR_ARG, "null argument");
    return eval_impl(c, d_coords, box, out_e, d_forces, out_virial, false);
}

namespace {

// Offsets of the adjoint accumulators for the uploaded system; allocates on
// first use or when the system grew.
int ensure_adjoint(mdgpu_ctx* c) {
    const Topo& t = c->topo;
    mdg::AdjointBuf& b = c->adj;
    const size_t n = static_cast<size_t>(t.n_atoms), nlj = static_cast<size_t>(std::max(t.n_lj, 0));
    size_t o = 0;
    b.q_dir = o; o += n;
    b.q_excl = o; o += n;
    b.q_14 = o; o += n;
    b.lj_a = o; o += 2 * nlj;
    b.lj_b = o; o += 2 * nlj;
    b.lj14_a = o; o += nlj;
    b.lj14_b = o; o += nlj;
    b.bonded = o;
    o += 2 * static_cast<size_t>(t.n_bond_types + t.n_ub_types + t.n_angle_types + t.n_dih_types + t.n_imp_types);
    b.cmap = o;
    if (t.n_cmaps > 0) o += static_cast<size_t>(t.cmap_n_sets) * t.cmap_res * t.cmap_res * 16;
    b.flag = o; o += 1;
    b.total = o;
    if (b.cap < o) {
        if (b.acc) (void)hipFree(b.acc);
        b.acc = nullptr;
        b.cap = 0;
        HIP_TRY(hipMalloc(reinterpret_cast<void**>(&b.acc), o * sizeof(unsigned long long)), "alloc adjoint");
        b.cap = o;
    }
    if (b.cap_lj_d < 2 * nlj + 1) {
        if (b.lj_d) (void)hipFree(b.lj_d);
        b.lj_d = nullptr;
        b.cap_lj_d = 0;
        HIP_TRY(hipMalloc(reinterpret_cast<void**>(&b.lj_d), (2 * nlj + 1) * sizeof(double)), "alloc adjoint");
        b.cap_lj_d = 2 * nlj + 1;
    }
    if (b.cap_pot < n) {
        if (b.pot_rec) (void)hipFree(b.pot_rec);
        b.pot_rec = nullptr;
        b.cap_pot = 0;
        HIP_TRY(hipMalloc(reinterpret_cast<void**>(&b.pot_rec), std::max<size_t>(n, 1) * sizeof(double)),
                "alloc adjoint");
        b.cap_pot = n;
    }
    return MDGPU_OK;
}

}  // namespace

int mdgpu_adjoint(mdgpu_ctx* c, const double* coords, const double box[3], const mdgpu_energies* weights,
                  mdgpu_param_grads* out) {
    if (!c || !coords || !box || !weights || !out) return set_error(c, MDGPU_ERR_ARG, "null argument");
    if (!c->have_system || !c->have_params) return set_error(c, MDGPU_ERR_STATE, "upload a system and set params first");
    const mdgpu_params& p = c->params;
    if (p.semantics != MDGPU_SEM_AMBER)
        return set_error(c, MDGPU_ERR_UNSUPPORTED, "the adjoint is defined for AMBER semantics only");
    if (!(c->angles[0] == 90.0 && c->angles[1] == 90.0 && c->angles[2] == 90.0))
        return set_error(c, MDGPU_ERR_UNSUPPORTED, "the adjoint is defined for orthorhombic cells only");
    HIP_TRY(hipSetDevice(c->device), "hipSetDevice");
    const bool mixed = p.precision == MDGPU_PREC_MIXED;
    const Topo& t = c->topo;
    const int n = t.n_atoms;
    int st = ensure_adjoint(c);
    if (st != MDGPU_OK) return st;
    mdg::AdjointBuf& b = c->adj;
    std::vector<unsigned long long> acc(b.total);
    std::vector<double> pot(n, 0.0);
    std::vector<double> ljd(2 * static_cast<size_t>(std::max(t.n_lj, 0)), 0.0);
    const bool pme = c->work.pme_k[0] > 0 || p.pme_grid[0] > 0;
    bool done = false;
    for (int attempt = 0; attempt < 4 && !done; ++attempt) {
        st = prepare(c, box);
        if (st != MDGPU_OK) return st;
        Work& w = c->work;
        HIP_TRY(hipMemcpyAsync(w.coords, coords, 3 * static_cast<size_t>(n) * sizeof(double), hipMemcpyHostToDevice,
                               c->stream), "upload coords");
        c->cur_x = w.coords;
        HIP_TRY(hipMemsetAsync(b.acc, 0, b.total * sizeof(unsigned long long), c->stream), "clear adjoint");
        HIP_TRY(hipMemsetAsync(b.lj_d, 0, b.cap_lj_d * sizeof(double), c->stream), "clear adjoint");
        // the adjoint builds its own list: a Verlet list of the evaluations
        // is rebuilt at their next call
        c->verlet_force = true;
        HIP_TRY(mdg::launch_sort(c, w.coords, mixed, t.has_excl, nullptr), "sort");
        HIP_TRY(mdg::launch_pairsearch(c, mixed, true, nullptr), "pair search");
        HIP_TRY(mdg::launch_pair_adjoint(c, mixed), "pair adjoint");
        HIP_TRY(mdg::launch_corrections_adjoint(c), "corrections adjoint");
        HIP_TRY(mdg::launch_bonded_adjoint(c, w.coords, mixed), "bonded adjoint");
        if (pme) {
            // the reciprocal potential in FP64 in both forms: the FP32 grid
            // transforms leave ~1e-6 of the potential, which the charge
            // gradient (a potential, with cancellation between the direct and
            // reciprocal parts and between frames) would inherit; the mixed
            // form first sets up the FP64 spline records with a second sort
            if (mixed) HIP_TRY(mdg::launch_sort(c, w.coords, false, t.has_excl, nullptr), "sort");
            HIP_TRY(mdg::launch_pme_potential(c, false, 0, b.pot_rec), "pme potential");
        }
        unsigned int flags[mdg::kFlags];
        HIP_TRY(hipMemcpyAsync(flags, w.flags, sizeof flags, hipMemcpyDeviceToHost, c->stream), "read flags");
        HIP_TRY(hipMemcpyAsync(acc.data(), b.acc, b.total * sizeof(unsigned long long), hipMemcpyDeviceToHost,
                               c->stream), "read adjoint");
        if (pme)
            HIP_TRY(hipMemcpyAsync(pot.data(), b.pot_rec, static_cast<size_t>(n) * sizeof(double),
                                   hipMemcpyDeviceToHost, c->stream), "read potential");
        if (!mixed)
            HIP_TRY(hipMemcpyAsync(ljd.data(), b.lj_d, ljd.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream),
                    "read LJ sums");
        HIP_TRY(hipMemsetAsync(w.flags, 0, sizeof flags, c->stream), "clear flags");
        HIP_TRY(hipMemsetAsync(w.excl_rmax, 0, sizeof(float), c->stream), "clear flags");
        HIP_TRY(hipStreamSynchronize(c->stream), "adjoint");
        if (flags[4]) {
            c->search_general = true;
            continue;
        }
        if (flags[2]) return set_error(c, MDGPU_ERR_UNSUPPORTED, "a sort column holds more than 2048 atoms");
        if (flags[3]) return set_error(c, MDGPU_ERR_UNSUPPORTED, "too many tiles near excluded pairs in one cluster");
        if (flags[1]) {
            w.list_cap *= 2;
            continue;
        }
        done = true;
    }
    if (!done) return set_error(c, MDGPU_ERR_STATE, "pair list did not settle");
    if (acc[b.flag]) return set_error(c, MDGPU_ERR_OVERFLOW, "adjoint fixed-point accumulator overflow");

    // weights: each term's own weight plus the weight of the total
    mdgpu_energies wt = *weights;
    const double wtot = weights->total;
    wt.bond += wtot; wt.angle += wtot; wt.dihedral += wtot; wt.urey_bradley += wtot; wt.improper += wtot;
    wt.cmap += wtot; wt.vdw += wtot; wt.ewald_direct += wtot; wt.ewald_self += wtot; wt.ewald_recip += wtot;
    wt.ewald_excl += wtot; wt.vdw14 += wtot; wt.ee14 += wtot;
    auto val = [&](size_t k, double scale) { return static_cast<double>(static_cast<long long>(acc[k])) / scale; };
    if (out->charge) {
        const double self = -2.0 * mdg::kCoulomb * p.ewald_alpha / std::sqrt(mdg::kPi);
        for (int i = 0; i < n; ++i)
            out->charge[i] = wt.ewald_direct * val(b.q_dir + i, mdg::kAdjScaleQ) +
                             wt.ewald_excl * val(b.q_excl + i, mdg::kAdjScaleQ) +
                             wt.ee14 * val(b.q_14 + i, mdg::kAdjScaleQ) + wt.ewald_self * self * c->host_charge[i] +
                             (pme ? wt.ewald_recip * pot[i] : 0.0);
    }
    auto fill = [&](double* dst, size_t off, int cnt, int stride, double wgt, double scale) {
        if (!dst) return;
        for (int k = 0; k < cnt; ++k) dst[k] = wgt * val(off + static_cast<size_t>(stride) * k, scale);
    };
    if (mixed) {
        // two words per entry: signed high word at 2^16, remainder at 2^48
        auto two = [&](double* dst, size_t off) {
            if (!dst) return;
            for (int k = 0; k < t.n_lj; ++k)
                dst[k] = wt.vdw * (val(off + 2 * static_cast<size_t>(k), mdg::kAdjScaleLJhi) +
                                   static_cast<double>(acc[off + 2 * static_cast<size_t>(k) + 1]) / mdg::kAdjScaleLJlo);
        };
        two(out->lj_a, b.lj_a);
        two(out->lj_b, b.lj_b);
    } else {
        for (int k = 0; k < t.n_lj; ++k) {
            if (out->lj_a) out->lj_a[k] = wt.vdw * ljd[k];
            if (out->lj_b) out->lj_b[k] = wt.vdw * ljd[t.n_lj + k];
        }
    }
    fill(out->lj14_a, b.lj14_a, t.n_lj, 1, wt.vdw14, mdg::kAdjScaleLJ);
    fill(out->lj14_b, b.lj14_b, t.n_lj, 1, wt.vdw14, mdg::kAdjScaleLJ);
    size_t o = b.bonded;
    fill(out->bond_k, o, t.n_bond_types, 2, wt.bond, mdg::kAdjScaleB);
    fill(out->bond_r0, o + 1, t.n_bond_types, 2, wt.bond, mdg::kAdjScaleB);
    o += 2 * static_cast<size_t>(t.n_bond_types);
    fill(out->ub_k, o, t.n_ub_types, 2, wt.urey_bradley, mdg::kAdjScaleB);
    fill(out->ub_r0, o + 1, t.n_ub_types, 2, wt.urey_bradley, mdg::kAdjScaleB);
    o += 2 * static_cast<size_t>(t.n_ub_types);
    fill(out->angle_k, o, t.n_angle_types, 2, wt.angle, mdg::kAdjScaleB);
    fill(out->angle_theta0, o + 1, t.n_angle_types, 2, wt.angle, mdg::kAdjScaleB);
    o += 2 * static_cast<size_t>(t.n_angle_types);
    fill(out->dihedral_k, o, t.n_dih_types, 2, wt.dihedral, mdg::kAdjScaleB);
    fill(out->dihedral_phase, o + 1, t.n_dih_types, 2, wt.dihedral, mdg::kAdjScaleB);
    o += 2 * static_cast<size_t>(t.n_dih_types);
    fill(out->improper_k, o, t.n_imp_types, 2, wt.improper, mdg::kAdjScaleB);
    fill(out->improper_psi0, o + 1, t.n_imp_types, 2, wt.improper, mdg::kAdjScaleB);
    if (out->cmap_coeff) {
        const int nc = t.n_cmaps > 0 ? t.cmap_n_sets * t.cmap_res * t.cmap_res * 16 : 0;
        fill(out->cmap_coeff, b.cmap, nc, 1, wt.cmap, mdg::kAdjScaleLJ);
    }
    return MDGPU_OK;
}

int mdgpu_set_timing(mdgpu_ctx* c, int mode) {
    if (!c) return MDGPU_ERR_ARG;
    if (mode < 0 || mode > 3) return set_error(c, MDGPU_ERR_ARG, "timing mode must be 0, 1, 2 or 3");
    c->timing = mdg::Timing{};
    c->timing.mode = mode;
    return MDGPU_OK;
}

int mdgpu_get_timing(mdgpu_ctx* c, int k, double* total_ms, int64_t* launches) {
    if (!c || k < 0 || k >= MDGPU_K_COUNT) return MDGPU_ERR_ARG;
    if (total_ms) *total_ms = c->timing.total_ms[k];
    if (launches) *launches = c->timing.launches[k];
    return MDGPU_OK;
}

int mdgpu_get_counts(mdgpu_ctx* c, int64_t counts[4]) {
    if (!c || !counts) return MDGPU_ERR_ARG;
    for (int k = 0; k < 4; ++k) counts[k] = c->counts[k];
    return MDGPU_OK;
}

int mdgpu_set_graph(mdgpu_ctx* c, int enable) {
    if (!c) return MDGPU_ERR_ARG;
    c->use_graph = enable != 0;
    return MDGPU_OK;
}

int mdgpu_graph_state(const mdgpu_ctx* c) { return c ? c->graph_state : -1; }

int mdgpu_set_shapes(mdgpu_ctx* c, int split, int fft_lines) {
    if (!c) return MDGPU_ERR_ARG;
    if (split < 1 || split > 8 || fft_lines < 0 || fft_lines > 16 || (fft_lines & (fft_lines - 1)) != 0)
        return set_error(c, MDGPU_ERR_ARG, "split 1..8, fft_lines 0 or a power of two <= 16");
    if (c->graph_state == 1 && (split != c->split || fft_lines != c->fft_lines))
        return set_error(c, MDGPU_ERR_STATE, "launch shapes are fixed while the context replays its graph");
    c->split = split;
    c->fft_lines = fft_lines;
    return MDGPU_OK;
}

int mdgpu_get_shapes(const mdgpu_ctx* c, int* split, int* fft_lines) {
    if (!c) return MDGPU_ERR_ARG;
    if (split) *split = c->split;
    if (fft_lines) *fft_lines = c->fft_lines;
    return MDGPU_OK;
}

int mdgpu_tune(mdgpu_ctx* c, const double* d_coords, const double box[3], double* d_forces, int reps, int* split,
               int* fft_lines, double* ms) {
    if (!c || !d_coords || !box || reps < 1) return set_error(c, MDGPU_ERR_ARG, "bad argument");
    if (!c->have_system || !c->have_params) return set_error(c, MDGPU_ERR_STATE, "upload a system and set params first");
    if (c->graph_state != 0) return set_error(c, MDGPU_ERR_STATE, "tune before the first graph evaluation");
    // Every candidate is timed with direct launches (no capture while tuning)
    // and without timing events; both settings come back afterwards.
    const bool use_graph = c->use_graph;
    const mdg::Timing timing = c->timing;
    c->use_graph = false;
    c->timing.mode = 0;
    const int split0 = c->split, fft0 = c->fft_lines;
    const int cand_split[] = {1, 2, 4, 8};
    const int cand_fft[] = {0, 8, 4};
    constexpr int kCand = 12;
    std::vector<double> t(static_cast<size_t>(reps));
    double med[kCand];
    int st = MDGPU_OK;
    // two rounds over the candidates, each candidate's time the lower of its
    // two medians: a slow spell of the GPU or the host then hits one round
    for (int round = 0; round < 2 && st == MDGPU_OK; ++round) {
        for (int k = 0; k < kCand && st == MDGPU_OK; ++k) {
            c->split = cand_split[k / 3];
            c->fft_lines = cand_fft[k % 3];
            st = eval_impl(c, d_coords, box, nullptr, d_forces, nullptr, false);  // settles the lists
            for (int r = 0; r < reps && st == MDGPU_OK; ++r) {
                const auto t0 = std::chrono::steady_clock::now();
                st = eval_impl(c, d_coords, box, nullptr, d_forces, nullptr, false);
                t[r] = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
            }
            if (st != MDGPU_OK) break;
            // the median evaluation: robust to a stray slow one
            std::nth_element(t.begin(), t.begin() + reps / 2, t.end());
            med[k] = round == 0 ? t[reps / 2] : std::min(med[k], t[reps / 2]);
        }
    }
    c->use_graph = use_graph;
    c->timing = timing;
    if (st != MDGPU_OK) {
        c->split = split0;
        c->fft_lines = fft0;
        return st;
    }
    int best_k = 0;
    for (int k = 1; k < kCand; ++k)
        if (med[k] < med[best_k]) best_k = k;
    const int best_p = cand_split[best_k / 3], best_f = cand_fft[best_k % 3];
    const double best = med[best_k];
    c->split = best_p;
    c->fft_lines = best_f;
    if (split) *split = best_p;
    if (fft_lines) *fft_lines = best_f;
    if (ms) *ms = best;
    return MDGPU_OK;
}

int mdgpu_set_virial(mdgpu_ctx* c, int enable) {
    if (!c) return MDGPU_ERR_ARG;
    c->virial = enable != 0;
    return MDGPU_OK;
}

int mdgpu_get_virial(const mdgpu_ctx* c, double w[9]) {
    if (!c || !w) return MDGPU_ERR_ARG;
    if (!c->have_virial) return MDGPU_ERR_STATE;
    const double* v = c->last_virial;  // xx, yy, zz, xy, xz, yz
    const double t[9] = {v[0], v[3], v[4], v[3], v[1], v[5], v[4], v[5], v[2]};
    for (int k = 0; k < 9; ++k) w[k] = t[k];
    return MDGPU_OK;
}

int mdgpu_set_verlet(mdgpu_ctx* c, double skin) {
    if (!c) return MDGPU_ERR_ARG;
    if (!(skin >= 0.0) || !std::isfinite(skin)) return set_error(c, MDGPU_ERR_ARG, "skin must be finite and >= 0");
    c->skin = skin;
    c->verlet_force = true;
    ++c->gen;
    return MDGPU_OK;
}

int mdgpu_set_box_angles(mdgpu_ctx* c, const double angles[3]) {
    if (!c || !angles) return MDGPU_ERR_ARG;
    for (int d = 0; d < 3; ++d)
        if (!(angles[d] > 0.0 && angles[d] < 180.0))
            return set_error(c, MDGPU_ERR_ARG, "cell angles must lie strictly between 0 and 180 degrees");
    const double unit[3] = {1.0, 1.0, 1.0};
    mdg::Cell cell;
    double face[3];
    if (!make_cell(unit, angles, cell, face)) return set_error(c, MDGPU_ERR_ARG, "the cell angles do not form a cell");
    bool same = true;
    for (int d = 0; d < 3; ++d) same = same && c->angles[d] == angles[d];
    if (same) return MDGPU_OK;
    for (int d = 0; d < 3; ++d) c->angles[d] = angles[d];
    c->verlet_force = true;  // the kept list holds shifts of the old cell
    return MDGPU_OK;
}

int mdgpu_get_search_mode(const mdgpu_ctx* c) {
    if (!c) return -1;
    return c->search_general || !c->search_fast ? 1 : 0;
}

int mdgpu_get_box_angles(const mdgpu_ctx* c, double angles[3]) {
    if (!c || !angles) return MDGPU_ERR_ARG;
    for (int d = 0; d < 3; ++d) angles[d] = c->angles[d];
    return MDGPU_OK;
}

int mdgpu_get_list_prunes(const mdgpu_ctx* c, int64_t* prunes) {
    if (!c || !prunes) return MDGPU_ERR_ARG;
    *prunes = c->list_prunes;
    return MDGPU_OK;
}

int mdgpu_get_list_builds(const mdgpu_ctx* c, int64_t* builds, int64_t* evals) {
    if (!c || !builds || !evals) return MDGPU_ERR_ARG;
    *builds = c->list_builds;
    *evals = c->list_evals;
    return MDGPU_OK;
}

int mdgpu_pairs_need_rebuild(mdgpu_ctx* c, const double* d_coords, int* needs) {
    if (!c || !d_coords || !needs) return MDGPU_ERR_ARG;
    if (!c->have_system || !c->have_params) return set_error(c, MDGPU_ERR_STATE, "upload a system and set params first");
    // without a kept list (no skin, none built yet, or a build pending) the
    // next evaluation builds one anyway
    if (c->skin <= 0 || c->list_builds == 0 || c->verlet_force) {
        *needs = 1;
        return MDGPU_OK;
    }
    HIP_TRY(hipSetDevice(c->device), "hipSetDevice");
    HIP_TRY(mdg::launch_disp_query(c, d_coords), "displacement query");
    unsigned int q = 0;
    HIP_TRY(hipMemcpyAsync(&q, &c->work.verlet->query, sizeof q, hipMemcpyDeviceToHost, c->stream), "query copy");
    HIP_TRY(hipStreamSynchronize(c->stream), "query");
    *needs = q != 0u ? 1 : 0;
    return MDGPU_OK;
}

int mdgpu_set_counting(mdgpu_ctx* c, int enable) {
    if (!c) return MDGPU_ERR_ARG;
    c->counting = enable != 0;
    return MDGPU_OK;
}

}  // extern "C"
