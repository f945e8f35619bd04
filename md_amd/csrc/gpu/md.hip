// Molecular dynamics on the GPU: velocity Verlet in the NVE ensemble.
//
// The reference's README describes an engine that integrates Newton's
// equations, but its code stops after one evaluation (apps/main.cpp:23-28,
// with sys.minimize() commented out); it has no integrator to follow.  This
// is the textbook velocity Verlet (Swope, Andersen, Berens, Wilson, J. Chem.
// Phys. 76, 637 (1982)):
//   v(t + dt/2) = v(t) + dt/2 a(t),   x(t + dt) = x(t) + dt v(t + dt/2),
//   a(t + dt) from the evaluation at x(t + dt),
//   v(t + dt)   = v(t + dt/2) + dt/2 a(t + dt),
// with a = F / m in AMBER units (kcal/mol/A over amu is 418.4 A/ps^2).
//
// Every step runs on the device with no host synchronisation: a kick-drift
// kernel, the full evaluation (its PME chain forked as in a single
// evaluation, its Verlet-list decision made on the device), a kick kernel
// that also forms the kinetic energy in fixed-order partial sums, and a
// one-block kernel that appends the step's energy row, flags and kinetic
// energy to a trace.  The host reads the trace once at the end; a step whose
// flags ask for a redo (list over capacity, a cluster too wide for the fast
// search) restarts the whole run from saved copies with the adjusted
// context.
#include "ctx.h"

#include <cmath>
#include <cstring>
#include <vector>

namespace mdg {

namespace {

constexpr double kAccel = 418.4;  // (kcal/mol/A) / amu in A/ps^2
constexpr int kMdThreads = 256;
constexpr int kMdMaxBlocks = 1024;
constexpr int kRow = E_NSLOTS + kFlags;  // an evaluation's energies and flags
constexpr int kTraceRow = kRow + 1;      // ... and the kinetic energy

// v += dt/2 F / m, then x += dt v, per coordinate.
__global__ __launch_bounds__(kMdThreads) void k_kick_drift(double* __restrict__ x, double* __restrict__ v,
                                                           const double* __restrict__ f,
                                                           const double* __restrict__ inv_m, double hdt, double dt,
                                                           int n3) {
    for (int i = blockIdx.x * kMdThreads + threadIdx.x; i < n3; i += gridDim.x * kMdThreads) {
        const double vi = v[i] + hdt * (kAccel * f[i] * inv_m[i / 3]);
        v[i] = vi;
        x[i] = x[i] + dt * vi;
    }
}

// v += dt/2 F / m; each block stores its share of sum m v^2 / 2 (in
// kcal/mol) in a fixed order (per-thread running sums, wave tree, then the
// waves in order), so the kinetic energy is the same on every run.
__global__ __launch_bounds__(kMdThreads) void k_kick(double* __restrict__ v, const double* __restrict__ f,
                                                     const double* __restrict__ inv_m,
                                                     const double* __restrict__ mass, double hdt, int n3,
                                                     double* __restrict__ ke_part) {
    __shared__ double red[kMdThreads / kWave];
    double ke = 0.0;
    for (int i = blockIdx.x * kMdThreads + threadIdx.x; i < n3; i += gridDim.x * kMdThreads) {
        const double vi = v[i] + hdt * (kAccel * f[i] * inv_m[i / 3]);
        v[i] = vi;
        ke += mass[i / 3] * vi * vi;
    }
    ke = wave_sum(ke);
    if ((threadIdx.x & (kWave - 1)) == 0) red[threadIdx.x / kWave] = ke;
    __syncthreads();
    if (threadIdx.x == 0) {
        double s = 0.0;
        for (int q = 0; q < kMdThreads / kWave; ++q) s += red[q];
        ke_part[blockIdx.x] = s * (0.5 / kAccel);
    }
}

// Appends the step's energy row (energies then flags, as the final reduction
// left them) and the kinetic energy to the trace, at the step counter, and
// advances the counter.  One block.
__global__ __launch_bounds__(kMdThreads) void k_record(const double* __restrict__ ke_part, int nb,
                                                       const double* __restrict__ row, double* __restrict__ trace,
                                                       int* __restrict__ step, int max_steps) {
    __shared__ double red[kMdThreads / kWave];
    double s = 0.0;
    for (int b = threadIdx.x; b < nb; b += kMdThreads) s += ke_part[b];
    s = wave_sum(s);
    if ((threadIdx.x & (kWave - 1)) == 0) red[threadIdx.x / kWave] = s;
    __syncthreads();
    const int k = *step;
    if (k < max_steps) {
        double* out = trace + static_cast<size_t>(k) * kTraceRow;
        for (int t = threadIdx.x; t < kRow; t += kMdThreads) out[t] = row[t];
        if (threadIdx.x == 0) {
            double ke = 0.0;
            for (int q = 0; q < kMdThreads / kWave; ++q) ke += red[q];
            out[kRow] = ke;
        }
    }
    __syncthreads();
    if (threadIdx.x == 0) *step = k + 1;
}

int md_blocks(int n3) { return std::max(1, std::min(kMdMaxBlocks, (n3 + kMdThreads - 1) / kMdThreads)); }

}  // namespace

// Buffers of the MD path: masses, saved state, kinetic-energy partials,
// the trace and its step counter.
struct MdBuf {
    int n = 0;
    double* mass = nullptr;
    double* inv_mass = nullptr;
    double* save = nullptr;     // [9N] x, v, f at the start of a run
    double* ke_part = nullptr;  // [kMdMaxBlocks]
    int* step = nullptr;
    double* trace = nullptr;    // [cap_steps * kTraceRow]
    int cap_steps = 0;
};

}  // namespace mdg

using mdg::MdBuf;

namespace {

MdBuf* md_of(mdgpu_ctx* c) { return static_cast<MdBuf*>(c->md); }

#define MD_TRY(expr, what)                                                                \
    do {                                                                                  \
        hipError_t e_ = (expr);                                                           \
        if (e_ != hipSuccess) return mdg::ctx_error(c, MDGPU_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e_)); \
    } while (0)

}  // namespace

namespace mdg {
void md_free(mdgpu_ctx* c) {
    MdBuf* m = md_of(c);
    if (!m) return;
    void* ptrs[] = {m->mass, m->inv_mass, m->save, m->ke_part, m->step, m->trace};
    for (void* p : ptrs)
        if (p) (void)hipFree(p);
    delete m;
    c->md = nullptr;
}
}  // namespace mdg

extern "C" {

int mdgpu_md_set_masses(mdgpu_ctx* c, const double* masses) {
    if (!c || !masses) return MDGPU_ERR_ARG;
    if (!c->have_system) return mdg::ctx_error(c, MDGPU_ERR_STATE, "upload a system first");
    const int n = c->topo.n_atoms;
    std::vector<double> inv(n);
    for (int i = 0; i < n; ++i) {
        if (!(masses[i] > 0.0) || !std::isfinite(masses[i]))
            return mdg::ctx_error(c, MDGPU_ERR_ARG, "masses must be positive and finite");
        inv[i] = 1.0 / masses[i];
    }
    MD_TRY(hipSetDevice(c->device), "hipSetDevice");
    MdBuf* m = md_of(c);
    if (m && m->n != n) {
        mdg::md_free(c);
        m = nullptr;
    }
    if (!m) {
        m = new MdBuf();
        c->md = m;
        m->n = n;
        MD_TRY(hipMalloc(reinterpret_cast<void**>(&m->mass), n * sizeof(double)), "alloc");
        MD_TRY(hipMalloc(reinterpret_cast<void**>(&m->inv_mass), n * sizeof(double)), "alloc");
        MD_TRY(hipMalloc(reinterpret_cast<void**>(&m->save), 9 * static_cast<size_t>(n) * sizeof(double)), "alloc");
        MD_TRY(hipMalloc(reinterpret_cast<void**>(&m->ke_part), mdg::kMdMaxBlocks * sizeof(double)), "alloc");
        MD_TRY(hipMalloc(reinterpret_cast<void**>(&m->step), sizeof(int)), "alloc");
    }
    MD_TRY(hipMemcpy(m->mass, masses, n * sizeof(double), hipMemcpyHostToDevice), "masses");
    MD_TRY(hipMemcpy(m->inv_mass, inv.data(), n * sizeof(double), hipMemcpyHostToDevice), "masses");
    return MDGPU_OK;
}

int mdgpu_md_run(mdgpu_ctx* c, double* d_x, double* d_v, double* d_f, const double box[3], double dt, int32_t nsteps,
                 mdgpu_md_energy* trace_out, mdgpu_energies* last) {
    if (!c || !d_x || !d_v || !d_f || !box) return MDGPU_ERR_ARG;
    if (!c->have_system || !c->have_params) return mdg::ctx_error(c, MDGPU_ERR_STATE, "upload a system and set params first");
    MdBuf* m = md_of(c);
    if (!m || m->n != c->topo.n_atoms) return mdg::ctx_error(c, MDGPU_ERR_STATE, "set the masses first");
    if (!(dt > 0.0) || !std::isfinite(dt) || nsteps < 0) return mdg::ctx_error(c, MDGPU_ERR_ARG, "dt must be positive, nsteps >= 0");
    if (nsteps == 0) return MDGPU_OK;
    MD_TRY(hipSetDevice(c->device), "hipSetDevice");
    const int n3 = 3 * c->topo.n_atoms;
    const size_t bytes = static_cast<size_t>(n3) * sizeof(double);
    if (m->cap_steps < nsteps) {
        if (m->trace) (void)hipFree(m->trace);
        m->trace = nullptr;
        m->cap_steps = 0;
        MD_TRY(hipMalloc(reinterpret_cast<void**>(&m->trace),
                         static_cast<size_t>(nsteps) * mdg::kTraceRow * sizeof(double)), "alloc trace");
        m->cap_steps = nsteps;
    }
    const int nb = mdg::md_blocks(n3);
    const double hdt = 0.5 * dt;
    std::vector<double> tr(static_cast<size_t>(nsteps) * mdg::kTraceRow);
    for (int attempt = 0; attempt < 4; ++attempt) {
        int st = mdg::ctx_prepare(c, box);
        if (st != MDGPU_OK) return st;
        st = mdg::ctx_verlet_prologue(c);
        if (st != MDGPU_OK) return st;
        hipStream_t s = c->stream;
        MD_TRY(hipMemcpyAsync(m->save, d_x, bytes, hipMemcpyDeviceToDevice, s), "save");
        MD_TRY(hipMemcpyAsync(m->save + n3, d_v, bytes, hipMemcpyDeviceToDevice, s), "save");
        MD_TRY(hipMemcpyAsync(m->save + 2 * n3, d_f, bytes, hipMemcpyDeviceToDevice, s), "save");
        MD_TRY(hipMemsetAsync(m->step, 0, sizeof(int), s), "step counter");
        c->timing.record = false;
        // back to the saved entry state (the stream's earlier work included)
        auto restore = [&]() -> hipError_t {
            hipError_t e = hipStreamSynchronize(s);
            if (e == hipSuccess) e = hipMemcpyAsync(d_x, m->save, bytes, hipMemcpyDeviceToDevice, s);
            if (e == hipSuccess) e = hipMemcpyAsync(d_v, m->save + n3, bytes, hipMemcpyDeviceToDevice, s);
            if (e == hipSuccess) e = hipMemcpyAsync(d_f, m->save + 2 * n3, bytes, hipMemcpyDeviceToDevice, s);
            if (e == hipSuccess) e = hipStreamSynchronize(s);
            return e;
        };
        for (int k = 0; k < nsteps; ++k) {
            hipLaunchKernelGGL(mdg::k_kick_drift, dim3(nb), dim3(mdg::kMdThreads), 0, s, d_x, d_v, d_f, m->inv_mass,
                               hdt, dt, n3);
            st = mdg::ctx_enqueue_step(c, d_x, d_f);
            if (st == MDGPU_OK) {
                hipLaunchKernelGGL(mdg::k_kick, dim3(nb), dim3(mdg::kMdThreads), 0, s, d_v, d_f, m->inv_mass,
                                   m->mass, hdt, n3, m->ke_part);
                hipLaunchKernelGGL(mdg::k_record, dim3(1), dim3(mdg::kMdThreads), 0, s, m->ke_part, nb,
                                   c->work.energies, m->trace, m->step, nsteps);
                const hipError_t e = hipGetLastError();
                if (e != hipSuccess) st = mdg::ctx_error(c, MDGPU_ERR_HIP, std::string("md step: ") + hipGetErrorString(e));
            }
            if (st != MDGPU_OK) {
                // a step that could not be enqueued: the steps before it may
                // have moved d_x, d_v and d_f, so put the entry state back
                // (the context keeps the error message of the failing call)
                const std::string msg = c->error;
                (void)restore();
                c->verlet_force = true;
                c->error = msg;
                return st;
            }
        }
        MD_TRY(hipMemcpyAsync(tr.data(), m->trace, tr.size() * sizeof(double), hipMemcpyDeviceToHost, s), "trace");
        MD_TRY(hipStreamSynchronize(s), "md run");
        bool redo = false;
        int err = MDGPU_OK;
        for (int k = 0; k < nsteps && !redo && err == MDGPU_OK; ++k) {
            st = mdg::ctx_check(c, &tr[static_cast<size_t>(k) * mdg::kTraceRow]);
            if (st == mdg::kEvalRedo) redo = true;
            else if (st != MDGPU_OK) err = st;
        }
        if (redo || err != MDGPU_OK) {
            // back to the saved state: a redo runs again with the adjusted
            // context (which builds a new list); an error leaves the entry state
            MD_TRY(restore(), "restore");
            c->verlet_force = true;
            if (err != MDGPU_OK) return err;
            continue;
        }
        mdg::ctx_steps_done(c, tr.data(), mdg::kTraceRow, nsteps);
        // pressure (bar) from the virial when the evaluations formed it:
        // P = (2 KE + trace W) / (3 V), 1 kcal/mol/A^3 = 69476.95 bar
        const bool stats = c->counting || c->virial;
        const double vol = mdg::cell_volume(c->grid.cell);
        for (int k = 0; k < nsteps; ++k) {
            const double* row = &tr[static_cast<size_t>(k) * mdg::kTraceRow];
            mdgpu_energies e{};
            mdg::ctx_outputs(c, row, &e, nullptr);
            if (trace_out) {
                trace_out[k].potential = e.total;
                trace_out[k].kinetic = row[mdg::kRow];
                double w = 0.0;
                for (int d = 0; d < 3; ++d) w += row[mdg::E_VIR_XX + d] + row[mdg::E_VSR_XX + d];
                trace_out[k].pressure =
                    stats ? (2.0 * row[mdg::kRow] + w) / (3.0 * vol) * 69476.95 : std::nan("");
            }
            if (last && k == nsteps - 1) *last = e;
        }
        return MDGPU_OK;
    }
    return mdg::ctx_error(c, MDGPU_ERR_STATE, "md run: repeated redo");
}

}  // extern "C"
