// Energy minimisation on the GPU: limited-memory BFGS.
//
// The reference leaves sys.minimize() commented out (apps/main.cpp:28) and
// names L-BFGS as the solver its force routine is meant for (System.h:84,
// 113); it has no implementation to follow.  This is L-BFGS as in Nocedal &
// Wright, Numerical Optimization (2nd ed.), Algorithm 7.4 (two-loop
// recursion) and 7.5, with a backtracking Armijo line search:
//   p = -H g (two-loop over the last m pairs s = x+ - x, y = g+ - g, with
//   H0 = (s.y / y.y) I, steepest descent while no pair is kept),
//   a step of at most max_step A for any atom, halved until
//   E(x + a p) <= E(x) + 1e-4 a g.p,
//   a pair kept only if s.y > 0 (so H stays positive definite).
// g = -F.  Every vector operation runs on the device: dot products are
// deterministic (fixed per-block partial sums, the last block adds them in
// order) and leave their results in a small device array that the update
// kernels read, so the two-loop recursion needs no host round trip.  The host
// reads back a handful of scalars per iteration (the energy, g.p, the step
// bound, the force norms) to steer the line search and the stopping test.
#include "ctx.h"

#include <algorithm>
#include <cmath>
#include <vector>

namespace mdg {

namespace {

constexpr int kMinThreads = 256;
constexpr int kMinBlocks = 240;
constexpr int kMaxHistory = 16;

// scalar slots of the device array
constexpr int kRho = 0;                     // [m] 1 / (y_i . s_i)
constexpr int kAlpha = kMaxHistory;         // [m] two-loop alpha_i
constexpr int kTmp = 2 * kMaxHistory;       // temporaries
constexpr int kNumScalars = 2 * kMaxHistory + 8;
constexpr int kGp = kTmp + 0;        // g . p
constexpr int kPmax2 = kTmp + 1;     // max over atoms of |p_atom|^2
constexpr int kF2 = kTmp + 2;        // F . F
constexpr int kFmax2 = kTmp + 3;     // max over atoms of |F_atom|^2
constexpr int kYs = kTmp + 4;        // y . s of the newest pair
constexpr int kYy = kTmp + 5;        // y . y of the newest pair
constexpr int kBeta = kTmp + 6;      // two-loop beta

// sum or max of a per-element value over a vector, deterministic: each
// block stores its share, the last block to finish combines them in order
// and writes sc[out] (times sc[scale] when scale >= 0), then clears the
// ticket.  op 0: sum a[i] b[i] over i < n; op 1: max over atoms (n = atoms,
// vectors of 3 n) of |a_atom|^2.
__global__ __launch_bounds__(kMinThreads) void k_reduce_vec(const double* __restrict__ a,
                                                            const double* __restrict__ b, int n, int op,
                                                            double* __restrict__ part, unsigned int* ticket,
                                                            double* __restrict__ sc, int out, int scale) {
    __shared__ double red[kMinThreads / kWave];
    __shared__ bool last;
    double v = 0.0;
    for (int i = blockIdx.x * kMinThreads + threadIdx.x; i < n; i += gridDim.x * kMinThreads) {
        if (op == 0) {
            v += a[i] * b[i];
        } else {
            const double t = a[3 * i] * a[3 * i] + a[3 * i + 1] * a[3 * i + 1] + a[3 * i + 2] * a[3 * i + 2];
            v = fmax(v, t);
        }
    }
    for (int m = 32; m >= 1; m >>= 1) {
        const double o = __shfl_xor(v, m, kWave);
        v = op == 0 ? v + o : fmax(v, o);
    }
    if ((threadIdx.x & (kWave - 1)) == 0) red[threadIdx.x / kWave] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        double s = red[0];
        for (int q = 1; q < kMinThreads / kWave; ++q) s = op == 0 ? s + red[q] : fmax(s, red[q]);
        part[blockIdx.x] = s;
        __threadfence();
        last = atomicAdd(ticket, 1u) == gridDim.x - 1;
    }
    __syncthreads();
    if (!last || threadIdx.x != 0) return;
    __threadfence();
    double s = 0.0;
    for (int k = 0; k < static_cast<int>(gridDim.x); ++k) {
        const double pk = __hip_atomic_load(part + k, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        s = op == 0 ? s + pk : fmax(s, pk);
    }
    sc[out] = scale >= 0 ? s * sc[scale] : s;
    atomicExch(ticket, 0u);
}

// y[i] = cy y[i] + c x[i] with c = ca sc[ia] - cb sc[ib] (a missing index
// counts as 0) or, with inv_ia >= 0, c = ca / (sc[inv_ia] sc[inv_ib]).
struct Coef {
    double cy;
    double ca;
    int ia;
    double cb;
    int ib;
    int inv_ia, inv_ib;
};
__global__ __launch_bounds__(kMinThreads) void k_update(double* __restrict__ y, const double* __restrict__ x, int n,
                                                        const double* __restrict__ sc, const Coef k) {
    double c;
    if (k.inv_ia >= 0) {
        c = k.ca / (sc[k.inv_ia] * sc[k.inv_ib]);
    } else {
        c = (k.ia >= 0 ? k.ca * sc[k.ia] : 0.0) - (k.ib >= 0 ? k.cb * sc[k.ib] : 0.0);
    }
    for (int i = blockIdx.x * kMinThreads + threadIdx.x; i < n; i += gridDim.x * kMinThreads)
        y[i] = k.cy * y[i] + c * x[i];
}

// z[i] = a[i] + t b[i], or t b[i] without a (trial point, step and gradient
// differences)
__global__ __launch_bounds__(kMinThreads) void k_axpy(double* __restrict__ z, const double* __restrict__ a,
                                                      const double* __restrict__ b, double t, int n) {
    for (int i = blockIdx.x * kMinThreads + threadIdx.x; i < n; i += gridDim.x * kMinThreads)
        z[i] = a ? a[i] + t * b[i] : t * b[i];
}

// sc[dst] = 1 / sc[src]
__global__ void k_recip(double* sc, int dst, int src) {
    if (threadIdx.x == 0 && blockIdx.x == 0) sc[dst] = 1.0 / sc[src];
}

int min_blocks(int n) { return std::max(1, std::min(kMinBlocks, (n + kMinThreads - 1) / kMinThreads)); }

}  // namespace

struct MinBuf {
    int n3 = 0, m = 0;
    double* x_t = nullptr;   // trial point
    double* f = nullptr;     // forces at x
    double* f_t = nullptr;   // forces at the trial point
    double* p = nullptr;     // search direction
    double* q = nullptr;     // two-loop vector
    double* s = nullptr;     // [m][n3]
    double* y = nullptr;     // [m][n3]
    double* sc = nullptr;    // [kNumScalars]
    double* part = nullptr;  // [kMinBlocks]
    unsigned int* ticket = nullptr;
    double* x_own = nullptr; // coordinates of host-buffer minimisations
};

void min_free(mdgpu_ctx* c) {
    MinBuf* b = static_cast<MinBuf*>(c->minbuf);
    if (!b) return;
    void* ptrs[] = {b->x_t, b->f, b->f_t, b->p, b->q, b->s, b->y, b->sc, b->part, b->ticket, b->x_own};
    for (void* p : ptrs)
        if (p) (void)hipFree(p);
    delete b;
    c->minbuf = nullptr;
}

}  // namespace mdg

namespace {

#define MIN_TRY(expr, what)                                                                                 \
    do {                                                                                                    \
        hipError_t e_ = (expr);                                                                             \
        if (e_ != hipSuccess) return mdg::ctx_error(c, MDGPU_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e_)); \
    } while (0)

int ensure_min(mdgpu_ctx* c, int m) {
    const int n3 = 3 * c->topo.n_atoms;
    auto* b = static_cast<mdg::MinBuf*>(c->minbuf);
    if (b && b->n3 == n3 && b->m >= m) return MDGPU_OK;
    mdg::min_free(c);
    b = new mdg::MinBuf();
    c->minbuf = b;
    b->n3 = n3;
    b->m = m;
    const size_t v = static_cast<size_t>(n3) * sizeof(double);
    MIN_TRY(hipMalloc(reinterpret_cast<void**>(&b->x_t), v), "alloc");
    MIN_TRY(hipMalloc(reinterpret_cast<void**>(&b->f), v), "alloc");
    MIN_TRY(hipMalloc(reinterpret_cast<void**>(&b->f_t), v), "alloc");
    MIN_TRY(hipMalloc(reinterpret_cast<void**>(&b->p), v), "alloc");
    MIN_TRY(hipMalloc(reinterpret_cast<void**>(&b->q), v), "alloc");
    MIN_TRY(hipMalloc(reinterpret_cast<void**>(&b->s), m * v), "alloc");
    MIN_TRY(hipMalloc(reinterpret_cast<void**>(&b->y), m * v), "alloc");
    MIN_TRY(hipMalloc(reinterpret_cast<void**>(&b->sc), mdg::kNumScalars * sizeof(double)), "alloc");
    MIN_TRY(hipMalloc(reinterpret_cast<void**>(&b->part), mdg::kMinBlocks * sizeof(double)), "alloc");
    MIN_TRY(hipMalloc(reinterpret_cast<void**>(&b->ticket), sizeof(unsigned int)), "alloc");
    MIN_TRY(hipMalloc(reinterpret_cast<void**>(&b->x_own), v), "alloc");
    MIN_TRY(hipMemset(b->ticket, 0, sizeof(unsigned int)), "clear");
    MIN_TRY(hipMemset(b->sc, 0, mdg::kNumScalars * sizeof(double)), "clear");
    return MDGPU_OK;
}

}  // namespace

extern "C" {

int mdgpu_minimize(mdgpu_ctx* c, double* d_x, const double box[3], const mdgpu_min_opts* opts,
                   mdgpu_min_result* res) {
    using namespace mdg;
    if (!c || !d_x || !box) return MDGPU_ERR_ARG;
    if (!c->have_system || !c->have_params) return ctx_error(c, MDGPU_ERR_STATE, "upload a system and set params first");
    mdgpu_min_opts o{500, 1e-2, 0.2, 8};
    if (opts) o = *opts;
    if (o.max_iter < 0 || !(o.rms_force_tol >= 0.0) || !(o.max_step > 0.0) || o.history < 1 ||
        o.history > kMaxHistory)
        return ctx_error(c, MDGPU_ERR_ARG, "minimize: max_iter >= 0, tol >= 0, max_step > 0, 1 <= history <= 16");
    MIN_TRY(hipSetDevice(c->device), "hipSetDevice");
    int st = ensure_min(c, o.history);
    if (st != MDGPU_OK) return st;
    auto* b = static_cast<MinBuf*>(c->minbuf);
    const int n = c->topo.n_atoms, n3 = 3 * n, m = o.history;
    const int nb = min_blocks(n3);
    const size_t vbytes = static_cast<size_t>(n3) * sizeof(double);
    hipStream_t s = c->stream;
    mdgpu_min_result r{};
    mdgpu_energies e{};
    auto reduce = [&](const double* a, const double* bb, int len, int op, int out, int scale) {
        hipLaunchKernelGGL(k_reduce_vec, dim3(min_blocks(len)), dim3(kMinThreads), 0, s, a, bb, len, op, b->part,
                           b->ticket, b->sc, out, scale);
    };
    auto update = [&](double* yv, const double* xv, const Coef& k) {
        hipLaunchKernelGGL(k_update, dim3(nb), dim3(kMinThreads), 0, s, yv, xv, n3, b->sc, k);
    };
    auto scalars = [&](double* h) -> int {
        MIN_TRY(hipMemcpyAsync(h, b->sc, kNumScalars * sizeof(double), hipMemcpyDeviceToHost, s), "scalars");
        MIN_TRY(hipStreamSynchronize(s), "scalars");
        return MDGPU_OK;
    };
    st = ctx_eval_device(c, d_x, box, &e, b->f);
    if (st != MDGPU_OK) return st;
    r.evaluations = 1;
    double energy = e.total;
    int kept = 0, newest = -1;  // pairs in the ring, slot of the newest
    std::vector<double> h(kNumScalars);
    for (int it = 0;; ++it) {
        // force norms for the stopping test
        reduce(b->f, b->f, n3, 0, kF2, -1);
        reduce(b->f, nullptr, n, 1, kFmax2, -1);
        if ((st = scalars(h.data())) != MDGPU_OK) return st;
        r.rms_force = std::sqrt(h[kF2] / n3);
        r.max_force = std::sqrt(h[kFmax2]);
        r.energy = energy;
        r.iterations = it;
        if (r.rms_force <= o.rms_force_tol) {
            r.converged = 1;
            break;
        }
        if (it >= o.max_iter) break;
        // direction: two-loop recursion over the kept pairs, newest first
        MIN_TRY(hipMemcpyAsync(b->q, b->f, vbytes, hipMemcpyDeviceToDevice, s), "q");  // q = -g = F
        for (int k = 0; k < kept; ++k) {
            const int i = (newest - k + m) % m;
            reduce(b->s + static_cast<size_t>(i) * n3, b->q, n3, 0, kAlpha + i, kRho + i);  // alpha_i = rho s.q
            // q and r carry F = -g, so the alphas and betas come out negated
            // and the updates keep the signs of the textbook recursion
            update(b->q, b->y + static_cast<size_t>(i) * n3, Coef{1.0, 0.0, -1, 1.0, kAlpha + i, -1, -1});
        }
        // r = H0 q, H0 = (s.y / y.y) I = 1 / (rho y.y)
        if (kept > 0) {
            MIN_TRY(hipMemcpyAsync(b->p, b->q, vbytes, hipMemcpyDeviceToDevice, s), "p");
            update(b->p, b->q, Coef{0.0, 1.0, -1, 0.0, -1, kRho + newest, kYy});
            for (int k = kept - 1; k >= 0; --k) {
                const int i = (newest - k + m) % m;
                reduce(b->y + static_cast<size_t>(i) * n3, b->p, n3, 0, kBeta, kRho + i);  // beta = rho y.r
                // r += s (alpha - beta); with r = -H g this is the direction p
                update(b->p, b->s + static_cast<size_t>(i) * n3, Coef{1.0, 1.0, kAlpha + i, 1.0, kBeta, -1, -1});
            }
        } else {
            MIN_TRY(hipMemcpyAsync(b->p, b->q, vbytes, hipMemcpyDeviceToDevice, s), "p");
        }
        // g.p = -F.p and the largest atom step
        reduce(b->f, b->p, n3, 0, kGp, -1);
        reduce(b->p, nullptr, n, 1, kPmax2, -1);
        if ((st = scalars(h.data())) != MDGPU_OK) return st;
        double gp = -h[kGp], pmax = std::sqrt(h[kPmax2]);
        if (!(gp < 0.0)) {
            // not a descent direction: forget the pairs, steepest descent
            kept = 0;
            MIN_TRY(hipMemcpyAsync(b->p, b->f, vbytes, hipMemcpyDeviceToDevice, s), "p");
            gp = -h[kF2];
            pmax = std::sqrt(h[kFmax2]);
        }
        double a = 1.0;
        if (a * pmax > o.max_step) a = o.max_step / pmax;
        bool accepted = false;
        double e_t = 0.0;
        for (int ls = 0; ls < 30; ++ls) {
            hipLaunchKernelGGL(k_axpy, dim3(nb), dim3(kMinThreads), 0, s, b->x_t, d_x, b->p, a, n3);
            st = ctx_eval_device(c, b->x_t, box, &e, b->f_t);
            ++r.evaluations;
            if (st == MDGPU_ERR_OVERFLOW) {
                // a force beyond the fixed-point range at the trial point: an
                // energy too high to accept, so the step is halved
                a *= 0.5;
                continue;
            }
            if (st != MDGPU_OK) return st;
            e_t = e.total;
            if (e_t <= energy + 1e-4 * a * gp) {
                accepted = true;
                break;
            }
            a *= 0.5;
        }
        if (!accepted) break;  // no decrease along p at any step tried: stop here
        // new pair: s = a p, y = g_t - g = F - F_t
        const int slot = (newest + 1) % m;
        double* sv = b->s + static_cast<size_t>(slot) * n3;
        double* yv = b->y + static_cast<size_t>(slot) * n3;
        hipLaunchKernelGGL(k_axpy, dim3(nb), dim3(kMinThreads), 0, s, sv, nullptr, b->p, a, n3);
        hipLaunchKernelGGL(k_axpy, dim3(nb), dim3(kMinThreads), 0, s, yv, b->f, b->f_t, -1.0, n3);
        reduce(yv, sv, n3, 0, kYs, -1);
        MIN_TRY(hipMemcpyAsync(d_x, b->x_t, vbytes, hipMemcpyDeviceToDevice, s), "x");
        MIN_TRY(hipMemcpyAsync(b->f, b->f_t, vbytes, hipMemcpyDeviceToDevice, s), "f");
        energy = e_t;
        if ((st = scalars(h.data())) != MDGPU_OK) return st;
        if (h[kYs] > 0.0) {
            hipLaunchKernelGGL(k_recip, dim3(1), dim3(1), 0, s, b->sc, kRho + slot, kYs);
            reduce(yv, yv, n3, 0, kYy, -1);
            newest = slot;
            kept = std::min(kept + 1, m);
        }
    }
    MIN_TRY(hipStreamSynchronize(s), "minimize");
    if (res) *res = r;
    return MDGPU_OK;
}

int mdgpu_minimize_host(mdgpu_ctx* c, double* x, const double box[3], const mdgpu_min_opts* opts,
                        mdgpu_min_result* res) {
    using namespace mdg;
    if (!c || !x || !box) return MDGPU_ERR_ARG;
    if (!c->have_system) return ctx_error(c, MDGPU_ERR_STATE, "upload a system first");
    MIN_TRY(hipSetDevice(c->device), "hipSetDevice");
    const int m = opts ? opts->history : 8;
    int st = ensure_min(c, m >= 1 && m <= kMaxHistory ? m : 8);
    if (st != MDGPU_OK) return st;
    auto* b = static_cast<MinBuf*>(c->minbuf);
    const size_t bytes = static_cast<size_t>(b->n3) * sizeof(double);
    MIN_TRY(hipMemcpy(b->x_own, x, bytes, hipMemcpyHostToDevice), "coordinates in");
    st = mdgpu_minimize(c, b->x_own, box, opts, res);
    if (st != MDGPU_OK) return st;
    MIN_TRY(hipMemcpy(x, b->x_own, bytes, hipMemcpyDeviceToHost), "coordinates out");
    return MDGPU_OK;
}

}  // extern "C"
