// Batched evaluation: many independent systems, one hipGraph launch.
//
// The reference evaluates one system per process (apps/main.cpp:23-27).  A
// batch keeps one context per system (its own topology, parameters, box and
// work buffers, all allocated before capture) and captures every system's
// kernel sequence once into a graph whose branches run on a few concurrent
// streams: systems are dealt to the streams largest first, each to the
// least-loaded stream (atoms as the load).  A last kernel gathers every
// system's energy row and flags into one buffer, read back with one copy.
// Replays need no host work per system.  A system whose flags ask for a redo
// (pair list over capacity, cluster too wide for the minimum-image search) is
// evaluated again on its own, and from then on the batch issues the graph's
// launches directly on the same streams (mode 2), without a graph.
//
// Each system keeps its own pair-tile split and FFT block shape.  The
// autotuner sets them per size class (systems ordered by atom count, cut into
// up to four classes of equal count), timing for each candidate the same
// launches on the same streams as the graph, issued directly (mode 2).
#include "ctx.h"

#include <array>
#include <chrono>
#include <numeric>

namespace {

constexpr int kMaxStreams = 8;
constexpr int kRow = mdg::E_NSLOTS + mdg::kFlags;  // energies + flags of one system

__global__ void k_collect(const double* const* __restrict__ src, double* __restrict__ dst, int n_sys) {
    const int s = blockIdx.x;
    if (s >= n_sys) return;
    for (int k = threadIdx.x; k < kRow; k += blockDim.x) dst[static_cast<size_t>(s) * kRow + k] = src[s][k];
}

}  // namespace

struct mdgpu_batch {
    int device = 0;
    std::string error;
    hipStream_t s0 = nullptr;
    hipStream_t workers[kMaxStreams] = {};
    hipEvent_t fork = nullptr;
    hipEvent_t join[kMaxStreams] = {};
    hipEvent_t t0 = nullptr, t1 = nullptr;
    std::vector<mdgpu_ctx*> sys;
    std::vector<std::array<double, 3>> box;
    std::vector<int64_t> off;
    int64_t n_atoms = 0;
    int mode = 0, streams = 4, split = 4, fft_lines = 8;  // split and fft_lines as a single context
    std::vector<int> sys_split, sys_fft;                  // per system (set by configure or the tuner)
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    const double* cap_x = nullptr;
    double* cap_f = nullptr;
    bool dirty = true;
    // The graph is captured once per batch.  Capturing a second time on the
    // same batch (after destroying the first graph) ended the process with a
    // host segmentation fault on the MI355X and the cause is not found, so a
    // batch whose captured graph went stale issues the same launches on the
    // same streams directly instead (mode 2).
    int graph_state = 0;              // 0 none yet, 1 captured and used, 2 stale: direct launches from now on
    const double** d_src = nullptr;   // per-system energy rows (device pointers)
    double* d_all = nullptr;          // [n_sys x kRow]
    double* h_all = nullptr;          // pinned copy
    size_t cap_sys = 0;
    double* d_x = nullptr;            // host-path staging
    double* d_f = nullptr;
    int64_t cap_atoms = 0;
};

namespace {

int berr(mdgpu_batch* b, int code, const std::string& msg) {
    if (b) b->error = msg;
    return code;
}

#define B_TRY(expr, what)                                                                       \
    do {                                                                                        \
        hipError_t e_ = (expr);                                                                 \
        if (e_ != hipSuccess) return berr(b, MDGPU_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e_)); \
    } while (0)

void drop_graph(mdgpu_batch* b) {
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    b->exec = nullptr;
    b->graph = nullptr;
    b->dirty = true;
}

int ensure_rows(mdgpu_batch* b) {
    const size_t n = b->sys.size();
    if (b->cap_sys >= n) return MDGPU_OK;
    if (b->d_src) (void)hipFree(b->d_src);
    if (b->d_all) (void)hipFree(b->d_all);
    if (b->h_all) (void)hipHostFree(b->h_all);
    b->d_src = nullptr;
    b->d_all = nullptr;
    b->h_all = nullptr;
    b->cap_sys = 0;
    B_TRY(hipMalloc(reinterpret_cast<void**>(&b->d_src), n * sizeof(double*)), "alloc batch rows");
    B_TRY(hipMalloc(reinterpret_cast<void**>(&b->d_all), n * kRow * sizeof(double)), "alloc batch rows");
    B_TRY(hipHostMalloc(reinterpret_cast<void**>(&b->h_all), n * kRow * sizeof(double), hipHostMallocDefault),
          "alloc batch rows");
    b->cap_sys = n;
    return MDGPU_OK;
}

// The launch layout of a multi-stream pass: every system prepared with its
// own settings, the row table uploaded, and the systems in issue order
// (largest first) with their streams (each to the least-loaded stream).
int layout(mdgpu_batch* b, std::vector<int>& order, std::vector<int>& stream_of, int& ns) {
    const int n = static_cast<int>(b->sys.size());
    int st = ensure_rows(b);
    if (st != MDGPU_OK) return st;
    std::vector<const double*> src(n);
    for (int i = 0; i < n; ++i) {
        mdgpu_ctx* c = b->sys[i];
        c->split = b->sys_split[i];
        c->fft_lines = b->sys_fft[i];
        c->timing.mode = 0;
        st = mdg::ctx_prepare(c, b->box[i].data());
        if (st != MDGPU_OK) return berr(b, st, "system " + std::to_string(i) + ": " + c->error);
        src[i] = c->work.energies;
    }
    B_TRY(hipMemcpy(b->d_src, src.data(), n * sizeof(double*), hipMemcpyHostToDevice), "upload row table");
    order.resize(n);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(),
                     [&](int a, int c) { return b->sys[a]->topo.n_atoms > b->sys[c]->topo.n_atoms; });
    // never more streams than systems: every branch of a graph holds kernels
    ns = std::max(1, std::min({b->streams, kMaxStreams, n}));
    std::vector<int64_t> load(ns, 0);
    stream_of.assign(n, 0);
    for (int i : order) {
        const int k = static_cast<int>(std::min_element(load.begin(), load.end()) - load.begin());
        load[k] += b->sys[i]->topo.n_atoms;
        stream_of[i] = k;
    }
    return MDGPU_OK;
}

// Acts on the gathered rows of a finished multi-stream pass: outputs, or a
// redo alone for a system whose flags ask for one.
int finish_rows(mdgpu_batch* b, const double* x, mdgpu_energies* out_e, double* f) {
    for (size_t i = 0; i < b->sys.size(); ++i) {
        mdgpu_ctx* c = b->sys[i];
        const double* row = b->h_all + i * kRow;
        int st = mdg::ctx_check(c, row);
        if (st == mdg::kEvalRedo) {
            // adjusted (larger list or general search): redo alone; a graph is stale
            b->dirty = true;
            st = mdg::ctx_eval_device(c, x + 3 * b->off[i], b->box[i].data(), out_e ? &out_e[i] : nullptr,
                                      f ? f + 3 * b->off[i] : nullptr);
            if (st != MDGPU_OK) return berr(b, st, "system " + std::to_string(i) + ": " + c->error);
            continue;
        }
        if (st != MDGPU_OK) return berr(b, st, "system " + std::to_string(i) + ": " + c->error);
        mdg::ctx_outputs(c, row, out_e ? &out_e[i] : nullptr, nullptr);
    }
    return MDGPU_OK;
}

// One pass with the graph's launches and streams, issued directly (mode 2).
int run_streams(mdgpu_batch* b, const double* x, mdgpu_energies* out_e, double* f) {
    std::vector<int> order, stream_of;
    int ns = 1;
    int st = layout(b, order, stream_of, ns);
    if (st != MDGPU_OK) return st;
    const int n = static_cast<int>(b->sys.size());
    B_TRY(hipEventRecord(b->fork, b->s0), "fork");
    for (int k = 0; k < ns; ++k) B_TRY(hipStreamWaitEvent(b->workers[k], b->fork, 0), "fork");
    for (int i : order) {
        mdgpu_ctx* c = b->sys[i];
        hipStream_t own = c->stream;
        c->stream = b->workers[stream_of[i]];
        st = mdg::ctx_enqueue(c, x + 3 * b->off[i], f ? f + 3 * b->off[i] : nullptr);
        c->stream = own;
        if (st != MDGPU_OK) {
            (void)hipDeviceSynchronize();
            return berr(b, st, "system " + std::to_string(i) + ": " + c->error);
        }
    }
    for (int k = 0; k < ns; ++k) {
        B_TRY(hipEventRecord(b->join[k], b->workers[k]), "join");
        B_TRY(hipStreamWaitEvent(b->s0, b->join[k], 0), "join");
    }
    hipLaunchKernelGGL(k_collect, dim3(n), dim3(64), 0, b->s0, b->d_src, b->d_all, n);
    B_TRY(hipGetLastError(), "collect");
    B_TRY(hipMemcpyAsync(b->h_all, b->d_all, n * kRow * sizeof(double), hipMemcpyDeviceToHost, b->s0), "read rows");
    B_TRY(hipStreamSynchronize(b->s0), "batch");
    return finish_rows(b, x, out_e, f);
}

// Captures every system's evaluation on x / f into one graph.
int capture(mdgpu_batch* b, const double* x, double* f) {
    drop_graph(b);
    const int n = static_cast<int>(b->sys.size());
    std::vector<int> order, stream_of;
    int ns = 1;
    int st = layout(b, order, stream_of, ns);
    if (st != MDGPU_OK) return st;
    B_TRY(hipStreamBeginCapture(b->s0, hipStreamCaptureModeThreadLocal), "begin capture");
    auto abort_capture = [&](int code, const std::string& msg) {
        hipGraph_t g = nullptr;
        (void)hipStreamEndCapture(b->s0, &g);
        if (g) (void)hipGraphDestroy(g);
        (void)hipGetLastError();
        return berr(b, code, msg);
    };
    if (hipEventRecord(b->fork, b->s0) != hipSuccess) return abort_capture(MDGPU_ERR_HIP, "fork");
    for (int k = 0; k < ns; ++k)
        if (hipStreamWaitEvent(b->workers[k], b->fork, 0) != hipSuccess) return abort_capture(MDGPU_ERR_HIP, "fork");
    for (int i : order) {
        mdgpu_ctx* c = b->sys[i];
        hipStream_t own = c->stream;
        c->stream = b->workers[stream_of[i]];
        st = mdg::ctx_enqueue(c, x + 3 * b->off[i], f ? f + 3 * b->off[i] : nullptr);
        c->stream = own;
        if (st != MDGPU_OK) return abort_capture(st, "system " + std::to_string(i) + ": " + c->error);
    }
    for (int k = 0; k < ns; ++k) {
        if (hipEventRecord(b->join[k], b->workers[k]) != hipSuccess) return abort_capture(MDGPU_ERR_HIP, "join");
        if (hipStreamWaitEvent(b->s0, b->join[k], 0) != hipSuccess) return abort_capture(MDGPU_ERR_HIP, "join");
    }
    hipLaunchKernelGGL(k_collect, dim3(n), dim3(64), 0, b->s0, b->d_src, b->d_all, n);
    if (hipGetLastError() != hipSuccess) return abort_capture(MDGPU_ERR_HIP, "collect");
    B_TRY(hipStreamEndCapture(b->s0, &b->graph), "end capture");
    B_TRY(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0), "instantiate graph");
    b->cap_x = x;
    b->cap_f = f;
    b->dirty = false;
    return MDGPU_OK;
}

int run_serial(mdgpu_batch* b, const double* x, mdgpu_energies* out_e, double* f) {
    for (size_t i = 0; i < b->sys.size(); ++i) {
        mdgpu_ctx* c = b->sys[i];
        c->split = b->sys_split[i];
        c->fft_lines = b->sys_fft[i];
        const int st = mdg::ctx_eval_device(c, x + 3 * b->off[i], b->box[i].data(), out_e ? &out_e[i] : nullptr,
                                            f ? f + 3 * b->off[i] : nullptr);
        if (st != MDGPU_OK) return berr(b, st, "system " + std::to_string(i) + ": " + c->error);
    }
    return MDGPU_OK;
}

int run_graph(mdgpu_batch* b, const double* x, mdgpu_energies* out_e, double* f) {
    if (b->graph_state == 1 && (b->dirty || !b->exec || x != b->cap_x || f != b->cap_f)) b->graph_state = 2;
    if (b->graph_state == 2) return run_streams(b, x, out_e, f);
    if (b->graph_state == 0) {
        // one serial pass sizes every pair list before the single capture
        int st = run_serial(b, x, nullptr, f);
        if (st == MDGPU_OK) st = capture(b, x, f);
        if (st != MDGPU_OK) return st;
        b->graph_state = 1;
    }
    const size_t n = b->sys.size();
    B_TRY(hipGraphLaunch(b->exec, b->s0), "graph launch");
    B_TRY(hipMemcpyAsync(b->h_all, b->d_all, n * kRow * sizeof(double), hipMemcpyDeviceToHost, b->s0), "read rows");
    B_TRY(hipStreamSynchronize(b->s0), "batch");
    return finish_rows(b, x, out_e, f);
}

}  // namespace

extern "C" {

int mdgpu_batch_create(int device, mdgpu_batch** out) {
    if (!out) return MDGPU_ERR_ARG;
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || device < 0 || device >= ndev) return MDGPU_ERR_ARG;
    auto* b = new (std::nothrow) mdgpu_batch();
    if (!b) return MDGPU_ERR_HIP;
    b->device = device;
    hipError_t e = hipSetDevice(device);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&b->s0, hipStreamNonBlocking);
    for (int k = 0; k < kMaxStreams && e == hipSuccess; ++k) {
        e = hipStreamCreateWithFlags(&b->workers[k], hipStreamNonBlocking);
        if (e == hipSuccess) e = hipEventCreateWithFlags(&b->join[k], hipEventDisableTiming);
    }
    if (e == hipSuccess) e = hipEventCreateWithFlags(&b->fork, hipEventDisableTiming);
    if (e == hipSuccess) e = hipEventCreate(&b->t0);
    if (e == hipSuccess) e = hipEventCreate(&b->t1);
    if (e != hipSuccess) {
        mdgpu_batch_destroy(b);
        return MDGPU_ERR_HIP;
    }
    *out = b;
    return MDGPU_OK;
}

int mdgpu_batch_destroy(mdgpu_batch* b) {
    if (!b) return MDGPU_OK;
    (void)hipSetDevice(b->device);
    if (b->s0) (void)hipStreamSynchronize(b->s0);
    drop_graph(b);
    for (mdgpu_ctx* c : b->sys) mdgpu_destroy(c);
    if (b->d_src) (void)hipFree(b->d_src);
    if (b->d_all) (void)hipFree(b->d_all);
    if (b->h_all) (void)hipHostFree(b->h_all);
    if (b->d_x) (void)hipFree(b->d_x);
    if (b->d_f) (void)hipFree(b->d_f);
    for (int k = 0; k < kMaxStreams; ++k) {
        if (b->workers[k]) (void)hipStreamDestroy(b->workers[k]);
        if (b->join[k]) (void)hipEventDestroy(b->join[k]);
    }
    if (b->fork) (void)hipEventDestroy(b->fork);
    if (b->t0) (void)hipEventDestroy(b->t0);
    if (b->t1) (void)hipEventDestroy(b->t1);
    if (b->s0) (void)hipStreamDestroy(b->s0);
    delete b;
    return MDGPU_OK;
}

const char* mdgpu_batch_last_error(const mdgpu_batch* b) { return b ? b->error.c_str() : "null batch"; }

int mdgpu_batch_add(mdgpu_batch* b, const mdgpu_system_desc* sys, const mdgpu_params* p, const double box[3],
                    int* index) {
    if (!b || !sys || !p || !box) return berr(b, MDGPU_ERR_ARG, "null argument");
    mdgpu_ctx* c = nullptr;
    int st = mdgpu_create(b->device, &c);
    if (st != MDGPU_OK) return berr(b, st, "mdgpu_create failed");
    // systems of a batch overlap each other on the batch's streams; they do
    // not fork streams of their own
    c->fork = false;
    st = mdgpu_upload_system(c, sys);
    if (st == MDGPU_OK) st = mdgpu_set_params(c, p);
    if (st == MDGPU_OK) {
        c->split = b->split;
        c->fft_lines = b->fft_lines;
        st = mdg::ctx_prepare(c, box);
    }
    if (st != MDGPU_OK) {
        const std::string msg = c->error;
        mdgpu_destroy(c);
        return berr(b, st, "system " + std::to_string(b->sys.size()) + ": " + msg);
    }
    if (index) *index = static_cast<int>(b->sys.size());
    b->sys.push_back(c);
    b->sys_split.push_back(b->split);
    b->sys_fft.push_back(b->fft_lines);
    b->box.push_back({box[0], box[1], box[2]});
    b->off.push_back(b->n_atoms);
    b->n_atoms += sys->n_atoms;
    if (b->graph_state == 1) b->graph_state = 2;  // the captured graph lacks the new system
    return MDGPU_OK;
}

int mdgpu_batch_eval_device(mdgpu_batch* b, const double* d_coords, mdgpu_energies* out_e, double* d_forces) {
    if (!b || !d_coords) return berr(b, MDGPU_ERR_ARG, "null argument");
    if (b->sys.empty()) return berr(b, MDGPU_ERR_STATE, "the batch has no systems");
    B_TRY(hipSetDevice(b->device), "hipSetDevice");
    if (b->mode == 1) return run_serial(b, d_coords, out_e, d_forces);
    if (b->mode == 2) return run_streams(b, d_coords, out_e, d_forces);
    return run_graph(b, d_coords, out_e, d_forces);
}

int mdgpu_batch_eval(mdgpu_batch* b, const double* coords, mdgpu_energies* out_e, double* out_forces) {
    if (!b || !coords) return berr(b, MDGPU_ERR_ARG, "null argument");
    if (b->sys.empty()) return berr(b, MDGPU_ERR_STATE, "the batch has no systems");
    B_TRY(hipSetDevice(b->device), "hipSetDevice");
    const size_t bytes = 3 * static_cast<size_t>(b->n_atoms) * sizeof(double);
    if (b->cap_atoms < b->n_atoms) {
        if (b->d_x) (void)hipFree(b->d_x);
        if (b->d_f) (void)hipFree(b->d_f);
        b->d_x = nullptr;
        b->d_f = nullptr;
        b->cap_atoms = 0;
        B_TRY(hipMalloc(reinterpret_cast<void**>(&b->d_x), bytes), "alloc staging");
        B_TRY(hipMalloc(reinterpret_cast<void**>(&b->d_f), bytes), "alloc staging");
        b->cap_atoms = b->n_atoms;
    }
    B_TRY(hipMemcpyAsync(b->d_x, coords, bytes, hipMemcpyHostToDevice, b->s0), "upload coords");
    B_TRY(hipStreamSynchronize(b->s0), "upload coords");
    const int st = mdgpu_batch_eval_device(b, b->d_x, out_e, out_forces ? b->d_f : nullptr);
    if (st != MDGPU_OK) return st;
    if (out_forces) B_TRY(hipMemcpy(out_forces, b->d_f, bytes, hipMemcpyDeviceToHost), "read forces");
    return MDGPU_OK;
}

int mdgpu_batch_configure(mdgpu_batch* b, int mode, int streams, int split, int fft_lines) {
    if (!b) return MDGPU_ERR_ARG;
    if (mode < 0 || mode > 2 || streams < 1 || streams > kMaxStreams || split < 1 || split > 8 || fft_lines < 0 ||
        fft_lines > 16 || (fft_lines & (fft_lines - 1)) != 0)
        return berr(b, MDGPU_ERR_ARG, "mode 0..2, streams 1..8, split 1..8, fft_lines 0 or a power of two <= 16");
    if (b->graph_state != 0 && (streams != b->streams || split != b->split || fft_lines != b->fft_lines))
        return berr(b, MDGPU_ERR_STATE, "launch settings are fixed once the batch graph exists; create a new batch");
    const bool reset = split != b->split || fft_lines != b->fft_lines;
    b->mode = mode;
    b->streams = streams;
    b->split = split;
    b->fft_lines = fft_lines;
    // new batch-wide settings replace per-class ones; unchanged ones keep them
    if (reset) {
        std::fill(b->sys_split.begin(), b->sys_split.end(), split);
        std::fill(b->sys_fft.begin(), b->sys_fft.end(), fft_lines);
    }
    return MDGPU_OK;
}

int mdgpu_batch_tune(mdgpu_batch* b, const double* d_coords, double* d_forces, int reps, int* streams, int* split,
                     int* fft_lines, double* ms) {
    if (!b || !d_coords || reps < 1) return berr(b, MDGPU_ERR_ARG, "bad argument");
    if (b->sys.empty()) return berr(b, MDGPU_ERR_STATE, "the batch has no systems");
    if (b->graph_state != 0) return berr(b, MDGPU_ERR_STATE, "tune before the first graph evaluation");
    B_TRY(hipSetDevice(b->device), "hipSetDevice");
    const int n = static_cast<int>(b->sys.size());
    b->streams = std::min(kMaxStreams, n);
    // size classes: systems by atom count, cut into up to kClasses of equal count
    constexpr int kClasses = 4;
    std::vector<int> order(n);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(),
                     [&](int a, int c) { return b->sys[a]->topo.n_atoms < b->sys[c]->topo.n_atoms; });
    const int nclass = std::min(kClasses, n);
    std::vector<std::vector<int>> cls(nclass);
    for (int k = 0; k < n; ++k) cls[static_cast<size_t>(k) * nclass / n].push_back(order[k]);
    // one pass with the graph's launches and streams, timed
    int st = run_serial(b, d_coords, nullptr, d_forces);  // settles list sizes
    if (st != MDGPU_OK) return st;
    auto timed = [&](double& per) {
        int s = run_streams(b, d_coords, nullptr, d_forces);  // settles lists for these settings
        const auto t0 = std::chrono::steady_clock::now();
        for (int r = 0; r < reps && s == MDGPU_OK; ++r) s = run_streams(b, d_coords, nullptr, d_forces);
        per = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count() / reps;
        return s;
    };
    double best = 0;
    st = timed(best);
    if (st != MDGPU_OK) return st;
    // one sweep over the classes, largest first: each tries every pair-tile
    // split (1, 2, 4, 8) and FFT block shape (16, 8, 4 lines) with the other
    // classes at their current choice, and keeps the fastest
    const int cand_split[] = {1, 2, 4, 8};
    const int cand_fft[] = {0, 8, 4};
    for (int k = nclass - 1; k >= 0; --k) {
        const int bp0 = b->sys_split[cls[k][0]], bf0 = b->sys_fft[cls[k][0]];
        int bp = bp0, bf = bf0;
        for (int p : cand_split)
            for (int f : cand_fft) {
                if (p == bp0 && f == bf0) continue;
                for (int i : cls[k]) {
                    b->sys_split[i] = p;
                    b->sys_fft[i] = f;
                }
                double per = 0;
                st = timed(per);
                if (st != MDGPU_OK) return st;
                if (per < best) {
                    best = per;
                    bp = p;
                    bf = f;
                }
            }
        for (int i : cls[k]) {
            b->sys_split[i] = bp;
            b->sys_fft[i] = bf;
        }
    }
    // the largest class's choice is reported as the batch's
    const int big = cls[nclass - 1][0];
    b->split = b->sys_split[big];
    b->fft_lines = b->sys_fft[big];
    // graph or direct launches: capture the graph once with these shapes,
    // time both forms and keep the faster as the batch's mode
    if (b->mode != 1) {
        st = capture(b, d_coords, d_forces);
        if (st != MDGPU_OK) return st;
        b->graph_state = 1;
        auto timed_mode = [&](int mode, double& per) {
            b->mode = mode;
            int s = mdgpu_batch_eval_device(b, d_coords, nullptr, d_forces);
            const auto t0 = std::chrono::steady_clock::now();
            for (int r = 0; r < reps && s == MDGPU_OK; ++r) s = mdgpu_batch_eval_device(b, d_coords, nullptr, d_forces);
            per = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count() / reps;
            return s;
        };
        double t_graph = 0, t_direct = 0;
        st = timed_mode(0, t_graph);
        if (st == MDGPU_OK) st = timed_mode(2, t_direct);
        if (st != MDGPU_OK) return st;
        b->mode = t_graph <= t_direct ? 0 : 2;
        best = std::min(t_graph, t_direct);
    }
    if (streams) *streams = b->streams;
    if (split) *split = b->split;
    if (fft_lines) *fft_lines = b->fft_lines;
    if (ms) *ms = best;
    return MDGPU_OK;
}

int mdgpu_batch_mode(const mdgpu_batch* b, int* mode) {
    if (!b || !mode) return MDGPU_ERR_ARG;
    *mode = b->mode;
    return MDGPU_OK;
}

int mdgpu_batch_system_config(const mdgpu_batch* b, int index, int* split, int* fft_lines) {
    if (!b || index < 0 || index >= static_cast<int>(b->sys.size())) return MDGPU_ERR_ARG;
    if (split) *split = b->sys_split[index];
    if (fft_lines) *fft_lines = b->sys_fft[index];
    return MDGPU_OK;
}

int mdgpu_batch_graph_state(const mdgpu_batch* b, int* state) {
    if (!b || !state) return MDGPU_ERR_ARG;
    *state = b->graph_state;
    return MDGPU_OK;
}

int mdgpu_batch_size(const mdgpu_batch* b, int* n_systems, int64_t* n_atoms) {
    if (!b) return MDGPU_ERR_ARG;
    if (n_systems) *n_systems = static_cast<int>(b->sys.size());
    if (n_atoms) *n_atoms = b->n_atoms;
    return MDGPU_OK;
}

}  // extern "C"
