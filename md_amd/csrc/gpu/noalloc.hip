// Allocation-free rebuild query: mdgpu_pairs_need_rebuild_noalloc.
//
// The same decision as mdgpu_pairs_need_rebuild (the reference's
// ljpairs_needs_rebuild, NonBondedUtils.cpp:87-100: some atom's squared
// minimum-image displacement since the reference positions exceeds
// (skin/2)^2), for callers that cannot allocate device memory: code under HIP
// graph capture and serving loops that keep their own reference positions.
// It takes no context.  Its kernels' scratch (one partial maximum per block
// and a ticket counter) is the file-scope device array `scratch` below, one
// copy per GPU, held until the process ends; its device address is looked up
// once per device with hipGetSymbolAddress and only that address is passed to
// kernels.  Calls that overlap on one GPU share the array and race; calls on
// one stream do not.
#include "ctx.h"

#include <cmath>
#include <mutex>

namespace {

constexpr int kNaThreads = 256;
constexpr int kNaMaxBlocks = 4096;                              // one partial maximum per block
constexpr int kNaMaxAtoms = kNaThreads * kNaMaxBlocks;          // 1,048,576 atoms
constexpr int kNaWords = kNaMaxBlocks + 1;                      // + the ticket
constexpr int kNaMaxDevices = 64;
static_assert(kNaMaxAtoms == MDGPU_NOALLOC_MAX_ATOMS, "header and kernels disagree on the maximum");

// [0, kNaMaxBlocks): bits of each block's largest squared displacement;
// [kNaMaxBlocks]: the ticket counter (low 32 bits), zero at the start of a call.
static __device__ __align__(256) unsigned long long scratch[kNaWords];

std::once_flag g_once[kNaMaxDevices];
unsigned long long* g_scratch[kNaMaxDevices] = {};
hipError_t g_status[kNaMaxDevices] = {};

// The scratch address on `device`, looked up once per device with that device
// current (the caller's current device is restored).
hipError_t scratch_on(int device, unsigned long long** out) {
    std::call_once(g_once[device], [device]() {
        int prev = 0;
        hipError_t e = hipGetDevice(&prev);
        if (e == hipSuccess) e = hipSetDevice(device);
        void* p = nullptr;
        if (e == hipSuccess) e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e == hipSuccess && p == nullptr) e = hipErrorInvalidSymbol;
        g_scratch[device] = static_cast<unsigned long long*>(p);
        g_status[device] = e;
        if (prev != device) (void)hipSetDevice(prev);
    });
    *out = g_scratch[device];
    return g_status[device];
}

__global__ void k_na_clear(unsigned long long* s) { s[kNaMaxBlocks] = 0ull; }

// One thread per atom; the minimum image and the squared length are the
// arithmetic of the context's k_disp (no contraction into FMAs), so the
// decision is the same bit for bit.  Each block stores its maximum; the last
// block to finish reads them all and publishes the decision.
__global__ __launch_bounds__(kNaThreads) void k_na_disp(const double* __restrict__ x, const double* __restrict__ xref,
                                                        int n, double lx, double ly, double lz, double lim2,
                                                        unsigned long long* s, int32_t* needs, double* max_d2) {
#pragma clang fp contract(off)
    __shared__ double wmax[kNaThreads / mdg::kWave];
    __shared__ int last;
    const int i = blockIdx.x * kNaThreads + threadIdx.x;
    double m = 0.0;
    if (i < n) {
        double dx = x[3 * i] - xref[3 * i];
        double dy = x[3 * i + 1] - xref[3 * i + 1];
        double dz = x[3 * i + 2] - xref[3 * i + 2];
        dx = dx - lx * round(dx / lx);
        dy = dy - ly * round(dy / ly);
        dz = dz - lz * round(dz / lz);
        const double d2 = dx * dx + dy * dy + dz * dz;
        if (d2 > m) m = d2;
    }
    for (int k = 32; k >= 1; k >>= 1) {
        const double o = __shfl_xor(m, k, mdg::kWave);
        if (o > m) m = o;
    }
    if ((threadIdx.x & (mdg::kWave - 1)) == 0) wmax[threadIdx.x / mdg::kWave] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int q = 1; q < kNaThreads / mdg::kWave; ++q)
            if (wmax[q] > m) m = wmax[q];
        atomicExch(&s[blockIdx.x], static_cast<unsigned long long>(__double_as_longlong(m)));
        __threadfence();
        last = atomicAdd(reinterpret_cast<unsigned int*>(s + kNaMaxBlocks), 1u) == gridDim.x - 1;
    }
    __syncthreads();
    if (!last) return;
    __threadfence();
    double mx = 0.0;
    for (int b = threadIdx.x; b < static_cast<int>(gridDim.x); b += kNaThreads) {
        const double v = __longlong_as_double(static_cast<long long>(atomicAdd(&s[b], 0ull)));
        if (v > mx) mx = v;
    }
    for (int k = 32; k >= 1; k >>= 1) {
        const double o = __shfl_xor(mx, k, mdg::kWave);
        if (o > mx) mx = o;
    }
    __syncthreads();
    if ((threadIdx.x & (mdg::kWave - 1)) == 0) wmax[threadIdx.x / mdg::kWave] = mx;
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int q = 1; q < kNaThreads / mdg::kWave; ++q)
            if (wmax[q] > mx) mx = wmax[q];
        *needs = mx > lim2 ? 1 : 0;
        if (max_d2) *max_d2 = mx;
        atomicExch(reinterpret_cast<unsigned int*>(s + kNaMaxBlocks), 0u);
    }
}

}  // namespace

extern "C" {

int mdgpu_pairs_need_rebuild_noalloc(int device, const double* d_coords, const double* d_coords_ref, int32_t n_atoms,
                                     const double box[3], double skin, void* stream, int32_t* d_needs,
                                     double* d_max_d2) {
    if (!d_coords || !d_coords_ref || !box || !d_needs || n_atoms < 0) return MDGPU_ERR_ARG;
    if (n_atoms > kNaMaxAtoms) return MDGPU_ERR_ARG;  // beyond the scratch array's partial maxima
    for (int d = 0; d < 3; ++d)
        if (!(box[d] > 0.0) || !std::isfinite(box[d])) return MDGPU_ERR_ARG;
    if (!(skin >= 0.0) || !std::isfinite(skin)) return MDGPU_ERR_ARG;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || device < 0 || device >= ndev || device >= kNaMaxDevices)
        return MDGPU_ERR_ARG;
    unsigned long long* s = nullptr;
    if (scratch_on(device, &s) != hipSuccess) return MDGPU_ERR_HIP;
    int prev = 0;
    if (hipGetDevice(&prev) != hipSuccess) return MDGPU_ERR_HIP;
    if (prev != device && hipSetDevice(device) != hipSuccess) return MDGPU_ERR_HIP;
    hipStream_t st = static_cast<hipStream_t>(stream);
    const double lim2 = (0.5 * skin) * (0.5 * skin);
    const int nb = n_atoms > 0 ? (n_atoms + kNaThreads - 1) / kNaThreads : 1;
    hipLaunchKernelGGL(k_na_clear, dim3(1), dim3(1), 0, st, s);
    hipLaunchKernelGGL(k_na_disp, dim3(nb), dim3(kNaThreads), 0, st, d_coords, d_coords_ref, n_atoms, box[0], box[1],
                       box[2], lim2, s, d_needs, d_max_d2);
    const hipError_t e = hipGetLastError();
    if (prev != device) (void)hipSetDevice(prev);
    return e == hipSuccess ? MDGPU_OK : MDGPU_ERR_HIP;
}

}  // extern "C"
