// Shared device definitions for the gfx950 kernels.
#ifndef MD_AMD_GPU_COMMON_H
#define MD_AMD_GPU_COMMON_H

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace mdg {

constexpr int kWave = 64;          // CDNA wavefront
constexpr int kCluster = 8;        // atoms per cluster; a tile is 8 x 8 = one wave
// 18.2206899283247^2, the reference's Coulomb constant (CoulombicEE.cpp:12)
constexpr double kCoulomb = 331.99354146411776;
constexpr double kPi = 3.14159265358979323846;
// A padding slot of a cluster sits here; every real atom is far beyond the
// cutoff from it, and it has zero charge and LJ parameters.
constexpr double kFarAway = 1.0e5;

// Bit of a tile's 64-bit pair mask for i-slot ti and j-slot tj.  It is the
// lane that computes the pair in the pair kernels (lane = 8 tj + ti), so a
// tile's mask is a lane mask of the wave and is applied with scalar ANDs.
__host__ __device__ constexpr int mask_bit(int ti, int tj) { return kCluster * tj + ti; }
// Mask of a cluster's tile with itself: the pairs with tj > ti.
constexpr unsigned long long self_tile_mask() {
    unsigned long long m = 0;
    for (int ti = 0; ti < kCluster; ++ti)
        for (int tj = ti + 1; tj < kCluster; ++tj) m |= 1ull << mask_bit(ti, tj);
    return m;
}
constexpr unsigned long long kSelfMask = self_tile_mask();
static_assert(kSelfMask == 0x7F3F1F0F07030100ull, "self-tile mask");

// Periodic cell.  Orthorhombic (tri == 0): edge lengths l.  Triclinic: edge
// vectors a = (ax, 0, 0), b = (bx, by, 0), c = (cx, cy, cz) (the columns of
// H) and the upper-triangular R = H^-1, so that s = R x are the fractional
// coordinates; l holds |a|, |b|, |c|.
struct Cell {
    double l[3];
    int tri;
    double ax, bx, by, cx, cy, cz;
    double r00, r01, r02, r11, r12, r22;
};

__host__ __device__ __forceinline__ double cell_volume(const Cell& b) {
    return b.tri ? b.ax * b.by * b.cz : b.l[0] * b.l[1] * b.l[2];
}

// Fractional coordinates s = R x of a triclinic cell.
// Written without contraction into fused multiply-adds, so that host and
// device code round alike.
__host__ __device__ __forceinline__ void cell_frac(const Cell& b, double x, double y, double z, double s[3]) {
#pragma clang fp contract(off)
    s[2] = b.r22 * z;
    s[1] = b.r11 * y + b.r12 * z;
    s[0] = b.r00 * x + b.r01 * y + b.r02 * z;
}

// Minimum image of a separation d in a triclinic cell: d - H round(R d).
// When the cutoff is below half the smallest distance between opposite
// faces, this is the only image that can lie within it.
__host__ __device__ __forceinline__ void cell_min_image_tri(const Cell& b, double& dx, double& dy, double& dz) {
#pragma clang fp contract(off)
    double s[3];
    cell_frac(b, dx, dy, dz, s);
    const double n0 = rint(s[0]), n1 = rint(s[1]), n2 = rint(s[2]);
    dx = dx - (b.ax * n0 + b.bx * n1 + b.cx * n2);
    dy = dy - (b.by * n1 + b.cy * n2);
    dz = dz - b.cz * n2;
}

// Energy accumulator slots of the partial-sum buffers.
enum EnergySlot {
    E_BOND = 0, E_ANGLE, E_DIHEDRAL, E_UB, E_IMPROPER, E_CMAP,
    E_VDW, E_EE_ALL, E_EE_LIST, E_EE_RF, E_EW_DIRECT,
    E_EW_RECIP, E_EW_EXCL, E_VDW14, E_EE14,
    E_VIR_XX, E_VIR_YY, E_VIR_ZZ, E_VIR_XY, E_VIR_XZ, E_VIR_YZ,   // reciprocal-space virial
    // virial of the pair, bonded and correction forces (statistics runs only):
    // W_ab = sum over interactions of (relative position)_a (force)_b
    E_VSR_XX, E_VSR_YY, E_VSR_ZZ, E_VSR_XY, E_VSR_XZ, E_VSR_YZ,
    E_PAIRS,   // number of atom pairs inside the cutoff (work count)
    E_TILES,   // number of cluster-pair tiles (work count)
    E_NSLOTS
};

// ---------------------------------------------------------------------------
// 64-bit fixed-point force accumulation.  A contribution f is added as
// llrint(f * 2^s); integer addition is associative, so the sum is independent
// of the order in which atomics land (bit-reproducible).  A contribution that
// does not fit sets the overflow flag instead.
// ---------------------------------------------------------------------------
struct FixedAcc {
    unsigned long long* f;   // [3N]
    double scale;            // 2^s
    unsigned int* overflow;
};

__device__ __forceinline__ long long to_fixed(double v, double scale, unsigned int* overflow) {
    const double s = v * scale;
    if (!(fabs(s) < 9.0e18)) {   // also catches NaN
        atomicOr(overflow, 1u);
        return 0;
    }
    return __double2ll_rn(s);
}

__device__ __forceinline__ void acc_add(const FixedAcc& a, int idx, double v) {
    const long long q = to_fixed(v, a.scale, a.overflow);
    if (q != 0) atomicAdd(a.f + idx, static_cast<unsigned long long>(q));
}

// The conversion at scale 2^32 in FP32 arithmetic for a float contribution
// (|v| < 2^31): with h = floor(v) the high word of v 2^32 is h and the low
// word (v - h) 2^32, so no 64-bit shifts are needed; the result is
// deterministic and within 2^-24 |v| of v 2^32.
__device__ __forceinline__ long long fixed32_f32(float v, unsigned int* overflow) {
    if (!(fabsf(v) < 2147483520.0f)) {  // also catches NaN
        atomicOr(overflow, 1u);
        return 0;
    }
    const float h = floorf(v);
    const unsigned int lo = static_cast<unsigned int>(fminf((v - h) * 4294967296.0f, 4294967040.0f));
    const unsigned long long hi = static_cast<unsigned long long>(static_cast<unsigned int>(static_cast<int>(h)));
    return static_cast<long long>((hi << 32) | lo);
}

__device__ __forceinline__ void acc_add(double* f, int idx, double v) {
    if (v != 0.0) atomicAdd(f + idx, v);
}

// ---------------------------------------------------------------------------
// Buffer atomics switched off per lane through the offset.  A lane whose byte
// offset is kOffOOB lies beyond the resource's size and the hardware drops its
// operation, so every lane issues the instruction and no branch surrounds it:
// the compiler then counts the operation as always issued when it places the
// waits for earlier loads.
// ---------------------------------------------------------------------------
using BufRsrc = __amdgpu_buffer_rsrc_t;
constexpr unsigned int kOffOOB = 0x80000000u;
// word 3 of a gfx9 buffer descriptor: 32-bit data format, no swizzle
constexpr int kBufWord3 = 0x00020000;

__device__ long long llvm_buf_atomic_add_i64(long long v, BufRsrc r, unsigned int off, int soff, int aux)
    __asm("llvm.amdgcn.raw.ptr.buffer.atomic.add.i64");
__device__ double llvm_buf_atomic_fadd_f64(double v, BufRsrc r, unsigned int off, int soff, int aux)
    __asm("llvm.amdgcn.raw.ptr.buffer.atomic.fadd.f64");

// Resource over `bytes` bytes from p; bytes must stay below kOffOOB.
__device__ __forceinline__ BufRsrc buf_rsrc(void* p, unsigned int bytes) {
    return __builtin_amdgcn_make_buffer_rsrc(p, 0, static_cast<int>(bytes), kBufWord3);
}
// s_waitcnt operand of gfx9: vmcnt(0), expcnt and lgkmcnt not waited for
constexpr int kWaitVmcnt0 = 0x0F70;
// Loads through a buffer resource with a scalar byte offset (soff, uniform
// over the wave) and a vector one (off).
typedef float mdg_v4f __attribute__((ext_vector_type(4)));
typedef int mdg_v2i __attribute__((ext_vector_type(2)));
__device__ mdg_v4f llvm_buf_load_v4f32(BufRsrc r, unsigned int off, int soff, int aux)
    __asm("llvm.amdgcn.raw.ptr.buffer.load.v4f32");
__device__ mdg_v2i llvm_buf_load_v2i32(BufRsrc r, unsigned int off, int soff, int aux)
    __asm("llvm.amdgcn.raw.ptr.buffer.load.v2i32");
__device__ __forceinline__ float4 buf_load_f4(BufRsrc r, unsigned int off, int soff) {
    const mdg_v4f v = llvm_buf_load_v4f32(r, off, soff, 0);
    return make_float4(v.x, v.y, v.z, v.w);
}
__device__ __forceinline__ int2 buf_load_i2(BufRsrc r, unsigned int off, int soff) {
    const mdg_v2i v = llvm_buf_load_v2i32(r, off, soff, 0);
    return make_int2(v.x, v.y);
}
__device__ __forceinline__ void buf_add(BufRsrc r, unsigned int off, long long q) {
    llvm_buf_atomic_add_i64(q, r, off, 0, 0);
}
__device__ __forceinline__ void buf_add(BufRsrc r, unsigned int off, double v) {
    llvm_buf_atomic_fadd_f64(v, r, off, 0, 0);
}

// ---------------------------------------------------------------------------
// wave-level helpers
// ---------------------------------------------------------------------------
template <typename T>
__device__ __forceinline__ T shfl_xor(T v, int mask) {
    return __shfl_xor(v, mask, kWave);
}

__device__ __forceinline__ double wave_sum(double v) {
    for (int m = 32; m >= 1; m >>= 1) v += __shfl_xor(v, m, kWave);
    return v;
}

// Block-wide sum of several doubles into row `row` of the partial-sum buffer,
// stored slot-major (partial[slot * ld + row], ld = rows allocated, so the
// final reduction reads each slot contiguously): slot slots[k] receives the
// block total of v[k], every other slot of the row receives 0 (so a row never
// keeps a value from an earlier evaluation).  blockDim.x must be a multiple of
// 64, >= E_NSLOTS and <= 1024.
template <int NV>
__device__ __forceinline__ void block_sum_store(double (&v)[NV], double* partial, int ld, int row, const int* slots) {
    __shared__ double red[16][NV];
    __shared__ double tot[NV];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k < NV; ++k) v[k] = wave_sum(v[k]);
    if (lane == 0)
#pragma unroll
        for (int k = 0; k < NV; ++k) red[w][k] = v[k];
    __syncthreads();
    if (threadIdx.x < NV) {
        const int nw = (blockDim.x + 63) >> 6;
        double s = 0.0;
        for (int q = 0; q < nw; ++q) s += red[q][threadIdx.x];
        tot[threadIdx.x] = s;
    }
    __syncthreads();
    if (threadIdx.x < E_NSLOTS) {
        double s = 0.0;
#pragma unroll
        for (int k = 0; k < NV; ++k)
            if (slots[k] == static_cast<int>(threadIdx.x)) s = tot[k];
        partial[static_cast<size_t>(threadIdx.x) * ld + row] = s;
    }
}

// XCD-aware block order: consecutive logical blocks land on one XCD (the
// hardware deals workgroups round-robin over the 8 XCDs), so neighbouring
// clusters share that XCD's L2.
// The map is a bijection of [0, nblocks): XCD x receives full + (x < rem)
// blocks, placed after the blocks of XCDs 0..x-1.
__device__ __forceinline__ int xcd_block(int b, int nblocks) {
    constexpr int kXcd = 8;
    const int full = nblocks / kXcd, rem = nblocks % kXcd;
    const int xcd = b % kXcd, k = b / kXcd;
    return xcd * full + (xcd < rem ? xcd : rem) + k;
}

}  // namespace mdg

#endif
