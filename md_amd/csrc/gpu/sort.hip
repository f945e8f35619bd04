// Spatial sort into clusters of 8 atoms and the cluster-pair search.
//
// Replaces System::build_lj_pairlist / build_ee_pairlist (reference
// src/System/NonBondedUtils.cpp:185-381: linked-list cells of size rc + skin,
// 27 neighbour cells, one record of 16 B per atom pair).  Here:
//   1. atoms are wrapped into the box in FP64 (PBC.cpp:78-81) and binned into
//      x-y columns about (8/rho)^(1/3) wide;
//   2. each column is sorted by z (ties by atom index, so the order is
//      deterministic) and cut into clusters of 8 (the last one padded);
//   3. each cluster gets a bounding box; its atoms are stored relative to the
//      box centre (FP32 offsets stay small, which keeps the mixed form's pair
//      distances accurate);
//   4. one wavefront per i-cluster tests candidate (j-cluster, periodic shift)
//      bounding boxes against the cutoff and writes one list entry per
//      surviving pair of clusters: j, the FP64-derived origin difference, and
//      a 64-bit mask (self-tile triangle, and in AMBER semantics the excluded
//      atom pairs).
// Each unordered atom pair lies in exactly one tile (j-cluster > i-cluster, or
// the self tile without shift), which is the reference's half list.
#include "ctx.h"

#include <algorithm>
#include <cmath>

namespace mdg {

namespace {

constexpr int kColMax = 2048;  // atoms per column the in-LDS sort handles

// Column binning without global atomics: each block takes a contiguous chunk
// of atoms, histograms their columns in LDS and stores its histogram row;
// the scan derives every (block, column) offset from those rows, and the
// scatter repeats the per-block LDS counting to place the atoms.
constexpr int kBinThreads = 1024;
constexpr int kMaxCols = 8192;

__global__ __launch_bounds__(kBinThreads) void k_bin(const double* __restrict__ x, int n, int chunk, double lx,
                                                     double ly, double lz, int ncx, int ncy, double sx, double sy,
                                                     double* __restrict__ wrapped, int* __restrict__ col_of,
                                                     int* __restrict__ block_hist) {
    __shared__ int hist[kMaxCols];
    const int ncol = ncx * ncy;
    for (int c = threadIdx.x; c < ncol; c += blockDim.x) hist[c] = 0;
    __syncthreads();
    const int i0 = blockIdx.x * chunk, i1 = min(n, i0 + chunk);
    for (int i = i0 + threadIdx.x; i < i1; i += blockDim.x) {
        // fractional wrapping in FP64: x - L floor(x / L)
        const double wx = x[3 * i] - lx * floor(x[3 * i] / lx);
        const double wy = x[3 * i + 1] - ly * floor(x[3 * i + 1] / ly);
        const double wz = x[3 * i + 2] - lz * floor(x[3 * i + 2] / lz);
        wrapped[3 * i] = wx;
        wrapped[3 * i + 1] = wy;
        wrapped[3 * i + 2] = wz;
        int cx = static_cast<int>(wx / sx), cy = static_cast<int>(wy / sy);
        cx = cx < 0 ? 0 : (cx >= ncx ? ncx - 1 : cx);
        cy = cy < 0 ? 0 : (cy >= ncy ? ncy - 1 : cy);
        const int c = cx + ncx * cy;
        col_of[i] = c;
        atomicAdd(hist + c, 1);
    }
    __syncthreads();
    for (int c = threadIdx.x; c < ncol; c += blockDim.x) block_hist[static_cast<size_t>(blockIdx.x) * ncol + c] = hist[c];
}

// One block: column totals, exclusive scans of atom and cluster counts, and
// the start of every (block, column) run.
__global__ __launch_bounds__(1024) void k_col_scan(const int* __restrict__ block_hist, int nblk, int ncol,
                                                   int* __restrict__ col_start, int* __restrict__ col_cstart,
                                                   int* __restrict__ block_off, unsigned int* flags) {
    __shared__ int s_a[1024], s_c[1024];
    __shared__ int carry_a, carry_c;
    if (threadIdx.x == 0) { carry_a = 0; carry_c = 0; }
    __syncthreads();
    for (int base = 0; base < ncol; base += 1024) {
        const int i = base + threadIdx.x;
        int a = 0;
        if (i < ncol)
            for (int b = 0; b < nblk; ++b) a += block_hist[static_cast<size_t>(b) * ncol + i];
        if (a > kColMax) atomicOr(flags + 2, 1u);
        s_a[threadIdx.x] = a;
        s_c[threadIdx.x] = (a + kCluster - 1) / kCluster;
        __syncthreads();
        for (int off = 1; off < 1024; off <<= 1) {
            const int va = threadIdx.x >= off ? s_a[threadIdx.x - off] : 0;
            const int vc = threadIdx.x >= off ? s_c[threadIdx.x - off] : 0;
            __syncthreads();
            s_a[threadIdx.x] += va;
            s_c[threadIdx.x] += vc;
            __syncthreads();
        }
        if (i < ncol) {
            col_start[i + 1] = carry_a + s_a[threadIdx.x];
            col_cstart[i + 1] = carry_c + s_c[threadIdx.x];
            int run = carry_a + s_a[threadIdx.x] - a;  // exclusive start of column i
            for (int b = 0; b < nblk; ++b) {
                block_off[static_cast<size_t>(b) * ncol + i] = run;
                run += block_hist[static_cast<size_t>(b) * ncol + i];
            }
        }
        __syncthreads();
        if (threadIdx.x == 1023) {
            carry_a += s_a[1023];
            carry_c += s_c[1023];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        col_start[0] = 0;
        col_cstart[0] = 0;
    }
}

__global__ __launch_bounds__(kBinThreads) void k_scatter(int n, int chunk, int ncol, const int* __restrict__ col_of,
                                                         const int* __restrict__ block_off,
                                                         int* __restrict__ col_tmp) {
    __shared__ int off[kMaxCols];
    for (int c = threadIdx.x; c < ncol; c += blockDim.x) off[c] = block_off[static_cast<size_t>(blockIdx.x) * ncol + c];
    __syncthreads();
    const int i0 = blockIdx.x * chunk, i1 = min(n, i0 + chunk);
    for (int i = i0 + threadIdx.x; i < i1; i += blockDim.x) col_tmp[atomicAdd(off + col_of[i], 1)] = i;
}

// One block per column: rank the column's atoms by (z, index) and write the
// slot -> atom map; the padding slots of the column's last cluster get -1.
__global__ void k_col_sort(const int* __restrict__ col_start, const int* __restrict__ col_cstart,
                           const int* __restrict__ col_tmp, const double* __restrict__ wrapped,
                           int* __restrict__ slot_atom, int* __restrict__ cluster_col) {
    __shared__ double s_z[kColMax];
    __shared__ int s_i[kColMax];
    const int c = blockIdx.x;
    const int a0 = col_start[c], n = min(col_start[c + 1] - a0, kColMax);
    const int cl0 = col_cstart[c], ncl = col_cstart[c + 1] - cl0;
    for (int k = threadIdx.x; k < n; k += blockDim.x) {
        const int atom = col_tmp[a0 + k];
        s_i[k] = atom;
        s_z[k] = wrapped[3 * atom + 2];
    }
    __syncthreads();
    for (int k = threadIdx.x; k < n; k += blockDim.x) {
        const double z = s_z[k];
        const int id = s_i[k];
        int rank = 0;
        for (int m = 0; m < n; ++m) {
            const double zm = s_z[m];
            rank += (zm < z) || (zm == z && s_i[m] < id);
        }
        slot_atom[cl0 * kCluster + rank] = id;
    }
    for (int k = n + threadIdx.x; k < ncl * kCluster; k += blockDim.x) slot_atom[cl0 * kCluster + k] = -1;
    for (int k = threadIdx.x; k < ncl; k += blockDim.x) cluster_col[cl0 + k] = c;
}

template <typename real>
struct Vec4;
template <>
struct Vec4<float> { using T = float4; };
template <>
struct Vec4<double> { using T = double4; };

// Eight lanes per cluster: bounding box by shuffles, then the relative
// coordinates.  Padding slots get the padding type, zero charge, and an
// offset of kFarAway.
// PME geometry for the per-slot spline records (k[0] == 0: no records).
struct PmeSlots {
    int k[3];
    double l[3];
    int* fl;
    void* w;   // double4 / float4 per slot: fractions (x, y, z) and charge
};

template <typename real>
__global__ void k_cluster_setup(const int* __restrict__ ncl_dev, int max_clusters, const int* __restrict__ slot_atom,
                                const double* __restrict__ wrapped, const double* __restrict__ charge,
                                const int* __restrict__ type, int pad_type, typename Vec4<real>::T* __restrict__ pos,
                                int2* __restrict__ slot_meta, double4* __restrict__ origin,
                                float4* __restrict__ half, const PmeSlots pme) {
    const int slot = blockIdx.x * blockDim.x + threadIdx.x;
    const int cl = slot / kCluster;
    const int ncl = *ncl_dev;
    const bool live = cl < ncl && cl < max_clusters;
    const int atom = live ? slot_atom[slot] : -1;
    double x = 0, y = 0, z = 0;
    if (atom >= 0) {
        x = wrapped[3 * atom];
        y = wrapped[3 * atom + 1];
        z = wrapped[3 * atom + 2];
    }
    double lo[3] = {atom >= 0 ? x : 1e300, atom >= 0 ? y : 1e300, atom >= 0 ? z : 1e300};
    double hi[3] = {atom >= 0 ? x : -1e300, atom >= 0 ? y : -1e300, atom >= 0 ? z : -1e300};
#pragma unroll
    for (int m = 1; m < kCluster; m <<= 1) {
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            lo[d] = fmin(lo[d], __shfl_xor(lo[d], m, kWave));
            hi[d] = fmax(hi[d], __shfl_xor(hi[d], m, kWave));
        }
    }
    if (!live) return;
    const double ox = 0.5 * (lo[0] + hi[0]), oy = 0.5 * (lo[1] + hi[1]), oz = 0.5 * (lo[2] + hi[2]);
    typename Vec4<real>::T p;
    if (atom >= 0) {
        p.x = static_cast<real>(x - ox);
        p.y = static_cast<real>(y - oy);
        p.z = static_cast<real>(z - oz);
        p.w = static_cast<real>(charge[atom]);
        slot_meta[slot] = make_int2(type[atom], atom);
    } else {
        p.x = p.y = p.z = static_cast<real>(kFarAway);
        p.w = 0;
        slot_meta[slot] = make_int2(pad_type, -1);
    }
    pos[slot] = p;
    if (pme.k[0] > 0) {
        // u = K x / L in FP64; fl = floor(u) taken mod K (the stencil wraps
        // anyway), w = u - floor(u)
        int packed = 0;
        typename Vec4<real>::T wq = {0, 0, 0, 0};
        if (atom >= 0) {
            const double xs[3] = {x, y, z};
            real wv[3];
            for (int d = 0; d < 3; ++d) {
                const double u = xs[d] * pme.k[d] / pme.l[d];
                const double f = floor(u);
                int fl = static_cast<int>(f) % pme.k[d];
                fl += fl < 0 ? pme.k[d] : 0;
                packed |= fl << (10 * d);
                wv[d] = static_cast<real>(u - f);
            }
            wq.x = wv[0];
            wq.y = wv[1];
            wq.z = wv[2];
            wq.w = static_cast<real>(charge[atom]);
        }
        pme.fl[slot] = packed;
        reinterpret_cast<typename Vec4<real>::T*>(pme.w)[slot] = wq;
    }
    if ((slot & (kCluster - 1)) == 0) {
        origin[cl] = make_double4(ox, oy, oz, 0.0);
        // round the half extents up so the float box still contains every atom
        half[cl] = make_float4(static_cast<float>(0.5 * (hi[0] - lo[0])) * 1.0001f + 1e-4f,
                               static_cast<float>(0.5 * (hi[1] - lo[1])) * 1.0001f + 1e-4f,
                               static_cast<float>(0.5 * (hi[2] - lo[2])) * 1.0001f + 1e-4f, 0.f);
    }
}

__device__ __forceinline__ int floordiv(int a, int b) { return a >= 0 ? a / b : -((-a + b - 1) / b); }

__device__ __forceinline__ bool pair_excluded(const int* __restrict__ start, const int* __restrict__ list, int a,
                                              int b) {
    int lo = start[a], hi = start[a + 1] - 1;
    while (lo <= hi) {
        const int mid = (lo + hi) >> 1;
        const int v = list[mid];
        if (v == b) return true;
        if (v < b) lo = mid + 1;
        else hi = mid - 1;
    }
    return false;
}

struct SearchArgs {
    const int* ncl_dev;
    int max_clusters;
    const int* cluster_col;
    const int* col_cstart;
    const double4* origin;
    const float4* half;
    const int* slot_atom;
    int ncx, ncy, dcx, dcy;
    double sx, sy, lx, ly, lz;
    double rl2;             // list cutoff squared
    int cap;
    int* list_count;
    int* list_j;
    unsigned long long* list_mask;
    void* list_delta;
    unsigned int* flags;
    // exclusions (AMBER)
    int amber;
    const int* excl_start;
    const int* excl_list;
    const float* excl_rmax;
};

constexpr int kNearCap = 256;
constexpr int kSearchWaves = 4;

template <typename real>
__global__ __launch_bounds__(kSearchWaves * kWave) void k_search(const SearchArgs a) {
    using R4 = typename Vec4<real>::T;
    __shared__ int2 s_near[kSearchWaves][kNearCap];  // (entry, j-cluster)
    const int lane = threadIdx.x & (kWave - 1), wv = threadIdx.x >> 6;
    const int ci = blockIdx.x * kSearchWaves + wv;
    const int ncl = min(*a.ncl_dev, a.max_clusters);
    if (ci >= ncl) return;  // whole wave leaves; the kernel has no block barrier
    const int col = a.cluster_col[ci];
    const int cx = col % a.ncx, cy = col / a.ncx;
    const double4 oi = a.origin[ci];
    const float4 hi = a.half[ci];
    const size_t base_i = static_cast<size_t>(ci) * a.cap;
    R4* delta = reinterpret_cast<R4*>(a.list_delta);
    // largest excluded-pair separation, with margin for the float boxes
    double rex = 0.0;
    if (a.amber) rex = static_cast<double>(*a.excl_rmax) + 0.01;
    const double rex2 = rex * rex;
    int count = 0, nnear = 0;
    const unsigned long long self_mask = 0x0080C0E0F0F8FCFEull;  // bit 8*ti+tj set iff tj > ti
    for (int dy = -a.dcy; dy <= a.dcy; ++dy) {
        const int jyr = cy + dy, ky = floordiv(jyr, a.ncy), jy = jyr - ky * a.ncy;
        const double gy = fmax(0.0, (abs(dy) - 1) * a.sy);
        if (gy * gy >= a.rl2) continue;
        for (int dx = -a.dcx; dx <= a.dcx; ++dx) {
            const int jxr = cx + dx, kx = floordiv(jxr, a.ncx), jx = jxr - kx * a.ncx;
            const double gx = fmax(0.0, (abs(dx) - 1) * a.sx);
            if (gx * gx + gy * gy >= a.rl2) continue;
            const int jcol = jx + a.ncx * jy;
            const int c0 = a.col_cstart[jcol], c1 = a.col_cstart[jcol + 1];
            const double shx = kx * a.lx, shy = ky * a.ly;
            for (int b = c0; b < c1; b += kWave) {
                const int cj = b + lane;
                const bool valid = cj < c1;
                double4 oj = make_double4(0, 0, 0, 0);
                float4 hj = make_float4(0, 0, 0, 0);
                if (valid) {
                    oj = a.origin[cj];
                    hj = a.half[cj];
                }
                for (int kz = -1; kz <= 1; ++kz) {
                    const double ddx = oj.x + shx - oi.x, ddy = oj.y + shy - oi.y, ddz = oj.z + kz * a.lz - oi.z;
                    const double qx = fmax(0.0, fabs(ddx) - (double)(hi.x + hj.x));
                    const double qy = fmax(0.0, fabs(ddy) - (double)(hi.y + hj.y));
                    const double qz = fmax(0.0, fabs(ddz) - (double)(hi.z + hj.z));
                    const double d2 = qx * qx + qy * qy + qz * qz;
                    const bool noshift = kx == 0 && ky == 0 && kz == 0;
                    const bool cand = valid && (cj > ci || (cj == ci && noshift)) && d2 < a.rl2;
                    const unsigned long long bal = __ballot(cand);
                    if (bal == 0) continue;
                    const int rank = __popcll(bal & ((1ull << lane) - 1ull));
                    const int pos = count + rank;
                    const bool near = a.amber && cand && d2 < rex2;
                    if (cand && pos < a.cap) {
                        const size_t e = base_i + pos;
                        a.list_j[e] = cj;
                        // masks of tiles that may hold excluded pairs are written below
                        if (!near) a.list_mask[e] = (cj == ci) ? self_mask : ~0ull;
                        R4 dl;
                        dl.x = static_cast<real>(-ddx);
                        dl.y = static_cast<real>(-ddy);
                        dl.z = static_cast<real>(-ddz);
                        dl.w = 0;
                        delta[e] = dl;
                    }
                    const unsigned long long nbal = __ballot(near && pos < a.cap);
                    if (near && pos < a.cap) {
                        const int nr = nnear + __popcll(nbal & ((1ull << lane) - 1ull));
                        if (nr < kNearCap) s_near[wv][nr] = make_int2(pos, cj);
                    }
                    nnear += __popcll(nbal);
                    count += __popcll(bal);
                }
            }
        }
    }
    if (count > a.cap) {
        if (lane == 0) atomicOr(a.flags + 1, 1u);
        count = a.cap;
    }
    if (lane == 0) a.list_count[ci] = count;
    if (!a.amber) return;
    if (nnear > kNearCap) {
        if (lane == 0) atomicOr(a.flags + 3, 1u);
        nnear = kNearCap;
    }
    // exclusion masks of the tiles whose boxes are close enough to hold an
    // excluded pair: lane (ti, tj) looks its pair up in the CSR list.  The
    // entries come from LDS (written by this wave above), not back from memory.
    const int ti = lane >> 3, tj = lane & 7;
    const int ai = a.slot_atom[ci * kCluster + ti];
    __builtin_amdgcn_wave_barrier();
    for (int q = 0; q < nnear; ++q) {
        const int2 en = s_near[wv][q];
        const int aj = a.slot_atom[en.y * kCluster + tj];
        const bool ex = ai >= 0 && aj >= 0 && pair_excluded(a.excl_start, a.excl_list, ai, aj);
        const unsigned long long bal = __ballot(ex);
        if (lane == 0) a.list_mask[base_i + en.x] = ((en.y == ci) ? self_mask : ~0ull) & ~bal;
    }
}

// Fast search: every candidate j-cluster is taken at its minimum-image shift
// (per lane, FP64 rint), so one pass over the clusters of each row of
// columns in reach replaces the per-column, per-shift enumeration.  It is
// exact when every cluster half-extent h satisfies 2h <= L/2 - r_list in each
// dimension (then no second image of a cluster can come within r_list);
// a cluster that breaks the bound sets flags[4] and the host falls back to
// k_search.
template <typename real>
__global__ __launch_bounds__(kSearchWaves * kWave) void k_search_mi(const SearchArgs a, double hlx, double hly,
                                                                     double hlz) {
    using R4 = typename Vec4<real>::T;
    __shared__ int2 s_near[kSearchWaves][kNearCap];
    const int lane = threadIdx.x & (kWave - 1), wv = threadIdx.x >> 6;
    const int ci = __builtin_amdgcn_readfirstlane(blockIdx.x * kSearchWaves + wv);
    const int ncl = min(*a.ncl_dev, a.max_clusters);
    if (ci >= ncl) return;  // whole wave leaves; the kernel has no block barrier
    const int col = a.cluster_col[ci];
    const int cx = col % a.ncx, cy = col / a.ncx;
    const double4 oi = a.origin[ci];
    const float4 hi = a.half[ci];
    if (hi.x > hlx || hi.y > hly || hi.z > hlz) {
        if (lane == 0) {
            atomicOr(a.flags + 4, 1u);
            a.list_count[ci] = 0;  // the pair kernel of this attempt reads an empty list
        }
        return;
    }
    const double ilx = 1.0 / a.lx, ily = 1.0 / a.ly, ilz = 1.0 / a.lz;
    const size_t base_i = static_cast<size_t>(ci) * a.cap;
    R4* delta = reinterpret_cast<R4*>(a.list_delta);
    double rex = 0.0;
    if (a.amber) rex = static_cast<double>(*a.excl_rmax) + 0.01;
    const double rex2 = rex * rex;
    int count = 0, nnear = 0;
    const unsigned long long self_mask = 0x0080C0E0F0F8FCFEull;  // bit 8*ti+tj set iff tj > ti
    const int ny = min(2 * a.dcy + 1, a.ncy);
    const bool yall = ny == a.ncy;
    const int y0 = yall ? 0 : cy - a.dcy;
    for (int ry = 0; ry < ny; ++ry) {
        int jy = (y0 + ry) % a.ncy;
        jy += jy < 0 ? a.ncy : 0;
        // columns of this row within reach: |dx| <= dxr, where the x gap
        // (|dx| - 1) sx and the y gap (|dy| - 1) sy stay inside r_list
        int dxr = a.dcx;
        if (!yall) {
            const int dy = ry - a.dcy;
            const double gy = fmax(0.0, (abs(dy) - 1) * a.sy);
            const double rem = a.rl2 - gy * gy;
            if (rem <= 0.0) continue;
            dxr = min(a.dcx, static_cast<int>(sqrt(rem) / a.sx) + 1);
        }
        const int nx = min(2 * dxr + 1, a.ncx);
        const int x0 = nx == a.ncx ? 0 : cx - dxr;
        // the x window of this row as at most two contiguous column segments
        int xa = x0 % a.ncx;
        xa += xa < 0 ? a.ncx : 0;
        const int first = min(nx, a.ncx - xa);
        for (int seg = 0; seg < 2; ++seg) {
            const int c_lo = seg == 0 ? xa : 0;
            const int c_n = seg == 0 ? first : nx - first;
            if (c_n <= 0) continue;
            const int rowc = a.ncx * jy;
            const int b0 = a.col_cstart[rowc + c_lo], b1 = a.col_cstart[rowc + c_lo + c_n];
            for (int b = b0; b < b1; b += kWave) {
                const int cj = b + lane;
                const bool valid = cj < b1 && cj >= ci;
                double ddx = 0, ddy = 0, ddz = 0, d2 = 1e300;
                if (valid) {
                    const double4 oj = a.origin[cj];
                    const float4 hj = a.half[cj];
                    ddx = oj.x - oi.x;
                    ddy = oj.y - oi.y;
                    ddz = oj.z - oi.z;
                    ddx -= a.lx * rint(ddx * ilx);
                    ddy -= a.ly * rint(ddy * ily);
                    ddz -= a.lz * rint(ddz * ilz);
                    const double qx = fmax(0.0, fabs(ddx) - (double)(hi.x + hj.x));
                    const double qy = fmax(0.0, fabs(ddy) - (double)(hi.y + hj.y));
                    const double qz = fmax(0.0, fabs(ddz) - (double)(hi.z + hj.z));
                    d2 = qx * qx + qy * qy + qz * qz;
                }
                const bool cand = valid && d2 < a.rl2;
                const unsigned long long bal = __ballot(cand);
                if (bal == 0) continue;
                const int pos = count + __popcll(bal & ((1ull << lane) - 1ull));
                const bool near = a.amber && cand && d2 < rex2;
                if (cand && pos < a.cap) {
                    const size_t e = base_i + pos;
                    a.list_j[e] = cj;
                    if (!near) a.list_mask[e] = (cj == ci) ? self_mask : ~0ull;
                    R4 dl;
                    dl.x = static_cast<real>(-ddx);
                    dl.y = static_cast<real>(-ddy);
                    dl.z = static_cast<real>(-ddz);
                    dl.w = 0;
                    delta[e] = dl;
                }
                const unsigned long long nbal = __ballot(near && pos < a.cap);
                if (near && pos < a.cap) {
                    const int nr = nnear + __popcll(nbal & ((1ull << lane) - 1ull));
                    if (nr < kNearCap) s_near[wv][nr] = make_int2(pos, cj);
                }
                nnear += __popcll(nbal);
                count += __popcll(bal);
            }
        }
    }
    if (count > a.cap) {
        if (lane == 0) atomicOr(a.flags + 1, 1u);
        count = a.cap;
    }
    if (lane == 0) a.list_count[ci] = count;
    if (!a.amber) return;
    if (nnear > kNearCap) {
        if (lane == 0) atomicOr(a.flags + 3, 1u);
        nnear = kNearCap;
    }
    const int ti = lane >> 3, tj = lane & 7;
    const int ai = a.slot_atom[ci * kCluster + ti];
    __builtin_amdgcn_wave_barrier();
    for (int q = 0; q < nnear; ++q) {
        const int2 en = s_near[wv][q];
        const int aj = a.slot_atom[en.y * kCluster + tj];
        const bool ex = ai >= 0 && aj >= 0 && pair_excluded(a.excl_start, a.excl_list, ai, aj);
        const unsigned long long bal = __ballot(ex);
        if (lane == 0) a.list_mask[base_i + en.x] = ((en.y == ci) ? self_mask : ~0ull) & ~bal;
    }
}

// Largest minimum-image distance over the excluded and 1-4 pairs (bounds the
// tiles that can hold an excluded pair).
__global__ void k_excl_extent(const double* __restrict__ x, int n1, const int2* __restrict__ p1, int n2,
                              const int2* __restrict__ p2, double lx, double ly, double lz, float* out) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    float r = 0.f;
    if (t < n1 + n2) {
        const int2 p = t < n1 ? p1[t] : p2[t - n1];
        double dx = x[3 * p.x] - x[3 * p.y], dy = x[3 * p.x + 1] - x[3 * p.y + 1], dz = x[3 * p.x + 2] - x[3 * p.y + 2];
        dx -= lx * rint(dx / lx);
        dy -= ly * rint(dy / ly);
        dz -= lz * rint(dz / lz);
        r = static_cast<float>(sqrt(dx * dx + dy * dy + dz * dz)) * 1.0001f;
    }
    for (int m = 32; m >= 1; m >>= 1) r = fmaxf(r, __shfl_xor(r, m, kWave));
    if ((threadIdx.x & 63) == 0 && r > 0.f) atomicMax(reinterpret_cast<unsigned int*>(out), __float_as_uint(r));
}

}  // namespace

int sort_blocks(int n) { return std::max(1, std::min(64, (n + 2047) / 2048)); }

hipError_t launch_excl_extent(mdgpu_ctx* c) {
    const Topo& t = c->topo;
    const int n = t.n_excl_pairs + t.n_pairs14;
    if (n == 0) return hipSuccess;
    const double* x = c->cur_x;
    hipLaunchKernelGGL(k_excl_extent, dim3((n + 255) / 256), dim3(256), 0, c->stream, x, t.n_excl_pairs,
                       t.excl_pairs, t.n_pairs14, t.pairs14, c->grid.box[0], c->grid.box[1], c->grid.box[2],
                       c->work.excl_rmax);
    return hipGetLastError();
}

hipError_t launch_sort(mdgpu_ctx* c, const double* coords, bool mixed) {
    Work& w = c->work;
    const Grid& g = c->grid;
    const int n = c->topo.n_atoms;
    const int ncol = g.ncx * g.ncy;
    if (ncol > kMaxCols) return hipErrorInvalidValue;
    const int nblk = sort_blocks(n), chunk = (n + nblk - 1) / nblk;
    hipLaunchKernelGGL(k_bin, dim3(nblk), dim3(kBinThreads), 0, c->stream, coords, n, chunk, g.box[0], g.box[1],
                       g.box[2], g.ncx, g.ncy, g.sx, g.sy, w.wrapped, w.col_of, w.block_hist);
    hipLaunchKernelGGL(k_col_scan, dim3(1), dim3(1024), 0, c->stream, w.block_hist, nblk, ncol, w.col_start,
                       w.col_cstart, w.block_off, w.flags);
    hipLaunchKernelGGL(k_scatter, dim3(nblk), dim3(kBinThreads), 0, c->stream, n, chunk, ncol, w.col_of, w.block_off,
                       w.col_tmp);
    hipLaunchKernelGGL(k_col_sort, dim3(ncol), dim3(256), 0, c->stream, w.col_start, w.col_cstart, w.col_tmp,
                       w.wrapped, w.slot_atom, w.cluster_col);
    const int* ncl_dev = w.col_cstart + ncol;
    const int max_cl = c->n_clusters;
    const int nslots = max_cl * kCluster;
    PmeSlots pme{};
    const bool with_pme = c->params.semantics == MDGPU_SEM_AMBER && w.pme_k[0] > 0;
    for (int d = 0; d < 3; ++d) {
        pme.k[d] = with_pme ? w.pme_k[d] : 0;
        pme.l[d] = g.box[d];
    }
    pme.fl = w.slot_pme_fl;
    pme.w = mixed ? static_cast<void*>(w.slot_pme_wf) : static_cast<void*>(w.slot_pme_wd);
    if (mixed) {
        hipLaunchKernelGGL(k_cluster_setup<float>, dim3((nslots + 255) / 256), dim3(256), 0, c->stream, ncl_dev,
                           max_cl, w.slot_atom, w.wrapped, c->topo.charge, c->topo.type, c->topo.n_types,
                           w.slot_pos_f, w.slot_meta, w.cl_origin, w.cl_half, pme);
    } else {
        hipLaunchKernelGGL(k_cluster_setup<double>, dim3((nslots + 255) / 256), dim3(256), 0, c->stream, ncl_dev,
                           max_cl, w.slot_atom, w.wrapped, c->topo.charge, c->topo.type, c->topo.n_types,
                           w.slot_pos_d, w.slot_meta, w.cl_origin, w.cl_half, pme);
    }
    return hipGetLastError();
}

hipError_t launch_pairsearch(mdgpu_ctx* c, bool mixed, bool amber) {
    Work& w = c->work;
    const Grid& g = c->grid;
    const int ncol = g.ncx * g.ncy;
    SearchArgs a{};
    a.ncl_dev = w.col_cstart + ncol;
    a.max_clusters = c->n_clusters;
    a.cluster_col = w.cluster_col;
    a.col_cstart = w.col_cstart;
    a.origin = w.cl_origin;
    a.half = w.cl_half;
    a.slot_atom = w.slot_atom;
    a.ncx = g.ncx;
    a.ncy = g.ncy;
    a.dcx = g.dcx;
    a.dcy = g.dcy;
    a.sx = g.sx;
    a.sy = g.sy;
    a.lx = g.box[0];
    a.ly = g.box[1];
    a.lz = g.box[2];
    const double rl = std::max(c->params.lj_cutoff, c->params.ee_cutoff);
    a.rl2 = rl * rl;
    a.cap = w.list_cap;
    a.list_count = w.list_count;
    a.list_j = w.list_j;
    a.list_mask = w.list_mask;
    a.list_delta = mixed ? static_cast<void*>(w.list_delta_f) : static_cast<void*>(w.list_delta_d);
    a.flags = w.flags;
    a.amber = amber && c->topo.has_excl ? 1 : 0;
    a.excl_start = c->topo.excl_start;
    a.excl_list = c->topo.excl_list;
    a.excl_rmax = w.excl_rmax;
    const int nb = (c->n_clusters + kSearchWaves - 1) / kSearchWaves;
    // half-extent bound of the minimum-image search in each dimension
    auto hlim = [&](double L) { return 0.5 * (0.5 * L - rl); };
    const double hx = hlim(g.box[0]), hy = hlim(g.box[1]), hz = hlim(g.box[2]);
    const bool fast = !c->search_general && hx > 1.0 && hy > 1.0 && hz > 1.0;
    if (fast) {
        if (mixed)
            hipLaunchKernelGGL(k_search_mi<float>, dim3(nb), dim3(kSearchWaves * kWave), 0, c->stream, a, hx, hy, hz);
        else
            hipLaunchKernelGGL(k_search_mi<double>, dim3(nb), dim3(kSearchWaves * kWave), 0, c->stream, a, hx, hy, hz);
    } else {
        if (mixed) hipLaunchKernelGGL(k_search<float>, dim3(nb), dim3(kSearchWaves * kWave), 0, c->stream, a);
        else hipLaunchKernelGGL(k_search<double>, dim3(nb), dim3(kSearchWaves * kWave), 0, c->stream, a);
    }
    return hipGetLastError();
}

}  // namespace mdg
