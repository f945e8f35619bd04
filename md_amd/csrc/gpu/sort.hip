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

constexpr int kColMax = kColBucket;  // atoms per column the in-LDS sort handles

// Column binning in two launches.
//   k_bin:     each block wraps a contiguous chunk of atoms, histograms their
//              columns in LDS, reserves its run in every column it touches
//              with one global atomic per (block, column) and writes its
//              atoms' sort keys into the runs.  Each column owns a bucket of
//              kColMax keys, so no scan over the columns is needed first; the
//              order of the runs inside a bucket is whatever the atomics give,
//              and the column sort below makes the result independent of it.
//              The same blocks also measure the excluded and 1-4 pair
//              distances.
//   k_col_sort: one block per column sums the totals of the columns before
//              its own (its cluster offset), ranks its atoms by (z, index),
//              writes the slot map and sets up the column's clusters.  The
//              last block to have read the totals clears them for the next
//              evaluation; block 0 also folds the exclusion extent.
constexpr int kBinThreads = 1024;
constexpr int kMaxCols = 8192;
constexpr int kMaxSortBlocks = 256;
constexpr int kAtomsPerSortBlock = 512;

// Excluded and 1-4 pairs whose minimum-image distances bound the tiles that
// can hold an excluded pair (n1 = n2 = 0: not needed).
struct ExclPairs {
    const int2* p1;
    int n1;
    const int2* p2;
    int n2;
    int chunk;         // pairs per sort block
    float* block_max;  // [sort blocks]
    float* out;        // largest distance x 1.0001
};

// The sort key of an atom: its wrapped z as 32-bit fixed point of the box
// height (zscale = 2^32 / L_z) above its index, so one unsigned comparison
// orders atoms by z and breaks ties by index.
__device__ __forceinline__ unsigned long long z_key(double z, double zscale, int atom) {
    const double q = floor(z * zscale);
    const unsigned long long zq = q <= 0.0 ? 0ull : (q >= 4294967295.0 ? 4294967295ull : static_cast<unsigned long long>(q));
    return (zq << 32) | static_cast<unsigned int>(atom);
}

__global__ __launch_bounds__(kBinThreads) void k_bin(const double* __restrict__ x, int n, int chunk, const Cell cell,
                                                     int ncx, int ncy, double sx, double sy, double zscale,
                                                     double* __restrict__ wrapped, int* __restrict__ col_of,
                                                     int* __restrict__ col_count,
                                                     unsigned long long* __restrict__ col_key, unsigned int* flags,
                                                     const ExclPairs ex, double* __restrict__ xref,
                                                     const VerletState* __restrict__ vs) {
    __shared__ int hist[kMaxCols];
    __shared__ float wmax[kBinThreads / kWave];
    if (vs && !vs->build) return;  // Verlet list kept: nothing to sort
    const int ncol = ncx * ncy;
    for (int c = threadIdx.x; c < ncol; c += blockDim.x) hist[c] = 0;
    __syncthreads();
    const double lx = cell.l[0], ly = cell.l[1], lz = cell.l[2];
    const int i0 = blockIdx.x * chunk, i1 = min(n, i0 + chunk);
    for (int i = i0 + threadIdx.x; i < i1; i += blockDim.x) {
        double wx, wy, wz;
        int cx, cy;
        if (!cell.tri) {
            // fractional wrapping in FP64: x - L floor(x / L)
            wx = x[3 * i] - lx * floor(x[3 * i] / lx);
            wy = x[3 * i + 1] - ly * floor(x[3 * i + 1] / ly);
            wz = x[3 * i + 2] - lz * floor(x[3 * i + 2] / lz);
            cx = static_cast<int>(wx / sx);
            cy = static_cast<int>(wy / sy);
        } else {
            // triclinic: wrap the fractional coordinates into [0, 1) and bin
            // by them; the wrapped position is H s
            double f[3];
            cell_frac(cell, x[3 * i], x[3 * i + 1], x[3 * i + 2], f);
            for (int d = 0; d < 3; ++d) f[d] -= floor(f[d]);
            wx = cell.ax * f[0] + cell.bx * f[1] + cell.cx * f[2];
            wy = cell.by * f[1] + cell.cy * f[2];
            wz = cell.cz * f[2];
            cx = static_cast<int>(f[0] * ncx);
            cy = static_cast<int>(f[1] * ncy);
        }
        wrapped[3 * i] = wx;
        wrapped[3 * i + 1] = wy;
        wrapped[3 * i + 2] = wz;
        if (xref) {
            // the positions the Verlet list is built from
            xref[3 * i] = x[3 * i];
            xref[3 * i + 1] = x[3 * i + 1];
            xref[3 * i + 2] = x[3 * i + 2];
        }
        cx = cx < 0 ? 0 : (cx >= ncx ? ncx - 1 : cx);
        cy = cy < 0 ? 0 : (cy >= ncy ? ncy - 1 : cy);
        const int c = cx + ncx * cy;
        col_of[i] = c;
        atomicAdd(hist + c, 1);
    }
    // excluded / 1-4 pair distances of this block's share
    float r = 0.f;
    const int ne = ex.n1 + ex.n2;
    const int e0 = blockIdx.x * ex.chunk, e1 = min(ne, e0 + ex.chunk);
    for (int e = e0 + threadIdx.x; e < e1; e += blockDim.x) {
        const int2 p = e < ex.n1 ? ex.p1[e] : ex.p2[e - ex.n1];
        double dx = x[3 * p.x] - x[3 * p.y], dy = x[3 * p.x + 1] - x[3 * p.y + 1], dz = x[3 * p.x + 2] - x[3 * p.y + 2];
        if (!cell.tri) {
            dx -= lx * rint(dx / lx);
            dy -= ly * rint(dy / ly);
            dz -= lz * rint(dz / lz);
        } else {
            cell_min_image_tri(cell, dx, dy, dz);
        }
        r = fmaxf(r, static_cast<float>(sqrt(dx * dx + dy * dy + dz * dz)) * 1.0001f);
    }
    for (int m = 32; m >= 1; m >>= 1) r = fmaxf(r, __shfl_xor(r, m, kWave));
    if ((threadIdx.x & 63) == 0) wmax[threadIdx.x >> 6] = r;
    __syncthreads();
    if (threadIdx.x == 0 && ne > 0) {
        float m = 0.f;
        for (int q = 0; q < kBinThreads / kWave; ++q) m = fmaxf(m, wmax[q]);
        ex.block_max[blockIdx.x] = m;
    }
    // reserve this block's run in each column it touches; hist[c] becomes
    // the run's start in the column
    for (int c = threadIdx.x; c < ncol; c += blockDim.x) {
        const int h = hist[c];
        hist[c] = h > 0 ? atomicAdd(col_count + c, h) : 0;
    }
    __syncthreads();
    // place this block's atoms in their columns' buckets (each thread walks
    // the atoms it wrapped above, so it reads back its own writes)
    bool over = false;
    for (int i = i0 + threadIdx.x; i < i1; i += blockDim.x) {
        const int c = col_of[i];
        const int p = atomicAdd(hist + c, 1);
        if (p < kColMax) col_key[static_cast<size_t>(c) * kColMax + p] = z_key(wrapped[3 * i + 2], zscale, i);
        else over = true;
    }
    if (over) atomicOr(flags + 2, 1u);  // a column holds more atoms than the column sort handles
}

template <typename real>
struct Vec4;
template <>
struct Vec4<float> { using T = float4; };
template <>
struct Vec4<double> { using T = double4; };

// PME geometry for the per-slot spline records (k[0] == 0: no records).
struct PmeSlots {
    int k[3];
    Cell cell;
    int* fl;
    void* w;   // double4 / float4 per slot: fractions (x, y, z) and charge
};

// PME spline record of a position: u = K x / L in FP64 (u = K s, s = R x,
// in a triclinic cell); the base index
// floor(u) taken mod K (the stencil wraps anyway), packed 10 bits per
// dimension, and the fractions u - floor(u).  The position may lie outside
// the box (a Verlet-list atom that moved across a face since the build).
template <typename real>
__device__ __forceinline__ int pme_record(const double xs[3], const PmeSlots& pme, real wv[3]) {
    int packed = 0;
    double fr[3] = {0.0, 0.0, 0.0};
    if (pme.cell.tri) cell_frac(pme.cell, xs[0], xs[1], xs[2], fr);
    for (int d = 0; d < 3; ++d) {
        const double u = pme.cell.tri ? fr[d] * pme.k[d] : xs[d] * pme.k[d] / pme.cell.l[d];
        const double f = floor(u);
        int fl = static_cast<int>(f) % pme.k[d];
        fl += fl < 0 ? pme.k[d] : 0;
        packed |= fl << (10 * d);
        wv[d] = static_cast<real>(u - f);
    }
    return packed;
}

// Per-cluster outputs of the column sort.
template <typename real>
struct ClusterOut {
    int max_clusters;
    const double* charge;
    const int* type;
    int pad_type;
    int* slot_atom;
    int* atom_slot;          // inverse map, atom -> slot
    int* cluster_col;
    typename Vec4<real>::T* pos;
    int2* slot_meta;
    double4* origin;
    float4* half;
    PmeSlots pme;
    int n_atoms;             // atoms of the system (slot records outside [0, n_atoms) are padding)
};

constexpr int kSortThreads = 256;

// One block per column: rank the column's atoms by their sort key (z, then
// index), then set up its clusters with eight lanes per cluster: bounding box
// by shuffles, the coordinates relative to the box centre, type and charge,
// and the PME spline record.  Padding slots get -1, the padding type, zero
// charge and an offset of kFarAway.
template <typename real>
__global__ __launch_bounds__(kSortThreads) void k_col_sort(int ncol, int* __restrict__ col_count,
                                                           unsigned int* __restrict__ ticket,
                                                           int* __restrict__ col_cstart,
                                                           const unsigned long long* __restrict__ col_key,
                                                           const double* __restrict__ wrapped, const ExclPairs ex,
                                                           int nbin, const ClusterOut<real> o,
                                                           const VerletState* __restrict__ vs) {
    constexpr int kPer = kColMax / kSortThreads;  // atoms per thread at most
    __shared__ unsigned long long s_key[kColMax];  // the column's sort keys
    __shared__ int s_sorted[kColMax];              // atom by rank
    __shared__ int s_red[kSortThreads / kWave];
    __shared__ int s_n, s_last;
    if (vs && !vs->build) return;
    const int c = blockIdx.x;
    // cluster offset of this column: clusters of the columns before it (a
    // column beyond kColMax atoms keeps kColMax of them; k_bin has flagged it)
    int before = 0;
    for (int q = threadIdx.x; q < c; q += kSortThreads) before += (min(col_count[q], kColMax) + kCluster - 1) / kCluster;
    for (int m = kWave / 2; m >= 1; m >>= 1) before += __shfl_xor(before, m, kWave);
    if ((threadIdx.x & (kWave - 1)) == 0) s_red[threadIdx.x / kWave] = before;
    // read into LDS, so the load has completed before the ticket below
    if (threadIdx.x == 0) s_n = min(col_count[c], kColMax);
    __syncthreads();
    int cl0 = 0;
    for (int w = 0; w < kSortThreads / kWave; ++w) cl0 += s_red[w];
    const int n = s_n;
    const int ncl = (n + kCluster - 1) / kCluster;
    if (threadIdx.x == 0) {
        col_cstart[c] = cl0;
        if (c == ncol - 1) col_cstart[ncol] = cl0 + ncl;
        if (c == 0 && ex.n1 + ex.n2 > 0) {
            float m = 0.f;
            for (int b = 0; b < nbin; ++b) m = fmaxf(m, ex.block_max[b]);
            *ex.out = m;
        }
        // every block has read the totals once it has taken its ticket; the
        // last one clears them (and the ticket) for the next evaluation
        __threadfence();
        s_last = atomicAdd(ticket, 1u) == gridDim.x - 1 ? 1 : 0;
    }
    __syncthreads();
    if (s_last) {
        for (int q = threadIdx.x; q < ncol; q += kSortThreads) col_count[q] = 0;
        if (threadIdx.x == 0) atomicExch(ticket, 0u);
    }
    const unsigned long long* bucket = col_key + static_cast<size_t>(c) * kColMax;
    unsigned long long key[kPer];
#pragma unroll
    for (int p = 0; p < kPer; ++p) {
        const int k = static_cast<int>(threadIdx.x) + p * kSortThreads;
        key[p] = k < n ? bucket[k] : 0ull;
        if (k < n) s_key[k] = key[p];
    }
    __syncthreads();
    int rank[kPer];
#pragma unroll
    for (int p = 0; p < kPer; ++p) {
        const int k = static_cast<int>(threadIdx.x) + p * kSortThreads;
        rank[p] = 0;
        if (k < n) {
            int r = 0;
            for (int m = 0; m < n; ++m) r += s_key[m] < key[p] ? 1 : 0;  // keys are distinct (they hold the index)
            rank[p] = r;
        }
    }
    __syncthreads();
#pragma unroll
    for (int p = 0; p < kPer; ++p) {
        const int k = static_cast<int>(threadIdx.x) + p * kSortThreads;
        if (k < n) s_sorted[rank[p]] = static_cast<int>(key[p] & 0xffffffffull);
    }
    __syncthreads();
    const int nslots = min(ncl * kCluster, kColMax);
    // whole waves step together so the 8 lanes of a cluster stay in one wave
    for (int k0 = 0; k0 < nslots; k0 += blockDim.x) {
        const int k = k0 + static_cast<int>(threadIdx.x);
        const int slot = cl0 * kCluster + k;
        const int cl = slot / kCluster;
        const bool live = k < nslots && cl < o.max_clusters;
        int atom = live && k < n ? s_sorted[k] : -1;
        // the ranks are a permutation of the column's keys (each atom's key
        // is placed once); an index outside the system never reaches memory
        if (atom >= o.n_atoms) atom = -1;
        // the atom's coordinates, charge and type in one round trip
        double x = 0, y = 0, z = 0, q = 0;
        int type = o.pad_type;
        if (atom >= 0) {
            x = wrapped[3 * atom];
            y = wrapped[3 * atom + 1];
            z = wrapped[3 * atom + 2];
            q = o.charge[atom];
            type = o.type[atom];
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
        if (!live) continue;
        o.slot_atom[slot] = atom;
        const double ox = 0.5 * (lo[0] + hi[0]), oy = 0.5 * (lo[1] + hi[1]), oz = 0.5 * (lo[2] + hi[2]);
        typename Vec4<real>::T p;
        if (atom >= 0) {
            o.atom_slot[atom] = slot;
            p.x = static_cast<real>(x - ox);
            p.y = static_cast<real>(y - oy);
            p.z = static_cast<real>(z - oz);
            p.w = static_cast<real>(q);
            o.slot_meta[slot] = make_int2(type, atom);
        } else {
            p.x = p.y = p.z = static_cast<real>(kFarAway);
            p.w = 0;
            o.slot_meta[slot] = make_int2(o.pad_type, -1);
        }
        o.pos[slot] = p;
        if (o.pme.k[0] > 0) {
            int packed = 0;
            typename Vec4<real>::T wq = {0, 0, 0, 0};
            if (atom >= 0) {
                const double xs[3] = {x, y, z};
                real wv[3];
                packed = pme_record<real>(xs, o.pme, wv);
                wq.x = wv[0];
                wq.y = wv[1];
                wq.z = wv[2];
                wq.w = static_cast<real>(q);
            }
            o.pme.fl[slot] = packed;
            reinterpret_cast<typename Vec4<real>::T*>(o.pme.w)[slot] = wq;
        }
        if ((slot & (kCluster - 1)) == 0) {
            o.cluster_col[cl] = c;
            o.origin[cl] = make_double4(ox, oy, oz, 0.0);
            // round the half extents up so the float box still contains every atom
            o.half[cl] = make_float4(static_cast<float>(0.5 * (hi[0] - lo[0])) * 1.0001f + 1e-4f,
                                     static_cast<float>(0.5 * (hi[1] - lo[1])) * 1.0001f + 1e-4f,
                                     static_cast<float>(0.5 * (hi[2] - lo[2])) * 1.0001f + 1e-4f, 0.f);
        }
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

// Half list: a pair of distinct clusters (i, j) is listed with i when j lies
// in the circular window of indices that starts at i and spans half of the n
// clusters (for even n, the pair half way round goes to the lower index), so
// every pair is listed once.  Cluster indices run column by column, so this
// window holds the columns ahead of i, wrapping past the last column: every
// cluster gets about half of its neighbours.  (The rule j >= i gave the first
// columns nearly full lists and the last ones nearly empty lists, since their
// periodic neighbours wrap to low indices; the pair kernel's waves then
// differed by 5x in work.)
__device__ __forceinline__ int half_window(int ci, int n) {
    return (n & 1) ? (n + 1) / 2 : n / 2 + (ci < n / 2 ? 1 : 0);
}
__device__ __forceinline__ bool in_half(int ci, int cj, int n) {
    int off = cj - ci;
    off += off < 0 ? n : 0;
    return off < half_window(ci, n);
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
    double sx, sy;          // distance between a column's opposite faces
    double lx, ly, lz;      // edge lengths (lz: the height cz of a triclinic cell)
    Cell cell;
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
    // slot positions relative to the cluster origin (float4 / double4) for the
    // atom-box test of the fast search
    const void* pos;
    double rp2;             // (r_list + margin)^2 for that test
    const VerletState* vs;  // Verlet mode: search only when the list is rebuilt
};

constexpr int kNearCap = 256;
constexpr int kSearchWaves = 4;

template <typename real>
__global__ __launch_bounds__(kSearchWaves * kWave) void k_search(const SearchArgs a) {
    using R4 = typename Vec4<real>::T;
    __shared__ int2 s_near[kSearchWaves][kNearCap];  // (entry, j-cluster)
    if (a.vs && !a.vs->build) return;
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
    for (int dy = -a.dcy; dy <= a.dcy; ++dy) {
        const int jyr = cy + dy, ky = floordiv(jyr, a.ncy), jy = jyr - ky * a.ncy;
        const double gy = fmax(0.0, (abs(dy) - 1) * a.sy);
        if (gy * gy >= a.rl2) continue;
        for (int dx = -a.dcx; dx <= a.dcx; ++dx) {
            const int jxr = cx + dx, kx = floordiv(jxr, a.ncx), jx = jxr - kx * a.ncx;
            const double gx = fmax(0.0, (abs(dx) - 1) * a.sx);
            // the column gaps bound the distance along the face normals; in a
            // triclinic cell the two normals are not orthogonal
            if ((a.cell.tri ? fmax(gx * gx, gy * gy) : gx * gx + gy * gy) >= a.rl2) continue;
            const int jcol = jx + a.ncx * jy;
            const int c0 = a.col_cstart[jcol], c1 = a.col_cstart[jcol + 1];
            const double shx = a.cell.tri ? kx * a.cell.ax + ky * a.cell.bx : kx * a.lx;
            const double shy = a.cell.tri ? ky * a.cell.by : ky * a.ly;
            // half list: the j-clusters in the half window after ci
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
                    const double tx = a.cell.tri ? shx + kz * a.cell.cx : shx;
                    const double ty = a.cell.tri ? shy + kz * a.cell.cy : shy;
                    const double ddx = oj.x + tx - oi.x, ddy = oj.y + ty - oi.y, ddz = oj.z + kz * a.lz - oi.z;
                    const double qx = fmax(0.0, fabs(ddx) - (double)(hi.x + hj.x));
                    const double qy = fmax(0.0, fabs(ddy) - (double)(hi.y + hj.y));
                    const double qz = fmax(0.0, fabs(ddz) - (double)(hi.z + hj.z));
                    const double d2 = qx * qx + qy * qy + qz * qz;
                    const bool noshift = kx == 0 && ky == 0 && kz == 0;
                    // a cluster with its own image: one of each pair of opposite
                    // shifts (the other holds the same atom pairs), all 64 pairs
                    const bool self_up = kz > 0 || (kz == 0 && (ky > 0 || (ky == 0 && kx > 0)));
                    const bool cand =
                        valid && (cj == ci ? (noshift || self_up) : in_half(ci, cj, ncl)) && d2 < a.rl2;
                    const bool triangle = cj == ci && noshift;  // the cluster with itself: j > i only
                    const unsigned long long bal = __ballot(cand);
                    if (bal == 0) continue;
                    const int rank = __popcll(bal & ((1ull << lane) - 1ull));
                    const int pos = count + rank;
                    const bool near = a.amber && cand && d2 < rex2;
                    if (cand && pos < a.cap) {
                        const size_t e = base_i + pos;
                        a.list_j[e] = cj;
                        // masks of tiles that may hold excluded pairs are written below
                        if (!near) a.list_mask[e] = triangle ? kSelfMask : ~0ull;
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
                        // bit 30 of the entry: a shifted image of the cluster itself
                        if (nr < kNearCap) s_near[wv][nr] = make_int2(pos | (cj == ci && !noshift ? 1 << 30 : 0), cj);
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
    const int ti = lane & 7, tj = lane >> 3;  // the ballot's bit of the pair is mask_bit(ti, tj)
    const int ai = a.slot_atom[ci * kCluster + ti];
    __builtin_amdgcn_wave_barrier();
    for (int q = 0; q < nnear; ++q) {
        const int2 en = s_near[wv][q];
        const int aj = a.slot_atom[en.y * kCluster + tj];
        const bool ex = ai >= 0 && aj >= 0 && pair_excluded(a.excl_start, a.excl_list, ai, aj);
        const unsigned long long bal = __ballot(ex);
        const bool shifted = (en.x >> 30) & 1;
        const int pos = en.x & ((1 << 30) - 1);
        if (lane == 0) a.list_mask[base_i + pos] = ((en.y == ci && !shifted) ? kSelfMask : ~0ull) & ~bal;
    }
}

// Exclusion masks of the near tiles of one wave's i-cluster.  The sorted
// exclusion lists of the cluster's eight atoms are staged in LDS (a list
// longer than kExclCap is searched in memory instead).  Eight tiles are
// handled per step: one load brings their j atoms (lane 8u + tj holds atom tj
// of tile u), and lane (ti, tj) scans atom ti's list once, comparing every
// entry with its j atom in all eight tiles.
constexpr int kExclCap = 32;

// The i-cluster's exclusion lists as registers, loaded in three dependent
// steps (the atoms of the cluster, their list bounds, the entries) that the
// search issues beside its own loads, so that each rides on a round trip the
// search makes anyway: the cluster's own records, the column ranges, and the
// near tiles' j atoms.
struct ExclRegs {
    int a_l, e0, en;
    int v[kCluster];  // entry `lane` of atom t's list
};

// second step: list bounds of lane's atom (lanes 0..7; a_l < 0 elsewhere),
// kept as loaded (en holds the end until the third step)
__device__ __forceinline__ void excl_bounds(const SearchArgs& a, ExclRegs& r) {
    const int k = r.a_l >= 0 ? r.a_l : 0;
    r.e0 = a.excl_start[k];
    r.en = a.excl_start[k + 1];
}

// third step: entry `lane` of each atom's list (the first kExclCap)
__device__ __forceinline__ void excl_entries(const SearchArgs& a, ExclRegs& r, int lane) {
    r.en = r.a_l >= 0 ? r.en - r.e0 : 0;
    r.e0 = r.a_l >= 0 ? r.e0 : 0;
#pragma unroll
    for (int t = 0; t < kCluster; ++t) {
        const int n_t = min(__shfl(r.en, t, kWave), kExclCap), s_t = __shfl(r.e0, t, kWave);
        r.v[t] = lane < n_t ? a.excl_list[s_t + lane] : -1;
    }
}

// The j atoms of the first kNearPre groups of eight near tiles are loaded at
// once, so their latency is paid once rather than once per group.
constexpr int kNearPre = 6;

__device__ void near_tile_masks(const SearchArgs& a, int ci, int lane, const ExclRegs& xr, int (*s_ex)[kExclCap],
                                const int2* near, int nnear, size_t base_i) {
    const int a_l = xr.a_l, en = xr.en;
    const int ti = lane & 7, tj = lane >> 3;  // the ballot's bit of the pair is mask_bit(ti, tj)
    // lane 8u + t: atom t of near tile 8g + u
    int aj_pre[kNearPre];
#pragma unroll
    for (int g = 0; g < kNearPre; ++g) {
        const int qq = kCluster * g + (lane >> 3);
        aj_pre[g] = qq < nnear ? a.slot_atom[near[qq].y * kCluster + (lane & 7)] : -1;
    }
#pragma unroll
    for (int t = 0; t < kCluster; ++t)
        if (lane < kExclCap) s_ex[t][lane] = xr.v[t];
    __builtin_amdgcn_wave_barrier();
    const int ai = __shfl(a_l, ti, kWave), my_n = __shfl(en, ti, kWave);
    const bool in_lds = my_n <= kExclCap;
    // longest staged list of the cluster (uniform loop bound)
    int nmax = in_lds ? my_n : 0;
    for (int m = 32; m >= 1; m >>= 1) nmax = max(nmax, __shfl_xor(nmax, m, kWave));
    // the masks of near tiles q0 .. q0 + 7, from this lane's j atom aj_l
    auto group = [&](int q0, int aj_l) {
        int aj[kCluster];
#pragma unroll
        for (int u = 0; u < kCluster; ++u) aj[u] = __shfl(aj_l, kCluster * u + tj, kWave);
        unsigned int exbits = 0;  // bit u: this lane's pair is excluded in tile q0 + u
        if (ai >= 0) {
            if (in_lds) {
                for (int k = 0; k < nmax; ++k) {
                    const int x = k < my_n ? s_ex[ti][k] : -2;
#pragma unroll
                    for (int u = 0; u < kCluster; ++u) exbits |= (x == aj[u] ? 1u : 0u) << u;
                }
            } else {
#pragma unroll
                for (int u = 0; u < kCluster; ++u)
                    if (aj[u] >= 0 && pair_excluded(a.excl_start, a.excl_list, ai, aj[u])) exbits |= 1u << u;
            }
        }
#pragma unroll
        for (int u = 0; u < kCluster; ++u) {
            if (q0 + u >= nnear) break;
            const unsigned long long bal = __ballot((exbits >> u) & 1u);
            const int2 en_q = near[q0 + u];
            if (lane == 0) a.list_mask[base_i + en_q.x] = ((en_q.y == ci) ? kSelfMask : ~0ull) & ~bal;
        }
    };
#pragma unroll
    for (int g = 0; g < kNearPre; ++g)
        if (kCluster * g < nnear) group(kCluster * g, aj_pre[g]);
    for (int q0 = kCluster * kNearPre; q0 < nnear; q0 += kCluster) {
        const int qq = q0 + (lane >> 3);
        group(q0, qq < nnear ? a.slot_atom[near[qq].y * kCluster + (lane & 7)] : -1);
    }
}

__device__ __forceinline__ float bcast_lane(float v, int l) {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), l));
}
__device__ __forceinline__ double bcast_lane(double v, int l) {
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane(static_cast<int>(b & 0xffffffffll), l);
    const int hi = __builtin_amdgcn_readlane(static_cast<int>(b >> 32), l);
    return __longlong_as_double((static_cast<long long>(hi) << 32) | static_cast<unsigned int>(lo));
}

// Fast search: every candidate j-cluster is taken at its minimum-image shift
// (per lane, FP64 rint), so one pass over the clusters of each row of
// columns in reach replaces the per-column, per-shift enumeration.  It is
// exact when every cluster half-extent h satisfies 2h <= L/2 - r_list in each
// dimension (then no second image of a cluster can come within r_list);
// a cluster that breaks the bound sets flags[4] and the host falls back to
// k_search.  In a triclinic cell the separation is rounded in fractional
// coordinates (d - H rint(R d)) and the bound holds each cluster's half
// diagonal to half of (half the smallest face distance - r_list).
//
// The candidate clusters of the rows in reach form at most two contiguous
// index ranges per row (the row's x window, split where it wraps).  Lane r
// loads the bounds of range r, clipped to the circular half window; the
// ranges are then walked as one stream of candidates, 64 per step, with the
// next step's cluster boxes loaded while the current step is tested.  The
// stream's cluster indices are written once into LDS (streams of up to
// kCandCap candidates; a longer one maps each index through the ranges).
// 768 holds the longest streams of a water-density system at rc 12 A (467)
// or 10 A (675) and keeps a block's LDS at 24.6 KB, so six blocks fit a CU
// (as many as the registers allow) where 1,024 allowed five.
constexpr int kCandCap = 768;
template <typename real>
__global__ __launch_bounds__(kSearchWaves * kWave) void k_search_mi(const SearchArgs a, double hlx, double hly,
                                                                     double hlz) {
    using R4 = typename Vec4<real>::T;
    __shared__ int2 s_near[kSearchWaves][kNearCap];
    __shared__ int s_ex[kSearchWaves][kCluster][kExclCap];
    __shared__ int s_cand[kSearchWaves][kCandCap];
    const int lane = threadIdx.x & (kWave - 1), wv = threadIdx.x >> 6;
    const int ci_w = __builtin_amdgcn_readfirstlane(blockIdx.x * kSearchWaves + wv);
    // First round trip: the Verlet build flag and the cluster count with
    // everything indexed by the cluster alone.  The arrays hold max_clusters
    // entries and the index is clamped to them, so these loads stay in
    // bounds for a wave past the clusters, and none of their values is used
    // as an index before the checks below.
    const int ci = min(ci_w, a.max_clusters - 1);
    const R4* pos = static_cast<const R4*>(a.pos);
    const int build = a.vs ? a.vs->build : 1;
    const int ncl_dev = *a.ncl_dev;
    const int col = a.cluster_col[ci];
    const double4 oi = a.origin[ci];
    const float4 hi = a.half[ci];
    const R4 pi_l = pos[ci * kCluster + (lane & 7)];  // the i-cluster's atoms relative to its origin
    ExclRegs xr{};
    xr.a_l = a.amber && lane < kCluster ? a.slot_atom[ci * kCluster + lane] : -1;
    const float excl_rmax = a.amber ? *a.excl_rmax : 0.f;
    const int ncl = min(ncl_dev, a.max_clusters);
    // the whole wave leaves (the kernel has no block barrier): a kept
    // Verlet list, or a wave past the clusters
    if (!build || ci_w >= ncl) return;
    const int cx = col % a.ncx, cy = col / a.ncx;
    const int ny = min(2 * a.dcy + 1, a.ncy);
    if (hi.x > hlx || hi.y > hly || hi.z > hlz || 2 * ny > kWave) {
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
    if (a.amber) rex = static_cast<double>(excl_rmax) + 0.01;
    const double rex2 = rex * rex;
    // the i-cluster's atoms relative to its origin, in scalar registers (read
    // before the next loads are issued: the counter that a wait for a load
    // checks also covers every load issued after it)
    real pix[kCluster], piy[kCluster], piz[kCluster];
#pragma unroll
    for (int k = 0; k < kCluster; ++k) {
        pix[k] = bcast_lane(pi_l.x, k);
        piy[k] = bcast_lane(pi_l.y, k);
        piz[k] = bcast_lane(pi_l.z, k);
    }
    // second round trip: the exclusion-list bounds beside the column ranges
    if (a.amber) excl_bounds(a, xr);

    // ranges: lane r = 2 * row + segment, each clipped to the half window
    // [ci, ci + hw) (circular)
    int r_b0 = 0, r_len = 0;
    bool straddle = false;
    {
        const int hw = half_window(ci, ncl);
        const bool yall = ny == a.ncy;
        const int y0 = yall ? 0 : cy - a.dcy;
        const int ry = lane >> 1, seg = lane & 1;
        int dxr = a.dcx;
        bool use = lane < 2 * ny;
        if (use && !yall) {
            // columns of this row within reach: |dx| <= dxr, where the x gap
            // (|dx| - 1) sx and the y gap (|dy| - 1) sy stay inside r_list
            const int dy = ry - a.dcy;
            const double gy = fmax(0.0, (abs(dy) - 1) * a.sy);
            const double rem = a.rl2 - gy * gy;
            if (rem <= 0.0) use = false;
            // a triclinic cell's face normals are not orthogonal: no narrowing
            else if (!a.cell.tri) dxr = min(a.dcx, static_cast<int>(sqrt(rem) / a.sx) + 1);
        }
        if (use) {
            int jy = (y0 + ry) % a.ncy;
            jy += jy < 0 ? a.ncy : 0;
            const int nx = min(2 * dxr + 1, a.ncx);
            const int x0 = nx == a.ncx ? 0 : cx - dxr;
            int xa = x0 % a.ncx;
            xa += xa < 0 ? a.ncx : 0;
            const int first = min(nx, a.ncx - xa);
            const int c_lo = seg == 0 ? xa : 0;
            const int c_n = seg == 0 ? first : nx - first;
            if (c_n > 0) {
                const int rowc = a.ncx * jy;
                const int c0 = a.col_cstart[rowc + c_lo], c1 = a.col_cstart[rowc + c_lo + c_n];
                int b0 = max(c0, ci), b1 = min(c1, ci + hw);
                if (ci + hw > ncl) {
                    // the window wraps: [ci, ncl) and [0, w2)
                    const int w2 = ci + hw - ncl;
                    const bool hi_part = c1 > ci, lo_part = c0 < w2;
                    straddle = hi_part && lo_part;  // a range as long as half the system
                    if (lo_part && !hi_part) {
                        b0 = c0;
                        b1 = min(c1, w2);
                    } else {
                        b1 = c1;
                    }
                }
                if (b1 > b0) {
                    r_b0 = b0;
                    r_len = b1 - b0;
                }
            }
        }
    }
    if (__ballot(straddle) != 0ull) {
        // a row range reaching both ends of the window (tiny systems): the
        // general search handles it
        if (lane == 0) {
            atomicOr(a.flags + 4, 1u);
            a.list_count[ci] = 0;
        }
        return;
    }
    int r_inc = r_len;
    for (int d = 1; d < kWave; d <<= 1) {
        const int o = __shfl_up(r_inc, d, kWave);
        if (lane >= d) r_inc += o;
    }
    const int r_pre = r_inc - r_len;
    const int total = __shfl(r_inc, kWave - 1, kWave);
    const int nr = 2 * ny;
    // the stream in LDS: range k occupies [r_pre_k, r_pre_k + r_len_k)
    int* cand_s = s_cand[wv];
    const bool staged = total <= kCandCap;
    if (staged) {
        for (int k = 0; k < nr; ++k) {
            const int b0 = __builtin_amdgcn_readlane(r_b0, k), len = __builtin_amdgcn_readlane(r_len, k);
            const int pre = __builtin_amdgcn_readlane(r_pre, k);
            for (int q = lane; q < len; q += kWave) cand_s[pre + q] = b0 + q;
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    }
    // candidate q of the stream -> cluster index (-1 past the end).  Called
    // with every lane active: the shuffles read the range lanes, which must
    // not be masked off.
    auto cand_of = [&](int q) {
        if (staged) return q < total ? cand_s[q] : -1;
        int r = 0;
        for (int k = 1; k < nr; ++k) {
            const int pk = __builtin_amdgcn_readlane(r_pre, k), lk = __builtin_amdgcn_readlane(r_len, k);
            r = (lk > 0 && pk <= q) ? k : r;
        }
        const int cj = __shfl(r_b0, r, kWave) + (q - __shfl(r_pre, r, kWave));
        return q < total ? cj : -1;
    };

    const real rp2 = static_cast<real>(a.rp2);

    int count = 0, nnear = 0;
    // one step of 64 candidates (cj, origin, half of this lane's candidate)
    auto step = [&](const int cj, const double4& oj, const float4& hj) {
        const bool valid = cj >= 0;
        double ddx = 0, ddy = 0, ddz = 0, d2 = 1e300;
        if (valid) {
            ddx = oj.x - oi.x;
            ddy = oj.y - oi.y;
            ddz = oj.z - oi.z;
            if (!a.cell.tri) {
                ddx -= a.lx * rint(ddx * ilx);
                ddy -= a.ly * rint(ddy * ily);
                ddz -= a.lz * rint(ddz * ilz);
            } else {
                cell_min_image_tri(a.cell, ddx, ddy, ddz);
            }
            const double qx = fmax(0.0, fabs(ddx) - (double)(hi.x + hj.x));
            const double qy = fmax(0.0, fabs(ddy) - (double)(hi.y + hj.y));
            const double qz = fmax(0.0, fabs(ddz) - (double)(hi.z + hj.z));
            d2 = qx * qx + qy * qy + qz * qz;
        }
        bool cand = valid && d2 < a.rl2;
        unsigned long long bal = __ballot(cand);
        if (bal == 0) return;
        // second test: a tile can only hold a pair within r_list if one of the
        // i-atoms lies within r_list (plus a margin) of the j-cluster's box
        if (cand) {
            const real cx = static_cast<real>(ddx), cy = static_cast<real>(ddy), cz = static_cast<real>(ddz);
            const real hx = static_cast<real>(hj.x), hy = static_cast<real>(hj.y), hz = static_cast<real>(hj.z);
            bool any = false;
#pragma unroll
            for (int k = 0; k < kCluster; ++k) {
                const real qx = fmax(static_cast<real>(0), fabs(cx - pix[k]) - hx);
                const real qy = fmax(static_cast<real>(0), fabs(cy - piy[k]) - hy);
                const real qz = fmax(static_cast<real>(0), fabs(cz - piz[k]) - hz);
                any = any || qx * qx + qy * qy + qz * qz < rp2;
            }
            cand = any;
        }
        bal = __ballot(cand);
        if (bal == 0) return;
        const int pos = count + __popcll(bal & ((1ull << lane) - 1ull));
        const bool near = a.amber && cand && d2 < rex2;
        if (cand && pos < a.cap) {
            const size_t e = base_i + pos;
            a.list_j[e] = cj;
            if (!near) a.list_mask[e] = (cj == ci) ? kSelfMask : ~0ull;
            R4 dl;
            dl.x = static_cast<real>(-ddx);
            dl.y = static_cast<real>(-ddy);
            dl.z = static_cast<real>(-ddz);
            dl.w = 0;
            delta[e] = dl;
        }
        const unsigned long long nbal = __ballot(near && pos < a.cap);
        if (near && pos < a.cap) {
            const int nrk = nnear + __popcll(nbal & ((1ull << lane) - 1ull));
            if (nrk < kNearCap) s_near[wv][nrk] = make_int2(pos, cj);
        }
        nnear += __popcll(nbal);
        count += __popcll(bal);
    };
    // the boxes of candidate q of the stream (nothing past the end)
    auto load = [&](int q, int& cj, double4& oj, float4& hj) {
        cj = cand_of(q);
        if (cj >= 0) {
            oj = a.origin[cj];
            hj = a.half[cj];
        }
    };
    // Two register sets, each loaded two steps before its use, so that a
    // step's loads have a whole step to arrive.
    int cjA = -1, cjB = -1;
    double4 oA = make_double4(0, 0, 0, 0), oB = oA;
    float4 hA = make_float4(0, 0, 0, 0), hB = hA;
    load(lane, cjA, oA, hA);
    load(kWave + lane, cjB, oB, hB);
    for (int b = 0; b < total; b += 2 * kWave) {
        step(cjA, oA, hA);
        load(b + 2 * kWave + lane, cjA, oA, hA);
        if (b + kWave < total) {
            step(cjB, oB, hB);
            load(b + 3 * kWave + lane, cjB, oB, hB);
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
    // the exclusion-list entries, issued beside the near tiles' j atoms
    // (one round trip for both)
    excl_entries(a, xr, lane);
    __builtin_amdgcn_wave_barrier();
    near_tile_masks(a, ci, lane, xr, s_ex[wv], s_near[wv], nnear, base_i);
}

// Verlet list criterion (System::ljpairs_needs_rebuild, reference
// src/System/NonBondedUtils.cpp:87-100): the list is rebuilt when some atom's
// squared minimum-image displacement since the build exceeds (skin/2)^2.
// The displacement is formed as the reference forms it (d - L round(d / L),
// then dx*dx + dy*dy + dz*dz without contraction), so the decision is the
// same bit for bit; as in the reference a NaN never becomes the maximum.
// Each block folds its maximum into vs->max_d2; the last block to finish
// decides (also when the host has set vs->force), clears the state for the
// next evaluation and publishes the decision: in vs->build and flags[5] for
// an evaluation, in vs->query for a query.
constexpr int kDispThreads = 256;
constexpr int kDispMaxBlocks = 1024;
__global__ __launch_bounds__(kDispThreads) void k_disp(const double* __restrict__ x, const double* __restrict__ xref,
                                                       const double* __restrict__ xprune, int n, const Cell cell,
                                                       double lim2, double limp2, int query, VerletState* vs,
                                                       unsigned int* flags) {
#pragma clang fp contract(off)
    __shared__ double wmax[kDispThreads / kWave], wmaxp[kDispThreads / kWave];
    const double lx = cell.l[0], ly = cell.l[1], lz = cell.l[2];
    auto min_image = [&](double& dx, double& dy, double& dz) {
        if (!cell.tri) {
            dx = dx - lx * round(dx / lx);
            dy = dy - ly * round(dy / ly);
            dz = dz - lz * round(dz / lz);
        } else {
            cell_min_image_tri(cell, dx, dy, dz);
        }
    };
    double m = 0.0, mp = 0.0;
    for (int i = blockIdx.x * kDispThreads + threadIdx.x; i < n; i += gridDim.x * kDispThreads) {
        double dx = x[3 * i] - xref[3 * i];
        double dy = x[3 * i + 1] - xref[3 * i + 1];
        double dz = x[3 * i + 2] - xref[3 * i + 2];
        min_image(dx, dy, dz);
        const double d2 = dx * dx + dy * dy + dz * dz;
        if (d2 > m) m = d2;
        // displacement since the last prune of the kept list
        double px = x[3 * i] - xprune[3 * i];
        double py = x[3 * i + 1] - xprune[3 * i + 1];
        double pz = x[3 * i + 2] - xprune[3 * i + 2];
        min_image(px, py, pz);
        const double p2 = px * px + py * py + pz * pz;
        if (p2 > mp) mp = p2;
    }
    for (int k = 32; k >= 1; k >>= 1) {
        const double o = __shfl_xor(m, k, kWave);
        if (o > m) m = o;
        const double op = __shfl_xor(mp, k, kWave);
        if (op > mp) mp = op;
    }
    if ((threadIdx.x & (kWave - 1)) == 0) {
        wmax[threadIdx.x / kWave] = m;
        wmaxp[threadIdx.x / kWave] = mp;
    }
    __syncthreads();
    if (threadIdx.x != 0) return;
    for (int q = 1; q < kDispThreads / kWave; ++q) {
        if (wmax[q] > m) m = wmax[q];
        if (wmaxp[q] > mp) mp = wmaxp[q];
    }
    if (m > 0.0) atomicMax(&vs->max_d2, static_cast<unsigned long long>(__double_as_longlong(m)));
    if (mp > 0.0) atomicMax(&vs->max_p2, static_cast<unsigned long long>(__double_as_longlong(mp)));
    __threadfence();
    if (atomicAdd(&vs->ticket, 1u) != gridDim.x - 1) return;
    __threadfence();
    const double mx = __longlong_as_double(static_cast<long long>(atomicAdd(&vs->max_d2, 0ull)));
    const double mpx = __longlong_as_double(static_cast<long long>(atomicAdd(&vs->max_p2, 0ull)));
    const unsigned int build = (atomicAdd(&vs->force, 0u) != 0u || mx > lim2) ? 1u : 0u;
    atomicExch(&vs->max_d2, 0ull);
    atomicExch(&vs->max_p2, 0ull);
    atomicExch(&vs->ticket, 0u);
    if (query) {
        atomicExch(&vs->query, build);
    } else {
        const unsigned int prune = (build || mpx > limp2) ? 1u : 0u;
        atomicExch(&vs->build, build);
        atomicExch(&vs->prune, prune);
        atomicExch(&vs->force, 0u);
        if (build) atomicOr(flags + 5, 1u);
        if (prune) atomicOr(flags + 6, 1u);
    }
}

// Verlet list kept: every slot of the kept clusters follows its atom.  The
// new position is the build's wrapped position plus the minimum-image
// displacement since the build, so it stays within skin / 2 of the cluster
// it was sorted into; the offset from the (unchanged) cluster origin and the
// PME spline record are formed from it as the column sort forms them.
constexpr int kRefreshThreads = 256;
template <typename real>
__global__ __launch_bounds__(kRefreshThreads) void k_refresh(const VerletState* __restrict__ vs,
                                                             const int* __restrict__ ncl_dev,
                                                             const double* __restrict__ x,
                                                             const double* __restrict__ xref,
                                                             const double* __restrict__ wrapped,
                                                             const ClusterOut<real> o, const Cell cell) {
    if (vs->build) return;
    const int s = blockIdx.x * kRefreshThreads + threadIdx.x;
    if (s >= min(*ncl_dev, o.max_clusters) * kCluster) return;
    const int atom = o.slot_atom[s];
    if (atom < 0) return;
    double dd[3];
    for (int d = 0; d < 3; ++d) dd[d] = x[3 * atom + d] - xref[3 * atom + d];
    if (!cell.tri) {
        for (int d = 0; d < 3; ++d) dd[d] = dd[d] - cell.l[d] * round(dd[d] / cell.l[d]);
    } else {
        cell_min_image_tri(cell, dd[0], dd[1], dd[2]);
    }
    double xs[3];
    for (int d = 0; d < 3; ++d) xs[d] = wrapped[3 * atom + d] + dd[d];
    const double4 org = o.origin[s / kCluster];
    typename Vec4<real>::T p;
    p.x = static_cast<real>(xs[0] - org.x);
    p.y = static_cast<real>(xs[1] - org.y);
    p.z = static_cast<real>(xs[2] - org.z);
    p.w = static_cast<real>(o.charge[atom]);
    o.pos[s] = p;
    if (o.pme.k[0] > 0) {
        real wv[3];
        o.pme.fl[s] = pme_record<real>(xs, o.pme, wv);
        typename Vec4<real>::T wq;
        wq.x = wv[0];
        wq.y = wv[1];
        wq.z = wv[2];
        wq.w = static_cast<real>(o.charge[atom]);
        reinterpret_cast<typename Vec4<real>::T*>(o.pme.w)[s] = wq;
    }
}

// Verlet mode, when the displacement check set vs->prune (a build, or some
// atom moved more than kPruneBuffer / 2 since the last prune): one wave per
// i-cluster copies, in order, the entries of its kept list whose tile holds
// an unmasked pair within cutoff + kPruneBuffer at the current slot
// positions, computed as the pair kernel computes them, and records the
// cluster's atom positions.  Until the next prune no pair can close in by
// more than kPruneBuffer, and the pair kernel tests every pair against the
// cutoff, so the pruned list gives the same energies and forces with fewer
// tiles.  Each lane takes one entry: its eight j-atom loads are in flight
// together, and the i-cluster's atoms are broadcast from registers.
struct PruneArgs {
    const int* ncl_dev;
    int max_clusters;
    int cap;
    const int* count;
    const int* list_j;
    const unsigned long long* list_mask;
    const void* list_delta;
    const void* pos;
    int* out_count;
    int* out_j;
    unsigned long long* out_mask;
    void* out_delta;
    double rc2;
    const VerletState* vs;   // prune only when vs->prune is set
    const double* x;         // raw coordinates, recorded as the prune's positions
    double* x_prune;
    const int* slot_atom;
};

constexpr int kPruneWaves = 4;

template <typename real>
__global__ __launch_bounds__(kPruneWaves * kWave) void k_prune(const PruneArgs a) {
    using R4 = typename Vec4<real>::T;
    const int lane = threadIdx.x & (kWave - 1);
    const int ci = __builtin_amdgcn_readfirstlane(blockIdx.x * kPruneWaves + (threadIdx.x >> 6));
    const int ncl = min(*a.ncl_dev, a.max_clusters);
    if (ci >= ncl || !a.vs->prune) return;  // whole wave leaves; no block barrier follows
    if (lane < kCluster) {
        // the positions this prune is valid for
        const int atom = a.slot_atom[ci * kCluster + lane];
        if (atom >= 0)
            for (int d = 0; d < 3; ++d) a.x_prune[3 * atom + d] = a.x[3 * atom + d];
    }
    const R4* pos = static_cast<const R4*>(a.pos);
    const R4* din = static_cast<const R4*>(a.list_delta);
    R4* dout = static_cast<R4*>(a.out_delta);
    const R4 pi_l = pos[ci * kCluster + (lane & 7)];
    real pix[kCluster], piy[kCluster], piz[kCluster];
#pragma unroll
    for (int k = 0; k < kCluster; ++k) {
        pix[k] = bcast_lane(pi_l.x, k);
        piy[k] = bcast_lane(pi_l.y, k);
        piz[k] = bcast_lane(pi_l.z, k);
    }
    const real rc2 = static_cast<real>(a.rc2);
    const int cnt = min(a.count[ci], a.cap);
    const size_t base = static_cast<size_t>(ci) * a.cap;
    int n = 0;
    for (int kb = 0; kb < cnt; kb += kWave) {
        const bool have = kb + lane < cnt;
        int e_j = 0;
        unsigned long long e_m = 0;
        R4 e_d = {0, 0, 0, 0};
        bool live = false;
        if (have) {
            e_j = a.list_j[base + kb + lane];
            e_m = a.list_mask[base + kb + lane];
            e_d = din[base + kb + lane];
            R4 pj[kCluster];
#pragma unroll
            for (int u = 0; u < kCluster; ++u) pj[u] = pos[e_j * kCluster + u];
#pragma unroll
            for (int u = 0; u < kCluster; ++u) {
#pragma unroll
                for (int k = 0; k < kCluster; ++k) {
                    // k the i-slot and u the j-slot of the pair
                    const real dx = pix[k] - pj[u].x + e_d.x;
                    const real dy = piy[k] - pj[u].y + e_d.y;
                    const real dz = piz[k] - pj[u].z + e_d.z;
                    const bool near = dx * dx + dy * dy + dz * dz <= rc2;
                    live |= near & static_cast<bool>((e_m >> mask_bit(k, u)) & 1ull);
                }
            }
        }
        const unsigned long long keep = __ballot(live);
        if (live) {
            const size_t o = base + n + __popcll(keep & ((1ull << lane) - 1ull));
            a.out_j[o] = e_j;
            a.out_mask[o] = e_m;
            dout[o] = e_d;
        }
        n += __popcll(keep);
    }
    if (lane == 0) a.out_count[ci] = n;
}

}  // namespace

hipError_t launch_prune(mdgpu_ctx* c, bool mixed, const double* coords, const VerletState* vs) {
    Work& w = c->work;
    const Grid& g = c->grid;
    PruneArgs a{};
    a.ncl_dev = w.col_cstart + g.ncx * g.ncy;
    a.max_clusters = c->n_clusters;
    a.cap = w.list_cap;
    a.count = w.list_count;
    a.list_j = w.list_j;
    a.list_mask = w.list_mask;
    a.list_delta = mixed ? static_cast<const void*>(w.list_delta_f) : static_cast<const void*>(w.list_delta_d);
    a.pos = mixed ? static_cast<const void*>(w.slot_pos_f) : static_cast<const void*>(w.slot_pos_d);
    a.out_count = w.inner_count;
    a.out_j = w.inner_j;
    a.out_mask = w.inner_mask;
    a.out_delta = mixed ? static_cast<void*>(w.inner_delta_f) : static_cast<void*>(w.inner_delta_d);
    const double rc = std::max(c->params.lj_cutoff, c->params.ee_cutoff) + kPruneBuffer;
    a.rc2 = rc * rc;
    a.vs = vs;
    a.x = coords;
    a.x_prune = w.x_prune;
    a.slot_atom = w.slot_atom;
    const int nb = (c->n_clusters + kPruneWaves - 1) / kPruneWaves;
    if (mixed) hipLaunchKernelGGL(k_prune<float>, dim3(nb), dim3(kPruneWaves * kWave), 0, c->stream, a);
    else hipLaunchKernelGGL(k_prune<double>, dim3(nb), dim3(kPruneWaves * kWave), 0, c->stream, a);
    return hipGetLastError();
}

int sort_blocks(int n) {
    return std::max(1, std::min(kMaxSortBlocks, (n + kAtomsPerSortBlock - 1) / kAtomsPerSortBlock));
}

hipError_t launch_sort(mdgpu_ctx* c, const double* coords, bool mixed, bool excl, VerletState* vs) {
    Work& w = c->work;
    const Grid& g = c->grid;
    const Topo& t = c->topo;
    const int n = t.n_atoms;
    const int ncol = g.ncx * g.ncy;
    if (ncol > kMaxCols) return hipErrorInvalidValue;
    const int nblk = sort_blocks(n), chunk = (n + nblk - 1) / nblk;
    ExclPairs ex{t.excl_pairs, excl ? t.n_excl_pairs : 0, t.pairs14, excl ? t.n_pairs14 : 0, 0, w.block_emax,
                 w.excl_rmax};
    ex.chunk = (ex.n1 + ex.n2 + nblk - 1) / nblk;
    if (vs) {
        const int db = std::max(1, std::min(kDispMaxBlocks, (n + kDispThreads - 1) / kDispThreads));
        hipLaunchKernelGGL(k_disp, dim3(db), dim3(kDispThreads), 0, c->stream, coords, w.x_ref, w.x_prune, n,
                           g.cell, (0.5 * c->skin) * (0.5 * c->skin), (0.5 * kPruneBuffer) * (0.5 * kPruneBuffer), 0,
                           vs, w.flags);
    }
    // the sort key is the wrapped z over the cell height
    const double zscale = 4294967296.0 / (g.cell.tri ? g.cell.cz : g.box[2]);
    hipLaunchKernelGGL(k_bin, dim3(nblk), dim3(kBinThreads), 0, c->stream, coords, n, chunk, g.cell, g.ncx, g.ncy,
                       g.sx, g.sy, zscale, w.wrapped, w.col_of, w.col_count, w.col_key, w.flags, ex,
                       vs ? w.x_ref : nullptr, vs);
    PmeSlots pme{};
    const bool with_pme = c->params.semantics == MDGPU_SEM_AMBER && w.pme_k[0] > 0;
    for (int d = 0; d < 3; ++d) pme.k[d] = with_pme ? w.pme_k[d] : 0;
    pme.cell = g.cell;
    pme.fl = w.slot_pme_fl;
    const int rb = (c->n_clusters * kCluster + kRefreshThreads - 1) / kRefreshThreads;
    if (mixed) {
        pme.w = w.slot_pme_wf;
        const ClusterOut<float> o{c->n_clusters, t.charge, t.type, t.n_types, w.slot_atom, w.atom_slot, w.cluster_col,
                                  w.slot_pos_f, w.slot_meta, w.cl_origin, w.cl_half, pme, n};
        hipLaunchKernelGGL(k_col_sort<float>, dim3(ncol), dim3(kSortThreads), 0, c->stream, ncol, w.col_count,
                           w.col_ticket, w.col_cstart, w.col_key, w.wrapped, ex, nblk, o, vs);
        if (vs)
            hipLaunchKernelGGL(k_refresh<float>, dim3(rb), dim3(kRefreshThreads), 0, c->stream, vs,
                               w.col_cstart + ncol, coords, w.x_ref, w.wrapped, o, g.cell);
    } else {
        pme.w = w.slot_pme_wd;
        const ClusterOut<double> o{c->n_clusters, t.charge, t.type, t.n_types, w.slot_atom, w.atom_slot, w.cluster_col,
                                   w.slot_pos_d, w.slot_meta, w.cl_origin, w.cl_half, pme, n};
        hipLaunchKernelGGL(k_col_sort<double>, dim3(ncol), dim3(kSortThreads), 0, c->stream, ncol, w.col_count,
                           w.col_ticket, w.col_cstart, w.col_key, w.wrapped, ex, nblk, o, vs);
        if (vs)
            hipLaunchKernelGGL(k_refresh<double>, dim3(rb), dim3(kRefreshThreads), 0, c->stream, vs,
                               w.col_cstart + ncol, coords, w.x_ref, w.wrapped, o, g.cell);
    }
    return hipGetLastError();
}

hipError_t launch_disp_query(mdgpu_ctx* c, const double* coords) {
    Work& w = c->work;
    const Grid& g = c->grid;
    const int n = c->topo.n_atoms;
    const int db = std::max(1, std::min(kDispMaxBlocks, (n + kDispThreads - 1) / kDispThreads));
    hipLaunchKernelGGL(k_disp, dim3(db), dim3(kDispThreads), 0, c->stream, coords, w.x_ref, w.x_prune, n, g.cell,
                       (0.5 * c->skin) * (0.5 * c->skin), (0.5 * kPruneBuffer) * (0.5 * kPruneBuffer), 1, w.verlet,
                       w.flags);
    return hipGetLastError();
}

hipError_t launch_pairsearch(mdgpu_ctx* c, bool mixed, bool amber, const VerletState* vs) {
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
    a.pos = mixed ? static_cast<const void*>(w.slot_pos_f) : static_cast<const void*>(w.slot_pos_d);
    a.ncx = g.ncx;
    a.ncy = g.ncy;
    a.dcx = g.dcx;
    a.dcy = g.dcy;
    a.sx = g.wx;
    a.sy = g.wy;
    a.lx = g.box[0];
    a.ly = g.box[1];
    a.lz = g.cell.tri ? g.cell.cz : g.box[2];
    a.cell = g.cell;
    const double rl = g.r_list;
    a.vs = vs;
    a.rl2 = rl * rl;
    a.rp2 = (rl + 1e-3) * (rl + 1e-3);
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
    // half-extent bound of the minimum-image search in each dimension.  A
    // triclinic cell bounds each cluster's half diagonal by half of
    // (half the smallest face distance - r_list): two clusters whose boxes
    // come within r_list then have origins closer than half every face
    // distance, and the fractional rounding of their separation is that image.
    auto hlim = [&](double L) { return 0.5 * (0.5 * L - rl); };
    double hx = hlim(g.box[0]), hy = hlim(g.box[1]), hz = hlim(g.box[2]);
    if (g.cell.tri) hx = hy = hz = hlim(std::min({g.face[0], g.face[1], g.face[2]})) / std::sqrt(3.0);
    const bool fast = !c->search_general && hx > 1.0 && hy > 1.0 && hz > 1.0;
    c->search_fast = fast;
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
