// Nonbonded cluster-pair kernel: LJ 12-6 (switched or force-shifted) and the
// Coulomb forms, one 8 x 8 tile per wavefront step.
//
// Replaces the six pair-list passes and the O(N^2) pass of the reference:
// calculate_LJ_energy_pairlist (NonBondedEnergy.cpp:79-131),
// calculate_EE_energy (:133-174), calculate_EE_energy_pairlist (:176-204),
// calculate_EE_energy_pairlist_with_cutoff (:206-234),
// calculate_EE_ewald_direct_term (:236-268), calculate_forces_LJ_pairlist
// (NonBondedForces.cpp:92-154) and calculate_forces_EE_pairlist (:207-242).
//
// Lane l of a wave computes the pair (i = 8c_i + l%8, j = 8c_j + l/8).  The
// i-forces stay in registers over the wave's list entries; the j-forces of a
// tile are summed over the 8 lanes sharing j with DPP moves and then added
// by 24 lanes at once (one component each).  Forces go to 64-bit fixed point
// (mixed form) or FP64 atomics (FP64 form); energies to per-block FP64 rows.
//
// REF semantics: every pair within the cutoff, LJ energy U*S and force
// (S G - U dS/dr / r) d, energies of the plain, reaction-field and erfc
// Coulomb forms, plain Coulomb force k qq d / r^3.
// AMBER semantics: masked tiles (exclusions and 1-4 pairs removed), erfc
// Ewald real-space energy and force.
#include "bonded_terms.h"
#include "ctx.h"
#include "erfc_f32.h"

#include <type_traits>

namespace mdg {

namespace {

template <typename real>
struct V4;
template <>
struct V4<float> { using T = float4; };
template <>
struct V4<double> { using T = double4; };

// waves per block: 8 for the FP32 kernel (measured faster next to the
// concurrent PME chain), 4 for the FP64 kernel (fewer registers per lane left
// for a larger block)
template <typename real>
constexpr int kWavesOf = sizeof(real) == 4 ? 8 : 4;

template <typename real>
struct NbArgs {
    const int* ncl_dev;
    int max_clusters;
    int split;
    int cap;
    const int* list_count;
    const int* list_j;
    const unsigned long long* list_mask;
    const typename V4<real>::T* list_delta;
    const typename V4<real>::T* pos;
    const int2* meta;                  // per slot (LJ type, original atom or -1)
    const typename V4<real>::T* lj;    // (T+1)^2 x (A, B, U(rc), G(rc))
    const unsigned char* listed;       // (T+1)^2, 1 if the type pair has an LJ entry
    int ntp;                           // T + 1
    real rc2_lj, rc2_ee, rc_lj, rc_ee, rs, inv_sw;
    real alpha, kc, krf_crf;           // krf_crf: 1/rc_ee (reaction field with eps = 1)
    double* force_d;                   // slot-ordered force accumulators (3 per slot)
    FixedAcc fx;
    double* partial;
    int ld;                            // rows allocated in the slot-major partial buffer
    int row0;
    int n3;                            // 3 x slots: entries of the force arrays
};

template <typename real>
struct Math;
template <>
struct Math<float> {
    static __device__ __forceinline__ float rsqrt(float x) { return __builtin_amdgcn_rsqf(x); }
    static __device__ __forceinline__ float rcp(float x) { return __builtin_amdgcn_rcpf(x); }
    static __device__ __forceinline__ void erfc_exp(float x, float& ec, float& ex) {
        ex = __expf(-x * x);
        ec = erfc_f32(x, ex);
    }
};
template <>
struct Math<double> {
    static __device__ __forceinline__ double rsqrt(double x) { return 1.0 / sqrt(x); }
    static __device__ __forceinline__ double rcp(double x) { return 1.0 / x; }
    static __device__ __forceinline__ void erfc_exp(double x, double& ec, double& ex) {
        ex = exp(-x * x);
        ec = erfc(x);
    }
};

// Force contribution of lanes with on set: FP64 atomics (FP64 form); 64-bit
// fixed point (mixed form), converted in FP32 arithmetic for AMBER semantics
// (|F| < 2^31) and in FP64 for REF semantics, whose bonded neighbours interact
// with full LJ.  Every lane issues the buffer atomic; lanes that add nothing
// point it past the end of the array.  force_payload returns the value to add
// and sets the byte offset (kOffOOB for nothing).
template <typename real>
using Payload = typename std::conditional<sizeof(real) == 8, double, long long>::type;

template <typename real, bool kAmber>
__device__ __forceinline__ Payload<real> force_payload(const NbArgs<real>& a, bool on, int idx, real v,
                                                       unsigned int& off) {
    const unsigned int o = static_cast<unsigned int>(idx) * 8u;
    if constexpr (sizeof(real) == 8) {
        off = on && v != 0.0 ? o : kOffOOB;
        return v;
    } else {
        long long q;
        if constexpr (kAmber) q = fixed32_f32(v, a.fx.overflow);  // AMBER scale 2^32
        else q = to_fixed(static_cast<double>(v), a.fx.scale, a.fx.overflow);
        off = on && q != 0 ? o : kOffOOB;
        return q;
    }
}

template <typename real, bool kAmber>
__device__ __forceinline__ void add_force(const NbArgs<real>& a, BufRsrc r, bool on, int idx, real v) {
    unsigned int off;
    const Payload<real> q = force_payload<real, kAmber>(a, on, idx, v, off);
    buf_add(r, off, q);
}

// Sum over the 8 consecutive lanes of a group with three DPP moves
// (quad_perm [1,0,3,2], quad_perm [2,3,0,1], row_half_mirror); every lane of
// the group ends with the group total.
__device__ __forceinline__ int dpp_xor1(int v) { return __builtin_amdgcn_update_dpp(0, v, 0xB1, 0xF, 0xF, true); }
__device__ __forceinline__ int dpp_xor2(int v) { return __builtin_amdgcn_update_dpp(0, v, 0x4E, 0xF, 0xF, true); }
__device__ __forceinline__ int dpp_hmir(int v) { return __builtin_amdgcn_update_dpp(0, v, 0x141, 0xF, 0xF, true); }

__device__ __forceinline__ float sum8(float v) {
    v += __int_as_float(dpp_xor1(__float_as_int(v)));
    v += __int_as_float(dpp_xor2(__float_as_int(v)));
    v += __int_as_float(dpp_hmir(__float_as_int(v)));
    return v;
}
__device__ __forceinline__ double dpp_d(double v, int which) {
    const long long b = __double_as_longlong(v);
    int lo = static_cast<int>(b & 0xffffffffll), hi = static_cast<int>(b >> 32);
    if (which == 0) { lo = dpp_xor1(lo); hi = dpp_xor1(hi); }
    else if (which == 1) { lo = dpp_xor2(lo); hi = dpp_xor2(hi); }
    else { lo = dpp_hmir(lo); hi = dpp_hmir(hi); }
    return __longlong_as_double((static_cast<long long>(hi) << 32) | static_cast<unsigned int>(lo));
}
__device__ __forceinline__ double sum8(double v) {
    v += dpp_d(v, 0);
    v += dpp_d(v, 1);
    v += dpp_d(v, 2);
    return v;
}

// j-force sums of a tile by reduce-scatter over the eight lanes of a row
// (lanes 8 tj .. 8 tj + 7): lanes ti = 0, 1, 2 of the row end with the x, y
// and z totals, the other lanes with partial sums.  Round 1 pairs lanes
// l, l ^ 1 (lanes with e keep x and z, the others y), round 2 pairs l, l ^ 2
// (one value kept each), round 3 adds the mirrored lane 7 - l of the row's
// half, which holds the same component: the roles in the upper four lanes of
// the eight are swapped (the e and h patterns below) so that lanes 0, 1, 2
// meet lanes 7, 6, 5.  Four DPP adds and four selects instead of nine DPP
// adds and two selects for three full sums.
struct JSum {
    bool e, h;
};
__device__ __forceinline__ JSum jsum_lanes(int lane) {
    return {((lane ^ (lane >> 2)) & 1) == 0, (((lane >> 1) ^ (lane >> 2)) & 1) != 0};
}
__device__ __forceinline__ float dpp_f(float v, int which) {
    const int b = __float_as_int(v);
    return __int_as_float(which == 0 ? dpp_xor1(b) : (which == 1 ? dpp_xor2(b) : dpp_hmir(b)));
}
__device__ __forceinline__ double dpp_f(double v, int which) { return dpp_d(v, which); }
template <typename real>
__device__ __forceinline__ real jsum3(real x, real y, real z, const JSum& s) {
    const real k1 = s.e ? x : y, s1 = s.e ? y : x;
    const real a = k1 + dpp_f(s1, 0);
    const real b = z + dpp_f(z, 0);
    const real k2 = s.h ? b : a, s2 = s.h ? a : b;
    const real c = k2 + dpp_f(s2, 1);
    return c + dpp_f(c, 2);
}
// lanes ti = 0, 1, 2 of every row: the lanes that add a j-force component
constexpr unsigned long long kJLanes = 0x0707070707070707ull;

// fixed32_f32 without a branch: a contribution that does not fit adds 0 and
// marks its lane in `bad` (lanes outside `lanes` are not checked).
__device__ __forceinline__ long long fixed32_f32_flag(float v, unsigned long long& bad, unsigned long long lanes) {
    const bool ok = fabsf(v) < 2147483520.0f;  // false for NaN
    bad |= __builtin_amdgcn_ballot_w64(!ok) & lanes;
    const float w = ok ? v : 0.0f;
    const float h = floorf(w);
    const unsigned int lo = static_cast<unsigned int>(fminf((w - h) * 4294967296.0f, 4294967040.0f));
    const unsigned long long hi = static_cast<unsigned long long>(static_cast<unsigned int>(static_cast<int>(h)));
    return static_cast<long long>((hi << 32) | lo);
}

template <typename real>
__device__ __forceinline__ real bcast(real v, int lane);
template <>
__device__ __forceinline__ float bcast<float>(float v, int lane) {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), lane));
}
template <>
__device__ __forceinline__ double bcast<double>(double v, int lane) {
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane(static_cast<int>(b & 0xffffffffll), lane);
    const int hi = __builtin_amdgcn_readlane(static_cast<int>(b >> 32), lane);
    return __longlong_as_double((static_cast<long long>(hi) << 32) | static_cast<unsigned int>(lo));
}

// Lane l handles i = l % 8 and j = l / 8 of the tile, so the per-tile j-force
// sum runs over 8 consecutive lanes (DPP) and the i-force sum, done once per
// work item, over lanes 8 apart.  List entries are fetched 64 at a time, one
// per lane, and broadcast with readlane; the j-cluster data of a tile is
// loaded two tiles ahead into one of two register sets.  The force atomics
// are buffer atomics issued by every lane (see add_force), so the wait for a
// prefetched set never has to cover the atomic of the tile just finished.
//
// One wave per work item (i-cluster, list part); blocks are ordered so that
// consecutive clusters run on one XCD and share its L2.
//
// The launch also carries the bonded terms and the AMBER corrections as nbb +
// ncb further blocks after the npb pair blocks: they are dispatched last and
// fill the CUs that the pair blocks leave idle at the end of the launch.
// The FP32 AMBER instances without statistics are held to 6 waves per SIMD
// (at most 80 VGPRs, no spills; 24 waves per CU instead of 16): more waves
// hide the load and atomic latency of the tile loop.  The FP64, REF and
// statistics instances need more registers and keep the compiler's choice.
template <typename real, bool kAmber, bool kSwitch, bool kCount>
__global__ __launch_bounds__(kWavesOf<real> * kWave)
__attribute__((amdgpu_waves_per_eu(sizeof(real) == 4 && kAmber && !kCount ? 6 : 1))) void k_nonbonded(const NbArgs<real> a, const BArgs<real> ba,
                                                                      const CArgs ca, int npb, int nbb) {
    constexpr int kWaves = kWavesOf<real>;
    if (static_cast<int>(blockIdx.x) >= npb) {
        const int b = static_cast<int>(blockIdx.x) - npb;
        const int k = b * static_cast<int>(blockDim.x) + static_cast<int>(threadIdx.x);
        if (b < nbb) bonded_terms(ba, k, ba.row0 + b);
        else correction_pairs(ca, k - nbb * static_cast<int>(blockDim.x), ca.row0 + (b - nbb));
        return;
    }
    using R4 = typename V4<real>::T;
    using M = Math<real>;
    extern __shared__ unsigned char smem[];
    R4* s_lj = reinterpret_cast<R4*>(smem);
    unsigned char* s_listed = smem + sizeof(R4) * a.ntp * a.ntp;
    const int ntp2 = a.ntp * a.ntp;
    for (int k = threadIdx.x; k < ntp2; k += blockDim.x) {
        s_lj[k] = a.lj[k];
        if (!kAmber) s_listed[k] = a.listed[k];
    }
    __syncthreads();

    const int lane = threadIdx.x & 63;
    const int ti = lane & 7, tj = lane >> 3;
    const int ncl = min(*a.ncl_dev, a.max_clusters);
    const int mbit = mask_bit(ti, tj);  // mask bit of this lane's pair: the lane itself
    const JSum js = jsum_lanes(lane);

    // energy accumulators: FP64 form in double; mixed form in float, flushed to
    // double every 32 tiles
    double e_vdw = 0, e_all = 0, e_list = 0, e_rf = 0, e_ew = 0, n_pairs = 0;
    real s_vdw = 0, s_all = 0, s_list = 0, s_rf = 0, s_ew = 0;
    // statistics instance: virial of the pair forces, sum f d_a d_b
    double w_d[6] = {0, 0, 0, 0, 0, 0};
    real w_s[6] = {0, 0, 0, 0, 0, 0};
    const real tsp = static_cast<real>(1.1283791670955126) * a.alpha;  // 2 alpha / sqrt(pi)
    const real rc2 = a.rc2_lj > a.rc2_ee ? a.rc2_lj : a.rc2_ee;
    int flush = 0;
    // forces go to slot-ordered arrays: the 24 adds of a tile's j-cluster
    // fall on 192 consecutive bytes
    const BufRsrc rsrc = sizeof(real) == 8 ? buf_rsrc(a.force_d, 8u * a.n3) : buf_rsrc(a.fx.f, 8u * a.n3);
    const int blk = xcd_block(blockIdx.x, npb);
    const int item = __builtin_amdgcn_readfirstlane(blk * kWaves + (threadIdx.x >> 6));
    if (item < ncl * a.split) {
        const int ci = item / a.split, part = item % a.split;
        const int cnt = a.list_count[ci];
        const int k0 = static_cast<int>((static_cast<long long>(cnt) * part) / a.split);
        const int k1 = static_cast<int>((static_cast<long long>(cnt) * (part + 1)) / a.split);
        const R4 pi = a.pos[ci * kCluster + ti];
        const int2 mi = a.meta[ci * kCluster + ti];
        const int tyi = mi.x * a.ntp;
        const real qi = pi.w * a.kc;
        real fix = 0, fiy = 0, fiz = 0;
        const size_t base = static_cast<size_t>(ci) * a.cap;
        for (int kb = k0; kb < k1; kb += kWave) {
            const int nb = min(kWave, k1 - kb);
            // this lane's list entry of the batch
            int e_j = 0;
            unsigned long long e_m = 0;
            R4 e_d = {0, 0, 0, 0};
            if (lane < nb) {
                e_j = a.list_j[base + kb + lane];
                e_m = a.list_mask[base + kb + lane];
                e_d = a.list_delta[base + kb + lane];
            }
            const int e_mlo = static_cast<int>(e_m & 0xffffffffull), e_mhi = static_cast<int>(e_m >> 32);

            // j-cluster data of tile t (clamped to the batch, so every fetch
            // is issued and reads a valid cluster)
            auto fetch = [&](int t, R4& p, int2& m) {
                const int cj = __builtin_amdgcn_readlane(e_j, min(t, nb - 1));
                p = a.pos[cj * kCluster + tj];
                m = a.meta[cj * kCluster + tj];
            };
            // tile t of the batch; live = false runs it with an empty mask
            auto tile = [&](const R4& pcur, const int2& mcur, int t, bool live) {
                unsigned long long mask =
                    (static_cast<unsigned long long>(static_cast<unsigned int>(__builtin_amdgcn_readlane(e_mhi, t))) << 32) |
                    static_cast<unsigned int>(__builtin_amdgcn_readlane(e_mlo, t));
                mask = live ? mask : 0ull;
                const real dx = pi.x - pcur.x + bcast<real>(e_d.x, t);
                const real dy = pi.y - pcur.y + bcast<real>(e_d.y, t);
                const real dz = pi.z - pcur.z + bcast<real>(e_d.z, t);
                const real r2 = dx * dx + dy * dy + dz * dz;
                const bool act = ((mask >> mbit) & 1ull) && r2 <= rc2 && r2 >= static_cast<real>(1e-12);
                real f = 0;
                if (act) {
                    const real rinv = M::rsqrt(r2);
                    const real r = r2 * rinv;
                    const real r2inv = rinv * rinv;
                    const int tp = tyi + mcur.x;
                    const R4 lj = s_lj[tp];
                    // work count: real (non-padding) pairs inside the cutoff
                    if (kCount && mcur.y >= 0 && mi.y >= 0) n_pairs += 1.0;
                    if (r2 <= a.rc2_lj) {
                        const real r6 = r2inv * r2inv * r2inv;
                        const real ar6 = lj.x * r6;
                        const real u = (ar6 - lj.y) * r6;  // A/r^12 - B/r^6
                        const real g = (static_cast<real>(12) * ar6 - static_cast<real>(6) * lj.y) * r6 * r2inv;  // -U'/r
                        real el, fl;
                        if (kSwitch) {
                            real xs = (r - a.rs) * a.inv_sw;
                            xs = xs < 0 ? static_cast<real>(0) : xs;
                            const real x2 = xs * xs;
                            const real sw = 1 + x2 * xs * (static_cast<real>(-10) + xs * (static_cast<real>(15) - 6 * xs));
                            const real dsw = x2 * (static_cast<real>(-30) + xs * (static_cast<real>(60) - 30 * xs)) * a.inv_sw;
                            el = u * sw;
                            fl = sw * g - u * dsw * rinv;
                        } else {
                            el = u - lj.z + (r - a.rc_lj) * (lj.w * a.rc_lj);
                            fl = g - lj.w * a.rc_lj * rinv;
                        }
                        if (!kAmber && !s_listed[tp]) { el = 0; fl = 0; }
                        s_vdw += el;
                        f += fl;
                    }
                    if (r2 <= a.rc2_ee) {
                        const real qq = qi * pcur.w;
                        real ec, ex;
                        M::erfc_exp(a.alpha * r, ec, ex);
                        const real e_c = qq * rinv;
                        if (kAmber) {
                            s_ew += e_c * ec;
                            f += qq * (ec * rinv + tsp * ex) * r2inv;
                        } else {
                            const real lst = s_listed[tp] ? static_cast<real>(1) : static_cast<real>(0);
                            s_all += e_c;
                            s_list += lst * e_c;
                            s_rf += lst * qq * (rinv - a.krf_crf);
                            s_ew += lst * e_c * ec;
                            f += lst * e_c * r2inv;
                        }
                    }
                }
                const real fx = f * dx, fy = f * dy, fz = f * dz;
                if constexpr (kCount) {
                    w_s[0] += fx * dx;
                    w_s[1] += fy * dy;
                    w_s[2] += fz * dz;
                    w_s[3] += fx * dy;
                    w_s[4] += fx * dz;
                    w_s[5] += fy * dz;
                }
                fix += fx;
                fiy += fy;
                fiz += fz;
                // j-forces: every tile issues its atomic (an empty tile adds
                // nothing and skips the sums and the conversion), so the waits
                // for the prefetched clusters stay short
                Payload<real> q = 0;
                unsigned int off = kOffOOB;
                if (__ballot(act) != 0ull) {
                    const real gj = jsum3(fx, fy, fz, js);
                    const int cjt = __builtin_amdgcn_readlane(e_j, t);
                    q = force_payload<real, kAmber>(a, ti < 3 && mcur.y >= 0, 3 * (cjt * kCluster + tj) + ti, -gj,
                                                    off);
                }
                buf_add(rsrc, off, q);
                if (++flush == 32) {
                    flush = 0;
                    e_vdw += s_vdw; e_all += s_all; e_list += s_list; e_rf += s_rf; e_ew += s_ew;
                    s_vdw = s_all = s_list = s_rf = s_ew = 0;
                    if constexpr (kCount) {
                        for (int q = 0; q < 6; ++q) {
                            w_d[q] += static_cast<double>(w_s[q]);
                            w_s[q] = 0;
                        }
                    }
                }
            };

            // Two register sets, each fetched two tiles before its use: the
            // wait for a set then covers operations issued a tile earlier.
            R4 pA, pB;
            int2 mA, mB;
            fetch(0, pA, mA);
            fetch(1, pB, mB);
            // complete before the loop (see k_nonbonded_c)
            __builtin_amdgcn_s_waitcnt(kWaitVmcnt0);
            for (int t = 0; t < nb; t += 2) {
                tile(pA, mA, t, true);
                fetch(t + 2, pA, mA);
                tile(pB, mB, t + 1, t + 1 < nb);
                fetch(t + 3, pB, mB);
            }
        }
        // i-forces of this item: sum over the 8 lanes sharing i (8 apart)
        fix += shfl_xor(fix, 8);  fiy += shfl_xor(fiy, 8);  fiz += shfl_xor(fiz, 8);
        fix += shfl_xor(fix, 16); fiy += shfl_xor(fiy, 16); fiz += shfl_xor(fiz, 16);
        fix += shfl_xor(fix, 32); fiy += shfl_xor(fiy, 32); fiz += shfl_xor(fiz, 32);
        add_force<real, kAmber>(a, rsrc, tj < 3 && k1 > k0 && mi.y >= 0, 3 * (ci * kCluster + ti) + tj,
                                tj == 0 ? fix : (tj == 1 ? fiy : fiz));
    }
    e_vdw += s_vdw; e_all += s_all; e_list += s_list; e_rf += s_rf; e_ew += s_ew;
    // row by logical block: the 16 rows sharing a cache line of a slot come
    // from blocks of one XCD
    if constexpr (kCount) {
        double v[12] = {e_vdw, e_all, e_list, e_rf, e_ew, n_pairs};
        for (int q = 0; q < 6; ++q) v[6 + q] = w_d[q] + static_cast<double>(w_s[q]);
        const int slots[12] = {E_VDW, E_EE_ALL, E_EE_LIST, E_EE_RF, E_EW_DIRECT, E_PAIRS,
                               E_VSR_XX, E_VSR_YY, E_VSR_ZZ, E_VSR_XY, E_VSR_XZ, E_VSR_YZ};
        block_sum_store<12>(v, a.partial, a.ld, a.row0 + blk, slots);
    } else {
        double v[6] = {e_vdw, e_all, e_list, e_rf, e_ew, n_pairs};
        const int slots[6] = {E_VDW, E_EE_ALL, E_EE_LIST, E_EE_RF, E_EW_DIRECT, E_PAIRS};
        block_sum_store<6>(v, a.partial, a.ld, a.row0 + blk, slots);
    }
}

// The FP32 AMBER pair kernel with in-range pairs queued.  38 % of a tile's 64
// pairs lie within the cutoff on average, and k_nonbonded runs the whole
// LJ + switch + erfc path for a tile once any of its pairs does.  Here a wave
// takes two tiles per step: every lane forms the separation of its pair in
// both tiles (as in k_nonbonded), and the in-range pairs are appended to a
// per-wave ring in LDS (separation, q_i q_j, type pair).  The pair terms run
// over 64 queued pairs at a time, whenever 64 are queued, so nearly every
// lane of a pass does useful work; the pass writes each pair's force back
// over its separation.  A step's tiles are finished (i-forces, j-force sums
// and the j-force atomic) in the next step, whose passes first take what is
// left of them.  So the ring never holds more than the pairs of two steps
// (256), and the finish order is the tile order.  Each
// pair's force is k_nonbonded's: the same separation, r^2 formed by the same
// fused operations, the same pair arithmetic.  Energies are summed by the lane
// that runs the pair; the statistics instance (kCount: pair count and virial)
// forms them in the same pass, so its energies and forces are bit-identical.
//
// A wave's share of an i-cluster's list is every split-th entry (part, part
// + split, ...): the list runs from the near columns to the far ones, so
// contiguous quarters differed by 2.4x in pairs and the waves of a block
// waited for the densest.
constexpr int kRing = 256;  // queued pairs per wave (a power of two)
// LDS of one block of the queued kernel at most (else k_nonbonded runs:
// about 40 LJ types)
constexpr size_t kPackMaxLds = 64 * 1024;

// LDS bytes of a block of the queued kernel for ntp LJ types (padding type
// included): the pair table, then per wave the ring of (separation, q_i q_j)
// and its type-pair indices.
// The switched form reads only (A, B) of each entry, so its table is staged
// as float2 (three blocks of eight waves then fit one CU's LDS).
constexpr size_t compact_lds(int ntp, bool sw) {
    return (sw ? sizeof(float2) : sizeof(float4)) * static_cast<size_t>(ntp) * ntp +
           kWavesOf<float> * kRing * (sizeof(float4) + sizeof(unsigned short));
}

// Orders this wave's LDS stores before its later LDS loads by other lanes.
__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// r^2 of a separation, in one fixed order of fused operations (the tile test
// and the pair pass form it alike).
__device__ __forceinline__ float sep2(float dx, float dy, float dz) {
    return __builtin_fmaf(dz, dz, __builtin_fmaf(dx, dx, dy * dy));
}

// Register budget of eight waves per SIMD (64 VGPRs): the pair loop fits
// without spills (the FP64 bonded and correction blocks of the same launch
// spill a few values; they are a small share of it).  LDS still holds a CU
// to three blocks of eight waves for 30 LJ types; the tighter allocation
// measured as fast as six waves' (80 VGPRs) at 29k atoms and faster at 100k.
// Seven waves' (72 VGPRs) was slower.
template <bool kSwitch, bool kCount>
__global__ __launch_bounds__(kWavesOf<float> * kWave) __attribute__((amdgpu_waves_per_eu(kCount ? 1 : 8))) void
k_nonbonded_c(const NbArgs<float> a, const BArgs<float> ba, const CArgs ca, int npb, int nbb) {
    constexpr int kWaves = kWavesOf<float>;
    if (static_cast<int>(blockIdx.x) >= npb) {
        const int b = static_cast<int>(blockIdx.x) - npb;
        const int k = b * static_cast<int>(blockDim.x) + static_cast<int>(threadIdx.x);
        if (b < nbb) bonded_terms(ba, k, ba.row0 + b);
        else correction_pairs(ca, k - nbb * static_cast<int>(blockDim.x), ca.row0 + (b - nbb));
        return;
    }
    extern __shared__ float4 smem4[];
    const int ntp2 = a.ntp * a.ntp;
    // the pair table: (A, B) for the switched form, (A, B, U(rc), G(rc)) else
    float2* s_lj2 = reinterpret_cast<float2*>(smem4);
    float4* s_lj4 = smem4;
    for (int k = threadIdx.x; k < ntp2; k += blockDim.x) {
        const float4 v = a.lj[k];
        if (kSwitch) s_lj2[k] = make_float2(v.x, v.y);
        else s_lj4[k] = v;
    }
    const int wv = threadIdx.x >> 6;
    // this wave's ring: (dx, dy, dz, q_i q_j) per queued pair, overwritten by
    // (fx, fy, fz, -) in its pass; type-pair indices beside it
    float4* ring0 = smem4 + (kSwitch ? (ntp2 + 1) / 2 : ntp2);
    float4* s_g = ring0 + wv * kRing;
    unsigned short* s_t = reinterpret_cast<unsigned short*>(ring0 + kWaves * kRing) + wv * kRing;
    __syncthreads();

    const int lane = threadIdx.x & 63;
    const int ti = lane & 7, tj = lane >> 3;
    const int ncl = min(*a.ncl_dev, a.max_clusters);
    const JSum js = jsum_lanes(lane);
    double e_vdw = 0, e_ew = 0, n_pairs = 0;
    float s_vdw = 0, s_ew = 0;
    double w_d[6] = {0, 0, 0, 0, 0, 0};
    float w_s[6] = {0, 0, 0, 0, 0, 0};
    const float tsp = 1.1283791670955126f * a.alpha;
    const float rc2 = a.rc2_lj > a.rc2_ee ? a.rc2_lj : a.rc2_ee;
    int flush = 0;
    const BufRsrc rsrc = buf_rsrc(a.fx.f, 8u * a.n3);
    // j-cluster data through buffer loads: the cluster's offset is a scalar,
    // the lane's slot a constant vector offset
    const unsigned int n_slots = static_cast<unsigned int>(a.n3 / 3);
    const BufRsrc pos_rs = buf_rsrc(const_cast<float4*>(a.pos), 16u * n_slots);
    const BufRsrc meta_rs = buf_rsrc(const_cast<int2*>(a.meta), 8u * n_slots);
    const unsigned int pos_off = 16u * tj, meta_off = 8u * tj;
    // byte offset of this lane's j-force component within its j-cluster's
    // 24 accumulators (lanes ti = 0, 1, 2 add x, y, z); the other lanes add
    // nothing
    const unsigned int joff = ti < 3 ? 8u * static_cast<unsigned int>(3 * tj + ti) : kOffOOB;
    unsigned long long bad = 0;  // lanes whose j-force did not fit the fixed point
    const int blk = xcd_block(blockIdx.x, npb);
    const int item = __builtin_amdgcn_readfirstlane(blk * kWaves + wv);
    if (item < ncl * a.split) {
        const int ci = item / a.split, part = item % a.split;
        const int cnt = a.list_count[ci];
        const int nmine = cnt > part ? (cnt - part + a.split - 1) / a.split : 0;  // entries part + split k
        const float4 pi = a.pos[ci * kCluster + ti];
        const int2 mi = a.meta[ci * kCluster + ti];
        const int tyi = mi.x * a.ntp;
        const float qi = pi.w * a.kc;
        float fix = 0, fiy = 0, fiz = 0;
        const size_t base = static_cast<size_t>(ci) * a.cap + part;

        // the pair terms of queued pair p (ring index); its force goes back
        // over its separation
        auto pair_at = [&](int p) {
            const float4 g = s_g[p];
            const int tp = s_t[p];
            const float r2 = sep2(g.x, g.y, g.z), qq = g.w;
            const float rinv = __builtin_amdgcn_rsqf(r2);
            const float r = r2 * rinv;
            const float r2inv = rinv * rinv;
            float4 lj;
            if (kSwitch) {
                const float2 ab = s_lj2[tp];
                lj = make_float4(ab.x, ab.y, 0.0f, 0.0f);
            } else {
                lj = s_lj4[tp];
            }
            float f = 0;
            if (r2 <= a.rc2_lj) {
                const float r6 = r2inv * r2inv * r2inv;
                const float ar6 = lj.x * r6;
                const float u = (ar6 - lj.y) * r6;
                const float gg = (12.0f * ar6 - 6.0f * lj.y) * r6 * r2inv;
                float el, fl;
                if (kSwitch) {
                    float xs = (r - a.rs) * a.inv_sw;
                    xs = xs < 0 ? 0.0f : xs;
                    const float x2 = xs * xs;
                    const float sw = 1 + x2 * xs * (-10.0f + xs * (15.0f - 6 * xs));
                    const float dsw = x2 * (-30.0f + xs * (60.0f - 30 * xs)) * a.inv_sw;
                    el = u * sw;
                    fl = sw * gg - u * dsw * rinv;
                } else {
                    el = u - lj.z + (r - a.rc_lj) * (lj.w * a.rc_lj);
                    fl = gg - lj.w * a.rc_lj * rinv;
                }
                s_vdw += el;
                f += fl;
            }
            if (r2 <= a.rc2_ee) {
                float ec, ex;
                Math<float>::erfc_exp(a.alpha * r, ec, ex);
                const float e_c = qq * rinv;
                s_ew += e_c * ec;
                f += qq * (ec * rinv + tsp * ex) * r2inv;
            }
            const float fx = f * g.x, fy = f * g.y, fz = f * g.z;
            if constexpr (kCount) {
                n_pairs += 1.0;
                w_s[0] += fx * g.x;
                w_s[1] += fy * g.y;
                w_s[2] += fz * g.z;
                w_s[3] += fx * g.y;
                w_s[4] += fx * g.z;
                w_s[5] += fy * g.z;
            }
            s_g[p] = make_float4(fx, fy, fz, qq);
        };
        // queue state: pairs [head, tail) wait for a pass (ring index = count
        // mod kRing); counters are uniform over the wave
        int head = 0, tail = 0;
        // one pass over the next n <= 64 queued pairs
        auto pass = [&](int n) {
            if (lane < n) pair_at((head + lane) & (kRing - 1));
            head += n;
        };
        // a tile waiting for its pairs' forces: lanes with a queued pair, the
        // pair's ring index, its j-cluster
        struct Wait {
            unsigned long long act;
            int pos;
            int cj;
        };
        // finish a tile from its queued forces
        auto finish = [&](const Wait& w) {
            float fx = 0, fy = 0, fz = 0;
            if (__builtin_amdgcn_inverse_ballot_w64(w.act)) {
                const float4 g = s_g[w.pos];
                fx = g.x;
                fy = g.y;
                fz = g.z;
            }
            fix += fx;
            fiy += fy;
            fiz += fz;
            long long q = 0;
            unsigned int off = kOffOOB;
            if (w.act != 0ull) {
                // a padding j-slot has no pair in range, so its sum is 0 and
                // it adds nothing
                const float gj = jsum3(fx, fy, fz, js);
                q = fixed32_f32_flag(-gj, bad, kJLanes);
                off = q != 0 ? joff + static_cast<unsigned int>(w.cj) * (3u * kCluster * 8u) : kOffOOB;
            }
            buf_add(rsrc, off, q);
        };
        // the previous step's tiles (none before the first step): each step
        // finishes exactly these two, so every step issues two j-force atomics
        // and the compiler's waits for the prefetched clusters count exactly
        // the operations issued after them
        Wait wA{0ull, 0, 0}, wB{0ull, 0, 0};
        int prev_end = 0;  // end of their queued pairs

        for (int kb = 0; kb < nmine; kb += kWave) {
            const int nb = min(kWave, nmine - kb);
            int e_j = 0;
            unsigned long long e_m = 0;
            float4 e_d = {0, 0, 0, 0};
            if (lane < nb) {
                const size_t e = base + static_cast<size_t>(kb + lane) * a.split;
                e_j = a.list_j[e];
                e_m = a.list_mask[e];
                e_d = a.list_delta[e];
            }
            const int e_mlo = static_cast<int>(e_m & 0xffffffffull), e_mhi = static_cast<int>(e_m >> 32);
            // j-cluster data of tile t (clamped to the batch, so every fetch
            // is issued and reads a valid cluster)
            auto fetch = [&](int t, float4& p, int2& m) {
                const int cj = __builtin_amdgcn_readlane(e_j, min(t, nb - 1));
                p = buf_load_f4(pos_rs, pos_off, cj * static_cast<int>(16 * kCluster));
                m = buf_load_i2(meta_rs, meta_off, cj * static_cast<int>(8 * kCluster));
            };
            // tile t: the separation of this lane's pair, queued if the pair is
            // unmasked and within the cutoff (the tile's mask is a lane mask,
            // so the in-range test combines with it in scalar registers)
            auto queue = [&](const float4& pc, const int2& mc, int t, bool live, int rank0, Wait& w) {
                unsigned long long mask =
                    (static_cast<unsigned long long>(static_cast<unsigned int>(__builtin_amdgcn_readlane(e_mhi, t))) << 32) |
                    static_cast<unsigned int>(__builtin_amdgcn_readlane(e_mlo, t));
                mask = live ? mask : 0ull;
                const float dx = pi.x - pc.x + bcast<float>(e_d.x, t);
                const float dy = pi.y - pc.y + bcast<float>(e_d.y, t);
                const float dz = pi.z - pc.z + bcast<float>(e_d.z, t);
                const float r2 = sep2(dx, dy, dz);
                w.act = __builtin_amdgcn_ballot_w64(r2 <= rc2) & __builtin_amdgcn_ballot_w64(r2 >= 1e-12f) & mask;
                w.cj = __builtin_amdgcn_readlane(e_j, t);
                const int rank = rank0 + __builtin_amdgcn_mbcnt_hi(static_cast<unsigned int>(w.act >> 32),
                                                                  __builtin_amdgcn_mbcnt_lo(static_cast<unsigned int>(w.act), 0u));
                w.pos = (tail + rank) & (kRing - 1);
                if (__builtin_amdgcn_inverse_ballot_w64(w.act)) {
                    s_g[w.pos] = make_float4(dx, dy, dz, qi * pc.w);
                    s_t[w.pos] = static_cast<unsigned short>(tyi + mc.x);
                }
                return __popcll(w.act);
            };

            float4 pA, pB;
            int2 mA, mB;
            fetch(0, pA, mA);
            fetch(1, pB, mB);
            // The first two tiles' data, complete before the loop: the
            // compiler places its waits for the loop's registers from the
            // state at the loop header, and a load still in flight there
            // would put a wait for every vector memory operation but the
            // newest four (the previous step's atomics included) into every
            // step.
            __builtin_amdgcn_s_waitcnt(kWaitVmcnt0);
            for (int t = 0; t < nb; t += 2) {
                Wait cA, cB;
                const int nA = queue(pA, mA, t, true, 0, cA);
                fetch(t + 2, pA, mA);
                const int nB = queue(pB, mB, t + 1, t + 1 < nb, nA, cB);
                fetch(t + 3, pB, mB);
                tail += nA + nB;
                wave_sync();
                // passes: every full 64, and first whatever the previous
                // step's tiles still lack
                while (tail - head >= kWave || head < prev_end) pass(min(kWave, tail - head));
                wave_sync();
                finish(wA);
                finish(wB);
                wA = cA;
                wB = cB;
                prev_end = tail;
                flush += 2;
                if (flush >= 32) {
                    flush = 0;
                    e_vdw += s_vdw;
                    e_ew += s_ew;
                    s_vdw = s_ew = 0;
                    if constexpr (kCount) {
                        for (int q = 0; q < 6; ++q) {
                            w_d[q] += static_cast<double>(w_s[q]);
                            w_s[q] = 0;
                        }
                    }
                }
            }
        }
        // the pairs still queued, then the last step's tiles
        if (head < tail) {
            pass(tail - head);
            wave_sync();
        }
        finish(wA);
        finish(wB);
        fix += shfl_xor(fix, 8);  fiy += shfl_xor(fiy, 8);  fiz += shfl_xor(fiz, 8);
        fix += shfl_xor(fix, 16); fiy += shfl_xor(fiy, 16); fiz += shfl_xor(fiz, 16);
        fix += shfl_xor(fix, 32); fiy += shfl_xor(fiy, 32); fiz += shfl_xor(fiz, 32);
        add_force<float, true>(a, rsrc, tj < 3 && nmine > 0 && mi.y >= 0, 3 * (ci * kCluster + ti) + tj,
                               tj == 0 ? fix : (tj == 1 ? fiy : fiz));
    }
    if (bad != 0ull && lane == 0) atomicOr(a.fx.overflow, 1u);
    e_vdw += s_vdw;
    e_ew += s_ew;
    if constexpr (kCount) {
        double v[12] = {e_vdw, 0.0, 0.0, 0.0, e_ew, n_pairs};
        for (int q = 0; q < 6; ++q) v[6 + q] = w_d[q] + static_cast<double>(w_s[q]);
        const int slots[12] = {E_VDW, E_EE_ALL, E_EE_LIST, E_EE_RF, E_EW_DIRECT, E_PAIRS,
                               E_VSR_XX, E_VSR_YY, E_VSR_ZZ, E_VSR_XY, E_VSR_XZ, E_VSR_YZ};
        block_sum_store<12>(v, a.partial, a.ld, a.row0 + blk, slots);
    } else {
        double v[6] = {e_vdw, 0.0, 0.0, 0.0, e_ew, n_pairs};
        const int slots[6] = {E_VDW, E_EE_ALL, E_EE_LIST, E_EE_RF, E_EW_DIRECT, E_PAIRS};
        block_sum_store<6>(v, a.partial, a.ld, a.row0 + blk, slots);
    }
}

template <typename real>
hipError_t run(mdgpu_ctx* c, bool amber, int row0, const double* coords) {
    Work& w = c->work;
    const Topo& t = c->topo;
    const mdgpu_params& p = c->params;
    NbArgs<real> a{};
    const int ncol = c->grid.ncx * c->grid.ncy;
    a.ncl_dev = w.col_cstart + ncol;
    a.max_clusters = c->n_clusters;
    a.split = c->split;
    a.cap = w.list_cap;
    // Verlet mode: the kept list pruned to the cutoff for this evaluation
    const bool pruned = c->skin > 0;
    a.list_count = pruned ? w.inner_count : w.list_count;
    a.list_j = pruned ? w.inner_j : w.list_j;
    a.list_mask = pruned ? w.inner_mask : w.list_mask;
    a.ntp = t.n_types + 1;
    if constexpr (sizeof(real) == 4) {
        a.list_delta = pruned ? w.inner_delta_f : w.list_delta_f;
        a.pos = w.slot_pos_f;
        a.lj = t.lj_f;
    } else {
        a.list_delta = pruned ? w.inner_delta_d : w.list_delta_d;
        a.pos = w.slot_pos_d;
        a.lj = t.lj_d;
    }
    a.listed = reinterpret_cast<const unsigned char*>(t.lj_listed);
    a.meta = w.slot_meta;
    a.rc2_lj = static_cast<real>(p.lj_cutoff * p.lj_cutoff);
    a.rc2_ee = static_cast<real>(p.ee_cutoff * p.ee_cutoff);
    a.rc_lj = static_cast<real>(p.lj_cutoff);
    a.rc_ee = static_cast<real>(p.ee_cutoff);
    const bool sw = p.lj_switch > 0.0;
    a.rs = static_cast<real>(sw ? p.lj_switch : 0.0);
    a.inv_sw = static_cast<real>(sw ? 1.0 / (p.lj_cutoff - p.lj_switch) : 0.0);
    a.alpha = static_cast<real>(p.ewald_alpha);
    a.kc = static_cast<real>(kCoulomb);
    a.krf_crf = static_cast<real>(1.0 / p.ee_cutoff);
    a.force_d = w.force_d_s;
    a.fx = FixedAcc{w.force_fx_s, fixed_scale(p.semantics), w.flags};
    a.partial = w.partial;
    a.ld = static_cast<int>(w.cap_partial);
    a.row0 = row0;
    a.n3 = 3 * c->n_clusters * kCluster;
    const int npb = nonbonded_rows(c);
    // bonded terms and corrections: rows after the pair blocks' rows
    const int nbb = bonded_rows(c);
    const int ncb = amber ? corrections_rows(c) : 0;
    const FixedAcc fxa{w.force_fx, fixed_scale(p.semantics), w.flags};
    // statistics runs (pair counting, virial) use the counting instance of the pair kernel
    const bool stats = c->counting || c->virial;
    const BArgs<real> ba{coords, t, w.force_d, fxa, w.partial, static_cast<int>(w.cap_partial), row0 + npb, amber ? 1 : 0,
                         stats ? 1 : 0};
    CArgs ca{};
    ca.x = coords;
    ca.lx = c->grid.box[0];
    ca.ly = c->grid.box[1];
    ca.lz = c->grid.box[2];
    ca.cell = c->grid.cell;
    ca.n_excl = t.n_excl_pairs;
    ca.excl = t.excl_pairs;
    ca.n14 = t.n_pairs14;
    ca.p14 = t.pairs14;
    ca.scee = t.scee14;
    ca.charge = t.charge;
    ca.type = t.type;
    ca.nb_index = t.nb_index;
    ca.lj14 = t.lj14;
    ca.ntypes = t.n_types;
    ca.alpha = p.ewald_alpha;
    ca.mixed = sizeof(real) == 4 ? 1 : 0;
    ca.force_d = w.force_d;
    ca.fx = fxa;
    ca.partial = w.partial;
    ca.ld = static_cast<int>(w.cap_partial);
    ca.row0 = row0 + npb + nbb;
    ca.virial = stats ? 1 : 0;
    const size_t shm = (sizeof(typename V4<real>::T) + 1) * a.ntp * a.ntp;
    dim3 grid(npb + nbb + ncb), block(kWavesOf<real> * kWave);
    if constexpr (sizeof(real) == 4) {
        if (amber) {
            // the compacted kernel: LJ table, then one pack buffer per wave
            const size_t shc = compact_lds(a.ntp, sw);
            if (shc <= kPackMaxLds) {
                if (stats) {
                    if (sw) hipLaunchKernelGGL((k_nonbonded_c<true, true>), grid, block, shc, c->stream, a, ba, ca, npb, nbb);
                    else hipLaunchKernelGGL((k_nonbonded_c<false, true>), grid, block, shc, c->stream, a, ba, ca, npb, nbb);
                } else {
                    if (sw) hipLaunchKernelGGL((k_nonbonded_c<true, false>), grid, block, shc, c->stream, a, ba, ca, npb, nbb);
                    else hipLaunchKernelGGL((k_nonbonded_c<false, false>), grid, block, shc, c->stream, a, ba, ca, npb, nbb);
                }
                return hipGetLastError();
            }
        }
    }
#define MDG_NB_LAUNCH(AM, SW)                                                                                  \
    do {                                                                                                      \
        if (stats)                                                                                            \
            hipLaunchKernelGGL((k_nonbonded<real, AM, SW, true>), grid, block, shm, c->stream, a, ba, ca, npb, \
                               nbb);                                                                          \
        else                                                                                                  \
            hipLaunchKernelGGL((k_nonbonded<real, AM, SW, false>), grid, block, shm, c->stream, a, ba, ca, npb, \
                               nbb);                                                                          \
    } while (0)
    if (amber) {
        if (sw) MDG_NB_LAUNCH(true, true);
        else MDG_NB_LAUNCH(true, false);
    } else {
        if (sw) MDG_NB_LAUNCH(false, true);
        else MDG_NB_LAUNCH(false, false);
    }
#undef MDG_NB_LAUNCH
    return hipGetLastError();
}

}  // namespace

template <typename real, bool kAmber, bool kSwitch>
const void* pair_fn(bool counting) {
    return counting ? reinterpret_cast<const void*>(&k_nonbonded<real, kAmber, kSwitch, true>)
                    : reinterpret_cast<const void*>(&k_nonbonded<real, kAmber, kSwitch, false>);
}

template <typename real>
const void* pair_fn(bool amber, bool sw, bool counting) {
    if (amber) return sw ? pair_fn<real, true, true>(counting) : pair_fn<real, true, false>(counting);
    return sw ? pair_fn<real, false, true>(counting) : pair_fn<real, false, false>(counting);
}

const void* pair_kernel_fn(const mdgpu_ctx* c) {
    const bool amber = c->params.semantics == MDGPU_SEM_AMBER, sw = c->params.lj_switch > 0.0;
    const bool stats = c->counting || c->virial;
    const bool mixed = c->params.precision == MDGPU_PREC_MIXED;
    if (mixed && amber && compact_lds(c->topo.n_types + 1, sw) <= kPackMaxLds) {
        if (stats)
            return sw ? reinterpret_cast<const void*>(&k_nonbonded_c<true, true>)
                      : reinterpret_cast<const void*>(&k_nonbonded_c<false, true>);
        return sw ? reinterpret_cast<const void*>(&k_nonbonded_c<true, false>)
                  : reinterpret_cast<const void*>(&k_nonbonded_c<false, false>);
    }
    return mixed ? pair_fn<float>(amber, sw, stats) : pair_fn<double>(amber, sw, stats);
}

int pair_block_threads(const mdgpu_ctx* c) {
    return kWave * (c->params.precision == MDGPU_PREC_MIXED ? kWavesOf<float> : kWavesOf<double>);
}

int nonbonded_rows(const mdgpu_ctx* c) {
    const int w = pair_block_threads(c) / kWave;
    return (c->n_clusters * c->split + w - 1) / w;
}

hipError_t launch_nonbonded(mdgpu_ctx* c, bool mixed, bool amber, int row0, const double* coords) {
    return mixed ? run<float>(c, amber, row0, coords) : run<double>(c, amber, row0, coords);
}

}  // namespace mdg
