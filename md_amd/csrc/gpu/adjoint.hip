// Adjoint kernels: derivatives of the AMBER energy terms with respect to the
// force-field parameters at one configuration.
//
// The reference computes no parameter derivatives.  Every quantity here is
// the derivative of an energy the evaluation kernels form, term by term:
//   LJ (switched or force-shifted, NonBondedEnergy.cpp:79-131) is linear in
//     the table entries A_p, B_p, so dE/dA_p and dE/dB_p are the pair
//     energies at (A, B) = (1, 0) and (0, 1) summed over the pairs of entry p;
//   Ewald direct k q_i q_j erfc(a r)/r, the exclusion term -k q_i q_j erf(a r)/r
//     and the 1-4 Coulomb k q_i q_j (1/(scee r) - erf(a r)/r) give dE/dq_i as
//     a potential sum_j c_ij q_j;
//   bonded terms give dE/dk and dE/d(equilibrium) per parameter type;
//   CMAP gives dE/da per patch coefficient (u^i v^j of the cell).
// The pair list and the pair values are recomputed, nothing is stored by an
// evaluation.  Sums go to 64-bit fixed-point accumulators (integer adds are
// associative: the result does not depend on the order of the atomics),
// except the LJ sums of the FP64 form, which use FP64 atomics: one pair of
// atoms without LJ parameters can sit a fraction of an angstrom apart, and
// its 1/r^12 then exceeds every fixed-point range fine enough for the FP64
// bar (the mixed form uses 2^32).
#include "ctx.h"
#include "erfc_f32.h"

namespace mdg {

namespace {

template <typename real>
struct V4;
template <>
struct V4<float> { using T = float4; };
template <>
struct V4<double> { using T = double4; };

constexpr int kWaves = 4;
constexpr int kBlock = 256;

__device__ __forceinline__ void fx_add(unsigned long long* acc, double v, double scale, unsigned int* ovf) {
    const long long q = to_fixed(v, scale, ovf);
    if (q != 0) atomicAdd(acc, static_cast<unsigned long long>(q));
}

// erfc of the mixed form: the pair kernel's relative-accuracy fit (erfc_f32.h)
__device__ __forceinline__ float erfc_r(float x) { return erfc_f32(x, __expf(-x * x)); }
__device__ __forceinline__ double erfc_r(double x) { return erfc(x); }
__device__ __forceinline__ float rsqrt_r(float x) { return __builtin_amdgcn_rsqf(x); }
__device__ __forceinline__ double rsqrt_r(double x) { return 1.0 / sqrt(x); }

// Sum over the 8 lanes l, l^1, ..., of a group (i = lane % 8 fixed j).
template <typename T>
__device__ __forceinline__ T sum_over_i(T v) {
    v += __shfl_xor(v, 1, kWave);
    v += __shfl_xor(v, 2, kWave);
    v += __shfl_xor(v, 4, kWave);
    return v;
}

template <typename real>
struct PairAdjArgs {
    const int* ncl_dev;
    int max_clusters, split, cap;
    const int* list_count;
    const int* list_j;
    const unsigned long long* list_mask;
    const typename V4<real>::T* list_delta;
    const typename V4<real>::T* pos;   // cluster-relative position + charge
    const int2* meta;                  // (LJ type, atom or -1)
    const int* pidx;                   // (T+1)^2 LJ table entry or -1
    int ntp, n_lj;
    real rc2_lj, rc2_ee, rc_lj, rs, inv_sw;
    real c12, g12, c6, g6;             // force-shifted LJ: rc^-12, 12 rc^-13, rc^-6, 6 rc^-7
    real alpha, kc;
    unsigned long long* q_dir;         // [N] sum_j k erfc(a r)/r q_j
    unsigned long long* lj_a;          // [n_lj] fixed point (mixed form)
    unsigned long long* lj_b;
    double* lj_ad;                     // [n_lj] FP64 (FP64 form)
    double* lj_bd;
    unsigned int* ovf;
};

// LJ derivative bins: FP64 in the FP64 form; in the mixed form two 64-bit
// fixed-point words per entry, deterministic like every other sum of the
// mixed adjoint.  A contribution v is split into hi = floor(v 2^16) and the
// remainder lo = round((v 2^16 - hi) 2^32), so the entry is
// sum hi / 2^16 + sum lo / 2^48: the resolution of a 2^48 fixed point with
// the range of a 2^16 one (|v| < 1.4e14, r > 0.066 A for 1/r^12).  The sums of
// the A and B entries cancel in the derivatives with respect to sigma and
// epsilon, so a single 2^32 word (the resolution of the other sums) would
// quantise the nearest pairs' 1/r^12 to ~5e-5 each.
template <typename real>
struct LjBin;
template <>
struct LjBin<double> {
    using T = double;
    static constexpr int kWords = 1;
    static __device__ __forceinline__ void add(double* s, double v, unsigned int*) { atomicAdd(s, v); }
    static __device__ __forceinline__ void flush(const double* v, double* gd, unsigned long long*) {
        if (*v != 0.0) atomicAdd(gd, *v);
    }
};
template <>
struct LjBin<float> {
    using T = unsigned long long;
    static constexpr int kWords = 2;
    static __device__ __forceinline__ void add(unsigned long long* s, double v, unsigned int* ovf) {
        const double x = v * kAdjScaleLJhi;
        if (!(fabs(x) < 9.0e18)) {  // also catches NaN
            atomicOr(ovf, 1u);
            return;
        }
        const double h = floor(x);
        const long long hi = static_cast<long long>(h);
        const unsigned long long lo = static_cast<unsigned long long>(__double2ll_rn((x - h) * 4294967296.0));
        if (hi) atomicAdd(s, static_cast<unsigned long long>(hi));
        if (lo) atomicAdd(s + 1, lo);
    }
    static __device__ __forceinline__ void flush(const unsigned long long* v, double*, unsigned long long* gf) {
        if (v[0]) atomicAdd(gf, v[0]);
        if (v[1]) atomicAdd(gf + 1, v[1]);
    }
};

// 1/r^2 from r^2 in FP32: the hardware reciprocal and one Newton step, so
// that 1/r^6 and 1/r^12 carry unbiased rounding only (their sums over the
// nearest pairs set the sigma and epsilon derivatives)
__device__ __forceinline__ float inv_r2(float r2) {
    const float y = __builtin_amdgcn_rcpf(r2);
    return fmaf(y, fmaf(-r2, y, 1.0f), y);
}
__device__ __forceinline__ double inv_r2(double r2) { return 1.0 / r2; }

// One wave per (i-cluster, list part) as in the pair kernel, lane = 8 tj + ti.
// LJ derivative sums go to per-block LDS bins, flushed once per block.
template <typename real, bool kSwitch>
__global__ __launch_bounds__(kWaves * kWave) void k_pair_adjoint(const PairAdjArgs<real> a) {
    using R4 = typename V4<real>::T;
    using Bin = LjBin<real>;
    using BT = typename Bin::T;
    extern __shared__ unsigned char smem[];
    constexpr int kW = Bin::kWords;
    BT* s_a = reinterpret_cast<BT*>(smem);
    BT* s_b = s_a + kW * a.n_lj;
    int* s_p = reinterpret_cast<int*>(s_b + kW * a.n_lj);
    const int ntp2 = a.ntp * a.ntp;
    for (int k = threadIdx.x; k < kW * a.n_lj; k += blockDim.x) {
        s_a[k] = 0;
        s_b[k] = 0;
    }
    for (int k = threadIdx.x; k < ntp2; k += blockDim.x) s_p[k] = a.pidx[k];
    __syncthreads();

    const int lane = threadIdx.x & 63, ti = lane & 7, tj = lane >> 3;
    const int ncl = min(*a.ncl_dev, a.max_clusters);
    const int nwork = ncl * a.split;
    const int mbit = mask_bit(ti, tj);
    for (int gw0 = blockIdx.x * kWaves; gw0 < nwork; gw0 += gridDim.x * kWaves) {
        const int gw = __builtin_amdgcn_readfirstlane(gw0 + (threadIdx.x >> 6));
        if (gw >= nwork) continue;  // uniform over the wave; no barrier inside the loop
        const int ci = gw / a.split, part = gw % a.split;
        const int cnt = a.list_count[ci];
        const int k0 = static_cast<int>((static_cast<long long>(cnt) * part) / a.split);
        const int k1 = static_cast<int>((static_cast<long long>(cnt) * (part + 1)) / a.split);
        const R4 pi = a.pos[ci * kCluster + ti];
        const int2 mi = a.meta[ci * kCluster + ti];
        const int tyi = mi.x * a.ntp;
        double phi_i = 0.0;
        const size_t base = static_cast<size_t>(ci) * a.cap;
        for (int k = k0; k < k1; ++k) {
            const int cj = a.list_j[base + k];
            const unsigned long long mask = a.list_mask[base + k];
            const R4 dl = a.list_delta[base + k];
            const R4 pj = a.pos[cj * kCluster + tj];
            const int2 mj = a.meta[cj * kCluster + tj];
            const real dx = pi.x - pj.x + dl.x, dy = pi.y - pj.y + dl.y, dz = pi.z - pj.z + dl.z;
            const real r2 = dx * dx + dy * dy + dz * dz;
            const bool act = ((mask >> mbit) & 1ull) && mi.y >= 0 && mj.y >= 0 && r2 >= static_cast<real>(1e-12);
            double phi_j = 0.0;
            if (act) {
                const real rinv = rsqrt_r(r2), r = r2 * rinv;
                if (r2 <= a.rc2_lj) {
                    const int p = s_p[tyi + mj.x];
                    if (p >= 0) {
                        const real i2 = inv_r2(r2), i6 = i2 * i2 * i2, i12 = i6 * i6;
                        real ea, eb;
                        if (kSwitch) {
                            real xs = (r - a.rs) * a.inv_sw;
                            xs = xs < 0 ? static_cast<real>(0) : xs;
                            const real sw = 1 + xs * xs * xs * (static_cast<real>(-10) + xs * (static_cast<real>(15) - 6 * xs));
                            ea = sw * i12;
                            eb = -sw * i6;
                        } else {
                            ea = i12 - a.c12 + (r - a.rc_lj) * a.g12;
                            eb = -(i6 - a.c6) - (r - a.rc_lj) * a.g6;
                        }
                        Bin::add(s_a + kW * p, static_cast<double>(ea), a.ovf);
                        Bin::add(s_b + kW * p, static_cast<double>(eb), a.ovf);
                    }
                }
                if (r2 <= a.rc2_ee) {
                    const real cc = a.kc * erfc_r(a.alpha * r) * rinv;
                    phi_i += static_cast<double>(cc * pj.w);
                    phi_j = static_cast<double>(cc * pi.w);
                }
            }
            if (__ballot(act) != 0ull) {
                const double sj = sum_over_i(phi_j);
                if (ti == 0 && mj.y >= 0) fx_add(a.q_dir + mj.y, sj, kAdjScaleQ, a.ovf);
            }
        }
        if (mi.y >= 0 && k1 > k0) fx_add(a.q_dir + mi.y, phi_i, kAdjScaleQ, a.ovf);
    }
    __syncthreads();
    for (int k = threadIdx.x; k < a.n_lj; k += blockDim.x) {
        Bin::flush(s_a + kW * k, a.lj_ad + k, a.lj_a + kW * k);
        Bin::flush(s_b + kW * k, a.lj_bd + k, a.lj_b + kW * k);
    }
}

// Adds v to acc[key] for every lane with key >= 0, one atomic per distinct
// key in the wave: the lanes sharing the first pending key are summed with a
// fixed butterfly, then removed.  Every lane of the wave must call it.
__device__ __forceinline__ void wave_keyed_add(unsigned long long* acc, int key, double v, double scale,
                                               unsigned int* ovf) {
    const int lane = threadIdx.x & 63;
    unsigned long long pending = __ballot(key >= 0);
    while (pending) {
        const int leader = __ffsll(static_cast<long long>(pending)) - 1;
        const int kk = __shfl(key, leader, kWave);
        const bool mine = key == kk && ((pending >> lane) & 1ull);
        const double s = wave_sum(mine ? v : 0.0);
        if (lane == leader) fx_add(acc + kk, s, scale, ovf);
        pending &= ~__ballot(mine);
    }
}

template <typename real>
struct R3 {
    real x, y, z;
};
template <typename real>
__device__ __forceinline__ R3<real> diff(const double* __restrict__ x, int a, int b) {
    return {static_cast<real>(x[3 * a] - x[3 * b]), static_cast<real>(x[3 * a + 1] - x[3 * b + 1]),
            static_cast<real>(x[3 * a + 2] - x[3 * b + 2])};
}
template <typename real>
__device__ __forceinline__ R3<real> cross(const R3<real>& a, const R3<real>& b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
template <typename real>
__device__ __forceinline__ real dot(const R3<real>& a, const R3<real>& b) {
    return a.x * b.x + a.y * b.y + a.z * b.z;
}
template <typename real>
__device__ __forceinline__ real dihedral_angle(const R3<real>& b1, const R3<real>& b2, const R3<real>& b3) {
    const R3<real> n1 = cross(b1, b2), n2 = cross(b2, b3);
    const real m1 = dot(n1, n1), m2 = dot(n2, n2), l2 = sqrt(dot(b2, b2));
    if (m1 < static_cast<real>(1e-28) || m2 < static_cast<real>(1e-28) || l2 < static_cast<real>(1e-14))
        return static_cast<real>(NAN);
    return atan2(dot(b1, n2) * l2, dot(n1, n2));
}
template <typename real>
__device__ __forceinline__ real wrap_pi(real d) {
    const real pi = static_cast<real>(kPi), two = static_cast<real>(2 * kPi);
    for (int it = 0; it < 8 && d <= -pi; ++it) d += two;
    for (int it = 0; it < 8 && d > pi; ++it) d -= two;
    return d;
}

struct BondAdjArgs {
    const double* x;
    Topo t;
    unsigned long long* bonded;  // [2 (k, eq)] per type: bonds, UB, angles, dihedrals, impropers
    int off_ub, off_ang, off_dih, off_imp;
    unsigned long long* cmap;    // [sets x res^2 x 16]
    unsigned int* ovf;
};

// One thread per term (bonds, UB, angles, dihedrals, impropers, CMAP, in that
// order), as the bonded kernel walks them.  FP64 in both precision forms: each
// term forms only a few scalars (a length, angle or torsion and its difference
// from the equilibrium), and in FP32 r - r0 or theta - theta0 would keep only
// about 6e-8 r / (r - r0) of relative precision; the per-type sums then cancel
// further.  The bonded terms are a negligible share of the adjoint's work.
__global__ __launch_bounds__(kBlock) void k_bonded_adjoint(const BondAdjArgs a) {
    using real = double;
    const Topo& t = a.t;
    const double* x = a.x;
    const int nb = t.n_bonds, nu = nb + t.n_ubs, na = nu + t.n_angles, nd = na + t.n_dihedrals,
              ni = nd + t.n_impropers, nc = ni + t.n_cmaps;
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    int key = -1;
    double gk = 0.0, ge = 0.0;
    if (k < nu) {
        const bool ub = k >= nb;
        const int q = ub ? k - nb : k;
        const int2 at = ub ? t.ub_atoms[q] : t.bond_atoms[q];
        const int ty = ub ? t.ub_type[q] : t.bond_type[q];
        const double2 par = ub ? t.ub_par[ty] : t.bond_par[ty];
        const R3<real> d = diff<real>(x, at.x, at.y);
        const real dr = sqrt(dot(d, d)) - static_cast<real>(par.y);
        gk = static_cast<double>(dr * dr);
        ge = static_cast<double>(static_cast<real>(-2.0 * par.x) * dr);
        key = 2 * ty + (ub ? a.off_ub : 0);
    } else if (k < na) {
        const int q = k - nu;
        const int4 at = t.angle_atoms[q];
        const int ty = t.angle_type[q];
        const double2 par = t.angle_par[ty];
        const R3<real> ba = diff<real>(x, at.x, at.y), bc = diff<real>(x, at.z, at.y);
        const real ab2 = dot(ba, ba), bc2 = dot(bc, bc);
        real cs = dot(ba, bc) * (static_cast<real>(1) / sqrt(ab2 * bc2));
        cs = cs < -1 ? static_cast<real>(-1) : (cs > 1 ? static_cast<real>(1) : cs);
        const real dth = acos(cs) - static_cast<real>(par.y);
        gk = static_cast<double>(dth * dth);
        ge = static_cast<double>(static_cast<real>(-2.0 * par.x) * dth);
        key = a.off_ang + 2 * ty;
    } else if (k < ni) {
        const bool imp = k >= nd;
        const int q = imp ? k - nd : k - na;
        const int4 at = imp ? t.imp_atoms[q] : t.dih_atoms[q];
        const R3<real> b1 = diff<real>(x, at.y, at.x), b2 = diff<real>(x, at.z, at.y), b3 = diff<real>(x, at.w, at.z);
        const real phi = dihedral_angle(b1, b2, b3);
        if (!imp) {
            const int ty = t.dih_type[q];
            const double4 par = t.dih_par[ty];
            const real arg = static_cast<real>(par.y) * phi - static_cast<real>(par.z);
            gk = static_cast<double>(1 + cos(arg));
            ge = static_cast<double>(static_cast<real>(par.x) * sin(arg));
            key = a.off_dih + 2 * ty;
        } else {
            const int ty = t.imp_type[q];
            const double2 par = t.imp_par[ty];
            const real dw = wrap_pi(phi - static_cast<real>(par.y));
            gk = static_cast<double>(dw * dw);
            ge = static_cast<double>(static_cast<real>(-2.0 * par.x) * dw);
            key = a.off_imp + 2 * ty;
        }
    } else if (k < nc) {
        const int q = k - ni;
        const int* at = t.cmap_atoms + 5 * q;
        const R3<real> b1 = diff<real>(x, at[1], at[0]), b2 = diff<real>(x, at[2], at[1]),
                       b3 = diff<real>(x, at[3], at[2]), b4 = diff<real>(x, at[4], at[3]);
        const real phi = dihedral_angle(b1, b2, b3), psi = dihedral_angle(b2, b3, b4);
        const int res = t.cmap_res;
        const real sp = static_cast<real>(2 * kPi) / static_cast<real>(res);
        const real gp = (wrap_pi(phi) + static_cast<real>(kPi)) / sp, gs = (wrap_pi(psi) + static_cast<real>(kPi)) / sp;
        int row = static_cast<int>(gp), col = static_cast<int>(gs);
        const real u = gp - static_cast<real>(row), v = gs - static_cast<real>(col);
        row %= res;
        col %= res;
        if (row < 0) row += res;
        if (col < 0) col += res;
        if (phi == phi && psi == psi) {
            unsigned long long* g = a.cmap + (static_cast<size_t>(t.cmap_set[q] - 1) * res * res + row * res + col) * 16;
            const real up[4] = {1, u, u * u, u * u * u}, vp[4] = {1, v, v * v, v * v * v};
            for (int i = 0; i < 4; ++i)
                for (int j = 0; j < 4; ++j) fx_add(g + 4 * i + j, static_cast<double>(up[i] * vp[j]), kAdjScaleLJ, a.ovf);
        }
    }
    wave_keyed_add(a.bonded, key, gk, kAdjScaleB, a.ovf);
    wave_keyed_add(a.bonded, key < 0 ? -1 : key + 1, ge, kAdjScaleB, a.ovf);
}

struct CorrAdjArgs {
    const double* x;
    double lx, ly, lz;
    int n_excl;
    const int2* excl;
    int n14;
    const int2* p14;
    const double* scee;
    const double* charge;
    const int* type;
    const int* nb_index;
    int ntypes;
    double alpha;
    unsigned long long* q_excl;
    unsigned long long* q_14;
    unsigned long long* lj14_a;
    unsigned long long* lj14_b;
    unsigned int* ovf;
};

// FP64 in both forms, minimum image, no cutoff (as k_corrections).
__global__ __launch_bounds__(kBlock) void k_corrections_adjoint(const CorrAdjArgs a) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= a.n_excl + a.n14) return;
    const bool is14 = k >= a.n_excl;
    const int2 p = is14 ? a.p14[k - a.n_excl] : a.excl[k];
    double dx = a.x[3 * p.x] - a.x[3 * p.y], dy = a.x[3 * p.x + 1] - a.x[3 * p.y + 1],
           dz = a.x[3 * p.x + 2] - a.x[3 * p.y + 2];
    dx -= a.lx * rint(dx / a.lx);
    dy -= a.ly * rint(dy / a.ly);
    dz -= a.lz * rint(dz / a.lz);
    const double r2 = dx * dx + dy * dy + dz * dz, r = sqrt(r2);
    const double ef = erf(a.alpha * r);
    double c;
    unsigned long long* q;
    if (!is14) {
        c = -kCoulomb * ef / r;
        q = a.q_excl;
    } else {
        c = kCoulomb * (1.0 / (a.scee[k - a.n_excl] * r) - ef / r);
        q = a.q_14;
        const int nb = a.nb_index[a.type[p.x] * a.ntypes + a.type[p.y]];
        if (nb > 0) {
            const double i6 = 1.0 / (r2 * r2 * r2);
            fx_add(a.lj14_a + nb - 1, i6 * i6, kAdjScaleLJ, a.ovf);
            fx_add(a.lj14_b + nb - 1, -i6, kAdjScaleLJ, a.ovf);
        }
    }
    fx_add(q + p.x, c * a.charge[p.y], kAdjScaleQ, a.ovf);
    fx_add(q + p.y, c * a.charge[p.x], kAdjScaleQ, a.ovf);
}

template <typename real>
hipError_t run_pair(mdgpu_ctx* c) {
    Work& w = c->work;
    const Topo& t = c->topo;
    const mdgpu_params& p = c->params;
    AdjointBuf& b = c->adj;
    PairAdjArgs<real> a{};
    const int ncol = c->grid.ncx * c->grid.ncy;
    a.ncl_dev = w.col_cstart + ncol;
    a.max_clusters = c->n_clusters;
    a.split = c->split;
    a.cap = w.list_cap;
    a.list_count = w.list_count;
    a.list_j = w.list_j;
    a.list_mask = w.list_mask;
    if constexpr (sizeof(real) == 4) {
        a.list_delta = w.list_delta_f;
        a.pos = w.slot_pos_f;
    } else {
        a.list_delta = w.list_delta_d;
        a.pos = w.slot_pos_d;
    }
    a.meta = w.slot_meta;
    a.pidx = t.lj_pidx;
    a.ntp = t.n_types + 1;
    a.n_lj = t.n_lj;
    const double rc = p.lj_cutoff;
    a.rc2_lj = static_cast<real>(rc * rc);
    a.rc2_ee = static_cast<real>(p.ee_cutoff * p.ee_cutoff);
    a.rc_lj = static_cast<real>(rc);
    const bool sw = p.lj_switch > 0.0;
    a.rs = static_cast<real>(sw ? p.lj_switch : 0.0);
    a.inv_sw = static_cast<real>(sw ? 1.0 / (rc - p.lj_switch) : 0.0);
    const double r6 = 1.0 / (rc * rc * rc * rc * rc * rc);
    a.c12 = static_cast<real>(r6 * r6);
    a.g12 = static_cast<real>(12.0 * r6 * r6 / rc);
    a.c6 = static_cast<real>(r6);
    a.g6 = static_cast<real>(6.0 * r6 / rc);
    a.alpha = static_cast<real>(p.ewald_alpha);
    a.kc = static_cast<real>(kCoulomb);
    a.q_dir = b.acc + b.q_dir;
    a.lj_a = b.acc + b.lj_a;
    a.lj_b = b.acc + b.lj_b;
    a.lj_ad = b.lj_d;
    a.lj_bd = b.lj_d + t.n_lj;
    a.ovf = reinterpret_cast<unsigned int*>(b.acc + b.flag);
    const size_t shm = 2 * LjBin<real>::kWords * sizeof(unsigned long long) * t.n_lj + sizeof(int) * a.ntp * a.ntp;
    if (shm > 60 * 1024) return hipErrorInvalidValue;
    const int nwork = c->n_clusters * c->split;
    const int nblk = std::max(1, std::min(1024, (nwork + kWaves - 1) / kWaves));
    if (sw)
        hipLaunchKernelGGL((k_pair_adjoint<real, true>), dim3(nblk), dim3(kWaves * kWave), shm, c->stream, a);
    else
        hipLaunchKernelGGL((k_pair_adjoint<real, false>), dim3(nblk), dim3(kWaves * kWave), shm, c->stream, a);
    return hipGetLastError();
}

}  // namespace

hipError_t launch_pair_adjoint(mdgpu_ctx* c, bool mixed) { return mixed ? run_pair<float>(c) : run_pair<double>(c); }

hipError_t launch_bonded_adjoint(mdgpu_ctx* c, const double* coords, bool mixed) {
    const Topo& t = c->topo;
    const int n = t.n_bonds + t.n_ubs + t.n_angles + t.n_dihedrals + t.n_impropers + t.n_cmaps;
    if (n == 0) return hipSuccess;
    AdjointBuf& b = c->adj;
    const int off_ub = 2 * t.n_bond_types, off_ang = off_ub + 2 * t.n_ub_types, off_dih = off_ang + 2 * t.n_angle_types,
              off_imp = off_dih + 2 * t.n_dih_types;
    unsigned int* ovf = reinterpret_cast<unsigned int*>(b.acc + b.flag);
    const dim3 grid((n + kBlock - 1) / kBlock), block(kBlock);
    (void)mixed;  // one FP64 kernel serves both forms
    const BondAdjArgs a{coords, t, b.acc + b.bonded, off_ub, off_ang, off_dih, off_imp, b.acc + b.cmap, ovf};
    hipLaunchKernelGGL(k_bonded_adjoint, grid, block, 0, c->stream, a);
    return hipGetLastError();
}

hipError_t launch_corrections_adjoint(mdgpu_ctx* c) {
    const Topo& t = c->topo;
    const int n = t.n_excl_pairs + t.n_pairs14;
    if (n == 0) return hipSuccess;
    AdjointBuf& b = c->adj;
    CorrAdjArgs a{};
    a.x = c->cur_x;
    a.lx = c->grid.box[0];
    a.ly = c->grid.box[1];
    a.lz = c->grid.box[2];
    a.n_excl = t.n_excl_pairs;
    a.excl = t.excl_pairs;
    a.n14 = t.n_pairs14;
    a.p14 = t.pairs14;
    a.scee = t.scee14;
    a.charge = t.charge;
    a.type = t.type;
    a.nb_index = t.nb_index;
    a.ntypes = t.n_types;
    a.alpha = c->params.ewald_alpha;
    a.q_excl = b.acc + b.q_excl;
    a.q_14 = b.acc + b.q_14;
    a.lj14_a = b.acc + b.lj14_a;
    a.lj14_b = b.acc + b.lj14_b;
    a.ovf = reinterpret_cast<unsigned int*>(b.acc + b.flag);
    hipLaunchKernelGGL(k_corrections_adjoint, dim3((n + kBlock - 1) / kBlock), dim3(kBlock), 0, c->stream, a);
    return hipGetLastError();
}

}  // namespace mdg
