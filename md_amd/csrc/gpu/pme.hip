// Smooth PME reciprocal space (AMBER semantics): B-spline charge spreading,
// a hand-written 3D FFT built from radix-4/2/3 Stockham passes in LDS, the
// influence-function convolution fused into the x pass together with the
// energy and virial, and the force gather.
//
// The reference's PME chain (src/System/PBC.cpp:151-653: spread_charge_pme,
// FFTW r2c/c2r plans, solve_poisson_kspace, gather_forces_pme) never reaches
// its outputs (its forces are cleared by calculate_forces, System.cpp:28, and
// no reciprocal energy is formed).  This implements the smooth PME of
// Essmann et al., J. Chem. Phys. 103, 8577 (1995) with cardinal B-splines of
// order 4:
//   Q(k) = sum_i q_i prod_d M4(u_id - k_d),  u = K x / L (FP64 wrap),
//   E = 1/2 sum_{m != 0} G(m) |FFT(Q)(m)|^2,
//   G(m) = kC exp(-pi^2 m^2 / a^2) / (pi V m^2) B(m),  m = (m_x/L_x, ...),
//   F_i = -q_i sum_k grad_i theta(k) (IFFT(G FFT(Q)))(k).
// Grid layout: real [Kx][Ky][Kz] (z fastest), complex [Kx][Ky][Kz/2+1].
// Passes: spread fused with z r2c -> y -> x (+ G, E, virial, x^-1) -> y^-1
// -> z c2r.
#include "ctx.h"

#include <type_traits>

namespace mdg {

namespace {

template <typename real>
struct C2;
template <>
struct C2<float> { using T = float2; };
template <>
struct C2<double> { using T = double2; };

constexpr int kThreads = 256;
constexpr double kGridScale = 1099511627776.0;  // 2^40, fixed-point charge grid

// v 2^40 as a 64-bit integer from FP32 arithmetic (|v| < 2^23): with
// t = v 2^8 = h + f, h = floor(t), the high word is h and the low word
// f 2^32 (f = t - h exact to FP32 rounding); the result is deterministic.
__device__ __forceinline__ long long fixed40_f32(float v) {
    const float t = v * 256.0f;
    const float h = floorf(t);
    const unsigned int lo = static_cast<unsigned int>(fminf((t - h) * 4294967296.0f, 4294967040.0f));
    const unsigned long long hi = static_cast<unsigned long long>(static_cast<unsigned int>(static_cast<int>(h)));
    return static_cast<long long>((hi << 32) | lo);
}

// theta_j = M4(w + j), dtheta_j = M4'(w + j), j = 0..3 (grid point fl - j)
template <typename real>
__device__ __forceinline__ void bspline4(real w, real* th, real* dth) {
    const real w2 = w * w, w3 = w2 * w, om = 1 - w;
    const real s = static_cast<real>(1.0 / 6.0);
    th[0] = w3 * s;
    th[1] = (1 + 3 * w + 3 * w2 - 3 * w3) * s;
    th[2] = (4 - 6 * w2 + 3 * w3) * s;
    th[3] = om * om * om * s;
    dth[0] = static_cast<real>(0.5) * w2;
    dth[1] = static_cast<real>(0.5) * (1 + 2 * w - 3 * w2);
    dth[2] = static_cast<real>(0.5) * (3 * w2 - 4 * w);
    dth[3] = static_cast<real>(-0.5) * om * om;
}

struct PmeGeom {
    int kx, ky, kz, hz;
    double lx, ly, lz;
    Cell cell;  // triclinic cells: m = R^T (mx, my, mz), forces through R^T
};

// ------------------------------------------------------------ LDS FFT -----
// Batched Stockham FFT of nl lines (nl a power of two) of length n held in
// LDS as buf[i*nl + l].  Passes of radix 4, then 2, then 3; p is the length
// already transformed.  The twiddle table (e^{-2 pi i m/n}, m < n) is in LDS.
// Divisions by the runtime p use a float reciprocal with an exact integer
// correction (i < 2^16).  Each pass reads src and writes dst; the result ends
// in the returned buffer.  The radix is a template parameter of the pass, so
// a butterfly's inputs and outputs stay in registers (indexed by a runtime
// radix they would live in scratch memory).
__device__ __forceinline__ int div_small(int i, int p, float inv_p) {
    int q = static_cast<int>(static_cast<float>(i) * inv_p);
    q -= (q * p > i);
    q += ((q + 1) * p <= i);
    return q;
}

template <typename real, int R>
__device__ __forceinline__ void lds_fft_pass(const typename C2<real>::T* src, typename C2<real>::T* dst, int n,
                                             int lognl, const typename C2<real>::T* __restrict__ tw, int sign,
                                             int p) {
    using c2 = typename C2<real>::T;
    const int nl = 1 << lognl;
    const int nr = n / R, stride_tw = n / (p * R);
    const float inv_p = 1.0f / static_cast<float>(p);
    for (int t = threadIdx.x; t < nr * nl; t += blockDim.x) {
        const int l = t & (nl - 1), i = t >> lognl;
        const int k = i - div_small(i, p, inv_p) * p;
        c2 u[R];
#pragma unroll
        for (int q = 0; q < R; ++q) {
            c2 v = src[(i + q * nr) * nl + l];
            if (q > 0 && k > 0) {
                const c2 w = tw[q * k * stride_tw];   // q k n/(p r) < n
                const real wi = sign < 0 ? w.y : -w.y;
                v = {v.x * w.x - v.y * wi, v.x * wi + v.y * w.x};
            }
            u[q] = v;
        }
        c2 o[R];
        if constexpr (R == 2) {
            o[0] = {u[0].x + u[1].x, u[0].y + u[1].y};
            o[1] = {u[0].x - u[1].x, u[0].y - u[1].y};
        } else if constexpr (R == 4) {
            const c2 s02 = {u[0].x + u[2].x, u[0].y + u[2].y}, d02 = {u[0].x - u[2].x, u[0].y - u[2].y};
            const c2 s13 = {u[1].x + u[3].x, u[1].y + u[3].y}, d13 = {u[1].x - u[3].x, u[1].y - u[3].y};
            // i * sign * d13
            const c2 jd = {-static_cast<real>(sign) * d13.y, static_cast<real>(sign) * d13.x};
            o[0] = {s02.x + s13.x, s02.y + s13.y};
            o[2] = {s02.x - s13.x, s02.y - s13.y};
            o[1] = {d02.x + jd.x, d02.y + jd.y};
            o[3] = {d02.x - jd.x, d02.y - jd.y};
        } else {
            static_assert(R == 3, "radix 2, 3 or 4");
            const real c = static_cast<real>(-0.5);
            const real s = static_cast<real>(sign) * static_cast<real>(0.86602540378443864676);
            const c2 a = {u[1].x + u[2].x, u[1].y + u[2].y}, b = {u[1].x - u[2].x, u[1].y - u[2].y};
            o[0] = {u[0].x + a.x, u[0].y + a.y};
            const c2 m = {u[0].x + c * a.x, u[0].y + c * a.y};
            const c2 isb = {-s * b.y, s * b.x};  // i s b
            o[1] = {m.x + isb.x, m.y + isb.y};
            o[2] = {m.x - isb.x, m.y - isb.y};
        }
        const int j = (i - k) * R + k;
#pragma unroll
        for (int q = 0; q < R; ++q) dst[(j + q * p) * nl + l] = o[q];
    }
}

template <typename real>
__device__ typename C2<real>::T* lds_fft(typename C2<real>::T* src, typename C2<real>::T* dst, int n, int lognl,
                                         const typename C2<real>::T* __restrict__ tw, int sign) {
    using c2 = typename C2<real>::T;
    int rem = n, p = 1;
    while (rem > 1) {
        const int r = (rem % 4 == 0) ? 4 : ((rem % 2 == 0) ? 2 : 3);
        if (r == 4) lds_fft_pass<real, 4>(src, dst, n, lognl, tw, sign, p);
        else if (r == 2) lds_fft_pass<real, 2>(src, dst, n, lognl, tw, sign, p);
        else lds_fft_pass<real, 3>(src, dst, n, lognl, tw, sign, p);
        __syncthreads();
        c2* tmp = src;
        src = dst;
        dst = tmp;
        p *= r;
        rem /= r;
    }
    return src;
}

// Twiddles of one axis into LDS (after the two line buffers).
template <typename real>
__device__ __forceinline__ const typename C2<real>::T* stage_twiddles(typename C2<real>::T* s_tw,
                                                                      const typename C2<real>::T* __restrict__ tw,
                                                                      int n) {
    for (int m = threadIdx.x; m < n; m += blockDim.x) s_tw[m] = tw[m];
    return s_tw;
}

// ------------------------------------------------- spread + z r2c --------
// Owner-computes charge spreading fused with the z transform.  A block owns
// by z-lines (x = X0, y in [Y0, Y0+by), all z) and builds them in LDS.  It
// visits the slots of the sort columns that can hold atoms reaching its lines
// (an atom with base index f contributes to points f - 0..3, so these lines
// receive f_x in [X0, X0+3] and f_y in [Y0, Y0+by+2], i.e. x in
// [X0 hx, (X0+4) hx) and y in [Y0 hy, (Y0+by+3) hy)), reads each slot's
// spline record (base indices, fractions, charge; written once per evaluation
// by the cluster set-up), and adds q theta_x theta_y theta_z with 64-bit
// integer LDS atomics (2^40 fixed point: the sum does not depend on their
// order).  It then converts the lines and transforms them along z in place.
// Every grid point is written by one block, so the real grid never goes to
// memory and needs no clearing.
//
// Each wave takes 64 candidate slots at a time, in two phases.  (A) Each
// lane tests its slot; a slot that reaches the block's lines gets a record in
// the wave's LDS buffer: q theta_x(a) for its one x point a, theta_y(0..3),
// theta_z(0..3), f_y - Y0 and f_z.  (B) 16 lanes per record, four records
// per step, lane (b, c) adds the point (a, b, c) when f_y - b lies in the
// block.  The products are formed as (q theta_x) theta_y, then theta_z.
struct SpreadCols {
    int ncx, ncy;
    double sx, sy;
    const int* col_cstart;   // [ncol+1] cluster offsets
    const int* fl;           // [slots] packed base indices
    const void* w;           // [slots] fractions + charge (double4 / float4)
    double margin;           // how far an atom may sit outside its column (Verlet list: skin / 2, in
                             // units of s_x |a| and s_y |b| for a triclinic cell)
};

constexpr int kSpreadCols = 256;  // columns per pass of the candidate loop
constexpr int kSpreadLines = 8;   // z-lines per spreading block (one x, eight y)
constexpr int kSpreadRec = 9;     // reals of an LDS spreading record

template <typename real>
__global__ __launch_bounds__(kThreads) void k_spread_fft_z(PmeGeom g, SpreadCols sc, int by, int lognl,
                                                           typename C2<real>::T* __restrict__ cg,
                                                           const typename C2<real>::T* __restrict__ tw_g,
                                                           unsigned int* overflow) {
    using c2 = typename C2<real>::T;
    using R4 = typename std::conditional<sizeof(real) == 8, double4, float4>::type;
    extern __shared__ unsigned char smem[];
    __shared__ int s_first[kSpreadCols], s_pre[kSpreadCols + 1];
    // records per candidate of a wave's chunk: [qa, theta_y 0..3, theta_z 0..3]
    // and dyp | flz << 16
    __shared__ real s_rec[kThreads / kWave][kWave * kSpreadRec];
    __shared__ int s_pk[kThreads / kWave][kWave];
    const int nl = 1 << lognl;  // == by
    c2* A = reinterpret_cast<c2*>(smem);
    c2* B = A + g.kz * nl;
    unsigned long long* acc = reinterpret_cast<unsigned long long*>(B);  // [kz][nl], aliases B
    const c2* tw = stage_twiddles<real>(B + g.kz * nl, tw_g, g.kz);
    for (int t = threadIdx.x; t < nl * g.kz; t += blockDim.x) acc[t] = 0ull;
    const int X0 = blockIdx.x, Y0 = blockIdx.y * by;
    const int lane = threadIdx.x & (kWave - 1), wv = threadIdx.x / kWave;
    real* rec = s_rec[wv];
    int* pks = s_pk[wv];
    const double hx = g.lx / g.kx, hy = g.ly / g.ky, tol = 1e-6 + sc.margin;
    // columns overlapping [X0 hx, (X0+4) hx) x [Y0 hy, (Y0+by+3) hy), with
    // 1e-6 A for the rounding of floor(u) against the column binning, grown
    // by how far an atom of a kept Verlet list may have left its column
    int cx0 = static_cast<int>(floor((X0 * hx - tol) / sc.sx));
    int ncxr = static_cast<int>(floor(((X0 + 4) * hx + tol) / sc.sx)) - cx0 + 1;
    int cy0 = static_cast<int>(floor((Y0 * hy - tol) / sc.sy));
    int ncyr = static_cast<int>(floor(((Y0 + by + 3) * hy + tol) / sc.sy)) - cy0 + 1;
    if (ncxr >= sc.ncx) { cx0 = 0; ncxr = sc.ncx; }
    if (ncyr >= sc.ncy) { cy0 = 0; ncyr = sc.ncy; }
    const int ncand = ncxr * ncyr;
    const R4* wrec = static_cast<const R4*>(sc.w);
    // phase B lane roles
    const int rk = lane >> 4, rb = (lane >> 2) & 3, rc = lane & 3;
    for (int cb = 0; cb < ncand; cb += kSpreadCols) {
        const int nc = min(kSpreadCols, ncand - cb);
        __syncthreads();  // the previous pass is done with s_first / s_pre
        if (threadIdx.x < nc) {
            const int k = cb + threadIdx.x;
            const int cy = ((cy0 + k / ncxr) % sc.ncy + sc.ncy) % sc.ncy;
            const int cx = ((cx0 + k % ncxr) % sc.ncx + sc.ncx) % sc.ncx;
            const int col = cx + sc.ncx * cy;
            s_first[threadIdx.x] = sc.col_cstart[col] * kCluster;
            s_pre[threadIdx.x + 1] = (sc.col_cstart[col + 1] - sc.col_cstart[col]) * kCluster;
        }
        if (threadIdx.x == 0) s_pre[0] = 0;
        __syncthreads();
        if (threadIdx.x == 0)
            for (int k = 0; k < nc; ++k) s_pre[k + 1] += s_pre[k];
        __syncthreads();
        const int total = s_pre[nc];
        // each wave takes chunks of 64 candidates, round robin
        for (int t0 = wv * kWave; t0 < total; t0 += blockDim.x) {
            // phase A: this lane's candidate
            const int t = t0 + lane;
            bool take = false;
            real th[3][4], dth[3][4];
            real qa = 0;
            int dyp = 0, flz = 0;
            if (t < total) {
                // column of candidate t: the last k with s_pre[k] <= t
                int lo = 0, hi = nc - 1;
                while (lo < hi) {
                    const int mid = (lo + hi + 1) >> 1;
                    if (s_pre[mid] <= t) lo = mid;
                    else hi = mid - 1;
                }
                const int s = s_first[lo] + (t - s_pre[lo]);
                const int fl = sc.fl[s];
                const R4 wq = wrec[s];
                const int flx = fl & 1023, fly = (fl >> 10) & 1023;
                flz = fl >> 20;
                int dxp = flx - X0;
                dyp = fly - Y0;
                dxp += dxp < 0 ? g.kx : 0;
                dyp += dyp < 0 ? g.ky : 0;
                // padding slots and uncharged atoms add nothing
                take = wq.w != 0 && dxp <= 3 && dyp <= by + 2;
                if (take && sizeof(real) == 4 &&
                    (!(fabs(wq.w) < static_cast<real>(8.0e6)) ||
                     !(fabs(wq.x) + fabs(wq.y) + fabs(wq.z) < static_cast<real>(4.0)))) {
                    // FP32 products and conversion need charges far below
                    // 2^23 e; a non-finite record raises the overflow flag
                    atomicOr(overflow, 1u);
                    take = false;
                }
                if (take) {
                    bspline4<real>(wq.x, th[0], dth[0]);
                    bspline4<real>(wq.y, th[1], dth[1]);
                    bspline4<real>(wq.z, th[2], dth[2]);
                    const real tx = dxp == 0 ? th[0][0] : (dxp == 1 ? th[0][1] : (dxp == 2 ? th[0][2] : th[0][3]));
                    qa = wq.w * tx;
                }
            }
            const unsigned long long bal = __ballot(take);
            const int nrec = __popcll(bal);
            if (nrec == 0) continue;
            if (take) {
                const int r = __builtin_amdgcn_mbcnt_hi(static_cast<unsigned int>(bal >> 32),
                                                        __builtin_amdgcn_mbcnt_lo(static_cast<unsigned int>(bal), 0u));
                real* e = rec + r * kSpreadRec;
                e[0] = qa;
                for (int q = 0; q < 4; ++q) {
                    e[1 + q] = th[1][q];
                    e[5 + q] = th[2][q];
                }
                pks[r] = dyp | (flz << 16);
            }
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            // phase B: four records per step, lane (rk, rb, rc)
            for (int r0 = 0; r0 < nrec; r0 += 4) {
                const int r = r0 + rk;
                if (r < nrec) {
                    const real* e = rec + r * kSpreadRec;
                    const int pk = pks[r];
                    const int lj = (pk & 0xffff) - rb;
                    if (lj >= 0 && lj < by && Y0 + lj < g.ky) {
                        int iz = (pk >> 16) - rc;
                        iz += iz < 0 ? g.kz : 0;
                        const real qab = e[0] * e[1 + rb];
                        long long v;
                        if constexpr (sizeof(real) == 4) v = fixed40_f32(qab * e[5 + rc]);
                        else v = to_fixed(qab * static_cast<double>(e[5 + rc]), kGridScale, overflow);
                        if (v != 0) atomicAdd(acc + iz * nl + lj, static_cast<unsigned long long>(v));
                    }
                }
            }
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        }
    }
    __syncthreads();
    for (int t = threadIdx.x; t < nl * g.kz; t += blockDim.x)
        A[t] = {static_cast<real>(static_cast<double>(static_cast<long long>(acc[t])) * (1.0 / kGridScale)),
                static_cast<real>(0)};
    __syncthreads();
    c2* R = lds_fft<real>(A, B, g.kz, lognl, tw, -1);
    for (int t = threadIdx.x; t < nl * g.hz; t += blockDim.x) {
        const int lj = t / g.hz, k = t - (t / g.hz) * g.hz;
        if (Y0 + lj < g.ky) cg[(static_cast<size_t>(X0) * g.ky + Y0 + lj) * g.hz + k] = R[k * nl + lj];
    }
}

// --------------------------------------------------------------- y pass ---
template <typename real>
__global__ void k_fft_y(PmeGeom g, int lognl, typename C2<real>::T* __restrict__ cg,
                        const typename C2<real>::T* __restrict__ tw_g, int sign) {
    using c2 = typename C2<real>::T;
    extern __shared__ unsigned char smem[];
    const int nl = 1 << lognl;
    c2* A = reinterpret_cast<c2*>(smem);
    c2* B = A + g.ky * nl;
    const c2* tw = stage_twiddles<real>(B + g.ky * nl, tw_g, g.ky);
    const int x = blockIdx.y, kz0 = blockIdx.x * nl;
    const int nv = min(nl, g.hz - kz0);
    for (int t = threadIdx.x; t < nl * g.ky; t += blockDim.x) {
        const int l = t % nl, y = t / nl;
        A[y * nl + l] = l < nv ? cg[(static_cast<size_t>(x) * g.ky + y) * g.hz + kz0 + l] : c2{0, 0};
    }
    __syncthreads();
    c2* R = lds_fft<real>(A, B, g.ky, lognl, tw, sign);
    for (int t = threadIdx.x; t < nl * g.ky; t += blockDim.x) {
        const int l = t % nl, y = t / nl;
        if (l < nv) cg[(static_cast<size_t>(x) * g.ky + y) * g.hz + kz0 + l] = R[y * nl + l];
    }
}

// ----------------------------------------------- x pass + convolution -----
template <typename real>
__global__ void k_fft_x_conv(PmeGeom g, int lognl, typename C2<real>::T* __restrict__ cg,
                             const typename C2<real>::T* __restrict__ tw_g, const double* __restrict__ bmod,
                             double alpha, double* partial, int ld, int row0) {
    using c2 = typename C2<real>::T;
    extern __shared__ unsigned char smem[];
    const int nl = 1 << lognl;
    c2* A = reinterpret_cast<c2*>(smem);
    c2* B = A + g.kx * nl;
    const c2* tw = stage_twiddles<real>(B + g.kx * nl, tw_g, g.kx);
    const int y = blockIdx.y, kz0 = blockIdx.x * nl;
    const int nv = min(nl, g.hz - kz0);
    for (int t = threadIdx.x; t < nl * g.kx; t += blockDim.x) {
        const int l = t % nl, x = t / nl;
        A[x * nl + l] = l < nv ? cg[(static_cast<size_t>(x) * g.ky + y) * g.hz + kz0 + l] : c2{0, 0};
    }
    __syncthreads();
    c2* S = lds_fft<real>(A, B, g.kx, lognl, tw, -1);
    c2* other = S == A ? B : A;
    // G(m) |S|^2, virial, and S <- G S
    const double V = g.cell.tri ? cell_volume(g.cell) : g.lx * g.ly * g.lz;
    const double pa2 = kPi * kPi / (alpha * alpha);
    const int my = y <= g.ky / 2 ? y : y - g.ky;
    const double* bx = bmod;
    const double* by = bmod + g.kx;
    const double* bz = bmod + g.kx + g.ky;
    double e[7] = {0, 0, 0, 0, 0, 0, 0};
    for (int t = threadIdx.x; t < nl * g.kx; t += blockDim.x) {
        const int l = t % nl, x = t / nl;
        const int kz = kz0 + l;
        c2 s = S[x * nl + l];
        if (l < nv) {
            const int mx = x <= g.kx / 2 ? x : x - g.kx;
            if (mx == 0 && my == 0 && kz == 0) {
                s = {0, 0};
            } else {
                double fx, fy, fz;
                if (!g.cell.tri) {
                    fx = mx / g.lx;
                    fy = my / g.ly;
                    fz = kz / g.lz;
                } else {
                    // the conjugate half (weight 2) is taken at -m, also on
                    // the Nyquist planes
                    fx = mx * g.cell.r00;
                    fy = mx * g.cell.r01 + my * g.cell.r11;
                    fz = mx * g.cell.r02 + my * g.cell.r12 + kz * g.cell.r22;
                }
                const double m2 = fx * fx + fy * fy + fz * fz;
                const double gm = kCoulomb * exp(-pa2 * m2) / (kPi * V * m2) * bx[x] * by[y] * bz[kz];
                const double wgt = (kz == 0 || 2 * kz == g.kz) ? 1.0 : 2.0;
                const double em = 0.5 * wgt * gm * (static_cast<double>(s.x) * s.x + static_cast<double>(s.y) * s.y);
                const double fac = 2.0 * (1.0 + pa2 * m2) / m2;
                e[0] += em;
                e[1] += em * (1.0 - fac * fx * fx);
                e[2] += em * (1.0 - fac * fy * fy);
                e[3] += em * (1.0 - fac * fz * fz);
                e[4] += em * (-fac * fx * fy);
                e[5] += em * (-fac * fx * fz);
                e[6] += em * (-fac * fy * fz);
                s = {static_cast<real>(s.x * gm), static_cast<real>(s.y * gm)};
            }
        }
        S[x * nl + l] = s;
    }
    __syncthreads();
    c2* R = lds_fft<real>(S, other, g.kx, lognl, tw, +1);
    for (int t = threadIdx.x; t < nl * g.kx; t += blockDim.x) {
        const int l = t % nl, x = t / nl;
        if (l < nv) cg[(static_cast<size_t>(x) * g.ky + y) * g.hz + kz0 + l] = R[x * nl + l];
    }
    const int slots[7] = {E_EW_RECIP, E_VIR_XX, E_VIR_YY, E_VIR_ZZ, E_VIR_XY, E_VIR_XZ, E_VIR_YZ};
    const int row = row0 + blockIdx.y * gridDim.x + blockIdx.x;
    block_sum_store<7>(e, partial, ld, row, slots);
}

// --------------------------------------------------------------- z c2r ----
template <typename real>
__global__ void k_fft_z_c2r(PmeGeom g, int lognl, const typename C2<real>::T* __restrict__ cg,
                            const typename C2<real>::T* __restrict__ tw_g, real* __restrict__ phi) {
    using c2 = typename C2<real>::T;
    extern __shared__ unsigned char smem[];
    const int nl = 1 << lognl;
    c2* A = reinterpret_cast<c2*>(smem);
    c2* B = A + g.kz * nl;
    const c2* tw = stage_twiddles<real>(B + g.kz * nl, tw_g, g.kz);
    const int nlines = g.kx * g.ky;
    const int l0 = blockIdx.x * nl;
    const int nv = min(nl, nlines - l0);
    for (int t = threadIdx.x; t < nl * g.kz; t += blockDim.x) {
        const int l = t / g.kz, k = t % g.kz;
        c2 v = {0, 0};
        if (l < nv) {
            const size_t row = static_cast<size_t>(l0 + l) * g.hz;
            if (k < g.hz) {
                v = cg[row + k];
            } else {
                const c2 m = cg[row + (g.kz - k)];
                v = {m.x, -m.y};
            }
        }
        A[k * nl + l] = v;
    }
    __syncthreads();
    c2* R = lds_fft<real>(A, B, g.kz, lognl, tw, +1);
    for (int t = threadIdx.x; t < nl * g.kz; t += blockDim.x) {
        const int l = t / g.kz, i = t % g.kz;
        if (l < nv) phi[static_cast<size_t>(l0) * g.kz + t] = R[i * nl + l].x;
    }
}

// ---------------------------------------------------------------- gather --
// One thread per slot, from the same spline records as the spreading; the
// forces go to the slot-ordered accumulators (3 per slot).
template <typename real>
__global__ void k_gather(const int* ncl_dev, int max_slots, const int* __restrict__ slot_atom,
                         const int* __restrict__ fl_rec, const void* w_rec, PmeGeom g,
                         const real* __restrict__ phi, double* force_d, FixedAcc fxa, double* pot) {
    using R4 = typename std::conditional<sizeof(real) == 8, double4, float4>::type;
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= max_slots || s >= *ncl_dev * kCluster) return;
    const int i = slot_atom[s];
    if (i < 0) return;  // padding slot
    const R4 wq = static_cast<const R4*>(w_rec)[s];
    if (wq.w == 0 && !pot) return;  // no force on an uncharged atom; its potential is still needed
    const int fl = fl_rec[s];
    const int flx = fl & 1023, fly = (fl >> 10) & 1023, flz = fl >> 20;
    real th[3][4], dth[3][4];
    bspline4<real>(wq.x, th[0], dth[0]);
    bspline4<real>(wq.y, th[1], dth[1]);
    bspline4<real>(wq.z, th[2], dth[2]);
    // per (x, y) row the z sums sum_c theta_z phi and sum_c theta_z' phi, then
    // the x and y weights; one x plane at a time keeps 16 loads in flight
    int iz[4];
    for (int c = 0; c < 4; ++c) {
        iz[c] = flz - c;
        iz[c] += iz[c] < 0 ? g.kz : 0;
    }
    real gx = 0, gy = 0, gz = 0, ph = 0;
#pragma unroll 1
    for (int a = 0; a < 4; ++a) {
        int ix = flx - a;
        ix += ix < 0 ? g.kx : 0;
#pragma unroll
        for (int b = 0; b < 4; ++b) {
            int iy = fly - b;
            iy += iy < 0 ? g.ky : 0;
            const real* row = phi + (static_cast<size_t>(ix) * g.ky + iy) * g.kz;
            real sz = 0, dz = 0;
#pragma unroll
            for (int c = 0; c < 4; ++c) {
                const real p = row[iz[c]];
                sz += th[2][c] * p;
                dz += dth[2][c] * p;
            }
            const real wxy = th[0][a] * th[1][b];
            gx += dth[0][a] * th[1][b] * sz;
            gy += th[0][a] * dth[1][b] * sz;
            gz += wxy * dz;
            ph += wxy * sz;
        }
    }
    if (pot) {
        // the adjoint of the spreading: dE_rec/dq_i = sum_k theta_i(k) phi(k)
        pot[i] = static_cast<double>(ph);
        return;
    }
    const double q = static_cast<double>(wq.w);
    double fx, fy, fz;
    if (!g.cell.tri) {
        fx = -q * static_cast<double>(gx) * g.kx / g.lx;
        fy = -q * static_cast<double>(gy) * g.ky / g.ly;
        fz = -q * static_cast<double>(gz) * g.kz / g.lz;
    } else {
        // dE/dx = R^T dE/ds, dE/ds_d = K_d dE/du_d
        const double ux = static_cast<double>(gx) * g.kx, uy = static_cast<double>(gy) * g.ky;
        const double uz = static_cast<double>(gz) * g.kz;
        fx = -q * (ux * g.cell.r00);
        fy = -q * (ux * g.cell.r01 + uy * g.cell.r11);
        fz = -q * (ux * g.cell.r02 + uy * g.cell.r12 + uz * g.cell.r22);
    }
    if constexpr (sizeof(real) == 8) {
        acc_add(force_d, 3 * s, fx);
        acc_add(force_d, 3 * s + 1, fy);
        acc_add(force_d, 3 * s + 2, fz);
    } else {
        acc_add(fxa, 3 * s, fx);
        acc_add(fxa, 3 * s + 1, fy);
        acc_add(fxa, 3 * s + 2, fz);
    }
}

// Lines per FFT block: two LDS buffers of n x nl complex values plus n
// twiddles within 64 KB, at most `cap` (a power of two; 0 = 16).
int lines_per_block(int n, size_t elem, int cap = 0) {
    int nl = cap > 0 ? cap : 16;
    while (nl > 1 && (2 * static_cast<size_t>(n) * nl + n) * elem > 65536) nl >>= 1;
    return nl;
}

int ilog2(int v) {
    int l = 0;
    while ((1 << (l + 1)) <= v) ++l;
    return l;
}

template <typename real>
hipError_t run_pme(mdgpu_ctx* c, int row0, int* rows_used, double* pot) {
    using c2 = typename C2<real>::T;
    Work& w = c->work;
    const int n = c->topo.n_atoms;
    PmeGeom g{w.pme_k[0], w.pme_k[1], w.pme_k[2], w.pme_k[2] / 2 + 1, c->grid.cell.l[0], c->grid.cell.l[1],
              c->grid.cell.l[2], c->grid.cell};
    const bool dbl = sizeof(real) == 8;
    c2* cg = static_cast<c2*>(w.cgrid);
    const c2* tw = dbl ? reinterpret_cast<const c2*>(w.twiddle_d) : reinterpret_cast<const c2*>(w.twiddle_f);
    const c2* twx = tw;
    const c2* twy = tw + g.kx;
    const c2* twz = tw + g.kx + g.ky;
    auto& T = c->timing;
    auto mark = [&](int k, bool start) {
        if (T.mode == 1) (void)record_timing_event(c, c->ev[2 * k + (start ? 0 : 1)]);
    };
    const int max_slots = c->n_clusters * kCluster;
    const int ncol = c->grid.ncx * c->grid.ncy;
    const void* wrec = dbl ? static_cast<const void*>(w.slot_pme_wd) : static_cast<const void*>(w.slot_pme_wf);
    // spreading fused with the z transform (timed as the spread)
    mark(MDGPU_K_SPREAD, true);
    {
        // up to 8 z-lines per block (1 x 8 in x-y): measured faster than 4 or
        // 16 next to the concurrent pair kernel
        const int nl = lines_per_block(g.kz, sizeof(c2), kSpreadLines);
        const size_t shm = (2 * static_cast<size_t>(g.kz) * nl + g.kz) * sizeof(c2);
        const SpreadCols sc{c->grid.ncx, c->grid.ncy, c->grid.sx, c->grid.sy, w.col_cstart, w.slot_pme_fl, wrec,
                            c->skin > 0 ? 0.5 * c->skin * c->grid.frac_scale : 0.0};
        const dim3 grid(g.kx, (g.ky + nl - 1) / nl);
        hipLaunchKernelGGL(k_spread_fft_z<real>, grid, dim3(kThreads), shm, c->stream, g, sc, nl, ilog2(nl), cg,
                           twz, w.flags);
    }
    mark(MDGPU_K_SPREAD, false);
    mark(MDGPU_K_FFT, true);
    const int nly = lines_per_block(g.ky, sizeof(c2), c->fft_lines);
    const size_t shy = (2 * static_cast<size_t>(g.ky) * nly + g.ky) * sizeof(c2);
    hipLaunchKernelGGL(k_fft_y<real>, dim3((g.hz + nly - 1) / nly, g.kx), dim3(kThreads), shy, c->stream, g,
                       ilog2(nly), cg, twy, -1);
    mark(MDGPU_K_FFT, false);
    mark(MDGPU_K_CONV, true);
    const int nlx = lines_per_block(g.kx, sizeof(c2), c->fft_lines);
    const size_t shx = (2 * static_cast<size_t>(g.kx) * nlx + g.kx) * sizeof(c2);
    const dim3 gx((g.hz + nlx - 1) / nlx, g.ky);
    hipLaunchKernelGGL(k_fft_x_conv<real>, gx, dim3(kThreads), shx, c->stream, g, ilog2(nlx), cg, twx, w.bsp_mod,
                       c->params.ewald_alpha, w.partial, static_cast<int>(w.cap_partial), row0);
    mark(MDGPU_K_CONV, false);
    *rows_used = static_cast<int>(gx.x * gx.y);
    mark(MDGPU_K_COUNT, true);  // inverse passes: the extra event pair
    hipLaunchKernelGGL(k_fft_y<real>, dim3((g.hz + nly - 1) / nly, g.kx), dim3(kThreads), shy, c->stream, g,
                       ilog2(nly), cg, twy, +1);
    real* phi = dbl ? reinterpret_cast<real*>(w.grid_phi_d) : reinterpret_cast<real*>(w.grid_f);
    {
        const int nl = lines_per_block(g.kz, sizeof(c2), c->fft_lines);
        const size_t shm = (2 * static_cast<size_t>(g.kz) * nl + g.kz) * sizeof(c2);
        hipLaunchKernelGGL(k_fft_z_c2r<real>, dim3((g.kx * g.ky + nl - 1) / nl), dim3(kThreads), shm, c->stream, g,
                           ilog2(nl), cg, twz, phi);
    }
    mark(MDGPU_K_COUNT, false);
    mark(MDGPU_K_GATHER, true);
    hipLaunchKernelGGL(k_gather<real>, dim3((max_slots + kThreads - 1) / kThreads), dim3(kThreads), 0, c->stream,
                       w.col_cstart + ncol, max_slots, w.slot_atom, w.slot_pme_fl, wrec, g, phi, w.force_d_s,
                       FixedAcc{w.force_fx_s, fixed_scale(c->params.semantics), w.flags}, pot);
    mark(MDGPU_K_GATHER, false);
    return hipGetLastError();
}

}  // namespace

int pme_rows(const mdgpu_ctx* c) {
    const int* k = c->work.pme_k;
    if (k[0] == 0) return 0;
    const int hz = k[2] / 2 + 1;
    // rows of the smaller of the two complex sizes bound both forms
    const int nlx = std::min(lines_per_block(k[0], 16, c->fft_lines), lines_per_block(k[0], 8, c->fft_lines));
    return ((hz + nlx - 1) / nlx) * k[1] * 2;
}

hipError_t launch_pme(mdgpu_ctx* c, bool mixed, int row0, int* rows_used) {
    *rows_used = 0;
    if (c->work.pme_k[0] == 0) return hipSuccess;
    return mixed ? run_pme<float>(c, row0, rows_used, nullptr) : run_pme<double>(c, row0, rows_used, nullptr);
}

hipError_t launch_pme_potential(mdgpu_ctx* c, bool mixed, int row0, double* pot) {
    int used = 0;
    if (c->work.pme_k[0] == 0) return hipSuccess;
    return mixed ? run_pme<float>(c, row0, &used, pot) : run_pme<double>(c, row0, &used, pot);
}

}  // namespace mdg
