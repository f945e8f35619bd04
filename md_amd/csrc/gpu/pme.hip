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
// Passes: z r2c (also clears the charge grid) -> y -> x (+ G, E, virial, x^-1)
// -> y^-1 -> z c2r.
#include "ctx.h"

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
};

__device__ __forceinline__ void frac(const double* wrapped, int i, const PmeGeom& g, int* fl, double* w) {
    const double u[3] = {wrapped[3 * i] * g.kx / g.lx, wrapped[3 * i + 1] * g.ky / g.ly,
                         wrapped[3 * i + 2] * g.kz / g.lz};
    for (int d = 0; d < 3; ++d) {
        fl[d] = static_cast<int>(floor(u[d]));
        w[d] = u[d] - fl[d];
    }
}

__device__ __forceinline__ int wrapk(int k, int n) {
    k %= n;
    return k < 0 ? k + n : k;
}

// ---------------------------------------------------------------- spread --
template <typename real>
__global__ void k_spread(int n, const double* __restrict__ wrapped, const double* __restrict__ charge, PmeGeom g,
                         double* __restrict__ grid_d, unsigned long long* __restrict__ grid_fx,
                         unsigned int* overflow) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double q = charge[i];
    if (q == 0.0) return;
    int fl[3];
    double w[3];
    frac(wrapped, i, g, fl, w);
    real th[3][4], dth[3][4];
    for (int d = 0; d < 3; ++d) bspline4<real>(static_cast<real>(w[d]), th[d], dth[d]);
    const real qr = static_cast<real>(q);
    for (int a = 0; a < 4; ++a) {
        const int ix = wrapk(fl[0] - a, g.kx);
        for (int b = 0; b < 4; ++b) {
            const int iy = wrapk(fl[1] - b, g.ky);
            const real qab = qr * th[0][a] * th[1][b];
            const size_t row = (static_cast<size_t>(ix) * g.ky + iy) * g.kz;
            for (int c = 0; c < 4; ++c) {
                const int iz = wrapk(fl[2] - c, g.kz);
                const real v = qab * th[2][c];
                if constexpr (sizeof(real) == 8) {
                    atomicAdd(grid_d + row + iz, v);
                } else {
                    const long long fx = to_fixed(static_cast<double>(v), kGridScale, overflow);
                    atomicAdd(grid_fx + row + iz, static_cast<unsigned long long>(fx));
                }
            }
        }
    }
}

// ------------------------------------------------------------ LDS FFT -----
// Batched Stockham FFT of nl lines (nl a power of two) of length n held in
// LDS as buf[i*nl + l].  Passes of radix 4, then 2, then 3; p is the length
// already transformed.  The twiddle table (e^{-2 pi i m/n}, m < n) is in LDS.
// Divisions by the runtime p use a float reciprocal with an exact integer
// correction (i < 2^16).  Each pass reads src and writes dst; the result ends
// in the returned buffer.
__device__ __forceinline__ int div_small(int i, int p, float inv_p) {
    int q = static_cast<int>(static_cast<float>(i) * inv_p);
    q -= (q * p > i);
    q += ((q + 1) * p <= i);
    return q;
}

template <typename real>
__device__ typename C2<real>::T* lds_fft(typename C2<real>::T* src, typename C2<real>::T* dst, int n, int lognl,
                                         const typename C2<real>::T* __restrict__ tw, int sign) {
    using c2 = typename C2<real>::T;
    const int nl = 1 << lognl;
    int rem = n, p = 1;
    while (rem > 1) {
        const int r = (rem % 4 == 0) ? 4 : ((rem % 2 == 0) ? 2 : 3);
        const int nr = n / r, stride_tw = n / (p * r);
        const float inv_p = 1.0f / static_cast<float>(p);
        for (int t = threadIdx.x; t < nr * nl; t += blockDim.x) {
            const int l = t & (nl - 1), i = t >> lognl;
            const int k = i - div_small(i, p, inv_p) * p;
            c2 u[4];
            for (int q = 0; q < r; ++q) {
                c2 v = src[(i + q * nr) * nl + l];
                if (q > 0 && k > 0) {
                    const c2 w = tw[q * k * stride_tw];   // q k n/(p r) < n
                    const real wi = sign < 0 ? w.y : -w.y;
                    v = {v.x * w.x - v.y * wi, v.x * wi + v.y * w.x};
                }
                u[q] = v;
            }
            c2 o[4];
            if (r == 2) {
                o[0] = {u[0].x + u[1].x, u[0].y + u[1].y};
                o[1] = {u[0].x - u[1].x, u[0].y - u[1].y};
            } else if (r == 4) {
                const c2 s02 = {u[0].x + u[2].x, u[0].y + u[2].y}, d02 = {u[0].x - u[2].x, u[0].y - u[2].y};
                const c2 s13 = {u[1].x + u[3].x, u[1].y + u[3].y}, d13 = {u[1].x - u[3].x, u[1].y - u[3].y};
                // i * sign * d13
                const c2 jd = {-static_cast<real>(sign) * d13.y, static_cast<real>(sign) * d13.x};
                o[0] = {s02.x + s13.x, s02.y + s13.y};
                o[2] = {s02.x - s13.x, s02.y - s13.y};
                o[1] = {d02.x + jd.x, d02.y + jd.y};
                o[3] = {d02.x - jd.x, d02.y - jd.y};
            } else {
                const real c = static_cast<real>(-0.5);
                const real s = static_cast<real>(sign) * static_cast<real>(0.86602540378443864676);
                const c2 a = {u[1].x + u[2].x, u[1].y + u[2].y}, b = {u[1].x - u[2].x, u[1].y - u[2].y};
                o[0] = {u[0].x + a.x, u[0].y + a.y};
                const c2 m = {u[0].x + c * a.x, u[0].y + c * a.y};
                const c2 isb = {-s * b.y, s * b.x};  // i s b
                o[1] = {m.x + isb.x, m.y + isb.y};
                o[2] = {m.x - isb.x, m.y - isb.y};
            }
            const int j = (i - k) * r + k;
            for (int q = 0; q < r; ++q) dst[(j + q * p) * nl + l] = o[q];
        }
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

// --------------------------------------------------------------- z r2c ----
template <typename real>
__global__ void k_fft_z_r2c(PmeGeom g, int lognl, double* grid_d, unsigned long long* grid_fx,
                            typename C2<real>::T* __restrict__ cg, const typename C2<real>::T* __restrict__ tw_g) {
    using c2 = typename C2<real>::T;
    extern __shared__ unsigned char smem[];
    const int nl = 1 << lognl;
    c2* A = reinterpret_cast<c2*>(smem);
    c2* B = A + g.kz * nl;
    const c2* tw = stage_twiddles<real>(B + g.kz * nl, tw_g, g.kz);
    const int nlines = g.kx * g.ky;
    const int l0 = blockIdx.x * nl;
    const int nv = min(nl, nlines - l0);
    const size_t base = static_cast<size_t>(l0) * g.kz;
    for (int t = threadIdx.x; t < nl * g.kz; t += blockDim.x) {
        const int l = t / g.kz, i = t % g.kz;
        real v = 0;
        if (l < nv) {
            if constexpr (sizeof(real) == 8) {
                v = grid_d[base + t];
                grid_d[base + t] = 0.0;
            } else {
                v = static_cast<real>(static_cast<double>(static_cast<long long>(grid_fx[base + t])) / kGridScale);
                grid_fx[base + t] = 0ull;
            }
        }
        A[i * nl + l] = {v, static_cast<real>(0)};
    }
    __syncthreads();
    c2* R = lds_fft<real>(A, B, g.kz, lognl, tw, -1);
    for (int t = threadIdx.x; t < nl * g.hz; t += blockDim.x) {
        const int l = t / g.hz, k = t % g.hz;
        if (l < nv) cg[static_cast<size_t>(l0 + l) * g.hz + k] = R[k * nl + l];
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
                             double alpha, double* partial, int row0) {
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
    const double V = g.lx * g.ly * g.lz;
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
                const double fx = mx / g.lx, fy = my / g.ly, fz = kz / g.lz;
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
    block_sum_store<7>(e, partial + static_cast<size_t>(row) * E_NSLOTS, slots);
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
template <typename real>
__global__ void k_gather(int n, const double* __restrict__ wrapped, const double* __restrict__ charge, PmeGeom g,
                         const real* __restrict__ phi, double* force_d, FixedAcc fxa) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double q = charge[i];
    if (q == 0.0) return;
    int fl[3];
    double w[3];
    frac(wrapped, i, g, fl, w);
    real th[3][4], dth[3][4];
    for (int d = 0; d < 3; ++d) bspline4<real>(static_cast<real>(w[d]), th[d], dth[d]);
    real gx = 0, gy = 0, gz = 0;
    for (int a = 0; a < 4; ++a) {
        const int ix = wrapk(fl[0] - a, g.kx);
        for (int b = 0; b < 4; ++b) {
            const int iy = wrapk(fl[1] - b, g.ky);
            const real* row = phi + (static_cast<size_t>(ix) * g.ky + iy) * g.kz;
            for (int c = 0; c < 4; ++c) {
                const real p = row[wrapk(fl[2] - c, g.kz)];
                gx += dth[0][a] * th[1][b] * th[2][c] * p;
                gy += th[0][a] * dth[1][b] * th[2][c] * p;
                gz += th[0][a] * th[1][b] * dth[2][c] * p;
            }
        }
    }
    const double fx = -q * static_cast<double>(gx) * g.kx / g.lx;
    const double fy = -q * static_cast<double>(gy) * g.ky / g.ly;
    const double fz = -q * static_cast<double>(gz) * g.kz / g.lz;
    if constexpr (sizeof(real) == 8) {
        acc_add(force_d, 3 * i, fx);
        acc_add(force_d, 3 * i + 1, fy);
        acc_add(force_d, 3 * i + 2, fz);
    } else {
        acc_add(fxa, 3 * i, fx);
        acc_add(fxa, 3 * i + 1, fy);
        acc_add(fxa, 3 * i + 2, fz);
    }
}

// Reciprocal potential at each atom, dE_rec/dq_i = sum_k theta_i(k) phi(k)
// (the adjoint of the spreading), one thread per atom, plain stores.
template <typename real>
__global__ void k_potential(int n, const double* __restrict__ wrapped, PmeGeom g, const real* __restrict__ phi,
                            double* __restrict__ pot) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int fl[3];
    double w[3];
    frac(wrapped, i, g, fl, w);
    real th[3][4], dth[3][4];
    for (int d = 0; d < 3; ++d) bspline4<real>(static_cast<real>(w[d]), th[d], dth[d]);
    double s = 0.0;
    for (int a = 0; a < 4; ++a) {
        const int ix = wrapk(fl[0] - a, g.kx);
        for (int b = 0; b < 4; ++b) {
            const int iy = wrapk(fl[1] - b, g.ky);
            const real* row = phi + (static_cast<size_t>(ix) * g.ky + iy) * g.kz;
            real sab = 0;
            for (int c = 0; c < 4; ++c) sab += th[2][c] * row[wrapk(fl[2] - c, g.kz)];
            s += static_cast<double>(th[0][a] * th[1][b] * sab);
        }
    }
    pot[i] = s;
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
    PmeGeom g{w.pme_k[0], w.pme_k[1], w.pme_k[2], w.pme_k[2] / 2 + 1, c->grid.box[0], c->grid.box[1],
              c->grid.box[2]};
    const bool dbl = sizeof(real) == 8;
    c2* cg = static_cast<c2*>(w.cgrid);
    const c2* tw = dbl ? reinterpret_cast<const c2*>(w.twiddle_d) : reinterpret_cast<const c2*>(w.twiddle_f);
    const c2* twx = tw;
    const c2* twy = tw + g.kx;
    const c2* twz = tw + g.kx + g.ky;
    auto& T = c->timing;
    auto mark = [&](int k, bool start) {
        if (T.mode == 1) (void)hipEventRecord(c->ev[2 * k + (start ? 0 : 1)], c->stream);
    };
    const int nb = (n + kThreads - 1) / kThreads;
    mark(MDGPU_K_SPREAD, true);
    hipLaunchKernelGGL(k_spread<real>, dim3(nb), dim3(kThreads), 0, c->stream, n, w.wrapped, c->topo.charge, g,
                       w.grid_d, w.grid_fx, w.flags);
    mark(MDGPU_K_SPREAD, false);
    mark(MDGPU_K_FFT, true);
    {
        const int nl = lines_per_block(g.kz, sizeof(c2), c->fft_lines);
        const size_t shm = (2 * static_cast<size_t>(g.kz) * nl + g.kz) * sizeof(c2);
        hipLaunchKernelGGL(k_fft_z_r2c<real>, dim3((g.kx * g.ky + nl - 1) / nl), dim3(kThreads), shm, c->stream, g,
                           ilog2(nl), w.grid_d, w.grid_fx, cg, twz);
    }
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
                       c->params.ewald_alpha, w.partial, row0);
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
    if (pot)
        hipLaunchKernelGGL(k_potential<real>, dim3(nb), dim3(kThreads), 0, c->stream, n, w.wrapped, g, phi, pot);
    else
        hipLaunchKernelGGL(k_gather<real>, dim3(nb), dim3(kThreads), 0, c->stream, n, w.wrapped, c->topo.charge, g,
                           phi, w.force_d, FixedAcc{w.force_fx, fixed_scale(c->params.semantics), w.flags});
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
