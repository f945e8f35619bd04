// Device code of the bonded terms and of the AMBER excluded-pair / 1-4
// corrections, shared by the launch that runs them as extra blocks of the
// pair kernel.  Reference routines: see bonded.hip.
#ifndef MD_AMD_GPU_BONDED_TERMS_H
#define MD_AMD_GPU_BONDED_TERMS_H

#include "ctx.h"

namespace mdg {

namespace {

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
__device__ __forceinline__ R3<real> cvt(const R3<double>& v) {
    return {static_cast<real>(v.x), static_cast<real>(v.y), static_cast<real>(v.z)};
}

template <typename real>
struct BArgs {
    const double* x;
    Topo t;
    double* force_d;
    FixedAcc fx;
    double* partial;
    int ld;            // rows allocated in the slot-major partial buffer
    int row0;
    int amber;
    int virial;        // also form the virial of the bonded forces
};

// Virial of forces f_k on the atoms of one term at positions r_k relative to
// one atom of the term (sum_k r_k f_k^T, symmetric for these terms): xx, yy,
// zz, xy, xz, yz.
template <typename real>
__device__ __forceinline__ void vir_add(double* w, const R3<real>& r, const R3<real>& f) {
    w[0] += static_cast<double>(r.x * f.x);
    w[1] += static_cast<double>(r.y * f.y);
    w[2] += static_cast<double>(r.z * f.z);
    w[3] += static_cast<double>(r.x * f.y);
    w[4] += static_cast<double>(r.x * f.z);
    w[5] += static_cast<double>(r.y * f.z);
}

template <typename real>
__device__ __forceinline__ void addf(const BArgs<real>& a, int atom, real fx, real fy, real fz) {
    if constexpr (sizeof(real) == 8) {
        acc_add(a.force_d, 3 * atom, fx);
        acc_add(a.force_d, 3 * atom + 1, fy);
        acc_add(a.force_d, 3 * atom + 2, fz);
    } else {
        acc_add(a.fx, 3 * atom, static_cast<double>(fx));
        acc_add(a.fx, 3 * atom + 1, static_cast<double>(fy));
        acc_add(a.fx, 3 * atom + 2, static_cast<double>(fz));
    }
}

// Metrics::dihedral (src/System/Metrics.cpp:42-82) from the three bond vectors
// b1 = x2 - x1, b2 = x3 - x2, b3 = x4 - x3; NaN when degenerate.
template <typename real>
__device__ __forceinline__ real dihedral_angle(const R3<real>& b1, const R3<real>& b2, const R3<real>& b3) {
    const R3<real> n1 = cross(b1, b2), n2 = cross(b2, b3);
    const real m1 = dot(n1, n1), m2 = dot(n2, n2), l2 = sqrt(dot(b2, b2));
    if (m1 < static_cast<real>(1e-28) || m2 < static_cast<real>(1e-28) || l2 < static_cast<real>(1e-14))
        return static_cast<real>(NAN);
    return atan2(dot(b1, n2) * l2, dot(n1, n2));
}

// Torque tau = -dE/dphi on (a, b, c, d); ba = x_a - x_b, cb = x_b - x_c,
// dc = x_c - x_d.  REF semantics distributes it as the reference does
// (BondedForces.cpp:195-239), which is not the gradient of the torsion
// energy; AMBER semantics uses the gradient (Bekker's form):
// F_a = -fa0, F_d = -fd0, F_b = (1 + pb) fa0 - pd fd0, F_c = (1 + pd) fd0 - pb fa0.
// Returns false (no force) when a cross product or the middle bond is
// degenerate.
template <typename real>
__device__ __forceinline__ bool torsion(const BArgs<real>& a, int ia, int ib, int ic, int id, const R3<real>& ba,
                                        const R3<real>& cb, const R3<real>& dc, real tau, bool apply,
                                        double* w) {
    const R3<real> m = cross(ba, cb), n = cross(cb, dc);
    const real m2 = dot(m, m), n2 = dot(n, n), c2 = dot(cb, cb);
    const real eps = static_cast<real>(1e-24);
    if (m2 < eps || n2 < eps || c2 < eps) return false;
    if (!apply) return true;
    const real lc = sqrt(c2);
    const real sa = tau * lc / m2, sd = -tau * lc / n2;
    const R3<real> fa = {sa * m.x, sa * m.y, sa * m.z}, fd = {sd * n.x, sd * n.y, sd * n.z};
    const real pb = dot(ba, cb) / c2, pd = dot(dc, cb) / c2;
    const real sg = a.amber ? static_cast<real>(-1) : static_cast<real>(1);
    const real cbf = a.amber ? pb + 1 : pb - 1, cdf = a.amber ? pd + 1 : pd - 1;
    const R3<real> Fa = {sg * fa.x, sg * fa.y, sg * fa.z}, Fd = {sg * fd.x, sg * fd.y, sg * fd.z};
    const R3<real> Fb = {cbf * fa.x - pd * fd.x, cbf * fa.y - pd * fd.y, cbf * fa.z - pd * fd.z};
    const R3<real> Fc = {cdf * fd.x - pb * fa.x, cdf * fd.y - pb * fa.y, cdf * fd.z - pb * fa.z};
    addf(a, ia, Fa.x, Fa.y, Fa.z);
    addf(a, id, Fd.x, Fd.y, Fd.z);
    addf(a, ib, Fb.x, Fb.y, Fb.z);
    addf(a, ic, Fc.x, Fc.y, Fc.z);
    if (a.virial) {
        // positions relative to atom b: x_a - x_b = ba, x_c - x_b = -cb, x_d - x_b = -(cb + dc)
        vir_add(w, ba, Fa);
        vir_add(w, R3<real>{-cb.x, -cb.y, -cb.z}, Fc);
        vir_add(w, R3<real>{-cb.x - dc.x, -cb.y - dc.y, -cb.z - dc.z}, Fd);
    }
    return true;
}

template <typename real>
__device__ __forceinline__ real wrap_pi(real d) {
    const real pi = static_cast<real>(kPi), two = static_cast<real>(2 * kPi);
    // the reference's loops (CMapGroup.cpp:88-91); bounded so that a
    // non-finite angle cannot keep a thread spinning
    for (int it = 0; it < 8 && d <= -pi; ++it) d += two;
    for (int it = 0; it < 8 && d > pi; ++it) d -= two;
    return d;
}

// Term k of the concatenated lists; the block's energies go to row `row`.
template <typename real>
__device__ __forceinline__ void bonded_terms(const BArgs<real>& a, int k, int row) {
    const Topo& t = a.t;
    const double* x = a.x;
    const int nb = t.n_bonds, nu = nb + t.n_ubs, na = nu + t.n_angles, nd = na + t.n_dihedrals,
              ni = nd + t.n_impropers, nc = ni + t.n_cmaps;
    double e[12] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};  // bond, angle, dihedral, UB, improper, CMAP, virial
    double* w = e + 6;
    // The geometry of a term (bond length, angle, torsion), its difference
    // from the equilibrium and the energy are formed in FP64 in both forms:
    // in FP32, r - r0 keeps only about 6e-8 r / (r - r0) of relative
    // precision.  The mixed form assembles the force vectors in FP32.
    if (k < nu) {
        // harmonic bond / Urey-Bradley: E = k (r - r0)^2, F_a = -2k (r - r0) d / r
        const bool ub = k >= nb;
        const int q = ub ? k - nb : k;
        const int2 at = ub ? t.ub_atoms[q] : t.bond_atoms[q];
        const double2 par = ub ? t.ub_par[t.ub_type[q]] : t.bond_par[t.bond_type[q]];
        const R3<double> dd = diff<double>(x, at.x, at.y);
        const double r = sqrt(dot(dd, dd));
        const double dr = r - par.y;
        e[ub ? 3 : 0] = par.x * dr * dr;
        if (r >= 1e-12) {
            const real s = static_cast<real>(-2.0 * par.x * dr / r);
            const R3<real> d = cvt<real>(dd);
            addf(a, at.x, s * d.x, s * d.y, s * d.z);
            addf(a, at.y, -s * d.x, -s * d.y, -s * d.z);
            if (a.virial) vir_add(w, d, R3<real>{s * d.x, s * d.y, s * d.z});
        }
    } else if (k < na) {
        // harmonic angle (BondedEnergy.cpp:30-49, BondedForces.cpp:73-153)
        const int q = k - nu;
        const int4 at = t.angle_atoms[q];
        const double2 par = t.angle_par[t.angle_type[q]];
        const R3<double> bad = diff<double>(x, at.x, at.y), bcd = diff<double>(x, at.z, at.y);
        const double ab2d = dot(bad, bad), bc2d = dot(bcd, bcd);
        double th;
        if (ab2d < 1e-14 || bc2d < 1e-14) {
            th = NAN;
        } else {
            double cs = dot(bad, bcd) * (1.0 / sqrt(ab2d * bc2d));
            cs = cs < -1 ? -1.0 : (cs > 1 ? 1.0 : cs);
            th = acos(cs);
        }
        const double dth = th - par.y;
        e[1] = par.x * dth * dth;
        const R3<real> ba = cvt<real>(bad), bc = cvt<real>(bcd);
        const real ab2 = dot(ba, ba), bc2 = dot(bc, bc), dt = dot(ba, bc);
        const real ab = sqrt(ab2), cb = sqrt(bc2);
        R3<real> pa = {dt * ba.x - ab2 * bc.x, dt * ba.y - ab2 * bc.y, dt * ba.z - ab2 * bc.z};
        R3<real> pc = {dt * bc.x - bc2 * ba.x, dt * bc.y - bc2 * ba.y, dt * bc.z - bc2 * ba.z};
        const real mpa = sqrt(dot(pa, pa)), mpc = sqrt(dot(pc, pc));
        const real tiny = static_cast<real>(1e-12);
        if (ab >= tiny && mpa >= tiny && cb >= tiny && mpc >= tiny) {
            const real fm = static_cast<real>(-2.0 * par.x * dth);
            const real sa = fm / ab / mpa, sc = fm / cb / mpc;
            const R3<real> fa = {sa * pa.x, sa * pa.y, sa * pa.z}, fc = {sc * pc.x, sc * pc.y, sc * pc.z};
            addf(a, at.x, fa.x, fa.y, fa.z);
            addf(a, at.z, fc.x, fc.y, fc.z);
            addf(a, at.y, -fa.x - fc.x, -fa.y - fc.y, -fa.z - fc.z);
            if (a.virial) {
                vir_add(w, ba, fa);
                vir_add(w, bc, fc);
            }
        }
    } else if (k < ni) {
        // cosine dihedral (k < nd) or harmonic improper
        const bool imp = k >= nd;
        const int q = imp ? k - nd : k - na;
        const int4 at = imp ? t.imp_atoms[q] : t.dih_atoms[q];
        const R3<double> b1 = diff<double>(x, at.y, at.x), b2 = diff<double>(x, at.z, at.y),
                         b3 = diff<double>(x, at.w, at.z);
        const double phi = dihedral_angle(b1, b2, b3);
        real tau;
        if (!imp) {
            const double4 par = t.dih_par[t.dih_type[q]];
            const double arg = par.y * phi - par.z;
            e[2] = par.x * (1 + cos(arg));
            tau = static_cast<real>(par.x * par.y * sin(arg));
        } else {
            const double2 par = t.imp_par[t.imp_type[q]];
            const double dl = phi - par.y;
            const double dw = wrap_pi(dl);
            e[4] = par.x * (a.amber ? dw * dw : dl * dl);
            tau = static_cast<real>(-2.0 * par.x * dw);
        }
        const R3<real> ba = cvt<real>({-b1.x, -b1.y, -b1.z}), cb = cvt<real>({-b2.x, -b2.y, -b2.z}),
                       dc = cvt<real>({-b3.x, -b3.y, -b3.z});
        torsion(a, at.x, at.y, at.z, at.w, ba, cb, dc, tau, true, w);
    } else if (k < nc) {
        // CMAP (BondedEnergy.cpp:124-166, CMapGroup.cpp:82-237,
        // BondedForces.cpp:334-479); the patch from the FP64 coefficients
        const int q = k - ni;
        const int* at = t.cmap_atoms + 5 * q;
        const R3<double> b1 = diff<double>(x, at[1], at[0]), b2 = diff<double>(x, at[2], at[1]),
                         b3 = diff<double>(x, at[3], at[2]), b4 = diff<double>(x, at[4], at[3]);
        const double phi = dihedral_angle(b1, b2, b3), psi = dihedral_angle(b2, b3, b4);
        const int res = t.cmap_res;
        const double sp = (2 * kPi) / static_cast<double>(res);
        const double pp = wrap_pi(phi) + kPi, ss = wrap_pi(psi) + kPi;
        const double gp = pp / sp, gs = ss / sp;
        int row = static_cast<int>(gp), col = static_cast<int>(gs);
        const double u = gp - static_cast<double>(row), v = gs - static_cast<double>(col);
        row %= res;
        col %= res;
        if (row < 0) row += res;
        if (col < 0) col += res;
        double en = 0, du = 0, dv = 0;
        if (phi == phi && psi == psi) {
            const double* c = t.cmap_coeff + (static_cast<size_t>(t.cmap_set[q] - 1) * res * res + row * res + col) * 16;
            const double up[4] = {1, u, u * u, u * u * u}, vp[4] = {1, v, v * v, v * v * v};
            const double dup[4] = {0, 1, 2 * u, 3 * u * u}, dvp[4] = {0, 1, 2 * v, 3 * v * v};
#pragma unroll
            for (int i = 0; i < 4; ++i)
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    const double cc = c[4 * i + j];
                    en += cc * up[i] * vp[j];
                    du += cc * dup[i] * vp[j];
                    dv += cc * up[i] * dvp[j];
                }
        } else {
            en = NAN;
        }
        e[5] = en;
        const real dphi = static_cast<real>(du / sp), dpsi = static_cast<real>(dv / sp);
        const R3<real> ba = cvt<real>({-b1.x, -b1.y, -b1.z}), cb = cvt<real>({-b2.x, -b2.y, -b2.z}),
                       dc = cvt<real>({-b3.x, -b3.y, -b3.z}), ed = cvt<real>({-b4.x, -b4.y, -b4.z});
        // the reference skips both halves if any of the five products is degenerate
        const bool ok = torsion(a, at[0], at[1], at[2], at[3], ba, cb, dc, -dphi, false, w) &&
                        torsion(a, at[1], at[2], at[3], at[4], cb, dc, ed, -dpsi, false, w);
        if (ok) {
            torsion(a, at[0], at[1], at[2], at[3], ba, cb, dc, -dphi, true, w);
            torsion(a, at[1], at[2], at[3], at[4], cb, dc, ed, -dpsi, true, w);
        }
    }
    const int slots[12] = {E_BOND, E_ANGLE, E_DIHEDRAL, E_UB, E_IMPROPER, E_CMAP,
                           E_VSR_XX, E_VSR_YY, E_VSR_ZZ, E_VSR_XY, E_VSR_XZ, E_VSR_YZ};
    block_sum_store<12>(e, a.partial, a.ld, row, slots);
}

struct CArgs {
    const double* x;
    double lx, ly, lz;
    Cell cell;
    int n_excl;
    const int2* excl;
    int n14;
    const int2* p14;
    const double* scee;
    const double* charge;
    const int* type;
    const int* nb_index;      // [T*T] 1-based
    const double2* lj14;      // [n_lj]
    int ntypes;
    double alpha;
    int mixed;
    double* force_d;
    FixedAcc fx;
    double* partial;
    int ld;
    int row0;
    int virial;               // also form the virial of the correction forces
};

__device__ __forceinline__ void addc(const CArgs& a, int atom, double fx, double fy, double fz) {
    if (a.mixed) {
        acc_add(a.fx, 3 * atom, fx);
        acc_add(a.fx, 3 * atom + 1, fy);
        acc_add(a.fx, 3 * atom + 2, fz);
    } else {
        acc_add(a.force_d, 3 * atom, fx);
        acc_add(a.force_d, 3 * atom + 1, fy);
        acc_add(a.force_d, 3 * atom + 2, fz);
    }
}

// Pair k of the excluded pairs followed by the 1-4 pairs.
__device__ __forceinline__ void correction_pairs(const CArgs& a, int k, int row) {
    double e[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};  // ewald_excl, vdw14, ee14, virial
    if (k < a.n_excl + a.n14) {
        const bool is14 = k >= a.n_excl;
        const int2 p = is14 ? a.p14[k - a.n_excl] : a.excl[k];
        double dx = a.x[3 * p.x] - a.x[3 * p.y], dy = a.x[3 * p.x + 1] - a.x[3 * p.y + 1],
               dz = a.x[3 * p.x + 2] - a.x[3 * p.y + 2];
        if (!a.cell.tri) {
            dx -= a.lx * rint(dx / a.lx);
            dy -= a.ly * rint(dy / a.ly);
            dz -= a.lz * rint(dz / a.lz);
        } else {
            cell_min_image_tri(a.cell, dx, dy, dz);
        }
        const double r2 = dx * dx + dy * dy + dz * dz, r = sqrt(r2);
        const double qq = kCoulomb * a.charge[p.x] * a.charge[p.y];
        const double ef = erf(a.alpha * r);
        const double g2 = 1.1283791670955126 * a.alpha * exp(-a.alpha * a.alpha * r2);
        double g;
        if (!is14) {
            e[0] = -qq * ef / r;
            g = qq * (g2 - ef / r) / r2;
        } else {
            const double is = 1.0 / a.scee[k - a.n_excl];
            e[2] = qq * (is / r - ef / r);
            g = qq * (is / r + g2 - ef / r) / r2;
            const int nb = a.nb_index[a.type[p.x] * a.ntypes + a.type[p.y]];
            if (nb > 0) {
                const double2 ab = a.lj14[nb - 1];
                const double i2 = 1.0 / r2, i6 = i2 * i2 * i2;
                e[1] = (ab.x * i6 - ab.y) * i6;
                g += (12.0 * ab.x * i6 - 6.0 * ab.y) * i6 * i2;
            }
        }
        addc(a, p.x, g * dx, g * dy, g * dz);
        addc(a, p.y, -g * dx, -g * dy, -g * dz);
        if (a.virial) {
            e[3] = g * dx * dx;
            e[4] = g * dy * dy;
            e[5] = g * dz * dz;
            e[6] = g * dx * dy;
            e[7] = g * dx * dz;
            e[8] = g * dy * dz;
        }
    }
    const int slots[9] = {E_EW_EXCL, E_VDW14, E_EE14, E_VSR_XX, E_VSR_YY, E_VSR_ZZ, E_VSR_XY, E_VSR_XZ, E_VSR_YZ};
    block_sum_store<9>(e, a.partial, a.ld, row, slots);
}

// One launch for both: blocks [0, nbb) take the bonded terms, the rest the
// corrections (both latency-bound, so they overlap instead of queueing).
// Block b writes energy row row0 + b.
}  // namespace

}  // namespace mdg

#endif
