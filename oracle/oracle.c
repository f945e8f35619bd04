/* oracle.c - CPU FP64 oracle (TEST INFRASTRUCTURE ONLY; see oracle.h).
 *
 * Citations are to the reference's files (prateekbansal97/MD).
 *
 * Pinning of each function (how a plausible mistake would be caught):
 *   bonded energies      per-term golden vectors from the compiled reference
 *                        parser + formulas (oracle/_ref, tests/golden)
 *   bonded forces        central finite differences of the bonded energies
 *   LJ / Coulomb forms   golden formula samples from oracle/_ref; FD of forces
 *   REF pair list        set equality with brute-force O(N^2) enumeration
 *   REF EE (O(N^2))      equals the list-based plain Coulomb sum (no nb=0 pairs)
 *   Ewald direct + self  alpha-independence of the total Ewald energy with the
 *   + exclusion terms    structure-factor reciprocal sum (a dropped term or a
 *                        wrong sign breaks the invariance)
 *   smooth PME           convergence to the structure-factor sum in grid size;
 *                        FD of its forces; finite-strain derivative of virial
 *   B-splines            partition of unity, closed form of M_4, and
 *                        derivative = finite difference
 *   L-BFGS minimiser     a harmonic bond and a bonded chain reach their known
 *                        minima (E = 0, r = r0) with the RMS force at the
 *                        tolerance, the energy never rises, and a single
 *                        quadratic step lands on the minimum once a pair is
 *                        kept (H0 = s.y / y.y is exact for one coordinate)
 *   velocity Verlet      the closed-form discrete solution of a harmonic bond
 *                        (x_n = A cos(n theta), cos theta = 1 - (w dt)^2 / 2),
 *                        time reversal, and the dt^2 scaling of the energy
 *                        fluctuation
 *   rebuild criterion    single-atom displacements just below / at / above
 *                        skin/2, across a periodic face (minimum image), a NaN
 *                        coordinate, and agreement with an exact-rational
 *                        brute force on random displacement sets
 *   triclinic cells      cell matrix lengths / angles / face distances; right
 *                        angles reproduce the orthorhombic path bit for bit;
 *                        an orthorhombic lattice in a sheared basis gives the
 *                        orthorhombic energies, forces and structure-factor
 *                        sum; triclinic PME converges to that sum at fourth
 *                        order, its forces are its FD gradient; alpha-
 *                        independence of the total Ewald energy; a cell
 *                        doubled along a doubles every term; diagonal virial
 *                        = FD strain derivative; rebuild criterion across
 *                        lattice vectors
 */
#include "oracle.h"

#include <float.h>
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

/* CoulombicEE.cpp:12 forms 18.2206899283247^2 in long double. */
#define SQRT_KC 18.2206899283247L

double oracle_coulomb_constant(void) { return (double)(SQRT_KC * SQRT_KC); }

/* ---------------------------------------------------------------- PBC ---- */

/* PBC.cpp:72-76 */
static double min_image_1d(double d, double L) { return d - L * round(d / L); }
/* PBC.cpp:78-81 */
static double wrap1(double x, double L) { return x - L * floor(x / L); }

static void min_image(double* d, const double box[3], int pbc) {
    if (!pbc) return;
    d[0] = min_image_1d(d[0], box[0]);
    d[1] = min_image_1d(d[1], box[1]);
    d[2] = min_image_1d(d[2], box[2]);
}

/* ---------------------------------------------------- triclinic cells ---- */

/* The reference's periodic boundary is orthorhombic only (PBC.cpp:72-81).
 * A triclinic cell has edge vectors a = (ax,0,0), b = (bx,by,0), c =
 * (cx,cy,cz), built from the edge lengths and the angles alpha (b,c), beta
 * (a,c), gamma (a,b) as on an rst7 box line; H has a, b, c as columns and
 * R = H^-1 is upper triangular, so s = R x are fractional coordinates.  An
 * angle of exactly 90 degrees has cosine 0; three of them make the cell
 * orthorhombic and every routine takes its orthorhombic path. */
typedef struct {
    int tri;
    double L[3];
    double ax, bx, by, cx, cy, cz;
    double r00, r01, r02, r11, r12, r22;
    double face[3]; /* distances between opposite faces */
} ocell;

static double g_angles[3] = {90.0, 90.0, 90.0};

int oracle_set_box_angles(const double ang[3]) {
    for (int d = 0; d < 3; ++d)
        if (!(ang[d] > 0.0 && ang[d] < 180.0)) return 1;
    for (int d = 0; d < 3; ++d) g_angles[d] = ang[d];
    return 0;
}

static double cos_deg(double a) { return a == 90.0 ? 0.0 : cos(a * (M_PI / 180.0)); }

/* Returns 0 when the lengths and angles form a cell. */
static int ocell_make(ocell* c, const double box[3]) {
    memset(c, 0, sizeof *c);
    for (int d = 0; d < 3; ++d) c->L[d] = c->face[d] = box[d];
    c->tri = !(g_angles[0] == 90.0 && g_angles[1] == 90.0 && g_angles[2] == 90.0);
    if (!c->tri) return 0;
    const double ca = cos_deg(g_angles[0]), cb = cos_deg(g_angles[1]), cg = cos_deg(g_angles[2]);
    const double sg = g_angles[2] == 90.0 ? 1.0 : sin(g_angles[2] * (M_PI / 180.0));
    if (!(sg > 1e-6)) return 1;
    c->ax = box[0];
    c->bx = box[1] * cg;
    c->by = box[1] * sg;
    c->cx = box[2] * cb;
    c->cy = box[2] * (ca - cb * cg) / sg;
    const double cz2 = box[2] * box[2] - c->cx * c->cx - c->cy * c->cy;
    if (!(cz2 > 1e-12 * box[2] * box[2])) return 1;
    c->cz = sqrt(cz2);
    c->r00 = 1.0 / c->ax;
    c->r11 = 1.0 / c->by;
    c->r22 = 1.0 / c->cz;
    c->r01 = -c->bx / (c->ax * c->by);
    c->r12 = -c->cy / (c->by * c->cz);
    c->r02 = (c->bx * c->cy - c->by * c->cx) / (c->ax * c->by * c->cz);
    c->face[0] = 1.0 / sqrt(c->r00 * c->r00 + c->r01 * c->r01 + c->r02 * c->r02);
    c->face[1] = 1.0 / sqrt(c->r11 * c->r11 + c->r12 * c->r12);
    c->face[2] = c->cz;
    return 0;
}

int oracle_cell_matrix(const double box[3], double h[9], double face[3]) {
    ocell c;
    if (ocell_make(&c, box)) return 1;
    const double m[9] = {c.tri ? c.ax : c.L[0], c.bx, c.cx, 0.0, c.tri ? c.by : c.L[1], c.cy, 0.0, 0.0,
                         c.tri ? c.cz : c.L[2]};
    for (int k = 0; k < 9; ++k) h[k] = m[k];
    for (int d = 0; d < 3; ++d) face[d] = c.face[d];
    return 0;
}

static double ocell_volume(const ocell* c) { return c->tri ? c->ax * c->by * c->cz : c->L[0] * c->L[1] * c->L[2]; }

static void ocell_frac(const ocell* c, const double* x, double* s) {
    s[2] = c->r22 * x[2];
    s[1] = c->r11 * x[1] + c->r12 * x[2];
    s[0] = c->r00 * x[0] + c->r01 * x[1] + c->r02 * x[2];
}

/* Minimum image d - H rint(R d) (triclinic) or the reference's per-axis
 * rounding (orthorhombic).  With the cutoff below half the smallest face
 * distance no other image of a pair can lie within it. */
static void ocell_min_image(const ocell* c, double* d) {
    if (!c->tri) {
        min_image(d, c->L, 1);
        return;
    }
    double s[3];
    ocell_frac(c, d, s);
    const double n0 = rint(s[0]), n1 = rint(s[1]), n2 = rint(s[2]);
    d[0] = d[0] - (c->ax * n0 + c->bx * n1 + c->cx * n2);
    d[1] = d[1] - (c->by * n1 + c->cy * n2);
    d[2] = d[2] - c->cz * n2;
}

/* Reciprocal vector m R (Cartesian) of the integer vector m. */
static void ocell_recip(const ocell* c, double ma, double mb, double mc, double* m) {
    if (!c->tri) {
        m[0] = ma / c->L[0];
        m[1] = mb / c->L[1];
        m[2] = mc / c->L[2];
        return;
    }
    m[0] = ma * c->r00;
    m[1] = ma * c->r01 + mb * c->r11;
    m[2] = ma * c->r02 + mb * c->r12 + mc * c->r22;
}

/* ------------------------------------------------------- Verlet list ---- */

/* setup_neighbor_rebuild_threshold_lj (NonBondedUtils.cpp:65-69) sets the
 * limit to (skin/2)^2; ljpairs_needs_rebuild (:87-100) keeps the running
 * maximum of the squared minimum-image displacements and answers 1 as soon
 * as it exceeds the limit. */
int oracle_pairs_need_rebuild(int n, const double* x, const double* xref, const double box[3], double skin) {
    const double half = 0.5 * skin, lim2 = half * half;
    double max_dr2 = 0.0;
    ocell cell;
    if (ocell_make(&cell, box)) return -1;
    for (int i = 0; i < n; ++i) {
        double d[3] = {x[3 * i] - xref[3 * i], x[3 * i + 1] - xref[3 * i + 1], x[3 * i + 2] - xref[3 * i + 2]};
        ocell_min_image(&cell, d);
        const double dr2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
        if (dr2 > max_dr2) max_dr2 = dr2;
        if (max_dr2 > lim2) return 1;
    }
    return 0;
}

/* ------------------------------------------------------------- metrics ---- */

/* Metrics.cpp:12-18 */
static double m_distance(const double* a, const double* b) {
    double dx = a[0] - b[0], dy = a[1] - b[1], dz = a[2] - b[2];
    return sqrt(dx * dx + dy * dy + dz * dz);
}

/* Metrics.cpp:20-40 */
static double m_angle(const double* a, const double* b, const double* c) {
    double d1[3] = {a[0] - b[0], a[1] - b[1], a[2] - b[2]};
    double d2[3] = {c[0] - b[0], c[1] - b[1], c[2] - b[2]};
    double dot = d1[0] * d2[0] + d1[1] * d2[1] + d1[2] * d2[2];
    double m1 = d1[0] * d1[0] + d1[1] * d1[1] + d1[2] * d1[2];
    double m2 = d2[0] * d2[0] + d2[1] * d2[1] + d2[2] * d2[2];
    if (m1 < 1e-14 || m2 < 1e-14) return NAN;
    double cs = dot * (1.0 / sqrt(m1 * m2));
    if (cs < -1.0) cs = -1.0;
    if (cs > 1.0) cs = 1.0;
    return acos(cs);
}

/* Metrics.cpp:42-82 */
static double m_dihedral(const double* p1, const double* p2, const double* p3, const double* p4) {
    double b1[3] = {p2[0] - p1[0], p2[1] - p1[1], p2[2] - p1[2]};
    double b2[3] = {p3[0] - p2[0], p3[1] - p2[1], p3[2] - p2[2]};
    double b3[3] = {p4[0] - p3[0], p4[1] - p3[1], p4[2] - p3[2]};
    double n1[3] = {b1[1] * b2[2] - b1[2] * b2[1], b2[0] * b1[2] - b1[0] * b2[2], b1[0] * b2[1] - b1[1] * b2[0]};
    double n2[3] = {b2[1] * b3[2] - b2[2] * b3[1], b3[0] * b2[2] - b2[0] * b3[2], b2[0] * b3[1] - b2[1] * b3[0]};
    double mn1 = n1[0] * n1[0] + n1[1] * n1[1] + n1[2] * n1[2];
    double mn2 = n2[0] * n2[0] + n2[1] * n2[1] + n2[2] * n2[2];
    double lb2 = sqrt(b2[0] * b2[0] + b2[1] * b2[1] + b2[2] * b2[2]);
    if (mn1 < 1e-28 || mn2 < 1e-28 || lb2 < 1e-14) return NAN;
    double xx = n1[0] * n2[0] + n1[1] * n2[1] + n1[2] * n2[2];
    double yy = (b1[0] * n2[0] + b1[1] * n2[1] + b1[2] * n2[2]) * lb2;
    return atan2(yy, xx);
}

/* ------------------------------------------------------------- bonded ----- */

static void addf(double* f, int i, double x, double y, double z) {
    f[3 * i] += x;
    f[3 * i + 1] += y;
    f[3 * i + 2] += z;
}

/* Pair virial sum g d_a d_b of the pair forces (full-virial evaluations):
 * the pair routines add into it when it is set. */
static double* g_wpair = NULL;

/* Force g d on a and -g d on b (d = x_a - x_b at its minimum image). */
static void pair_force(double* f, int a, int b, double g, const double* d) {
    addf(f, a, g * d[0], g * d[1], g * d[2]);
    addf(f, b, -g * d[0], -g * d[1], -g * d[2]);
    if (g_wpair)
        for (int u = 0; u < 3; ++u)
            for (int v = 0; v < 3; ++v) g_wpair[3 * u + v] += g * d[u] * d[v];
}

/* BondedEnergy.cpp:10-28 and BondedForces.cpp:10-39 (bonds); the same form
 * for Urey-Bradley, BondedEnergy.cpp:74-95 / BondedForces.cpp:41-71. */
static double harmonic_pairs(int n, const int32_t* at, const int32_t* ty, const double* k, const double* r0,
                             const double* x, double* f) {
    double e = 0.0;
    for (int t = 0; t < n; ++t) {
        int a = at[2 * t], b = at[2 * t + 1], p = ty[t];
        const double* xa = x + 3 * a;
        const double* xb = x + 3 * b;
        double d = m_distance(xa, xb);
        e += k[p] * (d - r0[p]) * (d - r0[p]);
        if (!f) continue;
        double dx = xa[0] - xb[0], dy = xa[1] - xb[1], dz = xa[2] - xb[2];
        double r = sqrt(dx * dx + dy * dy + dz * dz);
        if (r < 1e-12) continue;
        double fm = -2 * k[p] * (r - r0[p]);
        double fx = fm * dx / r, fy = fm * dy / r, fz = fm * dz / r;
        addf(f, a, fx, fy, fz);
        addf(f, b, -fx, -fy, -fz);
    }
    return e;
}

/* BondedEnergy.cpp:30-49 and BondedForces.cpp:73-153 */
static double angles(const mdgpu_system_desc* s, const double* x, double* f) {
    double e = 0.0;
    for (int t = 0; t < s->n_angles; ++t) {
        int a = s->angle_atoms[3 * t], b = s->angle_atoms[3 * t + 1], c = s->angle_atoms[3 * t + 2];
        int p = s->angle_type[t];
        double k = s->angle_k[p], t0 = s->angle_theta0[p];
        const double *xa = x + 3 * a, *xb = x + 3 * b, *xc = x + 3 * c;
        double th = m_angle(xa, xb, xc);
        e += k * (th - t0) * (th - t0);
        if (!f) continue;
        double ba[3] = {xa[0] - xb[0], xa[1] - xb[1], xa[2] - xb[2]};
        double bc[3] = {xc[0] - xb[0], xc[1] - xb[1], xc[2] - xb[2]};
        double ab2 = ba[0] * ba[0] + ba[1] * ba[1] + ba[2] * ba[2];
        double ab = sqrt(ab2);
        if (ab < 1e-12) continue;
        double dot = ba[0] * bc[0] + ba[1] * bc[1] + ba[2] * bc[2];
        double pa[3] = {dot * ba[0] - ab2 * bc[0], dot * ba[1] - ab2 * bc[1], dot * ba[2] - ab2 * bc[2]};
        double mpa = sqrt(pa[0] * pa[0] + pa[1] * pa[1] + pa[2] * pa[2]);
        if (mpa < 1e-12) continue;
        pa[0] /= mpa; pa[1] /= mpa; pa[2] /= mpa;
        double bc2 = bc[0] * bc[0] + bc[1] * bc[1] + bc[2] * bc[2];
        double cb = sqrt(bc2);
        if (cb < 1e-12) continue;
        double pc[3] = {dot * bc[0] - bc2 * ba[0], dot * bc[1] - bc2 * ba[1], dot * bc[2] - bc2 * ba[2]};
        double mpc = sqrt(pc[0] * pc[0] + pc[1] * pc[1] + pc[2] * pc[2]);
        if (mpc < 1e-12) continue;
        pc[0] /= mpc; pc[1] /= mpc; pc[2] /= mpc;
        double fm = -2 * k * (th - t0);
        double fa[3] = {fm / ab * pa[0], fm / ab * pa[1], fm / ab * pa[2]};
        double fc[3] = {fm / cb * pc[0], fm / cb * pc[1], fm / cb * pc[2]};
        addf(f, a, fa[0], fa[1], fa[2]);
        addf(f, c, fc[0], fc[1], fc[2]);
        addf(f, b, -fa[0] - fc[0], -fa[1] - fc[1], -fa[2] - fc[2]);
    }
    return e;
}

/* Torsion force distribution shared by the cosine dihedral, the improper and
 * each CMAP angle, for the torque tau = -dE/dphi on (a, b, c, d).
 * REF (amber = 0): the reference's distribution as written (BondedForces.cpp:
 * 173-239, 261-330, 372-477).  It is not -grad E: relative to the gradient its
 * forces on a and d have the opposite sign and b, c carry an extra 2 F_a,
 * 2 F_d (checked by finite differences in tests/test_oracle.py).
 * AMBER (amber = 1): the gradient, F_a = -fa0, F_d = -fd0,
 * F_b = (1 + pb) fa0 - pd fd0, F_c = (1 + pd) fd0 - pb fa0 (Bekker's form). */
static void torsion_forces_sem(const double* xa, const double* xb, const double* xc, const double* xd, double tau,
                               double* fa, double* fb, double* fc, double* fd, int* ok, int amber) {
    double ba[3] = {xa[0] - xb[0], xa[1] - xb[1], xa[2] - xb[2]};
    double cb[3] = {xb[0] - xc[0], xb[1] - xc[1], xb[2] - xc[2]};
    double dc[3] = {xc[0] - xd[0], xc[1] - xd[1], xc[2] - xd[2]};
    double m[3] = {ba[1] * cb[2] - ba[2] * cb[1], ba[2] * cb[0] - ba[0] * cb[2], ba[0] * cb[1] - ba[1] * cb[0]};
    double n[3] = {cb[1] * dc[2] - cb[2] * dc[1], cb[2] * dc[0] - cb[0] * dc[2], cb[0] * dc[1] - cb[1] * dc[0]};
    double m2 = m[0] * m[0] + m[1] * m[1] + m[2] * m[2];
    double n2 = n[0] * n[0] + n[1] * n[1] + n[2] * n[2];
    double cb2 = cb[0] * cb[0] + cb[1] * cb[1] + cb[2] * cb[2];
    *ok = !(m2 < 1e-24 || n2 < 1e-24 || cb2 < 1e-24);
    if (!*ok) return;
    double lcb = sqrt(cb2);
    double dba = ba[0] * cb[0] + ba[1] * cb[1] + ba[2] * cb[2];
    double ddc = dc[0] * cb[0] + dc[1] * cb[1] + dc[2] * cb[2];
    for (int q = 0; q < 3; ++q) {
        fa[q] = tau * lcb / m2 * m[q];
        fd[q] = -tau * lcb / n2 * n[q];
    }
    for (int q = 0; q < 3; ++q) {
        fb[q] = (dba / cb2 - 1) * fa[q] - ddc / cb2 * fd[q];
        fc[q] = (ddc / cb2 - 1) * fd[q] - dba / cb2 * fa[q];
    }
    if (amber)
        for (int q = 0; q < 3; ++q) {
            fb[q] += 2.0 * fa[q];
            fc[q] += 2.0 * fd[q];
            fa[q] = -fa[q];
            fd[q] = -fd[q];
        }
}

static int g_amber_torsion = 0;  /* set per evaluation by oracle_eval */

static void torsion_forces(const double* xa, const double* xb, const double* xc, const double* xd, double tau,
                           double* fa, double* fb, double* fc, double* fd, int* ok) {
    torsion_forces_sem(xa, xb, xc, xd, tau, fa, fb, fc, fd, ok, g_amber_torsion);
}

/* BondedEnergy.cpp:51-72 and BondedForces.cpp:155-241 */
static double dihedrals(const mdgpu_system_desc* s, const double* x, double* f) {
    double e = 0.0;
    for (int t = 0; t < s->n_dihedrals; ++t) {
        const int32_t* at = s->dihedral_atoms + 4 * t;
        int p = s->dihedral_type[t];
        double k = s->dihedral_k[p], nn = s->dihedral_n[p], ph = s->dihedral_phase[p];
        const double *xa = x + 3 * at[0], *xb = x + 3 * at[1], *xc = x + 3 * at[2], *xd = x + 3 * at[3];
        double phi = m_dihedral(xa, xb, xc, xd);
        e += k * (1.0 + cos(nn * phi - ph));
        if (!f) continue;
        double tau = k * nn * sin(nn * phi - ph);
        double fa[3], fb[3], fc[3], fd[3];
        int ok;
        torsion_forces(xa, xb, xc, xd, tau, fa, fb, fc, fd, &ok);
        if (!ok) continue;
        addf(f, at[0], fa[0], fa[1], fa[2]);
        addf(f, at[1], fb[0], fb[1], fb[2]);
        addf(f, at[2], fc[0], fc[1], fc[2]);
        addf(f, at[3], fd[0], fd[1], fd[2]);
    }
    return e;
}

static double wrap_pi(double d) {
    while (d <= -M_PI) d += 2.0 * M_PI;
    while (d > M_PI) d -= 2.0 * M_PI;
    return d;
}

/* BondedEnergy.cpp:97-122 (energy of the unwrapped difference in REF) and
 * BondedForces.cpp:243-332 (force of the wrapped difference). */
static double impropers(const mdgpu_system_desc* s, const double* x, double* f, int amber) {
    double e = 0.0;
    for (int t = 0; t < s->n_impropers; ++t) {
        const int32_t* at = s->improper_atoms + 4 * t;
        int p = s->improper_type[t];
        double k = s->improper_k[p], p0 = s->improper_psi0[p];
        const double *xa = x + 3 * at[0], *xb = x + 3 * at[1], *xc = x + 3 * at[2], *xd = x + 3 * at[3];
        double psi = m_dihedral(xa, xb, xc, xd);
        double de = amber ? wrap_pi(psi - p0) : psi - p0;
        e += k * de * de;
        if (!f) continue;
        double tau = -2.0 * k * wrap_pi(psi - p0);
        double fa[3], fb[3], fc[3], fd[3];
        int ok;
        torsion_forces(xa, xb, xc, xd, tau, fa, fb, fc, fd, &ok);
        if (!ok) continue;
        addf(f, at[0], fa[0], fa[1], fa[2]);
        addf(f, at[1], fb[0], fb[1], fb[2]);
        addf(f, at[2], fc[0], fc[1], fc[2]);
        addf(f, at[3], fd[0], fd[1], fd[2]);
    }
    return e;
}

/* CMapGroup.cpp:82-237: cell lookup and bicubic patch; returns E and
 * (dE/dphi, dE/dpsi). */
static double cmap_patch(const double* coeff, int set, int res, double phi, double psi, double* dphi,
                         double* dpsi) {
    phi = wrap_pi(phi) + M_PI;
    psi = wrap_pi(psi) + M_PI;
    double sp = (2.0 * M_PI) / res;
    double gp = phi / sp, gs = psi / sp;
    int row = (int)gp, col = (int)gs;
    double u = gp - (double)row, v = gs - (double)col;
    row %= res;
    col %= res;
    if (row < 0) row += res;
    if (col < 0) col += res;
    const double* a = coeff + (size_t)(set - 1) * res * res * 16 + (size_t)(row * res + col) * 16;
    double up[4] = {1.0, u, u * u, u * u * u}, vp[4] = {1.0, v, v * v, v * v * v};
    double dup[4] = {0.0, 1.0, 2.0 * u, 3.0 * u * u}, dvp[4] = {0.0, 1.0, 2.0 * v, 3.0 * v * v};
    double e = 0.0, du = 0.0, dv = 0.0;
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j) {
            double c = a[4 * i + j];
            e += c * up[i] * vp[j];
            du += c * dup[i] * vp[j];
            dv += c * up[i] * dvp[j];
        }
    *dphi = du * (1.0 / sp);
    *dpsi = dv * (1.0 / sp);
    return e;
}

/* BondedEnergy.cpp:124-166 and BondedForces.cpp:334-479 */
static double cmaps(const mdgpu_system_desc* s, const double* x, double* f) {
    double e = 0.0;
    for (int t = 0; t < s->n_cmaps; ++t) {
        const int32_t* at = s->cmap_atoms + 5 * t;
        const double *x1 = x + 3 * at[0], *x2 = x + 3 * at[1], *x3 = x + 3 * at[2], *x4 = x + 3 * at[3],
                     *x5 = x + 3 * at[4];
        double phi = m_dihedral(x1, x2, x3, x4);
        double psi = m_dihedral(x2, x3, x4, x5);
        double dphi, dpsi;
        e += cmap_patch(s->cmap_coeff, s->cmap_set[t], s->cmap_resolution, phi, psi, &dphi, &dpsi);
        if (!f) continue;
        /* the reference skips a CMAP force if any of its five cross products or
         * two bond vectors is degenerate (BondedForces.cpp:409-422) */
        double fa[3], fb[3], fc[3], fd[3], gb[3], gc[3], gd[3], ge[3];
        int ok1, ok2;
        torsion_forces(x1, x2, x3, x4, -dphi, fa, fb, fc, fd, &ok1);
        torsion_forces(x2, x3, x4, x5, -dpsi, gb, gc, gd, ge, &ok2);
        if (!ok1 || !ok2) continue;
        addf(f, at[0], fa[0], fa[1], fa[2]);
        addf(f, at[1], fb[0] + gb[0], fb[1] + gb[1], fb[2] + gb[2]);
        addf(f, at[2], fc[0] + gc[0], fc[1] + gc[1], fc[2] + gc[2]);
        addf(f, at[3], fd[0] + gd[0], fd[1] + gd[1], fd[2] + gd[2]);
        addf(f, at[4], ge[0], ge[1], ge[2]);
    }
    return e;
}

/* ------------------------------------------------------- exclusions ------ */

typedef struct {
    uint64_t* keys;
    double* vals;
    size_t cap;
} pairmap;

static uint64_t pkey(int a, int b) {
    if (a > b) { int t = a; a = b; b = t; }
    return ((uint64_t)(uint32_t)a << 32) | (uint32_t)b;
}
static size_t phash(uint64_t k, size_t cap) { return (size_t)((k * 0x9E3779B97F4A7C15ULL) >> 17) & (cap - 1); }

static void pm_init(pairmap* m, size_t n) {
    size_t cap = 16;
    while (cap < 2 * n + 16) cap <<= 1;
    m->cap = cap;
    m->keys = (uint64_t*)malloc(cap * sizeof(uint64_t));
    m->vals = (double*)malloc(cap * sizeof(double));
    for (size_t i = 0; i < cap; ++i) m->keys[i] = UINT64_MAX;
}
static void pm_put(pairmap* m, uint64_t k, double v) {
    size_t h = phash(k, m->cap);
    while (m->keys[h] != UINT64_MAX && m->keys[h] != k) h = (h + 1) & (m->cap - 1);
    m->keys[h] = k;
    m->vals[h] = v;
}
static int pm_get(const pairmap* m, uint64_t k, double* v) {
    size_t h = phash(k, m->cap);
    while (m->keys[h] != UINT64_MAX) {
        if (m->keys[h] == k) {
            if (v) *v = m->vals[h];
            return 1;
        }
        h = (h + 1) & (m->cap - 1);
    }
    return 0;
}
static void pm_free(pairmap* m) { free(m->keys); free(m->vals); }

static int excluded(const mdgpu_system_desc* s, int i, int j) {
    if (!s->excl_start) return 0;
    int lo = s->excl_start[i], hi = s->excl_start[i + 1] - 1;
    while (lo <= hi) {
        int mid = (lo + hi) / 2, v = s->excl_list[mid];
        if (v == j) return 1;
        if (v < j) lo = mid + 1; else hi = mid - 1;
    }
    return 0;
}

/* --------------------------------------------------------- LJ / switch --- */

/* LennardJones.cpp:9-27 */
static double lj_u(double r2, double a, double b) {
    double i2 = 1.0 / r2, i6 = i2 * i2 * i2, i12 = i6 * i6;
    return a * i12 - b * i6;
}
static double lj_g(double r2, double a, double b) {
    double i2 = 1.0 / r2, i4 = i2 * i2, i6 = i4 * i2, i8 = i4 * i4, i12 = i6 * i6, i14 = i12 * i2;
    return 12.0 * a * i14 - 6.0 * b * i8;
}

/* NonBondedUtils.cpp:117-132 */
static void lj_switch(double r, double rs, double rc, int use, double* S, double* dS) {
    if (!use || r <= rs) { *S = 1.0; *dS = 0.0; return; }
    if (r >= rc) { *S = 0.0; *dS = 0.0; return; }
    double inv = 1.0 / (rc - rs), xx = (r - rs) * inv;
    double x2 = xx * xx, x3 = x2 * xx, x4 = x2 * x2, x5 = x4 * xx;
    *S = 1.0 - 10.0 * x3 + 15.0 * x4 - 6.0 * x5;
    *dS = (-30.0 * x2 + 60.0 * x3 - 30.0 * x4) * inv;
}

/* --------------------------------------------------------- REF list ------ */

/* PBC.cpp:32-39 */
static int cell_of(double x, double L, int n) {
    int ix = (int)(x * n / L);
    if (ix >= n) ix = n - 1;
    if (ix < 0) ix = 0;
    return ix;
}

static int cmp_int(const void* a, const void* b) { return (*(const int*)a > *(const int*)b) - (*(const int*)a < *(const int*)b); }

/* NonBondedUtils.cpp:185-282: linked-list cells of size >= list cutoff, the 27
 * (deduplicated) neighbour cells with cell2 >= cell1, j > i inside a cell,
 * excluded non-1-4 pairs skipped, keep r^2 <= rl^2, r^2 > 1e-12, nb != 0. */
long oracle_ref_pairlist(const mdgpu_system_desc* s, const double* x, const double box[3], double rl, int* out,
                         long cap) {
    int n = s->n_atoms;
    if (!(box[0] > 0 && box[1] > 0 && box[2] > 0)) return 0; /* PBC off: no list (:187) */
    int nc[3];
    for (int d = 0; d < 3; ++d) {
        nc[d] = (int)floor(box[d] / rl);
        if (nc[d] < 1) nc[d] = 1;
    }
    int ncell = nc[0] * nc[1] * nc[2];
    int* head = (int*)malloc(sizeof(int) * ncell);
    int* next = (int*)malloc(sizeof(int) * n);
    for (int c = 0; c < ncell; ++c) head[c] = -1;
    for (int i = 0; i < n; ++i) {
        int ix = cell_of(wrap1(x[3 * i], box[0]), box[0], nc[0]);
        int iy = cell_of(wrap1(x[3 * i + 1], box[1]), box[1], nc[1]);
        int iz = cell_of(wrap1(x[3 * i + 2], box[2]), box[2], nc[2]);
        int c = ix + nc[0] * (iy + nc[1] * iz);
        next[i] = head[c];
        head[c] = i;
    }
    pairmap p14;
    pm_init(&p14, (size_t)s->n_pairs14);
    for (int t = 0; t < s->n_pairs14; ++t) pm_put(&p14, pkey(s->pairs14[2 * t], s->pairs14[2 * t + 1]), 1.0);
    double rl2 = rl * rl;
    long np = 0;
    for (int iz = 0; iz < nc[2]; ++iz)
        for (int iy = 0; iy < nc[1]; ++iy)
            for (int ix = 0; ix < nc[0]; ++ix) {
                int c1 = ix + nc[0] * (iy + nc[1] * iz);
                int neigh[27], m = 0;
                for (int dz = -1; dz <= 1; ++dz)
                    for (int dy = -1; dy <= 1; ++dy)
                        for (int dx = -1; dx <= 1; ++dx) {
                            int jx = (ix + dx + nc[0]) % nc[0], jy = (iy + dy + nc[1]) % nc[1],
                                jz = (iz + dz + nc[2]) % nc[2];
                            neigh[m++] = jx + nc[0] * (jy + nc[1] * jz);
                        }
                qsort(neigh, 27, sizeof(int), cmp_int);
                int u = 0;
                for (int q = 0; q < 27; ++q)
                    if (q == 0 || neigh[q] != neigh[q - 1]) neigh[u++] = neigh[q];
                for (int t = 0; t < u; ++t) {
                    int c2 = neigh[t];
                    if (c2 < c1) continue;
                    for (int i = head[c1]; i != -1; i = next[i])
                        for (int j = head[c2]; j != -1; j = next[j]) {
                            if (c1 == c2 && j <= i) continue;
                            int a = i < j ? i : j, b = i < j ? j : i;
                            int is14 = 0;
                            if (excluded(s, a, b)) {
                                is14 = pm_get(&p14, pkey(a, b), NULL);
                                if (!is14) continue;
                            }
                            double d[3] = {x[3 * a] - x[3 * b], x[3 * a + 1] - x[3 * b + 1], x[3 * a + 2] - x[3 * b + 2]};
                            min_image(d, box, 1);
                            double r2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
                            if (!(r2 <= rl2 && r2 > 1e-12)) continue;
                            int nb = s->nb_index[s->type[a] * s->n_types + s->type[b]];
                            if (nb == 0) continue;
                            if (np < cap && out) {
                                out[3 * np] = a;
                                out[3 * np + 1] = b;
                                out[3 * np + 2] = is14;
                            }
                            ++np;
                        }
                }
            }
    pm_free(&p14);
    free(head);
    free(next);
    return np;
}

/* -------------------------------------------------------- REF nonbonded -- */

static void ref_nonbonded(const mdgpu_system_desc* s, const mdgpu_params* p, const double* x, const double box[3],
                          const oracle_opts* o, mdgpu_energies* e, double* f) {
    const long double kc = SQRT_KC * SQRT_KC;
    int n = s->n_atoms;
    int pbc = box[0] > 0 && box[1] > 0 && box[2] > 0;
    double rl = (p->lj_cutoff > p->ee_cutoff ? p->lj_cutoff : p->ee_cutoff) + o->list_skin;
    long np = oracle_ref_pairlist(s, x, box, rl, NULL, 0);
    int* list = (int*)malloc(sizeof(int) * 3 * (size_t)(np > 0 ? np : 1));
    oracle_ref_pairlist(s, x, box, rl, list, np);
    pairmap scee;
    pm_init(&scee, (size_t)s->n_pairs14);
    for (int t = 0; t < s->n_pairs14; ++t)
        pm_put(&scee, pkey(s->pairs14[2 * t], s->pairs14[2 * t + 1]), s->scee14[t]);

    double rc_lj = p->lj_cutoff, rc2_lj = rc_lj * rc_lj;
    double rc_ee = p->ee_cutoff, rc2_ee = rc_ee * rc_ee;
    int use_sw = p->lj_switch > 0.0;
    double rs = p->lj_switch, alpha = p->ewald_alpha;

    /* precompute_lj_energy_force_shift (NonBondedUtils.cpp:162-183) is used
     * inline below through lj_u/lj_g at rc^2 */
    double evdw = 0.0, eee = 0.0, erf_ = 0.0, eew = 0.0;
    for (long t = 0; t < np; ++t) {
        int a = list[3 * t], b = list[3 * t + 1], is14 = list[3 * t + 2];
        double d[3] = {x[3 * a] - x[3 * b], x[3 * a + 1] - x[3 * b + 1], x[3 * a + 2] - x[3 * b + 2]};
        min_image(d, box, pbc);
        double r2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
        int pidx = s->nb_index[s->type[a] * s->n_types + s->type[b]] - 1;
        double A = is14 ? s->lj14_a[pidx] : s->lj_a[pidx];
        double B = is14 ? s->lj14_b[pidx] : s->lj_b[pidx];
        double scale = 1.0;
        if (is14) pm_get(&scee, pkey(a, b), &scale);
        /* calculate_LJ_energy_pairlist (NonBondedEnergy.cpp:79-131) and
         * calculate_forces_LJ_pairlist (NonBondedForces.cpp:92-154) */
        if (!(r2 > rc2_lj) && !(r2 < 1e-12)) {
            double r = sqrt(r2);
            double u = lj_u(r2, A, B);
            double g = lj_g(r2, A, B);
            if (use_sw) {
                double S, dS;
                lj_switch(r, rs, rc_lj, 1, &S, &dS);
                evdw += u * S;
                g = S * g - (u / r) * dS;
            } else {
                double ucut = lj_u(rc2_lj, A, B), gcut = lj_g(rc2_lj, A, B);
                evdw += u - ucut + (r - rc_lj) * (gcut * rc_lj);
                g -= gcut * (rc_lj / r);
            }
            if (f && r >= 1e-12) {
                pair_force(f, a, b, g, d);
            }
        }
        /* calculate_EE_energy_pairlist (:176-204), _with_cutoff (:206-234),
         * calculate_EE_ewald_direct_term (:236-268) and
         * calculate_forces_EE_pairlist (NonBondedForces.cpp:207-242) */
        if (!(r2 < 1e-12) && !(r2 > rc2_ee)) {
            double r = sqrt(r2);
            double qa = s->charge[a], qb = s->charge[b];
            eee += (1 / scale) * (double)(((qa * qb) / (r * 1.0)) * SQRT_KC * SQRT_KC);
            double irc = 1.0 / rc_ee;
            double krf = irc * irc * irc * ((1.0 - 1) / (2 * 1.0 + 1)), crf = irc * ((3 * 1.0) / (2 * 1.0 + 1));
            erf_ += (1 / scale) * (double)(((qa * qb / 1.0) * SQRT_KC * SQRT_KC) * (1.0 / r + krf * r2 - crf));
            if (is14) {
                double er = (double)((qa * qb) * SQRT_KC * SQRT_KC);
                eew += (1.0 / (scale * r) - erf(r) / r) * er; /* CoulombicEE.cpp:38-44 as written */
            } else {
                eew += (double)(((qa * qb) / r) * erfc(alpha * r) * SQRT_KC * SQRT_KC);
            }
            if (f && r >= 1e-12) {
                double g = (1 / scale) * (double)(((qa * qb) / (r * r * r * 1.0)) * SQRT_KC * SQRT_KC);
                pair_force(f, a, b, g, d);
            }
        }
    }
    /* exclusion correction of calculate_EE_ewald_direct_term (:271-304):
     * excluded partners j > i that are not 1-4, any distance */
    if (s->excl_start) {
        pairmap p14;
        pm_init(&p14, (size_t)s->n_pairs14);
        for (int t = 0; t < s->n_pairs14; ++t) pm_put(&p14, pkey(s->pairs14[2 * t], s->pairs14[2 * t + 1]), 1.0);
        for (int i = 0; i < n; ++i)
            for (int q = s->excl_start[i]; q < s->excl_start[i + 1]; ++q) {
                int j = s->excl_list[q];
                if (j <= i) continue;
                if (pm_get(&p14, pkey(i, j), NULL)) continue;
                double d[3] = {x[3 * i] - x[3 * j], x[3 * i + 1] - x[3 * j + 1], x[3 * i + 2] - x[3 * j + 2]};
                min_image(d, box, pbc);
                double re = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
                eew += (double)(-kc * (s->charge[i] * s->charge[j] / re) * erf(alpha * re));
            }
        pm_free(&p14);
    }
    /* calculate_EE_energy (NonBondedEnergy.cpp:133-174): all i < j, O(N^2) */
    double eall = 0.0;
    if (o->skip_all_pairs) {
        eall = NAN;
    } else {
        pairmap p14;
        pm_init(&p14, (size_t)s->n_pairs14);
        for (int t = 0; t < s->n_pairs14; ++t) pm_put(&p14, pkey(s->pairs14[2 * t], s->pairs14[2 * t + 1]), 1.0);
        for (int a = 0; a < n; ++a) {
            const double* xa = x + 3 * a;
            for (int b = a + 1; b < n; ++b) {
                int is14 = 0;
                if (s->excl_start && excluded(s, a, b)) {
                    is14 = pm_get(&p14, pkey(a, b), NULL);
                    if (!is14) continue;
                }
                double d[3] = {xa[0] - x[3 * b], xa[1] - x[3 * b + 1], xa[2] - x[3 * b + 2]};
                min_image(d, box, pbc);
                double r2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
                if (r2 < 1e-12 || r2 > rc2_ee) continue;
                double r = sqrt(r2);
                double scale = 1.0;
                if (is14) pm_get(&scee, pkey(a, b), &scale);
                eall += (1 / scale) * (double)(((s->charge[a] * s->charge[b]) / (r * 1.0)) * SQRT_KC * SQRT_KC);
            }
        }
        pm_free(&p14);
    }
    e->vdw = evdw;
    e->ee_pairlist = eee;
    e->ee_rf = erf_;
    e->ewald_direct = eew;
    e->ee_all = eall;
    pm_free(&scee);
    free(list);
}

/* ------------------------------------------------------ AMBER nonbonded -- */

static void amber_nonbonded(const mdgpu_system_desc* s, const mdgpu_params* p, const double* x, const ocell* cell,
                            mdgpu_energies* e, double* f) {
    const double* box = cell->L;
    const double kc = (double)(SQRT_KC * SQRT_KC);
    int n = s->n_atoms;
    double rc = p->lj_cutoff > p->ee_cutoff ? p->lj_cutoff : p->ee_cutoff;
    double rc2_lj = p->lj_cutoff * p->lj_cutoff, rc2_ee = p->ee_cutoff * p->ee_cutoff;
    double alpha = p->ewald_alpha, a2 = alpha * alpha, tsp = 2.0 * alpha / sqrt(M_PI);
    int use_sw = p->lj_switch > 0.0;
    /* cell grid over the wrapped coordinates; cells >= rc per side so the 27
     * neighbour cells contain every minimum-image partner within rc (a
     * triclinic cell is cut along its fractional coordinates, into slices at
     * least rc apart between faces) */
    int nc[3];
    for (int d = 0; d < 3; ++d) {
        nc[d] = (int)floor(cell->face[d] / rc);
        if (nc[d] < 1) nc[d] = 1;
    }
    int ncell = nc[0] * nc[1] * nc[2];
    int* head = (int*)malloc(sizeof(int) * ncell);
    int* next = (int*)malloc(sizeof(int) * n);
    int* cidx = (int*)malloc(sizeof(int) * n);
    for (int c = 0; c < ncell; ++c) head[c] = -1;
    for (int i = 0; i < n; ++i) {
        int ix, iy, iz;
        if (!cell->tri) {
            ix = cell_of(wrap1(x[3 * i], box[0]), box[0], nc[0]);
            iy = cell_of(wrap1(x[3 * i + 1], box[1]), box[1], nc[1]);
            iz = cell_of(wrap1(x[3 * i + 2], box[2]), box[2], nc[2]);
        } else {
            double sf[3];
            ocell_frac(cell, x + 3 * i, sf);
            ix = cell_of(wrap1(sf[0], 1.0), 1.0, nc[0]);
            iy = cell_of(wrap1(sf[1], 1.0), 1.0, nc[1]);
            iz = cell_of(wrap1(sf[2], 1.0), 1.0, nc[2]);
        }
        cidx[i] = ix + nc[0] * (iy + nc[1] * iz);
        next[i] = head[cidx[i]];
        head[cidx[i]] = i;
    }
    double evdw = 0.0, eew = 0.0;
    for (int i = 0; i < n; ++i) {
        int c = cidx[i];
        int ix = c % nc[0], iy = (c / nc[0]) % nc[1], iz = c / (nc[0] * nc[1]);
        int neigh[27], m = 0;
        for (int dz = -1; dz <= 1; ++dz)
            for (int dy = -1; dy <= 1; ++dy)
                for (int dx = -1; dx <= 1; ++dx)
                    neigh[m++] = (ix + dx + nc[0]) % nc[0] + nc[0] * ((iy + dy + nc[1]) % nc[1] + nc[1] * ((iz + dz + nc[2]) % nc[2]));
        qsort(neigh, 27, sizeof(int), cmp_int);
        for (int t = 0; t < 27; ++t) {
            if (t > 0 && neigh[t] == neigh[t - 1]) continue;
            for (int j = head[neigh[t]]; j != -1; j = next[j]) {
                if (j <= i) continue;
                if (excluded(s, i, j)) continue;
                double d[3] = {x[3 * i] - x[3 * j], x[3 * i + 1] - x[3 * j + 1], x[3 * i + 2] - x[3 * j + 2]};
                ocell_min_image(cell, d);
                double r2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
                if (r2 < 1e-12) continue;
                double g = 0.0;
                if (r2 <= rc2_lj) {
                    int pidx = s->nb_index[s->type[i] * s->n_types + s->type[j]] - 1;
                    if (pidx >= 0) {
                        double A = s->lj_a[pidx], B = s->lj_b[pidx];
                        double r = sqrt(r2), u = lj_u(r2, A, B), gg = lj_g(r2, A, B);
                        if (use_sw) {
                            double S, dS;
                            lj_switch(r, p->lj_switch, p->lj_cutoff, 1, &S, &dS);
                            evdw += u * S;
                            g += S * gg - (u / r) * dS;
                        } else {
                            double rcl = p->lj_cutoff;
                            double ucut = lj_u(rc2_lj, A, B), gcut = lj_g(rc2_lj, A, B);
                            evdw += u - ucut + (r - rcl) * (gcut * rcl);
                            g += gg - gcut * (rcl / r);
                        }
                    }
                }
                if (r2 <= rc2_ee) {
                    double r = sqrt(r2), qq = kc * s->charge[i] * s->charge[j];
                    double ec = erfc(alpha * r);
                    eew += qq * ec / r;
                    g += qq * (ec / r + tsp * exp(-a2 * r2)) / r2;
                }
                if (f) {
                    pair_force(f, i, j, g, d);
                }
            }
        }
    }
    free(head); free(next); free(cidx);
    e->vdw = evdw;
    e->ewald_direct = eew;

    /* pure exclusions: -k qq erf(a r)/r (reciprocal space counts them) */
    double eex = 0.0;
    for (int t = 0; t < s->n_excl_pairs; ++t) {
        int i = s->excl_pairs[2 * t], j = s->excl_pairs[2 * t + 1];
        double d[3] = {x[3 * i] - x[3 * j], x[3 * i + 1] - x[3 * j + 1], x[3 * i + 2] - x[3 * j + 2]};
        ocell_min_image(cell, d);
        double r2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2], r = sqrt(r2);
        double qq = kc * s->charge[i] * s->charge[j];
        double ef = erf(alpha * r);
        eex -= qq * ef / r;
        if (f) {
            double g = qq * (tsp * exp(-a2 * r2) - ef / r) / r2;
            pair_force(f, i, j, g, d);
        }
    }
    e->ewald_excl = eex;
    /* 1-4 pairs: LJ with the 1-4 tables and Coulomb / SCEE, minus the
     * reciprocal-space erf part (no cutoff, no switch) */
    double ev14 = 0.0, ee14 = 0.0;
    for (int t = 0; t < s->n_pairs14; ++t) {
        int i = s->pairs14[2 * t], j = s->pairs14[2 * t + 1];
        double d[3] = {x[3 * i] - x[3 * j], x[3 * i + 1] - x[3 * j + 1], x[3 * i + 2] - x[3 * j + 2]};
        ocell_min_image(cell, d);
        double r2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2], r = sqrt(r2);
        int pidx = s->nb_index[s->type[i] * s->n_types + s->type[j]] - 1;
        double g = 0.0;
        if (pidx >= 0) {
            double A = s->lj14_a[pidx], B = s->lj14_b[pidx];
            ev14 += lj_u(r2, A, B);
            g += lj_g(r2, A, B);
        }
        double qq = kc * s->charge[i] * s->charge[j];
        double ef = erf(alpha * r), is = 1.0 / s->scee14[t];
        ee14 += qq * (is / r - ef / r);
        g += qq * (is / r + tsp * exp(-a2 * r2) - ef / r) / r2;
        if (f) {
            pair_force(f, i, j, g, d);
        }
    }
    e->vdw14 = ev14;
    e->ee14 = ee14;
}

/* ------------------------------------------------------------- B-splines - */

/* Cardinal B-spline M_p (Essmann et al., J. Chem. Phys. 103, 8577 (1995),
 * eq. 4.1 recursion): theta[j] = M_p(w + j), dtheta[j] = M_p'(w + j). */
void oracle_bspline(int order, double w, double* theta, double* dtheta) {
    double m[16];
    /* M_2(w + j): j = 0 -> w, j = 1 -> 1 - w */
    memset(m, 0, sizeof m);
    m[0] = w;
    m[1] = 1.0 - w;
    for (int k = 3; k <= order; ++k) {
        double nm[16];
        memset(nm, 0, sizeof nm);
        if (k == order) {
            /* derivative M_k'(u) = M_{k-1}(u) - M_{k-1}(u - 1) */
            for (int j = 0; j < k; ++j) dtheta[j] = (j < k - 1 ? m[j] : 0.0) - (j >= 1 ? m[j - 1] : 0.0);
        }
        /* M_k(u) = (u M_{k-1}(u) + (k - u) M_{k-1}(u - 1)) / (k - 1) */
        for (int j = 0; j < k; ++j) {
            double u = w + j;
            double a = j < k - 1 ? m[j] : 0.0;
            double b = j >= 1 ? m[j - 1] : 0.0;
            nm[j] = (u * a + (k - u) * b) / (k - 1);
        }
        memcpy(m, nm, sizeof m);
    }
    for (int j = 0; j < order; ++j) theta[j] = m[j];
    if (order == 2) { dtheta[0] = 1.0; dtheta[1] = -1.0; }
}

/* ----------------------------------------------------------------- PME --- */

/* naive separable DFT along one axis of a complex 3D array (re, im) */
static void dft_axis(double* re, double* im, const int K[3], int axis, int sign) {
    int n = K[axis];
    int stride = axis == 2 ? 1 : (axis == 1 ? K[2] : K[1] * K[2]);
    int o1 = axis == 0 ? K[1] : K[0];
    int o2 = axis == 2 ? K[1] : K[2];
    double* cs = (double*)malloc(sizeof(double) * n);
    double* sn = (double*)malloc(sizeof(double) * n);
    for (int k = 0; k < n; ++k) {
        cs[k] = cos(2.0 * M_PI * k / n);
        sn[k] = sign * sin(2.0 * M_PI * k / n);
    }
    double* tr = (double*)malloc(sizeof(double) * n);
    double* ti = (double*)malloc(sizeof(double) * n);
    for (int a = 0; a < o1; ++a)
        for (int b = 0; b < o2; ++b) {
            size_t base;
            if (axis == 0) base = (size_t)a * K[2] + b;
            else if (axis == 1) base = (size_t)a * K[1] * K[2] + b;
            else base = ((size_t)a * K[1] + b) * K[2];
            for (int m = 0; m < n; ++m) {
                double sr = 0.0, si = 0.0;
                for (int k = 0; k < n; ++k) {
                    int q = (int)(((long)m * k) % n);
                    double xr = re[base + (size_t)k * stride], xi = im[base + (size_t)k * stride];
                    sr += xr * cs[q] - xi * sn[q];
                    si += xr * sn[q] + xi * cs[q];
                }
                tr[m] = sr;
                ti[m] = si;
            }
            for (int m = 0; m < n; ++m) {
                re[base + (size_t)m * stride] = tr[m];
                im[base + (size_t)m * stride] = ti[m];
            }
        }
    free(cs); free(sn); free(tr); free(ti);
}

/* Smooth PME reciprocal energy (Essmann 1995, eqs. 2.48-2.49, 3.4-3.6):
 * Q(k) = sum_i q_i prod_d M_p(u_id - k_d), u = K x / L (wrapped);
 * E = 1/2 sum_m G(m) |F(Q)(m)|^2, G(m) = kC exp(-pi^2 m^2/a^2) / (pi V m^2) B(m),
 * B(m) = prod_d 1/|sum_{k=0}^{p-2} M_p(k+1) e^{2 pi i m_d k / K_d}|^2;
 * F_i = -q_i sum_k grad theta(k) (G * Q)(k). */
static double pme_core(int n, const double* q, const double* x, const double box[3], double alpha, const int K[3],
                       int order, double* f, double* vir, double* pot);

double oracle_pme_recip(int n, const double* q, const double* x, const double box[3], double alpha, const int K[3],
                        int order, double* f, double* vir) {
    return pme_core(n, q, x, box, alpha, K, order, f, vir, NULL);
}

/* Reciprocal potential at each atom, dE_rec/dq_i = sum_k theta_i(k) (G * Q)(k):
 * E_rec = 1/2 Q^T C Q with C the convolution by G, so its q-gradient is the
 * convolved grid sampled with the spreading weights. */
double oracle_pme_potential(int n, const double* q, const double* x, const double box[3], double alpha, const int K[3],
                            int order, double* pot) {
    return pme_core(n, q, x, box, alpha, K, order, NULL, NULL, pot);
}

/* Fractional grid coordinate u = K s of atom i along dimension d (u = K x / L
 * wrapped, in an orthorhombic cell). */
static double pme_u(const ocell* cell, const double* xi, int d, int K) {
    if (!cell->tri) return K * wrap1(xi[d], cell->L[d]) / cell->L[d];
    double sf[3];
    ocell_frac(cell, xi, sf);
    return K * wrap1(sf[d], 1.0);
}

static double pme_core(int n, const double* q, const double* x, const double box[3], double alpha, const int K[3],
                       int order, double* f, double* vir, double* pot) {
    const double kc = (double)(SQRT_KC * SQRT_KC);
    ocell cell;
    if (ocell_make(&cell, box)) return NAN;
    size_t G = (size_t)K[0] * K[1] * K[2];
    double* re = (double*)calloc(G, sizeof(double));
    double* im = (double*)calloc(G, sizeof(double));
    double th[3][16], dth[3][16];
    int base[3];
    for (int i = 0; i < n; ++i) {
        for (int d = 0; d < 3; ++d) {
            double u = pme_u(&cell, x + 3 * i, d, K[d]);
            int fl = (int)floor(u);
            oracle_bspline(order, u - fl, th[d], dth[d]);
            base[d] = fl;
        }
        for (int a = 0; a < order; ++a) {
            int kx = ((base[0] - a) % K[0] + K[0]) % K[0];
            for (int b = 0; b < order; ++b) {
                int ky = ((base[1] - b) % K[1] + K[1]) % K[1];
                for (int c = 0; c < order; ++c) {
                    int kz = ((base[2] - c) % K[2] + K[2]) % K[2];
                    re[((size_t)kx * K[1] + ky) * K[2] + kz] += q[i] * th[0][a] * th[1][b] * th[2][c];
                }
            }
        }
    }
    for (int ax = 0; ax < 3; ++ax) dft_axis(re, im, K, ax, -1);
    /* B-spline moduli */
    double* bm[3];
    for (int d = 0; d < 3; ++d) {
        bm[d] = (double*)malloc(sizeof(double) * K[d]);
        double w[16], dw[16];
        oracle_bspline(order, 0.0, w, dw); /* w[j] = M_p(j) */
        for (int m = 0; m < K[d]; ++m) {
            double sr = 0.0, si = 0.0;
            for (int k = 0; k <= order - 2; ++k) {
                double ang = 2.0 * M_PI * m * k / K[d];
                sr += w[k + 1] * cos(ang);
                si += w[k + 1] * sin(ang);
            }
            bm[d][m] = 1.0 / (sr * sr + si * si);
        }
    }
    double V = ocell_volume(&cell);
    double e = 0.0, W[6] = {0, 0, 0, 0, 0, 0};
    for (int a = 0; a < K[0]; ++a) {
        int ma = a <= K[0] / 2 ? a : a - K[0];
        if (2 * a == K[0]) ma = a; /* Nyquist: either sign gives the same m^2 */
        for (int b = 0; b < K[1]; ++b) {
            int mb = b <= K[1] / 2 ? b : b - K[1];
            for (int c = 0; c < K[2]; ++c) {
                int mc = c <= K[2] / 2 ? c : c - K[2];
                size_t id = ((size_t)a * K[1] + b) * K[2] + c;
                if (a == 0 && b == 0 && c == 0) { re[id] = 0; im[id] = 0; continue; }
                double mv[3];
                if (!cell.tri || mc >= 0 || 2 * c == K[2]) {
                    ocell_recip(&cell, ma, mb, mc, mv);
                } else {
                    /* m_z < 0: the negative of the conjugate point's vector,
                     * so that m and -m pair up on the Nyquist planes too */
                    const int a2 = (K[0] - a) % K[0], b2 = (K[1] - b) % K[1];
                    int ma2 = a2 <= K[0] / 2 ? a2 : a2 - K[0];
                    int mb2 = b2 <= K[1] / 2 ? b2 : b2 - K[1];
                    ocell_recip(&cell, -ma2, -mb2, mc, mv);
                }
                double mx = mv[0], my = mv[1], mz = mv[2];
                double m2 = mx * mx + my * my + mz * mz;
                double g = kc * exp(-M_PI * M_PI * m2 / (alpha * alpha)) / (M_PI * V * m2) * bm[0][a] * bm[1][b] * bm[2][c];
                double s2 = re[id] * re[id] + im[id] * im[id];
                double em = 0.5 * g * s2;
                e += em;
                double fac = 2.0 * (1.0 + M_PI * M_PI * m2 / (alpha * alpha)) / m2;
                W[0] += em * (1.0 - fac * mx * mx);
                W[1] += em * (1.0 - fac * my * my);
                W[2] += em * (1.0 - fac * mz * mz);
                W[3] += em * (-fac * mx * my);
                W[4] += em * (-fac * mx * mz);
                W[5] += em * (-fac * my * mz);
                re[id] *= g;
                im[id] *= g;
            }
        }
    }
    if (vir) {
        vir[0] = W[0]; vir[4] = W[1]; vir[8] = W[2];
        vir[1] = vir[3] = W[3];
        vir[2] = vir[6] = W[4];
        vir[5] = vir[7] = W[5];
    }
    if (f || pot) {
        for (int ax = 0; ax < 3; ++ax) dft_axis(re, im, K, ax, +1);
        for (int i = 0; i < n; ++i) {
            for (int d = 0; d < 3; ++d) {
                double u = pme_u(&cell, x + 3 * i, d, K[d]);
                int fl = (int)floor(u);
                oracle_bspline(order, u - fl, th[d], dth[d]);
                base[d] = fl;
            }
            double fx = 0, fy = 0, fz = 0, ph = 0;
            for (int a = 0; a < order; ++a) {
                int kx = ((base[0] - a) % K[0] + K[0]) % K[0];
                for (int b = 0; b < order; ++b) {
                    int ky = ((base[1] - b) % K[1] + K[1]) % K[1];
                    for (int c = 0; c < order; ++c) {
                        int kz = ((base[2] - c) % K[2] + K[2]) % K[2];
                        double phi = re[((size_t)kx * K[1] + ky) * K[2] + kz];
                        fx += dth[0][a] * th[1][b] * th[2][c] * phi;
                        fy += th[0][a] * dth[1][b] * th[2][c] * phi;
                        fz += th[0][a] * th[1][b] * dth[2][c] * phi;
                        ph += th[0][a] * th[1][b] * th[2][c] * phi;
                    }
                }
            }
            if (pot) pot[i] = ph;
            if (f && !cell.tri) {
                f[3 * i] -= q[i] * fx * K[0] / box[0];
                f[3 * i + 1] -= q[i] * fy * K[1] / box[1];
                f[3 * i + 2] -= q[i] * fz * K[2] / box[2];
            } else if (f) {
                /* dE/dx = R^T dE/ds, dE/ds_d = K_d dE/du_d */
                const double ux = fx * K[0], uy = fy * K[1], uz = fz * K[2];
                f[3 * i] -= q[i] * (ux * cell.r00);
                f[3 * i + 1] -= q[i] * (ux * cell.r01 + uy * cell.r11);
                f[3 * i + 2] -= q[i] * (ux * cell.r02 + uy * cell.r12 + uz * cell.r22);
            }
        }
    }
    for (int d = 0; d < 3; ++d) free(bm[d]);
    free(re);
    free(im);
    return e;
}

double oracle_ewald_recip_dense(int n, const double* q, const double* x, const double box[3], double alpha, int mmax,
                                double* f, double* vir) {
    const double kc = (double)(SQRT_KC * SQRT_KC);
    ocell cell;
    if (ocell_make(&cell, box)) return NAN;
    double V = ocell_volume(&cell);
    double e = 0.0, W[6] = {0, 0, 0, 0, 0, 0};
    double* c = (double*)malloc(sizeof(double) * n);
    double* s = (double*)malloc(sizeof(double) * n);
    for (int a = -mmax; a <= mmax; ++a)
        for (int b = -mmax; b <= mmax; ++b)
            for (int cc = -mmax; cc <= mmax; ++cc) {
                if (a == 0 && b == 0 && cc == 0) continue;
                double mv[3];
                ocell_recip(&cell, a, b, cc, mv);
                double mx = mv[0], my = mv[1], mz = mv[2];
                double m2 = mx * mx + my * my + mz * mz;
                double g = kc * exp(-M_PI * M_PI * m2 / (alpha * alpha)) / (M_PI * V * m2);
                double sr = 0, si = 0;
                for (int i = 0; i < n; ++i) {
                    double ph = 2.0 * M_PI * (mx * x[3 * i] + my * x[3 * i + 1] + mz * x[3 * i + 2]);
                    c[i] = cos(ph);
                    s[i] = sin(ph);
                    sr += q[i] * c[i];
                    si += q[i] * s[i];
                }
                double em = 0.5 * g * (sr * sr + si * si);
                e += em;
                double fac = 2.0 * (1.0 + M_PI * M_PI * m2 / (alpha * alpha)) / m2;
                W[0] += em * (1.0 - fac * mx * mx);
                W[1] += em * (1.0 - fac * my * my);
                W[2] += em * (1.0 - fac * mz * mz);
                W[3] += em * (-fac * mx * my);
                W[4] += em * (-fac * mx * mz);
                W[5] += em * (-fac * my * mz);
                if (f)
                    for (int i = 0; i < n; ++i) {
                        /* dE/dx_i = g q_i 2 pi m (c_i si - s_i sr) */
                        double t = g * q[i] * 2.0 * M_PI * (c[i] * si - s[i] * sr);
                        f[3 * i] -= t * mx;
                        f[3 * i + 1] -= t * my;
                        f[3 * i + 2] -= t * mz;
                    }
            }
    if (vir) {
        vir[0] = W[0]; vir[4] = W[1]; vir[8] = W[2];
        vir[1] = vir[3] = W[3];
        vir[2] = vir[6] = W[4];
        vir[5] = vir[7] = W[5];
    }
    free(c);
    free(s);
    return e;
}

/* ------------------------------------------------------------ driver ----- */

int oracle_eval(const mdgpu_system_desc* s, const mdgpu_params* p, const double* x, const double box[3],
                const oracle_opts* opt, mdgpu_energies* e, double* f, double* vir) {
    oracle_opts dflt = {2.0, 0, 0};
    if (!opt) opt = &dflt;
    if (!s || !p || !x || !e) return 1;
    memset(e, 0, sizeof *e);
    int n = s->n_atoms;
    if (f) memset(f, 0, sizeof(double) * 3 * (size_t)n);
    if (vir) memset(vir, 0, sizeof(double) * 9);
    int amber = p->semantics == MDGPU_SEM_AMBER;
    ocell cell;
    if (ocell_make(&cell, box)) return 2;
    if (cell.tri && !amber) return 3; /* triclinic cells: AMBER semantics only */
    g_amber_torsion = amber;
    /* full virial: bonded forces into their own array (each term's forces sum
     * to zero, so sum_i x_i f_i is its virial with the coordinates as given),
     * pair forces through pair_force, the reciprocal part from the PME pass */
    const int full = opt->full_virial && vir && f;
    double* fb = full ? (double*)calloc(3 * (size_t)n, sizeof(double)) : f;
    double wpair[9] = {0};
    /* calculate_energies / calculate_forces (System.cpp:9-39) */
    e->bond = harmonic_pairs(s->n_bonds, s->bond_atoms, s->bond_type, s->bond_k, s->bond_r0, x, fb);
    e->angle = angles(s, x, fb);
    e->dihedral = dihedrals(s, x, fb);
    e->urey_bradley = harmonic_pairs(s->n_ubs, s->ub_atoms, s->ub_type, s->ub_k, s->ub_r0, x, fb);
    e->improper = impropers(s, x, fb, amber);
    e->cmap = cmaps(s, x, fb);
    if (full) {
        for (int i = 0; i < n; ++i)
            for (int u = 0; u < 3; ++u) {
                f[3 * i + u] += fb[3 * i + u];
                for (int v = 0; v < 3; ++v) vir[3 * u + v] += x[3 * i + u] * fb[3 * i + v];
            }
        free(fb);
        g_wpair = wpair;
    }
    /* calculate_EE_ewald_self_term (NonBondedEnergy.cpp:307-316) */
    {
        const double kc = (double)(SQRT_KC * SQRT_KC);
        double es = 0.0;
        for (int i = 0; i < n; ++i) es += -kc * (p->ewald_alpha / sqrt(M_PI)) * (s->charge[i] * s->charge[i]);
        e->ewald_self = es;
    }
    double bonded = e->bond + e->angle + e->dihedral + e->urey_bradley + e->improper + e->cmap;
    if (!amber) {
        ref_nonbonded(s, p, x, box, opt, e, f);
        e->total = bonded + e->vdw + e->ee_pairlist;
    } else {
        amber_nonbonded(s, p, x, &cell, e, f);
        if (p->pme_grid[0] > 0) {
            double wr[9];
            e->ewald_recip = oracle_pme_recip(n, s->charge, x, box, p->ewald_alpha, p->pme_grid, p->pme_order, f,
                                              vir ? wr : NULL);
            if (vir)
                for (int k = 0; k < 9; ++k) vir[k] += wr[k];
        }
        e->total = bonded + e->vdw + e->vdw14 + e->ewald_direct + e->ewald_recip + e->ewald_self + e->ewald_excl +
                   e->ee14;
    }
    if (full) {
        g_wpair = NULL;
        for (int k = 0; k < 9; ++k) vir[k] += wpair[k];
    }
    return 0;
}

/* --------------------------------------------------- parameter gradients - */

/* Weight of term t in a weighted sum: w.t plus w.total for every term of the
 * AMBER total. */
static void eff_weights(const mdgpu_energies* w, mdgpu_energies* o) {
    *o = *w;
    o->bond += w->total; o->angle += w->total; o->dihedral += w->total; o->urey_bradley += w->total;
    o->improper += w->total; o->cmap += w->total; o->vdw += w->total; o->ewald_direct += w->total;
    o->ewald_self += w->total; o->ewald_recip += w->total; o->ewald_excl += w->total;
    o->vdw14 += w->total; o->ee14 += w->total;
}

static void zero_if(double* a, size_t n) { if (a) memset(a, 0, n * sizeof(double)); }

/* Switched or force-shifted LJ energy of one pair for unit (A, B): the
 * energy is linear in A and B (NonBondedEnergy.cpp:79-131), so dE/dA and
 * dE/dB are the energies at (1, 0) and (0, 1). */
static void lj_unit(const mdgpu_params* p, double r2, double* ea, double* eb) {
    double r = sqrt(r2), i6 = 1.0 / (r2 * r2 * r2), i12 = i6 * i6;
    if (p->lj_switch > 0.0) {
        double S, dS;
        lj_switch(r, p->lj_switch, p->lj_cutoff, 1, &S, &dS);
        *ea = S * i12;
        *eb = -S * i6;
    } else {
        double rc = p->lj_cutoff, rc2 = rc * rc;
        double c6 = 1.0 / (rc2 * rc2 * rc2), c12 = c6 * c6;
        *ea = i12 - c12 + (r - rc) * rc * (12.0 * c12 / rc2);
        *eb = -(i6 - c6) - (r - rc) * rc * (6.0 * c6 / rc2);
    }
}

int oracle_param_grads(const mdgpu_system_desc* s, const mdgpu_params* p, const double* x, const double box[3],
                       const mdgpu_energies* weights, mdgpu_param_grads* g) {
    if (!s || !p || !x || !weights || !g || p->semantics != MDGPU_SEM_AMBER) return 1;
    if (!(g_angles[0] == 90.0 && g_angles[1] == 90.0 && g_angles[2] == 90.0)) return 1; /* orthorhombic only */
    const double kc = (double)(SQRT_KC * SQRT_KC);
    mdgpu_energies w;
    eff_weights(weights, &w);
    int n = s->n_atoms;
    zero_if(g->charge, (size_t)n);
    zero_if(g->lj_a, (size_t)s->n_lj); zero_if(g->lj_b, (size_t)s->n_lj);
    zero_if(g->lj14_a, (size_t)s->n_lj); zero_if(g->lj14_b, (size_t)s->n_lj);
    zero_if(g->bond_k, (size_t)s->n_bond_types); zero_if(g->bond_r0, (size_t)s->n_bond_types);
    zero_if(g->ub_k, (size_t)s->n_ub_types); zero_if(g->ub_r0, (size_t)s->n_ub_types);
    zero_if(g->angle_k, (size_t)s->n_angle_types); zero_if(g->angle_theta0, (size_t)s->n_angle_types);
    zero_if(g->dihedral_k, (size_t)s->n_dihedral_types); zero_if(g->dihedral_phase, (size_t)s->n_dihedral_types);
    zero_if(g->improper_k, (size_t)s->n_improper_types); zero_if(g->improper_psi0, (size_t)s->n_improper_types);
    size_t ncoef = (size_t)s->cmap_n_sets * s->cmap_resolution * s->cmap_resolution * 16;
    zero_if(g->cmap_coeff, ncoef);

    /* bonds and Urey-Bradley: E = k (r - r0)^2 */
    for (int pass = 0; pass < 2; ++pass) {
        int nt = pass ? s->n_ubs : s->n_bonds;
        const int32_t* at = pass ? s->ub_atoms : s->bond_atoms;
        const int32_t* ty = pass ? s->ub_type : s->bond_type;
        const double* kk = pass ? s->ub_k : s->bond_k;
        const double* r0 = pass ? s->ub_r0 : s->bond_r0;
        double* gk = pass ? g->ub_k : g->bond_k;
        double* gr = pass ? g->ub_r0 : g->bond_r0;
        double wt = pass ? w.urey_bradley : w.bond;
        for (int t = 0; t < nt; ++t) {
            int q = ty[t];
            double d = m_distance(x + 3 * at[2 * t], x + 3 * at[2 * t + 1]) - r0[q];
            if (gk) gk[q] += wt * d * d;
            if (gr) gr[q] += wt * (-2.0 * kk[q] * d);
        }
    }
    /* angles: E = k (theta - theta0)^2 */
    for (int t = 0; t < s->n_angles; ++t) {
        const int32_t* at = s->angle_atoms + 3 * t;
        int q = s->angle_type[t];
        double d = m_angle(x + 3 * at[0], x + 3 * at[1], x + 3 * at[2]) - s->angle_theta0[q];
        if (g->angle_k) g->angle_k[q] += w.angle * d * d;
        if (g->angle_theta0) g->angle_theta0[q] += w.angle * (-2.0 * s->angle_k[q] * d);
    }
    /* dihedrals: E = k (1 + cos(n phi - phase)) */
    for (int t = 0; t < s->n_dihedrals; ++t) {
        const int32_t* at = s->dihedral_atoms + 4 * t;
        int q = s->dihedral_type[t];
        double phi = m_dihedral(x + 3 * at[0], x + 3 * at[1], x + 3 * at[2], x + 3 * at[3]);
        double arg = s->dihedral_n[q] * phi - s->dihedral_phase[q];
        if (g->dihedral_k) g->dihedral_k[q] += w.dihedral * (1.0 + cos(arg));
        if (g->dihedral_phase) g->dihedral_phase[q] += w.dihedral * s->dihedral_k[q] * sin(arg);
    }
    /* impropers (AMBER: wrapped difference): E = k d^2 */
    for (int t = 0; t < s->n_impropers; ++t) {
        const int32_t* at = s->improper_atoms + 4 * t;
        int q = s->improper_type[t];
        double psi = m_dihedral(x + 3 * at[0], x + 3 * at[1], x + 3 * at[2], x + 3 * at[3]);
        double d = wrap_pi(psi - s->improper_psi0[q]);
        if (g->improper_k) g->improper_k[q] += w.improper * d * d;
        if (g->improper_psi0) g->improper_psi0[q] += w.improper * (-2.0 * s->improper_k[q] * d);
    }
    /* CMAP: E = sum_ij a_ij u^i v^j in the cell of (phi, psi) */
    if (g->cmap_coeff) {
        int res = s->cmap_resolution;
        for (int t = 0; t < s->n_cmaps; ++t) {
            const int32_t* at = s->cmap_atoms + 5 * t;
            double phi = m_dihedral(x + 3 * at[0], x + 3 * at[1], x + 3 * at[2], x + 3 * at[3]);
            double psi = m_dihedral(x + 3 * at[1], x + 3 * at[2], x + 3 * at[3], x + 3 * at[4]);
            phi = wrap_pi(phi) + M_PI;
            psi = wrap_pi(psi) + M_PI;
            double sp = (2.0 * M_PI) / res, gp = phi / sp, gs = psi / sp;
            int row = (int)gp, col = (int)gs;
            double u = gp - row, v = gs - col;
            row %= res; col %= res;
            if (row < 0) row += res;
            if (col < 0) col += res;
            double* gc = g->cmap_coeff + (size_t)(s->cmap_set[t] - 1) * res * res * 16 + (size_t)(row * res + col) * 16;
            double up[4] = {1.0, u, u * u, u * u * u}, vp[4] = {1.0, v, v * v, v * v * v};
            for (int i = 0; i < 4; ++i)
                for (int j = 0; j < 4; ++j) gc[4 * i + j] += w.cmap * up[i] * vp[j];
        }
    }

    /* pairs within the cutoffs (brute force over i < j with minimum image;
     * the oracle is used on small systems) */
    double alpha = p->ewald_alpha;
    double rc2_lj = p->lj_cutoff * p->lj_cutoff, rc2_ee = p->ee_cutoff * p->ee_cutoff;
    for (int i = 0; i < n; ++i)
        for (int j = i + 1; j < n; ++j) {
            if (excluded(s, i, j)) continue;
            double d[3] = {x[3 * i] - x[3 * j], x[3 * i + 1] - x[3 * j + 1], x[3 * i + 2] - x[3 * j + 2]};
            min_image(d, box, 1);
            double r2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
            if (r2 < 1e-12) continue;
            if (r2 <= rc2_lj) {
                int pidx = s->nb_index[s->type[i] * s->n_types + s->type[j]] - 1;
                if (pidx >= 0) {
                    double ea, eb;
                    lj_unit(p, r2, &ea, &eb);
                    if (g->lj_a) g->lj_a[pidx] += w.vdw * ea;
                    if (g->lj_b) g->lj_b[pidx] += w.vdw * eb;
                }
            }
            if (r2 <= rc2_ee && g->charge) {
                double r = sqrt(r2), ec = kc * erfc(alpha * r) / r * w.ewald_direct;
                g->charge[i] += ec * s->charge[j];
                g->charge[j] += ec * s->charge[i];
            }
        }
    /* pure exclusions: E = -k qq erf(a r)/r */
    for (int t = 0; t < s->n_excl_pairs; ++t) {
        int i = s->excl_pairs[2 * t], j = s->excl_pairs[2 * t + 1];
        double d[3] = {x[3 * i] - x[3 * j], x[3 * i + 1] - x[3 * j + 1], x[3 * i + 2] - x[3 * j + 2]};
        min_image(d, box, 1);
        double r = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
        double c = -kc * erf(alpha * r) / r * w.ewald_excl;
        if (g->charge) { g->charge[i] += c * s->charge[j]; g->charge[j] += c * s->charge[i]; }
    }
    /* 1-4 pairs: LJ-14 tables and k qq (1/(scee r) - erf(a r)/r) */
    for (int t = 0; t < s->n_pairs14; ++t) {
        int i = s->pairs14[2 * t], j = s->pairs14[2 * t + 1];
        double d[3] = {x[3 * i] - x[3 * j], x[3 * i + 1] - x[3 * j + 1], x[3 * i + 2] - x[3 * j + 2]};
        min_image(d, box, 1);
        double r2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2], r = sqrt(r2);
        int pidx = s->nb_index[s->type[i] * s->n_types + s->type[j]] - 1;
        if (pidx >= 0) {
            double i6 = 1.0 / (r2 * r2 * r2);
            if (g->lj14_a) g->lj14_a[pidx] += w.vdw14 * i6 * i6;
            if (g->lj14_b) g->lj14_b[pidx] += -w.vdw14 * i6;
        }
        double c = kc * (1.0 / (s->scee14[t] * r) - erf(alpha * r) / r) * w.ee14;
        if (g->charge) { g->charge[i] += c * s->charge[j]; g->charge[j] += c * s->charge[i]; }
    }
    if (g->charge) {
        /* self: E = -k a/sqrt(pi) sum q^2 */
        for (int i = 0; i < n; ++i) g->charge[i] += w.ewald_self * (-2.0 * kc * alpha / sqrt(M_PI)) * s->charge[i];
        if (p->pme_grid[0] > 0 && w.ewald_recip != 0.0) {
            double* pot = (double*)malloc(sizeof(double) * (size_t)n);
            pme_core(n, s->charge, x, box, alpha, p->pme_grid, p->pme_order, NULL, NULL, pot);
            for (int i = 0; i < n; ++i) g->charge[i] += w.ewald_recip * pot[i];
            free(pot);
        }
    }
    return 0;
}

/* ------------------------------------------------------------------ MD ---- */

/* The reference has no integrator (apps/main.cpp:23-28 stops after one
 * evaluation).  Textbook velocity Verlet: half kick, drift, new forces, half
 * kick; kinetic energy sum m v^2 / 2 converted to kcal/mol. */
int oracle_md_run(const mdgpu_system_desc* s, const mdgpu_params* p, const double* masses, double* x, double* v,
                  double* f, const double box[3], double dt, int nsteps, double* trace) {
    const double acc = 418.4, hdt = 0.5 * dt;
    const int n = s->n_atoms;
    oracle_opts o = {2.0, 1};
    mdgpu_energies e;
    for (int k = 0; k < nsteps; ++k) {
        for (int i = 0; i < 3 * n; ++i) {
            v[i] = v[i] + hdt * (acc * f[i] * (1.0 / masses[i / 3]));
            x[i] = x[i] + dt * v[i];
        }
        if (oracle_eval(s, p, x, box, &o, &e, f, NULL) != 0) return 1;
        double ke = 0.0;
        for (int i = 0; i < 3 * n; ++i) {
            v[i] = v[i] + hdt * (acc * f[i] * (1.0 / masses[i / 3]));
            ke += masses[i / 3] * v[i] * v[i];
        }
        if (trace) {
            trace[2 * k] = e.total;
            trace[2 * k + 1] = ke * (0.5 / acc);
        }
    }
    return 0;
}

/* ------------------------------------------------------------ minimise ---- */

/* The reference leaves sys.minimize() commented out (apps/main.cpp:28) and
 * names L-BFGS in its comments (System.h:84,113).  Nocedal & Wright,
 * Numerical Optimization, Algorithms 7.4 and 7.5, with backtracking Armijo. */
static double dotv(const double* a, const double* b, int n) {
    double s = 0.0;
    for (int i = 0; i < n; ++i) s += a[i] * b[i];
    return s;
}
static double max_atom2(const double* a, int n_atoms) {
    double m = 0.0;
    for (int i = 0; i < n_atoms; ++i) {
        double t = a[3 * i] * a[3 * i] + a[3 * i + 1] * a[3 * i + 1] + a[3 * i + 2] * a[3 * i + 2];
        if (t > m) m = t;
    }
    return m;
}

int oracle_minimize(const mdgpu_system_desc* s, const mdgpu_params* p, double* x, const double box[3], int max_iter,
                    double rms_tol, double max_step, int m, double* energies, double* out) {
    const int n = s->n_atoms, n3 = 3 * n;
    oracle_opts o = {2.0, 1};
    mdgpu_energies e;
    double* f = (double*)malloc(sizeof(double) * (size_t)n3 * (6 + 2 * (size_t)m));
    if (!f) return 1;
    double *ft = f + n3, *xt = ft + n3, *pv = xt + n3, *q = pv + n3, *sv = q + n3, *yv = sv + (size_t)m * n3;
    double rho[16], alpha[16];
    int kept = 0, newest = -1, it = 0, evals = 1, conv = 0;
    if (oracle_eval(s, p, x, box, &o, &e, f, NULL) != 0) { free(f); return 1; }
    double energy = e.total, rms = 0.0, fmax = 0.0;
    if (energies) energies[0] = energy;
    for (it = 0;; ++it) {
        double f2 = dotv(f, f, n3), fm2 = max_atom2(f, n);
        rms = sqrt(f2 / n3);
        fmax = sqrt(fm2);
        if (rms <= rms_tol) { conv = 1; break; }
        if (it >= max_iter) break;
        /* q and r carry F = -g; the alphas and betas come out negated */
        memcpy(q, f, sizeof(double) * n3);
        for (int k = 0; k < kept; ++k) {
            int i = (newest - k + m) % m;
            alpha[i] = dotv(sv + (size_t)i * n3, q, n3) * rho[i];
            for (int j = 0; j < n3; ++j) q[j] = q[j] + (-alpha[i]) * yv[(size_t)i * n3 + j];
        }
        if (kept > 0) {
            double yy = dotv(yv + (size_t)newest * n3, yv + (size_t)newest * n3, n3);
            double g0 = 1.0 / (rho[newest] * yy);
            for (int j = 0; j < n3; ++j) pv[j] = g0 * q[j];
            for (int k = kept - 1; k >= 0; --k) {
                int i = (newest - k + m) % m;
                double beta = dotv(yv + (size_t)i * n3, pv, n3) * rho[i];
                double c = alpha[i] - beta;
                for (int j = 0; j < n3; ++j) pv[j] = pv[j] + c * sv[(size_t)i * n3 + j];
            }
        } else {
            memcpy(pv, q, sizeof(double) * n3);
        }
        double gp = -dotv(f, pv, n3), pmax = sqrt(max_atom2(pv, n));
        if (!(gp < 0.0)) {
            kept = 0;
            memcpy(pv, f, sizeof(double) * n3);
            gp = -f2;
            pmax = fmax;
        }
        double a = 1.0, et = 0.0;
        if (a * pmax > max_step) a = max_step / pmax;
        int accepted = 0;
        for (int ls = 0; ls < 30; ++ls) {
            for (int j = 0; j < n3; ++j) xt[j] = x[j] + a * pv[j];
            if (oracle_eval(s, p, xt, box, &o, &e, ft, NULL) != 0) { free(f); return 1; }
            ++evals;
            et = e.total;
            if (et <= energy + 1e-4 * a * gp) { accepted = 1; break; }
            a *= 0.5;
        }
        if (!accepted) break;
        int slot = (newest + 1) % m;
        double* ss = sv + (size_t)slot * n3;
        double* yy = yv + (size_t)slot * n3;
        for (int j = 0; j < n3; ++j) {
            ss[j] = a * pv[j];
            yy[j] = f[j] + (-1.0) * ft[j];
        }
        double ys = dotv(yy, ss, n3);
        memcpy(x, xt, sizeof(double) * n3);
        memcpy(f, ft, sizeof(double) * n3);
        energy = et;
        if (energies) energies[it + 1] = energy;
        if (ys > 0.0) {
            rho[slot] = 1.0 / ys;
            newest = slot;
            kept = kept + 1 < m ? kept + 1 : m;
        }
    }
    if (out) {
        out[0] = it;
        out[1] = evals;
        out[2] = conv;
        out[3] = rms;
        out[4] = fmax;
    }
    free(f);
    return 0;
}
