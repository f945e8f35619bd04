// Scalar term formulas of the host API (the per-term members that the
// reference keeps private to md::System).  The GPU kernels evaluate the same
// formulas; these host versions serve md::System's API and its reports.
#include "Bonded/Terms.h"
#include "NonBonded/CoulombicEE.h"
#include "System/Metrics.h"

#include <algorithm>
#include <cmath>
#include <limits>

namespace md::Bonded {

// E = k (1 + cos(n phi - phase))  (src/Bonded/CosineDihedral.cpp:6-10)
double CosineDihedral::calculate_energy(double phi) const { return k_ * (1.0 + std::cos(n_ * phi - phase_)); }

namespace {
// Cell lookup shared by the CMAP energy and gradient (CMapGroup.cpp:82-121):
// angles are brought to (-pi, pi], shifted by +pi, divided by the grid spacing;
// the integer parts select the patch and the fractions are (u, v).
struct CmapCell {
    int offset;
    double u, v, inv_spacing;
};
CmapCell cmap_cell(double phi, double psi, int res, int set) {
    constexpr double kPi = 3.14159265358979323846;
    while (phi <= -kPi) phi += 2.0 * kPi;
    while (phi > kPi) phi -= 2.0 * kPi;
    while (psi <= -kPi) psi += 2.0 * kPi;
    while (psi > kPi) psi -= 2.0 * kPi;
    phi += kPi;
    psi += kPi;
    const double spacing = (2.0 * kPi) / res;
    const double gp = phi / spacing, gs = psi / spacing;
    int row = static_cast<int>(gp), col = static_cast<int>(gs);
    const double u = gp - row, v = gs - col;
    row %= res;
    col %= res;
    if (row < 0) row += res;
    if (col < 0) col += res;
    return {(set - 1) * res * res * 16 + (row * res + col) * 16, u, v, 1.0 / spacing};
}
}  // namespace

double CMapGroup::return_energy_bicubic(double phi, double psi, int res, const std::vector<double>& coeffs,
                                        const std::vector<double>&) const {
    const CmapCell c = cmap_cell(phi, psi, res, set_);
    const double up[4] = {1.0, c.u, c.u * c.u, c.u * c.u * c.u};
    const double vp[4] = {1.0, c.v, c.v * c.v, c.v * c.v * c.v};
    double e = 0.0;
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j) e += coeffs[c.offset + 4 * i + j] * up[i] * vp[j];
    return e;
}

std::pair<double, double> CMapGroup::return_gradient_bicubic(double phi, double psi, int res,
                                                             const std::vector<double>& coeffs,
                                                             const std::vector<double>&) const {
    const CmapCell c = cmap_cell(phi, psi, res, set_);
    const double up[4] = {1.0, c.u, c.u * c.u, c.u * c.u * c.u};
    const double vp[4] = {1.0, c.v, c.v * c.v, c.v * c.v * c.v};
    const double dup[4] = {0.0, 1.0, 2.0 * c.u, 3.0 * c.u * c.u};
    const double dvp[4] = {0.0, 1.0, 2.0 * c.v, 3.0 * c.v * c.v};
    double du = 0.0, dv = 0.0;
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j) {
            const double a = coeffs[c.offset + 4 * i + j];
            du += a * dup[i] * vp[j];
            dv += a * up[i] * dvp[j];
        }
    return {du * c.inv_spacing, dv * c.inv_spacing};
}

}  // namespace md::Bonded

namespace md::NonBonded {

namespace {
constexpr long double kSqrtCoulomb = 18.2206899283247L;
}

double CoulombicEE::CalculateEnergy(double r, double qa, double qb, double eps) {
    return static_cast<double>(((qa * qb) / (r * eps)) * kSqrtCoulomb * kSqrtCoulomb);
}
double CoulombicEE::CalculateGradient(double r, double qa, double qb, double eps) {
    return static_cast<double>(((qa * qb) / (r * r * r * eps)) * kSqrtCoulomb * kSqrtCoulomb);
}
// Reaction field (CoulombicEE.cpp:21-31); with eps = 1 it is qq k (1/r - 1/rc).
double CoulombicEE::CalculateEnergyCutoff(double r, double rc, double qa, double qb, double eps) {
    const double irc = 1.0 / rc;
    const double krf = irc * irc * irc * ((eps - 1) / (2 * eps + 1));
    const double crf = irc * ((3 * eps) / (2 * eps + 1));
    return static_cast<double>(((qa * qb) / eps) * kSqrtCoulomb * kSqrtCoulomb) * (1.0 / r + krf * r * r - crf);
}
double CoulombicEE::CalculateEnergyEwaldDirectTerm(double r, double qa, double qb, double alpha) {
    return static_cast<double>(((qa * qb) / r) * std::erfc(alpha * r) * kSqrtCoulomb * kSqrtCoulomb);
}
// The reference evaluates erf(r) here, not erf(alpha r) (CoulombicEE.cpp:38-44);
// kept as written for REF semantics.
double CoulombicEE::CalculateEnergyEwaldDirectTerm_14(double r, double qa, double qb, double, double scale) {
    const double er = static_cast<double>((qa * qb) * kSqrtCoulomb * kSqrtCoulomb);
    return (1.0 / (scale * r) - std::erf(r) / r) * er;
}

}  // namespace md::NonBonded

namespace Metrics {

double distance(double x1, double y1, double z1, double x2, double y2, double z2) {
    const double dx = x1 - x2, dy = y1 - y2, dz = z1 - z2;
    return std::sqrt(dx * dx + dy * dy + dz * dz);
}

double angle(double x1, double y1, double z1, double x2, double y2, double z2, double x3, double y3, double z3) {
    const double ax = x1 - x2, ay = y1 - y2, az = z1 - z2;
    const double bx = x3 - x2, by = y3 - y2, bz = z3 - z2;
    const double a2 = ax * ax + ay * ay + az * az, b2 = bx * bx + by * by + bz * bz;
    if (a2 < 1e-14 || b2 < 1e-14) return std::numeric_limits<double>::quiet_NaN();
    const double c = std::clamp((ax * bx + ay * by + az * bz) * (1.0 / std::sqrt(a2 * b2)), -1.0, 1.0);
    return std::acos(c);
}

double dihedral(double x1, double y1, double z1, double x2, double y2, double z2, double x3, double y3, double z3,
                double x4, double y4, double z4) {
    const double b1x = x2 - x1, b1y = y2 - y1, b1z = z2 - z1;
    const double b2x = x3 - x2, b2y = y3 - y2, b2z = z3 - z2;
    const double b3x = x4 - x3, b3y = y4 - y3, b3z = z4 - z3;
    const double n1x = b1y * b2z - b1z * b2y, n1y = b2x * b1z - b1x * b2z, n1z = b1x * b2y - b1y * b2x;
    const double n2x = b2y * b3z - b2z * b3y, n2y = b3x * b2z - b2x * b3z, n2z = b2x * b3y - b2y * b3x;
    const double m1 = n1x * n1x + n1y * n1y + n1z * n1z, m2 = n2x * n2x + n2y * n2y + n2z * n2z;
    const double lb2 = std::sqrt(b2x * b2x + b2y * b2y + b2z * b2z);
    if (m1 < 1e-28 || m2 < 1e-28 || lb2 < 1e-14) return std::numeric_limits<double>::quiet_NaN();
    const double x = n1x * n2x + n1y * n2y + n1z * n2z;
    const double y = (b1x * n2x + b1y * n2y + b1z * n2z) * lb2;
    return std::atan2(y, x);
}

}  // namespace Metrics
