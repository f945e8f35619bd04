// Scalar LJ 12-6 formulas (reference include/NonBonded/LennardJones.h:14-19,
// src/NonBonded/LennardJones.cpp:9-27): U = A/r^12 - B/r^6 and
// G = -U'(r)/r = 12A/r^14 - 6B/r^8, both from r^2.
#ifndef MD_AMD_LENNARDJONES_H
#define MD_AMD_LENNARDJONES_H

namespace md { class System; }

namespace md::NonBonded {
class LennardJones {
private:
    static double CalculateEnergy(double r2, double a, double b) {
        const double s2 = 1.0 / r2, s6 = s2 * s2 * s2;
        return a * s6 * s6 - b * s6;
    }
    static double CalculateGradient(double r2, double a, double b) {
        const double s2 = 1.0 / r2, s4 = s2 * s2, s6 = s4 * s2, s8 = s4 * s4;
        return 12.0 * a * s6 * s8 - 6.0 * b * s8;
    }
    friend class md::System;
};
}  // namespace md::NonBonded
#endif
