// Scalar Coulomb formulas (reference include/NonBonded/CoulombicEE.h:14-23,
// src/NonBonded/CoulombicEE.cpp:10-44).  kCoulomb = 18.2206899283247^2 formed
// in long double, as there.
#ifndef MD_AMD_COULOMBICEE_H
#define MD_AMD_COULOMBICEE_H

namespace md { class System; }

namespace md::NonBonded {
class CoulombicEE {
private:
    static double CalculateEnergy(double r, double qa, double qb, double eps);
    static double CalculateGradient(double r, double qa, double qb, double eps);
    static double CalculateEnergyCutoff(double r, double rc, double qa, double qb, double eps);
    static double CalculateEnergyEwaldDirectTerm(double r, double qa, double qb, double alpha);
    static double CalculateEnergyEwaldDirectTerm_14(double r, double qa, double qb, double alpha, double scale);
    friend class md::System;
};
}  // namespace md::NonBonded
#endif
