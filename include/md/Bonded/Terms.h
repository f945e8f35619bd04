// Bonded term records of the host topology.  Same class names, constructors
// and accessors as the reference's include/Bonded/*.h (HarmonicBond.h:8-63,
// harmonicUB.h:8-58, HarmonicAngle.h:8-64, CosineDihedral.h:8-93,
// HarmonicImproper.h:8-68, CMapGroup.h:12-63), so code written against those
// headers compiles unchanged.  The scalar energy of one term stays private and
// is reachable from md::System, as in the reference; the evaluation itself runs
// on the GPU.
#ifndef MD_AMD_BONDED_TERMS_H
#define MD_AMD_BONDED_TERMS_H

#include <utility>
#include <vector>

namespace md { class System; }

namespace md::Bonded {

class HarmonicBond {
public:
    HarmonicBond(double k, double r0) : k_(k), r0_(r0) {}
    HarmonicBond(double k, double r0, int a, int b, bool isH) : k_(k), r0_(r0), a_(a), b_(b), isH_(isH) {}
    HarmonicBond(int a, int b) : a_(a), b_(b) {}
    void set_energy(double e) { energy_ = e; }
    void set_type(int t) { type_ = t; }
    void set_Bond_force_constant(double k) { k_ = k; }
    void set_Bond_equil_length(double r0) { r0_ = r0; }
    void set_distance(double d) { distance_ = d; }
    [[nodiscard]] int get_type() const { return type_; }
    [[nodiscard]] double get_Bond_force_constant() const { return k_; }
    [[nodiscard]] double get_Bond_equil_length() const { return r0_; }
    [[nodiscard]] int get_atomA_index() const { return a_; }
    [[nodiscard]] int get_atomB_index() const { return b_; }
    [[nodiscard]] bool get_isH() const { return isH_; }
    [[nodiscard]] double get_distance() const { return distance_; }
    [[nodiscard]] double get_energy() const { return energy_; }
private:
    [[nodiscard]] double calculate_energy(double r) const { return k_ * (r - r0_) * (r - r0_); }
    double k_ = 0.0, r0_ = 0.0;
    int a_ = -1, b_ = -1, type_ = 0;
    bool isH_ = false;
    double distance_ = 0.0, energy_ = 0.0;
    friend class ::md::System;
};

class HarmonicUB {
public:
    HarmonicUB(double k, double r0) : k_(k), r0_(r0) {}
    HarmonicUB(double k, double r0, int a, int b) : k_(k), r0_(r0), a_(a), b_(b) {}
    HarmonicUB(int a, int b) : a_(a), b_(b) {}
    void set_energy(double e) { energy_ = e; }
    void set_type(int t) { type_ = t; }
    void set_UB_force_constant(double k) { k_ = k; }
    void set_UB_equil_value(double r0) { r0_ = r0; }
    void set_distance_value(double d) { distance_ = d; }
    [[nodiscard]] int get_type() const { return type_; }
    [[nodiscard]] double get_UB_force_constant() const { return k_; }
    [[nodiscard]] double get_UB_equil_value() const { return r0_; }
    [[nodiscard]] int get_atomA_index() const { return a_; }
    [[nodiscard]] int get_atomB_index() const { return b_; }
private:
    [[nodiscard]] double calculate_energy(double r) const { return k_ * (r - r0_) * (r - r0_); }
    double k_ = 0.0, r0_ = 0.0;
    int a_ = -1, b_ = -1, type_ = 0;
    double distance_ = 0.0, energy_ = 0.0;
    friend class ::md::System;
};

class HarmonicAngle {
public:
    HarmonicAngle(double k, double t0) : k_(k), t0_(t0) {}
    HarmonicAngle(double k, double t0, int a, int b, int c, bool isH)
        : k_(k), t0_(t0), a_(a), b_(b), c_(c), isH_(isH) {}
    HarmonicAngle(int a, int b, int c) : a_(a), b_(b), c_(c) {}
    void set_energy(double e) { energy_ = e; }
    void set_type(int t) { type_ = t; }
    void set_Angle_force_constant(double k) { k_ = k; }
    void set_Angle_equil_angle(double t0) { t0_ = t0; }
    void set_angle(double t) { angle_ = t; }
    [[nodiscard]] double get_Angle_equil_angle() const { return t0_; }
    [[nodiscard]] double get_Angle_force_constant() const { return k_; }
    [[nodiscard]] int get_type() const { return type_; }
    [[nodiscard]] int get_atomA_index() const { return a_; }
    [[nodiscard]] int get_atomB_index() const { return b_; }
    [[nodiscard]] int get_atomC_index() const { return c_; }
    [[nodiscard]] bool get_isH() const { return isH_; }
    [[nodiscard]] double get_angle() const { return angle_; }
    [[nodiscard]] double get_energy() const { return energy_; }
private:
    [[nodiscard]] double calculate_energy(double t) const { return k_ * (t - t0_) * (t - t0_); }
    double k_ = 0.0, t0_ = 0.0;
    int a_ = -1, b_ = -1, c_ = -1, type_ = 0;
    bool isH_ = false;
    double angle_ = 0.0, energy_ = 0.0;
    friend class ::md::System;
};

class CosineDihedral {
public:
    CosineDihedral(double k, double phase, double n, bool improper = false, bool exclude_14 = false)
        : k_(k), phase_(phase), n_(n), improper_(improper), exclude_14_(exclude_14) {}
    CosineDihedral(double k, double phase, double n, int a, int b, int c, int d, bool isH,
                   bool improper = false, bool exclude_14 = false)
        : k_(k), phase_(phase), n_(n), a_(a), b_(b), c_(c), d_(d), isH_(isH),
          improper_(improper), exclude_14_(exclude_14) {}
    CosineDihedral(int a, int b, int c, int d, bool improper = false, bool exclude_14 = false)
        : a_(a), b_(b), c_(c), d_(d), improper_(improper), exclude_14_(exclude_14) {}
    void set_energy(double e) { energy_ = e; }
    void set_type(int t) { type_ = t; }
    void set_Dihedral_force_constant(double k) { k_ = k; }
    void set_Dihedral_phase(double p) { phase_ = p; }
    void set_Dihedral_Periodicity(double n) { n_ = n; }
    void set_cosine_dihedral(double phi) { dihedral_ = phi; }
    [[nodiscard]] int get_type() const { return type_; }
    [[nodiscard]] double get_Dihedral_force_constant() const { return k_; }
    [[nodiscard]] double get_Dihedral_phase() const { return phase_; }
    [[nodiscard]] double get_Dihedral_Periodicity() const { return n_; }
    [[nodiscard]] int get_atomA_index() const { return a_; }
    [[nodiscard]] int get_atomB_index() const { return b_; }
    [[nodiscard]] int get_atomC_index() const { return c_; }
    [[nodiscard]] int get_atomD_index() const { return d_; }
    [[nodiscard]] bool get_isH() const { return isH_; }
    [[nodiscard]] bool get_improper() const { return improper_; }
    [[nodiscard]] bool get_exclude_14() const { return exclude_14_; }
    [[nodiscard]] double get_energy() const { return energy_; }
    [[nodiscard]] double get_dihedral() const { return dihedral_; }
private:
    [[nodiscard]] double calculate_energy(double phi) const;
    double k_ = 0.0, phase_ = 0.0, n_ = 0.0;
    int a_ = -1, b_ = -1, c_ = -1, d_ = -1, type_ = 0;
    bool isH_ = false, improper_ = false, exclude_14_ = false;
    double dihedral_ = 0.0, energy_ = 0.0;
    friend class ::md::System;
};

class HarmonicImproper {
public:
    HarmonicImproper(double k, double psi0) : k_(k), psi0_(psi0) {}
    HarmonicImproper(double k, double psi0, int a, int b, int c, int d)
        : k_(k), psi0_(psi0), a_(a), b_(b), c_(c), d_(d) {}
    HarmonicImproper(int a, int b, int c, int d) : a_(a), b_(b), c_(c), d_(d) {}
    void set_energy(double e) { energy_ = e; }
    void set_type(int t) { type_ = t; }
    void set_IMP_force_constant(double k) { k_ = k; }
    void set_IMP_phase_value(double p) { psi0_ = p; }
    void set_imp_dihedral(double d) { dihedral_ = d; }
    [[nodiscard]] int get_type() const { return type_; }
    [[nodiscard]] double get_IMP_force_constant() const { return k_; }
    [[nodiscard]] double get_IMP_phase_value() const { return psi0_; }
    [[nodiscard]] double get_imp_dihedral() const { return dihedral_; }
    [[nodiscard]] int get_atomA_index() const { return a_; }
    [[nodiscard]] int get_atomB_index() const { return b_; }
    [[nodiscard]] int get_atomC_index() const { return c_; }
    [[nodiscard]] int get_atomD_index() const { return d_; }
    [[nodiscard]] double get_energy() const { return energy_; }
private:
    [[nodiscard]] double calculate_energy(double psi) const { return k_ * (psi - psi0_) * (psi - psi0_); }
    double k_ = 0.0, psi0_ = 0.0;
    int a_ = -1, b_ = -1, c_ = -1, d_ = -1, type_ = 0;
    double dihedral_ = 0.0, energy_ = 0.0;
    friend class ::md::System;
};

class CMapGroup {
public:
    CMapGroup(int set, int a, int b, int c, int d, int e) : set_(set), a_(a), b_(b), c_(c), d_(d), e_(e) {}
    void set_energy(double e) { energy_ = e; }
    void set_parameter_set(int s) { set_ = s; }
    void set_angles(double phi, double psi) { phi_ = phi; psi_ = psi; }
    [[nodiscard]] double get_angle1() const { return phi_; }
    [[nodiscard]] double get_angle2() const { return psi_; }
    [[nodiscard]] int get_atomA_index() const { return a_; }
    [[nodiscard]] int get_atomB_index() const { return b_; }
    [[nodiscard]] int get_atomC_index() const { return c_; }
    [[nodiscard]] int get_atomD_index() const { return d_; }
    [[nodiscard]] int get_atomE_index() const { return e_; }
    [[nodiscard]] int get_parameter_set() const { return set_; }
    [[nodiscard]] double get_energy() const { return energy_; }
private:
    // Bicubic patch energy and (dE/dphi, dE/dpsi) from the 16-coefficient
    // tables (the grid argument is unused, as in the reference).
    [[nodiscard]] double return_energy_bicubic(double phi, double psi, int res,
                                               const std::vector<double>& coeffs,
                                               const std::vector<double>& grid) const;
    [[nodiscard]] std::pair<double, double> return_gradient_bicubic(double phi, double psi, int res,
                                                                    const std::vector<double>& coeffs,
                                                                    const std::vector<double>& grid) const;
    int set_ = 1;
    int a_ = -1, b_ = -1, c_ = -1, d_ = -1, e_ = -1;
    double phi_ = 0.0, psi_ = 0.0, energy_ = 0.0;
    friend class ::md::System;
};

}  // namespace md::Bonded

#endif
