// Atom record of the host topology (same API as the reference's
// include/ChemicalEntity/Atom.h:9-61).  The charge is kept as long double and
// is set from a double, as in the reference (Atom.h:23,41).
#ifndef MD_AMD_ATOM_H
#define MD_AMD_ATOM_H

#include <string>
#include <utility>
#include <vector>

class Atom {
public:
    Atom() = default;
    explicit Atom(std::string name) : name_(std::move(name)) {}

    void set_partial_charge(const double charge) { partial_charge_ = charge; }
    void set_type(const std::string& type) { atomtype_ = type; }
    void set_residue_name(const std::string& n) { residue_name_ = n; }
    void set_residue_number(const int n) { residue_number_ = static_cast<unsigned long>(n); }
    void set_atomic_number(const int z) { atomic_number_ = z; }
    void set_mass(const double m) { mass_ = m; }
    void set_element(const std::string& e) { element_ = e; }
    void set_atom_type_index(const unsigned long t) { atom_type_index_ = t; }
    void set_nExcluded_Atoms(const int n) { nExcluded_Atoms_ = n; }

    [[nodiscard]] const std::string& get_type() const { return atomtype_; }
    [[nodiscard]] const std::string& get_name() const { return name_; }
    [[nodiscard]] const std::string& get_residue_name() const { return residue_name_; }
    [[nodiscard]] unsigned long get_residue_number() const { return residue_number_; }
    [[nodiscard]] int get_atomic_number() const { return atomic_number_; }
    [[nodiscard]] int get_nExcluded_Atoms() const { return nExcluded_Atoms_; }
    [[nodiscard]] unsigned long get_atom_type_index() const { return atom_type_index_; }
    [[nodiscard]] long double get_partial_charge() const { return partial_charge_; }
    [[nodiscard]] std::vector<int>& get_excluded_atoms() { return excluded_atoms_; }
    [[nodiscard]] const std::vector<int>& get_excluded_atoms() const { return excluded_atoms_; }
    [[nodiscard]] double get_mass() const { return mass_; }
    [[nodiscard]] const std::string& get_element() const { return element_; }

private:
    long double partial_charge_ = 0.0L;
    std::string atomtype_;
    std::string name_;
    std::string residue_name_;
    unsigned long residue_number_ = 0;
    int atomic_number_ = 0;
    double mass_ = 0.0;
    std::string element_;
    unsigned long atom_type_index_ = 0;
    int nExcluded_Atoms_ = 0;
    std::vector<int> excluded_atoms_;
};

#endif
