// Amber/CHARMM prmtop + rst7 topology of the host.
//
// Same class, namespace and public accessors as the reference's
// include/AmberTopology/AmberTopology.h:21-70, so that
//   auto top = md::AmberTopology::AmberTopology::read_topology_coordinates(p, r);
//   md::System sys(top);
// works unchanged.  One addition: a Semantics argument.  Semantics::REF
// reproduces the term lists the reference actually builds (last entry of each
// list dropped by its `i + stride < size` loops, CreateBondedFromParsed.cpp:
// 11,74,100,126,152,178,220,262; exclusion lists left empty by the `auto` copy
// at CreateNonBondedFromParsed.cpp:33-36).  Semantics::AMBER builds complete
// lists and fills the exclusions.
#ifndef MD_AMD_AMBERTOPOLOGY_H
#define MD_AMD_AMBERTOPOLOGY_H

#include "ChemicalEntity/Atom.h"
#include "Bonded/Terms.h"

#include <array>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <unordered_map>
#include <unordered_set>
#include <vector>

namespace md {

enum class Semantics { REF = 0, AMBER = 1 };

namespace AmberTopology {

class AmberTopology {
public:
    AmberTopology() = default;
    explicit AmberTopology(const std::vector<Atom>& atoms) : atom_list_(atoms), nAtoms_(atoms.size()) {}

    // Reads a prmtop and (optionally) an rst7.  As in the reference
    // (AmberTopology.cpp:86-101, IO.cpp:45-53) a file that cannot be opened
    // prints "Error opening file: <path>" and exits with status 1.
    static AmberTopology read_topology_coordinates(const std::string& parmtop_path,
                                                   const std::string& coords_path = "",
                                                   Semantics semantics = Semantics::REF);
    // Same, but reports failures by throwing std::runtime_error.
    static AmberTopology load(const std::string& parmtop_path, const std::string& coords_path,
                              Semantics semantics);

    const std::vector<double>& get_coordinates() { return coordinates_; }
    [[nodiscard]] const std::vector<double>& get_coordinates() const { return coordinates_; }
    // addition: velocities (A/ps) of an AMBER-semantics rst7 that holds them, else empty
    [[nodiscard]] const std::vector<double>& get_velocities() const { return velocities_; }
    void set_coordinates(const std::vector<double>& c) { coordinates_ = c; }
    [[nodiscard]] std::vector<Bonded::HarmonicBond>& get_harmonic_bonds() { return bonds_; }
    [[nodiscard]] std::vector<Bonded::HarmonicAngle>& get_harmonic_angles() { return angles_; }
    [[nodiscard]] std::vector<Bonded::CosineDihedral>& get_cosine_dihedrals() { return dihedrals_; }
    [[nodiscard]] std::vector<Bonded::HarmonicUB>& get_harmonic_UBs() { return ubs_; }
    [[nodiscard]] std::vector<Bonded::HarmonicImproper>& get_harmonic_impropers() { return impropers_; }
    [[nodiscard]] std::vector<Bonded::CMapGroup>& get_cmaps() { return cmaps_; }
    [[nodiscard]] std::vector<Atom>& get_atom_list() { return atom_list_; }
    [[nodiscard]] const std::vector<Bonded::HarmonicBond>& get_harmonic_bonds() const { return bonds_; }
    [[nodiscard]] const std::vector<Bonded::HarmonicAngle>& get_harmonic_angles() const { return angles_; }
    [[nodiscard]] const std::vector<Bonded::CosineDihedral>& get_cosine_dihedrals() const { return dihedrals_; }
    [[nodiscard]] const std::vector<Bonded::HarmonicUB>& get_harmonic_UBs() const { return ubs_; }
    [[nodiscard]] const std::vector<Bonded::HarmonicImproper>& get_harmonic_impropers() const { return impropers_; }
    [[nodiscard]] const std::vector<Bonded::CMapGroup>& get_cmaps() const { return cmaps_; }
    [[nodiscard]] const std::vector<Atom>& get_atom_list() const { return atom_list_; }
    [[nodiscard]] unsigned long get_nAtoms() const { return nAtoms_; }
    [[nodiscard]] size_t get_num_atoms() const { return atom_list_.size(); }
    [[nodiscard]] const std::vector<int>& get_charmm_cmap_resolutions() const { return cmap_resolution_; }
    [[nodiscard]] const std::vector<double>& get_cmap_grid_data(int set_index) const {
        if (set_index < 1 || set_index > static_cast<int>(cmap_grids_.size()) || set_index > 5)
            throw std::out_of_range("CMAP set index must be between 1 and 5. Received: " +
                                    std::to_string(set_index));
        return cmap_grids_[static_cast<size_t>(set_index - 1)];
    }
    [[nodiscard]] std::vector<std::vector<unsigned long>>& get_nb_matrix() { return nbmatrix_; }
    [[nodiscard]] const std::vector<std::vector<unsigned long>>& get_nb_matrix() const { return nbmatrix_; }
    [[nodiscard]] std::vector<double>& get_lennard_jones_Acoefs_() { return lj_a_; }
    [[nodiscard]] std::vector<double>& get_lennard_jones_Bcoefs_() { return lj_b_; }
    [[nodiscard]] std::vector<double>& get_lennard_jones_14_Acoefs_() { return lj14_a_; }
    [[nodiscard]] std::vector<double>& get_lennard_jones_14_Bcoefs_() { return lj14_b_; }
    [[nodiscard]] const std::vector<double>& get_lennard_jones_Acoefs_() const { return lj_a_; }
    [[nodiscard]] const std::vector<double>& get_lennard_jones_Bcoefs_() const { return lj_b_; }
    [[nodiscard]] const std::vector<double>& get_lennard_jones_14_Acoefs_() const { return lj14_a_; }
    [[nodiscard]] const std::vector<double>& get_lennard_jones_14_Bcoefs_() const { return lj14_b_; }
    [[nodiscard]] std::array<double, 3> get_box_dimensions() const noexcept { return {box_x_, box_y_, box_z_}; }
    [[nodiscard]] bool is_14_pair(int i, int j) const;
    [[nodiscard]] double get_scee_scale_for_pair(int i, int j) const;
    [[nodiscard]] std::vector<double>& get_Cmap_Coefficient_Matrix_bicubic_spline() { return cmap_coeff_; }
    [[nodiscard]] const std::vector<double>& get_Cmap_Coefficient_Matrix_bicubic_spline() const { return cmap_coeff_; }

    // Additions (not in the reference).
    [[nodiscard]] Semantics get_semantics() const { return semantics_; }
    // cell angles alpha, beta, gamma (degrees) from the coordinate file's box line
    [[nodiscard]] std::array<double, 3> get_box_angles() const noexcept { return {box_alpha_, box_beta_, box_gamma_}; }
    [[nodiscard]] int get_nTypes() const { return static_cast<int>(nTypes_); }
    // per-type parameter tables and the per-term type index of each list
    struct TypeTables {
        std::vector<double> bond_k, bond_r0, angle_k, angle_theta0, ub_k, ub_r0;
        std::vector<double> dih_k, dih_n, dih_phase, imp_k, imp_psi0;
        std::vector<int> bond_type, angle_type, ub_type, dih_type, imp_type;
    };
    [[nodiscard]] const TypeTables& get_type_tables() const { return types_; }
    // 1-4 pairs (a < d) in first-seen dihedral order with their SCEE divisor
    [[nodiscard]] const std::vector<std::array<int, 2>>& get_pairs14() const { return pairs14_; }
    [[nodiscard]] const std::vector<double>& get_pairs14_scee() const { return pairs14_scee_; }

private:
    void parse_prmtop(const std::string& path);
    void read_coords(const std::string& path);
    void build_terms();
    void build_exclusions();
    void build_cmap_coefficients();
    void build_14_map();

    Semantics semantics_ = Semantics::REF;
    std::vector<Atom> atom_list_;
    std::vector<double> coordinates_;
    unsigned long nAtoms_ = 0, nTypes_ = 0;
    std::unordered_map<std::string, std::vector<std::string>> sections_;

    std::vector<Bonded::HarmonicBond> bonds_;
    std::vector<Bonded::HarmonicAngle> angles_;
    std::vector<Bonded::CosineDihedral> dihedrals_;
    std::vector<Bonded::HarmonicUB> ubs_;
    std::vector<Bonded::HarmonicImproper> impropers_;
    std::vector<Bonded::CMapGroup> cmaps_;
    TypeTables types_;

    std::vector<std::vector<unsigned long>> nbmatrix_;
    std::vector<double> lj_a_, lj_b_, lj14_a_, lj14_b_;
    std::vector<double> scee_, scnb_;
    std::vector<int> cmap_resolution_;
    std::vector<std::vector<double>> cmap_grids_;
    std::vector<double> cmap_coeff_;

    std::unordered_set<uint64_t> lj14_pairs_;
    std::unordered_map<uint64_t, double> scee14_map_;
    std::vector<std::array<int, 2>> pairs14_;
    std::vector<double> pairs14_scee_;

    double box_x_ = 0.0, box_y_ = 0.0, box_z_ = 0.0;
    double box_alpha_ = 90.0, box_beta_ = 90.0, box_gamma_ = 90.0;
    std::vector<double> velocities_;  // 3N A/ps when an AMBER-semantics rst7 holds them, else empty
};

// CMAP patch coefficients of n_sets grids of res x res values (16 per cell,
// the reference's table map, CreateNonBondedFromParsed.cpp:46-117), and the
// transpose of that linear map (gradient with respect to grid values from the
// gradient with respect to coefficients).  Not in the reference's API.
void cmap_patch_coefficients(int res, int n_sets, const double* grid, double* coeff);
void cmap_patch_coefficients_adjoint(int res, int n_sets, const double* d_coeff, double* d_grid);

}  // namespace AmberTopology
}  // namespace md

#endif
