// Golden-vector harness around the reference's own parser and term formulas.
//
// Test infrastructure only: this file is compiled together with the
// reference's sources (AmberTopology/, Bonded/, NonBonded/, System/Metrics.cpp,
// util/, ChemicalEntity/) into oracle/_ref/ref_harness.  The reference's
// System/*.cpp translation units are not built: they include an absolute
// Homebrew FFTW path (include/System/System.h:13) and need libfftw3, neither of
// which exists in this image.  What is built is enough to pin
//   * the parsed topology in the reference's semantics (term lists with the
//     dropped last entries, empty exclusion lists, charge conversion, CMAP
//     coefficient tables, 1-4 pair map), and
//   * every per-term energy formula (bond, angle, UB, cosine dihedral,
//     improper, CMAP energy and gradient, LJ energy and gradient, the four
//     Coulomb forms),
// evaluated exactly as the reference's System loops call them
// (src/System/BondedEnergy.cpp:10-166).
//
// The harness TU is compiled with -Dprivate=public so it can call the
// reference's private formula members; no reference file is modified.
//
// Usage: ref_harness <prmtop> <rst7> <out.json>

#include "AmberTopology/AmberTopology.h"
#include "System/Metrics.h"
#include "NonBonded/LennardJones.h"
#include "NonBonded/CoulombicEE.h"

#include <cstdio>
#include <cmath>
#include <string>
#include <vector>

namespace {

struct Json {
    FILE* f;
    bool first = true;
    void key(const char* k) {
        std::fprintf(f, "%s\n\"%s\": ", first ? "" : ",", k);
        first = false;
    }
    void num(const char* k, double v) { key(k); std::fprintf(f, "%.17g", v); }
    void inum(const char* k, long long v) { key(k); std::fprintf(f, "%lld", v); }
    template <class T>
    void darr(const char* k, const std::vector<T>& v) {
        key(k);
        std::fputc('[', f);
        for (size_t i = 0; i < v.size(); ++i) std::fprintf(f, "%s%.17g", i ? "," : "", (double)v[i]);
        std::fputc(']', f);
    }
    template <class T>
    void iarr(const char* k, const std::vector<T>& v) {
        key(k);
        std::fputc('[', f);
        for (size_t i = 0; i < v.size(); ++i) std::fprintf(f, "%s%lld", i ? "," : "", (long long)v[i]);
        std::fputc(']', f);
    }
};

}  // namespace

int main(int argc, char** argv) {
    if (argc < 4) {
        std::fprintf(stderr, "usage: %s prmtop rst7 out.json\n", argv[0]);
        return 2;
    }
    using md::AmberTopology::AmberTopology;
    AmberTopology top = AmberTopology::read_topology_coordinates(argv[1], argv[2]);
    const std::vector<double>& c = top.get_coordinates();
    FILE* f = std::fopen(argv[3], "w");
    if (!f) return 3;
    std::fputc('{', f);
    Json j{f};

    const auto& atoms = top.get_atom_list();
    std::vector<double> q;
    std::vector<long long> type1, nexcl;
    for (const auto& a : atoms) {
        q.push_back((double)a.get_partial_charge());
        type1.push_back((long long)a.get_atom_type_index());
        nexcl.push_back((long long)a.get_excluded_atoms().size());
    }
    j.inum("n_atoms", (long long)top.get_nAtoms());
    auto box = top.get_box_dimensions();
    j.darr("box", std::vector<double>(box.begin(), box.end()));
    j.darr("charge", q);
    j.iarr("type_index", type1);
    j.iarr("n_excluded_after_parse", nexcl);
    std::vector<long long> nb;
    for (const auto& row : top.get_nb_matrix())
        for (auto v : row) nb.push_back((long long)v);
    j.iarr("nb_matrix", nb);
    j.darr("lj_a", top.get_lennard_jones_Acoefs_());
    j.darr("lj_b", top.get_lennard_jones_Bcoefs_());
    j.darr("lj14_a", top.get_lennard_jones_14_Acoefs_());
    j.darr("lj14_b", top.get_lennard_jones_14_Bcoefs_());

    // bonds: energy loop of BondedEnergy.cpp:10-28
    {
        std::vector<long long> idx;
        std::vector<double> k, r0, e;
        double tot = 0.0;
        for (auto& b : top.get_harmonic_bonds()) {
            int a1 = b.get_atomA_index(), a2 = b.get_atomB_index();
            idx.push_back(a1); idx.push_back(a2);
            k.push_back(b.get_Bond_force_constant()); r0.push_back(b.get_Bond_equil_length());
            double d = Metrics::distance(c[3*a1], c[3*a1+1], c[3*a1+2], c[3*a2], c[3*a2+1], c[3*a2+2]);
            double en = b.calculate_energy(d);
            e.push_back(en); tot += en;
        }
        j.iarr("bond_atoms", idx); j.darr("bond_k", k); j.darr("bond_r0", r0);
        j.darr("bond_e", e); j.num("E_bond", tot);
    }
    // Urey-Bradley: BondedEnergy.cpp:74-95
    {
        std::vector<long long> idx;
        std::vector<double> k, r0, e;
        double tot = 0.0;
        for (auto& u : top.get_harmonic_UBs()) {
            int a1 = u.get_atomA_index(), a2 = u.get_atomB_index();
            idx.push_back(a1); idx.push_back(a2);
            k.push_back(u.get_UB_force_constant()); r0.push_back(u.get_UB_equil_value());
            double d = Metrics::distance(c[3*a1], c[3*a1+1], c[3*a1+2], c[3*a2], c[3*a2+1], c[3*a2+2]);
            double en = u.calculate_energy(d);
            e.push_back(en); tot += en;
        }
        j.iarr("ub_atoms", idx); j.darr("ub_k", k); j.darr("ub_r0", r0);
        j.darr("ub_e", e); j.num("E_ub", tot);
    }
    // angles: BondedEnergy.cpp:30-49
    {
        std::vector<long long> idx;
        std::vector<double> k, t0, e;
        double tot = 0.0;
        for (auto& a : top.get_harmonic_angles()) {
            int a1 = a.get_atomA_index(), a2 = a.get_atomB_index(), a3 = a.get_atomC_index();
            idx.push_back(a1); idx.push_back(a2); idx.push_back(a3);
            k.push_back(a.get_Angle_force_constant()); t0.push_back(a.get_Angle_equil_angle());
            double th = Metrics::angle(c[3*a1], c[3*a1+1], c[3*a1+2], c[3*a2], c[3*a2+1], c[3*a2+2],
                                       c[3*a3], c[3*a3+1], c[3*a3+2]);
            double en = a.calculate_energy(th);
            e.push_back(en); tot += en;
        }
        j.iarr("angle_atoms", idx); j.darr("angle_k", k); j.darr("angle_theta0", t0);
        j.darr("angle_e", e); j.num("E_angle", tot);
    }
    // cosine dihedrals: BondedEnergy.cpp:51-72
    {
        std::vector<long long> idx, flags, types;
        std::vector<double> k, ph, per, e;
        double tot = 0.0;
        for (auto& d : top.get_cosine_dihedrals()) {
            int a1 = d.get_atomA_index(), a2 = d.get_atomB_index(), a3 = d.get_atomC_index(),
                a4 = d.get_atomD_index();
            idx.push_back(a1); idx.push_back(a2); idx.push_back(a3); idx.push_back(a4);
            flags.push_back((d.get_improper() ? 2 : 0) | (d.get_exclude_14() ? 1 : 0));
            types.push_back(d.get_type());
            k.push_back(d.get_Dihedral_force_constant()); ph.push_back(d.get_Dihedral_phase());
            per.push_back(d.get_Dihedral_Periodicity());
            double phi = Metrics::dihedral(c[3*a1], c[3*a1+1], c[3*a1+2], c[3*a2], c[3*a2+1], c[3*a2+2],
                                           c[3*a3], c[3*a3+1], c[3*a3+2], c[3*a4], c[3*a4+1], c[3*a4+2]);
            double en = d.calculate_energy(phi);
            e.push_back(en); tot += en;
        }
        j.iarr("dih_atoms", idx); j.iarr("dih_flags", flags); j.iarr("dih_type", types);
        j.darr("dih_k", k); j.darr("dih_phase", ph); j.darr("dih_n", per);
        j.darr("dih_e", e); j.num("E_dihedral", tot);
    }
    // CHARMM impropers: BondedEnergy.cpp:97-122
    {
        std::vector<long long> idx;
        std::vector<double> k, p0, e;
        double tot = 0.0;
        for (auto& m : top.get_harmonic_impropers()) {
            int a1 = m.get_atomA_index(), a2 = m.get_atomB_index(), a3 = m.get_atomC_index(),
                a4 = m.get_atomD_index();
            idx.push_back(a1); idx.push_back(a2); idx.push_back(a3); idx.push_back(a4);
            k.push_back(m.get_IMP_force_constant()); p0.push_back(m.get_IMP_phase_value());
            double psi = Metrics::dihedral(c[3*a1], c[3*a1+1], c[3*a1+2], c[3*a2], c[3*a2+1], c[3*a2+2],
                                           c[3*a3], c[3*a3+1], c[3*a3+2], c[3*a4], c[3*a4+1], c[3*a4+2]);
            double en = m.calculate_energy(psi);
            e.push_back(en); tot += en;
        }
        j.iarr("imp_atoms", idx); j.darr("imp_k", k); j.darr("imp_psi0", p0);
        j.darr("imp_e", e); j.num("E_improper", tot);
    }
    // CMAP: BondedEnergy.cpp:124-166 and the gradient used by BondedForces.cpp:361-368
    {
        std::vector<long long> idx, sets;
        std::vector<double> e, dphi, dpsi, phis, psis;
        double tot = 0.0;
        const auto& res = top.get_charmm_cmap_resolutions();
        auto& coeff = top.get_Cmap_Coefficient_Matrix_bicubic_spline();
        for (auto& m : top.get_cmaps()) {
            int a[5] = {m.get_atomA_index(), m.get_atomB_index(), m.get_atomC_index(),
                        m.get_atomD_index(), m.get_atomE_index()};
            for (int t = 0; t < 5; ++t) idx.push_back(a[t]);
            sets.push_back(m.get_parameter_set());
            double phi = Metrics::dihedral(c[3*a[0]], c[3*a[0]+1], c[3*a[0]+2], c[3*a[1]], c[3*a[1]+1], c[3*a[1]+2],
                                           c[3*a[2]], c[3*a[2]+1], c[3*a[2]+2], c[3*a[3]], c[3*a[3]+1], c[3*a[3]+2]);
            double psi = Metrics::dihedral(c[3*a[1]], c[3*a[1]+1], c[3*a[1]+2], c[3*a[2]], c[3*a[2]+1], c[3*a[2]+2],
                                           c[3*a[3]], c[3*a[3]+1], c[3*a[3]+2], c[3*a[4]], c[3*a[4]+1], c[3*a[4]+2]);
            int r = res[m.get_parameter_set() - 1];
            const auto& grid = top.get_cmap_grid_data(m.get_parameter_set());
            double en = m.return_energy_bicubic(phi, psi, r, coeff, grid);
            auto g = m.return_gradient_bicubic(phi, psi, r, coeff, grid);
            phis.push_back(phi); psis.push_back(psi);
            e.push_back(en); dphi.push_back(g.first); dpsi.push_back(g.second);
            tot += en;
        }
        j.iarr("cmap_atoms", idx); j.iarr("cmap_set", sets);
        j.iarr("cmap_resolution", std::vector<int>(res.begin(), res.end()));
        j.darr("cmap_coeff", coeff);
        j.darr("cmap_phi", phis); j.darr("cmap_psi", psis);
        j.darr("cmap_e", e); j.darr("cmap_dEdphi", dphi); j.darr("cmap_dEdpsi", dpsi);
        j.num("E_cmap", tot);
    }
    // 1-4 pair map (AmberTopology.h:68-69, CreateNonBondedFromParsed.cpp:124-169)
    {
        std::vector<long long> pairs;
        std::vector<double> scee;
        for (auto& d : top.get_cosine_dihedrals()) {
            int a = d.get_atomA_index(), b = d.get_atomD_index();
            if (top.is_14_pair(a, b)) {
                pairs.push_back(a); pairs.push_back(b);
                scee.push_back(top.get_scee_scale_for_pair(a, b));
            }
        }
        j.iarr("pairs14_from_dihedrals", pairs);
        j.darr("scee14_from_dihedrals", scee);
    }
    // scalar formula samples (LennardJones.cpp:9-27, CoulombicEE.cpp:10-44)
    {
        const double r2s[] = {0.81, 1.7, 4.0, 7.3, 9.61, 16.0, 37.2, 63.9, 99.5, 143.0};
        const double abs_[][2] = {{582000.0, 595.0}, {1.3e6, 1020.0}, {7.9e4, 95.5}};
        std::vector<double> lje, ljg, ce, cg, ccut, cew, cew14, rs;
        for (double r2 : r2s) {
            double r = std::sqrt(r2);
            rs.push_back(r2);
            for (auto& ab : abs_) {
                lje.push_back(md::NonBonded::LennardJones::CalculateEnergy(r2, ab[0], ab[1]));
                ljg.push_back(md::NonBonded::LennardJones::CalculateGradient(r2, ab[0], ab[1]));
            }
            ce.push_back(md::NonBonded::CoulombicEE::CalculateEnergy(r, -0.834, 0.417, 1.0));
            cg.push_back(md::NonBonded::CoulombicEE::CalculateGradient(r, -0.834, 0.417, 1.0));
            ccut.push_back(md::NonBonded::CoulombicEE::CalculateEnergyCutoff(r, 10.0, -0.834, 0.417, 1.0));
            cew.push_back(md::NonBonded::CoulombicEE::CalculateEnergyEwaldDirectTerm(r, -0.834, 0.417, 0.3123));
            cew14.push_back(md::NonBonded::CoulombicEE::CalculateEnergyEwaldDirectTerm_14(r, -0.834, 0.417, 0.3123, 1.2));
        }
        j.darr("sample_r2", rs);
        j.darr("sample_lj_ab", std::vector<double>{582000.0, 595.0, 1.3e6, 1020.0, 7.9e4, 95.5});
        j.darr("sample_lj_e", lje); j.darr("sample_lj_g", ljg);
        j.darr("sample_coul_e", ce); j.darr("sample_coul_g", cg);
        j.darr("sample_coul_cut_e", ccut); j.darr("sample_ewald_e", cew);
        j.darr("sample_ewald14_e", cew14);
    }
    std::fputs("\n}\n", f);
    std::fclose(f);
    return 0;
}
