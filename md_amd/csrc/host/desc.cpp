// Flattening of the host topology into the mdgpu C-ABI system description
// (the arrays the GPU upload reads; the reference reads the same data from
// AmberTopology inside each System pass, e.g. NonBondedUtils.cpp:134-160).
#include "desc.h"

#include <algorithm>
#include <stdexcept>

namespace md::host {

void build_system_desc(const md::AmberTopology::AmberTopology& t, DescStorage* h) {
    const auto& atoms = t.get_atom_list();
    const int n = static_cast<int>(atoms.size());
    const int nt = t.get_nTypes();
    h->charge.resize(n);
    h->type.resize(n);
    h->mass.resize(n);
    for (int i = 0; i < n; ++i) {
        h->charge[i] = static_cast<double>(atoms[i].get_partial_charge());
        h->type[i] = static_cast<int32_t>(atoms[i].get_atom_type_index()) - 1;
        h->mass[i] = atoms[i].get_mass();
    }
    h->nb_index.assign(static_cast<size_t>(nt) * nt, 0);
    const auto& nb = t.get_nb_matrix();
    for (int i = 0; i < nt; ++i)
        for (int j = 0; j < nt; ++j) h->nb_index[static_cast<size_t>(i) * nt + j] = static_cast<int32_t>(nb[i][j]);
    h->lj_a = t.get_lennard_jones_Acoefs_();
    h->lj_b = t.get_lennard_jones_Bcoefs_();
    h->lj14_a = t.get_lennard_jones_14_Acoefs_();
    h->lj14_b = t.get_lennard_jones_14_Bcoefs_();
    const auto& ty = t.get_type_tables();
    for (const auto& b : t.get_harmonic_bonds()) {
        h->bond_atoms.push_back(b.get_atomA_index());
        h->bond_atoms.push_back(b.get_atomB_index());
        h->bond_type.push_back(b.get_type());
    }
    for (const auto& u : t.get_harmonic_UBs()) {
        h->ub_atoms.push_back(u.get_atomA_index());
        h->ub_atoms.push_back(u.get_atomB_index());
        h->ub_type.push_back(u.get_type());
    }
    for (const auto& a : t.get_harmonic_angles()) {
        h->angle_atoms.push_back(a.get_atomA_index());
        h->angle_atoms.push_back(a.get_atomB_index());
        h->angle_atoms.push_back(a.get_atomC_index());
        h->angle_type.push_back(a.get_type());
    }
    for (const auto& d : t.get_cosine_dihedrals()) {
        h->dih_atoms.push_back(d.get_atomA_index());
        h->dih_atoms.push_back(d.get_atomB_index());
        h->dih_atoms.push_back(d.get_atomC_index());
        h->dih_atoms.push_back(d.get_atomD_index());
        h->dih_type.push_back(d.get_type());
        h->dih_flags.push_back((d.get_improper() ? 2 : 0) | (d.get_exclude_14() ? 1 : 0));
    }
    for (const auto& m : t.get_harmonic_impropers()) {
        h->imp_atoms.push_back(m.get_atomA_index());
        h->imp_atoms.push_back(m.get_atomB_index());
        h->imp_atoms.push_back(m.get_atomC_index());
        h->imp_atoms.push_back(m.get_atomD_index());
        h->imp_type.push_back(m.get_type());
    }
    for (const auto& c : t.get_cmaps()) {
        h->cmap_atoms.push_back(c.get_atomA_index());
        h->cmap_atoms.push_back(c.get_atomB_index());
        h->cmap_atoms.push_back(c.get_atomC_index());
        h->cmap_atoms.push_back(c.get_atomD_index());
        h->cmap_atoms.push_back(c.get_atomE_index());
        h->cmap_set.push_back(c.get_parameter_set());
    }
    const auto& res = t.get_charmm_cmap_resolutions();
    int nsets = 0, cres = 0;
    for (int s = 0; s < 5; ++s) {
        try {
            const auto& g = t.get_cmap_grid_data(s + 1);
            if (s >= static_cast<int>(res.size())) break;
            if (cres == 0) cres = res[s];
            if (res[s] != cres) throw std::runtime_error("CMAP sets of different resolutions are not supported");
            h->cmap_grid.insert(h->cmap_grid.end(), g.begin(), g.begin() + static_cast<long>(cres) * cres);
            ++nsets;
        } catch (const std::out_of_range&) {
            break;
        }
    }
    h->cmap_coeff = t.get_Cmap_Coefficient_Matrix_bicubic_spline();
    h->cmap_coeff.resize(static_cast<size_t>(nsets) * cres * cres * 16);

    // exclusions: AMBER only (REF lists are empty by construction)
    const auto& p14 = t.get_pairs14();
    h->scee14 = t.get_pairs14_scee();
    for (const auto& pr : p14) {
        h->pairs14.push_back(pr[0]);
        h->pairs14.push_back(pr[1]);
    }
    bool any_excl = false;
    for (const auto& a : atoms) any_excl = any_excl || !a.get_excluded_atoms().empty();
    if (t.get_semantics() == md::Semantics::AMBER) {
        std::vector<std::vector<int32_t>> adj(n);
        for (int i = 0; i < n; ++i)
            for (int j : atoms[i].get_excluded_atoms()) {
                if (j == i || j < 0 || j >= n) continue;
                adj[i].push_back(j);
                adj[j].push_back(i);
            }
        for (const auto& pr : p14) {
            adj[pr[0]].push_back(pr[1]);
            adj[pr[1]].push_back(pr[0]);
        }
        h->excl_start.assign(n + 1, 0);
        for (int i = 0; i < n; ++i) {
            auto& v = adj[i];
            std::sort(v.begin(), v.end());
            v.erase(std::unique(v.begin(), v.end()), v.end());
            h->excl_start[i + 1] = h->excl_start[i] + static_cast<int32_t>(v.size());
            h->excl_list.insert(h->excl_list.end(), v.begin(), v.end());
        }
        for (int i = 0; i < n; ++i)
            for (int j : atoms[i].get_excluded_atoms()) {
                if (j <= i || j >= n) continue;
                if (t.is_14_pair(i, j)) continue;
                h->excl_pairs.push_back(i);
                h->excl_pairs.push_back(j);
            }
    }
    (void)any_excl;

    h->coords = t.get_coordinates();
    const auto b = t.get_box_dimensions();
    h->box[0] = b[0];
    h->box[1] = b[1];
    h->box[2] = b[2];

    auto& d = h->desc;
    d.n_atoms = n;
    d.n_types = nt;
    d.charge = h->charge.data();
    d.type = h->type.data();
    d.nb_index = h->nb_index.data();
    d.n_lj = static_cast<int32_t>(h->lj_a.size());
    d.lj_a = h->lj_a.data();
    d.lj_b = h->lj_b.data();
    d.lj14_a = h->lj14_a.data();
    d.lj14_b = h->lj14_b.data();
    d.n_bonds = static_cast<int32_t>(h->bond_type.size());
    d.bond_atoms = h->bond_atoms.data();
    d.bond_type = h->bond_type.data();
    d.n_bond_types = static_cast<int32_t>(ty.bond_k.size());
    d.bond_k = ty.bond_k.data();
    d.bond_r0 = ty.bond_r0.data();
    d.n_ubs = static_cast<int32_t>(h->ub_type.size());
    d.ub_atoms = h->ub_atoms.data();
    d.ub_type = h->ub_type.data();
    d.n_ub_types = static_cast<int32_t>(ty.ub_k.size());
    d.ub_k = ty.ub_k.data();
    d.ub_r0 = ty.ub_r0.data();
    d.n_angles = static_cast<int32_t>(h->angle_type.size());
    d.angle_atoms = h->angle_atoms.data();
    d.angle_type = h->angle_type.data();
    d.n_angle_types = static_cast<int32_t>(ty.angle_k.size());
    d.angle_k = ty.angle_k.data();
    d.angle_theta0 = ty.angle_theta0.data();
    d.n_dihedrals = static_cast<int32_t>(h->dih_type.size());
    d.dihedral_atoms = h->dih_atoms.data();
    d.dihedral_type = h->dih_type.data();
    d.n_dihedral_types = static_cast<int32_t>(ty.dih_k.size());
    d.dihedral_k = ty.dih_k.data();
    d.dihedral_n = ty.dih_n.data();
    d.dihedral_phase = ty.dih_phase.data();
    d.n_impropers = static_cast<int32_t>(h->imp_type.size());
    d.improper_atoms = h->imp_atoms.data();
    d.improper_type = h->imp_type.data();
    d.n_improper_types = static_cast<int32_t>(ty.imp_k.size());
    d.improper_k = ty.imp_k.data();
    d.improper_psi0 = ty.imp_psi0.data();
    d.n_cmaps = static_cast<int32_t>(h->cmap_set.size());
    d.cmap_atoms = h->cmap_atoms.data();
    d.cmap_set = h->cmap_set.data();
    d.cmap_n_sets = nsets;
    d.cmap_resolution = cres;
    d.cmap_coeff = h->cmap_coeff.data();
    d.cmap_grid = h->cmap_grid.data();
    d.excl_start = h->excl_start.empty() ? nullptr : h->excl_start.data();
    d.excl_list = h->excl_list.data();
    d.n_excl_pairs = static_cast<int32_t>(h->excl_pairs.size() / 2);
    d.excl_pairs = h->excl_pairs.data();
    d.n_pairs14 = static_cast<int32_t>(h->scee14.size());
    d.pairs14 = h->pairs14.data();
    d.scee14 = h->scee14.data();
}

}  // namespace md::host
