// Flattened system description built from the host topology.
#ifndef MD_AMD_HOST_DESC_H
#define MD_AMD_HOST_DESC_H

#include "AmberTopology/AmberTopology.h"
#include "mdgpu.h"

#include <cstdint>
#include <vector>

namespace md::host {

struct DescStorage {
    mdgpu_system_desc desc{};
    std::vector<double> charge, lj_a, lj_b, lj14_a, lj14_b, cmap_coeff, cmap_grid, scee14, coords, mass;
    std::vector<int32_t> type, nb_index, bond_atoms, bond_type, ub_atoms, ub_type, angle_atoms, angle_type,
        dih_atoms, dih_type, imp_atoms, imp_type, cmap_atoms, cmap_set, excl_start, excl_list, excl_pairs, pairs14,
        dih_flags;
    double box[3] = {0, 0, 0};
};

// Fills *h from t.  Pointers in h->desc refer to h's vectors and to t's
// parameter tables, so both must outlive any use of h->desc.
void build_system_desc(const md::AmberTopology::AmberTopology& t, DescStorage* h);

}  // namespace md::host

#endif
