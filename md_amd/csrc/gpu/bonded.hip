// Block counts of the bonded terms and the FP64 exclusion / 1-4 corrections.
//
// Both run as the last blocks of the pair launch (k_nonbonded): the bonded
// blocks after the pair blocks, one thread per term, walking the concatenated
// term lists (bonds, Urey-Bradley, angles, dihedrals, impropers, CMAP; the
// device code is bonded_terms() in bonded_terms.h, replacing
// calculate_{bond,angle,dihedral,UB,improper,CMAP}_energy,
// src/System/BondedEnergy.cpp:10-166, and calculate_forces_*,
// src/System/BondedForces.cpp:10-479), then the correction blocks, one thread
// per excluded or 1-4 pair (correction_pairs(); the exclusion loop of
// NonBondedEnergy.cpp:271-304 and the 1-4 terms).  These helpers size those
// block ranges.
#include "bonded_terms.h"

namespace mdg {

int bonded_rows(const mdgpu_ctx* c) {
    const Topo& t = c->topo;
    const int n = t.n_bonds + t.n_ubs + t.n_angles + t.n_dihedrals + t.n_impropers + t.n_cmaps;
    const int b = pair_block_threads(c);
    return (n + b - 1) / b;
}

int corrections_rows(const mdgpu_ctx* c) {
    const int b = pair_block_threads(c);
    return (c->topo.n_excl_pairs + c->topo.n_pairs14 + b - 1) / b;
}

}  // namespace mdg
