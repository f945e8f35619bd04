// Fused bonded kernel and the FP64 exclusion / 1-4 correction kernel.
//
// k_bonded replaces, in one launch with one thread per term, the reference's
// calculate_{bond,angle,dihedral,UB,improper,CMAP}_energy
// (src/System/BondedEnergy.cpp:10-166) and calculate_forces_{bonds,UBbonds,
// angles,cosinedihedrals,harmonicImpropers,cmap} (src/System/BondedForces.cpp:
// 10-479).  The threads of a block walk the term lists in the order bonds,
// Urey-Bradley, angles, dihedrals, impropers, CMAP.  Coordinate differences are
// taken in FP64 from the raw coordinates (the reference applies no minimum
// image to bonded terms); the mixed form does the rest in FP32.
//
// The corrections (AMBER semantics, FP64 in both forms; blocks after the
// bonded ones in the same launch) add, with minimum
// image and no cutoff, -k q_i q_j erf(a r)/r for every excluded pair (the
// reciprocal sum counts it; cf. the exclusion loop of NonBondedEnergy.cpp:
// 271-304) and, for every 1-4 pair, the 1-4 LJ (LENNARD_JONES_14 tables) plus
// k q_i q_j (1/(SCEE r) - erf(a r)/r).
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
