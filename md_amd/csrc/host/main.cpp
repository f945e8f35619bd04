// Command-line driver: md_main <prmtop> [rst7] [--print] [--minimize [iterations]] (the
// reference's apps/main.cpp:9-30), running the reference's System::init()
// sequence on the GPU.  --print appends, after init()'s report, every energy
// term, the Ewald alpha and all 3N forces at full precision (%.17g) between
// "#MDPRINT begin" and "#MDPRINT end" lines.  --minimize also runs the step
// the reference leaves commented out
// (sys.minimize(), apps/main.cpp:28) and reports the minimised energies.  The
// minimisation uses AMBER semantics on a second System: REF semantics keep no
// exclusions, so a bonded pair of opposite charges attracts without bound
// and its energy has no minimum.
#include "AmberTopology/AmberTopology.h"
#include "System/System.h"

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <iostream>
#include <vector>

namespace {

// Full-precision report of the last evaluation (energies, alpha, forces).
void print_full(const md::System& sys) {
    const mdgpu_energies& e = sys.energies();
    const std::vector<double>& f = sys.forces();
    const std::pair<const char*, double> terms[] = {
        {"bond", e.bond}, {"angle", e.angle}, {"dihedral", e.dihedral}, {"urey_bradley", e.urey_bradley},
        {"improper", e.improper}, {"cmap", e.cmap}, {"vdw", e.vdw}, {"ee_all", e.ee_all},
        {"ee_pairlist", e.ee_pairlist}, {"ee_rf", e.ee_rf}, {"ewald_direct", e.ewald_direct},
        {"ewald_self", e.ewald_self}, {"ewald_recip", e.ewald_recip}, {"ewald_excl", e.ewald_excl},
        {"vdw14", e.vdw14}, {"ee14", e.ee14}, {"total", e.total}};
    std::fflush(stdout);
    std::cout.flush();
    std::printf("#MDPRINT begin\n");
    for (const auto& t : terms) std::printf("term %s %.17g\n", t.first, t.second);
    std::printf("alpha %.17g\n", sys.get_ewald_alpha());
    std::printf("forces %zu\n", f.size() / 3);
    for (size_t i = 0; i + 2 < f.size(); i += 3) std::printf("%.17g %.17g %.17g\n", f[i], f[i + 1], f[i + 2]);
    std::printf("#MDPRINT end\n");
    std::fflush(stdout);
}

}  // namespace

int main(int argc, char** argv) {
    std::vector<const char*> files;
    bool minimize = false, print = false;
    int max_iter = 500;
    for (int i = 1; i < argc; ++i) {
        if (std::strcmp(argv[i], "--print") == 0) {
            print = true;
        } else if (std::strcmp(argv[i], "--minimize") == 0) {
            minimize = true;
            if (i + 1 < argc && argv[i + 1][0] >= '0' && argv[i + 1][0] <= '9') max_iter = std::atoi(argv[++i]);
        } else {
            files.push_back(argv[i]);
        }
    }
    if (files.empty()) {
        std::cerr << "usage: " << argv[0] << " <prmtop> [rst7] [--print] [--minimize [iterations]]\n";
        return 1;
    }
    const auto top = md::AmberTopology::AmberTopology::read_topology_coordinates(files[0], files.size() > 1 ? files[1] : "");
    md::System sys(top);
    try {
        sys.init();
        if (print) print_full(sys);
        if (minimize) {
            const auto top_amber = md::AmberTopology::AmberTopology::read_topology_coordinates(
                files[0], files.size() > 1 ? files[1] : "", md::Semantics::AMBER);
            md::System msys(top_amber);
            msys.set_lj_cutoff(10.0);
            msys.set_lj_switch_radius(9.0);
            msys.set_ee_cutoff(10.0);
            msys.set_lj_skin_(2.0);
            msys.set_ee_skin(2.0);
            msys.set_ewald_error_tolerance(1e-6);
            msys.set_minimize_options(mdgpu_min_opts{max_iter, 1e-2, 0.2, 8});
            msys.minimize();
            const mdgpu_min_result& r = msys.minimize_result();
            const mdgpu_energies& e = msys.energies();
            std::cout << "\nMinimized: iterations " << r.iterations << " evaluations " << r.evaluations
                      << " converged " << r.converged << "\nEnergy: " << r.energy << "\nRMS force: " << r.rms_force
                      << "\nMax force: " << r.max_force << "\nBond:" << e.bond << "\nAngle: " << e.angle
                      << "\nVDW: " << e.vdw << "\nEwald direct: " << e.ewald_direct
                      << "\nEwald reciprocal: " << e.ewald_recip << std::endl;
        }
    } catch (const std::exception& e) {
        std::cerr << e.what() << "\n";
        return 2;
    }
    return 0;
}
