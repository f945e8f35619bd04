// Command-line driver: md_main <prmtop> [rst7] [--minimize [iterations]] (the reference's
// apps/main.cpp:9-30), running the reference's System::init() sequence on the
// GPU.  --minimize also runs the step the reference leaves commented out
// (sys.minimize(), apps/main.cpp:28) and reports the minimised energies.  The
// minimisation uses AMBER semantics on a second System: REF semantics keep no
// exclusions, so a bonded pair of opposite charges attracts without bound
// and its energy has no minimum.
#include "AmberTopology/AmberTopology.h"
#include "System/System.h"

#include <cstdlib>
#include <cstring>
#include <iostream>
#include <vector>

int main(int argc, char** argv) {
    std::vector<const char*> files;
    bool minimize = false;
    int max_iter = 500;
    for (int i = 1; i < argc; ++i) {
        if (std::strcmp(argv[i], "--minimize") == 0) {
            minimize = true;
            if (i + 1 < argc && argv[i + 1][0] >= '0' && argv[i + 1][0] <= '9') max_iter = std::atoi(argv[++i]);
        } else {
            files.push_back(argv[i]);
        }
    }
    if (files.empty()) {
        std::cerr << "usage: " << argv[0] << " <prmtop> [rst7] [--minimize [iterations]]\n";
        return 1;
    }
    const auto top = md::AmberTopology::AmberTopology::read_topology_coordinates(files[0], files.size() > 1 ? files[1] : "");
    md::System sys(top);
    try {
        sys.init();
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
