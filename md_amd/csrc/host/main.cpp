// Command-line driver: md_main <prmtop> [rst7] (the reference's apps/main.cpp:
// 9-30), running the reference's System::init() sequence on the GPU.
#include "AmberTopology/AmberTopology.h"
#include "System/System.h"

#include <iostream>

int main(int argc, char** argv) {
    if (argc < 2) {
        std::cerr << "usage: " << argv[0] << " <prmtop> [rst7]\n";
        return 1;
    }
    const auto top = md::AmberTopology::AmberTopology::read_topology_coordinates(argv[1], argc > 2 ? argv[2] : "");
    md::System sys(top);
    try {
        sys.init();
    } catch (const std::exception& e) {
        std::cerr << e.what() << "\n";
        return 2;
    }
    return 0;
}
