// md::System: the evaluation engine of the host.
//
// Same class and public members as the reference's include/System/System.h:
// 33-75 (constructor from AmberTopology, init(), the cutoff/skin/switch/
// tolerance setters and getters, the cell helpers, the rebuild criteria and
// the PME entry points).  The evaluation itself is not done here: it is
// forwarded through the mdgpu C ABI (include/mdgpu.h) to hand-written HIP
// kernels on the GPU.
//
// Additions: set_semantics/set_precision/set_pme_grid/set_pme_order/
// set_device, evaluate(), energies(), forces(), virial(), and minimize() (the
// reference's commented-out next step, apps/main.cpp:28).
#ifndef MD_AMD_SYSTEM_H
#define MD_AMD_SYSTEM_H

#include "AmberTopology/AmberTopology.h"
#include "mdgpu.h"

#include <array>
#include <string_view>
#include <vector>

namespace md {

enum class Precision { FP64 = 0, MIXED = 1 };

class System {
public:
    // Runs the reference's evaluation sequence with its hard-coded defaults
    // (LJ/EE cutoff 10, skins 2, switch 9, Ewald tolerance 1e-6;
    // src/System/init.cpp:10-81) in REF semantics and prints the same report.
    void init();

    explicit System(AmberTopology::AmberTopology topology);
    System(const System&) = delete;
    System& operator=(const System&) = delete;
    ~System();

    [[nodiscard]] const AmberTopology::AmberTopology& get_topology() const { return topology_; }

    void set_lj_cutoff(double rc) { lj_cutoff_ = rc; lj_cutoff2_ = rc * rc; }
    void set_box(double Lx, double Ly, double Lz);
    void set_lj_switch_radius(double rswitch);
    void set_use_lj_switch(bool use) { lj_use_switch_ = use; }
    void set_lj_skin_(double skin);
    void set_lj_list_cutoff(double cutoff);
    void set_ee_cutoff(double cutoff);
    void set_ee_skin(double skin);
    void set_ewald_error_tolerance(double tolerance) { ewald_error_tolerance_ = tolerance; }

    [[nodiscard]] double get_lj_list_cutoff() const { return lj_list_cutoff_; }
    [[nodiscard]] double get_ee_list_cutoff() const { return ee_list_cutoff_; }
    [[nodiscard]] double get_lj_skin() const { return lj_skin_; }
    [[nodiscard]] int get_Cell_ID_lj(int ix, int iy, int iz) const;
    [[nodiscard]] int get_Cell_ID_ee(int ix, int iy, int iz) const;
    [[nodiscard]] static int cell_number_along_dim(double x, double L, int ncelldim);
    [[nodiscard]] bool ljpairs_needs_rebuild();
    [[nodiscard]] bool eepairs_needs_rebuild();

    // The reference plans FFTW here (PBC.cpp:404-444).  The GPU pipeline owns
    // its FFT; init_pme_fft() validates and enables the PME grid and
    // perform_forward_fft() runs one evaluation.
    void init_pme_fft();
    void perform_forward_fft();

    // ---- additions -------------------------------------------------------
    void set_semantics(Semantics s) { semantics_ = s; }
    void set_precision(Precision p) { precision_ = p; }
    void set_pme_grid(int nx, int ny, int nz) { pme_grid_ = {nx, ny, nz}; }
    void set_pme_order(int order) { pme_order_ = order; }
    void set_ewald_alpha(double alpha) { ewald_alpha_ = alpha; alpha_fixed_ = true; }
    void set_device(int device) { device_ = device; }
    void set_coordinates(const std::vector<double>& xyz);
    // One energy + force evaluation on the GPU with the current settings.
    void evaluate();
    // L-BFGS energy minimisation of the coordinates on the GPU (the
    // reference's commented-out sys.minimize(), apps/main.cpp:28), then one
    // evaluation at the minimised coordinates.  With a skin set
    // (set_lj_skin_ / set_ee_skin) the pair list is a Verlet list kept across
    // the minimiser's evaluations.
    void minimize();
    void set_minimize_options(const mdgpu_min_opts& o) { min_opts_ = o; }
    [[nodiscard]] const mdgpu_min_result& minimize_result() const { return min_result_; }
    [[nodiscard]] const mdgpu_energies& energies() const { return energies_; }
    [[nodiscard]] const std::vector<double>& forces() const { return forces_; }
    [[nodiscard]] const std::array<double, 9>& virial() const { return virial_; }
    [[nodiscard]] double get_ewald_alpha() const { return ewald_alpha_; }
    [[nodiscard]] std::array<int, 3> get_pme_grid() const { return pme_grid_; }

private:
    void init_box();
    void calculate_ewald_alpha();
    void calculate_ewald_nnodes();
    void setup_neighbor_rebuild_threshold_lj();
    void setup_neighbor_rebuild_threshold_ee();
    void copy_ref_coords_lj();
    void copy_ref_coords_ee();
    void apply_min_image(double& dx, double& dy, double& dz) const;
    static double min_image_1d(double d, double L);
    double calculate_center_of_mass_along_direction(std::string_view direction);
    void ensure_device_system();
    void apply_settings();
    [[noreturn]] void fail(int status, const char* what) const;

    AmberTopology::AmberTopology topology_;
    std::vector<double> forces_;
    mdgpu_energies energies_{};
    std::array<double, 9> virial_{};
    mdgpu_min_opts min_opts_{500, 1e-2, 0.2, 8};
    mdgpu_min_result min_result_{};

    Semantics semantics_;
    Precision precision_ = Precision::FP64;
    int device_ = 0;
    mdgpu_ctx* ctx_ = nullptr;
    bool uploaded_ = false;

    double lj_cutoff_ = 10.0, lj_cutoff2_ = 100.0;
    double lj_skin_ = 0.0, lj_list_cutoff_ = 0.0, lj_list_cutoff2_ = 0.0;
    double ee_cutoff_ = 10.0, ee_cutoff2_ = 100.0;
    double ee_skin_ = 0.0, ee_list_cutoff_ = 0.0, ee_list_cutoff2_ = 0.0;
    double lj_switch_ = 0.0;
    bool lj_use_switch_ = false;
    double ewald_error_tolerance_ = 1e-6;
    double ewald_alpha_ = 0.0;
    bool alpha_fixed_ = false;
    std::array<int, 3> pme_grid_{0, 0, 0};
    int pme_order_ = 4;

    bool pbc_enabled_ = false;
    double boxLx_ = 0.0, boxLy_ = 0.0, boxLz_ = 0.0;
    int nCells_x_lj_ = 0, nCells_y_lj_ = 0, nCells_z_lj_ = 0;
    int nCells_x_ee_ = 0, nCells_y_ee_ = 0, nCells_z_ee_ = 0;
    double rebuild_disp2_lj_ = 0.0, rebuild_disp2_ee_ = 0.0;
    std::vector<double> coordinates_ref_lj_, coordinates_ref_ee_;
};

}  // namespace md

#endif
