// md::System: host side of the evaluation.  Parameter handling mirrors the
// reference (src/System/NonBondedUtils.cpp:11-47, PBC.cpp:12-99, init.cpp);
// every energy and force is computed by the GPU through the mdgpu C ABI.
#include "System/System.h"

#include "desc.h"

#include <algorithm>
#include <cctype>
#include <cmath>
#include <iostream>
#include <limits>
#include <stdexcept>
#include <string>

namespace md {

namespace {
// The GPU pipeline keeps one upload per System; the flattened description
// must outlive the upload call only.
bool is_2a3b(int n) {
    if (n < 2) return false;
    while (n % 2 == 0) n /= 2;
    while (n % 3 == 0) n /= 3;
    return n == 1;
}
// the smallest 2^a 3^b >= n (the lengths the GPU FFT handles)
int next_2a3b(int n) {
    int k = std::max(n, 2);
    while (!is_2a3b(k)) ++k;
    return k;
}
}  // namespace

System::System(AmberTopology::AmberTopology topology)
    : topology_(std::move(topology)), semantics_(topology_.get_semantics()) {
    forces_.assign(3 * topology_.get_nAtoms(), 0.0);
    const auto b = topology_.get_box_dimensions();
    set_box(b[0], b[1], b[2]);
}

System::~System() { mdgpu_destroy(ctx_); }

void System::fail(int status, const char* what) const {
    std::string msg = std::string(what) + " failed (status " + std::to_string(status) + ")";
    if (ctx_) msg += ": " + std::string(mdgpu_last_error(ctx_));
    throw std::runtime_error(msg);
}

void System::set_box(double Lx, double Ly, double Lz) {
    boxLx_ = Lx;
    boxLy_ = Ly;
    boxLz_ = Lz;
    pbc_enabled_ = Lx > 0 && Ly > 0 && Lz > 0;
}

void System::set_lj_switch_radius(double rswitch) {
    lj_switch_ = rswitch;
    lj_use_switch_ = rswitch > 0.0;
    if (lj_use_switch_ && !(lj_switch_ < lj_cutoff_)) throw std::runtime_error("lj_switch must be < lj_cutoff");
}

void System::set_lj_skin_(double skin) {
    lj_skin_ = skin;
    lj_list_cutoff_ = lj_skin_ + lj_cutoff_;
    lj_list_cutoff2_ = lj_list_cutoff_ * lj_list_cutoff_;
}

void System::set_lj_list_cutoff(double cutoff) {
    lj_list_cutoff_ = cutoff;
    lj_list_cutoff2_ = cutoff * cutoff;
}

void System::set_ee_cutoff(double cutoff) {
    ee_cutoff_ = cutoff;
    ee_cutoff2_ = cutoff * cutoff;
}

void System::set_ee_skin(double skin) {
    ee_skin_ = skin;
    ee_list_cutoff_ = ee_skin_ + ee_cutoff_;
    ee_list_cutoff2_ = ee_list_cutoff_ * ee_list_cutoff_;
}

int System::get_Cell_ID_lj(int ix, int iy, int iz) const { return ix + nCells_x_lj_ * (iy + nCells_y_lj_ * iz); }
int System::get_Cell_ID_ee(int ix, int iy, int iz) const { return ix + nCells_x_ee_ * (iy + nCells_y_ee_ * iz); }

int System::cell_number_along_dim(double x, double L, int n) {
    int ix = static_cast<int>(x * n / L);
    if (ix >= n) ix = n - 1;
    if (ix < 0) ix = 0;
    return ix;
}

void System::init_box() {
    if (!pbc_enabled_) return;
    auto cells = [](double L, double rl) { return std::max(1, static_cast<int>(std::floor(L / rl))); };
    nCells_x_lj_ = cells(boxLx_, lj_list_cutoff_);
    nCells_y_lj_ = cells(boxLy_, lj_list_cutoff_);
    nCells_z_lj_ = cells(boxLz_, lj_list_cutoff_);
    nCells_x_ee_ = cells(boxLx_, ee_list_cutoff_);
    nCells_y_ee_ = cells(boxLy_, ee_list_cutoff_);
    nCells_z_ee_ = cells(boxLz_, ee_list_cutoff_);
}

double System::min_image_1d(double d, double L) { return d - L * std::round(d / L); }

void System::apply_min_image(double& dx, double& dy, double& dz) const {
    if (!pbc_enabled_) return;
    dx = min_image_1d(dx, boxLx_);
    dy = min_image_1d(dy, boxLy_);
    dz = min_image_1d(dz, boxLz_);
}

void System::setup_neighbor_rebuild_threshold_lj() {
    rebuild_disp2_lj_ = 0.25 * lj_skin_ * lj_skin_;
    coordinates_ref_lj_.resize(3 * topology_.get_nAtoms());
}
void System::setup_neighbor_rebuild_threshold_ee() {
    rebuild_disp2_ee_ = 0.25 * ee_skin_ * ee_skin_;
    coordinates_ref_ee_.resize(3 * topology_.get_nAtoms());
}
void System::copy_ref_coords_lj() { coordinates_ref_lj_ = topology_.get_coordinates(); }
void System::copy_ref_coords_ee() { coordinates_ref_ee_ = topology_.get_coordinates(); }

// Largest minimum-image displacement since the last list build against
// (skin/2)^2 (NonBondedUtils.cpp:87-115).
static bool displaced(const std::vector<double>& c, const std::vector<double>& ref, double lim2,
                      const System* self, void (System::*mi)(double&, double&, double&) const) {
    if (ref.size() != c.size()) return true;
    for (size_t i = 0; i + 2 < c.size(); i += 3) {
        double dx = c[i] - ref[i], dy = c[i + 1] - ref[i + 1], dz = c[i + 2] - ref[i + 2];
        (self->*mi)(dx, dy, dz);
        if (dx * dx + dy * dy + dz * dz > lim2) return true;
    }
    return false;
}

bool System::ljpairs_needs_rebuild() {
    return displaced(topology_.get_coordinates(), coordinates_ref_lj_, rebuild_disp2_lj_, this, &System::apply_min_image);
}
bool System::eepairs_needs_rebuild() {
    return displaced(topology_.get_coordinates(), coordinates_ref_ee_, rebuild_disp2_ee_, this, &System::apply_min_image);
}

// alpha = sqrt(-ln(2 tol)) / rc_ee (PBC.cpp:88-91)
void System::calculate_ewald_alpha() {
    if (!alpha_fixed_) ewald_alpha_ = std::sqrt(-std::log(2 * ewald_error_tolerance_)) / ee_cutoff_;
}

// K = int(2 alpha L / (3 tol^0.2)) per axis (PBC.cpp:93-99)
void System::calculate_ewald_nnodes() {
    auto k = [&](double L) {
        return static_cast<int>((2 * ewald_alpha_ * L) / (3 * std::pow(ewald_error_tolerance_, 0.2)));
    };
    if (pme_grid_[0] == 0) pme_grid_ = {k(boxLx_), k(boxLy_), k(boxLz_)};
}

void System::init_pme_fft() {
    for (int k : pme_grid_)
        if (!is_2a3b(k))
            throw std::runtime_error("PME grid dimensions must be of the form 2^a 3^b, got " + std::to_string(k));
}

void System::perform_forward_fft() { evaluate(); }

double System::calculate_center_of_mass_along_direction(std::string_view direction) {
    if (direction.empty()) return std::numeric_limits<double>::quiet_NaN();
    const char ax = static_cast<char>(std::toupper(static_cast<unsigned char>(direction[0])));
    const int off = ax == 'X' ? 0 : (ax == 'Y' ? 1 : (ax == 'Z' ? 2 : -1));
    if (off < 0) return std::numeric_limits<double>::quiet_NaN();
    const auto& c = topology_.get_coordinates();
    const auto& atoms = topology_.get_atom_list();
    double m = 0.0, s = 0.0;
    for (size_t i = 0; i < atoms.size(); ++i) {
        m += atoms[i].get_mass();
        s += c[3 * i + off] * atoms[i].get_mass();
    }
    return m == 0.0 ? std::numeric_limits<double>::quiet_NaN() : s / m;
}

void System::set_coordinates(const std::vector<double>& xyz) {
    if (xyz.size() != 3 * topology_.get_nAtoms()) throw std::runtime_error("set_coordinates: size mismatch");
    topology_.set_coordinates(xyz);
}

void System::ensure_device_system() {
    if (!ctx_) {
        int st = mdgpu_create(device_, &ctx_);
        if (st != MDGPU_OK) {
            ctx_ = nullptr;
            throw std::runtime_error("mdgpu_create failed (status " + std::to_string(st) + ")");
        }
    }
    if (!uploaded_) {
        host::DescStorage d;
        host::build_system_desc(topology_, &d);
        int st = mdgpu_upload_system(ctx_, &d.desc);
        if (st != MDGPU_OK) fail(st, "mdgpu_upload_system");
        uploaded_ = true;
    }
}

// Parameters of the current settings to the device context; the largest
// skin makes the pair list a Verlet list (built at cutoff + skin, rebuilt on
// the GPU once an atom has moved more than skin / 2).
void System::apply_settings() {
    if (!pbc_enabled_) throw std::runtime_error("evaluate: the GPU path needs a periodic box");
    calculate_ewald_alpha();
    ensure_device_system();
    mdgpu_params p{};
    p.semantics = semantics_ == Semantics::AMBER ? MDGPU_SEM_AMBER : MDGPU_SEM_REF;
    p.precision = precision_ == Precision::MIXED ? MDGPU_PREC_MIXED : MDGPU_PREC_FP64;
    p.lj_cutoff = lj_cutoff_;
    p.ee_cutoff = ee_cutoff_;
    p.lj_switch = lj_use_switch_ ? lj_switch_ : 0.0;
    p.ewald_alpha = ewald_alpha_;
    if (semantics_ == Semantics::AMBER) {
        if (pme_grid_[0] == 0) {
            // the reference's formula, each length rounded up to a length the
            // GPU FFT handles (2^a 3^b): a finer grid, never a coarser one
            calculate_ewald_nnodes();
            for (int& k : pme_grid_) k = next_2a3b(k);
        }
        init_pme_fft();
        p.pme_grid[0] = pme_grid_[0];
        p.pme_grid[1] = pme_grid_[1];
        p.pme_grid[2] = pme_grid_[2];
    }
    p.pme_order = pme_order_;
    int st = mdgpu_set_params(ctx_, &p);
    if (st != MDGPU_OK) fail(st, "mdgpu_set_params");
    st = mdgpu_set_verlet(ctx_, std::max({0.0, lj_skin_, ee_skin_}));
    if (st != MDGPU_OK) fail(st, "mdgpu_set_verlet");
    // AMBER semantics take the cell angles of the coordinate file; REF
    // semantics keep the reference's orthorhombic cell (PBC.cpp:72-81)
    const std::array<double, 3> right = {90.0, 90.0, 90.0};
    const std::array<double, 3> ang = semantics_ == Semantics::AMBER ? topology_.get_box_angles() : right;
    st = mdgpu_set_box_angles(ctx_, ang.data());
    if (st != MDGPU_OK) fail(st, "mdgpu_set_box_angles");
}

void System::evaluate() {
    apply_settings();
    const double box[3] = {boxLx_, boxLy_, boxLz_};
    const int st = mdgpu_eval(ctx_, topology_.get_coordinates().data(), box, &energies_, forces_.data(), virial_.data());
    if (st != MDGPU_OK) fail(st, "mdgpu_eval");
}

void System::minimize() {
    apply_settings();
    const double box[3] = {boxLx_, boxLy_, boxLz_};
    std::vector<double> x = topology_.get_coordinates();
    const int st = mdgpu_minimize_host(ctx_, x.data(), box, &min_opts_, &min_result_);
    if (st != MDGPU_OK) fail(st, "mdgpu_minimize_host");
    topology_.set_coordinates(x);
    evaluate();
}

// The reference's System::init() sequence and report (src/System/init.cpp:
// 10-81), in REF semantics with its hard-coded parameters.
void System::init() {
    const auto b = topology_.get_box_dimensions();
    set_box(b[0], b[1], b[2]);
    set_lj_cutoff(10.0);
    set_lj_skin_(2.0);
    set_lj_switch_radius(9.0);
    set_ee_cutoff(10.0);
    set_ee_skin(2.0);
    setup_neighbor_rebuild_threshold_lj();
    setup_neighbor_rebuild_threshold_ee();
    init_box();
    if (const double halfmin = 0.5 * std::min({boxLx_, boxLy_, boxLz_}); pbc_enabled_ && lj_cutoff_ > halfmin)
        std::cerr << "LJ cutoff too large for minimum image.\n";
    copy_ref_coords_lj();
    copy_ref_coords_ee();
    const double comx = calculate_center_of_mass_along_direction("x");
    const double comy = calculate_center_of_mass_along_direction("y");
    const double comz = calculate_center_of_mass_along_direction("z");
    std::cout << "Center of mass along x: " << comx << "\n"
              << "Center of mass along y: " << comy << "\n"
              << "Center of mass along z: " << comz << std::endl;
    // move_box edits a copy in the reference (PBC.cpp:136-149): report only
    const double com[3] = {comx, comy, comz}, L[3] = {boxLx_, boxLy_, boxLz_};
    const char* ax[3] = {"x", "y", "z"};
    for (int d = 0; d < 3; ++d) {
        if (std::abs(com[d] - L[d] / 2) < 1) std::cout << "Not moving box along " << ax[d] << ".\n";
        else std::cout << "Moving box along " << ax[d] << " by " << com[d] - L[d] / 2 << std::endl;
    }
    set_ewald_error_tolerance(1e-6);
    alpha_fixed_ = false;
    calculate_ewald_alpha();
    const auto saved_grid = pme_grid_;
    pme_grid_ = {0, 0, 0};
    calculate_ewald_nnodes();
    std::cout << "nX: " << pme_grid_[0] << " nY: " << pme_grid_[1] << " nZ: " << pme_grid_[2] << std::endl;
    pme_grid_ = saved_grid;
    const Semantics saved = semantics_;
    semantics_ = Semantics::REF;
    evaluate();
    semantics_ = saved;
    const auto& e = energies_;
    std::cout << "\nBond:" << e.bond << "\nAngle: " << e.angle << "\nCosineDihedral: " << e.dihedral
              << "\nUrey-Bradley: " << e.urey_bradley << "\nImpropers: " << e.improper << "\nCMAP energy: " << e.cmap
              << "\nVDW: " << e.vdw << "\nEE: " << e.ee_all << "\nEE pairlist: " << e.ee_pairlist << "\n"
              << " EE_energy_pairlist_cutoff_" << e.ee_rf << "\n"
              << "EE_energy_pairlist_ewald_direct_ " << e.ewald_direct << "\n"
              << "EE_energy_pairlist_ewald_self_" << e.ewald_self << "\n";
    std::cout << "Calculated Forces! \n"
              << forces_[0] << " " << forces_[1] << " " << forces_[2] << "\n"
              << forces_[3] << " " << forces_[4] << " " << forces_[5] << "\n"
              << forces_[6] << " " << forces_[7] << " " << forces_[8] << "\n";
}

}  // namespace md
