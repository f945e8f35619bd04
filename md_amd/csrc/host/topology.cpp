// prmtop / rst7 reader of the host topology.
//
// Sections are read into token lists first (numeric sections split on blanks,
// as the reference does in src/util/IO.cpp:109-136; name sections in 4-column
// fields, IO.cpp:76-83) and the term lists are assembled afterwards, so the
// result does not depend on section order.  Semantics::REF reproduces the
// reference's term-list construction (src/AmberTopology/CreateBondedFromParsed.
// cpp, CreateNonBondedFromParsed.cpp); Semantics::AMBER builds complete lists.
#include "AmberTopology/AmberTopology.h"
#include "IO/Restart.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <iostream>
#include <sstream>

namespace md::AmberTopology {

namespace {

// prmtop CHARGE unit: charges are divided by 18.2206899283247 in long double
// before being stored as double (TopologyParser.cpp:49-55, Atom.h:23).
constexpr long double kChargeUnit = 18.2206899283247L;

bool is_fixed_width_section(const std::string& name) {
    return name == "ATOM_NAME" || name == "RESIDUE_LABEL" || name == "AMBER_ATOM_TYPE" ||
           name == "TITLE";
}

std::string trim_right(std::string s) {
    while (!s.empty() && std::isspace(static_cast<unsigned char>(s.back()))) s.pop_back();
    return s;
}

bool blank(const std::string& s) {
    return std::all_of(s.begin(), s.end(), [](unsigned char c) { return std::isspace(c) != 0; });
}

// Field width of a "%FORMAT(...)" line: (10I8) -> 8, (5E16.8) -> 16,
// (20a4) -> 4, (8(F9.5)) -> 9; 0 when none can be read.
int format_width(const std::string& line) {
    size_t i = line.find('(');
    if (i == std::string::npos) return 0;
    ++i;
    while (i < line.size() && std::isdigit(static_cast<unsigned char>(line[i]))) ++i;
    if (i < line.size() && line[i] == '(') ++i;
    if (i >= line.size() || !std::isalpha(static_cast<unsigned char>(line[i]))) return 0;
    ++i;
    int w = 0;
    while (i < line.size() && std::isdigit(static_cast<unsigned char>(line[i]))) w = 10 * w + (line[i++] - '0');
    return w;
}

// The values of one numeric line of a section with field width w.  When
// every blank-separated run of characters ends on a field boundary (the
// right-justified layout the FORMAT describes), the line is cut into fields
// of w columns, which also separates fields that touch ("-1234567-1234567");
// otherwise (a FORMAT that does not describe the line) the runs are the values.
void numeric_fields(const std::string& line, int w, std::vector<std::string>& tok) {
    bool aligned = w > 0;
    for (size_t i = 0; aligned && i < line.size();) {
        if (std::isspace(static_cast<unsigned char>(line[i]))) {
            ++i;
            continue;
        }
        size_t j = i;
        while (j < line.size() && !std::isspace(static_cast<unsigned char>(line[j]))) ++j;
        aligned = j % static_cast<size_t>(w) == 0;
        i = j;
    }
    if (aligned) {
        for (size_t i = 0; i < line.size(); i += static_cast<size_t>(w)) {
            std::string f = line.substr(i, static_cast<size_t>(w));
            f.erase(std::remove_if(f.begin(), f.end(), [](unsigned char c) { return std::isspace(c); }), f.end());
            if (!f.empty()) tok.push_back(f);
        }
        return;
    }
    std::istringstream ss(line);
    std::string t;
    while (ss >> t) tok.push_back(t);
}

uint64_t pair_key(int a, int b) {
    if (a > b) std::swap(a, b);
    return (static_cast<uint64_t>(static_cast<uint32_t>(a)) << 32) | static_cast<uint32_t>(b);
}

// Patch-coefficient table of the reference's CMAP (CreateNonBondedFromParsed.
// cpp:48-64): a = T * e / 16 with e the 4x4 grid neighbourhood ordered
// e[4*c + r] = grid(row + r - 1, col + c - 1).  It is the definition of the
// CMAP energy surface and is reproduced exactly.
constexpr int kCmapTable[16][16] = {
    {0, 0, 0, 0, 0, 16, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0},
    {0, 0, 0, 0, -8, 0, 8, 0, 0, 0, 0, 0, 0, 0, 0, 0},
    {0, 0, 0, 0, 16, -40, 32, -8, 0, 0, 0, 0, 0, 0, 0, 0},
    {0, 0, 0, 0, -8, 24, -24, 8, 0, 0, 0, 0, 0, 0, 0, 0},
    {0, -8, 0, 0, 0, 0, 0, 0, 0, 8, 0, 0, 0, 0, 0, 0},
    {0, -4, 0, 0, -4, 4, 0, 0, 0, 0, 4, 0, 0, 0, 0, 0},
    {0, 32, -20, 0, 8, -4, -4, 0, 0, -24, 16, -4, 0, 0, 0, 0},
    {0, -20, 12, 0, -4, 0, 4, 0, 0, 16, -12, 4, 0, 0, 0, 0},
    {0, 16, 0, 0, 0, -40, 0, 0, 0, 32, 0, 0, 0, -8, 0, 0},
    {0, 8, 0, 0, 32, -4, -24, 0, -20, -4, 0, 0, 0, 0, -4, 0},
    {0, -64, 40, 0, -64, 96, -68, 24, 40, -68, -16, -16, 0, 24, -16, 0},
    {0, 40, -24, 0, 32, -52, 52, -24, -20, 40, 16, 16, 0, -16, 12, 0},
    {0, -8, 0, 0, 0, 24, 0, 0, 0, -24, 0, 0, 0, 8, 0, 4},
    {0, -4, 0, 0, -20, 0, 16, 0, 12, 4, -12, 0, 0, 0, 4, -4},
    {0, 32, -20, 0, 40, -52, 40, -16, -24, 52, -52, 12, 0, -24, 16, -4},
    {0, -20, 12, 0, -20, 28, -32, 16, 12, -32, 40, -12, 0, 16, -12, 4}};

}  // namespace

AmberTopology AmberTopology::read_topology_coordinates(const std::string& parmtop_path,
                                                       const std::string& coords_path,
                                                       Semantics semantics) {
    for (const std::string* p : {&parmtop_path, &coords_path}) {
        if (p->empty()) continue;
        std::ifstream probe(*p);
        if (!probe) {
            std::cerr << "Error opening file: " << *p << "\n";
            std::exit(1);
        }
    }
    return load(parmtop_path, coords_path, semantics);
}

AmberTopology AmberTopology::load(const std::string& parmtop_path, const std::string& coords_path,
                                  Semantics semantics) {
    AmberTopology top;
    top.semantics_ = semantics;
    top.parse_prmtop(parmtop_path);
    top.build_terms();
    top.build_exclusions();
    top.build_cmap_coefficients();
    top.build_14_map();
    if (!coords_path.empty()) {
        top.read_coords(coords_path);
    } else {
        top.coordinates_.assign(top.atom_list_.size() * 3, 0.0);
    }
    top.sections_.clear();
    return top;
}

void AmberTopology::parse_prmtop(const std::string& path) {
    std::ifstream in(path);
    if (!in) throw std::runtime_error("Error opening file: " + path);
    std::string line;
    if (!std::getline(in, line) || line.rfind("%VERSION", 0) != 0)
        throw std::runtime_error("Invalid File Format. File does not start with %VERSION.");
    std::string current;
    int width = 0;  // field width of the current section's %FORMAT
    while (std::getline(in, line)) {
        if (!line.empty() && line.back() == '\r') line.pop_back();
        if (blank(line)) continue;
        if (line[0] == '%') {
            if (line.rfind("%FLAG", 0) == 0) {
                const size_t sp = line.find(' ');
                current = sp == std::string::npos ? std::string() : trim_right(line.substr(sp + 1));
                while (!current.empty() && current.front() == ' ') current.erase(current.begin());
                sections_[current];
                width = 0;
            } else if (line.rfind("%FORMAT", 0) == 0) {
                width = format_width(line);
            }
            continue;
        }
        if (current.empty()) continue;
        auto& tok = sections_[current];
        if (is_fixed_width_section(current)) {
            for (size_t i = 0; i < line.size(); i += 4) {
                std::string f = line.substr(i, 4);
                f.erase(std::remove_if(f.begin(), f.end(), [](unsigned char c) { return std::isspace(c); }),
                        f.end());
                tok.push_back(f);
            }
        } else if (semantics_ == Semantics::AMBER) {
            // fixed columns per the section's FORMAT where they describe the line
            numeric_fields(line, width, tok);
        } else {
            // blank-separated tokens, as the reference splits (IO.cpp:109-136)
            std::istringstream ss(line);
            std::string t;
            while (ss >> t) tok.push_back(t);
        }
    }

    auto sec = [&](const char* n) -> const std::vector<std::string>& {
        static const std::vector<std::string> empty;
        auto it = sections_.find(n);
        return it == sections_.end() ? empty : it->second;
    };
    const auto& ptr = sec("POINTERS");
    if (ptr.size() < 29)
        throw std::runtime_error("Insufficient POINTERS data: expected at least 29 entries, got " +
                                 std::to_string(ptr.size()));
    nAtoms_ = std::stoul(ptr[0]);
    nTypes_ = std::stoul(ptr[1]);

    const auto& names = sec("ATOM_NAME");
    atom_list_.clear();
    for (size_t i = 0; i < names.size() && i < nAtoms_; ++i) atom_list_.emplace_back(names[i]);
    if (atom_list_.size() != nAtoms_)
        throw std::runtime_error("ATOM_NAME has " + std::to_string(atom_list_.size()) +
                                 " entries, POINTERS says " + std::to_string(nAtoms_));
    const auto& chg = sec("CHARGE");
    for (size_t i = 0; i < chg.size() && i < nAtoms_; ++i) {
        const long double q = std::stold(chg[i]) / kChargeUnit;
        atom_list_[i].set_partial_charge(static_cast<double>(q));
    }
    const auto& z = sec("ATOMIC_NUMBER");
    for (size_t i = 0; i < z.size() && i < nAtoms_; ++i) atom_list_[i].set_atomic_number(std::stoi(z[i]));
    const auto& m = sec("MASS");
    for (size_t i = 0; i < m.size() && i < nAtoms_; ++i) atom_list_[i].set_mass(std::stod(m[i]));
    const auto& ty = sec("ATOM_TYPE_INDEX");
    if (ty.size() < nAtoms_) throw std::runtime_error("ATOM_TYPE_INDEX shorter than the atom count");
    for (size_t i = 0; i < nAtoms_; ++i) atom_list_[i].set_atom_type_index(std::stoul(ty[i]));
    const auto& nex = sec("NUMBER_EXCLUDED_ATOMS");
    for (size_t i = 0; i < nex.size() && i < nAtoms_; ++i) atom_list_[i].set_nExcluded_Atoms(std::stoi(nex[i]));
    const auto& at = sec("AMBER_ATOM_TYPE");
    for (size_t i = 0; i < at.size() && i < nAtoms_; ++i) atom_list_[i].set_type(at[i]);

    nbmatrix_.assign(nTypes_, std::vector<unsigned long>(nTypes_, 0));
    const auto& nbi = sec("NONBONDED_PARM_INDEX");
    for (size_t k = 0; k < nbi.size() && k < nTypes_ * nTypes_; ++k) {
        const long v = std::stol(nbi[k]);
        nbmatrix_[k / nTypes_][k % nTypes_] = v > 0 ? static_cast<unsigned long>(v) : 0UL;
    }
    auto reals = [&](const char* n, std::vector<double>& out) {
        out.clear();
        for (const auto& t : sec(n)) out.push_back(std::stod(t));
    };
    reals("LENNARD_JONES_ACOEF", lj_a_);
    reals("LENNARD_JONES_BCOEF", lj_b_);
    reals("LENNARD_JONES_14_ACOEF", lj14_a_);
    reals("LENNARD_JONES_14_BCOEF", lj14_b_);
    if (lj14_a_.empty()) lj14_a_ = lj_a_;
    if (lj14_b_.empty()) lj14_b_ = lj_b_;
    reals("BOND_FORCE_CONSTANT", types_.bond_k);
    reals("BOND_EQUIL_VALUE", types_.bond_r0);
    reals("ANGLE_FORCE_CONSTANT", types_.angle_k);
    reals("ANGLE_EQUIL_VALUE", types_.angle_theta0);
    reals("CHARMM_UREY_BRADLEY_FORCE_CONSTANT", types_.ub_k);
    reals("CHARMM_UREY_BRADLEY_EQUIL_VALUE", types_.ub_r0);
    reals("DIHEDRAL_FORCE_CONSTANT", types_.dih_k);
    reals("DIHEDRAL_PERIODICITY", types_.dih_n);
    reals("DIHEDRAL_PHASE", types_.dih_phase);
    reals("CHARMM_IMPROPER_FORCE_CONSTANT", types_.imp_k);
    reals("CHARMM_IMPROPER_PHASE", types_.imp_psi0);
    reals("SCEE_SCALE_FACTOR", scee_);
    reals("SCNB_SCALE_FACTOR", scnb_);

    cmap_resolution_.clear();
    for (const char* n : {"CHARMM_CMAP_RESOLUTION", "CMAP_RESOLUTION"})
        for (const auto& t : sec(n)) cmap_resolution_.push_back(std::stoi(t));
    cmap_grids_.assign(5, {});
    for (int s = 0; s < 5; ++s) {
        char a[40], b[40];
        std::snprintf(a, sizeof a, "CHARMM_CMAP_PARAMETER_%02d", s + 1);
        std::snprintf(b, sizeof b, "CMAP_PARAMETER_%02d", s + 1);
        for (const char* n : {static_cast<const char*>(a), static_cast<const char*>(b)})
            for (const auto& t : sec(n)) cmap_grids_[s].push_back(std::stod(t));
    }
    while (!cmap_grids_.empty() && cmap_grids_.back().empty()) cmap_grids_.pop_back();

    const auto& bd = sec("BOX_DIMENSIONS");
    if (bd.size() >= 4) {
        box_beta_ = std::stod(bd[0]);
        box_x_ = std::stod(bd[1]);
        box_y_ = std::stod(bd[2]);
        box_z_ = std::stod(bd[3]);
    }
}

void AmberTopology::build_terms() {
    auto ints = [&](const char* n) {
        std::vector<long> v;
        auto it = sections_.find(n);
        if (it != sections_.end())
            for (const auto& t : it->second) v.push_back(std::stol(t));
        return v;
    };
    const bool ref = semantics_ == Semantics::REF;
    // The reference's loops run while i + stride < size, which skips the last
    // record of every list except the CHARMM impropers.
    auto count = [&](size_t size, size_t stride) -> size_t {
        const size_t full = size / stride;
        return (ref && full > 0 && full * stride == size) ? full - 1 : full;
    };
    auto check_atom = [&](long idx) {
        if (idx < 0 || static_cast<unsigned long>(idx) >= nAtoms_)
            throw std::runtime_error("atom index out of range in a term list");
    };
    bonds_.clear(); types_.bond_type.clear();
    for (const char* n : {"BONDS_INC_HYDROGEN", "BONDS_WITHOUT_HYDROGEN"}) {
        const auto v = ints(n);
        const bool h = n[6] == 'I';
        for (size_t r = 0, c = count(v.size(), 3); r < c; ++r) {
            const long a = v[3 * r] / 3, b = v[3 * r + 1] / 3, t = v[3 * r + 2] - 1;
            check_atom(a); check_atom(b);
            if (t < 0 || static_cast<size_t>(t) >= types_.bond_k.size()) throw std::runtime_error("bad bond type");
            bonds_.emplace_back(types_.bond_k[t], types_.bond_r0[t], int(a), int(b), h);
            bonds_.back().set_type(int(t));
            types_.bond_type.push_back(int(t));
        }
    }
    angles_.clear(); types_.angle_type.clear();
    for (const char* n : {"ANGLES_INC_HYDROGEN", "ANGLES_WITHOUT_HYDROGEN"}) {
        const auto v = ints(n);
        const bool h = n[7] == 'I';
        for (size_t r = 0, c = count(v.size(), 4); r < c; ++r) {
            const long a = v[4 * r] / 3, b = v[4 * r + 1] / 3, cc = v[4 * r + 2] / 3, t = v[4 * r + 3] - 1;
            check_atom(a); check_atom(b); check_atom(cc);
            if (t < 0 || static_cast<size_t>(t) >= types_.angle_k.size()) throw std::runtime_error("bad angle type");
            angles_.emplace_back(types_.angle_k[t], types_.angle_theta0[t], int(a), int(b), int(cc), h);
            angles_.back().set_type(int(t));
            types_.angle_type.push_back(int(t));
        }
    }
    dihedrals_.clear(); types_.dih_type.clear();
    for (const char* n : {"DIHEDRALS_INC_HYDROGEN", "DIHEDRALS_WITHOUT_HYDROGEN"}) {
        const auto v = ints(n);
        const bool h = n[10] == 'I';
        for (size_t r = 0, c = count(v.size(), 5); r < c; ++r) {
            const long a = v[5 * r] / 3, b = v[5 * r + 1] / 3;
            long cc = v[5 * r + 2] / 3, d = v[5 * r + 3] / 3;
            const long t = v[5 * r + 4] - 1;
            const bool no14 = cc < 0, improper = d < 0;
            cc = std::labs(cc); d = std::labs(d);
            check_atom(a); check_atom(b); check_atom(cc); check_atom(d);
            if (t < 0 || static_cast<size_t>(t) >= types_.dih_k.size()) throw std::runtime_error("bad dihedral type");
            dihedrals_.emplace_back(types_.dih_k[t], types_.dih_phase[t], types_.dih_n[t], int(a), int(b),
                                    int(cc), int(d), h, improper, no14);
            dihedrals_.back().set_type(int(t));
            types_.dih_type.push_back(int(t));
        }
    }
    ubs_.clear(); types_.ub_type.clear();
    {
        const auto v = ints("CHARMM_UREY_BRADLEY");
        for (size_t r = 0, c = count(v.size(), 3); r < c; ++r) {
            const long a = v[3 * r] - 1, b = v[3 * r + 1] - 1, t = v[3 * r + 2] - 1;
            check_atom(a); check_atom(b);
            if (t < 0 || static_cast<size_t>(t) >= types_.ub_k.size()) throw std::runtime_error("bad UB type");
            ubs_.emplace_back(types_.ub_k[t], types_.ub_r0[t], int(a), int(b));
            ubs_.back().set_type(int(t));
            types_.ub_type.push_back(int(t));
        }
    }
    impropers_.clear(); types_.imp_type.clear();
    {
        // CHARMM impropers keep their last record (CreateBondedFromParsed.cpp:41).
        const auto v = ints("CHARMM_IMPROPERS");
        for (size_t r = 0; r < v.size() / 5; ++r) {
            const long a = v[5 * r] - 1, b = v[5 * r + 1] - 1, cc = v[5 * r + 2] - 1, d = v[5 * r + 3] - 1;
            const long t = v[5 * r + 4] - 1;
            check_atom(a); check_atom(b); check_atom(cc); check_atom(d);
            if (t < 0 || static_cast<size_t>(t) >= types_.imp_k.size()) throw std::runtime_error("bad improper type");
            impropers_.emplace_back(types_.imp_k[t], types_.imp_psi0[t], int(a), int(b), int(cc), int(d));
            impropers_.back().set_type(int(t));
            types_.imp_type.push_back(int(t));
        }
    }
    cmaps_.clear();
    {
        auto v = ints("CHARMM_CMAP_INDEX");
        if (v.empty()) v = ints("CMAP_INDEX");
        for (size_t r = 0, c = count(v.size(), 6); r < c; ++r) {
            int at[5];
            for (int k = 0; k < 5; ++k) {
                at[k] = int(v[6 * r + k] - 1);
                check_atom(at[k]);
            }
            const int set = int(v[6 * r + 5]);
            if (set < 1 || set > static_cast<int>(cmap_grids_.size()))
                throw std::runtime_error("CMAP parameter set out of range");
            cmaps_.emplace_back(set, at[0], at[1], at[2], at[3], at[4]);
        }
    }
}

void AmberTopology::build_exclusions() {
    for (auto& a : atom_list_) a.get_excluded_atoms().clear();
    if (semantics_ == Semantics::REF) return;  // the reference's lists stay empty
    auto it = sections_.find("EXCLUDED_ATOMS_LIST");
    if (it == sections_.end()) return;
    const auto& v = it->second;
    size_t start = 0;
    for (auto& a : atom_list_) {
        const size_t n = static_cast<size_t>(std::max(0, a.get_nExcluded_Atoms()));
        if (start + n > v.size()) throw std::runtime_error("Excluded atoms list is shorter than expected");
        auto& ex = a.get_excluded_atoms();
        for (size_t p = start; p < start + n; ++p) {
            const long j = std::stol(v[p]);
            if (j > 0) ex.push_back(int(j - 1));
        }
        std::sort(ex.begin(), ex.end());
        ex.erase(std::unique(ex.begin(), ex.end()), ex.end());
        start += n;
    }
}

void AmberTopology::build_cmap_coefficients() {
    cmap_coeff_.clear();
    const size_t nsets = std::min(cmap_grids_.size(), cmap_resolution_.size());
    for (size_t s = 0; s < nsets; ++s) {
        const int res = cmap_resolution_[s];
        const auto& g = cmap_grids_[s];
        if (res <= 0 || g.size() < static_cast<size_t>(res) * res)
            throw std::runtime_error("CMAP grid shorter than resolution^2");
        const size_t off = cmap_coeff_.size();
        cmap_coeff_.resize(off + static_cast<size_t>(res) * res * 16);
        cmap_patch_coefficients(res, 1, g.data(), cmap_coeff_.data() + off);
    }
}

// Cell (row, col) of a res x res grid: a = T e / 16 with e[4c + r] =
// grid(row + r - 1, col + c - 1), indices periodic.
void cmap_patch_coefficients(int res, int n_sets, const double* grid, double* coeff) {
    auto wrap = [res](int i) { return i < 0 ? i + res : (i >= res ? i - res : i); };
    for (int s = 0; s < n_sets; ++s) {
        const double* g = grid + static_cast<size_t>(s) * res * res;
        for (int row = 0; row < res; ++row)
            for (int col = 0; col < res; ++col) {
                double e[16];
                for (int c = 0; c < 4; ++c)
                    for (int r = 0; r < 4; ++r)
                        e[4 * c + r] = g[static_cast<size_t>(wrap(row + r - 1)) * res + wrap(col + c - 1)];
                double* a = coeff + ((static_cast<size_t>(s) * res + row) * res + col) * 16;
                for (int k = 0; k < 16; ++k) {
                    double acc = 0.0;
                    for (int n = 0; n < 16; ++n) acc += kCmapTable[k][n] * e[n] / 16;
                    a[k] = acc;
                }
            }
    }
}

// Transpose of cmap_patch_coefficients: d_grid = sum over cells of T^T d_a / 16
// scattered back to the 4 x 4 neighbourhood of each cell.
void cmap_patch_coefficients_adjoint(int res, int n_sets, const double* d_coeff, double* d_grid) {
    auto wrap = [res](int i) { return i < 0 ? i + res : (i >= res ? i - res : i); };
    for (int s = 0; s < n_sets; ++s) {
        double* g = d_grid + static_cast<size_t>(s) * res * res;
        for (int i = 0; i < res * res; ++i) g[i] = 0.0;
        for (int row = 0; row < res; ++row)
            for (int col = 0; col < res; ++col) {
                const double* da = d_coeff + ((static_cast<size_t>(s) * res + row) * res + col) * 16;
                for (int n = 0; n < 16; ++n) {
                    double acc = 0.0;
                    for (int k = 0; k < 16; ++k) acc += kCmapTable[k][n] * da[k] / 16;
                    const int c = n / 4, r = n % 4;
                    g[static_cast<size_t>(wrap(row + r - 1)) * res + wrap(col + c - 1)] += acc;
                }
            }
    }
}

void AmberTopology::build_14_map() {
    lj14_pairs_.clear();
    scee14_map_.clear();
    pairs14_.clear();
    pairs14_scee_.clear();
    std::unordered_map<uint64_t, size_t> order;
    for (const auto& d : dihedrals_) {
        if (d.get_improper() || d.get_exclude_14()) continue;
        const int a = d.get_atomA_index(), b = d.get_atomD_index();
        const uint64_t key = pair_key(a, b);
        lj14_pairs_.insert(key);
        const int t = d.get_type();
        const double scale = (t >= 0 && static_cast<size_t>(t) < scee_.size()) ? scee_[t] : 1.0;
        auto it = scee14_map_.find(key);
        if (it == scee14_map_.end()) {
            scee14_map_.emplace(key, scale);
            order.emplace(key, pairs14_.size());
            pairs14_.push_back({std::min(a, b), std::max(a, b)});
            pairs14_scee_.push_back(scale);
        } else if (std::abs(it->second - scale) > 1e-12) {
            // conflicting factors for one pair keep the larger divisor
            // (CreateNonBondedFromParsed.cpp:150-156)
            it->second = std::max(it->second, scale);
            pairs14_scee_[order[key]] = it->second;
        }
    }
}

bool AmberTopology::is_14_pair(int i, int j) const { return lj14_pairs_.count(pair_key(i, j)) != 0; }

double AmberTopology::get_scee_scale_for_pair(int i, int j) const {
    auto it = scee14_map_.find(pair_key(i, j));
    return it == scee14_map_.end() ? 1.0 : it->second;
}

void AmberTopology::read_coords(const std::string& path) {
    if (semantics_ == Semantics::AMBER) {
        // fixed-column numbers, velocities and a three- or six-value box
        const io::Restart r = io::read_rst7(path);
        if (static_cast<size_t>(r.n_atoms) != atom_list_.size())
            throw std::runtime_error("rst7 holds " + std::to_string(r.n_atoms) + " atoms, topology has " +
                                     std::to_string(atom_list_.size()));
        coordinates_ = r.coords;
        velocities_ = r.velocities;
        if (!r.has_box) {
            box_x_ = box_y_ = box_z_ = 0.0;
            box_alpha_ = box_beta_ = box_gamma_ = 90.0;
            return;
        }
        box_x_ = r.box[0]; box_y_ = r.box[1]; box_z_ = r.box[2];
        box_alpha_ = r.box[3]; box_beta_ = r.box[4]; box_gamma_ = r.box[5];
        return;
    }
    std::ifstream in(path);
    if (!in) throw std::runtime_error("Error opening file: " + path);
    std::string line;
    coordinates_.assign(atom_list_.size() * 3, 0.0);
    // title and atom-count lines (AmberTopology.cpp:394-395)
    if (!std::getline(in, line) || !std::getline(in, line)) return;
    const size_t need = atom_list_.size() * 3;
    size_t got = 0;
    std::vector<double> extra;
    while (std::getline(in, line)) {
        std::istringstream ss(line);
        std::string t;
        while (ss >> t) {
            if (got < need) coordinates_[got++] = std::stod(t);
            else extra.push_back(std::stod(t));
        }
    }
    if (got < need)
        throw std::runtime_error("rst7 holds " + std::to_string(got / 3) + " atoms, topology has " +
                                 std::to_string(atom_list_.size()));
    // values after the 3N coordinates are the box (AmberTopology.cpp:423-438)
    if (extra.size() < 6) {
        box_x_ = box_y_ = box_z_ = 0.0;
        box_alpha_ = box_beta_ = box_gamma_ = 90.0;
        return;
    }
    box_x_ = extra[0]; box_y_ = extra[1]; box_z_ = extra[2];
    box_alpha_ = extra[3]; box_beta_ = extra[4]; box_gamma_ = extra[5];
}

}  // namespace md::AmberTopology
