// C entry points of the host library (libmdhost.so): parse a prmtop/rst7 with
// the host topology reader and expose it as an mdgpu_system_desc, so that
// non-C++ callers (the Python binding) share the C++ reader's semantics.
#include "mdhost.h"

#include "AmberTopology/AmberTopology.h"
#include "IO/Restart.h"
#include "desc.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

namespace {
thread_local std::string g_error;
}

struct mdhost_rst7 {
    md::io::Restart r;
};

struct mdhost_system {
    md::AmberTopology::AmberTopology top;
    md::host::DescStorage d;
};

extern "C" {

int mdhost_load(const char* prmtop, const char* rst7, int semantics, mdhost_system** out) {
    if (!prmtop || !out) {
        g_error = "mdhost_load: null argument";
        return MDGPU_ERR_ARG;
    }
    try {
        auto h = std::make_unique<mdhost_system>();
        h->top = md::AmberTopology::AmberTopology::load(
            prmtop, rst7 ? rst7 : "", semantics == MDGPU_SEM_AMBER ? md::Semantics::AMBER : md::Semantics::REF);
        md::host::build_system_desc(h->top, &h->d);
        *out = h.release();
        return MDGPU_OK;
    } catch (const std::exception& e) {
        g_error = e.what();
        return MDGPU_ERR_ARG;
    }
}

int mdhost_desc(const mdhost_system* h, mdgpu_system_desc* out) {
    if (!h || !out) return MDGPU_ERR_ARG;
    *out = h->d.desc;
    return MDGPU_OK;
}

const double* mdhost_coords(const mdhost_system* h) { return h ? h->d.coords.data() : nullptr; }
const double* mdhost_masses(const mdhost_system* h) { return h ? h->d.mass.data() : nullptr; }

int mdhost_box(const mdhost_system* h, double box[3]) {
    if (!h || !box) return MDGPU_ERR_ARG;
    std::memcpy(box, h->d.box, sizeof h->d.box);
    return MDGPU_OK;
}

const int32_t* mdhost_dihedral_flags(const mdhost_system* h) { return h ? h->d.dih_flags.data() : nullptr; }

void mdhost_free(mdhost_system* h) { delete h; }

int mdhost_cmap_coefficients(int res, int n_sets, const double* grid, double* coeff) {
    if (res <= 0 || n_sets < 0 || (!grid && n_sets) || (!coeff && n_sets)) return MDGPU_ERR_ARG;
    md::AmberTopology::cmap_patch_coefficients(res, n_sets, grid, coeff);
    return MDGPU_OK;
}

int mdhost_cmap_grid_grad(int res, int n_sets, const double* d_coeff, double* d_grid) {
    if (res <= 0 || n_sets < 0 || (!d_coeff && n_sets) || (!d_grid && n_sets)) return MDGPU_ERR_ARG;
    md::AmberTopology::cmap_patch_coefficients_adjoint(res, n_sets, d_coeff, d_grid);
    return MDGPU_OK;
}

const char* mdhost_last_error(void) { return g_error.c_str(); }

int mdhost_rst7_read(const char* path, mdhost_rst7** out) {
    if (!path || !out) {
        g_error = "mdhost_rst7_read: null argument";
        return MDGPU_ERR_ARG;
    }
    try {
        auto h = std::make_unique<mdhost_rst7>();
        h->r = md::io::read_rst7(path);
        *out = h.release();
        return MDGPU_OK;
    } catch (const std::exception& e) {
        g_error = e.what();
        return MDGPU_ERR_ARG;
    }
}

int mdhost_rst7_info(const mdhost_rst7* h, int32_t* n_atoms, int32_t* has_velocities, int32_t* has_box,
                     double* time) {
    if (!h) return MDGPU_ERR_ARG;
    if (n_atoms) *n_atoms = h->r.n_atoms;
    if (has_velocities) *has_velocities = h->r.velocities.empty() ? 0 : 1;
    if (has_box) *has_box = h->r.has_box ? 1 : 0;
    if (time) *time = h->r.has_time ? h->r.time : std::nan("");
    return MDGPU_OK;
}

const double* mdhost_rst7_coords(const mdhost_rst7* h) { return h ? h->r.coords.data() : nullptr; }
const double* mdhost_rst7_velocities(const mdhost_rst7* h) {
    return h && !h->r.velocities.empty() ? h->r.velocities.data() : nullptr;
}
int mdhost_rst7_box(const mdhost_rst7* h, double box[6]) {
    if (!h || !box) return MDGPU_ERR_ARG;
    for (int k = 0; k < 6; ++k) box[k] = h->r.box[k];
    return MDGPU_OK;
}
const char* mdhost_rst7_title(const mdhost_rst7* h) { return h ? h->r.title.c_str() : ""; }
void mdhost_rst7_free(mdhost_rst7* h) { delete h; }

int mdhost_rst7_write(const char* path, const char* title, int32_t n_atoms, const double* coords,
                      const double* velocities, const double* box6, double time, int32_t has_time) {
    if (!path || !coords || n_atoms <= 0) {
        g_error = "mdhost_rst7_write: null argument or no atoms";
        return MDGPU_ERR_ARG;
    }
    try {
        md::io::Restart r;
        r.title = title ? title : "";
        r.n_atoms = n_atoms;
        r.coords.assign(coords, coords + 3 * static_cast<size_t>(n_atoms));
        if (velocities) r.velocities.assign(velocities, velocities + 3 * static_cast<size_t>(n_atoms));
        if (box6) {
            r.has_box = true;
            for (int k = 0; k < 6; ++k) r.box[k] = box6[k];
        }
        r.has_time = has_time != 0;
        r.time = time;
        md::io::write_rst7(path, r);
        return MDGPU_OK;
    } catch (const std::exception& e) {
        g_error = e.what();
        return MDGPU_ERR_ARG;
    }
}

}  // extern "C"
