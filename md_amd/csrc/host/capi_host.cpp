// C entry points of the host library (libmdhost.so): parse a prmtop/rst7 with
// the host topology reader and expose it as an mdgpu_system_desc, so that
// non-C++ callers (the Python binding) share the C++ reader's semantics.
#include "mdhost.h"

#include "AmberTopology/AmberTopology.h"
#include "desc.h"

#include <algorithm>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

namespace {
thread_local std::string g_error;
}

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

}  // extern "C"
