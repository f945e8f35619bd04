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

const char* mdhost_last_error(void) { return g_error.c_str(); }

}  // extern "C"
