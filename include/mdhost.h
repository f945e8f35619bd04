/* mdhost.h - C entry points of the host topology library (libmdhost.so).
 *
 * They expose the C++ prmtop/rst7 reader (md::AmberTopology::AmberTopology,
 * which keeps the reference's parsing semantics, the src/AmberTopology sources) to C and
 * Python callers as an mdgpu_system_desc ready for mdgpu_upload_system().
 * No counterpart exists in the reference, whose topology is reachable only
 * from C++ (AmberTopology.h:33).
 */
#ifndef MDHOST_H
#define MDHOST_H

#include "mdgpu.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct mdhost_system mdhost_system;

/* Parse prmtop (+ optional rst7; NULL or "" for none) in MDGPU_SEM_REF or
 * MDGPU_SEM_AMBER semantics.  Returns MDGPU_OK or MDGPU_ERR_ARG. */
int mdhost_load(const char* prmtop, const char* rst7, int semantics, mdhost_system** out);
/* Fill *out with pointers into the handle (valid until mdhost_free). */
int mdhost_desc(const mdhost_system* h, mdgpu_system_desc* out);
/* 3N coordinates read from the rst7 (zeros without one). */
const double* mdhost_coords(const mdhost_system* h);
/* Atom masses [N]. */
const double* mdhost_masses(const mdhost_system* h);
/* Box lengths (0,0,0 when the rst7 has no box). */
int mdhost_box(const mdhost_system* h, double box[3]);
/* Per-dihedral flags: bit 0 = no 1-4 (negative 3rd index), bit 1 = improper. */
const int32_t* mdhost_dihedral_flags(const mdhost_system* h);
void mdhost_free(mdhost_system* h);
/* CMAP patch coefficients [n_sets*res*res*16] from grids [n_sets*res*res]
 * (the reference's 16 x 16 table map, CreateNonBondedFromParsed.cpp:46-117),
 * and the transposed map: d_grid from d_coeff (for mdgpu_adjoint's CMAP
 * output).  No counterpart in the reference's API. */
int mdhost_cmap_coefficients(int res, int n_sets, const double* grid, double* coeff);
int mdhost_cmap_grid_grad(int res, int n_sets, const double* d_coeff, double* d_grid);
/* Amber restart files (no counterpart in the reference's API: it reads an
 * rst7 inside AmberTopology, src/AmberTopology/AmberTopology.cpp:388-438,
 * taking any values after the 3N coordinates as the box, and writes none).
 * mdhost_rst7_read: title, atom count and time, 3N coordinates, optional 3N
 * velocities (returned in A/ps) and an optional box (3 lengths, or lengths
 * and angles in degrees), read from 12-column fields when every field of a
 * line parses, else from blank-separated tokens.  MDGPU_ERR_ARG (message in
 * mdhost_last_error) when the file is unreadable or its value count fits no
 * layout; the time is NaN when the file has none.  mdhost_rst7_write: the Amber ASCII layout (title, I6 count and
 * E15.7 time, 6F12.7 coordinates, velocities and box); velocities (A/ps,
 * or NULL) and box6 (or NULL) optional. */
typedef struct mdhost_rst7 mdhost_rst7;
int mdhost_rst7_read(const char* path, mdhost_rst7** out);
int mdhost_rst7_info(const mdhost_rst7* h, int32_t* n_atoms, int32_t* has_velocities, int32_t* has_box,
                     double* time);
const double* mdhost_rst7_coords(const mdhost_rst7* h);
const double* mdhost_rst7_velocities(const mdhost_rst7* h);  /* NULL without velocities */
int mdhost_rst7_box(const mdhost_rst7* h, double box[6]);
const char* mdhost_rst7_title(const mdhost_rst7* h);
void mdhost_rst7_free(mdhost_rst7* h);
int mdhost_rst7_write(const char* path, const char* title, int32_t n_atoms, const double* coords,
                      const double* velocities, const double* box6, double time, int32_t has_time);

/* Message of the last failure on this thread. */
const char* mdhost_last_error(void);

#ifdef __cplusplus
}
#endif

#endif
