/* oracle.h - CPU FP64 oracle of the energy-and-force evaluation.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product (md_amd/, include/) links,
 * loads or calls this code; only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py use it, as the checker.
 *
 * It restates, single-threaded and in plain C, the reference's evaluation
 * (files cited per function in oracle.c) in the two semantics of include/
 * mdgpu.h, and adds two independent references used to pin it: a
 * structure-factor Ewald sum and the cardinal B-splines.
 */
#ifndef MD_ORACLE_H
#define MD_ORACLE_H

#include "../include/mdgpu.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
    double list_skin;      /* REF pair-list skin (reference default 2 A)       */
    int skip_all_pairs;    /* 1: skip the O(N^2) "EE" pass (ee_all = NaN)      */
    int full_virial;       /* 1: virial = bonded + pair + correction + recip   */
} oracle_opts;

/* Cell angles alpha, beta, gamma (degrees) for every later call, with the
 * box arguments as edge lengths (default 90, 90, 90: orthorhombic).  A
 * triclinic cell is AMBER semantics only; oracle_param_grads refuses it.
 * File-scope state: the oracle is single-threaded.  Returns 0, or 1 for an
 * angle outside (0, 180). */
int oracle_set_box_angles(const double angles[3]);
/* The cell matrix H (row-major, edge vectors as columns) and the distances
 * between opposite faces for edge lengths box and the current angles.
 * Returns 1 when they form no cell. */
int oracle_cell_matrix(const double box[3], double h[9], double face[3]);

/* One evaluation.  Returns 0 on success. */
int oracle_eval(const mdgpu_system_desc* sys, const mdgpu_params* p, const double* x,
                const double box[3], const oracle_opts* opt, mdgpu_energies* e,
                double* forces, double* virial);

/* The REF pair list (NonBondedUtils.cpp:185-282) as (i, j, is14) triples.
 * Returns the number of pairs; writes at most cap triples. */
long oracle_ref_pairlist(const mdgpu_system_desc* sys, const double* x, const double box[3],
                         double list_cutoff, int* out, long cap);

/* The Verlet rebuild criterion (NonBondedUtils.cpp:65-75, 87-100): 1 when
 * some atom's squared minimum-image displacement from xref exceeds
 * (skin/2)^2, else 0. */
int oracle_pairs_need_rebuild(int n, const double* x, const double* xref, const double box[3], double skin);

/* Velocity Verlet (NVE) for nsteps steps of length dt (ps) from (x, v, f),
 * f being the forces at x; masses in amu, a = 418.4 F / m A/ps^2.  Updates
 * x, v, f in place; trace[2k], trace[2k+1] (or NULL) receive step k's
 * potential energy (the evaluation's total) and kinetic energy.  Forces come
 * from oracle_eval (without the O(N^2) REF "EE" pass).  Returns 0 on success. */
int oracle_md_run(const mdgpu_system_desc* sys, const mdgpu_params* p, const double* masses, double* x, double* v,
                  double* f, const double box[3], double dt, int nsteps, double* trace);

/* L-BFGS minimisation (two-loop recursion, H0 = (s.y/y.y) I, pairs kept
 * only when s.y > 0, backtracking Armijo line search with c1 = 1e-4 and
 * halving, steps capped at max_step A per atom, steepest descent when the
 * direction is not a descent one), the algorithm mdgpu_minimize runs.
 * Minimises x in place; energies[k] (or NULL, up to max_iter + 1 entries)
 * receives the energy after k accepted iterations.  out[0..4] = iterations,
 * evaluations, converged, rms force, max atom force. */
int oracle_minimize(const mdgpu_system_desc* sys, const mdgpu_params* p, double* x, const double box[3],
                    int max_iter, double rms_tol, double max_step, int history, double* energies, double* out);

/* Structure-factor Ewald reciprocal energy, forces (added) and virial over
 * all integer vectors with |m_x|,|m_y|,|m_z| <= mmax, m != 0. */
double oracle_ewald_recip_dense(int n, const double* q, const double* x, const double box[3],
                                double alpha, int mmax, double* forces, double* virial);

/* Smooth-PME reciprocal part alone (energy; forces added; virial). */
double oracle_pme_recip(int n, const double* q, const double* x, const double box[3], double alpha,
                        const int grid[3], int order, double* forces, double* virial);

/* Reciprocal potential dE_rec/dq_i at each atom (returns E_rec). */
double oracle_pme_potential(int n, const double* q, const double* x, const double box[3], double alpha,
                            const int grid[3], int order, double* pot);

/* Parameter gradients sum_t w_t dE_t/dtheta at one configuration (AMBER
 * semantics): per-term weights in *w (w->total adds to every term of the
 * total); arrays of *g as in mdgpu_param_grads (NULL skips, others are
 * overwritten).  Pairs are enumerated by brute force: small systems only. */
int oracle_param_grads(const mdgpu_system_desc* sys, const mdgpu_params* p, const double* x,
                       const double box[3], const mdgpu_energies* w, mdgpu_param_grads* g);

/* Cardinal B-spline weights M_p(w + j) and their derivatives, j = 0..p-1. */
void oracle_bspline(int order, double w, double* theta, double* dtheta);

/* kCoulomb = 18.2206899283247^2 as the reference forms it. */
double oracle_coulomb_constant(void);

#ifdef __cplusplus
}
#endif

#endif
