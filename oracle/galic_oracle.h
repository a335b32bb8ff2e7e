/*
 * galic_oracle.h -- TEST INFRASTRUCTURE ONLY.
 *
 * A scalar FP64 C restatement of the GALIC v1.1 equilibrium hot path, used as
 * the parity checker for the HIP implementation and as the CPU baseline.  It
 * follows the reference's operation order line by line; every function cites
 * the reference file:line it restates.  Nothing in the product (galic_amd/,
 * include/) links or calls it: only tests/, the smoke check and the
 * cpu_baseline leg of bench.py do.
 *
 * The reference itself cannot be compiled in this image (every translation
 * unit includes GSL headers, which are absent, and a generated config header),
 * so this restatement is pinned by closed forms, invariants and brute force in
 * tests/test_oracle_*.py.  Functions whose pinning is partial say so below.
 */
#ifndef GALIC_ORACLE_H
#define GALIC_ORACLE_H

#include "../include/galic_types.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct oracle_model oracle_model;

/* ---- grid geometry (grid.c:329-487, 830-850) ---- */
/* Newton solve for the log spacing factor; returns Fac, writes Rmin. */
double or_solve_fac(int nbin, double rin, double rmax, double *rmin_out);
/* Fill Rmax, DG/FG/EG Rmin and Fac of p from Halo_A and the two enclosed-mass fractions. */
int or_grid_create(galic_params *p, double outer_frac, double inner_frac);

/* ---- model lifetime ---- */
oracle_model *or_model_new(const galic_params *p);
void or_model_free(oracle_model *m);
const galic_params *or_model_params(const oracle_model *m);
/* Tables are copied.  pot/dpotdr/dpotdz: FG_Nbin^2 each; vstream: 4 pointers (type 1..3) or NULL. */
void or_set_force_tables(oracle_model *m, const double *pot, const double *dpotdr, const double *dpotdz);
void or_set_vstream(oracle_model *m, int type, const double *vstream);
/* Target stacks for one type (each of length Nstack of its grid); copied. */
void or_set_targets(oracle_model *m, int type, const double *dgs_mass_target,
                    const double *egs_mass_target, const double *egs_r, const double *egs_t,
                    const double *egs_p, const double *egs_q);
void or_set_particles(oracle_model *m, int n, const double *pos, const double *vel, const double *mass,
                      const double *tint, const int *type, const double *vesc);

/* ---- accessors ---- */
const double *or_cellsize(const oracle_model *m);
const double *or_distance(const oracle_model *m);
/* which: 0 DGs_MassResponse, 1 EGs_MassResponse, 2..5 EGs_EgyResponse_{r,t,p,q},
 *        6..9 EGs_EgyResponse{R,T,Q,P}S, 10 DGs_MassTarget */
const double *or_stack(const oracle_model *m, int type, int which);
/* which: 0 S, 1 Sdisp_r, 2 Sdisp_t, 3 Sdisp_p, 4 Sdisp_q, 5 Srelfac, 6 Srelfac_count, 7 Sall */
double or_fit_value(const oracle_model *m, int type, int which);
void or_get_particles(const oracle_model *m, double *vel, int *orbits, double *vr2, double *vt2,
                      double *vp2, double *vq2);

/* ---- lookups (grid.c:15-197, 494-534, 896-938; disp_fields.c:118-132; halo.c/bulge.c) ---- */
void   or_forcegrid_get_cell(const oracle_model *m, const double *pos, int *iR, int *iz, double *fR, double *fz);
double or_forcegrid_get_potential(const oracle_model *m, const double *pos);
void   or_forcegrid_get_acceleration(const oracle_model *m, const double *pos, double *acc);
void   or_densitygrid_get_cell(const oracle_model *m, const double *pos, int *iR, int *iz, double *fR, double *fz);
void   or_energygrid_get_cell(const oracle_model *m, const double *pos, int *iR, int *iz, double *fR, double *fz);
double or_get_vstream(const oracle_model *m, const double *pos, int type);
void   or_halo_get_acceleration(const galic_params *p, const double *pos, double *acc);
double or_halo_get_potential(const galic_params *p, const double *pos);
void   or_bulge_get_acceleration(const galic_params *p, const double *pos, double *acc);
double or_bulge_get_potential(const galic_params *p, const double *pos);
double or_get_timestep(const oracle_model *m, const double *pos, const double *vel, const double *acc, int icell);

/* ---- orbit response (orbit_response.c:33-151, 155-491) ---- */
/* Extended form: also returns the step count and the start/end energies.
 * Returns t (sum of dt), or -1 if the step cap was exceeded. */
double or_orbit_response_mod(const oracle_model *m, const double *pos, const double *vel,
                             double *mfield, double *egy_r, double *egy_t, double *egy_q, double *egy_p,
                             double mass, double timespan, int *orbitstaken, int type,
                             long long *steps_out, double *e0_out, double *e1_out);
double or_orbit_response_plain(const oracle_model *m, const double *pos, const double *vel, double *mfield,
                               double mass, double timespan, int *orbitstaken, long long *steps_out);

/* ---- quadtree stacks (grid.c:1086-1151, 1247-1265) ---- */
void   or_smooth_stack(double *data, int maxlevel);
double or_calc_stack_difference(const oracle_model *m, const double *d1, const double *d2, int l, int i, int j,
                                int maxlevel, const double *ref1, const double *ref2, double thresh,
                                const double *dist, int flag);
double or_calc_stack_difference_mod(const oracle_model *m, const double *d1, const double *d2, int l, int i, int j,
                                    int maxlevel, const double *ref1, const double *ref2, double thresh,
                                    const double *dist, int flag);

/* ---- velocity decomposition and merit (disp_fields.c:200-372, main.c:877-917, 1021-1217) ---- */
void   or_calc_disp_components(const oracle_model *m, const double *pos, int type, const double *vel,
                               double *vr2, double *vt2, double *vp2, double *vq2);
double or_eval_fit_mod(oracle_model *m, int n, const double *vel,
                       const double *mo_new, const double *mo_old, const double *r_new, const double *r_old,
                       const double *t_new, const double *t_old, const double *q_new, const double *q_old,
                       const double *p_new, const double *p_old);

/* ---- optimiser driver (main.c:120-362, 365-527, 808-874) ---- */
void or_calc_all_response_fields(oracle_model *m, int nthreads);
void or_calc_global_fit(oracle_model *m);
/* One rep: init_updates, optimize() for each listed particle with the given trial
 * velocity (3 per trial), commit_updates.  accept[i] receives 1 if taken.
 * dF (optional) receives F_try - F_base per trial. */
void or_rep(oracle_model *m, int ntrial, const int *idx, const double *vtry, int *accept, double *dF, int nthreads);
/* As or_rep but without the commit: evaluates every trial against the current global
 * state and returns F_base, F_try.  Used to check merit values one by one. */
void or_eval_trials(oracle_model *m, int ntrial, const int *idx, const double *vtry, double *fbase, double *ftry,
                    int nthreads);

/* ---- sampled targets (grid.c:536-617, 941-1076; halo.c:17-40; disk.c:17-51; bulge.c:17-38;
 *      disp_fields.c:26-198, 375-398) ---- */
/* One round of a component's fresh-coordinate draw with uniforms u[0..2] in the
 * reference's order; returns 1 if the point is kept, 0 if the reference draws again. */
int    or_sample_point(const galic_params *p, int type, double stretch, const double *u, double *pos);
/* Target dispersion tables of a type (copied; NULL where unused) and its beta parameter. */
void   or_set_disp_tables(oracle_model *m, int type, double beta, const double *disp_r, const double *disp_z,
                          const double *disp_phi, const double *tilt_r, const double *tilt_theta);
void   or_get_disp_rtp(const oracle_model *m, const double *pos, int type, double *dr, double *dt, double *dp,
                       double *dq);
/* Bins n points into DGs_MassTarget (dg, [DG_Nstack]) and the EGs targets (eg, [5][EG_Nstack]:
 * mass, r, t, p, q), either may be NULL; normalised by MType/count and smoothed.  Returns the count. */
long long or_bin_targets(const oracle_model *m, int type, long long n, const double *pos, double *dg, double *eg);

/* ---- Barnes-Hut tree (forcetree.c:139-722, forcetree_walk.c:12-222, gravtree.c:455-463) ---- */
/* Accelerations and potentials (already multiplied by G) at ntarget points due to nsrc sources.
 * The domain cube is the bounding cube of sources and targets (domain_box.c:20-65).
 * Returns the total number of interactions, or -1 on failure. */
long long or_tree_forces(const galic_params *p, int nsrc, const double *spos, const double *smass,
                         int ntarget, const double *tpos, double *acc, double *pot, int nthreads);
/* Direct summation with the same softening kernel (for the theta -> 0 check). */
void or_direct_forces(const galic_params *p, int nsrc, const double *spos, const double *smass,
                      int ntarget, const double *tpos, double *acc, double *pot, int nthreads);
/* Table fill (grid.c:680-750): targets at (R_j, z_k, phi_s) for 8 sections, results
 * projected on R and averaged.  Writes FG_Nbin^2 each of pot, dpotdr, dpotdz. */
long long or_table_fill(const galic_params *p, int nsrc, const double *spos, const double *smass,
                        double *pot, double *dpotdr, double *dpotdz, int nthreads);

#ifdef __cplusplus
}
#endif

#endif
