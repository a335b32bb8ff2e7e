/*
 * galic_hip.h -- C ABI of the MI355X (gfx950) implementation of GALIC's
 * equilibrium hot path.  Plain pointers and sizes only; every entry point that
 * can fail returns 0 on success or a negative GALIC_ERR_* code, with a message
 * available from galic_last_error().
 *
 * Layouts (all row-major, as in the reference):
 *   force tables   FG_Nbin*FG_Nbin doubles, index iz*FG_Nbin + iR   (grid.c:78-80,140-146)
 *   response grids 2^L * 2^L doubles,      index iz*2^L + iR        (orbit_response.c:283-286)
 *   stacks         (4^(L+1)-1)/3 doubles,  node (l,i,j) at GALIC_STACKOFFSET(l,i,j) (allvars.h:160)
 *   positions and velocities of n particles: 3*n doubles, particle-major (x0,y0,z0,x1,...)
 *   orbit fields   per orbit 5 grids in the order m, r, t, q, p of
 *                  produce_orbit_response_field_mod's arguments (orbit_response.c:155-157)
 * Component types are numbered as in the reference: 1 halo, 2 disk, 3 bulge.
 */
#ifndef GALIC_HIP_H
#define GALIC_HIP_H

#include "galic_types.h"

#ifdef __cplusplus
extern "C" {
#endif

#define GALIC_OK 0
#define GALIC_ERR_ARG (-1)     /* bad argument or size */
#define GALIC_ERR_HIP (-2)     /* HIP runtime error (message has the HIP error string) */
#define GALIC_ERR_STEPCAP (-3) /* an orbit exceeded MaxStepsPerOrbit; the reference exits (orbit_response.c:449-456) */
#define GALIC_ERR_STATE (-4)   /* call out of order (e.g. a rep before the initial response pass) */
#define GALIC_ERR_NOMEM (-5)   /* device allocation failed */

typedef struct galic_ctx galic_ctx;

const char *galic_last_error(void);
int galic_version(void);

/* ---- set-up geometry (host) ---- */

/* Replaces the geometry part of forcedensitygrid_create (grid.c:329-480): fills
 * p->Rmax and the DG/FG/EG Rmin and Fac from Halo_A, DG_MaxLevel, FG_Nbin,
 * EG_MaxLevel and the two enclosed-mass fractions. */
int galic_grid_create(galic_params *p, double outer_frac, double inner_frac);

/* DG_CellSize and DG_CellVol (grid.c:830-850; 4^L each) and the DGs_Distance
 * stack (grid.c:385-416; Nstack).  Any output pointer may be NULL. */
int galic_density_geometry(const galic_params *p, double *cellsize, double *cellvol, double *distance);

/* ---- context: one model resident on one GPU ---- */

/* Allocates the grids for p on HIP device `device`.  p->DG_MaxLevel must equal
 * p->EG_MaxLevel: the reference bins the energy fields with density-grid
 * indices (orbit_response.c:283-319) and is only consistent when they match. */
int galic_create(const galic_params *p, int device, galic_ctx **out);
int galic_destroy(galic_ctx *ctx);

/* Force and potential tables, FG_DPotDR/FG_DPotDz/FG_Pot (grid.c:204-216). */
int galic_set_force_tables(galic_ctx *ctx, const double *pot, const double *dpotdr, const double *dpotdz);
/* Streaming velocity table FG_Vstream[type] (disp_fields.c:118-132); NULL means zero. */
int galic_set_vstream(galic_ctx *ctx, int type, const double *vstream);
/* Target stacks of one component: DGs_MassTarget, EGs_MassTarget and
 * EGs_EgyTarget_{r,t,p,q} (grid.c:830-877, 941-1076). */
int galic_set_targets(galic_ctx *ctx, int type, const double *dgs_mass_target, const double *egs_mass_target,
                      const double *egs_r, const double *egs_t, const double *egs_p, const double *egs_q);
/* Target dispersion tables of one component, as get_disp_rtp reads them
 * (disp_fields.c:66-115): FG_Disp_r [FG_Nbin] for velocity structures 0 and 1
 * (disp_fields.c:375-398), FG_DispZ [FG_Nbin^2] for structure 2, FG_DispPhi
 * [FG_Nbin^2] for structures 2 and 3, FG_tilted_vR2_prime and
 * FG_tilted_vz2_prime [FG_Nbin^2] for structure 3 (disp_fields.c:134-198);
 * NULL where the structure does not use the table.  beta is
 * HaloBetaParameter / BulgeBetaParameter (>= 1: the slope-dependent beta of
 * disp_fields.c:26-64).  Copied to the device. */
int galic_set_disp_tables(galic_ctx *ctx, int type, double beta, const double *disp_r, const double *disp_z,
                          const double *disp_phi, const double *tilt_r, const double *tilt_theta);
/* Sampled target maps of one component, computed on the device:
 * densitygrid_sample_targetresponse (grid.c:536-617) into dgs_mass_target
 * [Nstack] and calc_energy_grid_mass_maps (grid.c:941-1076) into egs_targets
 * [5][Nstack] (EGs_MassTarget, EGs_EgyTarget_r, _t, _p, _q); either may be NULL.
 * Points come from halo_/disk_/bulge_get_fresh_coordinate (halo.c:17-40,
 * disk.c:17-51, bulge.c:17-38) driven by Philox draws of (sample id, round) under
 * the key (seed, stream) instead of one GSL stream; nsamples ids are drawn (at
 * most 2^30), those with sqrt(x^2+z^2) < Rmax are binned and counted
 * (*nbinned), and the maps are normalised by MType/count as the reference
 * normalises by its count.  stretch is HaloStretch / BulgeStretch.  The results
 * are returned, not installed: galic_set_targets installs them.  points
 * (optional, [nsamples][3]) receives the sample points (zeros for an id whose
 * draws never fell inside Rmax); kernel_ms (optional) the sampling kernel's time. */
int galic_sample_targets(galic_ctx *ctx, int type, long long nsamples, unsigned int seed, unsigned int stream,
                         double stretch, double *dgs_mass_target, double *egs_targets, double *points,
                         long long *nbinned, float *kernel_ms);
/* Particle state P[] (allvars.h:844-868): positions, velocities, mass, Tint,
 * type and Vesc.  Copied to device SoA arrays. */
int galic_set_particles(galic_ctx *ctx, int n, const double *pos, const double *vel, const double *mass,
                        const double *tint, const int *type, const double *vesc);
/* Target dispersions of each particle, P[].vr2_target, vt2_target, vp2_target,
 * vq2_target (set_particles.c:158-161): 4*n doubles, particle-major.  Used by
 * the device trial draws of galic_fp32_rep. */
int galic_set_dispersions(galic_ctx *ctx, int n, const double *disp4);
/* Reads back P[].Vel, P[].Orbits and P[].vr2..vq2; any pointer may be NULL. */
int galic_get_particles(galic_ctx *ctx, double *vel, int *orbits, double *vr2, double *vt2, double *vp2,
                        double *vq2);

/* ---- optimiser, plain FP64 path (the drop-in) ---- */

/* calc_all_response_fields (main.c:151-246): one orbit per particle, Loc grids,
 * instantaneous grids and the first calc_global_fit. */
int galic_calc_all_response_fields(galic_ctx *ctx);

/* One rep of the optimisation loop (main.c:96-100): init_updates, optimize()
 * for each particle idx[i] with trial velocity vtry[3*i..3*i+2] in place of the
 * GSL draw of main.c:407-435, then commit_updates.  accept[i] = 1 if taken;
 * dF[i] (optional) = F_try - F_base. */
int galic_rep(galic_ctx *ctx, int ntrial, const int *idx, const double *vtry, int *accept, double *dF);

/* Stacks: which = 0 DGs_MassResponse, 1 EGs_MassResponse, 2..5 EGs_EgyResponse_{r,t,p,q},
 * 6..9 EGs_EgyResponse{R,T,Q,P}S, 10 DGs_MassTarget, 11 EGs_MassTarget,
 * 12..15 EGs_EgyTarget_{r,t,p,q}.  out has Nstack doubles. */
int galic_get_stack(galic_ctx *ctx, int type, int which, double *out);

/* calc_global_fit results of one type: out[0..7] = S, Sdisp_r, Sdisp_t, Sdisp_p,
 * Sdisp_q, Srelfac, Srelfac_count, Sall (main.c:294-359, 1270). */
int galic_get_fit(galic_ctx *ctx, int type, double *out8);

/* ---- the reference's output files, from the current state ---- */
/* Appends one line of fit_<type>.txt as log_message writes it (main.c:1221-1288):
 * opt-frac, S, vavg, Sdisp_r, Sdisp_t, Sdisp_p, Sdisp_q, Sall, Chg-Frac.  opt_frac
 * is the caller's Noptimized / TotNumPart, tries and changes its counts of the type
 * since the last line; vavg is the rms speed of the type's particles.  Writes
 * nothing for a type without particles, as the reference. */
int galic_write_fit_line(galic_ctx *ctx, int type, const char *path, double opt_frac, long long tries,
                         long long changes);
/* densfield_<type>_<num>.dat as output_density_field writes it (io.c:36-457, the
 * build without VER_1_1_GNUPLOT_LOG): the DGs_MassTarget stack level by level,
 * calc_smoothed_stack maps (grid.c:1217-1244) of the targets and responses, the
 * fine DGs_MassResponse, the cell-centre coordinates of both grids. */
int galic_write_densfield(galic_ctx *ctx, int type, const char *path);
/* One-file Gadget format-1 snapshot as output_particles/write_file write it
 * (io.c:460-1240, SnapFormat 1, NumFilesPerSnapshot 1): header, then the POS,
 * VEL, ID (1..N in the particles' order, set_particles.c:65), MASS and VELTHEO
 * blocks, particles grouped by type.  Single precision unless double_precision
 * (OUTPUT_IN_DOUBLEPRECISION); vel_theo (3 per particle, the comparison
 * velocities of set_particles.c:94-109) may be NULL for zeros. */
int galic_write_snapshot(galic_ctx *ctx, const char *path, int double_precision, const double *vel_theo);

/* Orbit responses of many independent orbits (produce_orbit_response_field_mod,
 * orbit_response.c:155-491, once per orbit): orbit o starts at pos[3*o..] with
 * velocity vel[3*o..], mass[o], timespan tint[o] and component type[o].
 * fields (optional) receives 5*4^L doubles per orbit (m, r, t, q, p), scaled by
 * mass/t as the reference returns them.  t, orbits, steps, e0, e1 are optional
 * per-orbit outputs (integration time, pericentre passages, leapfrog steps,
 * start and end energy). */
int galic_orbit_responses(galic_ctx *ctx, int norbit, const double *pos, const double *vel, const double *mass,
                          const double *tint, const int *type, double *fields, double *t, int *orbits,
                          long long *steps, double *e0, double *e1);

/* The orbits of galic_orbit_responses integrated by the FP32 path (FP32 state
 * and tables, FP64 time): the same outputs, for measuring its response error
 * and energy drift against the FP64 path.  E0 and E1 use the FP64 potential. */
int galic_fp32_orbit_responses(galic_ctx *ctx, int norbit, const double *pos, const double *vel, const double *mass,
                               const double *tint, const int *type, double *fields, double *t, int *orbits,
                               long long *steps, double *e0, double *e1);

/* Lookups at n points (pos: 3*n): acceleration (3*n) and potential (n) of
 * forcegrid_get_acceleration / forcegrid_get_potential (grid.c:59-197). */
int galic_field_at(galic_ctx *ctx, int n, const double *pos, double *acc, double *pot);

/* The lookups of galic_field_at without allocating device memory: no counterpart in the
 * reference; for callers that cannot allocate (code under HIP graph capture, tight serving
 * loops).  The device staging is a static array of this library holding
 * GALIC_FIELD_STATIC_MAX points, one copy per GPU kept until the process ends; more points
 * are rejected with GALIC_ERR_ARG.  Calls that overlap on one GPU share the array and race. */
#define GALIC_FIELD_STATIC_MAX 65536
int galic_field_at_static(galic_ctx *ctx, int n, const double *pos, double *acc, double *pot);

/* ---- FP32 batched path ---- */

/* Initial response pass of the FP32 path: orbits in FP32, grids as 64-bit
 * fixed point (bit-identical run to run), merit sums in FP64. */
int galic_fp32_calc_all_response_fields(galic_ctx *ctx);

/* One FP32 rep over ntrial particles idx[] with k trial velocities each
 * (vtry: 3*k*ntrial, particle-major, trial-minor).  The best trial of a
 * particle is taken when it lowers the merit.  accept[i] = chosen trial + 1,
 * or 0.  vtry == NULL draws the trials on the device: Gaussian perturbations of
 * sigma_frac times the local dispersion targets, counter-based RNG keyed by seed. */
int galic_fp32_rep(galic_ctx *ctx, int ntrial, const int *idx, int k, const double *vtry, double sigma_frac,
                   unsigned long long seed, int *accept);

/* Of the k trials a particle gets from the device draws of galic_fp32_rep /
 * galic_fp32_iterate, the first nfresh are fresh draws from optimize()'s trial
 * distribution (main.c:383-435) and the rest perturbations of the current
 * velocity (default nfresh = 0). */
int galic_fp32_set_fresh_trials(galic_ctx *ctx, int nfresh);

/* ntrial trials in reps of rep_size (main.c:93-100: each rep evaluated against
 * the grids at its start, then committed), with the arguments of galic_fp32_rep.
 * Equivalent to galic_fp32_rep on consecutive slices of idx; the orbits of as
 * many reps as fit in device memory are integrated in one launch. */
int galic_fp32_iterate(galic_ctx *ctx, int ntrial, const int *idx, int k, int rep_size, const double *vtry,
                       double sigma_frac, unsigned long long seed, int *accept);

/* The merit change F_try - F_base of every trial of the last galic_fp32_rep /
 * galic_fp32_iterate call (main.c:464-465, one value per trial velocity):
 * dFk[i*k + j] for trial j of the i-th particle of that call's idx[].  ntrial
 * and k must be that call's. */
int galic_fp32_trial_merits(galic_ctx *ctx, int ntrial, int k, double *dFk);

/* Times the FP32 orbit kernel's persistent grid sizes (1 to all of the workgroups a
 * CU holds at once, and twice that) on nsample orbits of the model and keeps the
 * fastest for later reps; reports the choice (variant: the kernel variant, 0).
 * The launches give bit-identical results.  nsample should exceed the lanes the
 * GPU holds at once. */
int galic_fp32_autotune(galic_ctx *ctx, int nsample, int *variant, int *blocks_per_cu);

/* Form of the FP32 merit kernel: 0 automatic (form 4 where it applies: grids of 16^2 to
 * 1024^2 cells and K <= 16; otherwise LDS grids of the whole response grid where they fit,
 * LDS tiles of 32 x 32 cells, else two int64 grids per workgroup in HBM), 1 dense 32 x 32
 * tiles at every level, 2 never tiles, 3 32 x 32 tiles restricted to the 2x2 nodes the logs
 * touch (where their LDS fits: K up to 8 on 512^2 grids), 4 tiles of 16 x 16 cells in each
 * particle's own int32 fixed point (2^30 units for the largest total of its orbits).  Forms 1-3
 * give the same decisions as form 2 and bit-identical grids, merit sums equal to FP64
 * rounding; form 4's merit sums differ from them by its fixed point's rounding (about 2^-30 of
 * an orbit's total per record), its deltas are the same int64 sums. */
int galic_fp32_set_merit_form(galic_ctx *ctx, int form);

/* galic_fp32_iterate over trials idx[0, ntrial) (device-drawn trials, the same results bit for
 * bit), its reps split into six runs, two for each of three sharings of the CUs between the
 * orbit launches and the judging they overlap (orbit kernel workgroups per CU beside the
 * judging, merit kernel workgroups per CU): (autotuned / 2, all), (autotuned, all),
 * (autotuned / 2, 1).  The fastest (the better run, trials per second) is kept for later calls
 * and returned.  No counterpart in the
 * reference (an autotuner of this build); galic_fp32_autotune must run first. */
int galic_fp32_autotune_schedule(galic_ctx *ctx, int ntrial, const int *idx, int K, int rep_size,
                                 double sigma_frac, unsigned long long seed, int *accept, int *orbit_wg_per_cu,
                                 int *merit_wg_per_cu);

/* Workgroups per CU of the FP32 path's tile merit kernels (its persistent grid): at most
 * wg_per_cu, 0 (default) as many as the CU holds.  No counterpart in the reference; set by
 * galic_fp32_autotune.  The results do not depend on it. */
int galic_fp32_set_merit_launch(galic_ctx *ctx, int wg_per_cu);

/* Schedule of the FP32 reps (no counterpart in the reference, whose reps are one loop).
 * The trials of a call are integrated in groups of whole reps, one orbit launch per group;
 * max_group_trials > 0 caps a group at that many trials (0: as many as the log arena holds).
 * overlap = 1: the next group's orbits are integrated on a second stream while a group's reps
 * are judged (when the two groups' particles are distinct), 0: one group after the other,
 * -1 (default): overlap when a call's orbits need more than one launch.  graph = 1 (default):
 * with the 16x16-tile merit, the reps of a group (merit, choice, commit, fit) are captured in
 * one hipGraph and launched once, 0: launched rep by rep with the host between them.  The
 * results are the same bits whatever the schedule. */
int galic_fp32_set_schedule(galic_ctx *ctx, int max_group_trials, int overlap, int graph);

/* Timing aid of the sparse tile merit: on != 0 starts (or restarts) summing, per phase of
 * each particle, the wall-clock ticks (100 MHz) of the kernel's thread 0; out16 (optional)
 * first receives the sums so far: [0] tile lists, [1] fields and tiles, [2] choice, [3]
 * deltas of the taken trial, [5] tiles visited, [6] list entries, [7] particles, [8] taken. */
int galic_fp32_merit_profile(galic_ctx *ctx, int on, unsigned long long *out16);

/* Sets the FP32 orbit launch by hand (variant 0, 1..32 workgroups per CU). */
int galic_fp32_set_launch(galic_ctx *ctx, int variant, int blocks_per_cu);

/* ---- Barnes-Hut tree fill of the force tables ---- */

/* Walk modes, or-ed into the `precision` argument of the tree calls.  0 is the
 * reference's walk: FP64, monopole moments (forcetree.c:619-697), Gadget-2
 * spline softening of length ForceSoftening (forcetree_walk.c:158-191). */
#define GALIC_TREE_FP32 1        /* FP32 interactions and sums per target, FP64 moments */
#define GALIC_TREE_QUADRUPOLE 2  /* traceless quadrupole moments (FP64 sums) at accepted nodes; the reference has none */
#define GALIC_TREE_PLUMMER 4     /* Plummer softening of length ForceSoftening / 2.8 (the Softening parameter) */
#define GALIC_TREE_GROUP 8       /* with GALIC_TREE_FP32: 64 targets walk together, a node accepted only where the
                                    criterion holds for their whole bounding box (more interactions, none less
                                    accurate); table fills take the targets in 8 x 8 blocks of a section */

/* Accelerations and potentials (times G) at ntarget points from nsrc sources
 * (forcetree.c:41-722, forcetree_walk.c:12-222, gravtree.c:455-463).
 * precision: a combination of the GALIC_TREE_ modes (0: the reference's walk).
 * ninteractions (optional) receives the interaction count. */
int galic_tree_forces(const galic_params *p, int device, int nsrc, const double *spos, const double *smass,
                      int ntarget, const double *tpos, double *acc, double *pot, int precision,
                      long long *ninteractions);

/* forcedensitygrid_calculate's table fill (grid.c:680-750): targets at
 * (R_j, z_k) in FG_SECTIONS azimuthal sections, projected on R and averaged.
 * Writes FG_Nbin^2 doubles each of pot, dpotdr, dpotdz. */
int galic_tree_fill(const galic_params *p, int device, int nsrc, const double *spos, const double *smass,
                    double *pot, double *dpotdr, double *dpotdz, int precision, long long *ninteractions);

/* Direct summation in FP64 with the tree's softened interaction
 * (forcetree_walk.c:162-183), times G: the reference point of the tree's force
 * error.  O(nsrc * ntarget). */
int galic_direct_forces(const galic_params *p, int device, int nsrc, const double *spos, const double *smass,
                        int ntarget, const double *tpos, double *acc, double *pot);
/* The same with mode GALIC_TREE_PLUMMER for Plummer softening (0: the spline). */
int galic_direct_forces_mode(const galic_params *p, int device, int nsrc, const double *spos, const double *smass,
                             int ntarget, const double *tpos, double *acc, double *pot, int mode);

/* Device times of the last galic_tree_forces / galic_tree_fill call of this
 * process: out[0] tree build ms, out[1] walk ms, out[2] interactions. */
int galic_tree_last_timings(double *out3);

/* Sources of the table fill for the context, kept on its GPU: nsrc positions (3 each) and
 * masses, the fresh samples forcedensitygrid_calculate draws (grid.c:659-678). */
int galic_set_tree_sources(galic_ctx *ctx, int nsrc, const double *spos, const double *smass);
/* The sources drawn on the device: SampleForceNhalo / SampleForceNdisk / SampleForceNbulge
 * points of each component with mass (halo.c:17-40, disk.c:17-51, bulge.c:17-38), each of
 * mass MType / count, from Philox draws of (point id, round) under (seed, 200 + type). */
int galic_sample_tree_sources(galic_ctx *ctx, unsigned int seed);
/* forcedensitygrid_calculate's table fill (grid.c:621-757) from the context's tree sources,
 * installed as the context's force tables (FG_Pot, FG_DPotDR, FG_DPotDz and the FP32 copy)
 * without a host round trip: the per-iteration re-tabulation.  The targets and the domain are
 * those of galic_tree_fill, so the tables equal its output on the same sources.
 * precision: GALIC_TREE_ modes; ninteractions optional. */
int galic_retabulate(galic_ctx *ctx, int precision, long long *ninteractions);
/* Reads the context's force tables (any pointer may be NULL). */
int galic_get_force_tables(galic_ctx *ctx, double *pot, double *dpotdr, double *dpotdz);

/* ---- measurement ---- */

/* Device time of the last launches, measured with HIP events on the context's
 * stream: out[0] orbit kernel ms (last launch), out[1] merit kernel ms,
 * out[2] fit ms, out[3] accumulated orbit-kernel ms, out[4] orbit-kernel
 * launches, out[5] orbits integrated, out[6] leapfrog steps integrated,
 * out[7] accumulated merit-kernel ms, out[8] accumulated fit ms, out[9] orbit
 * launches repeated to give the response logs more room, out[10] graphs of reps
 * launched (their whole time, fit included, counts as merit time). */
int galic_timings(galic_ctx *ctx, double *out11);
/* A plain device copy of nbytes (reps times): the known-traffic kernel that
 * calibrates the HBM byte counters; ms_per_copy receives the mean duration. */
int galic_copy_probe(int device, size_t nbytes, int reps, double *ms_per_copy);
int galic_reset_timings(galic_ctx *ctx);

/* ---- drop-in entry points with the reference signatures ----
 * They act on the context made current by galic_bind(), as the reference's
 * versions act on its globals. */
int galic_bind(galic_ctx *ctx);

/* orbit_response.c:155-491 (proto.h:145) */
double produce_orbit_response_field_mod(double *pos, double *vel, int id, double *mfield, double *egyfield_r,
                                        double *egyfield_t, double *egyfield_q, double *egyfield_p, double mass,
                                        double timespan, int *orbitstaken, int type);
/* orbit_response.c:33-151 (proto.h:143): the mass field only */
double produce_orbit_response_field(double *pos, double *vel, int id, double *mfield, double mass, double timespan,
                                    int *orbitstaken);
/* grid.c:59-197 (proto.h:100-103) */
double forcegrid_get_potential(double *pos);
double forcegrid_get_escape_speed(double *pos);
void forcegrid_get_acceleration(double *pos, double *acc);
/* grid.c:1247-1265 (proto.h:120) */
void smooth_stack(double *data, int maxlevel);
/* grid.c:15-55, 494-534, 896-938 (proto.h): the cell and fractions of pos on the
 * force, density and energy grids of the bound context */
void forcegrid_get_cell(double *pos, int *iR, int *iz, double *fR, double *fz);
void densitygrid_get_cell(double *pos, int *iR, int *iz, double *fR, double *fz);
void energygrid_get_cell(double *pos, int *iR, int *iz, double *fR, double *fz);
/* orbit_response.c:16-26 (proto.h): with the bound context's DG_CellSize */
double get_timestep(double *pos, double *vel, double *acc, int icell);
/* grid.c:329-487 (proto.h:104): the density-grid geometry (DG_CellSize, DGs_Distance) of the
 * bound context recomputed from its parameters (galic_grid_create fills Rmax, Rmin, Fac) */
void forcedensitygrid_create(void);
/* grid.c:621-757 (proto.h:105): galic_retabulate of the bound context with the reference's
 * walk (FP64, monopole, spline); sources from galic_set_tree_sources, else drawn by
 * galic_sample_tree_sources with seed 42 */
void forcedensitygrid_calculate(void);

/* ---- forcetree module (forcetree.h:48-65, proto.h:195) ----
 * P[] of these drop-ins is bound by the caller, as the reference's global particle array:
 * n particles, positions (3 each), masses, types; Type-5 particles are the walk targets and
 * are not inserted (forcetree.c:331, gravtree.c:115-119), and the walks write their
 * GravAccel (3 each) and Potential.  The tree parameters (ErrTolTheta, ForceSoftening, G) and
 * the GPU are the bound context's (galic_bind). */
int galic_bind_tree_particles(int n, double *pos, double *mass, int *type, double *grav_accel, double *potential);
/* forcetree.c:41: builds the tree of P[0..npart) on the GPU (the reference's construction:
 * the unique insertion tree, FP64 monopole moments); returns the node count */
int force_treebuild(int npart, int optimized_domain_mapping);
/* forcetree.c:591: the moments and the walk threading of the tree recomputed from the root;
 * *last receives the particle that ends the walk order.  The device tree numbers its nodes
 * its own way: no must be the root, npart of the last force_treebuild (the reference's
 * Tree_MaxPart), with sib = father = -1 as forcetree.c:86 calls it. */
void force_update_node_recursive(int no, int sib, int father, int *last);
/* forcetree_walk.c:12: the walk for TargetList[target] (the target-th Type-5 particle), its
 * GravAccel and Potential written without the factor G; returns the interaction count.
 * mode 0 only (mode 1 walks targets imported from other MPI tasks) */
int force_treeevaluate(int target, int mode, int thread_id);
/* gravtree.c:35-50, 88-467: tree of P[], every Type-5 particle walked, results times G */
void gravity(void);

/* grid.c:1086-1112 and 1115-1151 (proto.h): R200 and LowerDispLimit of the bound context */
double calc_stack_difference(double *d1, double *d2, int l, int i, int j, int maxlevel, double *ref1, double *ref2,
                             double thresh, double *dist, int flag);
double calc_stack_difference_mod(double *d1, double *d2, int l, int i, int j, int maxlevel, double *ref1,
                                 double *ref2, double thresh, double *dist, int flag);

#ifdef __cplusplus
}
#endif

#endif /* GALIC_HIP_H */
