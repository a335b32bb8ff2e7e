/*
 * galic_oracle.c -- TEST INFRASTRUCTURE ONLY (see galic_oracle.h).
 *
 * Scalar FP64 restatement of the GALIC v1.1 hot path.  Operation order follows
 * the reference so that results agree to rounding; each function cites the
 * reference lines it follows.  Compiled with -ffp-contract=off so no fused
 * multiply-adds change the rounding relative to the reference's x86-64 build.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#include "galic_oracle.h"

#define OR_MAXITER 1000 /* allvars.h:270 */

static double or_dmax(double a, double b) { return (a > b) ? a : b; } /* system.c:38-44 */
static double or_dmin(double a, double b) { return (a < b) ? a : b; } /* system.c:46-52 */

struct oracle_model {
  galic_params p;
  int FG_Ngrid;
  int DG_Nbin, DG_Ngrid, DG_Nstack;
  int EG_Nbin, EG_Ngrid, EG_Nstack;
  double *FG_Pot, *FG_DPotDR, *FG_DPotDz;
  double *FG_Vstream[GALIC_NTYPES];
  double *DG_CellSize, *DG_CellVol, *DGs_Distance;

  double *DGs_MassTarget[GALIC_NTYPES], *DGs_MassResponse[GALIC_NTYPES];
  double *EGs_MassTarget[GALIC_NTYPES], *EGs_MassResponse[GALIC_NTYPES];
  double *EGs_EgyTarget_r[GALIC_NTYPES], *EGs_EgyTarget_t[GALIC_NTYPES];
  double *EGs_EgyTarget_p[GALIC_NTYPES], *EGs_EgyTarget_q[GALIC_NTYPES];
  double *EGs_EgyResponse_r[GALIC_NTYPES], *EGs_EgyResponse_t[GALIC_NTYPES];
  double *EGs_EgyResponse_p[GALIC_NTYPES], *EGs_EgyResponse_q[GALIC_NTYPES];
  double *EGs_EgyResponseRS[GALIC_NTYPES], *EGs_EgyResponseTS[GALIC_NTYPES];
  double *EGs_EgyResponseQS[GALIC_NTYPES], *EGs_EgyResponsePS[GALIC_NTYPES];

  double *DG_MassLoc[GALIC_NTYPES], *DG_MassLoc_delta[GALIC_NTYPES];
  double *EG_MassLoc[GALIC_NTYPES];
  double *EG_RLoc[GALIC_NTYPES], *EG_TLoc[GALIC_NTYPES], *EG_PLoc[GALIC_NTYPES], *EG_QLoc[GALIC_NTYPES];
  double *EG_RLoc_d[GALIC_NTYPES], *EG_TLoc_d[GALIC_NTYPES], *EG_PLoc_d[GALIC_NTYPES], *EG_QLoc_d[GALIC_NTYPES];
  double *EG_RLocS[GALIC_NTYPES], *EG_TLocS[GALIC_NTYPES], *EG_QLocS[GALIC_NTYPES], *EG_PLocS[GALIC_NTYPES];
  double *EG_RLocS_d[GALIC_NTYPES], *EG_TLocS_d[GALIC_NTYPES], *EG_QLocS_d[GALIC_NTYPES], *EG_PLocS_d[GALIC_NTYPES];

  double S[GALIC_NTYPES], Sdisp_r[GALIC_NTYPES], Sdisp_t[GALIC_NTYPES], Sdisp_p[GALIC_NTYPES], Sdisp_q[GALIC_NTYPES];
  double Srelfac[GALIC_NTYPES], Srelfac_count[GALIC_NTYPES];
  int firstrun[GALIC_NTYPES];

  int N;
  double *pos, *vel, *mass, *tint, *vesc;
  int *type, *orbits;
  double *vr2, *vt2, *vp2, *vq2;

  /* target dispersion tables (disp_fields.c): FG_Disp_r [FG_Nbin], FG_DispZ, FG_DispPhi,
   * FG_tilted_vR2_prime, FG_tilted_vz2_prime [FG_Nbin^2]; beta per type */
  double *Disp_r[GALIC_NTYPES], *DispZ[GALIC_NTYPES], *DispPhi[GALIC_NTYPES];
  double *TiltR[GALIC_NTYPES], *TiltTheta[GALIC_NTYPES];
  double beta[GALIC_NTYPES];
};

/* ------------------------------------------------------------------------- */
/* grid geometry                                                             */
/* ------------------------------------------------------------------------- */

/* Newton iteration for the spacing factor, grid.c:353-374 (DG), 424-445 (FG), 456-477 (EG). */
double or_solve_fac(int nbin, double rin, double rmax, double *rmin_out)
{
  int iter = 0;
  double dfac, fac = 1.02;
  double y = (rmax / rin);
  do {
    double f = log((pow(fac, nbin) - 1.0) / (y * (fac - 1.0)));
    double df = nbin / (fac - 1.0 / pow(fac, nbin - 1)) - 1.0 / (fac - 1.0);
    dfac = -f / df;
    fac += dfac;
    iter++;
    if(iter > OR_MAXITER)
      return -1;
  } while(fabs(dfac) > 1.0e-8 * fac);
  *rmin_out = rin / (fac - 1.0);
  return fac;
}

/* grid.c:334-345 (Rmax, Rin) followed by the three Newton solves. */
int or_grid_create(galic_params *p, double outer_frac, double inner_frac)
{
  double fac = outer_frac;
  p->Rmax = p->Halo_A * (fac + sqrt(fac)) / (1 - fac);
  fac = inner_frac;
  double rin = p->Halo_A * (fac + sqrt(fac)) / (1 - fac);
  p->DG_Fac = or_solve_fac(1 << p->DG_MaxLevel, rin, p->Rmax, &p->DG_Rmin);
  p->FG_Fac = or_solve_fac(p->FG_Nbin, rin, p->Rmax, &p->FG_Rmin);
  p->EG_Fac = or_solve_fac(1 << p->EG_MaxLevel, rin, p->Rmax, &p->EG_Rmin);
  if(p->DG_Fac < 0 || p->FG_Fac < 0 || p->EG_Fac < 0)
    return -1;
  return 0;
}

/* DG_CellSize/DG_CellVol (grid.c:830-850) and the DGs_Distance stack (grid.c:385-416). */
static void or_density_geometry(oracle_model *m)
{
  int L = m->p.DG_MaxLevel, nb = m->DG_Nbin, i, j, k, l;
  double *logR = calloc(m->DG_Nstack, sizeof(double));
  double *logZ = calloc(m->DG_Nstack, sizeof(double));

  for(k = 0; k < nb; k++)
    for(j = 0; j < nb; j++) {
      double z = m->p.DG_Rmin * 0.5 * (pow(m->p.DG_Fac, k) + pow(m->p.DG_Fac, k + 1) - 2.0);
      double R = m->p.DG_Rmin * 0.5 * (pow(m->p.DG_Fac, j) + pow(m->p.DG_Fac, j + 1) - 2.0);
      logR[GALIC_STACKOFFSET(L, k, j)] = log(R);
      logZ[GALIC_STACKOFFSET(L, k, j)] = log(z);
    }
  or_smooth_stack(logR, L);
  or_smooth_stack(logZ, L);
  for(l = L - 1; l >= 0; l--) {
    int n = (1 << l);
    int f = (1 << (L - l));
    for(i = 0; i < n; i++)
      for(j = 0; j < n; j++) {
        logR[GALIC_STACKOFFSET(l, i, j)] /= (f * f);
        logZ[GALIC_STACKOFFSET(l, i, j)] /= (f * f);
      }
  }
  for(l = L; l >= 0; l--) {
    int n = (1 << l);
    for(i = 0; i < n; i++)
      for(j = 0; j < n; j++)
        m->DGs_Distance[GALIC_STACKOFFSET(l, i, j)] =
            sqrt(pow(exp(logR[GALIC_STACKOFFSET(l, i, j)]), 2) + pow(exp(logZ[GALIC_STACKOFFSET(l, i, j)]), 2));
  }
  free(logR);
  free(logZ);

  for(k = 0; k < nb; k++)
    for(j = 0; j < nb; j++) {
      i = k * nb + j;
      double r1 = m->p.DG_Rmin * (pow(m->p.DG_Fac, j) - 1.0);
      double r2 = m->p.DG_Rmin * (pow(m->p.DG_Fac, j + 1) - 1.0);
      double z1 = m->p.DG_Rmin * (pow(m->p.DG_Fac, k) - 1.0);
      double z2 = m->p.DG_Rmin * (pow(m->p.DG_Fac, k + 1) - 1.0);
      double vol = M_PI * (r2 * r2 - r1 * r1) * (z2 - z1);
      vol *= 2;
      m->DG_CellSize[i] = or_dmax(r2 - r1, z2 - z1);
      m->DG_CellVol[i] = vol;
    }
}

/* ------------------------------------------------------------------------- */
/* model lifetime                                                            */
/* ------------------------------------------------------------------------- */

#define OR_ZALLOC(n) ((double *)calloc((size_t)(n), sizeof(double)))

oracle_model *or_model_new(const galic_params *p)
{
  oracle_model *m = calloc(1, sizeof(*m));
  int t;
  m->p = *p;
  m->FG_Ngrid = p->FG_Nbin * p->FG_Nbin;
  m->DG_Nbin = 1 << p->DG_MaxLevel;
  m->DG_Ngrid = m->DG_Nbin * m->DG_Nbin;
  m->DG_Nstack = GALIC_NSTACK(p->DG_MaxLevel);
  m->EG_Nbin = 1 << p->EG_MaxLevel;
  m->EG_Ngrid = m->EG_Nbin * m->EG_Nbin;
  m->EG_Nstack = GALIC_NSTACK(p->EG_MaxLevel);
  m->FG_Pot = OR_ZALLOC(m->FG_Ngrid);
  m->FG_DPotDR = OR_ZALLOC(m->FG_Ngrid);
  m->FG_DPotDz = OR_ZALLOC(m->FG_Ngrid);
  m->DG_CellSize = OR_ZALLOC(m->DG_Ngrid);
  m->DG_CellVol = OR_ZALLOC(m->DG_Ngrid);
  m->DGs_Distance = OR_ZALLOC(m->DG_Nstack);
  for(t = 1; t < GALIC_NTYPES; t++) {
    m->FG_Vstream[t] = OR_ZALLOC(m->FG_Ngrid);
    m->DGs_MassTarget[t] = OR_ZALLOC(m->DG_Nstack);
    m->DGs_MassResponse[t] = OR_ZALLOC(m->DG_Nstack);
    m->EGs_MassTarget[t] = OR_ZALLOC(m->EG_Nstack);
    m->EGs_MassResponse[t] = OR_ZALLOC(m->EG_Nstack);
    m->EGs_EgyTarget_r[t] = OR_ZALLOC(m->EG_Nstack);
    m->EGs_EgyTarget_t[t] = OR_ZALLOC(m->EG_Nstack);
    m->EGs_EgyTarget_p[t] = OR_ZALLOC(m->EG_Nstack);
    m->EGs_EgyTarget_q[t] = OR_ZALLOC(m->EG_Nstack);
    m->EGs_EgyResponse_r[t] = OR_ZALLOC(m->EG_Nstack);
    m->EGs_EgyResponse_t[t] = OR_ZALLOC(m->EG_Nstack);
    m->EGs_EgyResponse_p[t] = OR_ZALLOC(m->EG_Nstack);
    m->EGs_EgyResponse_q[t] = OR_ZALLOC(m->EG_Nstack);
    m->EGs_EgyResponseRS[t] = OR_ZALLOC(m->EG_Nstack);
    m->EGs_EgyResponseTS[t] = OR_ZALLOC(m->EG_Nstack);
    m->EGs_EgyResponseQS[t] = OR_ZALLOC(m->EG_Nstack);
    m->EGs_EgyResponsePS[t] = OR_ZALLOC(m->EG_Nstack);
    m->DG_MassLoc[t] = OR_ZALLOC(m->DG_Ngrid);
    m->DG_MassLoc_delta[t] = OR_ZALLOC(m->DG_Ngrid);
    m->EG_MassLoc[t] = OR_ZALLOC(m->EG_Ngrid);
    m->EG_RLoc[t] = OR_ZALLOC(m->EG_Ngrid);
    m->EG_TLoc[t] = OR_ZALLOC(m->EG_Ngrid);
    m->EG_PLoc[t] = OR_ZALLOC(m->EG_Ngrid);
    m->EG_QLoc[t] = OR_ZALLOC(m->EG_Ngrid);
    m->EG_RLoc_d[t] = OR_ZALLOC(m->EG_Ngrid);
    m->EG_TLoc_d[t] = OR_ZALLOC(m->EG_Ngrid);
    m->EG_PLoc_d[t] = OR_ZALLOC(m->EG_Ngrid);
    m->EG_QLoc_d[t] = OR_ZALLOC(m->EG_Ngrid);
    m->EG_RLocS[t] = OR_ZALLOC(m->EG_Ngrid);
    m->EG_TLocS[t] = OR_ZALLOC(m->EG_Ngrid);
    m->EG_QLocS[t] = OR_ZALLOC(m->EG_Ngrid);
    m->EG_PLocS[t] = OR_ZALLOC(m->EG_Ngrid);
    m->EG_RLocS_d[t] = OR_ZALLOC(m->EG_Ngrid);
    m->EG_TLocS_d[t] = OR_ZALLOC(m->EG_Ngrid);
    m->EG_QLocS_d[t] = OR_ZALLOC(m->EG_Ngrid);
    m->EG_PLocS_d[t] = OR_ZALLOC(m->EG_Ngrid);
  }
  or_density_geometry(m);
  return m;
}

void or_model_free(oracle_model *m)
{
  int t;
  if(!m)
    return;
  free(m->FG_Pot); free(m->FG_DPotDR); free(m->FG_DPotDz);
  free(m->DG_CellSize); free(m->DG_CellVol); free(m->DGs_Distance);
  for(t = 1; t < GALIC_NTYPES; t++) {
    free(m->FG_Vstream[t]);
    free(m->DGs_MassTarget[t]); free(m->DGs_MassResponse[t]);
    free(m->EGs_MassTarget[t]); free(m->EGs_MassResponse[t]);
    free(m->EGs_EgyTarget_r[t]); free(m->EGs_EgyTarget_t[t]); free(m->EGs_EgyTarget_p[t]); free(m->EGs_EgyTarget_q[t]);
    free(m->EGs_EgyResponse_r[t]); free(m->EGs_EgyResponse_t[t]); free(m->EGs_EgyResponse_p[t]); free(m->EGs_EgyResponse_q[t]);
    free(m->EGs_EgyResponseRS[t]); free(m->EGs_EgyResponseTS[t]); free(m->EGs_EgyResponseQS[t]); free(m->EGs_EgyResponsePS[t]);
    free(m->DG_MassLoc[t]); free(m->DG_MassLoc_delta[t]); free(m->EG_MassLoc[t]);
    free(m->EG_RLoc[t]); free(m->EG_TLoc[t]); free(m->EG_PLoc[t]); free(m->EG_QLoc[t]);
    free(m->EG_RLoc_d[t]); free(m->EG_TLoc_d[t]); free(m->EG_PLoc_d[t]); free(m->EG_QLoc_d[t]);
    free(m->EG_RLocS[t]); free(m->EG_TLocS[t]); free(m->EG_QLocS[t]); free(m->EG_PLocS[t]);
    free(m->EG_RLocS_d[t]); free(m->EG_TLocS_d[t]); free(m->EG_QLocS_d[t]); free(m->EG_PLocS_d[t]);
    free(m->Disp_r[t]); free(m->DispZ[t]); free(m->DispPhi[t]); free(m->TiltR[t]); free(m->TiltTheta[t]);
  }
  free(m->pos); free(m->vel); free(m->mass); free(m->tint); free(m->vesc);
  free(m->type); free(m->orbits);
  free(m->vr2); free(m->vt2); free(m->vp2); free(m->vq2);
  free(m);
}

const galic_params *or_model_params(const oracle_model *m) { return &m->p; }

void or_set_force_tables(oracle_model *m, const double *pot, const double *dpotdr, const double *dpotdz)
{
  memcpy(m->FG_Pot, pot, sizeof(double) * m->FG_Ngrid);
  memcpy(m->FG_DPotDR, dpotdr, sizeof(double) * m->FG_Ngrid);
  memcpy(m->FG_DPotDz, dpotdz, sizeof(double) * m->FG_Ngrid);
}

void or_set_vstream(oracle_model *m, int type, const double *vstream)
{
  if(vstream)
    memcpy(m->FG_Vstream[type], vstream, sizeof(double) * m->FG_Ngrid);
  else
    memset(m->FG_Vstream[type], 0, sizeof(double) * m->FG_Ngrid);
}

void or_set_targets(oracle_model *m, int type, const double *dgs_mass_target, const double *egs_mass_target,
                    const double *egs_r, const double *egs_t, const double *egs_p, const double *egs_q)
{
  memcpy(m->DGs_MassTarget[type], dgs_mass_target, sizeof(double) * m->DG_Nstack);
  memcpy(m->EGs_MassTarget[type], egs_mass_target, sizeof(double) * m->EG_Nstack);
  memcpy(m->EGs_EgyTarget_r[type], egs_r, sizeof(double) * m->EG_Nstack);
  memcpy(m->EGs_EgyTarget_t[type], egs_t, sizeof(double) * m->EG_Nstack);
  memcpy(m->EGs_EgyTarget_p[type], egs_p, sizeof(double) * m->EG_Nstack);
  memcpy(m->EGs_EgyTarget_q[type], egs_q, sizeof(double) * m->EG_Nstack);
}

void or_set_particles(oracle_model *m, int n, const double *pos, const double *vel, const double *mass,
                      const double *tint, const int *type, const double *vesc)
{
  free(m->pos); free(m->vel); free(m->mass); free(m->tint); free(m->vesc);
  free(m->type); free(m->orbits);
  free(m->vr2); free(m->vt2); free(m->vp2); free(m->vq2);
  m->N = n;
  m->pos = malloc(sizeof(double) * 3 * (size_t)n);
  m->vel = malloc(sizeof(double) * 3 * (size_t)n);
  m->mass = malloc(sizeof(double) * (size_t)n);
  m->tint = malloc(sizeof(double) * (size_t)n);
  m->vesc = malloc(sizeof(double) * (size_t)n);
  m->type = malloc(sizeof(int) * (size_t)n);
  m->orbits = calloc((size_t)n, sizeof(int));
  m->vr2 = OR_ZALLOC(n); m->vt2 = OR_ZALLOC(n); m->vp2 = OR_ZALLOC(n); m->vq2 = OR_ZALLOC(n);
  memcpy(m->pos, pos, sizeof(double) * 3 * (size_t)n);
  memcpy(m->vel, vel, sizeof(double) * 3 * (size_t)n);
  memcpy(m->mass, mass, sizeof(double) * (size_t)n);
  memcpy(m->tint, tint, sizeof(double) * (size_t)n);
  memcpy(m->type, type, sizeof(int) * (size_t)n);
  if(vesc)
    memcpy(m->vesc, vesc, sizeof(double) * (size_t)n);
  else
    memset(m->vesc, 0, sizeof(double) * (size_t)n);
}

const double *or_cellsize(const oracle_model *m) { return m->DG_CellSize; }
const double *or_distance(const oracle_model *m) { return m->DGs_Distance; }

const double *or_stack(const oracle_model *m, int type, int which)
{
  switch(which) {
    case 0: return m->DGs_MassResponse[type];
    case 1: return m->EGs_MassResponse[type];
    case 2: return m->EGs_EgyResponse_r[type];
    case 3: return m->EGs_EgyResponse_t[type];
    case 4: return m->EGs_EgyResponse_p[type];
    case 5: return m->EGs_EgyResponse_q[type];
    case 6: return m->EGs_EgyResponseRS[type];
    case 7: return m->EGs_EgyResponseTS[type];
    case 8: return m->EGs_EgyResponseQS[type];
    case 9: return m->EGs_EgyResponsePS[type];
    case 10: return m->DGs_MassTarget[type];
    case 11: return m->EGs_MassTarget[type];
    case 12: return m->EGs_EgyTarget_r[type];
    case 13: return m->EGs_EgyTarget_t[type];
    case 14: return m->EGs_EgyTarget_p[type];
    case 15: return m->EGs_EgyTarget_q[type];
  }
  return NULL;
}

double or_fit_value(const oracle_model *m, int t, int which)
{
  switch(which) {
    case 0: return m->S[t];
    case 1: return m->Sdisp_r[t];
    case 2: return m->Sdisp_t[t];
    case 3: return m->Sdisp_p[t];
    case 4: return m->Sdisp_q[t];
    case 5: return m->Srelfac[t];
    case 6: return m->Srelfac_count[t];
    case 7: /* logged Sall, main.c:1270 */
      return m->S[t] + (m->Sdisp_r[t] + m->Sdisp_t[t] + m->Sdisp_p[t] + m->Sdisp_q[t]) / m->Srelfac_count[t];
  }
  return 0;
}

void or_get_particles(const oracle_model *m, double *vel, int *orbits, double *vr2, double *vt2, double *vp2,
                      double *vq2)
{
  if(vel) memcpy(vel, m->vel, sizeof(double) * 3 * (size_t)m->N);
  if(orbits) memcpy(orbits, m->orbits, sizeof(int) * (size_t)m->N);
  if(vr2) memcpy(vr2, m->vr2, sizeof(double) * (size_t)m->N);
  if(vt2) memcpy(vt2, m->vt2, sizeof(double) * (size_t)m->N);
  if(vp2) memcpy(vp2, m->vp2, sizeof(double) * (size_t)m->N);
  if(vq2) memcpy(vq2, m->vq2, sizeof(double) * (size_t)m->N);
}

/* ------------------------------------------------------------------------- */
/* lookups                                                                   */
/* ------------------------------------------------------------------------- */

/* halo.c:63-84 */
double or_halo_get_potential(const galic_params *p, const double *pos)
{
  double r = sqrt(pos[0] * pos[0] + pos[1] * pos[1] + pos[2] * pos[2]);
  return -p->G * p->Halo_Mass / (r + p->Halo_A);
}

void or_halo_get_acceleration(const galic_params *p, const double *pos, double *acc)
{
  double r = sqrt(pos[0] * pos[0] + pos[1] * pos[1] + pos[2] * pos[2]);
  double fac = p->G * p->Halo_Mass / ((r + 1.0e-6 * p->Halo_A) * (r + p->Halo_A) * (r + p->Halo_A));
  acc[0] = -fac * pos[0];
  acc[1] = -fac * pos[1];
  acc[2] = -fac * pos[2];
}

/* bulge.c:64-85 */
double or_bulge_get_potential(const galic_params *p, const double *pos)
{
  double r = sqrt(pos[0] * pos[0] + pos[1] * pos[1] + pos[2] * pos[2]);
  return -p->G * p->Bulge_Mass / (r + p->Bulge_A);
}

void or_bulge_get_acceleration(const galic_params *p, const double *pos, double *acc)
{
  double r = sqrt(pos[0] * pos[0] + pos[1] * pos[1] + pos[2] * pos[2]);
  double fac = p->G * p->Bulge_Mass / ((r + 1.0e-6 * p->Bulge_A) * (r + p->Bulge_A) * (r + p->Bulge_A));
  acc[0] = -fac * pos[0];
  acc[1] = -fac * pos[1];
  acc[2] = -fac * pos[2];
}

static int or_uses_table(const galic_params *p)
{
  /* grid.c:63 and grid.c:124 */
  return (p->Disk_Mass > 0 || (p->Halo_Mass > 0 && p->SampleForceNhalo > 0) ||
          (p->Bulge_Mass > 0 && p->SampleForceNbulge > 0));
}

/* grid.c:15-55 (clamping only; the terminate() branches cannot trigger after clamping br, bz at 0) */
void or_forcegrid_get_cell(const oracle_model *m, const double *pos, int *iR, int *iz, double *fR, double *fz)
{
  double r = sqrt(pos[0] * pos[0] + pos[1] * pos[1]);
  double z = fabs(pos[2]);
  double br = (log(r / m->p.FG_Rmin + 1.0) / log(m->p.FG_Fac));
  double bz = (log(z / m->p.FG_Rmin + 1.0) / log(m->p.FG_Fac));
  int binr, binz;
  if(br < 0) br = 0;
  if(bz < 0) bz = 0;
  binr = (int)br;
  *fR = br - binr;
  binz = (int)bz;
  *fz = bz - binz;
  if(binr >= m->p.FG_Nbin - 1) { binr = m->p.FG_Nbin - 2; *fR = 1; }
  if(binz >= m->p.FG_Nbin - 1) { binz = m->p.FG_Nbin - 2; *fz = 1; }
  *iR = binr;
  *iz = binz;
}

/* grid.c:59-113 */
double or_forcegrid_get_potential(const oracle_model *m, const double *pos)
{
  const galic_params *p = &m->p;
  double pot;
  int nb = p->FG_Nbin;
  if(or_uses_table(p)) {
    double R = sqrt(pos[0] * pos[0] + pos[1] * pos[1]);
    double jdbl = log(R / p->FG_Rmin + 1.0) / log(p->FG_Fac);
    int jbin = (int)jdbl;
    double jfac = jdbl - jbin;
    double idbl = log(fabs(pos[2]) / p->FG_Rmin + 1.0) / log(p->FG_Fac);
    int ibin = (int)idbl;
    double ifac = idbl - ibin;
    if(ibin < nb - 1 && jbin < nb - 1) {
      pot = m->FG_Pot[ibin * nb + jbin] * (1 - ifac) * (1 - jfac) + m->FG_Pot[ibin * nb + (jbin + 1)] * (1 - ifac) * jfac +
            m->FG_Pot[(ibin + 1) * nb + jbin] * ifac * (1 - jfac) + m->FG_Pot[(ibin + 1) * nb + (jbin + 1)] * ifac * jfac;
    } else {
      pot = p->M200 / p->Halo_Mass * or_halo_get_potential(p, pos);
    }
  } else
    pot = 0;
  if(p->Halo_Mass > 0 && p->SampleForceNhalo == 0)
    pot += or_halo_get_potential(p, pos);
  if(p->Bulge_Mass > 0 && p->SampleForceNbulge == 0)
    pot += or_bulge_get_potential(p, pos);
  return pot;
}

/* grid.c:122-197 */
void or_forcegrid_get_acceleration(const oracle_model *m, const double *pos, double *acc)
{
  const galic_params *p = &m->p;
  int nb = p->FG_Nbin;
  if(or_uses_table(p)) {
    double phi = atan2(pos[1], pos[0]);
    double R = sqrt(pos[0] * pos[0] + pos[1] * pos[1]);
    double jdbl = log(R / p->FG_Rmin + 1.0) / log(p->FG_Fac);
    int jbin = (int)jdbl;
    double jfac = jdbl - jbin;
    double idbl = log(fabs(pos[2]) / p->FG_Rmin + 1.0) / log(p->FG_Fac);
    int ibin = (int)idbl;
    double ifac = idbl - ibin;
    if(ibin < nb - 1 && jbin < nb - 1) {
      double accR = m->FG_DPotDR[ibin * nb + jbin] * (1 - ifac) * (1 - jfac) +
                    m->FG_DPotDR[ibin * nb + (jbin + 1)] * (1 - ifac) * jfac +
                    m->FG_DPotDR[(ibin + 1) * nb + jbin] * ifac * (1 - jfac) +
                    m->FG_DPotDR[(ibin + 1) * nb + (jbin + 1)] * ifac * jfac;
      double accz = m->FG_DPotDz[ibin * nb + jbin] * (1 - ifac) * (1 - jfac) +
                    m->FG_DPotDz[ibin * nb + (jbin + 1)] * (1 - ifac) * jfac +
                    m->FG_DPotDz[(ibin + 1) * nb + jbin] * ifac * (1 - jfac) +
                    m->FG_DPotDz[(ibin + 1) * nb + (jbin + 1)] * ifac * jfac;
      if(pos[2] < 0)
        accz = -accz;
      acc[0] = accR * cos(phi);
      acc[1] = accR * sin(phi);
      acc[2] = accz;
    } else {
      int k;
      or_halo_get_acceleration(p, pos, acc);
      double fac = p->M200 / p->Halo_Mass;
      for(k = 0; k < 3; k++)
        acc[k] *= fac;
    }
  } else {
    acc[0] = acc[1] = acc[2] = 0;
  }
  if(p->Halo_Mass > 0 && p->SampleForceNhalo == 0) {
    double acc_h[3];
    or_halo_get_acceleration(p, pos, acc_h);
    acc[0] += acc_h[0];
    acc[1] += acc_h[1];
    acc[2] += acc_h[2];
  }
  if(p->Bulge_Mass > 0 && p->SampleForceNbulge == 0) {
    double acc_b[3];
    or_bulge_get_acceleration(p, pos, acc_b);
    acc[0] += acc_b[0];
    acc[1] += acc_b[1];
    acc[2] += acc_b[2];
  }
}

/* grid.c:494-534 */
void or_densitygrid_get_cell(const oracle_model *m, const double *pos, int *iR, int *iz, double *fR, double *fz)
{
  double r = sqrt(pos[0] * pos[0] + pos[1] * pos[1]);
  double z = fabs(pos[2]);
  double br = (log(r / m->p.DG_Rmin + 1.0) / log(m->p.DG_Fac)) - 0.5;
  double bz = (log(z / m->p.DG_Rmin + 1.0) / log(m->p.DG_Fac)) - 0.5;
  int binr, binz;
  if(br < 0) br = 0;
  if(bz < 0) bz = 0;
  binr = (int)br;
  *fR = br - binr;
  binz = (int)bz;
  *fz = bz - binz;
  if(binr >= m->DG_Nbin - 1) { binr = m->DG_Nbin - 2; *fR = 1; }
  if(binz >= m->DG_Nbin - 1) { binz = m->DG_Nbin - 2; *fz = 1; }
  *iR = binr;
  *iz = binz;
}

/* grid.c:896-938 */
void or_energygrid_get_cell(const oracle_model *m, const double *pos, int *iR, int *iz, double *fR, double *fz)
{
  double r = sqrt(pos[0] * pos[0] + pos[1] * pos[1]);
  double z = fabs(pos[2]);
  double br = (log(r / m->p.EG_Rmin + 1.0) / log(m->p.EG_Fac)) - 0.5;
  double bz = (log(z / m->p.EG_Rmin + 1.0) / log(m->p.EG_Fac)) - 0.5;
  int binr, binz;
  if(br < 0) br = 0;
  if(bz < 0) bz = 0;
  binr = (int)br;
  *fR = br - binr;
  binz = (int)bz;
  *fz = bz - binz;
  if(binr >= m->EG_Nbin - 1) { binr = m->EG_Nbin - 2; *fR = 1; }
  if(binz >= m->EG_Nbin - 1) { binz = m->EG_Nbin - 2; *fz = 1; }
  *iR = binr;
  *iz = binz;
}

/* disp_fields.c:118-132 */
double or_get_vstream(const oracle_model *m, const double *pos, int type)
{
  int iz, iR, nb = m->p.FG_Nbin;
  double fR, fz;
  const double *vs = m->FG_Vstream[type];
  or_forcegrid_get_cell(m, pos, &iR, &iz, &fR, &fz);
  return vs[iz * nb + iR] * (1 - fR) * (1 - fz) + vs[(iz + 1) * nb + iR] * (1 - fR) * (fz) +
         vs[iz * nb + (iR + 1)] * (fR) * (1 - fz) + vs[(iz + 1) * nb + (iR + 1)] * (fR) * (fz);
}

/* orbit_response.c:16-26 */
double or_get_timestep(const oracle_model *m, const double *pos, const double *vel, const double *acc, int icell)
{
  (void)pos;
  double v = sqrt(vel[0] * vel[0] + vel[1] * vel[1] + vel[2] * vel[2]);
  double aa = sqrt(acc[0] * acc[0] + acc[1] * acc[1] + acc[2] * acc[2]);
  double torbit = m->p.V200 / aa;
  double tcross = m->DG_CellSize[icell] / v;
  return or_dmin(m->p.TimeStepFactorOrbit * torbit, m->p.TimeStepFactorCellCross * tcross);
}

/* ------------------------------------------------------------------------- */
/* orbit response                                                            */
/* ------------------------------------------------------------------------- */

/* Velocity decomposition used at both ends of a step, orbit_response.c:222-273 / 338-391. */
static void or_orbit_decompose(const oracle_model *m, const double *x, const double *v, int tovs, int type,
                               double *vr2, double *vt2, double *vq2, double *vp2)
{
  double r2, v_dot_x, Q[3], T[3], q2, q, vq, v_dot_Q, t2, v_dot_T, vstr;
  const double Z[] = {0, 0, -1};
  if(tovs == 2) {
    r2 = x[0] * x[0] + x[1] * x[1];
    v_dot_x = v[0] * x[0] + v[1] * x[1];
    *vr2 = v_dot_x * v_dot_x / r2;
    Q[0] = -x[1];
    Q[1] = x[0];
    q2 = Q[0] * Q[0] + Q[1] * Q[1];
    v_dot_Q = v[0] * Q[0] + v[1] * Q[1];
    q = sqrt(q2);
    vq = v_dot_Q / q;
    *vq2 = vq * vq;
    vstr = or_get_vstream(m, x, type);
    *vp2 = (vq - vstr) * (vq - vstr);
    *vt2 = v[2] * v[2];
  } else {
    r2 = x[0] * x[0] + x[1] * x[1] + x[2] * x[2];
    v_dot_x = v[0] * x[0] + v[1] * x[1] + v[2] * x[2];
    *vr2 = v_dot_x * v_dot_x / r2;
    Q[0] = x[1] * Z[2] - x[2] * Z[1];
    Q[1] = x[2] * Z[0] - x[0] * Z[2];
    Q[2] = x[0] * Z[1] - x[1] * Z[0];
    q2 = Q[0] * Q[0] + Q[1] * Q[1] + Q[2] * Q[2];
    v_dot_Q = v[0] * Q[0] + v[1] * Q[1] + v[2] * Q[2];
    q = sqrt(q2);
    vq = v_dot_Q / q;
    *vq2 = vq * vq;
    vstr = or_get_vstream(m, x, type);
    *vp2 = (vq - vstr) * (vq - vstr);
    T[0] = x[1] * Q[2] - x[2] * Q[1];
    T[1] = x[2] * Q[0] - x[0] * Q[2];
    T[2] = x[0] * Q[1] - x[1] * Q[0];
    t2 = T[0] * T[0] + T[1] * T[1] + T[2] * T[2];
    v_dot_T = v[0] * T[0] + v[1] * T[1] + v[2] * T[2];
    *vt2 = v_dot_T * v_dot_T / t2;
  }
}

/* Trapezoid half-weight binning of the five fields, orbit_response.c:276-319 / 393-436. */
static void or_orbit_bin(int nb, int iR, int iz, double fR, double fz, double dt, double vr2, double vt2,
                         double vq2, double vp2, double *mf, double *er, double *et, double *eq, double *ep)
{
  double mm[2][2];
  int irz[2][2];
  mm[0][0] = 0.5 * dt * (1 - fR) * (1 - fz);
  mm[1][0] = 0.5 * dt * (fR) * (1 - fz);
  mm[0][1] = 0.5 * dt * (1 - fR) * (fz);
  mm[1][1] = 0.5 * dt * (fR) * (fz);
  irz[0][0] = iz * nb + iR;
  irz[1][0] = iz * nb + (iR + 1);
  irz[0][1] = (iz + 1) * nb + iR;
  irz[1][1] = (iz + 1) * nb + (iR + 1);
  mf[irz[0][0]] += mm[0][0]; mf[irz[1][0]] += mm[1][0]; mf[irz[0][1]] += mm[0][1]; mf[irz[1][1]] += mm[1][1];
  if(!er)
    return;
  er[irz[0][0]] += mm[0][0] * vr2; er[irz[1][0]] += mm[1][0] * vr2; er[irz[0][1]] += mm[0][1] * vr2; er[irz[1][1]] += mm[1][1] * vr2;
  et[irz[0][0]] += mm[0][0] * vt2; et[irz[1][0]] += mm[1][0] * vt2; et[irz[0][1]] += mm[0][1] * vt2; et[irz[1][1]] += mm[1][1] * vt2;
  eq[irz[0][0]] += mm[0][0] * vq2; eq[irz[1][0]] += mm[1][0] * vq2; eq[irz[0][1]] += mm[0][1] * vq2; eq[irz[1][1]] += mm[1][1] * vq2;
  ep[irz[0][0]] += mm[0][0] * vp2; ep[irz[1][0]] += mm[1][0] * vp2; ep[irz[0][1]] += mm[0][1] * vp2; ep[irz[1][1]] += mm[1][1] * vp2;
}

/* orbit_response.c:155-491 (egy fields present) and :33-151 (egy fields NULL). */
static double or_orbit_core(const oracle_model *m, const double *pos, const double *vel, double *mfield,
                            double *egy_r, double *egy_t, double *egy_q, double *egy_p, double mass,
                            double timespan, int *orbitstaken, int type, long long *steps_out, double *e0_out,
                            double *e1_out)
{
  const galic_params *p = &m->p;
  int tovs = p->TypeOfVelocityStructure[type];
  int i, norbit, icell, flag = 0, iR, iz, nb = m->DG_Nbin;
  double x[3], v[3], a[3], dt, tall, radsign_previous = 0, radsign, fR, fz;
  double vr2 = 0, vt2 = 0, vq2 = 0, vp2 = 0;
  long long steps = 0;
  long long maxsteps = p->MaxStepsPerOrbit > 0 ? p->MaxStepsPerOrbit : 100000000;

  for(i = 0; i < 3; i++) {
    x[i] = pos[i];
    v[i] = vel[i];
  }
  for(i = 0; i < m->DG_Ngrid; i++)
    mfield[i] = 0;
  if(egy_r)
    for(i = 0; i < m->EG_Ngrid; i++) {
      egy_r[i] = 0;
      egy_t[i] = 0;
      egy_q[i] = 0;
      egy_p[i] = 0;
    }
  norbit = 0;
  tall = 0;
  or_forcegrid_get_acceleration(m, x, a);
  or_densitygrid_get_cell(m, x, &iR, &iz, &fR, &fz);
  icell = iz * nb + iR;
  double E0 = 0.5 * (v[0] * v[0] + v[1] * v[1] + v[2] * v[2]) + or_forcegrid_get_potential(m, x);

  do {
    dt = or_get_timestep(m, x, v, a, icell);
    if(0 < timespan)
      if(timespan < dt + tall) {
        dt = timespan - tall;
        flag = 1;
      }
    if(egy_r)
      or_orbit_decompose(m, x, v, tovs, type, &vr2, &vt2, &vq2, &vp2);
    or_orbit_bin(nb, iR, iz, fR, fz, dt, vr2, vt2, vq2, vp2, mfield, egy_r, egy_t, egy_q, egy_p);

    for(i = 0; i < 3; i++)
      v[i] += 0.5 * dt * a[i];
    for(i = 0; i < 3; i++)
      x[i] += dt * v[i];
    or_forcegrid_get_acceleration(m, x, a);
    for(i = 0; i < 3; i++)
      v[i] += 0.5 * dt * a[i];
    or_densitygrid_get_cell(m, x, &iR, &iz, &fR, &fz);
    icell = iz * nb + iR;

    if(egy_r)
      or_orbit_decompose(m, x, v, tovs, type, &vr2, &vt2, &vq2, &vp2);
    or_orbit_bin(nb, iR, iz, fR, fz, dt, vr2, vt2, vq2, vp2, mfield, egy_r, egy_t, egy_q, egy_p);

    tall += dt;
    radsign = v[0] * x[0] + v[1] * x[1] + v[2] * x[2];
    if(radsign > 0 && radsign_previous < 0)
      norbit++;
    radsign_previous = radsign;
    steps++;
    if(steps > maxsteps) {
      if(steps_out) *steps_out = steps;
      return -1;
    }
  } while((timespan == 0 && norbit < 100000000) || (timespan != 0 && flag == 0));

  double E1 = 0.5 * (v[0] * v[0] + v[1] * v[1] + v[2] * v[2]) + or_forcegrid_get_potential(m, x);
  double fac = mass / tall;
  for(i = 0; i < m->DG_Ngrid; i++)
    mfield[i] *= fac;
  if(egy_r)
    for(i = 0; i < m->EG_Ngrid; i++) {
      egy_r[i] *= fac;
      egy_t[i] *= fac;
      egy_q[i] *= fac;
      egy_p[i] *= fac;
    }
  *orbitstaken = norbit;
  if(steps_out) *steps_out = steps;
  if(e0_out) *e0_out = E0;
  if(e1_out) *e1_out = E1;
  return tall;
}

double or_orbit_response_mod(const oracle_model *m, const double *pos, const double *vel, double *mfield,
                             double *egy_r, double *egy_t, double *egy_q, double *egy_p, double mass,
                             double timespan, int *orbitstaken, int type, long long *steps_out, double *e0_out,
                             double *e1_out)
{
  return or_orbit_core(m, pos, vel, mfield, egy_r, egy_t, egy_q, egy_p, mass, timespan, orbitstaken, type,
                       steps_out, e0_out, e1_out);
}

double or_orbit_response_plain(const oracle_model *m, const double *pos, const double *vel, double *mfield,
                               double mass, double timespan, int *orbitstaken, long long *steps_out)
{
  return or_orbit_core(m, pos, vel, mfield, NULL, NULL, NULL, NULL, mass, timespan, orbitstaken, 1, steps_out,
                       NULL, NULL);
}

/* ------------------------------------------------------------------------- */
/* quadtree stacks                                                           */
/* ------------------------------------------------------------------------- */

/* grid.c:1247-1265 */
void or_smooth_stack(double *data, int maxlevel)
{
  int l, i, j, n;
  for(l = maxlevel - 1; l >= 0; l--) {
    n = (1 << l);
    for(i = 0; i < n; i++)
      for(j = 0; j < n; j++)
        data[GALIC_STACKOFFSET(l, i, j)] =
            data[GALIC_STACKOFFSET(l + 1, 2 * i, 2 * j)] + data[GALIC_STACKOFFSET(l + 1, 2 * i + 1, 2 * j)] +
            data[GALIC_STACKOFFSET(l + 1, 2 * i, 2 * j + 1)] + data[GALIC_STACKOFFSET(l + 1, 2 * i + 1, 2 * j + 1)];
  }
}

/* grid.c:1086-1112 */
double or_calc_stack_difference(const oracle_model *m, const double *d1, const double *d2, int l, int i, int j,
                                int maxlevel, const double *ref1, const double *ref2, double thresh,
                                const double *dist, int flag)
{
  int o = GALIC_STACKOFFSET(l, i, j);
  if(l >= maxlevel || ref1[o] < thresh) {
    if(flag) {
      if(ref1[o] > 0 && ref2[o] > 0)
        return fabs(d1[o] / ref1[o] - d2[o] / ref2[o]) / or_dmax((d2[o] / ref2[o]), m->p.LowerDispLimit);
      else
        return 0;
    } else
      return fabs(d1[o] - d2[o]) / or_dmin(dist[o], m->p.R200 / 20.0);
  } else {
    double sum = 0;
    int ii, jj;
    for(ii = 2 * i; ii <= 2 * i + 1; ii++)
      for(jj = 2 * j; jj <= 2 * j + 1; jj++)
        sum += or_calc_stack_difference(m, d1, d2, l + 1, ii, jj, maxlevel, ref1, ref2, thresh, dist, flag);
    return sum;
  }
}

/* grid.c:1115-1151 */
double or_calc_stack_difference_mod(const oracle_model *m, const double *d1, const double *d2, int l, int i, int j,
                                    int maxlevel, const double *ref1, const double *ref2, double thresh,
                                    const double *dist, int flag)
{
  int o = GALIC_STACKOFFSET(l, i, j);
  if(l >= maxlevel || ref1[o] < thresh) {
    if(flag == 1) {
      if(ref1[o] > 0 && ref2[o] > 0)
        return fabs(d1[o] / ref1[o] - d2[o] / ref2[o]) / or_dmax((d2[o] / ref2[o]), m->p.LowerDispLimit);
      else
        return 0;
    } else if(flag == 2) {
      if(ref2[o] > 0)
        return fabs(d1[o] - d2[o]) * (ref1[o] / ref2[o]) / or_dmin(dist[o], m->p.R200 / 20.0);
      else
        return 0;
    } else
      return fabs(d1[o] - d2[o]) / or_dmin(dist[o], m->p.R200 / 20.0);
  } else {
    double sum = 0;
    int ii, jj;
    for(ii = 2 * i; ii <= 2 * i + 1; ii++)
      for(jj = 2 * j; jj <= 2 * j + 1; jj++)
        sum += or_calc_stack_difference_mod(m, d1, d2, l + 1, ii, jj, maxlevel, ref1, ref2, thresh, dist, flag);
    return sum;
  }
}

/* ------------------------------------------------------------------------- */
/* velocity decomposition, energy grid, merit                                */
/* ------------------------------------------------------------------------- */

/* disp_fields.c:200-372 */
void or_calc_disp_components(const oracle_model *m, const double *pos, int type, const double *vel, double *vr2,
                             double *vt2, double *vp2, double *vq2)
{
  int tovs = m->p.TypeOfVelocityStructure[type];
  if(tovs == 0 || tovs == 1 || tovs == 3) {
    double phi = atan2(pos[1], pos[0]);
    double theta = acos(pos[2] / sqrt(pos[0] * pos[0] + pos[1] * pos[1] + pos[2] * pos[2]));
    double er[3], ePhi[3], eTheta[3];
    er[0] = sin(theta) * cos(phi);
    er[1] = sin(theta) * sin(phi);
    er[2] = cos(theta);
    ePhi[0] = -sin(phi);
    ePhi[1] = cos(phi);
    ePhi[2] = 0;
    eTheta[0] = -cos(theta) * cos(phi);
    eTheta[1] = -cos(theta) * sin(phi);
    eTheta[2] = sin(theta);
    double vr = vel[0] * er[0] + vel[1] * er[1] + vel[2] * er[2];
    double vphi = vel[0] * ePhi[0] + vel[1] * ePhi[1] + vel[2] * ePhi[2];
    double vtheta = vel[0] * eTheta[0] + vel[1] * eTheta[1] + vel[2] * eTheta[2];
    double vstr = 0;
    if(tovs == 1 || tovs == 3)
      vstr = or_get_vstream(m, pos, type);
    *vr2 = vr * vr;
    *vt2 = vtheta * vtheta;
    *vp2 = (vphi - vstr) * (vphi - vstr);
    *vq2 = vphi * vphi;
  } else {
    double phi = atan2(pos[1], pos[0]);
    double eR[3], ePhi[3], eZ[3];
    eR[0] = cos(phi);
    eR[1] = sin(phi);
    eR[2] = 0;
    ePhi[0] = -sin(phi);
    ePhi[1] = cos(phi);
    ePhi[2] = 0;
    eZ[0] = 0;
    eZ[1] = 0;
    eZ[2] = 1;
    double vR = vel[0] * eR[0] + vel[1] * eR[1] + vel[2] * eR[2];
    double vphi = vel[0] * ePhi[0] + vel[1] * ePhi[1] + vel[2] * ePhi[2];
    double vZ = vel[0] * eZ[0] + vel[1] * eZ[1] + vel[2] * eZ[2];
    double vstr = or_get_vstream(m, pos, type);
    *vr2 = vR * vR;
    *vt2 = vZ * vZ;
    *vp2 = (vphi - vstr) * (vphi - vstr);
    *vq2 = vphi * vphi;
  }
}

/* main.c:877-917 */
static void or_add_to_energy_grid(const oracle_model *m, const double *pos, double mass, double vr2, double vt2,
                                  double vp2, double vq2, double *egyMass, double *er, double *et, double *ep,
                                  double *eq)
{
  int iR, iz, nb = m->EG_Nbin;
  double fR, fz;
  or_energygrid_get_cell(m, pos, &iR, &iz, &fR, &fz);
  vr2 *= mass;
  vt2 *= mass;
  vp2 *= mass;
  vq2 *= mass;
#define OFS(i, j) ((i) * nb + (j))
  if(egyMass) {
    egyMass[OFS(iz, iR)] += (1 - fR) * (1 - fz) * mass;
    egyMass[OFS(iz, iR + 1)] += (fR) * (1 - fz) * mass;
    egyMass[OFS(iz + 1, iR)] += (1 - fR) * (fz) * mass;
    egyMass[OFS(iz + 1, iR + 1)] += (fR) * (fz) * mass;
  }
  er[OFS(iz, iR)] += (1 - fR) * (1 - fz) * vr2; er[OFS(iz, iR + 1)] += (fR) * (1 - fz) * vr2;
  er[OFS(iz + 1, iR)] += (1 - fR) * (fz) * vr2; er[OFS(iz + 1, iR + 1)] += (fR) * (fz) * vr2;
  et[OFS(iz, iR)] += (1 - fR) * (1 - fz) * vt2; et[OFS(iz, iR + 1)] += (fR) * (1 - fz) * vt2;
  et[OFS(iz + 1, iR)] += (1 - fR) * (fz) * vt2; et[OFS(iz + 1, iR + 1)] += (fR) * (fz) * vt2;
  ep[OFS(iz, iR)] += (1 - fR) * (1 - fz) * vp2; ep[OFS(iz, iR + 1)] += (fR) * (1 - fz) * vp2;
  ep[OFS(iz + 1, iR)] += (1 - fR) * (fz) * vp2; ep[OFS(iz + 1, iR + 1)] += (fR) * (fz) * vp2;
  eq[OFS(iz, iR)] += (1 - fR) * (1 - fz) * vq2; eq[OFS(iz, iR + 1)] += (fR) * (1 - fz) * vq2;
  eq[OFS(iz + 1, iR)] += (1 - fR) * (fz) * vq2; eq[OFS(iz + 1, iR + 1)] += (fR) * (fz) * vq2;
#undef OFS
}

static double or_thresh_density(const oracle_model *m, int type)
{
  return m->p.MinParticlesPerBinForDensityMeasurement * m->p.MType[type] / m->p.NType[type];
}

static double or_thresh_disp(const oracle_model *m, int type)
{
  return m->p.MinParticlesPerBinForDispersionMeasurement * m->p.MType[type] / m->p.NType[type];
}

/* One of the four orbit-averaged S terms of eval_fit_mod, main.c:1052-1118. */
static double or_eval_sterm(oracle_model *m, int type, const double *global_s, const double *target,
                            const double *f_new, const double *f_old, double *tmp)
{
  int i, L = m->p.EG_MaxLevel;
  for(i = 0; i < m->EG_Ngrid; i++)
    tmp[GALIC_STACKOFFSET(m->p.DG_MaxLevel, 0, 0) + i] =
        global_s[GALIC_STACKOFFSET(L, 0, 0) + i] + f_new[i] - f_old[i];
  or_smooth_stack(tmp, L);
  return or_calc_stack_difference_mod(m, tmp, target, 0, 0, 0, L, m->DGs_MassTarget[type], target,
                                      or_thresh_density(m, type), m->DGs_Distance, 2) /
         m->p.MType[type];
}

/* main.c:1021-1217 */
double or_eval_fit_mod(oracle_model *m, int n, const double *vel, const double *mo_new, const double *mo_old,
                       const double *r_new, const double *r_old, const double *t_new, const double *t_old,
                       const double *q_new, const double *q_old, const double *p_new, const double *p_old)
{
  int i, type = m->type[n];
  int Ld = m->p.DG_MaxLevel, Le = m->p.EG_MaxLevel;
  double value, value_rs, value_ts, value_qs, value_ps, value_r, value_t, value_p, value_q;
  int nst = m->DG_Nstack > m->EG_Nstack ? m->DG_Nstack : m->EG_Nstack;
  double *tmp = malloc(sizeof(double) * nst);

  for(i = 0; i < m->DG_Ngrid; i++)
    tmp[GALIC_STACKOFFSET(Ld, 0, 0) + i] =
        m->DGs_MassResponse[type][GALIC_STACKOFFSET(Ld, 0, 0) + i] + mo_new[i] - mo_old[i];
  or_smooth_stack(tmp, Ld);
  value = or_calc_stack_difference(m, tmp, m->DGs_MassTarget[type], 0, 0, 0, Ld, m->DGs_MassTarget[type], NULL,
                                   or_thresh_density(m, type), m->DGs_Distance, 0);
  value *= 1.0 / m->p.MType[type];

  value_rs = or_eval_sterm(m, type, m->EGs_EgyResponseRS[type], m->EGs_EgyTarget_r[type], r_new, r_old, tmp);
  value_ts = or_eval_sterm(m, type, m->EGs_EgyResponseTS[type], m->EGs_EgyTarget_t[type], t_new, t_old, tmp);
  value_qs = or_eval_sterm(m, type, m->EGs_EgyResponseQS[type], m->EGs_EgyTarget_q[type], q_new, q_old, tmp);
  value_ps = or_eval_sterm(m, type, m->EGs_EgyResponsePS[type], m->EGs_EgyTarget_p[type], p_new, p_old, tmp);

  double mass = m->mass[n], vr2, vt2, vp2, vq2;
  or_calc_disp_components(m, &m->pos[3 * n], type, vel, &vr2, &vt2, &vp2, &vq2);
  double vr2_diff = (vr2 - m->vr2[n]);
  double vt2_diff = (vt2 - m->vt2[n]);
  double vp2_diff = (vp2 - m->vp2[n]);
  double vq2_diff = (vq2 - m->vq2[n]);

  double *ur = malloc(sizeof(double) * m->EG_Nstack);
  double *ut = malloc(sizeof(double) * m->EG_Nstack);
  double *up = malloc(sizeof(double) * m->EG_Nstack);
  double *uq = malloc(sizeof(double) * m->EG_Nstack);
  int off = GALIC_STACKOFFSET(Le, 0, 0);
  for(i = 0; i < m->EG_Ngrid; i++) {
    ur[off + i] = m->EGs_EgyResponse_r[type][off + i];
    ut[off + i] = m->EGs_EgyResponse_t[type][off + i];
    up[off + i] = m->EGs_EgyResponse_p[type][off + i];
    uq[off + i] = m->EGs_EgyResponse_q[type][off + i];
  }
  or_add_to_energy_grid(m, &m->pos[3 * n], mass, vr2_diff, vt2_diff, vp2_diff, vq2_diff, NULL, ur + off, ut + off,
                        up + off, uq + off);
  or_smooth_stack(ur, Le);
  or_smooth_stack(ut, Le);
  or_smooth_stack(up, Le);
  or_smooth_stack(uq, Le);
  double th = or_thresh_disp(m, type);
  value_r = or_calc_stack_difference(m, ur, m->EGs_EgyTarget_r[type], 0, 0, 0, Le, m->EGs_MassResponse[type],
                                     m->EGs_MassTarget[type], th, NULL, 1);
  value_t = or_calc_stack_difference(m, ut, m->EGs_EgyTarget_t[type], 0, 0, 0, Le, m->EGs_MassResponse[type],
                                     m->EGs_MassTarget[type], th, NULL, 1);
  value_p = or_calc_stack_difference(m, up, m->EGs_EgyTarget_p[type], 0, 0, 0, Le, m->EGs_MassResponse[type],
                                     m->EGs_MassTarget[type], th, NULL, 1);
  value_q = or_calc_stack_difference(m, uq, m->EGs_EgyTarget_q[type], 0, 0, 0, Le, m->EGs_MassResponse[type],
                                     m->EGs_MassTarget[type], th, NULL, 1);
  free(uq); free(up); free(ut); free(ur); free(tmp);

  int tovs = m->p.TypeOfVelocityStructure[type];
  if(tovs == 0 || type == 2)
    value_q = 0;

  return value + m->Srelfac[type] / m->Srelfac_count[type] * (value_r + value_t + value_p + value_q) +
         (m->p.fac_value_rs[type] * value_rs + m->p.fac_value_ts[type] * value_ts +
          m->p.fac_value_ps[type] * value_ps + m->p.fac_value_qs[type] * value_qs) / 4;
}

/* ------------------------------------------------------------------------- */
/* optimiser driver                                                          */
/* ------------------------------------------------------------------------- */

static int or_type_present(const oracle_model *m, int type) { return m->p.NType[type] > 0; }

/* main.c:251-362.  Types without particles are skipped: for them the reference
 * computes 0/0 thresholds and NaN merits that are never read. */
void or_calc_global_fit(oracle_model *m)
{
  int type, Ld = m->p.DG_MaxLevel, Le = m->p.EG_MaxLevel;
  for(type = 1; type <= 3; type++) {
    if(!or_type_present(m, type))
      continue;
    memcpy(&m->DGs_MassResponse[type][GALIC_STACKOFFSET(Ld, 0, 0)], m->DG_MassLoc[type], sizeof(double) * m->DG_Ngrid);
    or_smooth_stack(m->DGs_MassResponse[type], Ld);
    memcpy(&m->EGs_EgyResponseRS[type][GALIC_STACKOFFSET(Le, 0, 0)], m->EG_RLocS[type], sizeof(double) * m->EG_Ngrid);
    or_smooth_stack(m->EGs_EgyResponseRS[type], Le);
    memcpy(&m->EGs_EgyResponseTS[type][GALIC_STACKOFFSET(Le, 0, 0)], m->EG_TLocS[type], sizeof(double) * m->EG_Ngrid);
    or_smooth_stack(m->EGs_EgyResponseTS[type], Le);
    memcpy(&m->EGs_EgyResponseQS[type][GALIC_STACKOFFSET(Le, 0, 0)], m->EG_QLocS[type], sizeof(double) * m->EG_Ngrid);
    or_smooth_stack(m->EGs_EgyResponseQS[type], Le);
    memcpy(&m->EGs_EgyResponsePS[type][GALIC_STACKOFFSET(Le, 0, 0)], m->EG_PLocS[type], sizeof(double) * m->EG_Ngrid);
    or_smooth_stack(m->EGs_EgyResponsePS[type], Le);

    memcpy(&m->EGs_MassResponse[type][GALIC_STACKOFFSET(Le, 0, 0)], m->EG_MassLoc[type], sizeof(double) * m->EG_Ngrid);
    memcpy(&m->EGs_EgyResponse_r[type][GALIC_STACKOFFSET(Le, 0, 0)], m->EG_RLoc[type], sizeof(double) * m->EG_Ngrid);
    memcpy(&m->EGs_EgyResponse_t[type][GALIC_STACKOFFSET(Le, 0, 0)], m->EG_TLoc[type], sizeof(double) * m->EG_Ngrid);
    memcpy(&m->EGs_EgyResponse_p[type][GALIC_STACKOFFSET(Le, 0, 0)], m->EG_PLoc[type], sizeof(double) * m->EG_Ngrid);
    memcpy(&m->EGs_EgyResponse_q[type][GALIC_STACKOFFSET(Le, 0, 0)], m->EG_QLoc[type], sizeof(double) * m->EG_Ngrid);
    or_smooth_stack(m->EGs_MassResponse[type], Le);
    or_smooth_stack(m->EGs_EgyResponse_r[type], Le);
    or_smooth_stack(m->EGs_EgyResponse_t[type], Le);
    or_smooth_stack(m->EGs_EgyResponse_p[type], Le);
    or_smooth_stack(m->EGs_EgyResponse_q[type], Le);

    m->S[type] = or_calc_stack_difference(m, m->DGs_MassResponse[type], m->DGs_MassTarget[type], 0, 0, 0, Ld,
                                          m->DGs_MassTarget[type], NULL, or_thresh_density(m, type), m->DGs_Distance, 0);
    m->S[type] *= 1.0 / m->DGs_MassTarget[type][0];
    double th = or_thresh_disp(m, type);
    m->Sdisp_r[type] = or_calc_stack_difference(m, m->EGs_EgyResponse_r[type], m->EGs_EgyTarget_r[type], 0, 0, 0, Le,
                                                m->EGs_MassResponse[type], m->EGs_MassTarget[type], th, NULL, 1);
    m->Sdisp_t[type] = or_calc_stack_difference(m, m->EGs_EgyResponse_t[type], m->EGs_EgyTarget_t[type], 0, 0, 0, Le,
                                                m->EGs_MassResponse[type], m->EGs_MassTarget[type], th, NULL, 1);
    m->Sdisp_p[type] = or_calc_stack_difference(m, m->EGs_EgyResponse_p[type], m->EGs_EgyTarget_p[type], 0, 0, 0, Le,
                                                m->EGs_MassResponse[type], m->EGs_MassTarget[type], th, NULL, 1);
    m->Sdisp_q[type] = or_calc_stack_difference(m, m->EGs_EgyResponse_q[type], m->EGs_EgyTarget_q[type], 0, 0, 0, Le,
                                                m->EGs_MassResponse[type], m->EGs_MassTarget[type], th, NULL, 1);
    int tovs = m->p.TypeOfVelocityStructure[type];
    if(tovs == 0 || tovs == 1) {
      m->Srelfac_count[type] = 3.0;
      m->Sdisp_q[type] = 0;
    } else
      m->Srelfac_count[type] = 4.0;
    if(m->firstrun[type] == 0) {
      m->Srelfac[type] = m->S[type] / ((m->Sdisp_r[type] + m->Sdisp_t[type] + m->Sdisp_p[type] + m->Sdisp_q[type]) /
                                       m->Srelfac_count[type]);
      m->firstrun[type] = 1;
    }
    m->Sdisp_r[type] *= m->Srelfac[type];
    m->Sdisp_t[type] *= m->Srelfac[type];
    m->Sdisp_p[type] *= m->Srelfac[type];
    m->Sdisp_q[type] *= m->Srelfac[type];
  }
}

typedef struct {
  double *mo, *r, *t, *q, *p;
} or_fields;

static void or_fields_alloc(const oracle_model *m, or_fields *f)
{
  f->mo = malloc(sizeof(double) * m->DG_Ngrid);
  f->r = malloc(sizeof(double) * m->EG_Ngrid);
  f->t = malloc(sizeof(double) * m->EG_Ngrid);
  f->q = malloc(sizeof(double) * m->EG_Ngrid);
  f->p = malloc(sizeof(double) * m->EG_Ngrid);
}

static void or_fields_free(or_fields *f)
{
  free(f->mo); free(f->r); free(f->t); free(f->q); free(f->p);
}

/* main.c:151-246 */
void or_calc_all_response_fields(oracle_model *m, int nthreads)
{
  int i, n, type;
  for(type = 1; type <= 3; type++) {
    if(!or_type_present(m, type))
      continue;
    memset(m->DG_MassLoc[type], 0, sizeof(double) * m->DG_Ngrid);
    memset(m->EG_RLocS[type], 0, sizeof(double) * m->EG_Ngrid);
    memset(m->EG_TLocS[type], 0, sizeof(double) * m->EG_Ngrid);
    memset(m->EG_QLocS[type], 0, sizeof(double) * m->EG_Ngrid);
    memset(m->EG_PLocS[type], 0, sizeof(double) * m->EG_Ngrid);
    memset(m->EG_MassLoc[type], 0, sizeof(double) * m->EG_Ngrid);
    memset(m->EG_RLoc[type], 0, sizeof(double) * m->EG_Ngrid);
    memset(m->EG_TLoc[type], 0, sizeof(double) * m->EG_Ngrid);
    memset(m->EG_PLoc[type], 0, sizeof(double) * m->EG_Ngrid);
    memset(m->EG_QLoc[type], 0, sizeof(double) * m->EG_Ngrid);
  }
#ifdef _OPENMP
  if(nthreads > 0)
    omp_set_num_threads(nthreads);
#pragma omp parallel
#endif
  {
    or_fields f;
    or_fields_alloc(m, &f);
#ifdef _OPENMP
#pragma omp for ordered schedule(dynamic, 1)
#endif
    for(n = 0; n < m->N; n++) {
      int t = m->type[n];
      int orbits;
      or_orbit_core(m, &m->pos[3 * n], &m->vel[3 * n], f.mo, f.r, f.t, f.q, f.p, m->mass[n], m->tint[n], &orbits, t,
                    NULL, NULL, NULL);
#ifdef _OPENMP
#pragma omp ordered
#endif
      {
        m->orbits[n] = orbits;
        for(i = 0; i < m->DG_Ngrid; i++)
          m->DG_MassLoc[t][i] += f.mo[i];
        for(i = 0; i < m->EG_Ngrid; i++) {
          m->EG_RLocS[t][i] += f.r[i];
          m->EG_TLocS[t][i] += f.t[i];
          m->EG_QLocS[t][i] += f.q[i];
          m->EG_PLocS[t][i] += f.p[i];
        }
        or_calc_disp_components(m, &m->pos[3 * n], t, &m->vel[3 * n], &m->vr2[n], &m->vt2[n], &m->vp2[n], &m->vq2[n]);
        or_add_to_energy_grid(m, &m->pos[3 * n], m->mass[n], m->vr2[n], m->vt2[n], m->vp2[n], m->vq2[n],
                              m->EG_MassLoc[t], m->EG_RLoc[t], m->EG_TLoc[t], m->EG_PLoc[t], m->EG_QLoc[t]);
      }
    }
    or_fields_free(&f);
  }
  or_calc_global_fit(m);
}

/* main.c:844-874 */
static void or_init_updates(oracle_model *m)
{
  int type;
  for(type = 1; type <= 3; type++) {
    if(!or_type_present(m, type))
      continue;
    memset(m->DG_MassLoc_delta[type], 0, sizeof(double) * m->DG_Ngrid);
    memset(m->EG_RLocS_d[type], 0, sizeof(double) * m->EG_Ngrid);
    memset(m->EG_TLocS_d[type], 0, sizeof(double) * m->EG_Ngrid);
    memset(m->EG_QLocS_d[type], 0, sizeof(double) * m->EG_Ngrid);
    memset(m->EG_PLocS_d[type], 0, sizeof(double) * m->EG_Ngrid);
    memset(m->EG_RLoc_d[type], 0, sizeof(double) * m->EG_Ngrid);
    memset(m->EG_TLoc_d[type], 0, sizeof(double) * m->EG_Ngrid);
    memset(m->EG_PLoc_d[type], 0, sizeof(double) * m->EG_Ngrid);
    memset(m->EG_QLoc_d[type], 0, sizeof(double) * m->EG_Ngrid);
  }
}

/* main.c:808-841 */
static void or_commit_updates(oracle_model *m)
{
  int i, type;
  for(type = 1; type <= 3; type++) {
    if(!or_type_present(m, type))
      continue;
    for(i = 0; i < m->DG_Ngrid; i++)
      m->DG_MassLoc[type][i] += m->DG_MassLoc_delta[type][i];
    for(i = 0; i < m->EG_Ngrid; i++) {
      m->EG_RLocS[type][i] += m->EG_RLocS_d[type][i];
      m->EG_TLocS[type][i] += m->EG_TLocS_d[type][i];
      m->EG_QLocS[type][i] += m->EG_QLocS_d[type][i];
      m->EG_PLocS[type][i] += m->EG_PLocS_d[type][i];
    }
    for(i = 0; i < m->EG_Ngrid; i++) {
      m->EG_RLoc[type][i] += m->EG_RLoc_d[type][i];
      m->EG_TLoc[type][i] += m->EG_TLoc_d[type][i];
      m->EG_PLoc[type][i] += m->EG_PLoc_d[type][i];
      m->EG_QLoc[type][i] += m->EG_QLoc_d[type][i];
    }
  }
  or_calc_global_fit(m);
}

/* The body of optimize() (main.c:437-507) with the trial velocity supplied by the caller
 * instead of the GSL draw of main.c:407-435.  or_trial_compute integrates the two orbits
 * and evaluates the two merits; or_trial_apply is the write-back of main.c:460, 470-505. */
typedef struct {
  or_fields fo, fn;
  int orbits_old, orbits_try, accept;
  double F_base, F_try;
} or_trial;

static void or_trial_compute(oracle_model *m, int n, const double *vel_try, or_trial *tr)
{
  int type = m->type[n];
  or_fields *fo = &tr->fo, *fn = &tr->fn;
  or_orbit_core(m, &m->pos[3 * n], &m->vel[3 * n], fo->mo, fo->r, fo->t, fo->q, fo->p, m->mass[n], m->tint[n],
                &tr->orbits_old, type, NULL, NULL, NULL);
  or_orbit_core(m, &m->pos[3 * n], vel_try, fn->mo, fn->r, fn->t, fn->q, fn->p, m->mass[n], m->tint[n],
                &tr->orbits_try, type, NULL, NULL, NULL);
  tr->F_base = or_eval_fit_mod(m, n, &m->vel[3 * n], fo->mo, fo->mo, fo->r, fo->r, fo->t, fo->t, fo->q, fo->q,
                               fo->p, fo->p);
  tr->F_try = or_eval_fit_mod(m, n, vel_try, fn->mo, fo->mo, fn->r, fo->r, fn->t, fo->t, fn->q, fo->q, fn->p,
                              fo->p);
  tr->accept = (tr->F_try < tr->F_base);
}

static void or_trial_apply(oracle_model *m, int n, const double *vel_try, const or_trial *tr)
{
  int i, type = m->type[n];
  const or_fields *fo = &tr->fo, *fn = &tr->fn;
  m->orbits[n] = tr->orbits_old;
  if(!tr->accept)
    return;
  for(i = 0; i < 3; i++)
    m->vel[3 * n + i] = vel_try[i];
  m->orbits[n] = tr->orbits_try;
  for(i = 0; i < m->DG_Ngrid; i++)
    m->DG_MassLoc_delta[type][i] += fn->mo[i] - fo->mo[i];
  for(i = 0; i < m->EG_Ngrid; i++) {
    m->EG_RLocS_d[type][i] += fn->r[i] - fo->r[i];
    m->EG_TLocS_d[type][i] += fn->t[i] - fo->t[i];
    m->EG_QLocS_d[type][i] += fn->q[i] - fo->q[i];
    m->EG_PLocS_d[type][i] += fn->p[i] - fo->p[i];
  }
  double vr2, vt2, vp2, vq2;
  or_calc_disp_components(m, &m->pos[3 * n], type, vel_try, &vr2, &vt2, &vp2, &vq2);
  double vr2_diff = (vr2 - m->vr2[n]);
  double vt2_diff = (vt2 - m->vt2[n]);
  double vp2_diff = (vp2 - m->vp2[n]);
  double vq2_diff = (vq2 - m->vq2[n]);
  or_add_to_energy_grid(m, &m->pos[3 * n], m->mass[n], vr2_diff, vt2_diff, vp2_diff, vq2_diff, NULL,
                        m->EG_RLoc_d[type], m->EG_TLoc_d[type], m->EG_PLoc_d[type], m->EG_QLoc_d[type]);
  m->vr2[n] = vr2;
  m->vt2[n] = vt2;
  m->vp2[n] = vp2;
  m->vq2[n] = vq2;
}

/* main.c:93-102 for one rep: init_updates, optimize() per listed particle, commit_updates.
 * The trials of a rep only read global stacks that change at the commit, so they are
 * evaluated in parallel; the write-back runs in trial order, as in the serial reference. */
void or_rep(oracle_model *m, int ntrial, const int *idx, const double *vtry, int *accept, double *dF, int nthreads)
{
  int i;
  or_init_updates(m);
#ifdef _OPENMP
  if(nthreads > 0)
    omp_set_num_threads(nthreads);
#pragma omp parallel
#endif
  {
    or_trial tr;
    or_fields_alloc(m, &tr.fo);
    or_fields_alloc(m, &tr.fn);
#ifdef _OPENMP
#pragma omp for ordered schedule(dynamic, 1)
#endif
    for(i = 0; i < ntrial; i++) {
      or_trial_compute(m, idx[i], &vtry[3 * i], &tr);
#ifdef _OPENMP
#pragma omp ordered
#endif
      {
        accept[i] = tr.accept;
        if(dF)
          dF[i] = tr.F_try - tr.F_base;
        or_trial_apply(m, idx[i], &vtry[3 * i], &tr);
      }
    }
    or_fields_free(&tr.fo);
    or_fields_free(&tr.fn);
  }
  or_commit_updates(m);
}

void or_eval_trials(oracle_model *m, int ntrial, const int *idx, const double *vtry, double *fbase, double *ftry,
                    int nthreads)
{
  int i;
#ifdef _OPENMP
  if(nthreads > 0)
    omp_set_num_threads(nthreads);
#pragma omp parallel
#endif
  {
    or_trial tr;
    or_fields_alloc(m, &tr.fo);
    or_fields_alloc(m, &tr.fn);
#ifdef _OPENMP
#pragma omp for schedule(dynamic, 1)
#endif
    for(i = 0; i < ntrial; i++) {
      or_trial_compute(m, idx[i], &vtry[3 * i], &tr);
      fbase[i] = tr.F_base;
      ftry[i] = tr.F_try;
    }
    or_fields_free(&tr.fo);
    or_fields_free(&tr.fn);
  }
}

/* ------------------------------------------------------------------------- */
/* Barnes-Hut tree                                                           */
/* ------------------------------------------------------------------------- */

/* A node as the walk sees it (allvars.h:1068-1096): centre of mass and mass in
 * double, geometric centre and side length in float. */
typedef struct {
  double s[3], mass;
  int sibling, nextnode;
  float center[3], len;
  int suns[8];
} or_node;

typedef struct {
  int nsrc, maxnodes, numnodes;
  or_node *nodes;  /* index 0 is the root; node k has walk id nsrc + k */
  int *nextnode;   /* for particles */
  unsigned long long *intpos;
  double corner[3], center[3], len, inverse_len, bigfac;
  const double *pos, *mass;
} or_tree;

/* forcetree.h:33-38: the 52 mantissa bits of a double in [1,2) */
static unsigned long long or_double_to_int(double d)
{
  union { double d; unsigned long long ull; } u;
  u.d = d;
  return (u.ull & 0xFFFFFFFFFFFFFllu);
}

/* domain_box.c:20-65 over sources and targets together (the reference's P[] holds both). */
static void or_tree_extent(or_tree *t, int nsrc, const double *spos, int ntarget, const double *tpos)
{
  double xmin[3], xmax[3], len = 0;
  int i, j;
  for(j = 0; j < 3; j++) {
    xmin[j] = 1.0e30;
    xmax[j] = -1.0e30;
  }
  for(i = 0; i < nsrc + ntarget; i++) {
    const double *x = i < nsrc ? &spos[3 * i] : &tpos[3 * (i - nsrc)];
    for(j = 0; j < 3; j++) {
      if(xmin[j] > x[j]) xmin[j] = x[j];
      if(xmax[j] < x[j]) xmax[j] = x[j];
    }
  }
  for(j = 0; j < 3; j++)
    if(xmax[j] - xmin[j] > len)
      len = xmax[j] - xmin[j];
  len *= 1.00001;
  for(j = 0; j < 3; j++) {
    t->center[j] = 0.5 * (xmin[j] + xmax[j]);
    t->corner[j] = 0.5 * (xmin[j] + xmax[j]) - 0.5 * len;
  }
  t->len = len;
  t->inverse_len = 1.0 / len;
  t->bigfac = (len / (((long long)1) << 52));
}

/* forcetree.c:369-495 with a single top-level node (one task): th starts at the root. */
static int or_tree_insert(or_tree *t, int i)
{
  const int MAXP = t->nsrc;
  int j, parent = -1, th = MAXP; /* root */
  unsigned char subnode = 0, levels = 0;
  unsigned long long *ip = &t->intpos[3 * i];
  unsigned long long xxb = ip[0], yyb = ip[1], zzb = ip[2];
  unsigned long long mask = ((unsigned long long)1) << ((52 - 1) - levels);
  unsigned char shiftx = (52 - 1) - levels, shifty = (52 - 2) - levels, shiftz = (52 - 3) - levels;
  signed long long centermask = (signed long long)(0xFFF0000000000000llu);
  centermask >>= levels;
  while(1) {
    if(th >= MAXP) { /* internal node */
      or_node *nd = &t->nodes[th - MAXP];
      subnode = (((unsigned char)((xxb & mask) >> (shiftx--))) | ((unsigned char)((yyb & mask) >> (shifty--))) |
                 ((unsigned char)((zzb & mask) >> (shiftz--))));
      centermask >>= 1;
      mask >>= 1;
      levels++;
      if(levels > 30) { /* MAX_TREE_LEVEL (forcetree.h:16), forcetree.c:400-416 */
        for(j = 0; j < 8; j++) {
          if(nd->suns[subnode] < 0)
            break;
          subnode++;
          if(subnode >= 8)
            subnode = 7;
        }
      }
      int nn = nd->suns[subnode];
      if(nn >= 0) {
        parent = th;
        th = nn;
      } else {
        nd->suns[subnode] = i;
        break;
      }
    } else { /* leaf with one particle: split */
      if(t->numnodes >= t->maxnodes)
        return -1;
      int newn = MAXP + t->numnodes;
      t->nodes[parent - MAXP].suns[subnode] = newn;
      or_node *nf = &t->nodes[t->numnodes];
      double len = ((double)(mask << 1)) * t->bigfac;
      double cx = ((double)((xxb & centermask) | mask)) * t->bigfac + t->corner[0];
      double cy = ((double)((yyb & centermask) | mask)) * t->bigfac + t->corner[1];
      double cz = ((double)((zzb & centermask) | mask)) * t->bigfac + t->corner[2];
      nf->len = len;
      nf->center[0] = cx;
      nf->center[1] = cy;
      nf->center[2] = cz;
      for(j = 0; j < 8; j++)
        nf->suns[j] = -1;
      unsigned long long *ipp = &t->intpos[3 * th];
      subnode = (((unsigned char)((ipp[0] & mask) >> shiftx)) | ((unsigned char)((ipp[1] & mask) >> shifty)) |
                 ((unsigned char)((ipp[2] & mask) >> shiftz)));
      nf->suns[subnode] = th;
      th = newn;
      t->numnodes++;
    }
  }
  return 0;
}

/* forcetree.c:591-722: monopole moments and the nextnode/sibling threading. */
static void or_tree_update(or_tree *t, int no, int sib, int *last)
{
  const int MAXP = t->nsrc;
  int j, jj, p, pp, nextsib, suns[8];
  double s[3], mass;
  if(no >= MAXP) {
    or_node *nd = &t->nodes[no - MAXP];
    for(j = 0; j < 8; j++)
      suns[j] = nd->suns[j];
    if(*last >= 0) {
      if(*last >= MAXP)
        t->nodes[*last - MAXP].nextnode = no;
      else
        t->nextnode[*last] = no;
    }
    *last = no;
    mass = 0;
    s[0] = s[1] = s[2] = 0;
    for(j = 0; j < 8; j++) {
      if((p = suns[j]) >= 0) {
        for(jj = j + 1; jj < 8; jj++)
          if((pp = suns[jj]) >= 0)
            break;
        nextsib = (jj < 8) ? pp : sib;
        or_tree_update(t, p, nextsib, last);
        if(p < MAXP) {
          const double *pos = &t->pos[3 * p];
          mass += t->mass[p];
          s[0] += t->mass[p] * pos[0];
          s[1] += t->mass[p] * pos[1];
          s[2] += t->mass[p] * pos[2];
        } else {
          or_node *c = &t->nodes[p - MAXP];
          mass += c->mass;
          s[0] += c->mass * c->s[0];
          s[1] += c->mass * c->s[1];
          s[2] += c->mass * c->s[2];
        }
      }
    }
    if(mass) {
      s[0] /= mass;
      s[1] /= mass;
      s[2] /= mass;
    } else {
      s[0] = nd->center[0];
      s[1] = nd->center[1];
      s[2] = nd->center[2];
    }
    nd->mass = mass;
    nd->s[0] = s[0];
    nd->s[1] = s[1];
    nd->s[2] = s[2];
    nd->sibling = sib;
  } else {
    if(*last >= 0) {
      if(*last >= MAXP)
        t->nodes[*last - MAXP].nextnode = no;
      else
        t->nextnode[*last] = no;
    }
    *last = no;
  }
}

static int or_tree_build(or_tree *t, int nsrc, const double *spos, const double *smass, int ntarget,
                         const double *tpos)
{
  int i, j;
  memset(t, 0, sizeof(*t));
  t->nsrc = nsrc;
  t->pos = spos;
  t->mass = smass;
  or_tree_extent(t, nsrc, spos, ntarget, tpos);
  t->maxnodes = 2 * nsrc + 64;
  t->nodes = calloc((size_t)t->maxnodes, sizeof(or_node));
  t->nextnode = malloc(sizeof(int) * (size_t)(nsrc + 1));
  t->intpos = malloc(sizeof(unsigned long long) * 3 * (size_t)(nsrc + 1));
  /* root, forcetree.c:150-162 */
  for(j = 0; j < 8; j++)
    t->nodes[0].suns[j] = -1;
  t->nodes[0].len = t->len;
  for(j = 0; j < 3; j++)
    t->nodes[0].center[j] = t->center[j];
  t->numnodes = 1;
  /* integer positions, forcetree.c:197-210 */
  for(i = 0; i < nsrc; i++)
    for(j = 0; j < 3; j++)
      t->intpos[3 * i + j] = or_double_to_int(((spos[3 * i + j] - t->corner[j]) * t->inverse_len) + 1.0);
  for(i = 0; i < nsrc; i++)
    if(or_tree_insert(t, i) < 0)
      return -1;
  int last = -1;
  or_tree_update(t, nsrc, -1, &last);
  if(last >= nsrc)
    t->nodes[last - nsrc].nextnode = -1;
  else if(last >= 0)
    t->nextnode[last] = -1;
  return 0;
}

static void or_tree_free(or_tree *t)
{
  free(t->nodes);
  free(t->nextnode);
  free(t->intpos);
}

/* Gadget-2 cubic spline softened monopole, forcetree_walk.c:158-191. */
static void or_interact(double mass, double r2, double hmax, double *fac, double *wp)
{
  double r = sqrt(r2);
  if(r >= hmax) {
    *fac = mass / (r2 * r);
    *wp = -mass / r;
  } else {
    double h_inv = 1.0 / hmax;
    double h3_inv = h_inv * h_inv * h_inv;
    double u = r * h_inv;
    if(u < 0.5) {
      *fac = mass * h3_inv * (10.666666666667 + u * u * (32.0 * u - 38.4));
      *wp = mass * h_inv * (-2.8 + u * u * (5.333333333333 + u * u * (6.4 * u - 9.6)));
    } else {
      *fac = mass * h3_inv * (21.333333333333 - 48.0 * u + 38.4 * u * u - 10.666666666667 * u * u * u -
                              0.066666666667 / (u * u * u));
      *wp = mass * h_inv * (-3.2 + 0.066666666667 / u + u * u * (10.666666666667 + u * (-16.0 + u * (9.6 - 2.133333333333 * u))));
    }
  }
}

/* forcetree_walk.c:12-222, mode 0, one task. Returns the interaction count. */
static int or_tree_walk(const or_tree *t, const galic_params *p, const double *tp, double *acc, double *pot)
{
  const int MAXP = t->nsrc;
  double acc_x = 0, acc_y = 0, acc_z = 0, potv = 0, r2, dx, dy, dz, mass, fac, wp;
  double hmax = p->ForceSoftening;
  int no = MAXP, nint = 0;
  while(no >= 0) {
    if(no < MAXP) {
      dx = t->pos[3 * no + 0] - tp[0];
      dy = t->pos[3 * no + 1] - tp[1];
      dz = t->pos[3 * no + 2] - tp[2];
      r2 = dx * dx + dy * dy + dz * dz;
      mass = t->mass[no];
      no = t->nextnode[no];
    } else {
      const or_node *nop = &t->nodes[no - MAXP];
      mass = nop->mass;
      dx = nop->s[0] - tp[0];
      dy = nop->s[1] - tp[1];
      dz = nop->s[2] - tp[2];
      r2 = dx * dx + dy * dy + dz * dz;
      float len2 = nop->len * nop->len;
      if(len2 > r2 * p->ErrTolTheta * p->ErrTolTheta) {
        no = nop->nextnode;
        continue;
      }
      no = nop->sibling;
    }
    if(mass) {
      or_interact(mass, r2, hmax, &fac, &wp);
      acc_x += dx * fac;
      acc_y += dy * fac;
      acc_z += dz * fac;
      potv += wp;
      nint++;
    }
  }
  /* gravtree.c:455-463 */
  acc[0] = acc_x * p->G;
  acc[1] = acc_y * p->G;
  acc[2] = acc_z * p->G;
  *pot = potv * p->G;
  return nint;
}

long long or_tree_forces(const galic_params *p, int nsrc, const double *spos, const double *smass, int ntarget,
                         const double *tpos, double *acc, double *pot, int nthreads)
{
  or_tree t;
  long long nint = 0;
  int i;
  if(or_tree_build(&t, nsrc, spos, smass, ntarget, tpos) < 0) {
    or_tree_free(&t);
    return -1;
  }
#ifdef _OPENMP
  if(nthreads > 0)
    omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 64) reduction(+ : nint)
#endif
  for(i = 0; i < ntarget; i++)
    nint += or_tree_walk(&t, p, &tpos[3 * i], &acc[3 * i], &pot[i]);
  or_tree_free(&t);
  return nint;
}

void or_direct_forces(const galic_params *p, int nsrc, const double *spos, const double *smass, int ntarget,
                      const double *tpos, double *acc, double *pot, int nthreads)
{
  int i;
  (void)nthreads;
#ifdef _OPENMP
  if(nthreads > 0)
    omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 16)
#endif
  for(i = 0; i < ntarget; i++) {
    double ax = 0, ay = 0, az = 0, pv = 0, fac, wp;
    int k;
    for(k = 0; k < nsrc; k++) {
      double dx = spos[3 * k] - tpos[3 * i];
      double dy = spos[3 * k + 1] - tpos[3 * i + 1];
      double dz = spos[3 * k + 2] - tpos[3 * i + 2];
      if(smass[k] == 0)
        continue;
      or_interact(smass[k], dx * dx + dy * dy + dz * dz, p->ForceSoftening, &fac, &wp);
      ax += dx * fac;
      ay += dy * fac;
      az += dz * fac;
      pv += wp;
    }
    acc[3 * i] = ax * p->G;
    acc[3 * i + 1] = ay * p->G;
    acc[3 * i + 2] = az * p->G;
    pot[i] = pv * p->G;
  }
}

/* grid.c:680-750 */
long long or_table_fill(const galic_params *p, int nsrc, const double *spos, const double *smass, double *potg,
                        double *dpotdr, double *dpotdz, int nthreads)
{
  int nb = p->FG_Nbin, s, k, j, i, ns = GALIC_FG_SECTIONS * nb * nb;
  double *tpos = malloc(sizeof(double) * 3 * (size_t)ns);
  double *acc = malloc(sizeof(double) * 3 * (size_t)ns);
  double *pot = malloc(sizeof(double) * (size_t)ns);
  int n = 0;
  for(s = 0; s < GALIC_FG_SECTIONS; s++)
    for(k = 0; k < nb; k++)
      for(j = 0; j < nb; j++) {
        double phi = 2 * M_PI / GALIC_FG_SECTIONS * s;
        double r1 = p->FG_Rmin * (pow(p->FG_Fac, j) - 1.0);
        double z1 = p->FG_Rmin * (pow(p->FG_Fac, k) - 1.0);
        tpos[3 * n + 0] = r1 * cos(phi);
        tpos[3 * n + 1] = r1 * sin(phi);
        tpos[3 * n + 2] = z1;
        n++;
      }
  long long nint = or_tree_forces(p, nsrc, spos, smass, ns, tpos, acc, pot, nthreads);
  memset(dpotdr, 0, sizeof(double) * nb * nb);
  memset(dpotdz, 0, sizeof(double) * nb * nb);
  memset(potg, 0, sizeof(double) * nb * nb);
  for(n = 0; n < ns; n++) {
    s = n / (nb * nb);
    i = n - s * (nb * nb);
    k = i / nb;
    j = i - k * nb;
    i = k * nb + j;
    double phi = 2 * M_PI / GALIC_FG_SECTIONS * s;
    dpotdr[i] += acc[3 * n] * cos(phi) + acc[3 * n + 1] * sin(phi);
    dpotdz[i] += acc[3 * n + 2];
    potg[i] += pot[n];
  }
  for(i = 0; i < nb * nb; i++) {
    dpotdr[i] /= GALIC_FG_SECTIONS;
    dpotdz[i] /= GALIC_FG_SECTIONS;
    potg[i] /= GALIC_FG_SECTIONS;
  }
  free(tpos);
  free(acc);
  free(pot);
  return nint;
}

/* ------------------------------------------------------------------------- */
/* sampled targets                                                           */
/* ------------------------------------------------------------------------- */

/* One round of halo_get_fresh_coordinate (halo.c:17-40), disk_get_fresh_coordinate
 * (disk.c:17-51) or bulge_get_fresh_coordinate (bulge.c:17-38) with the round's
 * uniforms u[0..2] taken in the order the reference draws them.  Returns 1 if the
 * reference would keep the point, 0 if it would draw again.  The halo and the bulge
 * are spherical here (the stretch divides z as in the reference). */
int or_sample_point(const galic_params *p, int type, double stretch, const double *u, double *pos)
{
  if(type == 2) {
    double q, f, f_, R, Rold, phi;
    pos[2] = p->Disk_Z0 / 2 * log(u[0] / (1 - u[0]));
    q = u[1];
    R = 1.0;
    do {
      f = (1 + R) * exp(-R) + q - 1;
      f_ = -R * exp(-R);
      Rold = R;
      R = R - f / f_;
    } while(fabs(R - Rold) / R > 1e-7);
    R *= p->Disk_H;
    phi = u[2] * M_PI * 2;
    pos[0] = R * cos(phi);
    pos[1] = R * sin(phi);
    return !(pos[0] * pos[0] + pos[1] * pos[1] + pos[2] * pos[2] > p->Rmax * p->Rmax);
  } else {
    double a = type == 1 ? p->Halo_A : p->Bulge_A;
    double q = u[0], r, phi, theta;
    r = q > 0 ? a * (q + sqrt(q)) / (1 - q) : 0;
    if(type == 3 && r > p->Rmax)
      return 0;
    phi = u[1] * M_PI * 2;
    theta = acos(u[2] * 2 - 1);
    pos[0] = r * sin(theta) * cos(phi);
    pos[1] = r * sin(theta) * sin(phi);
    pos[2] = r * cos(theta) / stretch;
    if(type == 3)
      return 1;
    r = sqrt(pos[0] * pos[0] + pos[1] * pos[1] + pos[2] * pos[2]);
    return !(r > p->Rmax);
  }
}

static double *copy_or_null(const double *src, size_t n)
{
  double *d;
  if(!src)
    return NULL;
  d = (double *)malloc(n * sizeof(double));
  memcpy(d, src, n * sizeof(double));
  return d;
}

void or_set_disp_tables(oracle_model *m, int type, double beta, const double *disp_r, const double *disp_z,
                        const double *disp_phi, const double *tilt_r, const double *tilt_theta)
{
  size_t nb = (size_t)m->p.FG_Nbin, n2 = nb * nb;
  free(m->Disp_r[type]); free(m->DispZ[type]); free(m->DispPhi[type]); free(m->TiltR[type]); free(m->TiltTheta[type]);
  m->Disp_r[type] = copy_or_null(disp_r, nb);
  m->DispZ[type] = copy_or_null(disp_z, n2);
  m->DispPhi[type] = copy_or_null(disp_phi, n2);
  m->TiltR[type] = copy_or_null(tilt_r, n2);
  m->TiltTheta[type] = copy_or_null(tilt_theta, n2);
  m->beta[type] = beta;
}

/* bilinear value of a force-grid table, as get_z_disp_cylindrical (disp_fields.c:134-147) */
static double fg_table(const oracle_model *m, const double *T, const double *pos)
{
  int iz, iR, nb = m->p.FG_Nbin;
  double fR, fz;
  or_forcegrid_get_cell(m, pos, &iR, &iz, &fR, &fz);
  return T[iz * nb + iR] * (1 - fR) * (1 - fz) + T[(iz + 1) * nb + iR] * (1 - fR) * (fz) +
         T[iz * nb + (iR + 1)] * (fR) * (1 - fz) + T[(iz + 1) * nb + (iR + 1)] * (fR) * (fz);
}

/* disp_fields.c:375-398 */
static double radial_disp_spherical(const oracle_model *m, const double *pos, int type)
{
  double r = sqrt(pos[0] * pos[0] + pos[1] * pos[1] + pos[2] * pos[2]);
  double br = (log(r / m->p.FG_Rmin + 1.0) / log(m->p.FG_Fac));
  int binr;
  double fR;
  if(br < 0)
    br = 0;
  binr = (int)br;
  fR = br - binr;
  if(binr >= m->p.FG_Nbin - 1) {
    binr = m->p.FG_Nbin - 2;
    fR = 1;
  }
  return m->Disp_r[type][binr] * (1 - fR) + m->Disp_r[type][binr + 1] * (fR);
}

/* disp_fields.c:26-64 */
static double beta_of_type(const oracle_model *m, const double *pos, int type)
{
  double beta = 0;
  if(type == 1 || type == 3) {
    beta = m->beta[type];
    if(beta >= 1) {
      double a = type == 1 ? m->p.Halo_A : m->p.Bulge_A;
      double r = sqrt(pos[0] * pos[0] + pos[1] * pos[1] + pos[2] * pos[2]);
      double dlogrhodr = -1.0 - 3.0 * r / (r + a);
      beta = -0.15 - 0.20 * dlogrhodr;
    }
  }
  return beta;
}

/* disp_fields.c:66-115 */
void or_get_disp_rtp(const oracle_model *m, const double *pos, int type, double *dr, double *dt, double *dp,
                     double *dq)
{
  int tovs = m->p.TypeOfVelocityStructure[type];
  if(tovs == 0) {
    *dr = radial_disp_spherical(m, pos, type);
    *dt = *dr;
    *dp = *dr;
    *dq = *dr;
  } else if(tovs == 1) {
    *dr = radial_disp_spherical(m, pos, type);
    *dt = (1 - beta_of_type(m, pos, type)) * (*dr);
    *dp = *dt;
    *dq = *dt;
  } else {
    double vstr;
    if(tovs == 2) {
      *dt = fg_table(m, m->DispZ[type], pos);
      *dr = *dt;
    } else {
      *dr = fg_table(m, m->TiltR[type], pos);
      *dt = fg_table(m, m->TiltTheta[type], pos);
    }
    *dp = fg_table(m, m->DispPhi[type], pos);
    vstr = or_get_vstream(m, pos, type);
    *dq = (*dp) + vstr * vstr;
  }
}

/* The binning half of densitygrid_sample_targetresponse (grid.c:536-617) and
 * calc_energy_grid_mass_maps (grid.c:941-1076) for given points: points with
 * sqrt(x^2 + z^2) < Rmax are binned and counted, the maps are scaled by
 * MType/count and smoothed.  dg [DG_Nstack] and eg [5][EG_Nstack] may be NULL.
 * Returns the count. */
long long or_bin_targets(const oracle_model *m, int type, long long n, const double *pos, double *dg, double *eg)
{
  long long i, count = 0;
  int k, nb = m->DG_Nbin, ne = m->EG_Nbin;
  double *mfield = (double *)calloc((size_t)m->DG_Ngrid, sizeof(double));
  double *efield[5];
  for(k = 0; k < 5; k++)
    efield[k] = (double *)calloc((size_t)m->EG_Ngrid, sizeof(double));
  for(i = 0; i < n; i++) {
    const double *x = pos + 3 * i;
    double r = sqrt(x[0] * x[0] + x[2] * x[2]);
    int iR, iz;
    double fR, fz;
    if(!(r < m->p.Rmax))
      continue;
    if(dg) {
      or_densitygrid_get_cell(m, x, &iR, &iz, &fR, &fz);
      mfield[iz * nb + iR] += (1 - fR) * (1 - fz);
      mfield[iz * nb + (iR + 1)] += (fR) * (1 - fz);
      mfield[(iz + 1) * nb + iR] += (1 - fR) * (fz);
      mfield[(iz + 1) * nb + (iR + 1)] += (fR) * (fz);
    }
    if(eg) {
      double d[4];
      or_energygrid_get_cell(m, x, &iR, &iz, &fR, &fz);
      or_get_disp_rtp(m, x, type, &d[0], &d[1], &d[2], &d[3]);
      efield[0][iz * ne + iR] += (1 - fR) * (1 - fz);
      efield[0][iz * ne + (iR + 1)] += (fR) * (1 - fz);
      efield[0][(iz + 1) * ne + iR] += (1 - fR) * (fz);
      efield[0][(iz + 1) * ne + (iR + 1)] += (fR) * (fz);
      for(k = 0; k < 4; k++) {
        efield[1 + k][iz * ne + iR] += (1 - fR) * (1 - fz) * d[k];
        efield[1 + k][iz * ne + (iR + 1)] += (fR) * (1 - fz) * d[k];
        efield[1 + k][(iz + 1) * ne + iR] += (1 - fR) * (fz) * d[k];
        efield[1 + k][(iz + 1) * ne + (iR + 1)] += (fR) * (fz) * d[k];
      }
    }
    count++;
  }
  if(dg) {
    int off = GALIC_STACKOFFSET(m->p.DG_MaxLevel, 0, 0);
    memset(dg, 0, sizeof(double) * (size_t)m->DG_Nstack);
    for(i = 0; i < m->DG_Ngrid; i++)
      dg[off + i] = mfield[i] * (m->p.MType[type] / (double)count);
    or_smooth_stack(dg, m->p.DG_MaxLevel);
  }
  if(eg) {
    int off = GALIC_STACKOFFSET(m->p.EG_MaxLevel, 0, 0);
    for(k = 0; k < 5; k++) {
      double *st = eg + (size_t)k * m->EG_Nstack;
      memset(st, 0, sizeof(double) * (size_t)m->EG_Nstack);
      for(i = 0; i < m->EG_Ngrid; i++)
        st[off + i] = efield[k][i] * (m->p.MType[type] / (double)count);
      or_smooth_stack(st, m->p.EG_MaxLevel);
    }
  }
  free(mfield);
  for(k = 0; k < 5; k++)
    free(efield[k]);
  return count;
}

