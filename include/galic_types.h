/*
 * galic_types.h -- plain-C description of a GALIC model as the equilibrium
 * hot path reads it.
 *
 * The reference keeps this state in globals: the `All` parameter struct
 * (src/allvars.h:639-836), the force-grid globals FG_* (src/allvars.h:308-327),
 * and the density/energy grid globals DG_*, EG_* (src/allvars.h:328-407).
 * Here the scalars the hot path reads are gathered into one POD struct, so a
 * caller fills it once and hands it over the C ABI.  Component types follow
 * the reference numbering: 1 = halo, 2 = disk, 3 = bulge (index 0 unused).
 */
#ifndef GALIC_TYPES_H
#define GALIC_TYPES_H

#ifdef __cplusplus
extern "C" {
#endif

#define GALIC_NTYPES 4          /* types 1..3 used, index 0 unused */
#define GALIC_FG_SECTIONS 8     /* azimuthal sections of the table fill (allvars.h:123) */

typedef struct galic_params {
  /* structural scalars (structure.c:54-143, init.c:40-63) */
  double G;                     /* gravitational constant, internal units */
  double V200, R200, M200;
  double Halo_Mass, Halo_A;
  double Disk_Mass, Disk_H, Disk_Z0;
  double Bulge_Mass, Bulge_A;
  double Rmax;                  /* outer grid edge (grid.c:336) */
  double MType[GALIC_NTYPES];   /* mass per component (structure.c:106-108) */
  double NType[GALIC_NTYPES];   /* particle count per component (structure.c:110-112) */

  /* which components are represented by the tabulated field (grid.c:63,124) */
  int SampleForceNhalo, SampleForceNdisk, SampleForceNbulge;

  /* velocity structure per type: 0 isotropic, 1 beta, 2 f(E,Lz), 3 f(E,Lz,I3) */
  int TypeOfVelocityStructure[GALIC_NTYPES];

  /* orbit integration (orbit_response.c:16-26) */
  double TimeStepFactorOrbit;
  double TimeStepFactorCellCross;
  double MaxVelInUnitsVesc;
  int    MaxStepsPerOrbit;      /* step cap; the reference aborts above 1e8 (orbit_response.c:449) */

  /* merit (main.c:1021-1217, 251-362) */
  double LowerDispLimit;
  double MinParticlesPerBinForDensityMeasurement;
  double MinParticlesPerBinForDispersionMeasurement;
  double fac_value_rs[GALIC_NTYPES], fac_value_ts[GALIC_NTYPES];
  double fac_value_qs[GALIC_NTYPES], fac_value_ps[GALIC_NTYPES];

  /* force grid: FG_Nbin x FG_Nbin, log spaced, layout [iz*FG_Nbin + iR] */
  int    FG_Nbin;
  double FG_Rmin, FG_Fac;

  /* density and energy response grids: 2^L x 2^L, cell centred */
  int    DG_MaxLevel;
  double DG_Rmin, DG_Fac;
  int    EG_MaxLevel;
  double EG_Rmin, EG_Fac;

  /* tree (forcetree_walk.c:12-222, init.c:78-84) */
  double ErrTolTheta;
  double ForceSoftening;        /* = 2.8 * Softening */
} galic_params;

/* Number of elements of a quadtree "stack" of max level L: (4^(L+1)-1)/3 (allvars.h:160) */
#define GALIC_NSTACK(L) ((((1 << (2 + 2 * (L))) - 1)) / 3)
/* Offset of node (l,i,j) in a stack: i is the z row, j the R column (allvars.h:160) */
#define GALIC_STACKOFFSET(l, i, j) ((((1 << (2 * (l))) - 1) / 3) + ((i) * (1 << (l))) + (j))

#ifdef __cplusplus
}
#endif

#endif /* GALIC_TYPES_H */
