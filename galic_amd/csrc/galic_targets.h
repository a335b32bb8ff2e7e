// galic_targets.h -- sampled target maps on the device.
//
// The reference builds its density target and its dispersion targets from
// fresh random points of each component's density profile
// (densitygrid_sample_targetresponse, grid.c:536-617; calc_energy_grid_mass_maps,
// grid.c:941-1076), 1e8 points per type in its parameter files.  Here every
// sample id draws its own Philox numbers, so the points are independent of the
// launch and equal to the host module's draws, and the maps are summed in 64-bit
// fixed point, so they are bit-identical run to run.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "galic_device.h"

namespace galic {

struct TargetArgs {
  DevModel M;               // grid geometry and the streaming-velocity tables
  int type, tovs;           // component 1..3 and its velocity structure 0..3
  long long n;              // sample ids 0..n-1
  uint32_t seed, stream;    // Philox key
  double Rmax;              // outer grid edge
  double Disk_H, Disk_Z0;   // exponential scale length and sech^2 height of the disk
  double stretch;           // HaloStretch / BulgeStretch (z is divided by it)
  double beta;              // HaloBetaParameter / BulgeBetaParameter (>= 1: slope dependent)
  const double *disp_r;     // [FG_Nbin] FG_Disp_r (structures 0, 1)
  const double *disp_z;     // [FG_Nbin^2] FG_DispZ (structure 2)
  const double *disp_phi;   // [FG_Nbin^2] FG_DispPhi (structures 2, 3)
  const double *tilt_r;     // [FG_Nbin^2] FG_tilted_vR2_prime (structure 3)
  const double *tilt_theta; // [FG_Nbin^2] FG_tilted_vz2_prime (structure 3)
  long long *dg;            // [DG_Ngrid] fixed-point mass map, or null
  long long *eg;            // [5][EG_Ngrid] fixed-point maps of mass and the r, t, p, q dispersions, or null
  double qdg;               // fixed-point units per unit weight of dg
  double qeg[5];            // and of each eg map
  double *pos_out;          // [n][3] the sample points (0 for ids without an accepted draw), or null
  unsigned long long *counts;  // [0] points binned, [1] ids without an accepted draw
  int count;                // add to counts (one launch of a sampling does)
  int f0, f1;               // the eg maps this launch bins: f0 .. f1-1
  int lds;                  // sum the launch's maps per workgroup in LDS first (dynamic shared memory
                            // of 8 bytes per cell of each map: dg first, then eg maps f0 .. f1-1)
};

// Grid-stride launch over the sample ids on stream s.
int launch_sample_targets(const TargetArgs &A, int grid, size_t shmem, hipStream_t s);

}  // namespace galic
