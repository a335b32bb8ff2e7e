// galic_handle.h -- the opaque galic_ctx of the C ABI: device state plus the
// host copies the driver sequences and the file writers read.
#pragma once
#include <vector>

#include "../../include/galic_hip.h"
#include "galic_ctx.h"

struct galic_ctx {
  galic::Ctx c;
  // host copies of the targets (DGs_MassTarget, EGs_MassTarget, EGs_EgyTarget_r, _t, _p, _q)
  // and of the geometry, used to build the cuts
  std::vector<double> targets[GALIC_NTYPES][6];
  std::vector<double> distance, cellsize;
  std::vector<double> weights;  // [type][W_COUNT][Nstack]
  std::vector<int> cutB_of;     // [type][Ngrid]
  bool initial_done = false;
  bool fp32_initial_done = false;
};
