// Launch arguments shared by the host entry points and the kernels.
#pragma once

#include "hydro_device.hpp"

namespace ramses {

// One batched unsplit launch (see unsplit.hip).
struct UnsplitArgs {
  const double* uin;
  const double* grav;  // may be null: gravity zero
  double* flux;
  double* tmp;
  double dx[3];
  double dt;
  const double* d_dt;  // if not null, dt is read from device memory
  long ngroups;
  int ngrid_last;
  int nv;
  HydroParams hp;
};

}  // namespace ramses
