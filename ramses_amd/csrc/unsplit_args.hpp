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
  // Oct-tree sweeps only: per-batch refinement flags of the stencil cells
  // (nv x 216 bytes per batch, lane fastest); a face next to a flagged cell
  // gets a zero flux.  Null for the batched boundary call.
  const unsigned char* ok = nullptr;
  // Oct-tree sweeps only: at or below this many octs the level's unsplit
  // runs one face per thread instead of one row of faces per thread.
  int face_octs = 0;
};

}  // namespace ramses
