// Launch arguments of the oct-tree (AMR) level kernels, shared by the host
// entry points and amr.hip.
#pragma once

#include "hydro_device.hpp"

namespace ramses {

struct AmrTreeView {
  const int* son;      // son[icell], icell in 1..ncell (p[0] unused)
  const int* father;   // father[igrid]
  const int* nbor;     // nbor[(j-1)*ngridmax + igrid], j = 1..6
  int ncoarse, ngridmax, nx, ny, nz;
  long ncell;          // ncoarse + 8*ngridmax
};

// A level's oct count: the host value, or (when d is set) the value in
// device memory at kernel time, so that a launch sequence can be captured
// once and replayed while the mesh changes.
__device__ __forceinline__ int level_count(int n, const int* d) { return d ? *d : n; }

struct AmrLevelArgs {
  AmrTreeView t;
  const double* uold;  // uold[v*(ncell+1) + icell]
  double* unew;
  long ustride;        // ncell + 1
  const int* active;   // active[a], a = 0..noct-1 (grid indices)
  int noct;
  const int* dnoct;    // if set, noct is read from here
  const int* dnoct_coarse;  // if set, noct_coarse is read from here
  int ilevel;
  const int* pos;      // pos[igrid] = position of the grid in its level's active list
  double* uin;         // batches (nbatch, 5, 6,6,6, nv)
  unsigned char* ok;   // (nbatch, 6,6,6, nv) refined flags
  double* flux;        // (nbatch, 3, 5, 3,3,3, nv)
  int nv;
  int interpol_type, interpol_var;
  HydroParams hp;
  const int* active_coarse;  // grids of level ilevel-1 (reflux targets)
  int noct_coarse;
  int* error;                // set nonzero if the tree violates the refinement rule
};

// Every level's grid list and host oct count (index = level), for a kernel
// that walks several levels.
constexpr int kConsMaxLevels = 33;  // RAMSES_MAX_LEVELS + 1
struct ConsLevels {
  const int* active[kConsMaxLevels];
  int noct[kConsMaxLevels];
};

}  // namespace ramses
