// Oct-tree index helpers shared by the level sweep (amr.hip) and the
// refinement kernels (refine.hip).  Index conventions are the reference's:
// cell ind (1..8) of grid igrid is ncoarse + (ind-1)*ngridmax + igrid, the
// arrays are passed so that element i of the reference is p[i].
#pragma once

#include "hydro_device.hpp"
#include "amr_args.hpp"

namespace ramses {

__device__ __forceinline__ long cell_of(const AmrTreeView& t, int ind, int igrid) {
  return (long)t.ncoarse + (long)(ind - 1) * t.ngridmax + igrid;
}
__device__ __forceinline__ int nbor_of(const AmrTreeView& t, int igrid, int j) {
  return t.nbor[(long)(j - 1) * t.ngridmax + igrid];
}

// Neighbour father cell (i1,j1,k1 in 0..2) of father cell f of a level>1 oct:
// the get3cubepos chain (z hop, then y, then x) over son(nbor(...)),
// amr/nbors_utils.f90:199-300.
__device__ inline int nbor_father_cell(const AmrTreeView& t, long f, int i1, int j1, int k1, int ilevel) {
  if (ilevel == 1) {  // coarse grid: periodic index arithmetic (amr/nbors_utils.f90:28-140)
    const int nxny = t.nx * t.ny;
    const int iz = (int)((f - 1) / nxny), iy = (int)((f - 1 - (long)iz * nxny) / t.nx);
    const int ix = (int)(f - 1 - (long)iy * t.nx - (long)iz * nxny);
    int iix = ix + i1 - 1, iiy = iy + j1 - 1, iiz = iz + k1 - 1;
    if (iix < 0) iix = t.nx - 1;
    if (iix > t.nx - 1) iix = 0;
    if (iiy < 0) iiy = t.ny - 1;
    if (iiy > t.ny - 1) iiy = 0;
    if (iiz < 0) iiz = t.nz - 1;
    if (iiz > t.nz - 1) iiz = 0;
    return 1 + iix + iiy * t.nx + iiz * nxny;
  }
  // cell indices fit in 32 bits (checked at the entry points): 32-bit division
  const unsigned r = (unsigned)(f - t.ncoarse - 1);
  const int pos = (int)(r / (unsigned)t.ngridmax) + 1;
  const int g = (int)(r - (unsigned)(pos - 1) * (unsigned)t.ngridmax) + 1;
  const int px = (pos - 1) & 1, py = ((pos - 1) >> 1) & 1, pz = (pos - 1) >> 2;
  const int cx = px + i1 - 1, cy = py + j1 - 1, cz = pz + k1 - 1;
  int gg = g;
  if (cz != pz && (cz < 0 || cz > 1) && gg > 0) gg = t.son[nbor_of(t, gg, cz < 0 ? 5 : 6)];
  if (cy != py && (cy < 0 || cy > 1) && gg > 0) gg = t.son[nbor_of(t, gg, cy < 0 ? 3 : 4)];
  if (cx != px && (cx < 0 || cx > 1) && gg > 0) gg = t.son[nbor_of(t, gg, cx < 0 ? 1 : 2)];
  if (gg <= 0) return 0;
  const int ind = 1 + (cx & 1) + 2 * (cy & 1) + 4 * (cz & 1);
  return (int)cell_of(t, ind, gg);
}

// The 2*ndim+1 cells used to interpolate a buffer oct in father cell f
// (getnborfather, amr/nbors_utils.f90:404-525): f itself and its six
// neighbours at its own level, or the coarser neighbour of f's grid where a
// neighbour does not exist.
__device__ inline void nbor_father7(const AmrTreeView& t, long f, int ilevel, long out[7]) {
  out[0] = f;
  if (ilevel == 1) {
    const int nxny = t.nx * t.ny;
    const int ix3 = (int)((f - 1) / nxny);
    const int ix2 = (int)((f - 1 - (long)ix3 * nxny) / t.nx);
    const int ix1 = (int)(f - 1 - (long)ix2 * t.nx - (long)ix3 * nxny);
    const int ixs[3] = {ix1, ix2, ix3};
    const int ib[3] = {t.nx - 1, t.ny - 1, t.nz - 1};
    const int s1[3] = {1, t.nx, nxny};
    const int s2[3] = {t.nx - 1, (t.ny - 1) * t.nx, (t.nz - 1) * nxny};
    for (int d = 0; d < 3; ++d) {
      out[2 * d + 1] = ixs[d] > 0 ? f - s1[d] : f + s2[d];
      out[2 * d + 2] = ixs[d] < ib[d] ? f + s1[d] : f - s2[d];
    }
    return;
  }
  // cell indices fit in 32 bits (checked at the entry points): 32-bit division
  const unsigned r = (unsigned)(f - t.ncoarse - 1);
  const int pos = (int)(r / (unsigned)t.ngridmax) + 1;
  const int g = (int)(r - (unsigned)(pos - 1) * (unsigned)t.ngridmax) + 1;
  const int p[3] = {(pos - 1) & 1, ((pos - 1) >> 1) & 1, (pos - 1) >> 2};
  // The six neighbour links and their sons are loaded unconditionally (in
  // range for any grid g: nbor holds cells, and son[0] exists), so the
  // twelve loads issue as two independent rounds instead of a chain per
  // direction behind a branch.
  int nbc[6], nbs[6];
#pragma unroll
  for (int in = 1; in <= 6; ++in) nbc[in - 1] = nbor_of(t, g, in);
#pragma unroll
  for (int in = 1; in <= 6; ++in) nbs[in - 1] = t.son[nbc[in - 1]];
#pragma unroll
  for (int in = 1; in <= 6; ++in) {
    const int a = (in - 1) >> 1, plus = (in - 1) & 1;
    // getnborcells: the neighbour stays in the grid when it moves inward
    const bool inside = plus ? (p[a] == 0) : (p[a] == 1);
    int q[3] = {p[0], p[1], p[2]};
    q[a] = 1 - p[a];
    const int ind = 1 + q[0] + 2 * q[1] + 4 * q[2];
    const int gn = inside ? g : nbs[in - 1];
    out[in] = gn > 0 ? cell_of(t, ind, gn) : nbc[in - 1];
  }
}

// Slopes for the prolongation (compute_limiter_minmod / _central /
// compute_central, hydro/interpol_hydro.f90:449-637), a[0..6] = centre, -x,
// +x, -y, +y, -z, +z.
__device__ inline void interp_slopes(const double a[7], int itype, double w[3]) {
  if (itype == 1) {
    for (int d = 0; d < 3; ++d) {
      const double dl = 0.5 * (a[2 * d + 2] - a[0]);
      const double dr = 0.5 * (a[0] - a[2 * d + 1]);
      if (dl * dr <= 0.0)
        w[d] = 0.0;
      else
        w[d] = fmin_f(__builtin_fabs(dl), __builtin_fabs(dr)) * dl / __builtin_fabs(dl);
    }
    return;
  }
  for (int d = 0; d < 3; ++d) w[d] = 0.25 * (a[2 * d + 2] - a[2 * d + 1]);
  if (itype == 3) return;
  if (itype != 2) {
    w[0] = w[1] = w[2] = 0.0;
    return;
  }
  // monotonized central: limit by the extrema of the kernel and the corners
  double ac[8];
  for (int ind = 0; ind < 8; ++ind) ac[ind] = a[0];
  for (int d = 0; d < 3; ++d)
    for (int ind = 0; ind < 8; ++ind) {
      const double xxc = (double)((ind >> d) & 1) - 0.5;
      ac[ind] = ac[ind] + 2.0 * w[d] * xxc;
    }
  double corner = ac[0], kernel = a[1];
  for (int j = 1; j < 8; ++j) corner = fmax_f(corner, ac[j]);
  for (int j = 2; j <= 6; ++j) kernel = fmax_f(kernel, a[j]);
  double dk = a[0] - kernel, dc = a[0] - corner;
  double maxl = 0.0;
  if (dk * dc > 0.0) maxl = fmin_f(1.0, dk / dc);
  corner = ac[0];
  kernel = a[1];
  for (int j = 1; j < 8; ++j) corner = fmin_f(corner, ac[j]);
  for (int j = 2; j <= 6; ++j) kernel = fmin_f(kernel, a[j]);
  dk = a[0] - kernel;
  dc = a[0] - corner;
  double minl = 0.0;
  if (dk * dc > 0.0) minl = fmin_f(1.0, dk / dc);
  const double lim = fmin_f(minl, maxl);
  for (int d = 0; d < 3; ++d) w[d] = w[d] * lim;
}

// interpol_hydro for one buffer oct: u1[7][5] -> u2[8][5]
// (hydro/interpol_hydro.f90:268-444; interpol_var 0, 1, 2; types 0-4).
// Type 4 (with interpol_var 2 only) takes the unlimited central slope for
// the velocities and the monotonized central one for density and internal
// energy (:367-377).
__device__ inline void interpol_oct(double u1[7][5], double u2[8][5], int itype, int ivar_mode, double smallr) {
  if (ivar_mode == 1 || ivar_mode == 2) {
    for (int j = 0; j < 7; ++j) {
      double ekin = 0.0;
      for (int d = 0; d < 3; ++d) ekin = ekin + 0.5 * (u1[j][d + 1] * u1[j][d + 1]) / fmax_f(u1[j][0], smallr);
      const double erad = 0.0;
      u1[j][4] = u1[j][4] - ekin - erad;
      if (ivar_mode == 2)
        for (int d = 0; d < 3; ++d) u1[j][d + 1] = u1[j][d + 1] / fmax_f(u1[j][0], smallr);
    }
  }
  for (int v = 0; v < 5; ++v) {
    double a[7], w[3] = {0.0, 0.0, 0.0};
    for (int j = 0; j < 7; ++j) a[j] = u1[j][v];
    const int t = itype == 4 ? ((v >= 1 && v <= 3) ? 3 : 2) : itype;
    if (t >= 1 && t <= 3) interp_slopes(a, t, w);
    for (int ind = 0; ind < 8; ++ind) {
      double s = a[0];
      for (int d = 0; d < 3; ++d) s = s + w[d] * ((double)((ind >> d) & 1) - 0.5);
      u2[ind][v] = s;
    }
  }
  if (ivar_mode == 1 || ivar_mode == 2) {
    if (ivar_mode == 2) {
      for (int ind = 0; ind < 8; ++ind)
        for (int d = 0; d < 3; ++d) u2[ind][d + 1] = u2[ind][d + 1] * u2[ind][0];
      for (int d = 0; d < 3; ++d) {
        double mom = 0.0;
        for (int ind = 0; ind < 8; ++ind) mom = mom + u2[ind][d + 1] * (1.0 / 8.0);
        mom = mom - u1[0][d + 1] * u1[0][0];
        for (int ind = 0; ind < 8; ++ind) u2[ind][d + 1] = u2[ind][d + 1] - mom;
      }
    }
    for (int ind = 0; ind < 8; ++ind) {
      double ekin = 0.0;
      for (int d = 0; d < 3; ++d)
        ekin = ekin + 0.5 * (u2[ind][d + 1] * u2[ind][d + 1]) / fmax_f(u2[ind][0], smallr);
      const double erad = 0.0;
      u2[ind][4] = u2[ind][4] + ekin + erad;
    }
  }
}

// Cell `in` (1..6: -x,+x,-y,+y,-z,+z) next to cell ind of grid g at the same
// level, or 0 where that neighbour grid does not exist: getnborgrids +
// getnborcells (amr/nbors_utils.f90:530-555, 363-399).
__device__ __forceinline__ long nbor_cell(const AmrTreeView& t, int g, int ind, int in) {
  const int p[3] = {(ind - 1) & 1, ((ind - 1) >> 1) & 1, (ind - 1) >> 2};
  const int a = (in - 1) >> 1, plus = (in - 1) & 1;
  const bool inside = plus ? (p[a] == 0) : (p[a] == 1);
  int q[3] = {p[0], p[1], p[2]};
  q[a] = 1 - p[a];
  const int gn = inside ? g : t.son[nbor_of(t, g, in)];
  return gn > 0 ? cell_of(t, 1 + q[0] + 2 * q[1] + 4 * q[2], gn) : 0;
}

}  // namespace ramses
