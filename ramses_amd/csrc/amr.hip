// One level of an oct-tree (AMR) godunov sweep on the GPU:
// godfine1 (hydro/godunov_fine.f90:486-911) split into four launches over
// the level's active octs, all reading the tree (son / father / nbor) and
// the state (uold / unew) resident in HBM in the reference's memory model:
//   amr_gather   6x6x6 stencil of every oct in nvector-strided batches
//                (get3cubefather/get3cubepos, amr/nbors_utils.f90:5-300; a
//                missing neighbour oct is a buffer oct interpolated from its
//                father's 7-point neighbourhood: getnborfather :404-525 and
//                interpol_hydro, hydro/interpol_hydro.f90:268-637);
//   amr_unsplit  unsplit of every batch (the device code of unsplit.hip),
//                storing a zero flux at faces next to a refined cell
//                (:720-749);
//   amr_update   the level update unew += F(i) - F(i+1) (:751-792);
//   amr_reflux   the conservative correction of coarser leaf neighbours
//                (:798-908), gathered per coarse cell and summed in the
//                reference's order (batch, dimension, side, face) so the
//                result is bitwise and deterministic without atomics.
// Index conventions are the reference's (1-based cells and grids; arrays
// are passed so that element i of the reference is p[i]).
#include "hydro_device.hpp"
#include "unsplit_args.hpp"
#include "amr_args.hpp"
#include "amr_device.hpp"
#include "unsplit_device.hpp"

namespace ramses {
hipError_t launch_cons_finish(const double* partials, long nblocks, double* out, int accumulate, hipStream_t s);


namespace {

// One (batch b, father neighbour j in 0..26, lane) task.
__device__ __forceinline__ void gather_task(const AmrLevelArgs& a, int noct, int b, int j, int lane) {
  const int nv = a.nv;
  const int oct = b * nv + lane;
  if (oct >= noct) return;
  const AmrTreeView& t = a.t;
  const int igrid = a.active[oct];
  const long f = t.father[igrid];
  const int i1 = j % 3, j1 = (j / 3) % 3, k1 = j / 9;
  const long nf = nbor_father_cell(t, f, i1, j1, k1, a.ilevel);
  if (nf <= 0) {
    atomicOr(a.error, 1);
    return;
  }
  const int igrid_n = t.son[nf];
  const size_t us = (size_t)nv * 216;               // variable stride in uin
  double* uin = a.uin + (size_t)b * us * 5 + lane;
  unsigned char* ok = a.ok + (size_t)b * us + lane;
  // Every load of a branch is issued before its first store: the stores into
  // uin could alias uold as far as the compiler knows, so interleaving them
  // would serialise the 40 loads into a chain of memory latencies.
  if (igrid_n > 0) {
    double val[8][5];
    unsigned char okv[8];
#pragma unroll
    for (int ind = 1; ind <= 8; ++ind) {
      const long c = cell_of(t, ind, igrid_n);
#pragma unroll
      for (int v = 0; v < 5; ++v) val[ind - 1][v] = a.uold[(long)v * a.ustride + c];
      okv[ind - 1] = t.son[c] > 0 ? 1 : 0;
    }
#pragma unroll
    for (int ind = 1; ind <= 8; ++ind) {
      const int i2 = (ind - 1) & 1, j2 = ((ind - 1) >> 1) & 1, k2 = (ind - 1) >> 2;
      const int i3 = 1 + 2 * (i1 - 1) + i2, j3 = 1 + 2 * (j1 - 1) + j2, k3 = 1 + 2 * (k1 - 1) + k2;
      const size_t s = (size_t)((i3 + 1) + 6 * ((j3 + 1) + 6 * (k3 + 1))) * nv;
#pragma unroll
      for (int v = 0; v < 5; ++v) uin[s + (size_t)v * us] = val[ind - 1][v];
      ok[s] = okv[ind - 1];
    }
  } else {
    long nb7[7];
    nbor_father7(t, nf, a.ilevel, nb7);
    double u1[7][5], u2[8][5];
#pragma unroll
    for (int q = 0; q < 7; ++q)
#pragma unroll
      for (int v = 0; v < 5; ++v) u1[q][v] = a.uold[(long)v * a.ustride + nb7[q]];
    interpol_oct(u1, u2, a.interpol_type, a.interpol_var, a.hp.smallr);
#pragma unroll
    for (int ind = 1; ind <= 8; ++ind) {
      const int i2 = (ind - 1) & 1, j2 = ((ind - 1) >> 1) & 1, k2 = (ind - 1) >> 2;
      const int i3 = 1 + 2 * (i1 - 1) + i2, j3 = 1 + 2 * (j1 - 1) + j2, k3 = 1 + 2 * (k1 - 1) + k2;
      const size_t s = (size_t)((i3 + 1) + 6 * ((j3 + 1) + 6 * (k3 + 1))) * nv;
#pragma unroll
      for (int v = 0; v < 5; ++v) uin[s + (size_t)v * us] = u2[ind - 1][v];
      ok[s] = 0;
    }
  }
}

// Flat grid-stride over the 27*nv tasks of every batch, lane fastest, so a
// wavefront covers two neighbours j of 32 octs and every (cell, variable)
// store is one coalesced segment.  A level of a few thousand octs spreads
// over all CUs instead of one workgroup per batch: each task is a chain of
// dependent tree loads, and the chains of one CU share its memory pipeline.
__global__ __launch_bounds__(256) void amr_gather(AmrLevelArgs a) {
  const int noct = level_count(a.noct, a.dnoct);
  const int nv = a.nv;
  const long per_batch = 27L * nv;
  const long ntask = (long)((noct + nv - 1) / nv) * per_batch;
  for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < ntask; t += (long)gridDim.x * blockDim.x) {
    const int b = (int)(t / per_batch);
    const int rem = (int)(t - b * per_batch);
    gather_task(a, noct, b, rem / nv, rem % nv);
  }
}

__device__ __forceinline__ size_t fslot(int i3, int j3, int k3) { return (size_t)((i3 - 1) + 3 * ((j3 - 1) + 3 * (k3 - 1))); }

// Level update of one (cell ind, variable v) of lane `lane` of batch b, q =
// (ind-1)*5 + v: unew += flux(i3) - flux(i3+1) for idim = 1, 2, 3 in order
// (hydro/godunov_fine.f90:751-792).
__device__ __forceinline__ void update_cell(const AmrLevelArgs& a, int b, int q, int lane) {
  const int nv = a.nv;
  const size_t fvs = (size_t)nv * 27;
  const double* flux = a.flux + (size_t)b * fvs * 15;
  const int ind = q / 5 + 1, v = q % 5;
  const int i2 = (ind - 1) & 1, j2 = ((ind - 1) >> 1) & 1, k2 = (ind - 1) >> 2;
  const int i3 = 1 + i2, j3 = 1 + j2, k3 = 1 + k2;
  const long c = (long)v * a.ustride + cell_of(a.t, ind, a.active[b * nv + lane]);
  double acc = a.unew[c];
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    const int i0 = d == 0, j0 = d == 1, k0 = d == 2;
    const size_t base = (size_t)27 * (v + 5 * d);
    const double fl = flux[(fslot(i3, j3, k3) + base) * nv + lane];
    const double fr = flux[(fslot(i3 + i0, j3 + j0, k3 + k0) + base) * nv + lane];
    acc = acc + (fl - fr);
  }
  a.unew[c] = acc;
}

// The level's unsplit over all batches (workgroups of 128), the flux reset
// at faces next to a refined cell folded into the face stores: one (batch,
// face, lane) task per thread on levels of at most u.face_octs octs, where
// the chain of dependent steps matters more than the arithmetic, else one
// (batch, row, lane) task per thread.  Small levels spread over every CU.
// The levels carry no gravity array (u.grav is null), so the tasks are the
// gravity-free ones.
template <typename R, int RS, bool S27, int ST>
__global__ __launch_bounds__(768) void amr_unsplit(AmrLevelArgs a, UnsplitArgs u) {
  const int noct = level_count(a.noct, a.dnoct);
  const int nv = a.nv;
  const int nb = (noct + nv - 1) / nv;
  const int ngrid_last = noct - (nb - 1) * nv;
  const long stride = (long)gridDim.x * blockDim.x;
  if (noct <= u.face_octs) {
    const long ntask = (long)nb * 36 * nv;
    for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < ntask; t += stride)
      unsplit_dev::face_task<R, RS, S27, true, ST, false>(u, t, nb, ngrid_last);
  } else {
    const long ntask = (long)nb * 12 * nv;
    for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < ntask; t += stride)
      unsplit_dev::unsplit_task<R, RS, S27, true, ST, false>(u, t, nb, ngrid_last);
  }
}

// The level update, one (batch, cell, variable, lane) task per thread.
__global__ __launch_bounds__(256) void amr_update(AmrLevelArgs a) {
  const int noct = level_count(a.noct, a.dnoct);
  const int nv = a.nv;
  const long per_batch = 40L * nv;
  const long ntask = (long)((noct + nv - 1) / nv) * per_batch;
  for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < ntask; t += (long)gridDim.x * blockDim.x) {
    const int b = (int)(t / per_batch);
    const int rem = (int)(t - b * per_batch);
    const int lane = rem % nv;
    if (b * nv + lane < noct) update_cell(a, b, rem / nv, lane);
  }
}

// Minmod (the reference's default slope) with the limiter fixed at compile
// time; other slope types take it at run time.
template <typename R, int RS, bool S27>
void launch_solve_t(const AmrLevelArgs& a, const UnsplitArgs& u, unsigned grid, unsigned block, hipStream_t s) {
  if (!S27 && a.hp.slope_type == 1)
    hipLaunchKernelGGL((amr_unsplit<R, RS, S27, 1>), dim3(grid), dim3(block), 0, s, a, u);
  else
    hipLaunchKernelGGL((amr_unsplit<R, RS, S27, -1>), dim3(grid), dim3(block), 0, s, a, u);
}

template <typename R>
void launch_solve_r(const AmrLevelArgs& a, const UnsplitArgs& u, unsigned grid, unsigned block, hipStream_t s) {
  const bool s27 = a.hp.slope_type == 3;
  switch (a.hp.riemann) {
    case kHLLC:
      s27 ? launch_solve_t<R, kHLLC, true>(a, u, grid, block, s) : launch_solve_t<R, kHLLC, false>(a, u, grid, block, s);
      break;
    case kHLL:
      s27 ? launch_solve_t<R, kHLL, true>(a, u, grid, block, s) : launch_solve_t<R, kHLL, false>(a, u, grid, block, s);
      break;
    case kACOUSTIC:
    case kEXACT:
      s27 ? launch_solve_t<R, kDynamicRiemann, true>(a, u, grid, block, s)
          : launch_solve_t<R, kDynamicRiemann, false>(a, u, grid, block, s);
      break;
    default:
      s27 ? launch_solve_t<R, kLLF, true>(a, u, grid, block, s) : launch_solve_t<R, kLLF, false>(a, u, grid, block, s);
      break;
  }
}

// Conservative correction of the coarser leaf cells next to the level's
// octs.  Thread = (one cell of a coarse-level oct, one variable): the tree
// walk is repeated per variable, which keeps five times more threads in
// flight on the small coarse levels where this kernel is latency-bound.  A
// coarse leaf c receives from each fine oct o that sits across one of its
// faces the four (reset) fluxes of o's boundary faces on c's side, times
// -1/8 (c left of o) or +1/8 (c right of o).  The reference adds them in
// loop order: batch of o, then dimension, then side, then face (k3, j3, i3);
// the same order is rebuilt here from the position of o in the level's
// active list.
__device__ __forceinline__ void reflux_task(const AmrLevelArgs& a, long tix) {
  const long cidx = tix / 5;
  const int v = (int)(tix % 5);
  const AmrTreeView& t = a.t;
  const int g = a.active_coarse[cidx >> 3];
  const int ind = (int)(cidx & 7) + 1;
  const long c = cell_of(t, ind, g);
  if (t.son[c] != 0) return;  // only leaves receive the correction (:811,863)
  const int p[3] = {(ind - 1) & 1, ((ind - 1) >> 1) & 1, (ind - 1) >> 2};
  // the six neighbour cells first (independent loads), then their sons
  long nc[6];
#pragma unroll
  for (int in = 1; in <= 6; ++in) {
    const int ax = (in - 1) >> 1, plus = (in - 1) & 1;
    const bool inside = plus ? (p[ax] == 0) : (p[ax] == 1);
    int q[3] = {p[0], p[1], p[2]};
    q[ax] = 1 - p[ax];
    const int gn = inside ? g : t.son[nbor_of(t, g, in)];
    nc[in - 1] = gn > 0 ? cell_of(t, 1 + q[0] + 2 * q[1] + 4 * q[2], gn) : 0;
  }
  int so[6];
#pragma unroll
  for (int k = 0; k < 6; ++k) so[k] = nc[k] > 0 ? t.son[nc[k]] : 0;
  int pz[6];
#pragma unroll
  for (int k = 0; k < 6; ++k) pz[k] = so[k] > 0 ? a.pos[so[k]] : 0;
  // up to 6 contributing octs: key = batch*6 + 2*d + side
  int keys[6], posn[6], ne = 0;
#pragma unroll
  for (int in = 1; in <= 6; ++in) {
    if (so[in - 1] <= 0) continue;
    const int ax = (in - 1) >> 1, plus = (in - 1) & 1;
    // c is on o's minus side when n is c's plus neighbour
    const int side = plus ? 0 : 1;
    keys[ne] = (pz[in - 1] / a.nv) * 6 + 2 * ax + side;
    posn[ne] = pz[in - 1];
    ++ne;
  }
  if (ne == 0) return;
  for (int i = 1; i < ne; ++i)  // insertion sort, at most 6 entries
    for (int k = i; k > 0 && keys[k - 1] > keys[k]; --k) {
      const int tk = keys[k]; keys[k] = keys[k - 1]; keys[k - 1] = tk;
      const int tp = posn[k]; posn[k] = posn[k - 1]; posn[k - 1] = tp;
    }
  const int nv = a.nv;
  const size_t fvs = (size_t)nv * 27;
  const double oneontwotondim = 1.0 / 8.0;
  // All contributions are loaded first (independent loads, issued together),
  // then summed in the reference's order: contributor e by key, and within
  // it the 4 faces of the oct side in (k3, j3, i3) loop order.
  double fv[6][4];
#pragma unroll
  for (int e = 0; e < 6; ++e) {
    if (e >= ne) break;
    const int d = (keys[e] % 6) >> 1, side = keys[e] & 1;
    const int b = posn[e] / nv, lane = posn[e] % nv;
    const double* flux = a.flux + (size_t)b * fvs * 15 + lane + (size_t)27 * (v + 5 * d) * nv;
    // the side's 2x2 faces: the transverse coordinates run over {1,2} and the
    // normal one is 1 (side 0, the oct's minus face) or 3 (side 1)
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int t1 = 1 + (q & 1), t2 = 1 + (q >> 1);  // inner loop first
      int i3, j3, k3;
      if (d == 0) { i3 = side ? 3 : 1; j3 = t1; k3 = t2; }
      else if (d == 1) { i3 = t1; j3 = side ? 3 : 1; k3 = t2; }
      else { i3 = t1; j3 = t2; k3 = side ? 3 : 1; }
      fv[e][q] = flux[fslot(i3, j3, k3) * nv];
    }
  }
  double s = a.unew[(long)v * a.ustride + c];
#pragma unroll
  for (int e = 0; e < 6; ++e) {
    if (e >= ne) break;
    const int side = keys[e] & 1;
#pragma unroll
    for (int q = 0; q < 4; ++q) s = side ? s + fv[e][q] * oneontwotondim : s - fv[e][q] * oneontwotondim;
  }
  a.unew[(long)v * a.ustride + c] = s;
}

__global__ void amr_reflux(AmrLevelArgs a) {
  const long n = 40L * level_count(a.noct_coarse, a.dnoct_coarse);
  const long stride = (long)gridDim.x * blockDim.x;
  for (long tix = (long)blockIdx.x * blockDim.x + threadIdx.x; tix < n; tix += stride) reflux_task(a, tix);
}

// set_unew / set_uold over the level's active cells (hydro/godunov_fine.f90:58-65, 193-197).
// The per-cell kernels below take the oct count as a host value or, with
// dn set, from device memory (grid-stride loops over a persistent grid).
__global__ void amr_copy_level(const int* active, int noct_h, const int* dn, AmrTreeView t, const double* src,
                               double* dst, long ustride) {
  const int noct = level_count(noct_h, dn);
  const long n = 8L * noct, stride = (long)gridDim.x * blockDim.x;
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    const long c = cell_of(t, (int)(i / noct) + 1, active[i % noct]);
    for (int v = 0; v < 5; ++v) dst[(long)v * ustride + c] = src[(long)v * ustride + c];
  }
}

// upl for one refined cell c with son oct s: the average of its children
// (hydro/interpol_hydro.f90:72-263). Density with the smallr floor, then
// momenta and total energy, each summed over the children in order and /8;
// with interpol_var 1 or 2 the energy is instead rebuilt from the children's
// average internal energy plus the kinetic energy of the averaged state.
__device__ inline void upl_cell(const AmrTreeView& t, double* u, long ustride, long c, int s, double smallr,
                                int ivar_mode) {
  double getx = 0.0;
  for (int ind = 1; ind <= 8; ++ind) getx = getx + fmax_f(u[cell_of(t, ind, s)], smallr);
  double out[5];
  out[0] = getx / 8.0;
  for (int v = 1; v < 5; ++v) {
    getx = 0.0;
    for (int ind = 1; ind <= 8; ++ind) getx = getx + u[(long)v * ustride + cell_of(t, ind, s)];
    out[v] = getx / 8.0;
  }
  if (ivar_mode == 1 || ivar_mode == 2) {
    getx = 0.0;
    for (int ind = 1; ind <= 8; ++ind) {
      const long cs = cell_of(t, ind, s);
      const double rho = fmax_f(u[cs], smallr);
      double ekin = 0.0;
      for (int d = 1; d <= 3; ++d) {
        const double m = u[(long)d * ustride + cs];
        ekin = ekin + 0.5 * (m * m) / rho;
      }
      getx = getx + u[4 * ustride + cs] - ekin;
    }
    const double rho = fmax_f(out[0], smallr);
    double ekin = 0.0;
    for (int d = 1; d <= 3; ++d) ekin = ekin + 0.5 * (out[d] * out[d]) / rho;
    out[4] = getx / 8.0 + ekin;
  }
  for (int v = 0; v < 5; ++v) u[(long)v * ustride + c] = out[v];
}

// upload_fine: refined cells of the level get the average of their children
// (hydro/interpol_hydro.f90:5-67).
__global__ void amr_upload(const int* active, int noct_h, const int* dn, AmrTreeView t, double* u, long ustride,
                           double smallr, int ivar_mode) {
  const long n = 8L * level_count(noct_h, dn), stride = (long)gridDim.x * blockDim.x;
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    const long c = cell_of(t, (int)(i % 8) + 1, active[i / 8]);
    const int s = t.son[c];
    if (s > 0) upl_cell(t, u, ustride, c, s, smallr, ivar_mode);
  }
}

// set_uold followed by upload_fine of the level in one pass
// (hydro/godunov_fine.f90:135-232, hydro/interpol_hydro.f90:5-263): a leaf
// takes unew, a refined cell the average of its children's uold (set_uold
// writes every cell and upload_fine then rewrites the refined ones, so the
// result is the same).
__global__ void amr_finish(const int* active, int noct_h, const int* dn, AmrTreeView t, const double* unew, double* u,
                           long ustride, double smallr, int ivar_mode) {
  const long n = 8L * level_count(noct_h, dn), stride = (long)gridDim.x * blockDim.x;
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    const long c = cell_of(t, (int)(i % 8) + 1, active[i / 8]);
    const int s = t.son[c];
    if (s <= 0) {
#pragma unroll
      for (int v = 0; v < 5; ++v) u[(long)v * ustride + c] = unew[(long)v * ustride + c];
    } else {
      upl_cell(t, u, ustride, c, s, smallr, ivar_mode);
    }
  }
}

// courant_fine over the level's leaf cells (hydro/courant_fine.f90:44-131):
// block max of the cmpdt wave speed, block dt limit min-combined (exact,
// see cmpdt_w).
__global__ void amr_courant(const int* active, int noct_h, const int* dn, AmrTreeView t, const double* u,
                            long ustride, double dx, HydroParams hp, double* d_dt) {
  const long n = 8L * level_count(noct_h, dn), stride = (long)gridDim.x * blockDim.x;
  double w = 0.0;
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    const long c = cell_of(t, (int)(i % 8) + 1, active[i / 8]);
    if (t.son[c] == 0)
      w = fmax_f(w, cmpdt_w(u[c], u[ustride + c], u[2 * ustride + c], u[3 * ustride + c], u[4 * ustride + c], hp));
  }
  for (int off = 32; off > 0; off >>= 1) w = fmax_f(w, __shfl_xor(w, off, 64));
  __shared__ double red[4];
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = w;
  __syncthreads();
  if (threadIdx.x == 0) {
    double m = red[0];
    for (int k = 1; k < (int)(blockDim.x >> 6); ++k) m = fmax_f(m, red[k]);
    if (m > 0.0) atomicMin(reinterpret_cast<unsigned long long*>(d_dt), dt_bits(cmpdt_from_w(m, dx, hp)));
  }
}

// set_unew and courant_fine of a level in one pass over its cells
// (hydro/godunov_fine.f90:40-130, hydro/courant_fine.f90:44-131): every
// cell copies uold to unew, every leaf joins the block max of the cmpdt wave
// speed, and the block's dt limit is min-combined into *d_dt (exact: the dt
// limit is non-increasing in the wave speed).
__global__ void amr_begin(const int* active, int noct_h, const int* dn, AmrTreeView t, const double* u, double* unew,
                          long ustride, double dx, HydroParams hp, double* d_dt, double* partials) {
  const int noct = level_count(noct_h, dn);
  const long n = 8L * noct, stride = (long)gridDim.x * blockDim.x;
  double w = 0.0;
  ConsSums cs{0.0, 0.0, 0.0};  // courant_fine's sums over the leaves (partials set)
  const double vol = dx * dx * dx;
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    const long c = cell_of(t, (int)(i / noct) + 1, active[i % noct]);
    double q[5];
#pragma unroll
    for (int v = 0; v < 5; ++v) q[v] = u[(long)v * ustride + c];
#pragma unroll
    for (int v = 0; v < 5; ++v) unew[(long)v * ustride + c] = q[v];
    if (t.son[c] == 0) {
      w = fmax_f(w, cmpdt_w(q[0], q[1], q[2], q[3], q[4], hp));
      if (partials) cons_add(cs, q[0], q[1], q[2], q[3], q[4], vol, hp.smallr);
    }
  }
  if (partials) {
    __shared__ double cred[3 * 4];
    cons_block_store(cs, partials, blockIdx.x, cred);
  }
  for (int off = 32; off > 0; off >>= 1) w = fmax_f(w, __shfl_xor(w, off, 64));
  __shared__ double red[4];
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = w;
  __syncthreads();
  if (threadIdx.x == 0) {
    double m = red[0];
    for (int k = 1; k < (int)(blockDim.x >> 6); ++k) m = fmax_f(m, red[k]);
    if (m > 0.0) atomicMin(reinterpret_cast<unsigned long long*>(d_dt), dt_bits(cmpdt_from_w(m, dx, hp)));
  }
}

// The level time-step bookkeeping of amr_step (amr/amr_step.f90), one
// thread, so a coarse step keeps every dt on the device:
//   op 0  newdt_fine prologue: dtold(l) = dtnew(l); dtnew(l) = start, capped
//         by dtnew(l-1)/nsubcycle(l-1) below levelmin (:327-330; the courant
//         reduction then min-combines into it -- min is exact, so the order
//         of the three minima does not matter)
//   op 1  no finer level: dtold(l+1) = dtnew(l+1) = dtnew(l)/nsubcycle(l),
//         t += dtnew(l) (:353-362, amr/update_time.f90:217,331)
//   op 2  end of the step: dtnew(l-1) = dtnew(l) if nsubcycle(l-1) == 1, or
//         dtold(l) + dtnew(l) after the second substep (:573-576)
// With dn (the device oct counts, indexed by level) the operations are those
// of a level step that runs only when level l has octs, and op 1 only when
// no finer level does (amr/amr_step.f90:345-365); cnt, if set, counts
// update_time calls (cnt[0], nstep) and cell updates (cnt[1]).
__global__ void amr_dt_kernel(int op, double* dt, double* dtold, double* time, int l, int coarser, int nsub,
                              int icount, double start, const int* dn, long long* cnt) {
  if (dn) {
    if (dn[l] == 0) return;
    if (op == 1 && coarser && dn[l + 1] > 0) return;
    if (cnt && op == 0) cnt[1] += 8LL * dn[l];
    if (cnt && op == 1) cnt[0] += 1;
  }
  if (op == 0) {
    dtold[l] = dt[l];
    double v = start;
    if (coarser) v = fmin_f(dt[l - 1] / (double)nsub, v);
    dt[l] = v;
  } else if (op == 1) {
    if (coarser) {  // here: a finer level index exists
      dtold[l + 1] = dt[l] / (double)nsub;
      dt[l + 1] = dt[l] / (double)nsub;
    }
    time[0] = time[0] + dt[l];
  } else {
    if (nsub == 1) dt[l - 1] = dt[l];
    if (icount == 2) dt[l - 1] = dtold[l] + dt[l];
  }
}

// update_time's conservation baselines (amr/update_time.f90:229-246) at
// the first update_time of a coarse step: mass_tot_0 = mass_tot and einit =
// epot_tot + ekin_tot (epot_tot = 0 without gravity), each only while still
// zero.  mass_tot and ekin_tot are then the courant_fine sums of the levels
// the step's first descent has visited (amr/amr_step.f90:327-352): levelmin
// down to the first level with no finer octs, each level's sum added as one
// term (hydro/courant_fine.f90:152-153).  A level's sum runs in the
// reference's serial order -- batches of nvector grids, cell position ind,
// grid in the batch (:44-102) -- so the values have its rounding.  The
// workgroup gathers chunks of kConsChunk terms of that order into LDS, many
// threads loading in parallel, so the tree walks and loads of a chunk
// overlap.  Threads 0 and 64 add a chunk's mass and energy terms in order
// while the wavefronts after the first two gather the next chunk into the
// second buffer.  A cell
// that is not a leaf adds +0.0, which leaves the positive running sum
// unchanged.  With dn the level counts are read from device memory and
// `armed` (nonzero: nstep_coarse changed since the last update_time)
// decides at run time.
constexpr int kConsChunk = 2048;
constexpr int kConsThreads = 1024;


// Terms [q0, q0 + kConsChunk) of a level's serial order into tp (mass,
// energy), loaded by threads [first, kConsThreads); past the level's end,
// zeros.
__device__ __forceinline__ void cons_gather(const int* active, int n, int nv, AmrTreeView t, const double* u,
                                            long ustride, double vol, long q0, int first, double2* tp) {
  // term q: batch b0 = q / (8 nv) * nv (only the last batch is short, so
  // 8 b0 terms precede it), then slot s = q - 8 b0 = (ind - 1) * ng + grid
  const long nterm = 8L * n;
  for (int k = (int)threadIdx.x - first; k < kConsChunk; k += kConsThreads - first) {
    double vm = 0.0, ve = 0.0;
    if (q0 + k < nterm) {
      const int q = (int)(q0 + k);  // 8 n <= 8 ngridmax fits the reference's 32-bit cell indices
      const int b0 = q / (8 * nv) * nv;
      const int ng = min(nv, n - b0);
      const int sl = q - 8 * b0;
      const long c = cell_of(t, sl / ng + 1, active[b0 + sl % ng]);
      if (t.son[c] == 0) {
        vm = u[c] * vol;                 // hydro/courant_fine.f90:95-97
        ve = u[4 * ustride + c] * vol;   // :100-102
      }
    }
    tp[k] = make_double2(vm, ve);
  }
}

// Adds component `which` (0: mass, 1: energy) of the first cnt terms of tp
// to acc, in order, by one thread.  The LDS loads run ahead of the chain of
// additions.  Mass and energy are summed by threads of two different
// wavefronts: an FP64 addition occupies its SIMD for 8 cycles whatever the
// number of active lanes, so the two chains in one wavefront shared one
// SIMD (2.5 ms for the adaptive Sedov run's baselines, 2.0 ms split).
__device__ __forceinline__ void cons_add(const double2* tp, int cnt, int which, double& acc) {
  int k0 = 0;
  for (; k0 + 16 <= cnt; k0 += 16) {
    double v[16];
#pragma unroll
    for (int j = 0; j < 16; ++j) v[j] = which == 0 ? tp[k0 + j].x : tp[k0 + j].y;
#pragma unroll
    for (int j = 0; j < 16; ++j) acc = acc + v[j];
  }
  for (; k0 < cnt; ++k0) acc = acc + (which == 0 ? tp[k0].x : tp[k0].y);
}

__global__ __launch_bounds__(kConsThreads) void amr_cons0_kernel(ConsLevels lv, const int* dn, AmrTreeView t,
                                                                const double* u, long ustride, int levelmin,
                                                                int nlevelmax, double boxlen, int nv, const int* armed,
                                                                double* cons0) {
  __shared__ double2 tp[2][kConsChunk];  // double-buffered chunks of (mass, energy) terms
  if (armed && *armed == 0) return;
  if (cons0[0] != 0.0 && cons0[1] != 0.0) return;
  const bool adder_waves = threadIdx.x < 128;
  const int which = threadIdx.x == 0 ? 0 : 1;  // thread 0 sums the mass, thread 64 the energy
  double tot = 0.0;                            // threads 0 and 64: mass_tot, ekin_tot
  for (int l = levelmin; l <= nlevelmax; ++l) {
    const int n = dn ? dn[l] : lv.noct[l];
    if (n <= 0) break;
    const int* active = lv.active[l];
    const double dx = __builtin_ldexp(1.0, -l) * boxlen;
    const double vol = dx * dx * dx;
    const long nterm = 8L * n;
    const long nchunk = (nterm + kConsChunk - 1) / kConsChunk;
    double acc = 0.0;
    cons_gather(active, n, nv, t, u, ustride, vol, 0, 0, tp[0]);
    __syncthreads();
    // threads 0 and 64 add chunk c while the wavefronts after the first two
    // gather c + 1
    for (long c = 0; c < nchunk; ++c) {
      const int b = (int)(c & 1);
      if (threadIdx.x == 0 || threadIdx.x == 64)
        cons_add(tp[b], (int)min((long)kConsChunk, nterm - c * kConsChunk), which, acc);
      else if (!adder_waves && c + 1 < nchunk)
        cons_gather(active, n, nv, t, u, ustride, vol, (c + 1) * kConsChunk, 128, tp[b ^ 1]);
      __syncthreads();
    }
    tot = tot + acc;
    if (l == nlevelmax || (dn ? dn[l + 1] : lv.noct[l + 1]) == 0) break;
  }
  if (threadIdx.x == 0 && cons0[1] == 0.0) cons0[1] = tot;        // mass_tot_0
  if (threadIdx.x == 64 && cons0[0] == 0.0) cons0[0] = 0.0 + tot;  // einit = epot_tot + ekin_tot
}

__global__ void clear_bits(int* p, int mask) { p[0] &= ~mask; }

__global__ void set_ints(int* p, int n, int v) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

}  // namespace

// Grid of a per-cell launch over n items: sized to n for a host count, a
// fixed persistent grid for a count read from device memory.
inline unsigned cell_grid(long n, const int* dn) {
  return dn ? persist_blocks() : (unsigned)((n + 255) / 256);
}

// gather + unsplit + update + reflux of one level; dt is read from d_dt.
// With a.dnoct / a.dnoct_coarse the counts are read from device memory and
// the grids are persistent (batch loops), so the launches can be captured.
hipError_t launch_amr_godunov(const AmrLevelArgs& a, double dx, const double* d_dt, double* tmp, bool fp32,
                              hipStream_t s) {
  const bool dev = a.dnoct != nullptr;
  if (!dev && a.noct <= 0) return hipSuccess;
  const int nbatch = (a.noct + a.nv - 1) / a.nv;
  const unsigned gblocks = (unsigned)(((long)nbatch * 27 * a.nv + 255) / 256);
  hipLaunchKernelGGL(amr_gather, dim3(dev ? persist_blocks() : gblocks), dim3(256), 0, s, a);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  UnsplitArgs u;
  u.uin = a.uin;
  u.grav = nullptr;
  u.flux = a.flux;
  u.tmp = tmp;
  u.dx[0] = u.dx[1] = u.dx[2] = dx;
  u.dt = 0.0;
  u.d_dt = d_dt;
  u.ngroups = nbatch;
  u.ngrid_last = a.noct - (nbatch - 1) * a.nv;
  u.nv = a.nv;
  u.hp = a.hp;
  u.ok = a.ok;
  u.face_octs = amr_face_octs();
  const unsigned block = 128;
  const long tasks = (long)nbatch * (a.noct <= u.face_octs ? 36 : 12) * a.nv;
  // the fixed grids of the captured mode keep the 384 x persist_blocks()
  // threads of one batch per 384-thread workgroup
  const unsigned grid = dev ? 3 * persist_blocks() : (unsigned)((tasks + block - 1) / block);
  if (fp32)
    launch_solve_r<float>(a, u, grid, block, s);
  else
    launch_solve_r<double>(a, u, grid, block, s);
  e = hipGetLastError();
  if (e != hipSuccess) return e;
  const unsigned ugrid = dev ? 2 * persist_blocks() : (unsigned)(((long)nbatch * 40 * a.nv + 255) / 256);
  hipLaunchKernelGGL(amr_update, dim3(ugrid), dim3(256), 0, s, a);
  e = hipGetLastError();
  if (e != hipSuccess) return e;
  if (a.active_coarse && (a.dnoct_coarse || a.noct_coarse > 0)) {
    const long n = 40L * a.noct_coarse;
    hipLaunchKernelGGL(amr_reflux, dim3(cell_grid(n, a.dnoct_coarse)), dim3(256), 0, s, a);
  }
  return hipGetLastError();
}

hipError_t launch_amr_copy(const int* active, int noct, const AmrTreeView& t, const double* src, double* dst,
                           long ustride, hipStream_t s, const int* dn) {
  if (!dn && noct <= 0) return hipSuccess;
  hipLaunchKernelGGL(amr_copy_level, dim3(cell_grid(8L * noct, dn)), dim3(256), 0, s, active, noct, dn, t, src, dst,
                     ustride);
  return hipGetLastError();
}

hipError_t launch_amr_upload(const int* active, int noct, const AmrTreeView& t, double* u, long ustride,
                             double smallr, int ivar_mode, hipStream_t s, const int* dn) {
  if (!dn && noct <= 0) return hipSuccess;
  hipLaunchKernelGGL(amr_upload, dim3(cell_grid(8L * noct, dn)), dim3(256), 0, s, active, noct, dn, t, u, ustride,
                     smallr, ivar_mode);
  return hipGetLastError();
}

hipError_t launch_amr_courant(const int* active, int noct, const AmrTreeView& t, const double* u, long ustride,
                              double dx, const HydroParams& hp, double* d_dt, hipStream_t s, const int* dn) {
  if (!dn && noct <= 0) return hipSuccess;
  hipLaunchKernelGGL(amr_courant, dim3(cell_grid(8L * noct, dn)), dim3(256), 0, s, active, noct, dn, t, u, ustride,
                     dx, hp, d_dt);
  return hipGetLastError();
}

// With partials (room for amr_begin_blocks(noct, dn) workgroups) the
// level's conservation sums are added into sums[0..2] after the pass.
hipError_t launch_amr_begin(const int* active, int noct, const AmrTreeView& t, const double* u, double* unew,
                            long ustride, double dx, const HydroParams& hp, double* d_dt, hipStream_t s,
                            const int* dn, double* partials, double* sums) {
  if (!dn && noct <= 0) return hipSuccess;
  const unsigned nb = cell_grid(8L * noct, dn);
  hipLaunchKernelGGL(amr_begin, dim3(nb), dim3(256), 0, s, active, noct, dn, t, u, unew, ustride, dx, hp, d_dt,
                     partials);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess || !partials) return e;
  return launch_cons_finish(partials, nb, sums, 1, s);
}

size_t amr_begin_partials_bytes(int noct, bool persistent) {
  const size_t nb = persistent ? persist_blocks() : (size_t)((8L * noct + 255) / 256);
  return (nb > 0 ? nb : 1) * 3 * sizeof(double);
}

hipError_t launch_amr_finish(const int* active, int noct, const AmrTreeView& t, const double* unew, double* u,
                             long ustride, double smallr, int ivar_mode, hipStream_t s, const int* dn) {
  if (!dn && noct <= 0) return hipSuccess;
  hipLaunchKernelGGL(amr_finish, dim3(cell_grid(8L * noct, dn)), dim3(256), 0, s, active, noct, dn, t, unew, u,
                     ustride, smallr, ivar_mode);
  return hipGetLastError();
}

hipError_t launch_amr_dt(int op, double* dt, double* dtold, double* time, int l, int coarser, int nsub, int icount,
                         double start, hipStream_t s, const int* dn, long long* cnt) {
  hipLaunchKernelGGL(amr_dt_kernel, dim3(1), dim3(1), 0, s, op, dt, dtold, time, l, coarser, nsub, icount, start, dn,
                     cnt);
  return hipGetLastError();
}

hipError_t launch_amr_cons0(const ConsLevels& lv, const int* dn, const AmrTreeView& t, const double* u, long ustride,
                            int levelmin, int nlevelmax, double boxlen, int nv, const int* armed, double* cons0,
                            hipStream_t s) {
  hipLaunchKernelGGL(amr_cons0_kernel, dim3(1), dim3(kConsThreads), 0, s, lv, dn, t, u, ustride, levelmin, nlevelmax,
                     boxlen, nv, armed, cons0);
  return hipGetLastError();
}

hipError_t launch_clear_bits(int* p, int mask, hipStream_t s) {
  hipLaunchKernelGGL(clear_bits, dim3(1), dim3(1), 0, s, p, mask);
  return hipGetLastError();
}

hipError_t launch_set_ints(int* p, int n, int v, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(set_ints, dim3((n + 255) / 256), dim3(256), 0, s, p, n, v);
  return hipGetLastError();
}

}  // namespace ramses
