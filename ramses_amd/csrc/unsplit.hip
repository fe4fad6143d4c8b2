// Batched `unsplit` (hydro/umuscl.f90:22-171) for gfx950.
//
// Input and output keep the reference's Fortran layouts so that the kernel
// sits behind the reference's own call boundary:
//   uin   (nv,-1:4,-1:4,-1:4,5)   gravin (nv,-1:4,-1:4,-1:4,3)
//   flux  (nv,1:3,1:3,1:3,5,3)    tmp    (nv,1:3,1:3,1:3,2,3)
// repeated `ngroups` times back to back (one array set per oct batch).
//
// Work decomposition.  Per oct the reference computes fluxes on 36 faces:
// 12 per direction, arranged as 4 rows of 3 faces (x rows at j,k in {1,2},
// faces i=1..3; likewise y and z).  One thread owns one row of one oct; the
// 32 octs of a batch are 32 consecutive lanes, so every global load of a
// (cell, variable) pair is a 256-byte coalesced segment.  A row needs the
// traced states of its 4 cells (s = 0..3 along the row), each of which needs
// primitives of its 6 axial neighbours: 22 primitive evaluations per row,
// all in registers, no LDS and no scratch.  Stencil cells that no face of
// the oct depends on (the 6^3 corners) are never read.  Because every value
// is a deterministic function of its inputs, computing it once per row
// instead of once per 6^3 sweep gives the reference's values bit for bit.
//
// FP32-mixed with nvector 32 stages instead: a workgroup takes half a batch
// (16 octs x 12 rows), converts the 80 stencil cells its faces read to
// primitives once, into LDS, and its rows read them from there.  A row then
// evaluates no primitive itself, against 22 in the row kernel.
#include <type_traits>

#include "hydro_device.hpp"
#include "unsplit_args.hpp"
#include "unsplit_device.hpp"

namespace ramses {

// UnsplitArgs is declared in unsplit_args.hpp

namespace {

using unsplit_dev::unsplit_task;

// The dispatcher hands workgroup b to XCD b % 8, and each XCD has its own
// L2.  Workgroup b is given the logical index that makes every XCD run one
// contiguous range of logical workgroups (a bijection of [0, G)), so the
// workgroups that share one batch's stencil cells (a batch is 384 tasks,
// a workgroup 64..768) run on the same XCD and re-read those cells from
// its L2 rather than from HBM.
__device__ __forceinline__ unsigned xcd_contiguous(unsigned b, unsigned G) {
  constexpr unsigned X = 8;
  const unsigned x = b % X, q = G / X, r = G % X;
  return x * q + (x < r ? x : r) + b / X;
}

// One thread per (group, row, lane) task; G: a.grav is non-null.
template <typename R, int RS, bool S27, int ST, bool G>
__global__ __launch_bounds__(768) void unsplit_rows(UnsplitArgs a) {
  const long t = (long)xcd_contiguous(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
  if (t >= a.ngroups * 12L * a.nv) return;
  unsplit_task<R, RS, S27, false, ST, G>(a, t, a.ngroups, a.ngrid_last);
}

// The stencil cells some face of an oct depends on: a cell (i,j,k) in
// -1..4 is read when at most one coordinate lies outside 1..2, or when two
// do and both are 0 or 3 (a transverse neighbour of a row's end cell).
// That is 8 + 48 + 24 = 80 cells.  slot[cell6] numbers them in cell6 order
// (-1 for the unread cells) and cell[slot] inverts it.
struct StageMap {
  signed char slot[216];
  unsigned char cell[80];
};

constexpr StageMap make_stage_map() {
  StageMap m{};
  int n = 0;
  for (int c = 0; c < 216; ++c) {
    const int x[3] = {c % 6 - 1, (c / 6) % 6 - 1, c / 36 - 1};
    int out = 0, edge = 0;
    for (int a = 0; a < 3; ++a) {
      if (x[a] < 1 || x[a] > 2) ++out;
      if (x[a] == 0 || x[a] == 3) ++edge;
    }
    const bool read = out <= 1 || (out == 2 && edge == 2);
    m.slot[c] = read ? (signed char)n : (signed char)-1;
    if (read) m.cell[n++] = (unsigned char)c;
  }
  return m;
}

constexpr StageMap kStageHost = make_stage_map();
static_assert(kStageHost.slot[215] == -1 && kStageHost.cell[79] == 201, "80 stencil cells");
__constant__ StageMap kStage = make_stage_map();

constexpr int kStageLanes = 16;                  // octs per workgroup
constexpr int kStageCells = 80 * kStageLanes;    // (cell, lane) pairs staged
constexpr int kStageThreads = 12 * kStageLanes;  // one thread per (row, lane)

// Primitives of the staged cells, variable-major: [5][80][kStageLanes].
template <typename R>
struct StagedCtx {
  const R* pr;  // + lane
  int nv;
  __device__ __forceinline__ Prim<R> prim(int i, int j, int k) const {
    const R* q = pr + kStage.slot[unsplit_dev::cell6(i, j, k)] * kStageLanes;
    return Prim<R>{q[0], q[kStageCells], q[2 * kStageCells], q[3 * kStageCells], q[4 * kStageCells]};
  }
};

// Half a batch per workgroup (nv == 32, 7-point slopes): stage, then one
// row of faces per thread.  A workgroup's rows read only its own lanes.
template <typename R, int RS, int ST>
__global__ __launch_bounds__(kStageThreads) void unsplit_staged(UnsplitArgs a) {
  __shared__ R pr[5 * kStageCells];
  constexpr int nv = 32;
  const long gidx = blockIdx.x >> 1;
  const int l0 = (blockIdx.x & 1) * kStageLanes;
  const int t = threadIdx.x;
  const HydroParams hp = a.hp;
  const double dt = a.d_dt ? *a.d_dt : a.dt;
  const double dtxhalf = dt * 0.5;
  const double smalle = hp.smallc * hp.smallc / hp.gamma / (hp.gamma - 1.0);  // hydro/umuscl.f90:884
  const size_t vs = (size_t)nv * 216;
  const double* u = a.uin + gidx * vs * 5 + l0;
  const double* g = a.grav ? a.grav + gidx * vs * 3 + l0 : nullptr;
  for (int e = t; e < kStageCells; e += kStageThreads) {
    const int cell = e / kStageLanes, l = e - cell * kStageLanes;
    const size_t o = (size_t)kStage.cell[cell] * nv + l;
    double gx = 0.0, gy = 0.0, gz = 0.0;
    if (g) {
      gx = g[o];
      gy = g[o + vs];
      gz = g[o + 2 * vs];
    }
    const Prim<R> q = ctoprim_cell<R>(u[o], u[o + vs], u[o + 2 * vs], u[o + 3 * vs], u[o + 4 * vs], gx, gy, gz,
                                      dtxhalf, smalle, hp);
    pr[e] = q.r;
    pr[kStageCells + e] = q.u;
    pr[2 * kStageCells + e] = q.v;
    pr[3 * kStageCells + e] = q.w;
    pr[4 * kStageCells + e] = q.p;
  }
  __syncthreads();
  const int row = t / kStageLanes, l = t - row * kStageLanes;
  const int ngrid = (gidx == a.ngroups - 1) ? a.ngrid_last : nv;
  if (l0 + l >= ngrid) return;
  const int d = row >> 2, a1 = 1 + (row & 1), b1 = 1 + ((row >> 1) & 1);
  const StagedCtx<R> c{pr + l, nv};
  const SolverConsts<R> sc = make_consts<R>(hp);
  double* flux = a.flux + gidx * (size_t)nv * 27 * 15 + l0 + l;
  double* tmp = a.tmp + gidx * (size_t)nv * 27 * 6 + l0 + l;
  const size_t fvs = (size_t)nv * 27;
  using unsplit_dev::row_faces;
  if (d == 0)
    row_faces<R, 0, RS, false, false, ST>(c, sc, a1, b1, (R)(dt / a.dx[0]), flux, tmp, fvs, make_scale(dt, a.dx[0]),
                                      nullptr);
  else if (d == 1)
    row_faces<R, 1, RS, false, false, ST>(c, sc, a1, b1, (R)(dt / a.dx[1]), flux, tmp, fvs, make_scale(dt, a.dx[1]),
                                      nullptr);
  else
    row_faces<R, 2, RS, false, false, ST>(c, sc, a1, b1, (R)(dt / a.dx[2]), flux, tmp, fvs, make_scale(dt, a.dx[2]),
                                      nullptr);
}

}  // namespace

// Staging is used in FP32-mixed only.  On 2^20 octs it measured 1.27 ms
// against 1.32 ms for the row kernel (hllc and llf alike, the same bits);
// in FP64 it measured 2.25 ms against 1.72 ms: 51 KB of FP64 primitives per
// workgroup leave 2.25 waves per SIMD, and the FP64 kernel is not bound by
// its primitive evaluations (a staged workgroup of 8 or 32 octs was slower
// still in both precisions).
//
// Minmod, the reference's default slope (hydro/hydro_parameters.f90:76),
// has kernels with the limiter fixed at compile time (ST = 1); the other
// 7-point types take it at run time (ST = -1).  The row kernels without a
// gravity array (G = false) read no gravity and test no pointer per cell:
// 8% of the FP64 kernel's time on 2^20 octs.  The reference always passes
// a gravity array (zeros without Poisson), so minmod has a gravity instance
// too.
template <typename R, int RS, bool S27>
static void launch_t(const UnsplitArgs& a, long grid, int block, hipStream_t stream) {
  const bool st1 = !S27 && a.hp.slope_type == 1;
  if constexpr (!S27 && std::is_same<R, float>::value) {
    if (a.nv == 32) {
      const dim3 g((unsigned)(a.ngroups * 2)), b(kStageThreads);
      if (st1)
        hipLaunchKernelGGL((unsplit_staged<R, RS, 1>), g, b, 0, stream, a);
      else
        hipLaunchKernelGGL((unsplit_staged<R, RS, -1>), g, b, 0, stream, a);
      return;
    }
  }
  const dim3 g((unsigned)grid), b(block);
  if (a.grav) {
    if constexpr (!S27) {
      if (st1) {
        hipLaunchKernelGGL((unsplit_rows<R, RS, S27, 1, true>), g, b, 0, stream, a);
        return;
      }
    }
    hipLaunchKernelGGL((unsplit_rows<R, RS, S27, -1, true>), g, b, 0, stream, a);
    return;
  }
  if constexpr (!S27) {
    if (st1) {
      hipLaunchKernelGGL((unsplit_rows<R, RS, S27, 1, false>), g, b, 0, stream, a);
      return;
    }
  }
  hipLaunchKernelGGL((unsplit_rows<R, RS, S27, -1, false>), g, b, 0, stream, a);
}

template <typename R>
static void launch_r(const UnsplitArgs& a, long grid, int block, hipStream_t stream) {
  const bool s27 = a.hp.slope_type == 3;
  switch (a.hp.riemann) {
    case kHLLC: s27 ? launch_t<R, kHLLC, true>(a, grid, block, stream) : launch_t<R, kHLLC, false>(a, grid, block, stream); break;
    case kHLL: s27 ? launch_t<R, kHLL, true>(a, grid, block, stream) : launch_t<R, kHLL, false>(a, grid, block, stream); break;
    case kACOUSTIC:
    case kEXACT:
      s27 ? launch_t<R, kDynamicRiemann, true>(a, grid, block, stream)
          : launch_t<R, kDynamicRiemann, false>(a, grid, block, stream);
      break;
    default: s27 ? launch_t<R, kLLF, true>(a, grid, block, stream) : launch_t<R, kLLF, false>(a, grid, block, stream); break;
  }
}

hipError_t launch_unsplit(const UnsplitArgs& a, bool fp32, hipStream_t stream) {
  const long threads = a.ngroups * 12L * a.nv;
  if (threads <= 0) return hipSuccess;
  // Any multiple of 64 works with the flat (batch, row, lane) mapping.  384 =
  // 12 rows x 32 lanes puts exactly one nvector=32 batch in a workgroup, so
  // re-reads of its stencil cells stay in one XCD's L2; 256 spreads a batch
  // over two workgroups.  The autotuner picks per device.
  const int block = tuning().unsplit_block ? tuning().unsplit_block : 256;
  const long grid = (threads + block - 1) / block;
  if (fp32)
    launch_r<float>(a, grid, block, stream);
  else
    launch_r<double>(a, grid, block, stream);
  return hipGetLastError();
}

}  // namespace ramses
