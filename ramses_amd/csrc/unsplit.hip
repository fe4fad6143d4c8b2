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

// One thread per (group, row, lane) task (SH: partner rows share
// primitives across the wavefront, nvector 32 only).
template <typename R, int RS, bool S27, bool SH>
__global__ __launch_bounds__(768) void unsplit_rows(UnsplitArgs a) {
  const long t = (long)xcd_contiguous(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
  if (t >= a.ngroups * 12L * a.nv) return;
  unsplit_task<R, RS, S27, false, SH>(a, t, a.ngroups, a.ngrid_last);
}

}  // namespace

// Partner sharing is used in FP32 only: there it saves 4% (1.38 -> 1.33 ms
// on 2^20 octs); in FP64 it measured 1% slower (1.74 -> 1.76 ms) although
// it removes a quarter of the transverse primitive evaluations.
template <typename R, int RS, bool S27>
static void launch_t(const UnsplitArgs& a, long grid, int block, hipStream_t stream) {
  if constexpr (!S27 && std::is_same<R, float>::value) {
    if (a.nv == 32) {
      hipLaunchKernelGGL((unsplit_rows<R, RS, false, true>), dim3((unsigned)grid), dim3(block), 0, stream, a);
      return;
    }
  }
  hipLaunchKernelGGL((unsplit_rows<R, RS, S27, false>), dim3((unsigned)grid), dim3(block), 0, stream, a);
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
