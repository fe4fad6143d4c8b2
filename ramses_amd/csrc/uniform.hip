// Resident uniform periodic level: the godunov_fine -> godfine1 sweep
// (hydro/godunov_fine.f90:5-35, 486-911) with set_unew/set_uold replaced by
// ping-pong buffers, and the courant_fine dt reduction
// (hydro/courant_fine.f90:1-159) fused into the update epilogue.
//
// State: u[var][z][y][x], n^3 doubles per variable (SoA, x fastest).
//
// Kernel structure (2.5D blocking).  A workgroup owns an interior tile of
// TX x TY columns and marches along z through a chunk of TZ planes.  Its
// threads cover the tile and the ring of cells next to its four sides (the
// corners excluded), so the traced states of the halo cells that face the
// tile are computed by a thread like any other.  Per plane:
//   1. primitives of the next plane (tile + 2 cells of halo) -> LDS,
//   2. each thread computes slopes and traced states of its own cell from
//      its column (registers) and its x/y neighbours (LDS),
//   3. left states are exchanged through LDS and x/y face fluxes solved,
//   4. the z face below the cell is solved from the previous plane's traced
//      state kept in registers, completing the update of the plane below,
//   5. the new state is written and its cmpdt limit min-reduced (wavefront
//      shuffles, then LDS across waves, then one 64-bit atomic min on the
//      bit pattern of the positive double).
// Each primitive is computed by one ctoprim per cell per workgroup, each
// face flux once, and every value keeps the reference's operation order, so
// the FP64 path reproduces godfine1+unsplit on a uniform grid bit for bit.
#include <type_traits>

#include "hydro_device.hpp"

namespace ramses {

namespace {

struct UniformArgs {
  const double* u;
  double* unew;
  const double* d_dt;     // time step (device)
  double* d_dt_next;      // fused courant output or null
  double* partials;       // courant_fine sums per workgroup (3 doubles each) or null
  int n;
  double dx;
  HydroParams hp;
};

// Slot state layout (device doubles, see ramses_level_slot in the header).
enum SlotField : int { kTime = 0, kTend = 1, kDtUse = 2, kDtNext = 3, kDone = 4, kSteps = 5, kParity = 6, kStepped = 7 };

constexpr int kMaxSlots = 32;

struct BatchSlot {
  double* buf[2];
  double* state;
  int n;
  int tz;
  double dx;
};

struct BatchArgs {
  BatchSlot slot[kMaxSlots];
  int blk0[kMaxSlots + 1];
  int nslots;
  HydroParams hp;
};

__device__ __forceinline__ int wrapi(int a, int n) { return a < 0 ? a + n : (a >= n ? a - n : a); }

// Threads of a TX x TY tile: one per interior cell plus one per halo cell
// whose traced state a tile face needs (the 2*(TX+TY) non-corner cells of
// the one-cell ring), rounded up to whole wavefronts.
template <int TX, int TY>
constexpr int tile_threads() { return (TX * TY + 2 * (TX + TY) + 63) / 64 * 64; }

// One workgroup's tile of a level: columns [x0, x0+TX) x [y0, y0+TY),
// planes [z0, z0+TZ).  Shared by the single-level and the batched kernels.
// ST: the slope type fixed at compile time, or -1 (see slope1).
template <typename R, int TX, int TY, int RS, bool SUMS = false, int ST = -1>
__device__ __forceinline__ void level_tile(const double* __restrict__ u, double* __restrict__ unew, int n, double dx,
                                           double dt, double* d_dt_next, const HydroParams& hp, int x0, int y0,
                                           int z0, int TZ, double* partials = nullptr, long block = 0) {
  constexpr int EX = TX + 2, EY = TY + 2;      // tile with its one-cell ring
  constexpr int QX = TX + 4, QY = TY + 4;      // primitive plane with 2-cell halo
  constexpr int NQ = QX * QY;
  constexpr int NE = EX * EY;
  constexpr int NI = TX * TY;
  constexpr int NT = tile_threads<TX, TY>();
  static_assert(NI % 64 == 0, "interior cells fill whole wavefronts");

  __shared__ R qs[2][5][NQ];   // primitive planes (ring of 2)
  __shared__ __align__(16) R xch[10][NE];  // exchange: traced left states, then fluxes
  __shared__ double red[NT / 64];

  const long plane = (long)n * n;
  const long vstride = plane * n;
  const int t = threadIdx.x;
  // Threads [0, NI) own the interior cells, x fastest.  Threads [NI, NI +
  // 2*(TX+TY)) own the ring cells a tile face needs: the left column (its
  // qm_x), the right column (its qp_x and the face x = TX+1/2), the bottom
  // row (qm_y) and the top row (qp_y and the face y = TY+1/2).  Corner cells
  // feed no face and get no thread.  Since NI is a multiple of 64, the ring
  // cells fill waves of their own, which skip the z face, the update and the
  // courant epilogue.
  // (ex, ey) are coordinates in the tile with its ring, interior 1..TX, 1..TY.
  enum { kNone = -1, kInterior = 0, kLeft, kRight, kBottom, kTop };
  int ex = 0, ey = 0, kind = kNone;
  if (t < NI) {
    ex = 1 + t % TX;
    ey = 1 + t / TX;
    kind = kInterior;
  } else if (t < NI + 2 * (TX + TY)) {
    int h = t - NI;
    if (h < 2 * TY) {
      ex = h < TY ? 0 : TX + 1;
      ey = 1 + h % TY;
      kind = h < TY ? kLeft : kRight;
    } else {
      h -= 2 * TY;
      ex = 1 + h % TX;
      ey = h < TX ? 0 : TY + 1;
      kind = h < TX ? kBottom : kTop;
    }
  }
  const bool active = kind != kNone;
  const int e = ey * EX + ex;  // own slot in the exchange area
  const int xg = wrapi(x0 - 1 + ex, n), yg = wrapi(y0 - 1 + ey, n);
  const int qo = (ey + 1) * QX + (ex + 1);  // own cell in the primitive plane

  const SolverConsts<R> sc = make_consts<R>(hp);
  const double dtxhalf = dt * 0.5;
  const double smalle = hp.smallc * hp.smallc / hp.gamma / (hp.gamma - 1.0);
  const R dtdx = (R)(dt / dx);
  const FluxScale fs = make_scale(dt, dx);

  // cooperative ctoprim of plane zp (periodic) into slot s
  auto load_plane = [&](int zp, int s) {
    const int zg = wrapi(zp, n);
    static_assert(NQ <= 2 * NT, "a plane is at most two cells per thread");
    // cell t for every thread; the cells beyond NT go to the last threads
    // (the ring's wave, which solves fewer faces), in reverse order
    for (int k = 0; k < 2; ++k) {
      const int c = k == 0 ? t : 2 * NT - 1 - t;
      if (c >= NQ) continue;
      const int qx = c % QX, qy = c / QX;
      const int gx = wrapi(x0 - 2 + qx, n), gy = wrapi(y0 - 2 + qy, n);
      const long o = (long)zg * plane + (long)gy * n + gx;
      const Prim<R> q = ctoprim_cell<R>(u[o], u[o + vstride], u[o + 2 * vstride], u[o + 3 * vstride],
                                        u[o + 4 * vstride], 0.0, 0.0, 0.0, dtxhalf, smalle, hp);
      qs[s][0][c] = q.r;
      qs[s][1][c] = q.u;
      qs[s][2][c] = q.v;
      qs[s][3][c] = q.w;
      qs[s][4][c] = q.p;
    }
  };
  auto read_q = [&](int s, int c) {
    return Prim<R>{qs[s][0][c], qs[s][1][c], qs[s][2][c], qs[s][3][c], qs[s][4][c]};
  };

  // own-column primitives at p-1 and p
  Prim<R> qprev{}, qcur{};
  // values carried from the previous plane
  Face<R> fz_left{};          // rotated qm_z of plane p-1
  double fz_prev[5] = {0, 0, 0, 0, 0};  // flux through the z face below plane p-1 (scaled)
  double part_prev[5] = {0, 0, 0, 0, 0};  // (u + dFx) + dFy of plane p-1
  double wmax = 0.0;  // largest cmpdt wave speed of the new state
  ConsSums cs{0.0, 0.0, 0.0};  // courant_fine's sums of the new state (SUMS)
  const double vol = dx * dx * dx;

  // prologue: planes z0-2 and z0-1
  load_plane(z0 - 2, 0);
  __syncthreads();
  if (active) qprev = read_q(0, qo);
  __syncthreads();
  load_plane(z0 - 1, 1);
  __syncthreads();
  if (active) qcur = read_q(1, qo);

  // One plane of the march.  kWaves: 0 for a wave of interior cells, 1 for
  // the ring's wave; the cell kinds of a wave are then known at compile time
  // (the cells of a wave are all interior or all ring, NI being a multiple
  // of 64).
  const bool active_run = active;
  const int kind_run = kind;
  auto plane_step = [&](auto waves_tag, int p) {
    constexpr int kWaves = decltype(waves_tag)::value;
    const bool interior = kWaves == 0;
    const bool active = kWaves == 0 ? true : active_run;
    const int kind = kWaves == 0 ? (int)kInterior : kind_run;
    const int scur = (p - z0 + 2) & 1;  // slot of plane p (z0-2 -> 0, z0-1 -> 1)
    const int snext = scur ^ 1;
    const bool faces_xy = (p >= z0) && (p < z0 + TZ);
    // FP32-mixed: the own cell's conserved state for the x/y update is
    // loaded before the next plane, so its latency hides behind the plane's
    // work (in FP64 the 10 more live registers would spill)
    double upre[5] = {0, 0, 0, 0, 0};
    if constexpr (sizeof(R) == 4) {
      if (interior && faces_xy) {
        const long o = (long)wrapi(p, n) * plane + (long)yg * n + xg;
#pragma unroll
        for (int v = 0; v < 5; ++v) upre[v] = u[o + v * vstride];
      }
    }
    __syncthreads();
    load_plane(p + 1, snext);
    __syncthreads();
    Prim<R> qnext{};
    Prim<R> qp_x{}, qm_x{}, qp_y{}, qm_y{}, qp_z{}, qm_z{};
    if (active) {
      qnext = read_q(snext, qo);
      const Prim<R> xm = read_q(scur, qo - 1), xp = read_q(scur, qo + 1);
      const Prim<R> ym = read_q(scur, qo - QX), yp = read_q(scur, qo + QX);
      const Prim<R> dqx = slope_prim<R, ST>(xm, qcur, xp, sc.slope_type, sc.theta);
      const Prim<R> dqy = slope_prim<R, ST>(ym, qcur, yp, sc.slope_type, sc.theta);
      const Prim<R> dqz = slope_prim<R, ST>(qprev, qcur, qnext, sc.slope_type, sc.theta);
      const Source<R> src = trace_sources(qcur, dqx, dqy, dqz, sc.gamma);
      if constexpr (kWaves == 0) {
        trace_dir(qcur, dqx, src, dtdx, sc.smallr, qp_x, qm_x);
        trace_dir(qcur, dqy, src, dtdx, sc.smallr, qp_y, qm_y);
        trace_dir(qcur, dqz, src, dtdx, sc.smallr, qp_z, qm_z);
      } else {
        // a left or right ring cell feeds only an x face, a bottom or top one
        // only a y face, and none a z face
        const bool xdir = kind == kLeft || kind == kRight;
        trace_dir(qcur, xdir ? dqx : dqy, src, dtdx, sc.smallr, qp_x, qm_x);
        qp_y = qp_x;
        qm_y = qm_x;
      }
    }
    double part[5] = {0, 0, 0, 0, 0};
    if (faces_xy) {
      if (active) {
        xch[0][e] = qm_x.r; xch[1][e] = qm_x.u; xch[2][e] = qm_x.v; xch[3][e] = qm_x.w; xch[4][e] = qm_x.p;
        xch[5][e] = qm_y.r; xch[6][e] = qm_y.u; xch[7][e] = qm_y.v; xch[8][e] = qm_y.w; xch[9][e] = qm_y.p;
      }
      __syncthreads();
      double fxs[5] = {0, 0, 0, 0, 0}, fys[5] = {0, 0, 0, 0, 0};
      if (interior || kind == kRight) {  // x face at x-1/2
        const int lo = e - 1;
        const Prim<R> l{xch[0][lo], xch[1][lo], xch[2][lo], xch[3][lo], xch[4][lo]};
        const Flux6<R> g = riemann_fixed<R, RS>(rotate(l, 0), rotate(qp_x, 0), sc);
        R fc[5];
        unrotate(g, 0, fc);
#pragma unroll
        for (int v = 0; v < 5; ++v) fxs[v] = scale_flux(fc[v], fs);  // hydro/umuscl.f90:106
      }
      if (interior || kind == kTop) {  // y face at y-1/2
        const int lo = e - EX;
        const Prim<R> l{xch[5][lo], xch[6][lo], xch[7][lo], xch[8][lo], xch[9][lo]};
        const Flux6<R> g = riemann_fixed<R, RS>(rotate(l, 1), rotate(qp_y, 1), sc);
        R fc[5];
        unrotate(g, 1, fc);
#pragma unroll
        for (int v = 0; v < 5; ++v) fys[v] = scale_flux(fc[v], fs);  // :129
      }
      __syncthreads();
      // The fluxes, in double, go through the exchange area reinterpreted:
      // in FP64 it holds the x and the y fluxes side by side (one exchange),
      // in FP32 one direction at a time.
      double* fx_sh = reinterpret_cast<double*>(&xch[0][0]);
      constexpr int FSTRIDE = NE;  // doubles per variable
      static_assert(sizeof(R) * 10 * NE >= sizeof(double) * 5 * NE, "exchange area too small");
      if constexpr (sizeof(R) == 8) {
        double* fy_sh = fx_sh + 5 * FSTRIDE;
        if (interior || kind == kRight) {
#pragma unroll
          for (int v = 0; v < 5; ++v) fx_sh[v * FSTRIDE + e] = fxs[v];
        }
        if (interior || kind == kTop) {
#pragma unroll
          for (int v = 0; v < 5; ++v) fy_sh[v * FSTRIDE + e] = fys[v];
        }
        __syncthreads();
        if (interior) {
          // hydro/godunov_fine.f90:770-772: unew = uold + (flux(i) - flux(i+1)), idim=1 then 2
          const long o = (long)wrapi(p, n) * plane + (long)yg * n + xg;
#pragma unroll
          for (int v = 0; v < 5; ++v)
            part[v] = u[o + v * vstride] + (fx_sh[v * FSTRIDE + e] - fx_sh[v * FSTRIDE + e + 1]);
#pragma unroll
          for (int v = 0; v < 5; ++v) part[v] = part[v] + (fy_sh[v * FSTRIDE + e] - fy_sh[v * FSTRIDE + e + EX]);
        }
      } else {
        if (interior || kind == kRight) {
#pragma unroll
          for (int v = 0; v < 5; ++v) fx_sh[v * FSTRIDE + e] = fxs[v];
        }
        __syncthreads();
        if (interior) {
          // hydro/godunov_fine.f90:770-772, idim=1: unew = uold + (flux(i) - flux(i+1))
#pragma unroll
          for (int v = 0; v < 5; ++v) part[v] = upre[v] + (fx_sh[v * FSTRIDE + e] - fx_sh[v * FSTRIDE + e + 1]);
        }
        __syncthreads();
        if (interior || kind == kTop) {
#pragma unroll
          for (int v = 0; v < 5; ++v) fx_sh[v * FSTRIDE + e] = fys[v];
        }
        __syncthreads();
        if (interior) {
#pragma unroll
          for (int v = 0; v < 5; ++v) part[v] = part[v] + (fx_sh[v * FSTRIDE + e] - fx_sh[v * FSTRIDE + e + EX]);  // idim=2
        }
      }
    }
    // z face below plane p (between p-1 and p)
    double fz[5] = {0, 0, 0, 0, 0};
    if (interior && p >= z0) {
      const Flux6<R> g = riemann_fixed<R, RS>(fz_left, rotate(qp_z, 2), sc);
      R fc[5];
      unrotate(g, 2, fc);
#pragma unroll
      for (int v = 0; v < 5; ++v) fz[v] = scale_flux(fc[v], fs);  // :153
    }
    // plane p-1 is complete: hydro/godunov_fine.f90:751-792 in idim order x, y, z
    if (interior && p >= z0 + 1) {
      const int zg = wrapi(p - 1, n);
      const long o = (long)zg * plane + (long)yg * n + xg;
      double un[5];
#pragma unroll
      for (int v = 0; v < 5; ++v) {
        const double s = part_prev[v] + (fz_prev[v] - fz[v]);  // idim=3
        un[v] = s;
        unew[o + v * vstride] = s;
      }
      if (d_dt_next) {
        if constexpr (sizeof(R) == 8) wmax = fmax_f(wmax, cmpdt_w(un[0], un[1], un[2], un[3], un[4], hp));
        else wmax = fmax_f(wmax, cmpdt_w_fast(un[0], un[1], un[2], un[3], un[4], hp));
      }
      if constexpr (SUMS) cons_add(cs, un[0], un[1], un[2], un[3], un[4], vol, hp.smallr);
    }
    // carry to the next plane
    fz_left = rotate(qm_z, 2);
#pragma unroll
    for (int v = 0; v < 5; ++v) {
      fz_prev[v] = fz[v];
      part_prev[v] = part[v];
    }
    qprev = qcur;
    qcur = qnext;
  };
  // The interior waves and the ring's wave run copies of the march with
  // their cell kinds fixed, both meeting the same barriers in the same order
  // (at 512^3: FP64 5.5% faster, FP32-mixed 3.5% once held to 128 VGPRs,
  // where it takes 117).  A further copy for the planes between the first
  // two and the last, with their tests folded, spilled 48 B in FP64 and ran
  // 9% slower.
  using kInteriorWaves = std::integral_constant<int, 0>;
  using kRingWave = std::integral_constant<int, 1>;
  if (t >= NI) {
    for (int p = z0 - 1; p <= z0 + TZ; ++p) plane_step(kRingWave{}, p);
  } else {
    for (int p = z0 - 1; p <= z0 + TZ; ++p) plane_step(kInteriorWaves{}, p);
  }

  if (d_dt_next) {
    // wavefront max of w with shuffles, then across the block's waves; the
    // block's dt limit is cmpdt_from_w(max w), min-combined across blocks
    for (int off = 32; off > 0; off >>= 1) wmax = fmax_f(wmax, __shfl_xor(wmax, off, 64));
    const int wave = t >> 6, lane = t & 63;
    if (lane == 0) red[wave] = wmax;
    __syncthreads();
    if (t == 0) {
      double m = red[0];
      for (int w = 1; w < NT / 64; ++w) m = fmax_f(m, red[w]);
      if (m > 0.0) atomicMin(reinterpret_cast<unsigned long long*>(d_dt_next), dt_bits(cmpdt_from_w(m, dx, hp)));
    }
  }
  if constexpr (SUMS) {
    __shared__ double cred[3 * (NT / 64)];
    cons_block_store(cs, partials, block, cred);
  }
}


template <typename R, int TX, int TY, int RS, bool SUMS, int ST>
// The register budget: 3 waves per SIMD in FP64 (166 VGPRs), 4 in FP32-mixed
// (at most 128).
__global__ __launch_bounds__((tile_threads<TX, TY>()))
__attribute__((amdgpu_waves_per_eu(sizeof(R) == 4 ? 4 : 3, 8))) void uniform_godunov_kernel(UniformArgs a, int TZ,
                                                                                           int y0, int boff) {
  // XCD-aware tile order: the hardware deals workgroups round-robin to the
  // 8 XCDs, so workgroup b runs on XCD b%8.  Giving each XCD a contiguous
  // range of tiles (x fastest, then y, then z chunk) keeps the tiles that run
  // side by side on one XCD spatial neighbours, and their shared halo
  // columns are served by that XCD's L2 instead of HBM.  The launch covers
  // rows [y0, y0 + TY * gridDim.y); its workgroups' courant sums go to
  // partials[boff + block].
  const int ntx = gridDim.x, nty = gridDim.y;
  const int nb = ntx * nty * gridDim.z;
  const int b = blockIdx.x + ntx * (blockIdx.y + nty * blockIdx.z);
  const int per_xcd = (nb + 7) / 8;
  int logical = (b & 7) * per_xcd + (b >> 3);
  if ((nb & 7) != 0) logical = b;  // ragged grids keep the plain order
  const int tx = logical % ntx, ty = (logical / ntx) % nty, tzc = logical / (ntx * nty);
  level_tile<R, TX, TY, RS, SUMS, ST>(a.u, a.unew, a.n, a.dx, *a.d_dt, a.d_dt_next, a.hp, tx * TX, y0 + ty * TY,
                                  tzc * TZ, TZ, a.partials, boff + b);
}

// Many independent levels advanced by one launch (continuous batching).
// Workgroups [blk0[s], blk0[s+1]) belong to slot s; a slot whose state says
// `done` is skipped.  The slot's buffer parity selects input and output.
template <typename R, int TX, int TY, int RS, int ST>
__global__ __launch_bounds__((tile_threads<TX, TY>()))
__attribute__((amdgpu_waves_per_eu(sizeof(R) == 4 ? 4 : 3, 8))) void batch_godunov_kernel(BatchArgs a) {
  const int b = blockIdx.x;
  int s = 0;
  while (s + 1 < a.nslots && b >= a.blk0[s + 1]) ++s;
  const BatchSlot& sl = a.slot[s];
  double* st = sl.state;
  if (st[kDone] != 0.0) return;
  const int par = (int)st[kParity];
  const double* u = sl.buf[par];
  double* unew = sl.buf[par ^ 1];
  const int n = sl.n;
  const int ntx = n / TX, nty = n / TY;
  const int local = b - a.blk0[s];
  const int tx = local % ntx, ty = (local / ntx) % nty, tzc = local / (ntx * nty);
  level_tile<R, TX, TY, RS, false, ST>(u, unew, n, sl.dx, st[kDtUse], st + kDtNext, a.hp, tx * TX, ty * TY,
                                      tzc * sl.tz, sl.tz);
}

// Per-slot bookkeeping before each batched step (one thread per slot):
// finish the previous step (t += dt, flip the buffer parity), stop at t_end,
// and choose this step's dt = min(courant dt, t_end - t); then reset the
// fused courant accumulator to newdt_fine's start value
// (pm/newdt_fine.f90:51) capped by cmpdt's C*dx/smallc
// (hydro/godunov_utils.f90:114).  The end-time cap is this server's own
// contract (a request asks for the state at t_end) and has no counterpart in
// the reference: newdt_fine (pm/newdt_fine.f90:51-173) never looks at the
// output times, and amr_step writes an output once t >= tout
// (amr/amr_step.f90:151), after the step that crossed it.  Every other step
// uses the reference's courant dt unchanged.
__global__ void batch_advance_kernel(BatchArgs a) {
  const int s = threadIdx.x;
  if (s >= a.nslots) return;
  const BatchSlot& sl = a.slot[s];
  double* st = sl.state;
  if (st[kDone] != 0.0) return;
  if (st[kStepped] != 0.0) {
    st[kTime] = st[kTime] + st[kDtUse];
    st[kSteps] = st[kSteps] + 1.0;
    st[kParity] = 1.0 - st[kParity];
    st[kStepped] = 0.0;
  }
  if (st[kTime] >= st[kTend]) {
    st[kDone] = 1.0;
    return;
  }
  st[kDtUse] = fmin_f(st[kDtNext], st[kTend] - st[kTime]);
  const double boxlen = sl.n * sl.dx;
  st[kDtNext] = fmin_f(boxlen / a.hp.smallc, a.hp.courant_factor * sl.dx / a.hp.smallc);
  st[kStepped] = 1.0;
}

// courant_fine over a uniform level without gravity: max of the cmpdt wave
// speed per block, then the block's dt limit min-combined across blocks
// (exactly min over cells of cmpdt, see cmpdt_w).
__global__ void courant_kernel(const double* u, long ncell, double dx, HydroParams hp, double* d_dt,
                               double* partials) {
  double wmax = 0.0;
  ConsSums cs{0.0, 0.0, 0.0};
  const double vol = dx * dx * dx;
  for (long c = (long)blockIdx.x * blockDim.x + threadIdx.x; c < ncell; c += (long)gridDim.x * blockDim.x) {
    wmax = fmax_f(wmax, cmpdt_w(u[c], u[c + ncell], u[c + 2 * ncell], u[c + 3 * ncell], u[c + 4 * ncell], hp));
    if (partials) cons_add(cs, u[c], u[c + ncell], u[c + 2 * ncell], u[c + 3 * ncell], u[c + 4 * ncell], vol, hp.smallr);
  }
  if (partials) {
    __shared__ double cred[3 * 4];
    cons_block_store(cs, partials, blockIdx.x, cred);
  }
  for (int off = 32; off > 0; off >>= 1) wmax = fmax_f(wmax, __shfl_xor(wmax, off, 64));
  __shared__ double red[4];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  if (lane == 0) red[wave] = wmax;
  __syncthreads();
  if (threadIdx.x == 0) {
    double m = red[0];
    for (int w = 1; w < (int)(blockDim.x >> 6); ++w) m = fmax_f(m, red[w]);
    if (m > 0.0) atomicMin(reinterpret_cast<unsigned long long*>(d_dt), dt_bits(cmpdt_from_w(m, dx, hp)));
  }
}

__global__ void set_double(double* p, double v) { *p = v; }

__global__ void set_doubles(double* p, int n, double v) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

// Sum of the per-workgroup conservation sums in block order: thread t adds
// blocks t, t+256, ... in order, then a fixed pairwise tree; out[0..2] is
// overwritten, or incremented when `accumulate` (courant_fine's
// mass_tot = mass_tot + mass_all, hydro/courant_fine.f90:152-154).
__global__ __launch_bounds__(256) void cons_finish(const double* partials, long nblocks, double* out,
                                                  int accumulate) {
  __shared__ double sh[3][256];
  const int t = threadIdx.x;
  double m = 0.0, e = 0.0, ei = 0.0;
  for (long b = t; b < nblocks; b += 256) {
    m = m + partials[3 * b];
    e = e + partials[3 * b + 1];
    ei = ei + partials[3 * b + 2];
  }
  sh[0][t] = m;
  sh[1][t] = e;
  sh[2][t] = ei;
  __syncthreads();
  for (int h = 128; h > 0; h >>= 1) {
    if (t < h) {
      sh[0][t] = sh[0][t] + sh[0][t + h];
      sh[1][t] = sh[1][t] + sh[1][t + h];
      sh[2][t] = sh[2][t] + sh[2][t + h];
    }
    __syncthreads();
  }
  if (t == 0) {
    for (int k = 0; k < 3; ++k) out[k] = accumulate ? out[k] + sh[k][0] : sh[k][0];
  }
}

__global__ void init_slot_kernel(double* st, double t_end, double start) {
  st[kTime] = 0.0;
  st[kTend] = t_end;
  st[kDtUse] = 0.0;
  st[kDtNext] = start;
  st[kDone] = 0.0;
  st[kSteps] = 0.0;
  st[kParity] = 0.0;
  st[kStepped] = 0.0;
}

__global__ void cmpdt_group(const double* uu, const double* gg, int nv, int ncell, double dx, HydroParams hp,
                            double* d_dt) {
  // hydro/godunov_utils.f90:114-118: dt = min(C*dx/smallc, min over cells)
  double dt = hp.courant_factor * dx / hp.smallc;
  for (int k = threadIdx.x; k < ncell; k += blockDim.x) {
    const double gx = gg ? gg[k] : 0.0, gy = gg ? gg[k + nv] : 0.0, gz = gg ? gg[k + 2 * nv] : 0.0;
    dt = fmin_f(dt, cmpdt_cell(uu[k], uu[k + nv], uu[k + 2 * nv], uu[k + 3 * nv], uu[k + 4 * nv], gx, gy, gz, dx, hp));
  }
  for (int off = 32; off > 0; off >>= 1) dt = fmin_f(dt, __shfl_xor(dt, off, 64));
  __shared__ double red[4];
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = dt;
  __syncthreads();
  if (threadIdx.x == 0) {
    double m = red[0];
    for (int w = 1; w < (int)(blockDim.x >> 6); ++w) m = fmin_f(m, red[w]);
    *d_dt = m;
  }
}

// Tiles of rows [y0, y0 + ny), ny a multiple of TY; returns the number of
// workgroups launched.
template <typename R, int TX, int TY, bool SUMS, int ST>
long launch_band_st(const UniformArgs& a, int TZ, int y0, int ny, int boff, hipStream_t s) {
  constexpr int NT = tile_threads<TX, TY>();
  const dim3 grid(a.n / TX, ny / TY, a.n / TZ);
  switch (a.hp.riemann) {
    case kHLLC:
      hipLaunchKernelGGL((uniform_godunov_kernel<R, TX, TY, kHLLC, SUMS, ST>), grid, dim3(NT), 0, s, a, TZ, y0, boff);
      break;
    case kHLL:
      hipLaunchKernelGGL((uniform_godunov_kernel<R, TX, TY, kHLL, SUMS, ST>), grid, dim3(NT), 0, s, a, TZ, y0, boff);
      break;
    case kACOUSTIC:
    case kEXACT:
      hipLaunchKernelGGL((uniform_godunov_kernel<R, TX, TY, kDynamicRiemann, SUMS, ST>), grid, dim3(NT), 0, s, a, TZ,
                         y0, boff);
      break;
    default:
      hipLaunchKernelGGL((uniform_godunov_kernel<R, TX, TY, kLLF, SUMS, ST>), grid, dim3(NT), 0, s, a, TZ, y0, boff);
      break;
  }
  return (long)grid.x * grid.y * grid.z;
}

// Minmod, the reference's default slope (hydro/hydro_parameters.f90:76),
// has kernels of its own with the limiter fixed at compile time: without
// the branches on the slope type the compiler schedules the traces across
// what were separate blocks, which measured 8% faster at 512^3 in both
// precisions.  Other slope types take it at run time.
template <typename R, int TX, int TY, bool SUMS>
long launch_band(const UniformArgs& a, int TZ, int y0, int ny, int boff, hipStream_t s) {
  return a.hp.slope_type == 1 ? launch_band_st<R, TX, TY, SUMS, 1>(a, TZ, y0, ny, boff, s)
                              : launch_band_st<R, TX, TY, SUMS, -1>(a, TZ, y0, ny, boff, s);
}

// Rows [0, ny12) in 16x12 tiles, the rest in 16x8 tiles: ny12 is the
// largest multiple of 12 that leaves a multiple of 8, or -1 if there is
// none (n must be a multiple of 16).
inline int rows_in_16x12(int n) {
  for (int ny = n / 12 * 12; ny >= 12; ny -= 12)
    if ((n - ny) % 8 == 0) return ny;
  return -1;
}

// Launch the level kernel (16x12 tiles: two bands, see rows_in_16x12); with
// sums, then add the workgroups' conservation sums into d_sums[0..2] in
// block order.
template <typename R, int TX, int TY, bool SUMS>
hipError_t launch_tile_s(const UniformArgs& a, int TZ, double* d_sums, hipStream_t s) {
  long nblocks;
  if constexpr (TY == 12) {
    const int ny = rows_in_16x12(a.n);
    nblocks = launch_band<R, 16, 12, SUMS>(a, TZ, 0, ny, 0, s);
    if (ny < a.n) nblocks += launch_band<R, 16, 8, SUMS>(a, TZ, ny, a.n - ny, (int)nblocks, s);
  } else {
    nblocks = launch_band<R, TX, TY, SUMS>(a, TZ, 0, a.n, 0, s);
  }
  if constexpr (SUMS)
    hipLaunchKernelGGL(cons_finish, dim3(1), dim3(256), 0, s, (const double*)a.partials, nblocks, d_sums, 0);
  return hipGetLastError();
}

template <typename R, int TX, int TY>
hipError_t launch_tile(const UniformArgs& a, int TZ, double* d_sums, hipStream_t s) {
  return a.partials ? launch_tile_s<R, TX, TY, true>(a, TZ, d_sums, s) : launch_tile_s<R, TX, TY, false>(a, TZ, d_sums, s);
}

template <typename R, int RS>
void launch_batch_t(const BatchArgs& a, int nblocks, hipStream_t s) {
  constexpr int TX = 16, TY = 8, NT = tile_threads<TX, TY>();
  if (a.hp.slope_type == 1)  // minmod: see launch_band
    hipLaunchKernelGGL((batch_godunov_kernel<R, TX, TY, RS, 1>), dim3(nblocks), dim3(NT), 0, s, a);
  else
    hipLaunchKernelGGL((batch_godunov_kernel<R, TX, TY, RS, -1>), dim3(nblocks), dim3(NT), 0, s, a);
}

}  // namespace

// Continuous batching: one bookkeeping launch and one update launch advance
// every slot of `slots` by one step of its own dt.
hipError_t launch_level_batch(double* const* buf0, double* const* buf1, double* const* state, const int* n,
                              const double* dx, int nslots, const HydroParams& hp, bool fp32, hipStream_t s) {
  BatchArgs a{};
  a.nslots = nslots;
  a.hp = hp;
  int blocks = 0;
  for (int i = 0; i < nslots; ++i) {
    BatchSlot& sl = a.slot[i];
    sl.buf[0] = buf0[i];
    sl.buf[1] = buf1[i];
    sl.state = state[i];
    sl.n = n[i];
    sl.dx = dx[i];
    sl.tz = batch_tz(n[i]);
    a.blk0[i] = blocks;
    blocks += (n[i] / 16) * (n[i] / 8) * (n[i] / sl.tz);
  }
  a.blk0[nslots] = blocks;
  hipLaunchKernelGGL(batch_advance_kernel, dim3(1), dim3(64), 0, s, a);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess || blocks == 0) return e;
  if (fp32) {
    if (hp.riemann == kHLLC) launch_batch_t<float, kHLLC>(a, blocks, s);
    else if (hp.riemann == kHLL) launch_batch_t<float, kHLL>(a, blocks, s);
    else if (hp.riemann == kACOUSTIC || hp.riemann == kEXACT) launch_batch_t<float, kDynamicRiemann>(a, blocks, s);
    else launch_batch_t<float, kLLF>(a, blocks, s);
  } else {
    if (hp.riemann == kHLLC) launch_batch_t<double, kHLLC>(a, blocks, s);
    else if (hp.riemann == kHLL) launch_batch_t<double, kHLL>(a, blocks, s);
    else if (hp.riemann == kACOUSTIC || hp.riemann == kEXACT) launch_batch_t<double, kDynamicRiemann>(a, blocks, s);
    else launch_batch_t<double, kLLF>(a, blocks, s);
  }
  return hipGetLastError();
}

hipError_t launch_init_slot(double* state, double t_end, double start, hipStream_t s) {
  hipLaunchKernelGGL(init_slot_kernel, dim3(1), dim3(1), 0, s, state, t_end, start);
  return hipGetLastError();
}

hipError_t launch_cmpdt(const double* uu, const double* gg, int nv, int ncell, double dx, const HydroParams& hp,
                        double* d_dt, hipStream_t stream) {
  hipLaunchKernelGGL(cmpdt_group, dim3(1), dim3(256), 0, stream, uu, gg, nv, ncell, dx, hp, d_dt);
  return hipGetLastError();
}

hipError_t launch_cons_finish(const double* partials, long nblocks, double* out, int accumulate, hipStream_t s) {
  hipLaunchKernelGGL(cons_finish, dim3(1), dim3(256), 0, s, partials, nblocks, out, accumulate);
  return hipGetLastError();
}

hipError_t launch_set_doubles(double* p, int n, double v, hipStream_t s) {
  hipLaunchKernelGGL(set_doubles, dim3((n + 63) / 64), dim3(64), 0, s, p, n, v);
  return hipGetLastError();
}

hipError_t launch_set_double(double* p, double v, hipStream_t s) {
  hipLaunchKernelGGL(set_double, dim3(1), dim3(1), 0, s, p, v);
  return hipGetLastError();
}

hipError_t launch_courant(const double* u, int n, double dx, const HydroParams& hp, double* d_dt, hipStream_t s,
                          double* partials, double* d_sums) {
  const long ncell = (long)n * n * n;
  long blocks = (ncell + 255) / 256;
  if (blocks > 4096) blocks = 4096;
  hipLaunchKernelGGL(courant_kernel, dim3((unsigned)blocks), dim3(256), 0, s, u, ncell, dx, hp, d_dt, partials);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess || !partials) return e;
  hipLaunchKernelGGL(cons_finish, dim3(1), dim3(256), 0, s, (const double*)partials, blocks, d_sums, 0);
  return hipGetLastError();
}

// Workgroups of the largest level launch (8x8 tiles of 4 planes; the
// courant reduction uses at most as many), 3 doubles of sums each.
size_t uniform_sums_work_bytes(int n) {
  const long blocks = (long)n * n * n / 256;
  return (size_t)(blocks > 1 ? blocks : 1) * 3 * sizeof(double);
}

// Tile choice: the launch-shape knob when it fits n, else 16x12 where n
// allows it (three interior waves and one ring wave: the ring's work is a
// third of a tile's, against a half for 16x8), else the largest tile
// dividing n; and a z chunk that leaves enough workgroups to fill the chip.
hipError_t launch_uniform_godunov(const double* u, double* unew, int n, double dx, const double* d_dt,
                                  double* d_dt_next, const HydroParams& hp, bool fp32, hipStream_t s,
                                  double* partials, double* d_sums) {
  UniformArgs a{u, unew, d_dt, d_dt_next, partials, n, dx, hp};
  int TX, TY;
  const int tile = tuning().level_tile;
  const bool fits12 = n % 16 == 0 && rows_in_16x12(n) > 0;
  if (tile == 1 && n % 32 == 0) { TX = 32; TY = 8; }
  else if (tile == 2 && n % 16 == 0) { TX = 16; TY = 8; }
  else if (tile == 3) { TX = 8; TY = 8; }
  else if (tile == 4 && n % 32 == 0) { TX = 32; TY = 16; }
  else if (fits12) { TX = 16; TY = 12; }  // tile 5, and the default
  else if (n % 32 == 0) { TX = 32; TY = 8; }
  else if (n % 16 == 0) { TX = 16; TY = 8; }
  else { TX = 8; TY = 8; }
  const long tiles = (long)(n / TX) * (n / TY);
  int TZ = n;
  while (TZ > 4 && TZ % 2 == 0 && tiles * (n / TZ) < 2048) TZ /= 2;
  const int tz = tuning().level_tz;
  if (tz >= 4 && tz <= n && n % tz == 0) TZ = tz;
  if (fp32) {
    if (TY == 12) return launch_tile<float, 16, 12>(a, TZ, d_sums, s);
    if (TX == 32 && TY == 16) return launch_tile<float, 32, 16>(a, TZ, d_sums, s);
    if (TX == 32) return launch_tile<float, 32, 8>(a, TZ, d_sums, s);
    if (TX == 16) return launch_tile<float, 16, 8>(a, TZ, d_sums, s);
    return launch_tile<float, 8, 8>(a, TZ, d_sums, s);
  }
  if (TY == 12) return launch_tile<double, 16, 12>(a, TZ, d_sums, s);
  if (TX == 32 && TY == 16) return launch_tile<double, 32, 16>(a, TZ, d_sums, s);
  if (TX == 32) return launch_tile<double, 32, 8>(a, TZ, d_sums, s);
  if (TX == 16) return launch_tile<double, 16, 8>(a, TZ, d_sums, s);
  return launch_tile<double, 8, 8>(a, TZ, d_sums, s);
}

}  // namespace ramses
