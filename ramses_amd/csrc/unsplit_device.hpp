// Device code of the batched `unsplit` (hydro/umuscl.f90:22-171), shared by
// unsplit.hip (the batched boundary kernel) and amr.hip (the fused
// oct-tree level kernel).  See unsplit.hip for the work decomposition.
#pragma once

#include "hydro_device.hpp"
#include "unsplit_args.hpp"

namespace ramses {
namespace unsplit_dev {

// Fortran stencil index (i,j,k in -1..4) -> cell offset in units of nv.
__device__ __forceinline__ int cell6(int i, int j, int k) { return (i + 1) + 6 * ((j + 1) + 6 * (k + 1)); }

// G: whether the call may have a gravity array.  The kernels without one
// (G = false) carry no test of g, which frees registers (FP64 minmod rows:
// 126 VGPRs against 162, four waves per SIMD instead of three); zero
// acceleration enters ctoprim_cell either way, the same operations on the
// same values.  With G the test stays: loading g unconditionally made the
// FP64 kernels spill.
template <typename R, bool G>
struct RowCtx {
  const double* u;     // group base + lane
  const double* g;     // group gravity base + lane, or null
  size_t vs;           // variable stride = 216*nv
  int nv;
  double dtxhalf, smalle;
  const HydroParams* hp;

  __device__ __forceinline__ Prim<R> prim(int i, int j, int k) const {
    const size_t o = (size_t)cell6(i, j, k) * nv;
    const double u0 = u[o], u1 = u[o + vs], u2 = u[o + 2 * vs], u3 = u[o + 3 * vs], u4 = u[o + 4 * vs];
    double gx = 0.0, gy = 0.0, gz = 0.0;
    if (G && g) {
      gx = g[o];
      gy = g[o + vs];
      gz = g[o + 2 * vs];
    }
    return ctoprim_cell<R>(u0, u1, u2, u3, u4, gx, gy, gz, dtxhalf, smalle, *hp);
  }
};

// Cell coordinates of position s along a row of direction d at transverse
// coordinates (a, b).
__device__ __forceinline__ void row_cell(int d, int s, int a, int b, int& i, int& j, int& k) {
  if (d == 0) { i = s; j = a; k = b; }
  else if (d == 1) { i = a; j = s; k = b; }
  else { i = a; j = b; k = s; }
}

// Positivity-preserving 27-point slope (hydro/umuscl.f90:1328-1392) of the
// cell (i,j,k) for all five variables.
template <typename R, typename Ctx>
__device__ void slope_27(const Ctx& c, int i, int j, int k, Prim<R>& dqx, Prim<R>& dqy, Prim<R>& dqz) {
  const Prim<R> q0 = c.prim(i, j, k);
  R vmin[5], vmax[5];
  const R c0[5] = {q0.r, q0.u, q0.v, q0.w, q0.p};
#pragma unroll
  for (int n = 0; n < 5; ++n) { vmin[n] = (R)0; vmax[n] = (R)0; }  // dfmmm = 0 is in the set
  R xm[5], xp[5], ym[5], yp[5], zm[5], zp[5];
#pragma unroll 1
  for (int dk = -1; dk <= 1; ++dk)
#pragma unroll 1
    for (int dj = -1; dj <= 1; ++dj)
#pragma unroll 1
      for (int di = -1; di <= 1; ++di) {
        const Prim<R> qn = c.prim(i + di, j + dj, k + dk);
        const R cn[5] = {qn.r, qn.u, qn.v, qn.w, qn.p};
#pragma unroll
        for (int n = 0; n < 5; ++n) {
          const R df = cn[n] - c0[n];
          vmin[n] = fmin_f(vmin[n], df);
          vmax[n] = fmax_f(vmax[n], df);
          if (dj == 0 && dk == 0 && di == -1) xm[n] = cn[n];
          if (dj == 0 && dk == 0 && di == 1) xp[n] = cn[n];
          if (di == 0 && dk == 0 && dj == -1) ym[n] = cn[n];
          if (di == 0 && dk == 0 && dj == 1) yp[n] = cn[n];
          if (di == 0 && dj == 0 && dk == -1) zm[n] = cn[n];
          if (di == 0 && dj == 0 && dk == 1) zp[n] = cn[n];
        }
      }
  R ox[5], oy[5], oz[5];
#pragma unroll
  for (int n = 0; n < 5; ++n) {
    const R half = (R)0.5;
    const R dfx = half * (xp[n] - xm[n]);
    const R dfy = half * (yp[n] - ym[n]);
    const R dfz = half * (zp[n] - zm[n]);
    const R dff = half * (abs_w(dfx) + abs_w(dfy) + abs_w(dfz));
    R slop;
    if (dff > (R)0)
      slop = fmin_f((R)1, div_w(fmin_f(abs_w(vmin[n]), abs_w(vmax[n])), dff));
    else
      slop = (R)1;
    ox[n] = slop * dfx;
    oy[n] = slop * dfy;
    oz[n] = slop * dfz;
  }
  dqx = Prim<R>{ox[0], ox[1], ox[2], ox[3], ox[4]};
  dqy = Prim<R>{oy[0], oy[1], oy[2], oy[3], oy[4]};
  dqz = Prim<R>{oz[0], oz[1], oz[2], oz[3], oz[4]};
}

template <typename R>
__device__ __forceinline__ void store_face(double* flux, double* tmp, size_t fbase, size_t tbase, size_t fvs,
                                           int d, const Flux6<R>& g, const Face<R>& L, const Face<R>& Rt,
                                           const FluxScale& fs) {
  R fc[5];
  unrotate(g, d, fc);
#pragma unroll
  for (int v = 0; v < 5; ++v) flux[fbase + (size_t)v * fvs] = scale_flux(fc[v], fs);  // :106,129,153
  const R t1 = (R)0.5 * (L.un + Rt.un);                                                // :845
  tmp[tbase] = scale_flux(t1, fs);                                                    // :111
  tmp[tbase + fvs] = scale_flux(g.f[5], fs);
}

// One row of faces along direction D of one oct (see the file comment).
// With ZR, ok points at the batch's refinement flags (+ lane) and a face
// with a flagged cell on either side stores a zero flux: the value the
// reference's reset leaves there (hydro/godunov_fine.f90:720-749).
// Ctx supplies prim(i,j,k) and nv: RowCtx evaluates the primitives from
// uin, and the staged kernel of unsplit.hip reads them from LDS (the 27-point
// slopes need RowCtx).  ST: the 7-point slope type fixed at compile time, or
// -1 (see slope1).
template <typename R, int D, int RS, bool S27, bool ZR, int ST, typename Ctx>
__device__ __forceinline__ void row_faces(const Ctx& c, const SolverConsts<R>& sc, int a1, int b1,
                                          R dtdx, double* flux, double* tmp, size_t fvs, const FluxScale fs,
                                          const unsigned char* ok) {
  constexpr int T1 = (D == 0) ? 1 : 0;  // first transverse axis
  constexpr int T2 = (D == 2) ? 1 : 2;  // second transverse axis
  int ci, cj, ck;
  // Sliding window of primitives along the row: positions s-1, s, s+1.
  row_cell(D, -1, a1, b1, ci, cj, ck);
  Prim<R> qa = c.prim(ci, cj, ck);
  row_cell(D, 0, a1, b1, ci, cj, ck);
  Prim<R> qb = c.prim(ci, cj, ck);
  Face<R> fl_prev{};
#pragma unroll 1
  for (int s = 0; s <= 3; ++s) {
    row_cell(D, s + 1, a1, b1, ci, cj, ck);
    const Prim<R> qc = c.prim(ci, cj, ck);
    int i, j, k;
    row_cell(D, s, a1, b1, i, j, k);
    Prim<R> dqd, dq1, dq2;
    if constexpr (S27) {
      Prim<R> gx, gy, gz;
      slope_27(c, i, j, k, gx, gy, gz);
      dqd = (D == 0) ? gx : ((D == 1) ? gy : gz);
      dq1 = (T1 == 0) ? gx : gy;
      dq2 = (T2 == 1) ? gy : gz;
    } else {
      dqd = slope_prim<R, ST>(qa, qb, qc, sc.slope_type, sc.theta);
      const Prim<R> m1 = (T1 == 0) ? c.prim(i - 1, j, k) : c.prim(i, j - 1, k);
      const Prim<R> p1 = (T1 == 0) ? c.prim(i + 1, j, k) : c.prim(i, j + 1, k);
      dq1 = slope_prim<R, ST>(m1, qb, p1, sc.slope_type, sc.theta);
      const Prim<R> m2 = (T2 == 1) ? c.prim(i, j - 1, k) : c.prim(i, j, k - 1);
      const Prim<R> p2 = (T2 == 1) ? c.prim(i, j + 1, k) : c.prim(i, j, k + 1);
      dq2 = slope_prim<R, ST>(m2, qb, p2, sc.slope_type, sc.theta);
    }
    const Prim<R>& dqx = (D == 0) ? dqd : dq1;
    const Prim<R>& dqy = (D == 1) ? dqd : ((D == 0) ? dq1 : dq2);
    const Prim<R>& dqz = (D == 2) ? dqd : dq2;
    const Source<R> src = trace_sources(qb, dqx, dqy, dqz, sc.gamma);
    Prim<R> qp, qm;
    trace_dir(qb, dqd, src, dtdx, sc.smallr, qp, qm);
    if (s >= 1) {
      // face s of the row: left = qm of cell s-1, right = qp of cell s
      const Face<R> fr = rotate(qp, D);
      const Flux6<R> g = riemann_fixed<R, RS>(fl_prev, fr, sc);
      const size_t fslot = (size_t)((i - 1) + 3 * ((j - 1) + 3 * (k - 1))) * c.nv;
      store_face(flux, tmp, fslot + (size_t)D * 5 * fvs, fslot + (size_t)D * 2 * fvs, fvs, D, g, fl_prev, fr, fs);
      if (ZR) {
        int li = i, lj = j, lk = k;  // left cell of the face
        if (D == 0) --li; else if (D == 1) --lj; else --lk;
        if (ok[(size_t)cell6(li, lj, lk) * c.nv] || ok[(size_t)cell6(i, j, k) * c.nv]) {
#pragma unroll
          for (int v = 0; v < 5; ++v) flux[fslot + (size_t)D * 5 * fvs + (size_t)v * fvs] = 0.0;
        }
      }
    }
    fl_prev = rotate(qm, D);
    qa = qb;
    qb = qc;
  }
}

// Slopes of cell (i,j,k) and its traced states along D (qp on the minus
// side, qm on the plus side), exactly as row_faces computes them for a cell
// of a row: the same routines on the same primitives.
template <typename R, int D, bool S27, int ST, typename Ctx>
__device__ __forceinline__ void trace_cell(const Ctx& c, const SolverConsts<R>& sc, int i, int j, int k,
                                           R dtdx, Prim<R>& qp, Prim<R>& qm) {
  const Prim<R> q = c.prim(i, j, k);
  Prim<R> dqx, dqy, dqz;
  if (S27) {
    slope_27(c, i, j, k, dqx, dqy, dqz);
  } else {
    dqx = slope_prim<R, ST>(c.prim(i - 1, j, k), q, c.prim(i + 1, j, k), sc.slope_type, sc.theta);
    dqy = slope_prim<R, ST>(c.prim(i, j - 1, k), q, c.prim(i, j + 1, k), sc.slope_type, sc.theta);
    dqz = slope_prim<R, ST>(c.prim(i, j, k - 1), q, c.prim(i, j, k + 1), sc.slope_type, sc.theta);
  }
  const Source<R> src = trace_sources(q, dqx, dqy, dqz, sc.gamma);
  trace_dir(q, D == 0 ? dqx : (D == 1 ? dqy : dqz), src, dtdx, sc.smallr, qp, qm);
}

// The face of direction D whose right cell is (i,j,k) (ZR: see row_faces).
template <typename R, int D, int RS, bool S27, bool ZR, int ST, typename Ctx>
__device__ __forceinline__ void face_dir(const Ctx& c, const SolverConsts<R>& sc, int i, int j, int k,
                                         R dtdx, double* flux, double* tmp, size_t fvs, const FluxScale fs,
                                         const unsigned char* ok) {
  constexpr int di = D == 0, dj = D == 1, dk = D == 2;
  Prim<R> qpL, qmL, qpR, qmR;
  trace_cell<R, D, S27, ST>(c, sc, i - di, j - dj, k - dk, dtdx, qpL, qmL);
  trace_cell<R, D, S27, ST>(c, sc, i, j, k, dtdx, qpR, qmR);
  const Face<R> fl = rotate(qmL, D), fr = rotate(qpR, D);
  const Flux6<R> g = riemann_fixed<R, RS>(fl, fr, sc);
  const size_t fslot = (size_t)((i - 1) + 3 * ((j - 1) + 3 * (k - 1))) * c.nv;
  store_face(flux, tmp, fslot + (size_t)D * 5 * fvs, fslot + (size_t)D * 2 * fvs, fvs, D, g, fl, fr, fs);
  if (ZR && (ok[(size_t)cell6(i - di, j - dj, k - dk) * c.nv] || ok[(size_t)cell6(i, j, k) * c.nv])) {
#pragma unroll
    for (int v = 0; v < 5; ++v) flux[fslot + (size_t)D * 5 * fvs + (size_t)v * fvs] = 0.0;
  }
}

// One (group, face, lane) task: the 36 faces of an oct one per thread, face
// f = 12*D + r with the right cell (i,j,k) enumerated x fastest over the
// 3x2x2 (D=x), 2x3x2 (D=y) or 2x2x3 (D=z) face positions.  Each face
// evaluates its two cells and their neighbours (14 primitives) instead of
// sharing a row's sliding window: about 1.3x the arithmetic of the row
// decomposition for a third of its chain of dependent steps, which is the
// better trade where a launch holds few octs.  Same routines, same inputs,
// same bits.  G: a.grav is non-null (see RowCtx).
template <typename R, int RS, bool S27, bool ZR, int ST, bool G>
__device__ __forceinline__ void face_task(const UnsplitArgs& a, long t, long ngroups, int ngrid_last) {
  const int nv = a.nv;
  const long per_group = 36L * nv;
  const long gidx = t / per_group;
  const int rem = (int)(t - gidx * per_group);
  const int f = rem / nv;
  const int l = rem - f * nv;
  const int ngrid = (gidx == ngroups - 1) ? ngrid_last : nv;
  if (l >= ngrid) return;
  const int d = f / 12, r = f - d * 12;
  const int ni = 2 + (d == 0), nj = 2 + (d == 1);
  const int i = 1 + r % ni, j = 1 + (r / ni) % nj, k = 1 + r / (ni * nj);

  const size_t gsz_u = (size_t)nv * 216 * 5;
  const size_t gsz_g = (size_t)nv * 216 * 3;
  const size_t gsz_f = (size_t)nv * 27 * 5 * 3;
  const size_t gsz_t = (size_t)nv * 27 * 2 * 3;
  const HydroParams hp = a.hp;
  const double dt = a.d_dt ? *a.d_dt : a.dt;
  RowCtx<R, G> c;
  c.u = a.uin + gidx * gsz_u + l;
  c.g = G ? a.grav + gidx * gsz_g + l : nullptr;
  c.vs = (size_t)nv * 216;
  c.nv = nv;
  c.dtxhalf = dt * 0.5;
  c.smalle = hp.smallc * hp.smallc / hp.gamma / (hp.gamma - 1.0);  // hydro/umuscl.f90:884
  c.hp = &hp;
  const SolverConsts<R> sc = make_consts<R>(hp);
  double* flux = a.flux + gidx * gsz_f + l;
  double* tmp = a.tmp + gidx * gsz_t + l;
  const size_t fvs = (size_t)nv * 27;
  const unsigned char* ok = ZR ? a.ok + gidx * (size_t)nv * 216 + l : nullptr;
  if (d == 0)
    face_dir<R, 0, RS, S27, ZR, ST>(c, sc, i, j, k, (R)(dt / a.dx[0]), flux, tmp, fvs, make_scale(dt, a.dx[0]), ok);
  else if (d == 1)
    face_dir<R, 1, RS, S27, ZR, ST>(c, sc, i, j, k, (R)(dt / a.dx[1]), flux, tmp, fvs, make_scale(dt, a.dx[1]), ok);
  else
    face_dir<R, 2, RS, S27, ZR, ST>(c, sc, i, j, k, (R)(dt / a.dx[2]), flux, tmp, fvs, make_scale(dt, a.dx[2]), ok);
}

// One (group, row, lane) task of the flat decomposition (ZR, ST: see
// row_faces; G: a.grav is non-null, see RowCtx).
template <typename R, int RS, bool S27, bool ZR, int ST, bool G>
__device__ __forceinline__ void unsplit_task(const UnsplitArgs& a, long t, long ngroups, int ngrid_last) {
  const int nv = a.nv;
  const long per_group = 12L * nv;
  const long gidx = t / per_group;
  const int rem = (int)(t - gidx * per_group);
  const int row = rem / nv;
  const int l = rem - row * nv;
  const int ngrid = (gidx == ngroups - 1) ? ngrid_last : nv;
  if (l >= ngrid) return;

  const int d = row >> 2;               // direction of the faces
  const int a1 = 1 + (row & 1);         // first transverse coordinate
  const int b1 = 1 + ((row >> 1) & 1);  // second transverse coordinate

  const size_t gsz_u = (size_t)nv * 216 * 5;
  const size_t gsz_g = (size_t)nv * 216 * 3;
  const size_t gsz_f = (size_t)nv * 27 * 5 * 3;
  const size_t gsz_t = (size_t)nv * 27 * 2 * 3;

  const HydroParams hp = a.hp;
  const double dt = a.d_dt ? *a.d_dt : a.dt;
  RowCtx<R, G> c;
  c.u = a.uin + gidx * gsz_u + l;
  c.g = G ? a.grav + gidx * gsz_g + l : nullptr;
  c.vs = (size_t)nv * 216;
  c.nv = nv;
  c.dtxhalf = dt * 0.5;
  c.smalle = hp.smallc * hp.smallc / hp.gamma / (hp.gamma - 1.0);  // :884
  c.hp = &hp;
  const SolverConsts<R> sc = make_consts<R>(hp);

  double* flux = a.flux + gidx * gsz_f + l;
  double* tmp = a.tmp + gidx * gsz_t + l;
  const size_t fvs = (size_t)nv * 27;
  const unsigned char* ok = ZR ? a.ok + gidx * (size_t)nv * 216 + l : nullptr;
  if (d == 0)
    row_faces<R, 0, RS, S27, ZR, ST>(c, sc, a1, b1, (R)(dt / a.dx[0]), flux, tmp, fvs, make_scale(dt, a.dx[0]), ok);  // :516
  else if (d == 1)
    row_faces<R, 1, RS, S27, ZR, ST>(c, sc, a1, b1, (R)(dt / a.dx[1]), flux, tmp, fvs, make_scale(dt, a.dx[1]), ok);
  else
    row_faces<R, 2, RS, S27, ZR, ST>(c, sc, a1, b1, (R)(dt / a.dx[2]), flux, tmp, fvs, make_scale(dt, a.dx[2]), ok);
}

}  // namespace unsplit_dev
}  // namespace ramses
