// Device-side MUSCL-Hancock building blocks for gfx950, shared by every
// kernel of the library.  Each routine follows one reference routine (cited
// per function) and keeps its left-to-right evaluation order: the library is
// compiled with -ffp-contract=off, so in FP64 every value rounds exactly as
// the reference's own x86-64 build rounds it.
//
// Template parameter R is the arithmetic type of primitives, slopes, traced
// states and Riemann fluxes: double (FP64 path) or float (FP32-mixed path).
// Conservative state, E - kinetic recovery and the dt reduction are always
// double.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace ramses {

enum Riemann : int { kLLF = 0, kHLLC = 1, kHLL = 2, kACOUSTIC = 3, kEXACT = 4 };
// Template value of kernels that pick the solver at run time (the
// acoustic and exact solvers share one instantiation per kernel).
constexpr int kDynamicRiemann = 99;

// hydro/hydro_parameters.f90:74-85 resolved once on the host.
struct HydroParams {
  double gamma;
  double smallr;
  double smallc;
  double slope_theta;
  double courant_factor;
  int slope_type;  // 0,1,2,3,7,8
  int riemann;     // Riemann enum
  int niter_riemann;  // Newton iterations of the exact solver
};

// Launch-shape knobs set through ramses_tune (0 = built-in default).
struct Tuning {
  int unsplit_block;
  int level_tile;
  int level_tz;
  int persist_blocks;  // workgroups of a captured AMR step's fixed-grid kernels (0: default)
  int amr_face_octs;   // largest level run face-per-thread by the AMR unsplit (0: default, -1: never)
};
Tuning& tuning();
// Workgroups of the fixed grids of a captured AMR coarse step.
inline unsigned persist_blocks() { return tuning().persist_blocks > 0 ? (unsigned)tuning().persist_blocks : 1024u; }
// Levels of at most this many octs run the AMR unsplit one face per thread.
inline int amr_face_octs() { return tuning().amr_face_octs > 0 ? tuning().amr_face_octs : (tuning().amr_face_octs < 0 ? 0 : 8192); }

// Fortran MAX/MIN of two reals.  The hardware min/max (IEEE maxNum/minNum)
// returns the same value as the reference for every non-NaN pair; the only
// possible difference, the sign of a zero result when +0 and -0 tie, cannot
// reach any output of the path: every tie of that kind feeds either a sum
// with a nonzero term (SL/SR of hllc/hll) or is excluded by a dlft*drgt>0
// test (minmod), and every other MAX/MIN of the path has a positive operand.
__device__ __forceinline__ double fmax_f(double a, double b) { return __builtin_fmax(a, b); }
__device__ __forceinline__ double fmin_f(double a, double b) { return __builtin_fmin(a, b); }
__device__ __forceinline__ float fmax_f(float a, float b) { return __builtin_fmaxf(a, b); }
__device__ __forceinline__ float fmin_f(float a, float b) { return __builtin_fminf(a, b); }

// Division / square root in the working precision.  FP64 keeps IEEE
// correctly rounded division and sqrt (the reference's semantics); the FP32
// path uses the hardware reciprocal / sqrt (about 1 ulp), which is what makes
// it cheaper than the FP64 path on the divide-heavy Riemann solve.
__device__ __forceinline__ double div_w(double a, double b) { return a / b; }
__device__ __forceinline__ float div_w(float a, float b) { return a * __builtin_amdgcn_rcpf(b); }
__device__ __forceinline__ double sqrt_w(double a) { return __builtin_sqrt(a); }
__device__ __forceinline__ float sqrt_w(float a) { return __builtin_amdgcn_sqrtf(a); }
__device__ __forceinline__ double pow_w(double a, double b) { return pow(a, b); }
__device__ __forceinline__ float pow_w(float a, float b) { return powf(a, b); }
__device__ __forceinline__ double abs_w(double a) { return __builtin_fabs(a); }
__device__ __forceinline__ float abs_w(float a) { return __builtin_fabsf(a); }
__device__ __forceinline__ double copysign_w(double a, double b) { return __builtin_copysign(a, b); }
__device__ __forceinline__ float copysign_w(float a, float b) { return __builtin_copysignf(a, b); }

// Primitive state of one cell: rho, vx, vy, vz, p.
template <typename R>
struct Prim {
  R r, u, v, w, p;
};

// ctoprim for one cell (hydro/umuscl.f90:861-965, NENER=0, no sound speed:
// the MUSCL trace does not read `c`).  Density floor, 1/rho, velocities and
// E/rho - kinetic are evaluated in FP64; the result is rounded to R.
// FP64 reciprocal to about 1 ulp: hardware estimate refined by two Newton
// steps.  Used only by the FP32-mixed path, whose FP64 parts need FP64
// accuracy but not the reference's correctly rounded division.
__device__ __forceinline__ double rcp_fast(double x) {
  double r = __builtin_amdgcn_rcp(x);
  double e = __builtin_fma(-x, r, 1.0);
  r = __builtin_fma(r, e, r);
  e = __builtin_fma(-x, r, 1.0);
  return __builtin_fma(r, e, r);
}

template <typename R>
__device__ __forceinline__ double recip(double x) {
  if constexpr (sizeof(R) == 8) return 1.0 / x;  // hydro/umuscl.f90:897, correctly rounded
  else return rcp_fast(x);
}

template <typename R>
__device__ __forceinline__ Prim<R> ctoprim_cell(double u0, double u1, double u2, double u3, double u4,
                                                double gx, double gy, double gz, double dtxhalf,
                                                double smalle, const HydroParams& hp) {
  const double r = fmax_f(u0, hp.smallr);                  // :894
  const double oneoverrho = recip<R>(r);                    // :897
  const double vx = u1 * oneoverrho;                        // :898-904
  const double vy = u2 * oneoverrho;
  const double vz = u3 * oneoverrho;
  double eken = 0.5 * vx * vx;                              // :907-913
  eken = eken + 0.5 * vy * vy;
  eken = eken + 0.5 * vz * vz;
  const double erad = 0.0;
  const double eint = fmax_f(u4 * oneoverrho - eken - erad, smalle);  // :923
  const double p = (hp.gamma - 1.0) * r * eint;             // :924
  Prim<R> q;
  q.r = (R)r;
  q.p = (R)p;
  q.u = (R)(vx + gx * dtxhalf);                             // :936-942
  q.v = (R)(vy + gy * dtxhalf);
  q.w = (R)(vz + gz * dtxhalf);
  return q;
}

// One-dimensional slope of the 3D uslope branches with axial stencils
// (hydro/umuscl.f90:1241-1285 minmod, :1286-1327 moncen, :1393-1431 van
// Leer, :1432-1473 generalised moncen).
// ST >= 0 fixes the slope type at compile time (the kernel then holds one
// limiter and no branches on it); ST = -1 takes it from st at run time.
template <typename R, int ST = -1>
__device__ __forceinline__ R slope1(R qm1, R q0, R qp1, int st_run, R theta) {
  const int st = ST >= 0 ? ST : st_run;
  const R zero = (R)0;
  if (st == 1) {
    const R dlft = q0 - qm1;
    const R drgt = qp1 - q0;
    if ((dlft * drgt) <= zero) return zero;
    return (dlft > zero) ? fmin_f(dlft, drgt) : fmax_f(dlft, drgt);
  } else if (st == 2) {
    const R two = (R)2;
    const R dlft = two * (q0 - qm1);
    const R drgt = two * (qp1 - q0);
    const R dcen = (R)0.5 * (dlft + drgt) / two;
    const R dsgn = copysign_w((R)1, dcen);
    R dlim = fmin_f(abs_w(dlft), abs_w(drgt));
    if ((dlft * drgt) <= zero) dlim = zero;
    return dsgn * fmin_f(dlim, abs_w(dcen));
  } else if (st == 7) {
    const R dlft = (q0 - qm1);
    const R drgt = (qp1 - q0);
    if ((dlft * drgt) <= zero) return zero;
    return div_w((R)2 * dlft * drgt, (dlft + drgt));
  } else if (st == 8) {
    const R dlft = (q0 - qm1);
    const R drgt = (qp1 - q0);
    const R dcen = (R)0.5 * (dlft + drgt);
    const R dsgn = copysign_w((R)1, dcen);
    R dlim = fmin_f(theta * abs_w(dlft), theta * abs_w(drgt));
    if ((dlft * drgt) <= zero) dlim = zero;
    return dsgn * fmin_f(dlim, abs_w(dcen));
  }
  return zero;  // slope_type 0 (:1002-1005)
}

// Slopes of the five primitives in one direction (ST: see slope1).
template <typename R, int ST = -1>
__device__ __forceinline__ Prim<R> slope_prim(const Prim<R>& a, const Prim<R>& b, const Prim<R>& c,
                                              int st, R theta) {
  Prim<R> d;
  d.r = slope1<R, ST>(a.r, b.r, c.r, st, theta);
  d.u = slope1<R, ST>(a.u, b.u, c.u, st, theta);
  d.v = slope1<R, ST>(a.v, b.v, c.v, st, theta);
  d.w = slope1<R, ST>(a.w, b.w, c.w, st, theta);
  d.p = slope1<R, ST>(a.p, b.p, c.p, st, theta);
  return d;
}

// Source terms of the Hancock predictor (hydro/umuscl.f90:576-580).
template <typename R>
struct Source {
  R sr0, su0, sv0, sw0, sp0;
};

template <typename R>
__device__ __forceinline__ Source<R> trace_sources(const Prim<R>& q, const Prim<R>& dx_,
                                                   const Prim<R>& dy_, const Prim<R>& dz_, R gamma) {
  const R r = q.r, u = q.u, v = q.v, w = q.w, p = q.p;
  Source<R> s;
  s.sr0 = -u * dx_.r - v * dy_.r - w * dz_.r - (dx_.u + dy_.v + dz_.w) * r;
  s.sp0 = -u * dx_.p - v * dy_.p - w * dz_.p - (dx_.u + dy_.v + dz_.w) * gamma * p;
  s.su0 = -u * dx_.u - v * dy_.u - w * dz_.u - div_w((dx_.p), r);
  s.sv0 = -u * dx_.v - v * dy_.v - w * dz_.v - div_w((dy_.p), r);
  s.sw0 = -u * dx_.w - v * dy_.w - w * dz_.w - div_w((dz_.p), r);
  return s;
}

// Right (qp, minus side) and left (qm, plus side) traced states along one
// direction with slope dd and dt/dx = dtdx (hydro/umuscl.f90:591-673).
template <typename R>
__device__ __forceinline__ void trace_dir(const Prim<R>& q, const Prim<R>& dd, const Source<R>& s, R dtdx,
                                          R smallr, Prim<R>& qp, Prim<R>& qm) {
  const R half = (R)0.5;
  qp.r = q.r - half * dd.r + s.sr0 * dtdx * half;
  qp.p = q.p - half * dd.p + s.sp0 * dtdx * half;
  qp.u = q.u - half * dd.u + s.su0 * dtdx * half;
  qp.v = q.v - half * dd.v + s.sv0 * dtdx * half;
  qp.w = q.w - half * dd.w + s.sw0 * dtdx * half;
  if (qp.r < smallr) qp.r = q.r;
  qm.r = q.r + half * dd.r + s.sr0 * dtdx * half;
  qm.p = q.p + half * dd.p + s.sp0 * dtdx * half;
  qm.u = q.u + half * dd.u + s.su0 * dtdx * half;
  qm.v = q.v + half * dd.v + s.sv0 * dtdx * half;
  qm.w = q.w + half * dd.w + s.sw0 * dtdx * half;
  if (qm.r < smallr) qm.r = q.r;
}

// Interface state in the rotated order of cmpflxm (hydro/umuscl.f90:749-780):
// rho, normal velocity, pressure, tangential 1, tangential 2.
template <typename R>
struct Face {
  R r, un, p, ut1, ut2;
};

template <typename R>
__device__ __forceinline__ Face<R> rotate(const Prim<R>& q, int d) {
  Face<R> f;
  f.r = q.r;
  f.p = q.p;
  if (d == 0) { f.un = q.u; f.ut1 = q.v; f.ut2 = q.w; }
  else if (d == 1) { f.un = q.v; f.ut1 = q.u; f.ut2 = q.w; }
  else { f.un = q.w; f.ut1 = q.u; f.ut2 = q.v; }
  return f;
}

// Godunov flux in rotated order plus the internal-energy flux (fgdnv(1:6)).
template <typename R>
struct Flux6 {
  R f[6];
};

// riemann_llf: hydro/godunov_utils.f90:660-820
template <typename R>
__device__ __forceinline__ Flux6<R> riemann_llf(const Face<R>& L, const Face<R>& Rt, R gamma, R smallr,
                                                R smallp, R entho) {
  const R half = (R)0.5;
  const R rl = fmax_f(L.r, smallr), ul = L.un, pl = fmax_f(L.p, rl * smallp);
  R cl = gamma * pl;
  cl = sqrt_w(div_w(cl, rl));
  const R rr = fmax_f(Rt.r, smallr), ur = Rt.un, pr = fmax_f(Rt.p, rr * smallp);
  R cr = gamma * pr;
  cr = sqrt_w(div_w(cr, rr));
  const R cmax = fmax_f(abs_w(ul) + cl, abs_w(ur) + cr);
  R uL[6], uR[6], fL[6], fR[6];
  uL[0] = L.r;
  uR[0] = Rt.r;
  uL[1] = L.r * L.un;
  uR[1] = Rt.r * Rt.un;
  uL[2] = L.p * entho + half * L.r * (L.un * L.un);
  uR[2] = Rt.p * entho + half * Rt.r * (Rt.un * Rt.un);
  uL[2] = uL[2] + half * L.r * (L.ut1 * L.ut1);
  uR[2] = uR[2] + half * Rt.r * (Rt.ut1 * Rt.ut1);
  uL[2] = uL[2] + half * L.r * (L.ut2 * L.ut2);
  uR[2] = uR[2] + half * Rt.r * (Rt.ut2 * Rt.ut2);
  uL[3] = L.r * L.ut1;
  uR[3] = Rt.r * Rt.ut1;
  uL[4] = L.r * L.ut2;
  uR[4] = Rt.r * Rt.ut2;
  uL[5] = L.p * entho;
  uR[5] = Rt.p * entho;
  fL[0] = L.un * uL[0];
  fR[0] = Rt.un * uR[0];
  fL[1] = L.un * uL[1] + L.p;
  fR[1] = Rt.un * uR[1] + Rt.p;
  fL[2] = L.un * (uL[2] + L.p);
  fR[2] = Rt.un * (uR[2] + Rt.p);
#pragma unroll
  for (int n = 3; n < 6; ++n) {
    fL[n] = L.un * uL[n];
    fR[n] = Rt.un * uR[n];
  }
  Flux6<R> g;
#pragma unroll
  for (int n = 0; n < 6; ++n) g.f[n] = half * (fL[n] + fR[n] - cmax * (uR[n] - uL[n]));
  return g;
}

// riemann_hll: hydro/godunov_utils.f90:825-983
template <typename R>
__device__ __forceinline__ Flux6<R> riemann_hll(const Face<R>& L, const Face<R>& Rt, R gamma, R smallr,
                                                R smallp, R entho) {
  const R half = (R)0.5, zero = (R)0;
  const R rl = fmax_f(L.r, smallr), ul = L.un, pl = fmax_f(L.p, rl * smallp);
  R cl = gamma * pl;
  cl = sqrt_w(div_w(cl, rl));
  const R rr = fmax_f(Rt.r, smallr), ur = Rt.un, pr = fmax_f(Rt.p, rr * smallp);
  R cr = gamma * pr;
  cr = sqrt_w(div_w(cr, rr));
  const R SL = fmin_f(fmin_f(ul, ur) - fmax_f(cl, cr), zero);
  const R SR = fmax_f(fmax_f(ul, ur) + fmax_f(cl, cr), zero);
  R uL[6], uR[6], fL[6], fR[6];
  uL[0] = L.r;
  uR[0] = Rt.r;
  uL[1] = L.r * L.un;
  uR[1] = Rt.r * Rt.un;
  uL[2] = L.p * entho + half * L.r * (L.un * L.un);
  uR[2] = Rt.p * entho + half * Rt.r * (Rt.un * Rt.un);
  uL[2] = uL[2] + half * L.r * (L.ut1 * L.ut1);
  uR[2] = uR[2] + half * Rt.r * (Rt.ut1 * Rt.ut1);
  uL[2] = uL[2] + half * L.r * (L.ut2 * L.ut2);
  uR[2] = uR[2] + half * Rt.r * (Rt.ut2 * Rt.ut2);
  uL[3] = L.r * L.ut1;
  uR[3] = Rt.r * Rt.ut1;
  uL[4] = L.r * L.ut2;
  uR[4] = Rt.r * Rt.ut2;
  uL[5] = L.p * entho;
  uR[5] = Rt.p * entho;
  fL[0] = uL[1];
  fR[0] = uR[1];
  fL[1] = L.p + uL[1] * L.un;
  fR[1] = Rt.p + uR[1] * Rt.un;
  fL[2] = L.un * (uL[2] + L.p);
  fR[2] = Rt.un * (uR[2] + Rt.p);
#pragma unroll
  for (int n = 3; n < 6; ++n) {
    fL[n] = L.un * uL[n];
    fR[n] = Rt.un * uR[n];
  }
  Flux6<R> g;
  const R den = SR - SL;
#pragma unroll
  for (int n = 0; n < 6; ++n) g.f[n] = div_w(SR * fL[n] - SL * fR[n] + SR * SL * (uR[n] - uL[n]), den);
  return g;
}

// riemann_hllc: hydro/godunov_utils.f90:988-1209.  Only the star state of
// the sampled branch feeds the flux, so the other branch's divisions are
// skipped; the values of the sampled quantities are unchanged.
template <typename R>
__device__ __forceinline__ Flux6<R> riemann_hllc(const Face<R>& L, const Face<R>& Rt, R gamma, R smallr,
                                                 R smallp, R entho, R smallc2) {
  const R half = (R)0.5, zero = (R)0;
  const R rl = fmax_f(L.r, smallr);
  const R Pl = fmax_f(L.p, rl * smallp);
  const R ul = L.un;
  const R el = Pl * entho;
  R ecinl = half * rl * ul * ul;
  ecinl = ecinl + half * rl * (L.ut1 * L.ut1);
  ecinl = ecinl + half * rl * (L.ut2 * L.ut2);
  const R etotl = el + ecinl;
  const R Ptotl = Pl;
  const R rr = fmax_f(Rt.r, smallr);
  const R Pr = fmax_f(Rt.p, rr * smallp);
  const R ur = Rt.un;
  const R er = Pr * entho;
  R ecinr = half * rr * ur * ur;
  ecinr = ecinr + half * rr * (Rt.ut1 * Rt.ut1);
  ecinr = ecinr + half * rr * (Rt.ut2 * Rt.ut2);
  const R etotr = er + ecinr;
  const R Ptotr = Pr;
  R cfastl = gamma * Pl;
  cfastl = sqrt_w(fmax_f(div_w(cfastl, rl), smallc2));
  R cfastr = gamma * Pr;
  cfastr = sqrt_w(fmax_f(div_w(cfastr, rr), smallc2));
  const R SL = fmin_f(ul, ur) - fmax_f(cfastl, cfastr);
  const R SR = fmax_f(ul, ur) + fmax_f(cfastl, cfastr);
  const R rcl = rl * (ul - SL);
  const R rcr = rr * (SR - ur);
  const R ustar = div_w(rcr * ur + rcl * ul + (Ptotl - Ptotr), (rcr + rcl));
  R ro, uo, Ptoto, etoto, eo;
  if (SL > zero) {
    ro = rl; uo = ul; Ptoto = Ptotl; etoto = etotl; eo = el;
  } else if (SR > zero || ustar > zero) {
    const R Ptotstar = div_w(rcr * Ptotl + rcl * Ptotr + rcl * rcr * (ul - ur), (rcr + rcl));
    // left star state when ustar>0, right star state otherwise (SR>0)
    const bool left = ustar > zero;
    const R S = left ? SL : SR;
    const R us = left ? ul : ur;
    const R rs = left ? rl : rr;
    const R es = left ? el : er;
    const R ets = left ? etotl : etotr;
    const R Pts = left ? Ptotl : Ptotr;
    const R den = S - ustar;
    ro = div_w(rs * (S - us), den);
    etoto = div_w((S - us) * ets - Pts * us + Ptotstar * ustar, den);
    eo = div_w(es * (S - us), den);
    uo = ustar;
    Ptoto = Ptotstar;
  } else {
    ro = rr; uo = ur; Ptoto = Ptotr; etoto = etotr; eo = er;
  }
  Flux6<R> g;
  g.f[0] = ro * uo;
  g.f[1] = ro * uo * uo + Ptoto;
  g.f[2] = (etoto + Ptoto) * uo;
  if (ustar > zero) {
    g.f[3] = ro * uo * L.ut1;
    g.f[4] = ro * uo * L.ut2;
  } else {
    g.f[3] = ro * uo * Rt.ut1;
    g.f[4] = ro * uo * Rt.ut2;
  }
  g.f[5] = uo * eo;
  return g;
}

// Flux of a sampled state q = (rho, u_n, p, u_t1, u_t2, e_int), shared by
// the acoustic and exact solvers (hydro/godunov_utils.f90:468-492, 631-652).
template <typename R>
__device__ __forceinline__ Flux6<R> sampled_flux(const R q[6], R entho) {
  const R half = (R)0.5;
  Flux6<R> g;
  g.f[0] = q[0] * q[1];
  g.f[1] = q[2] + q[0] * (q[1] * q[1]);
  R etot = q[2] * entho + half * q[0] * (q[1] * q[1]);
  etot = etot + half * q[0] * (q[3] * q[3]);
  etot = etot + half * q[0] * (q[4] * q[4]);
  g.f[2] = q[1] * (etot + q[2]);
#pragma unroll
  for (int n = 3; n < 6; ++n) g.f[n] = g.f[0] * q[n];
  return g;
}

// riemann_acoustic: hydro/godunov_utils.f90:500-655 (linearised star state,
// rarefaction fan interpolated linearly).
template <typename R>
__device__ __forceinline__ Flux6<R> riemann_acoustic(const Face<R>& L, const Face<R>& Rt, R gamma, R smallr,
                                                     R smallp, R entho, R smallc) {
  const R half = (R)0.5, zero = (R)0, one = (R)1;
  const R rl = fmax_f(L.r, smallr), ul = L.un, pl = fmax_f(L.p, rl * smallp);
  const R rr = fmax_f(Rt.r, smallr), ur = Rt.un, pr = fmax_f(Rt.p, rr * smallp);
  const R cl = sqrt_w(div_w(gamma * pl, rl)), cr = sqrt_w(div_w(gamma * pr, rr));
  const R wl = cl * rl, wr = cr * rr;
  const R pstar = div_w((wr * pl + wl * pr) + wl * wr * (ul - ur), wl + wr);
  const R ustar = div_w((wr * ur + wl * ul) + (pl - pr), wl + wr);
  const R sgnm = copysign_w(one, ustar);
  const bool left = sgnm == one;
  const R ro = left ? rl : rr, uo = left ? ul : ur, po = left ? pl : pr, co = left ? cl : cr;
  R rstar = ro + div_w(pstar - po, co * co);
  rstar = fmax_f(rstar, smallr);
  R cstar = sqrt_w(abs_w(div_w(gamma * pstar, rstar)));
  cstar = fmax_f(cstar, smallc);
  R spout = co - sgnm * uo, spin = cstar - sgnm * ustar;
  R ushock = half * (spin + spout);
  ushock = fmax_f(ushock, -sgnm * ustar);
  if (pstar >= po) {
    spout = ushock;
    spin = spout;
  }
  R q[6];
  if (spout < zero) {
    q[0] = ro; q[1] = uo; q[2] = po;
  } else if (spin >= zero) {
    q[0] = rstar; q[1] = ustar; q[2] = pstar;
  } else {
    const R frac = div_w(spout, spout - spin);
    q[0] = frac * rstar + (one - frac) * ro;
    q[1] = frac * ustar + (one - frac) * uo;
    q[2] = frac * pstar + (one - frac) * po;
  }
  q[3] = left ? L.ut1 : Rt.ut1;
  q[4] = left ? L.ut2 : Rt.ut2;
  q[5] = div_w(po, ro) * entho;
  return sampled_flux(q, entho);
}

// riemann_approx ('exact'): hydro/godunov_utils.f90:268-495.  Newton
// iterations on the two-shock star pressure; the reference drops a face from
// the iteration once |delta p / (p + smallr*smallp)| <= 1e-6, i.e. per face:
// update, then stop if converged, at most niter updates.  The rarefaction
// branches use pow (libm's on the reference side, the device library's
// here: those branches can differ in the last bit).
template <typename R>
__device__ __forceinline__ Flux6<R> riemann_exact(const Face<R>& L, const Face<R>& Rt, R gamma, R smallr, R smallp,
                                                  R entho, R smallc, int niter) {
  const R half = (R)0.5, zero = (R)0, one = (R)1, two = (R)2;
  const R smallpp = smallr * smallp;
  const R gamma6 = div_w(gamma + one, two * gamma);
  const R rl = fmax_f(L.r, smallr), ul = L.un, pl = fmax_f(L.p, rl * smallp);
  const R rr = fmax_f(Rt.r, smallr), ur = Rt.un, pr = fmax_f(Rt.p, rr * smallp);
  const R cl = gamma * pl * rl, cr = gamma * pr * rr;
  R wl = sqrt_w(cl), wr = sqrt_w(cr);
  R pold = fmax_f(div_w((wr * pl + wl * pr) + wl * wr * (ul - ur), wl + wr), zero);
  for (int iter = 1; iter <= niter; ++iter) {
    const R wwl = sqrt_w(cl * (one + div_w(gamma6 * (pold - pl), pl)));
    const R wwr = sqrt_w(cr * (one + div_w(gamma6 * (pold - pr), pr)));
    const R ql = div_w(two * (wwl * wwl * wwl), wwl * wwl + cl);
    const R qr = div_w(two * (wwr * wwr * wwr), wwr * wwr + cr);
    const R usl = ul - div_w(pold - pl, wwl);
    const R usr = ur + div_w(pold - pr, wwr);
    const R delp = fmax_f(div_w(qr * ql, qr + ql) * (usl - usr), -pold);
    pold = pold + delp;
    if (!(abs_w(div_w(delp, pold + smallpp)) > (R)1e-06)) break;
  }
  const R pstar = pold;
  wl = sqrt_w(cl * (one + div_w(gamma6 * (pstar - pl), pl)));
  wr = sqrt_w(cr * (one + div_w(gamma6 * (pstar - pr), pr)));
  const R ustar = half * (ul + div_w(pl - pstar, wl) + ur - div_w(pr - pstar, wr));
  const R sgnm = copysign_w(one, ustar);
  const bool left = sgnm == one;
  const R ro = left ? rl : rr, uo = left ? ul : ur, po = left ? pl : pr, wo = left ? wl : wr;
  const R co = fmax_f(smallc, sqrt_w(abs_w(div_w(gamma * po, ro))));
  const R inv_gamma = div_w(one, gamma);
  R rstar = pstar >= po ? div_w(ro, one + div_w(ro * (po - pstar), wo * wo)) : ro * pow_w(div_w(pstar, po), inv_gamma);
  rstar = fmax_f(rstar, smallr);
  R cstar = sqrt_w(abs_w(div_w(gamma * pstar, rstar)));
  cstar = fmax_f(cstar, smallc);
  R spout = co - sgnm * uo, spin = cstar - sgnm * ustar;
  const R ushock = div_w(wo, ro) - sgnm * uo;
  if (pstar >= po) {
    spout = ushock;
    spin = spout;
  }
  R q[6];
  if (spout <= zero) {
    q[0] = ro; q[1] = uo; q[2] = po;
  } else if (spin >= zero) {
    q[0] = rstar; q[1] = ustar; q[2] = pstar;
  } else {
    const R frac = div_w(spout, spout - spin);
    q[1] = frac * ustar + (one - frac) * uo;
    q[2] = frac * pstar + (one - frac) * po;
    q[0] = ro * pow_w(div_w(q[2], po), inv_gamma);
  }
  q[3] = left ? L.ut1 : Rt.ut1;
  q[4] = left ? L.ut2 : Rt.ut2;
  q[5] = div_w(po, ro) * entho;
  return sampled_flux(q, entho);
}

// Scaling of a face flux by dt/dx as the reference writes it, (f*dt)/dx
// (hydro/umuscl.f90:106,129,153).  When dx is a power of two the division is
// the exact multiplication by 1/dx (inv_dx != 0), which the FP64 path uses
// to stay bitwise while avoiding a divide; the FP32-mixed path multiplies by
// dt/dx in FP64.
struct FluxScale {
  double dt, dx, inv_dx, dtdx;
};

__host__ __device__ inline FluxScale make_scale(double dt, double dx) {
  FluxScale s;
  s.dt = dt;
  s.dx = dx;
  s.dtdx = dt / dx;
  int e = 0;
  const double m = __builtin_frexp(dx, &e);
  s.inv_dx = (m == 0.5) ? 1.0 / dx : 0.0;
  return s;
}

template <typename R>
__device__ __forceinline__ double scale_flux(R f, const FluxScale& s) {
  if constexpr (sizeof(R) == 8) {
    const double fd = (double)f * s.dt;
    return s.inv_dx != 0.0 ? fd * s.inv_dx : fd / s.dx;
  } else {
    return (double)f * s.dtdx;
  }
}

// Solver constants in working precision.
template <typename R>
struct SolverConsts {
  R gamma, smallr, smallp, entho, smallc2, theta, smallc;
  int riemann, slope_type, niter;
};

template <typename R>
__host__ __device__ inline SolverConsts<R> make_consts(const HydroParams& hp) {
  SolverConsts<R> c;
  c.gamma = (R)hp.gamma;
  c.smallr = (R)hp.smallr;
  c.smallp = (R)(hp.smallc * hp.smallc / hp.gamma);  // :683, :1017
  c.entho = (R)(1.0 / (hp.gamma - 1.0));
  c.smallc2 = (R)(hp.smallc * hp.smallc);
  c.theta = (R)hp.slope_theta;
  c.riemann = hp.riemann;
  c.slope_type = hp.slope_type;
  c.smallc = (R)hp.smallc;
  c.niter = hp.niter_riemann;
  return c;
}

template <typename R>
__device__ __forceinline__ Flux6<R> riemann_solve(const Face<R>& L, const Face<R>& Rt, const SolverConsts<R>& c) {
  if (c.riemann == kHLLC) return riemann_hllc(L, Rt, c.gamma, c.smallr, c.smallp, c.entho, c.smallc2);
  if (c.riemann == kHLL) return riemann_hll(L, Rt, c.gamma, c.smallr, c.smallp, c.entho);
  if (c.riemann == kACOUSTIC) return riemann_acoustic(L, Rt, c.gamma, c.smallr, c.smallp, c.entho, c.smallc);
  if (c.riemann == kEXACT) return riemann_exact(L, Rt, c.gamma, c.smallr, c.smallp, c.entho, c.smallc, c.niter);
  return riemann_llf(L, Rt, c.gamma, c.smallr, c.smallp, c.entho);
}

template <typename R, int RS>
__device__ __forceinline__ Flux6<R> riemann_fixed(const Face<R>& L, const Face<R>& Rt, const SolverConsts<R>& c) {
  if constexpr (RS == kHLLC) return riemann_hllc(L, Rt, c.gamma, c.smallr, c.smallp, c.entho, c.smallc2);
  else if constexpr (RS == kHLL) return riemann_hll(L, Rt, c.gamma, c.smallr, c.smallp, c.entho);
  else if constexpr (RS == kDynamicRiemann) return riemann_solve(L, Rt, c);
  else return riemann_llf(L, Rt, c.gamma, c.smallr, c.smallp, c.entho);
}

// Conservative-variable flux of a face in (rho, mx, my, mz, E) order
// (cmpflxm unpacking, hydro/umuscl.f90:809-842).
template <typename R>
__device__ __forceinline__ void unrotate(const Flux6<R>& g, int d, R out[5]) {
  out[0] = g.f[0];
  out[4] = g.f[2];
  if (d == 0) { out[1] = g.f[1]; out[2] = g.f[3]; out[3] = g.f[4]; }
  else if (d == 1) { out[2] = g.f[1]; out[1] = g.f[3]; out[3] = g.f[4]; }
  else { out[3] = g.f[1]; out[1] = g.f[3]; out[2] = g.f[4]; }
}

// cmpdt for one cell (hydro/godunov_utils.f90:5-120): the per-cell dt limit
// dx/w*(sqrt(1+2*C*r)-1)/r with w = 3*c + |vx|+|vy|+|vz| and
// r = max(|g|_1*dx/w^2, 1e-4).  Always FP64.
__device__ __forceinline__ double cmpdt_cell(double u0, double u1, double u2, double u3, double u4,
                                             double gx, double gy, double gz, double dx,
                                             const HydroParams& hp) {
  const double gamma = hp.gamma;
  const double smallp = hp.smallc * hp.smallc / gamma;
  const double rho = fmax_f(u0, hp.smallr);
  const double vx = u1 / rho, vy = u2 / rho, vz = u3 / rho;
  double e = u4;
  e = e - 0.5 * rho * (vx * vx);
  e = e - 0.5 * rho * (vy * vy);
  e = e - 0.5 * rho * (vz * vz);
  double c = fmax_f((gamma - 1.0) * e, rho * smallp);
  c = gamma * c;
  c = __builtin_sqrt(c / rho);
  double w = 3.0 * c;
  w = w + __builtin_fabs(vx);
  w = w + __builtin_fabs(vy);
  w = w + __builtin_fabs(vz);
  double r = 0.0;
  r = r + __builtin_fabs(gx);
  r = r + __builtin_fabs(gy);
  r = r + __builtin_fabs(gz);
  r = r * dx / (w * w);
  r = fmax_f(r, 0.0001);
  return dx / w * (__builtin_sqrt(1.0 + 2.0 * hp.courant_factor * r) - 1.0) / r;
}

// Without gravity cmpdt's gravity ratio is max(0*dx/w^2, 1e-4) = 1e-4 for
// every cell, so the per-cell limit dx/w*(sqrt(1+2C*1e-4)-1)/1e-4 is a
// non-increasing function of the wave speed w alone (each IEEE operation is
// monotone), and min over cells equals its value at max over cells of w.
// cmpdt_w is the per-cell part (hydro/godunov_utils.f90:24-97), cmpdt_from_w
// the rest (:108-118), bitwise the reference when combined.
__device__ __forceinline__ double cmpdt_w(double u0, double u1, double u2, double u3, double u4,
                                          const HydroParams& hp) {
  const double gamma = hp.gamma;
  const double smallp = hp.smallc * hp.smallc / gamma;
  const double rho = fmax_f(u0, hp.smallr);
  const double vx = u1 / rho, vy = u2 / rho, vz = u3 / rho;
  double e = u4;
  e = e - 0.5 * rho * (vx * vx);
  e = e - 0.5 * rho * (vy * vy);
  e = e - 0.5 * rho * (vz * vz);
  double c = fmax_f((gamma - 1.0) * e, rho * smallp);
  c = gamma * c;
  c = __builtin_sqrt(c / rho);
  double w = 3.0 * c;
  w = w + __builtin_fabs(vx);
  w = w + __builtin_fabs(vy);
  w = w + __builtin_fabs(vz);
  return w;
}

// cmpdt_w of the FP32-mixed path: the same FP64 expression with the
// divisions and the square root to about 1 ulp (hardware reciprocal and
// reciprocal square root refined by Newton steps) instead of correctly
// rounded.  The dt stays an FP64 reduction; only the FP64 path needs the
// reference's bits.
__device__ __forceinline__ double sqrt_fast(double x) {  // x > 0, normal
  const double y = __builtin_amdgcn_rsq(x);
  double s = x * y;
  double h = 0.5 * y;
  double r = __builtin_fma(-s, h, 0.5);
  s = __builtin_fma(s, r, s);
  h = __builtin_fma(h, r, h);
  r = __builtin_fma(-s, s, x);
  return __builtin_fma(r, h, s);
}

__device__ __forceinline__ double cmpdt_w_fast(double u0, double u1, double u2, double u3, double u4,
                                               const HydroParams& hp) {
  const double gamma = hp.gamma;
  const double smallp = hp.smallc * hp.smallc / gamma;
  const double rho = fmax_f(u0, hp.smallr);
  const double ir = rcp_fast(rho);
  const double vx = u1 * ir, vy = u2 * ir, vz = u3 * ir;
  double e = u4;
  e = e - 0.5 * rho * (vx * vx);
  e = e - 0.5 * rho * (vy * vy);
  e = e - 0.5 * rho * (vz * vz);
  double c = fmax_f((gamma - 1.0) * e, rho * smallp);
  c = gamma * c;
  c = sqrt_fast(c * ir);
  double w = 3.0 * c;
  w = w + __builtin_fabs(vx);
  w = w + __builtin_fabs(vy);
  w = w + __builtin_fabs(vz);
  return w;
}

__device__ __forceinline__ double cmpdt_from_w(double w, double dx, const HydroParams& hp) {
  double r = 0.0;
  r = r * dx / (w * w);
  r = fmax_f(r, 0.0001);
  return dx / w * (__builtin_sqrt(1.0 + 2.0 * hp.courant_factor * r) - 1.0) / r;
}

// courant_fine's conservation sums over the leaf cells of a level
// (hydro/courant_fine.f90:95-119): mass += rho*vol, ekin += E*vol (the
// reference keeps the total energy under the name ekin), eint += E*vol minus
// 0.5*m_d^2/max(rho,smallr)*vol for d = x, y, z.  Accumulated per thread,
// reduced per workgroup in a fixed order (cons_block_store) and summed over
// workgroups in block order by cons_finish, so the result is the same on
// every run; its rounding differs from the reference's serial order.
struct ConsSums {
  double m, e, ei;
};

__device__ __forceinline__ void cons_add(ConsSums& s, double u0, double u1, double u2, double u3, double u4,
                                         double vol, double smallr) {
  s.m = s.m + u0 * vol;
  s.e = s.e + u4 * vol;
  const double r = fmax_f(u0, smallr);
  double ei = u4 * vol;
  ei = ei - 0.5 * (u1 * u1) / r * vol;
  ei = ei - 0.5 * (u2 * u2) / r * vol;
  ei = ei - 0.5 * (u3 * u3) / r * vol;
  s.ei = s.ei + ei;
}

// Workgroup reduction of the sums in a fixed order (xor shuffles inside each
// wavefront, then the wavefronts in order); thread 0 writes the three sums
// of workgroup `block` to partials[3*block .. 3*block+2].  `lds` holds 3
// doubles per wavefront.  Every thread of the workgroup must call it.
__device__ __forceinline__ void cons_block_store(ConsSums s, double* partials, long block, double* lds) {
  for (int off = 32; off > 0; off >>= 1) {
    s.m = s.m + __shfl_xor(s.m, off, 64);
    s.e = s.e + __shfl_xor(s.e, off, 64);
    s.ei = s.ei + __shfl_xor(s.ei, off, 64);
  }
  const int wave = threadIdx.x >> 6, nw = (int)((blockDim.x + 63) >> 6);
  if ((threadIdx.x & 63) == 0) {
    lds[3 * wave] = s.m;
    lds[3 * wave + 1] = s.e;
    lds[3 * wave + 2] = s.ei;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double m = lds[0], e = lds[1], ei = lds[2];
    for (int w = 1; w < nw; ++w) {
      m = m + lds[3 * w];
      e = e + lds[3 * w + 1];
      ei = ei + lds[3 * w + 2];
    }
    partials[3 * block] = m;
    partials[3 * block + 1] = e;
    partials[3 * block + 2] = ei;
  }
}

// z planes per workgroup of a batched level slot (n a multiple of 16): a
// divisor of n, at most 32.
__host__ __device__ inline int batch_tz(int n) { return n <= 32 ? n : (n % 32 == 0 ? 32 : 16); }

// Positive doubles order like their bit patterns, so a min-reduction of dt
// can use an integer atomic.
__device__ __forceinline__ unsigned long long dt_bits(double x) {
  return (unsigned long long)__double_as_longlong(x);
}

}  // namespace ramses
