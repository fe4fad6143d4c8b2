/*
 * TEST INFRASTRUCTURE ONLY -- CPU restatement (the oracle) of the RAMSES
 * unsplit MUSCL-Hancock hydro step.  It is the checker for the HIP product
 * path and never part of it.  Every routine cites the reference routine it
 * follows; expressions keep the reference's left-to-right evaluation order so
 * that, compiled with -ffp-contract=off, each value rounds exactly as the
 * reference's x86-64 build rounds it.
 *
 * Constants of the reference's `const` module (hydro/hydro_commons.f90:14-27):
 * zero=0, one=1, two=2, half=1/2.
 */
#include "hydro_ref.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define NVAR 5
#define NDIM 3
#define NC6 216 /* 6x6x6 stencil cells */
#define NF3 27  /* 3x3x3 face slots */

static hr_params P = {1.4, 1e-10, 1e-10, 1.5, 0.5, 0.0, 1, HR_RIEMANN_LLF, 10};

static const double zero = 0.0, one = 1.0, two = 2.0, half = 0.5;

/* Fortran MAX/MIN of two reals */
static inline double fmaxr(double a, double b) { return (a > b) ? a : b; }
static inline double fminr(double a, double b) { return (a < b) ? a : b; }

int hr_set_params(const hr_params *p) {
  if (!p) return -1;
  if (!(p->slope_type == 0 || p->slope_type == 1 || p->slope_type == 2 || p->slope_type == 3 ||
        p->slope_type == 7 || p->slope_type == 8))
    return -2; /* hydro/umuscl.f90:1474-1477 stops on other 3D slope types */
  if (p->riemann < 0 || p->riemann > 4) return -3;
  if (p->niter_riemann < 0) return -6;
  if (p->difmag != 0.0) return -4; /* cmpdivu/consup not restated */
  if (!(p->gamma > 1.0)) return -5;
  P = *p;
  return 0;
}

/* index of cell (i,j,k) in a (nv,-1:4,-1:4,-1:4) array, lane l (0-based) */
#define IC(l, i, j, k) ((size_t)(l) + (size_t)nv * (size_t)(((i) + 1) + 6 * (((j) + 1) + 6 * ((k) + 1))))
/* index of face slot (i,j,k) in a (nv,1:3,1:3,1:3) array */
#define IF(l, i, j, k) ((size_t)(l) + (size_t)nv * (size_t)(((i)-1) + 3 * (((j)-1) + 3 * ((k)-1))))

/* ---------------------------------------------------------------------- */
/* ctoprim: hydro/umuscl.f90:861-965                                        */
static void ctoprim(const double *uin, const double *gravin, double *q, double *c, double dt,
                    int ngrid, int nv) {
  const size_t vs = (size_t)nv * NC6;
  const double gamma = P.gamma;
  const double smalle = P.smallc * P.smallc / gamma / (gamma - one); /* :884 */
  const double dtxhalf = dt * half;
  for (int k = -1; k <= 4; ++k)
    for (int j = -1; j <= 4; ++j)
      for (int i = -1; i <= 4; ++i)
        for (int l = 0; l < ngrid; ++l) {
          const size_t o = IC(l, i, j, k);
          double r = fmaxr(uin[o], P.smallr); /* :894 */
          q[o] = r;
          double oneoverrho = one / r; /* :897 */
          double vx = uin[o + vs] * oneoverrho;
          double vy = uin[o + 2 * vs] * oneoverrho;
          double vz = uin[o + 3 * vs] * oneoverrho;
          double eken = half * vx * vx; /* :907-913 */
          eken = eken + half * vy * vy;
          eken = eken + half * vz * vz;
          double erad = zero;
          double eint = fmaxr(uin[o + 4 * vs] * oneoverrho - eken - erad, smalle); /* :923 */
          double p = (gamma - one) * r * eint;
          q[o + 4 * vs] = p;
          double cc = gamma * p; /* :927-933 */
          c[o] = sqrt(cc * oneoverrho);
          double gx = gravin ? gravin[o] : 0.0;
          double gy = gravin ? gravin[o + vs] : 0.0;
          double gz = gravin ? gravin[o + 2 * vs] : 0.0;
          q[o + vs] = vx + gx * dtxhalf; /* :936-942 */
          q[o + 2 * vs] = vy + gy * dtxhalf;
          q[o + 3 * vs] = vz + gz * dtxhalf;
        }
}

/* ---------------------------------------------------------------------- */
/* uslope, NDIM==3 branches: hydro/umuscl.f90:970-1480 (3D at :1240-1478)  */
static void uslope(const double *q, double *dq, int ngrid, int nv) {
  const size_t vs = (size_t)nv * NC6;
  const size_t ds = vs * NVAR;
  const int st = P.slope_type;
  if (st == 0) { /* :1002-1005 */
    memset(dq, 0, sizeof(double) * ds * NDIM);
    return;
  }
  const int sx = 1, sy = 6, sz = 36; /* strides in cells */
  const int strides[3] = {sx, sy, sz};
  for (int n = 0; n < NVAR; ++n)
    for (int k = 0; k <= 3; ++k)
      for (int j = 0; j <= 3; ++j)
        for (int i = 0; i <= 3; ++i)
          for (int l = 0; l < ngrid; ++l) {
            const size_t o = IC(l, i, j, k) + n * vs;
            if (st == 3) { /* positivity preserving 27-point slope :1328-1392 */
              const double qc = q[o];
              double df[27];
              int m = 0;
              /* order: k-1 plane (l,m,r in i) x (l,m,r in j) ... as in the reference */
              for (int dk = -1; dk <= 1; ++dk)
                for (int di = -1; di <= 1; ++di)
                  for (int dj = -1; dj <= 1; ++dj)
                    df[m++] = q[o + (size_t)nv * (di * sx + dj * sy + dk * sz)] - qc;
              double vmin = df[0], vmax = df[0];
              for (int t = 1; t < 27; ++t) {
                vmin = fminr(vmin, df[t]);
                vmax = fmaxr(vmax, df[t]);
              }
              double dfx = half * (q[o + (size_t)nv * sx] - q[o - (size_t)nv * sx]);
              double dfy = half * (q[o + (size_t)nv * sy] - q[o - (size_t)nv * sy]);
              double dfz = half * (q[o + (size_t)nv * sz] - q[o - (size_t)nv * sz]);
              double dff = half * (fabs(dfx) + fabs(dfy) + fabs(dfz));
              double slop;
              if (dff > zero)
                slop = fminr(one, fminr(fabs(vmin), fabs(vmax)) / dff);
              else
                slop = one;
              double dlim = slop;
              dq[o + 0 * ds] = dlim * dfx;
              dq[o + 1 * ds] = dlim * dfy;
              dq[o + 2 * ds] = dlim * dfz;
              continue;
            }
            for (int d = 0; d < NDIM; ++d) {
              const size_t s = (size_t)nv * strides[d];
              const double qm1 = q[o - s], q0 = q[o], qp1 = q[o + s];
              double res;
              if (st == 1) { /* minmod :1241-1285 */
                double dlft = q0 - qm1;
                double drgt = qp1 - q0;
                if ((dlft * drgt) <= zero)
                  res = zero;
                else if (dlft > 0)
                  res = fminr(dlft, drgt);
                else
                  res = fmaxr(dlft, drgt);
              } else if (st == 2) { /* moncen :1286-1327 (slope_type promoted to real) */
                const double stp = 2.0;
                double dlft = stp * (q0 - qm1);
                double drgt = stp * (qp1 - q0);
                double dcen = half * (dlft + drgt) / stp;
                double dsgn = copysign(one, dcen);
                double slop = fminr(fabs(dlft), fabs(drgt));
                double dlim = slop;
                if ((dlft * drgt) <= zero) dlim = zero;
                res = dsgn * fminr(dlim, fabs(dcen));
              } else if (st == 7) { /* van Leer :1393-1431 */
                double dlft = (q0 - qm1);
                double drgt = (qp1 - q0);
                if ((dlft * drgt) <= zero)
                  res = zero;
                else
                  res = (2.0 * dlft * drgt / (dlft + drgt));
              } else { /* st == 8, generalised moncen/minmod :1432-1473 */
                double dlft = (q0 - qm1);
                double drgt = (qp1 - q0);
                double dcen = half * (dlft + drgt);
                double dsgn = copysign(one, dcen);
                double slop = fminr(P.slope_theta * fabs(dlft), P.slope_theta * fabs(drgt));
                double dlim = slop;
                if ((dlft * drgt) <= zero) dlim = zero;
                res = dsgn * fminr(dlim, fabs(dcen));
              }
              dq[o + d * ds] = res;
            }
          }
}

/* ---------------------------------------------------------------------- */
/* trace3d: hydro/umuscl.f90:483-708                                        */
static void trace3d(const double *q, const double *dq, double *qm, double *qp, double dx,
                    double dy, double dz, double dt, int ngrid, int nv) {
  const size_t vs = (size_t)nv * NC6;
  const size_t ds = vs * NVAR;
  const double gamma = P.gamma, smallr = P.smallr;
  const double dtdx = dt / dx, dtdy = dt / dy, dtdz = dt / dz;
  enum { ir = 0, iu = 1, iv = 2, iw = 3, ip = 4 };
  for (int k = 0; k <= 3; ++k)
    for (int j = 0; j <= 3; ++j)
      for (int i = 0; i <= 3; ++i)
        for (int l = 0; l < ngrid; ++l) {
          const size_t o = IC(l, i, j, k);
          const double r = q[o + ir * vs], u = q[o + iu * vs], v = q[o + iv * vs];
          const double w = q[o + iw * vs], p = q[o + ip * vs];
#define DQ(var, d) dq[o + (var)*vs + (d)*ds]
          const double drx = DQ(ir, 0), dpx = DQ(ip, 0), dux = DQ(iu, 0), dvx = DQ(iv, 0), dwx = DQ(iw, 0);
          const double dry = DQ(ir, 1), dpy = DQ(ip, 1), duy = DQ(iu, 1), dvy = DQ(iv, 1), dwy = DQ(iw, 1);
          const double drz = DQ(ir, 2), dpz = DQ(ip, 2), duz = DQ(iu, 2), dvz = DQ(iv, 2), dwz = DQ(iw, 2);
#undef DQ
          /* source terms :576-580 */
          const double sr0 = -u * drx - v * dry - w * drz - (dux + dvy + dwz) * r;
          const double sp0 = -u * dpx - v * dpy - w * dpz - (dux + dvy + dwz) * gamma * p;
          const double su0 = -u * dux - v * duy - w * duz - (dpx) / r;
          const double sv0 = -u * dvx - v * dvy - w * dvz - (dpy) / r;
          const double sw0 = -u * dwx - v * dwy - w * dwz - (dpz) / r;
#define QP(var, d) qp[o + (var)*vs + (d)*ds]
#define QM(var, d) qm[o + (var)*vs + (d)*ds]
          /* x :592-617 */
          QP(ir, 0) = r - half * drx + sr0 * dtdx * half;
          QP(ip, 0) = p - half * dpx + sp0 * dtdx * half;
          QP(iu, 0) = u - half * dux + su0 * dtdx * half;
          QP(iv, 0) = v - half * dvx + sv0 * dtdx * half;
          QP(iw, 0) = w - half * dwx + sw0 * dtdx * half;
          if (QP(ir, 0) < smallr) QP(ir, 0) = r;
          QM(ir, 0) = r + half * drx + sr0 * dtdx * half;
          QM(ip, 0) = p + half * dpx + sp0 * dtdx * half;
          QM(iu, 0) = u + half * dux + su0 * dtdx * half;
          QM(iv, 0) = v + half * dvx + sv0 * dtdx * half;
          QM(iw, 0) = w + half * dwx + sw0 * dtdx * half;
          if (QM(ir, 0) < smallr) QM(ir, 0) = r;
          /* y :619-645 */
          QP(ir, 1) = r - half * dry + sr0 * dtdy * half;
          QP(ip, 1) = p - half * dpy + sp0 * dtdy * half;
          QP(iu, 1) = u - half * duy + su0 * dtdy * half;
          QP(iv, 1) = v - half * dvy + sv0 * dtdy * half;
          QP(iw, 1) = w - half * dwy + sw0 * dtdy * half;
          if (QP(ir, 1) < smallr) QP(ir, 1) = r;
          QM(ir, 1) = r + half * dry + sr0 * dtdy * half;
          QM(ip, 1) = p + half * dpy + sp0 * dtdy * half;
          QM(iu, 1) = u + half * duy + su0 * dtdy * half;
          QM(iv, 1) = v + half * dvy + sv0 * dtdy * half;
          QM(iw, 1) = w + half * dwy + sw0 * dtdy * half;
          if (QM(ir, 1) < smallr) QM(ir, 1) = r;
          /* z :647-673 */
          QP(ir, 2) = r - half * drz + sr0 * dtdz * half;
          QP(ip, 2) = p - half * dpz + sp0 * dtdz * half;
          QP(iu, 2) = u - half * duz + su0 * dtdz * half;
          QP(iv, 2) = v - half * dvz + sv0 * dtdz * half;
          QP(iw, 2) = w - half * dwz + sw0 * dtdz * half;
          if (QP(ir, 2) < smallr) QP(ir, 2) = r;
          QM(ir, 2) = r + half * drz + sr0 * dtdz * half;
          QM(ip, 2) = p + half * dpz + sp0 * dtdz * half;
          QM(iu, 2) = u + half * duz + su0 * dtdz * half;
          QM(iv, 2) = v + half * dvz + sv0 * dtdz * half;
          QM(iw, 2) = w + half * dwz + sw0 * dtdz * half;
          if (QM(ir, 2) < smallr) QM(ir, 2) = r;
#undef QP
#undef QM
        }
}

/* ---------------------------------------------------------------------- */
/* Riemann solvers, one interface.  ql/qr = (rho, u_n, p, u_t1, u_t2);
 * fg[0..5] = (mass, normal momentum, energy, t1 mom, t2 mom, e_int flux)   */

/* riemann_llf: hydro/godunov_utils.f90:660-820 */
static void riemann_llf(const double *ql, const double *qr, double *fg) {
  const double gamma = P.gamma;
  const double smallp = P.smallc * P.smallc / gamma; /* :683 */
  const double entho = one / (gamma - one);
  double rl = fmaxr(ql[0], P.smallr), ul = ql[1], pl = fmaxr(ql[2], rl * smallp);
  double cl = gamma * pl;
  cl = sqrt(cl / rl);
  double rr = fmaxr(qr[0], P.smallr), ur = qr[1], pr = fmaxr(qr[2], rr * smallp);
  double cr = gamma * pr;
  cr = sqrt(cr / rr);
  double cmax = fmaxr(fabs(ul) + cl, fabs(ur) + cr); /* :713 */
  double uL[6], uR[6], fL[6], fR[6];
  uL[0] = ql[0];
  uR[0] = qr[0];
  uL[1] = ql[0] * ql[1];
  uR[1] = qr[0] * qr[1];
  uL[2] = ql[2] * entho + half * ql[0] * (ql[1] * ql[1]); /* :727-736 */
  uR[2] = qr[2] * entho + half * qr[0] * (qr[1] * qr[1]);
  uL[2] = uL[2] + half * ql[0] * (ql[3] * ql[3]);
  uR[2] = uR[2] + half * qr[0] * (qr[3] * qr[3]);
  uL[2] = uL[2] + half * ql[0] * (ql[4] * ql[4]);
  uR[2] = uR[2] + half * qr[0] * (qr[4] * qr[4]);
  uL[3] = ql[0] * ql[3]; /* :746-751 */
  uR[3] = qr[0] * qr[3];
  uL[4] = ql[0] * ql[4];
  uR[4] = qr[0] * qr[4];
  uL[5] = ql[2] * entho; /* :772-775 */
  uR[5] = qr[2] * entho;
  fL[0] = ql[1] * uL[0]; /* :780-809 */
  fR[0] = qr[1] * uR[0];
  fL[1] = ql[1] * uL[1] + ql[2];
  fR[1] = qr[1] * uR[1] + qr[2];
  fL[2] = ql[1] * (uL[2] + ql[2]);
  fR[2] = qr[1] * (uR[2] + qr[2]);
  for (int n = 3; n < 6; ++n) {
    fL[n] = ql[1] * uL[n];
    fR[n] = qr[1] * uR[n];
  }
  for (int n = 0; n < 6; ++n) /* :814-818 */
    fg[n] = half * (fL[n] + fR[n] - cmax * (uR[n] - uL[n]));
}

/* riemann_hll: hydro/godunov_utils.f90:825-983 */
static void riemann_hll(const double *ql, const double *qr, double *fg) {
  const double gamma = P.gamma;
  const double smallp = P.smallc * P.smallc / gamma;
  const double entho = one / (gamma - one);
  double rl = fmaxr(ql[0], P.smallr), ul = ql[1], pl = fmaxr(ql[2], rl * smallp);
  double cl = gamma * pl;
  cl = sqrt(cl / rl);
  double rr = fmaxr(qr[0], P.smallr), ur = qr[1], pr = fmaxr(qr[2], rr * smallp);
  double cr = gamma * pr;
  cr = sqrt(cr / rr);
  double SL = fminr(fminr(ul, ur) - fmaxr(cl, cr), zero); /* :874-875 */
  double SR = fmaxr(fmaxr(ul, ur) + fmaxr(cl, cr), zero);
  double uL[6], uR[6], fL[6], fR[6];
  uL[0] = ql[0];
  uR[0] = qr[0];
  uL[1] = ql[0] * ql[1];
  uR[1] = qr[0] * qr[1];
  uL[2] = ql[2] * entho + half * ql[0] * (ql[1] * ql[1]);
  uR[2] = qr[2] * entho + half * qr[0] * (qr[1] * qr[1]);
  uL[2] = uL[2] + half * ql[0] * (ql[3] * ql[3]);
  uR[2] = uR[2] + half * qr[0] * (qr[3] * qr[3]);
  uL[2] = uL[2] + half * ql[0] * (ql[4] * ql[4]);
  uR[2] = uR[2] + half * qr[0] * (qr[4] * qr[4]);
  uL[3] = ql[0] * ql[3];
  uR[3] = qr[0] * qr[3];
  uL[4] = ql[0] * ql[4];
  uR[4] = qr[0] * qr[4];
  uL[5] = ql[2] * entho;
  uR[5] = qr[2] * entho;
  fL[0] = uL[1]; /* :942-971 */
  fR[0] = uR[1];
  fL[1] = ql[2] + uL[1] * ql[1];
  fR[1] = qr[2] + uR[1] * qr[1];
  fL[2] = ql[1] * (uL[2] + ql[2]);
  fR[2] = qr[1] * (uR[2] + qr[2]);
  for (int n = 3; n < 6; ++n) {
    fL[n] = ql[1] * uL[n];
    fR[n] = qr[1] * uR[n];
  }
  for (int n = 0; n < 6; ++n) /* :976-981 */
    fg[n] = (SR * fL[n] - SL * fR[n] + SR * SL * (uR[n] - uL[n])) / (SR - SL);
}

/* riemann_hllc: hydro/godunov_utils.f90:988-1209 */
static void riemann_hllc(const double *ql, const double *qr, double *fg) {
  const double gamma = P.gamma;
  const double smallp = P.smallc * P.smallc / gamma; /* :1017 */
  const double entho = one / (gamma - one);
  double rl = fmaxr(ql[0], P.smallr);
  double Pl = fmaxr(ql[2], rl * smallp);
  double ul = ql[1];
  double el = Pl * entho;
  double ecinl = half * rl * ul * ul;
  ecinl = ecinl + half * rl * (ql[3] * ql[3]);
  ecinl = ecinl + half * rl * (ql[4] * ql[4]);
  double etotl = el + ecinl;
  double Ptotl = Pl;
  double rr = fmaxr(qr[0], P.smallr);
  double Pr = fmaxr(qr[2], rr * smallp);
  double ur = qr[1];
  double er = Pr * entho;
  double ecinr = half * rr * ur * ur;
  ecinr = ecinr + half * rr * (qr[3] * qr[3]);
  ecinr = ecinr + half * rr * (qr[4] * qr[4]);
  double etotr = er + ecinr;
  double Ptotr = Pr;
  double cfastl = gamma * Pl; /* :1077-1091 */
  cfastl = sqrt(fmaxr(cfastl / rl, P.smallc * P.smallc));
  double cfastr = gamma * Pr;
  cfastr = sqrt(fmaxr(cfastr / rr, P.smallc * P.smallc));
  double SL = fminr(ul, ur) - fmaxr(cfastl, cfastr); /* :1094-1095 */
  double SR = fmaxr(ul, ur) + fmaxr(cfastl, cfastr);
  double rcl = rl * (ul - SL);
  double rcr = rr * (SR - ur);
  double ustar = (rcr * ur + rcl * ul + (Ptotl - Ptotr)) / (rcr + rcl); /* :1102-1103 */
  double Ptotstar = (rcr * Ptotl + rcl * Ptotr + rcl * rcr * (ul - ur)) / (rcr + rcl);
  double rstarl = rl * (SL - ul) / (SL - ustar); /* :1106-1118 */
  double etotstarl = ((SL - ul) * etotl - Ptotl * ul + Ptotstar * ustar) / (SL - ustar);
  double estarl = el * (SL - ul) / (SL - ustar);
  double rstarr = rr * (SR - ur) / (SR - ustar);
  double etotstarr = ((SR - ur) * etotr - Ptotr * ur + Ptotstar * ustar) / (SR - ustar);
  double estarr = er * (SR - ur) / (SR - ustar);
  double ro, uo, Ptoto, etoto, eo;
  if (SL > 0.0) { /* :1126-1170 */
    ro = rl; uo = ul; Ptoto = Ptotl; etoto = etotl; eo = el;
  } else if (ustar > 0.0) {
    ro = rstarl; uo = ustar; Ptoto = Ptotstar; etoto = etotstarl; eo = estarl;
  } else if (SR > 0.0) {
    ro = rstarr; uo = ustar; Ptoto = Ptotstar; etoto = etotstarr; eo = estarr;
  } else {
    ro = rr; uo = ur; Ptoto = Ptotr; etoto = etotr; eo = er;
  }
  fg[0] = ro * uo; /* :1175-1205 */
  fg[1] = ro * uo * uo + Ptoto;
  fg[2] = (etoto + Ptoto) * uo;
  for (int iv = 3; iv < 5; ++iv) {
    if (ustar > 0)
      fg[iv] = ro * uo * ql[iv];
    else
      fg[iv] = ro * uo * qr[iv];
  }
  fg[5] = uo * eo;
}

/* The flux of a sampled state q[0..5] = (rho, u_n, p, u_t1, u_t2, e_int)
 * shared by riemann_acoustic and riemann_approx
 * (hydro/godunov_utils.f90:468-492, 631-652). */
static void sampled_flux(const double *q, double *fg) {
  const double entho = one / (P.gamma - one);
  fg[0] = q[0] * q[1];
  fg[1] = q[2] + q[0] * (q[1] * q[1]);
  double etot = q[2] * entho + half * q[0] * (q[1] * q[1]);
  etot = etot + half * q[0] * (q[3] * q[3]);
  etot = etot + half * q[0] * (q[4] * q[4]);
  fg[2] = q[1] * (etot + q[2]);
  for (int n = 3; n < 6; ++n) fg[n] = fg[0] * q[n];
}

/* riemann_acoustic: hydro/godunov_utils.f90:500-655 */
static void riemann_acoustic(const double *ql, const double *qr, double *fg) {
  const double gamma = P.gamma;
  const double smallp = P.smallc * P.smallc / gamma; /* :528 */
  const double entho = one / (gamma - one);
  const double rl = fmaxr(ql[0], P.smallr), ul = ql[1], pl = fmaxr(ql[2], rl * smallp);
  const double rr = fmaxr(qr[0], P.smallr), ur = qr[1], pr = fmaxr(qr[2], rr * smallp);
  const double cl = sqrt(gamma * pl / rl), cr = sqrt(gamma * pr / rr); /* :542-553 */
  const double wl = cl * rl, wr = cr * rr;
  const double pstar = ((wr * pl + wl * pr) + wl * wr * (ul - ur)) / (wl + wr);
  const double ustar = ((wr * ur + wl * ul) + (pl - pr)) / (wl + wr);
  const double sgnm = copysign(one, ustar); /* :557-559 */
  double ro, uo, po, co;
  if (sgnm == one) {
    ro = rl; uo = ul; po = pl; co = cl;
  } else {
    ro = rr; uo = ur; po = pr; co = cr;
  }
  double rstar = ro + (pstar - po) / (co * co); /* :579-584 */
  rstar = fmaxr(rstar, P.smallr);
  double cstar = sqrt(fabs(gamma * pstar / rstar));
  cstar = fmaxr(cstar, P.smallc);
  double spout = co - sgnm * uo, spin = cstar - sgnm * ustar; /* :587-590 */
  double ushock = half * (spin + spout);
  ushock = fmaxr(ushock, -sgnm * ustar);
  if (pstar >= po) {
    spout = ushock;
    spin = spout;
  }
  double q[6];
  if (spout < zero) { /* :605-621 */
    q[0] = ro; q[1] = uo; q[2] = po;
  } else if (spin >= zero) {
    q[0] = rstar; q[1] = ustar; q[2] = pstar;
  } else {
    const double frac = spout / (spout - spin);
    q[0] = frac * rstar + (one - frac) * ro;
    q[1] = frac * ustar + (one - frac) * uo;
    q[2] = frac * pstar + (one - frac) * po;
  }
  q[3] = sgnm == one ? ql[3] : qr[3]; /* :624-634 */
  q[4] = sgnm == one ? ql[4] : qr[4];
  q[5] = po / ro * entho;
  sampled_flux(q, fg);
}

/* riemann_approx ('exact'): hydro/godunov_utils.f90:268-495.  The
 * reference iterates all faces together and drops a face from the Newton
 * loop once its relative pressure change is at most 1e-6; per face that is
 * "update, then stop if converged", up to niter_riemann updates. */
static void riemann_approx(const double *ql, const double *qr, double *fg) {
  const double gamma = P.gamma;
  const double smallp = P.smallc * P.smallc / gamma; /* :295-298 */
  const double smallpp = P.smallr * smallp;
  const double gamma6 = (gamma + one) / (two * gamma);
  const double entho = one / (gamma - one);
  const double rl = fmaxr(ql[0], P.smallr), ul = ql[1], pl = fmaxr(ql[2], rl * smallp);
  const double rr = fmaxr(qr[0], P.smallr), ur = qr[1], pr = fmaxr(qr[2], rr * smallp);
  const double cl = gamma * pl * rl, cr = gamma * pr * rr; /* :311-314 */
  double wl = sqrt(cl), wr = sqrt(cr);
  double pstar = ((wr * pl + wl * pr) + wl * wr * (ul - ur)) / (wl + wr);
  pstar = fmaxr(pstar, 0.0);
  double pold = pstar;
  for (int iter = 1; iter <= P.niter_riemann; ++iter) { /* :328-363 */
    const double wwl = sqrt(cl * (one + gamma6 * (pold - pl) / pl));
    const double wwr = sqrt(cr * (one + gamma6 * (pold - pr) / pr));
    const double ql_ = two * (wwl * wwl * wwl) / (wwl * wwl + cl);
    const double qr_ = two * (wwr * wwr * wwr) / (wwr * wwr + cr);
    const double usl = ul - (pold - pl) / wwl;
    const double usr = ur + (pold - pr) / wwr;
    const double delp = fmaxr(qr_ * ql_ / (qr_ + ql_) * (usl - usr), -pold);
    pold = pold + delp;
    const double uo = fabs(delp / (pold + smallpp));
    if (!(uo > 1e-06)) break;
  }
  pstar = pold; /* :367-374 */
  wl = sqrt(cl * (one + gamma6 * (pstar - pl) / pl));
  wr = sqrt(cr * (one + gamma6 * (pstar - pr) / pr));
  const double ustar = half * (ul + (pl - pstar) / wl + ur - (pr - pstar) / wr); /* :378-381 */
  const double sgnm = copysign(one, ustar);
  double ro, uo, po, wo;
  if (sgnm == one) {
    ro = rl; uo = ul; po = pl; wo = wl;
  } else {
    ro = rr; uo = ur; po = pr; wo = wr;
  }
  const double co = fmaxr(P.smallc, sqrt(fabs(gamma * po / ro))); /* :403-405 */
  double rstar;
  if (pstar >= po)
    rstar = ro / (one + ro * (po - pstar) / (wo * wo)); /* shock */
  else
    rstar = ro * pow(pstar / po, one / gamma); /* rarefaction */
  rstar = fmaxr(rstar, P.smallr); /* :419-430 */
  double cstar = sqrt(fabs(gamma * pstar / rstar));
  cstar = fmaxr(cstar, P.smallc);
  double spout = co - sgnm * uo, spin = cstar - sgnm * ustar;
  const double ushock = wo / ro - sgnm * uo;
  if (pstar >= po) {
    spout = ushock;
    spin = spout;
  }
  double q[6];
  if (spout <= zero) { /* :440-455 */
    q[0] = ro; q[1] = uo; q[2] = po;
  } else if (spin >= zero) {
    q[0] = rstar; q[1] = ustar; q[2] = pstar;
  } else {
    const double frac = spout / (spout - spin);
    q[1] = frac * ustar + (one - frac) * uo;
    q[2] = frac * pstar + (one - frac) * po;
    q[0] = ro * pow(q[2] / po, one / gamma);
  }
  q[3] = sgnm == one ? ql[3] : qr[3]; /* :458-468 */
  q[4] = sgnm == one ? ql[4] : qr[4];
  q[5] = po / ro * entho;
  sampled_flux(q, fg);
}

static inline void riemann(const double *ql, const double *qr, double *fg) {
  switch (P.riemann) { /* hydro/umuscl.f90:791-804 */
  case HR_RIEMANN_HLLC: riemann_hllc(ql, qr, fg); break;
  case HR_RIEMANN_HLL: riemann_hll(ql, qr, fg); break;
  case HR_RIEMANN_ACOUSTIC: riemann_acoustic(ql, qr, fg); break;
  case HR_RIEMANN_EXACT: riemann_approx(ql, qr, fg); break;
  default: riemann_llf(ql, qr, fg); break;
  }
}

/* ---------------------------------------------------------------------- */
/* cmpflxm + the flux scaling of unsplit: hydro/umuscl.f90:714-856, 96-164.
 * Direction d (0=x,1=y,2=z): face (i,j,k) takes qm of the cell one step
 * down along d and qp of cell (i,j,k).  ln/lt1/lt2 are 0-based variable
 * indices of the normal and the two tangential velocities.                 */
static void cmpflxm_dir(const double *qm, const double *qp, double *flux, double *tmp, int d,
                        double dt, double dxd, int ngrid, int nv) {
  const size_t vs = (size_t)nv * NC6;
  const size_t ds = vs * NVAR;
  const size_t fvs = (size_t)nv * NF3; /* flux var stride */
  const int ln = 1 + d, lt1 = (d == 0) ? 2 : 1, lt2 = (d == 2) ? 2 : 3;
  const int cstride[3] = {1, 6, 36};
  const int ilo = (d == 0) ? 1 : 1, ihi = (d == 0) ? 3 : 2;
  const int jlo = 1, jhi = (d == 1) ? 3 : 2;
  const int klo = 1, khi = (d == 2) ? 3 : 2;
  (void)ilo;
  for (int k = klo; k <= khi; ++k)
    for (int j = jlo; j <= jhi; ++j)
      for (int i = 1; i <= ihi; ++i)
        for (int l = 0; l < ngrid; ++l) {
          const size_t oR = IC(l, i, j, k);
          const size_t oL = oR - (size_t)nv * cstride[d];
          double qleft[5], qright[5], fg[6];
          qleft[0] = qm[oL + 0 * vs + d * ds];
          qright[0] = qp[oR + 0 * vs + d * ds];
          qleft[1] = qm[oL + ln * vs + d * ds];
          qright[1] = qp[oR + ln * vs + d * ds];
          qleft[2] = qm[oL + 4 * vs + d * ds];
          qright[2] = qp[oR + 4 * vs + d * ds];
          qleft[3] = qm[oL + lt1 * vs + d * ds];
          qright[3] = qp[oR + lt1 * vs + d * ds];
          qleft[4] = qm[oL + lt2 * vs + d * ds];
          qright[4] = qp[oR + lt2 * vs + d * ds];
          riemann(qleft, qright, fg);
          double flx[5];
          flx[0] = fg[0];
          flx[ln] = fg[1];
          flx[lt1] = fg[3];
          flx[lt2] = fg[4];
          flx[4] = fg[2];
          const size_t fo = IF(l, i, j, k);
          for (int v = 0; v < NVAR; ++v)
            flux[fo + v * fvs + (size_t)d * fvs * NVAR] = flx[v] * dt / dxd; /* :106 */
          double t1 = half * (qleft[1] + qright[1]); /* :845 */
          double t2 = fg[5];
          tmp[fo + 0 * fvs + (size_t)d * fvs * 2] = t1 * dt / dxd; /* :111 */
          tmp[fo + 1 * fvs + (size_t)d * fvs * 2] = t2 * dt / dxd;
        }
}

typedef struct {
  double *q, *c, *dq, *qm, *qp;
  int nv;
} hr_scratch;

static int scratch_alloc(hr_scratch *s, int nv) {
  const size_t n = (size_t)nv * NC6;
  s->nv = nv;
  s->q = (double *)calloc(n * NVAR, sizeof(double));
  s->c = (double *)calloc(n, sizeof(double));
  s->dq = (double *)calloc(n * NVAR * NDIM, sizeof(double));
  s->qm = (double *)calloc(n * NVAR * NDIM, sizeof(double));
  s->qp = (double *)calloc(n * NVAR * NDIM, sizeof(double));
  return (s->q && s->c && s->dq && s->qm && s->qp) ? 0 : -1;
}
static void scratch_free(hr_scratch *s) {
  free(s->q); free(s->c); free(s->dq); free(s->qm); free(s->qp);
}

/* unsplit: hydro/umuscl.f90:22-171 (scheme='muscl', difmag=0) */
static void unsplit_s(hr_scratch *s, const double *uin, const double *gravin, double *flux,
                      double *tmp, double dx, double dy, double dz, double dt, int ngrid) {
  const int nv = s->nv;
  ctoprim(uin, gravin, s->q, s->c, dt, ngrid, nv);           /* :67 */
  uslope(s->q, s->dq, ngrid, nv);                            /* :70 */
  trace3d(s->q, s->dq, s->qm, s->qp, dx, dy, dz, dt, ngrid, nv); /* :81 */
  cmpflxm_dir(s->qm, s->qp, flux, tmp, 0, dt, dx, ngrid, nv);    /* :97-116 */
  cmpflxm_dir(s->qm, s->qp, flux, tmp, 1, dt, dy, ngrid, nv);    /* :120-139 */
  cmpflxm_dir(s->qm, s->qp, flux, tmp, 2, dt, dz, ngrid, nv);    /* :144-163 */
}

int hr_unsplit(const double *uin, const double *gravin, double *flux, double *tmp, double dx,
               double dy, double dz, double dt, int ngrid, int nv) {
  if (ngrid < 0 || ngrid > nv || nv <= 0) return -1;
  hr_scratch s;
  if (scratch_alloc(&s, nv)) { scratch_free(&s); return -2; }
  unsplit_s(&s, uin, gravin, flux, tmp, dx, dy, dz, dt, ngrid);
  scratch_free(&s);
  return 0;
}

int hr_unsplit_batched(const double *uin, const double *gravin, double *flux, double *tmp,
                       double dx, double dt, long ngroups, int ngrid_last, int nv, int nthreads) {
  if (ngroups < 0 || nv <= 0 || ngrid_last < 0 || ngrid_last > nv) return -1;
  const size_t us = (size_t)nv * NC6 * NVAR, gs = (size_t)nv * NC6 * NDIM;
  const size_t fs = (size_t)nv * NF3 * NVAR * NDIM, ts = (size_t)nv * NF3 * 2 * NDIM;
  int err = 0;
#ifdef _OPENMP
  if (nthreads <= 0) nthreads = omp_get_max_threads();
#pragma omp parallel num_threads(nthreads) reduction(| : err)
#endif
  {
    hr_scratch s;
    if (scratch_alloc(&s, nv)) {
      err |= 1;
    } else {
#ifdef _OPENMP
#pragma omp for schedule(static)
#endif
      for (long g = 0; g < ngroups; ++g) {
        int ng = (g == ngroups - 1) ? ngrid_last : nv;
        unsplit_s(&s, uin + g * us, gravin ? gravin + g * gs : NULL, flux + g * fs, tmp + g * ts,
                  dx, dx, dx, dt, ng);
      }
    }
    scratch_free(&s);
  }
  (void)nthreads;
  return err ? -2 : 0;
}

/* ---------------------------------------------------------------------- */
/* cmpdt: hydro/godunov_utils.f90:5-120                                     */
double hr_cmpdt(double *uu, const double *gg, double dx, int ncell, int nv) {
  const double gamma = P.gamma;
  const double smallp = P.smallc * P.smallc / gamma; /* :21 */
  const size_t s = (size_t)nv;
  for (int k = 0; k < ncell; ++k) uu[k] = fmaxr(uu[k], P.smallr); /* :24-26 */
  for (int d = 0; d < NDIM; ++d)                                 /* :28-32 */
    for (int k = 0; k < ncell; ++k) uu[k + (d + 1) * s] = uu[k + (d + 1) * s] / uu[k];
  for (int d = 0; d < NDIM; ++d) /* :34-38 */
    for (int k = 0; k < ncell; ++k)
      uu[k + 4 * s] = uu[k + 4 * s] - half * uu[k] * (uu[k + (d + 1) * s] * uu[k + (d + 1) * s]);
  for (int k = 0; k < ncell; ++k) /* :63-65 */
    uu[k + 4 * s] = fmaxr((gamma - one) * uu[k + 4 * s], uu[k] * smallp);
  for (int k = 0; k < ncell; ++k) uu[k + 4 * s] = gamma * uu[k + 4 * s]; /* :75-77 */
  for (int k = 0; k < ncell; ++k) uu[k + 4 * s] = sqrt(uu[k + 4 * s] / uu[k]); /* :85-87 */
  for (int k = 0; k < ncell; ++k) uu[k + 4 * s] = (double)NDIM * uu[k + 4 * s]; /* :90-92 */
  for (int d = 0; d < NDIM; ++d) /* :93-97 */
    for (int k = 0; k < ncell; ++k) uu[k + 4 * s] = uu[k + 4 * s] + fabs(uu[k + (d + 1) * s]);
  for (int k = 0; k < ncell; ++k) uu[k] = zero; /* :100-102 */
  for (int d = 0; d < NDIM; ++d)
    for (int k = 0; k < ncell; ++k) uu[k] = uu[k] + fabs(gg ? gg[k + d * s] : 0.0);
  for (int k = 0; k < ncell; ++k) { /* :108-111 */
    uu[k] = uu[k] * dx / (uu[k + 4 * s] * uu[k + 4 * s]);
    uu[k] = fmaxr(uu[k], 0.0001);
  }
  double dt = P.courant_factor * dx / P.smallc; /* :114-118 */
  for (int k = 0; k < ncell; ++k) {
    double dtcell =
        dx / uu[k + 4 * s] * (sqrt(one + two * P.courant_factor * uu[k]) - one) / uu[k];
    dt = fminr(dt, dtcell);
  }
  return dt;
}

/* ---------------------------------------------------------------------- */
/* Uniform periodic level.                                                 */
static inline int wrap(int a, int n) { a %= n; return a < 0 ? a + n : a; }

int hr_uniform_godunov(const double *u, double *unew, int n, double dx, double dt, int nv,
                       int nthreads) {
  if (n < 2 || (n & 1) || nv <= 0) return -1;
  const int no = n / 2;                    /* octs per axis */
  const long noct = (long)no * no * no;
  const size_t ncell = (size_t)n * n * n;
  const long nbatch = (noct + nv - 1) / nv;
  /* set_unew: hydro/godunov_fine.f90:58-65 */
  memcpy(unew, u, sizeof(double) * ncell * NVAR);
  int err = 0;
#ifdef _OPENMP
  if (nthreads <= 0) nthreads = omp_get_max_threads();
#pragma omp parallel num_threads(nthreads) reduction(| : err)
#endif
  {
    hr_scratch s;
    const size_t us = (size_t)nv * NC6;
    double *uloc = (double *)calloc(us * NVAR, sizeof(double));
    double *flux = (double *)calloc((size_t)nv * NF3 * NVAR * NDIM, sizeof(double));
    double *tmp = (double *)calloc((size_t)nv * NF3 * 2 * NDIM, sizeof(double));
    if (scratch_alloc(&s, nv) || !uloc || !flux || !tmp) {
      err |= 1;
    } else {
#ifdef _OPENMP
#pragma omp for schedule(static)
#endif
      for (long b = 0; b < nbatch; ++b) {
        const long first = b * nv;
        const int ngrid = (int)((noct - first) < nv ? (noct - first) : nv);
        /* gather the 6x6x6 stencil: uloc(i3,j3,k3) is the cell at
         * (2*oct + i3 - 1) with periodic wrap (hydro/godunov_fine.f90:553-676
         * on a fully refined periodic level) */
        for (int l = 0; l < ngrid; ++l) {
          const long io = first + l;
          const int ox = (int)(io % no), oy = (int)((io / no) % no), oz = (int)(io / ((long)no * no));
          for (int k3 = -1; k3 <= 4; ++k3)
            for (int j3 = -1; j3 <= 4; ++j3)
              for (int i3 = -1; i3 <= 4; ++i3) {
                const int X = wrap(2 * ox + i3 - 1, n), Y = wrap(2 * oy + j3 - 1, n),
                          Z = wrap(2 * oz + k3 - 1, n);
                const size_t c = (size_t)X + (size_t)n * ((size_t)Y + (size_t)n * Z);
                for (int v = 0; v < NVAR; ++v) uloc[IC(l, i3, j3, k3) + v * us] = u[c + v * ncell];
              }
        }
        unsplit_s(&s, uloc, NULL, flux, tmp, dx, dx, dx, dt, ngrid); /* :681 */
        /* conservative update at the level: hydro/godunov_fine.f90:751-792 */
        const size_t fvs = (size_t)nv * NF3;
        for (int d = 0; d < NDIM; ++d) {
          const int i0 = (d == 0), j0 = (d == 1), k0 = (d == 2);
          for (int k2 = 0; k2 <= 1; ++k2)
            for (int j2 = 0; j2 <= 1; ++j2)
              for (int i2 = 0; i2 <= 1; ++i2) {
                const int i3 = 1 + i2, j3 = 1 + j2, k3 = 1 + k2;
                for (int v = 0; v < NVAR; ++v)
                  for (int l = 0; l < ngrid; ++l) {
                    const long io = first + l;
                    const int ox = (int)(io % no), oy = (int)((io / no) % no),
                              oz = (int)(io / ((long)no * no));
                    const size_t c = (size_t)(2 * ox + i2) +
                                     (size_t)n * ((size_t)(2 * oy + j2) + (size_t)n * (2 * oz + k2));
                    const size_t fb = (size_t)v * fvs + (size_t)d * fvs * NVAR;
                    unew[c + v * ncell] = unew[c + v * ncell] +
                                          (flux[IF(l, i3, j3, k3) + fb] -
                                           flux[IF(l, i3 + i0, j3 + j0, k3 + k0) + fb]);
                  }
              }
        }
      }
    }
    scratch_free(&s);
    free(uloc); free(flux); free(tmp);
  }
  (void)nthreads;
  return err ? -2 : 0;
}

double hr_uniform_courant(const double *u, int n, double dx, double dt_start, int nv,
                          double *mass, double *etot, double *eint) {
  const size_t ncell = (size_t)n * n * n;
  const double vol = dx * dx * dx; /* dx**ndim, hydro/courant_fine.f90:41 */
  double *uu = (double *)malloc(sizeof(double) * (size_t)nv * NVAR);
  double gg[1] = {0};
  (void)gg;
  double dt_loc = dt_start, m = 0, ek = 0, ei = 0;
  /* vector sweeps over the cells (every cell of a uniform level is a leaf) */
  for (size_t c0 = 0; c0 < ncell; c0 += (size_t)nv) {
    int nleaf = (int)((ncell - c0) < (size_t)nv ? (ncell - c0) : (size_t)nv);
    for (int v = 0; v < NVAR; ++v)
      for (int i = 0; i < nleaf; ++i) uu[i + (size_t)v * nv] = u[c0 + i + v * ncell];
    for (int i = 0; i < nleaf; ++i) m = m + uu[i] * vol; /* :95-97 */
    for (int i = 0; i < nleaf; ++i) ek = ek + uu[i + 4 * (size_t)nv] * vol;
    for (int i = 0; i < nleaf; ++i) ei = ei + uu[i + 4 * (size_t)nv] * vol;
    for (int d = 0; d < NDIM; ++d)
      for (int i = 0; i < nleaf; ++i) {
        double mv = uu[i + (size_t)(d + 1) * nv];
        ei = ei - 0.5 * (mv * mv) / fmaxr(uu[i], P.smallr) * vol;
      }
    double dt_lev = hr_cmpdt(uu, NULL, dx, nleaf, nv); /* :123 */
    dt_loc = fminr(dt_loc, dt_lev);
  }
  free(uu);
  if (mass) *mass = m;
  if (etot) *etot = ek;
  if (eint) *eint = ei;
  return dt_loc;
}

double hr_uniform_step(const double *u, double *unew, int n, double dx, int nv, int nthreads) {
  /* newdt_fine starts from boxlen/smallc (pm/newdt_fine.f90:51); boxlen=1 */
  double dt = hr_uniform_courant(u, n, dx, 1.0 / P.smallc, nv, NULL, NULL, NULL);
  if (hr_uniform_godunov(u, unew, n, dx, dt, nv, nthreads)) return -1.0;
  return dt;
}
