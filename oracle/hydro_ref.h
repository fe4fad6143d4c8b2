/*
 * TEST INFRASTRUCTURE ONLY -- the CPU oracle for the RAMSES unsplit
 * MUSCL-Hancock hydro step.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load it, and only as the checker.  The
 * product library (ramses_amd/) never links or calls it.
 *
 * Array layouts are the reference's Fortran layouts (column-major, the oct
 * lane `l` fastest with leading dimension `nv` = nvector):
 *   uin   (nv, -1:4, -1:4, -1:4, nvar)      hydro/umuscl.f90:33
 *   gravin(nv, -1:4, -1:4, -1:4, ndim)      hydro/umuscl.f90:34
 *   flux  (nv, 1:3, 1:3, 1:3, nvar, ndim)   hydro/umuscl.f90:37
 *   tmp   (nv, 1:3, 1:3, 1:3, 2, ndim)      hydro/umuscl.f90:38
 *   uu    (nv, nvar), gg (nv, ndim)         hydro/godunov_utils.f90:12-13
 * NDIM=3, NVAR=5, NENER=0 are fixed.
 */
#ifndef RAMSES_HYDRO_REF_H
#define RAMSES_HYDRO_REF_H

#ifdef __cplusplus
extern "C" {
#endif

enum { HR_RIEMANN_LLF = 0, HR_RIEMANN_HLLC = 1, HR_RIEMANN_HLL = 2, HR_RIEMANN_ACOUSTIC = 3, HR_RIEMANN_EXACT = 4 };

/* hydro/hydro_parameters.f90:74-85 (module globals of the reference) */
typedef struct {
  double gamma;
  double smallr;
  double smallc;
  double slope_theta;
  double courant_factor;
  double difmag;
  int slope_type;
  int riemann;
  int niter_riemann; /* Newton iterations of the 'exact' solver (default 10) */
} hr_params;

/* returns 0 on success, <0 if the configuration is not supported */
int hr_set_params(const hr_params *p);

/* hydro/umuscl.f90:22-171 */
int hr_unsplit(const double *uin, const double *gravin, double *flux, double *tmp,
               double dx, double dy, double dz, double dt, int ngrid, int nv);

/* `ngroups` independent unsplit calls on consecutive nvector-strided arrays,
 * the last one with `ngrid_last` lanes; OpenMP over groups */
int hr_unsplit_batched(const double *uin, const double *gravin, double *flux, double *tmp,
                       double dx, double dt, long ngroups, int ngrid_last, int nv, int nthreads);

/* hydro/godunov_utils.f90:5-120; uu is overwritten in place like the reference */
double hr_cmpdt(double *uu, const double *gg, double dx, int ncell, int nv);

/* Uniform periodic grid of n^3 cells, conserved state u[var][z][y][x].
 * godfine1 (hydro/godunov_fine.f90:486-911) restated for a single fully
 * refined periodic level: 6x6x6 gather per oct, unsplit per nvector batch,
 * unew = uold + sum_dim (flux_left - flux_right).  unew must not alias u. */
int hr_uniform_godunov(const double *u, double *unew, int n, double dx, double dt, int nv,
                       int nthreads);

/* courant_fine (hydro/courant_fine.f90:1-159) on the same grid: returns the
 * new dt (starting from dt_start) and the mass/energy/internal-energy sums. */
double hr_uniform_courant(const double *u, int n, double dx, double dt_start, int nv,
                          double *mass, double *etot, double *eint);

/* one full step: dt = courant, unew = godunov(u, dt); returns dt */
double hr_uniform_step(const double *u, double *unew, int n, double dx, int nv, int nthreads);

#ifdef __cplusplus
}
#endif
#endif
