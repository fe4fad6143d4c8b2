// C ABI of the library (include/ramses_hip.h): configuration, error
// reporting, and the host-memory / device-memory forms of `unsplit`.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <atomic>
#include <mutex>
#include <string>
#include <vector>
#include <cstdint>

#include "../../include/ramses_hip.h"
#include "hydro_device.hpp"
#include "amr_args.hpp"

namespace ramses {
struct UnsplitArgs;
hipError_t launch_unsplit(const UnsplitArgs& a, bool fp32, hipStream_t stream);
hipError_t launch_cmpdt(const double* uu, const double* gg, int nv, int ncell, double dx,
                        const HydroParams& hp, double* d_dt, hipStream_t stream);
hipError_t launch_set_double(double* p, double v, hipStream_t s);
hipError_t launch_courant(const double* u, int n, double dx, const HydroParams& hp, double* d_dt, hipStream_t s,
                          double* partials = nullptr, double* d_sums = nullptr);
size_t uniform_sums_work_bytes(int n);
size_t amr_begin_partials_bytes(int noct, bool persistent);
hipError_t launch_set_doubles(double* p, int n, double v, hipStream_t s);
hipError_t launch_uniform_godunov(const double* u, double* unew, int n, double dx, const double* d_dt,
                                  double* d_dt_next, const HydroParams& hp, bool fp32, hipStream_t s,
                                  double* partials = nullptr, double* d_sums = nullptr);
hipError_t launch_level_batch(double* const* buf0, double* const* buf1, double* const* state, const int* n,
                              const double* dx, int nslots, const HydroParams& hp, bool fp32, hipStream_t s);
hipError_t launch_init_slot(double* state, double t_end, double start, hipStream_t s);
hipError_t launch_amr_godunov(const AmrLevelArgs& a, double dx, const double* d_dt, double* tmp, bool fp32,
                              hipStream_t s);
hipError_t launch_amr_copy(const int* active, int noct, const AmrTreeView& t, const double* src, double* dst,
                           long ustride, hipStream_t s, const int* dn = nullptr);
hipError_t launch_amr_upload(const int* active, int noct, const AmrTreeView& t, double* u, long ustride,
                             double smallr, int ivar_mode, hipStream_t s, const int* dn = nullptr);
hipError_t launch_amr_courant(const int* active, int noct, const AmrTreeView& t, const double* u, long ustride,
                              double dx, const HydroParams& hp, double* d_dt, hipStream_t s, const int* dn = nullptr);
hipError_t launch_amr_dt(int op, double* dt, double* dtold, double* time, int l, int coarser, int nsub, int icount,
                         double start, hipStream_t s, const int* dn = nullptr, long long* cnt = nullptr);
hipError_t launch_set_ints(int* p, int n, int v, hipStream_t s);
hipError_t launch_amr_cons0(const ConsLevels& lv, const int* dn, const AmrTreeView& t, const double* u, long ustride,
                            int levelmin, int nlevelmax, double boxlen, int nv, const int* armed, double* cons0,
                            hipStream_t s);
hipError_t launch_clear_bits(int* p, int mask, hipStream_t s);
hipError_t launch_amr_begin(const int* active, int noct, const AmrTreeView& t, const double* u, double* unew,
                            long ustride, double dx, const HydroParams& hp, double* d_dt, hipStream_t s,
                            const int* dn = nullptr, double* partials = nullptr,
                            double* sums = nullptr);
hipError_t launch_amr_finish(const int* active, int noct, const AmrTreeView& t, const double* unew, double* u,
                             long ustride, double smallr, int ivar_mode, hipStream_t s, const int* dn = nullptr);
struct RefineParams {
  double err_grad_d, err_grad_u, err_grad_p;
  double floor_d, floor_u, floor_p;
  int nexpand;
};
struct TreeMut {
  int* son;
  int* father;
  int* nbor;
};
hipError_t launch_flag_fine(const int* active, int noct, const AmrTreeView& t, int ilevel, int icount, int levelmin,
                            int nlevelmax, const double* u, long us, int* flag1, int* flag2, const RefineParams& rp,
                            const HydroParams& hp, hipStream_t s, const int* dn = nullptr);
size_t refine_work_bytes(int noct, int noct_fine);
hipError_t launch_refine_fine_dev(const int* active, const int* dn_l, int ilevel, const AmrTreeView& t,
                                  const TreeMut& m, int* active_fine, int* dn_f, int* pos, int* ring, int* dfree,
                                  int* ctrl, int* flag1, int* flag2, double* xg, double* uold, double* unew,
                                  int interpol_type, int interpol_var, double smallr, void* work, int* d_error,
                                  int capacity, hipStream_t s);
hipError_t launch_refine_fine(const int* active, int noct, int ilevel, const AmrTreeView& t, const TreeMut& m,
                              int* active_fine, int* noct_fine, int* pos, int* ring, int* head, int* numbf,
                              int* used_mem, int* flag1, int* flag2, double* xg, double* uold, double* unew,
                              int interpol_type, int interpol_var, double smallr, void* work, int* d_error,
                              int* nofree, hipStream_t s);
Tuning& tuning() {
  static Tuning t{0, 0, 0, 0, 0};
  return t;
}
}  // namespace ramses

#include "unsplit_args.hpp"
#include "amr_args.hpp"

namespace {

thread_local std::string g_err;

// The whole process-wide configuration, read as one snapshot (current())
// by every entry point: hydro parameters, prolongation and refinement
// criteria.
struct Config {
  ramses::HydroParams hp{1.4, 1e-10, 1e-10, 1.5, 0.5, 1, ramses::kLLF, 10};
  int precision = RAMSES_FP64;
  int nvector = 32;
  bool set = false;
  int interpol_type = 1, interpol_var = 0;  // hydro/hydro_parameters.f90:88-89
  ramses::RefineParams rp{-1.0, -1.0, -1.0, 1e-10, 1e-10, 1e-10, 1};
};

std::mutex g_cfg_mu;
Config g_cfg;
// Bumped by every configuration call: a captured coarse step holds the
// configuration of its capture and is captured again after a change.
std::atomic<long> g_config_gen{0};

int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}

int hip_fail(hipError_t e, const char* where) {
  g_err = std::string(where) + ": " + hipGetErrorString(e);
  return (int)e;
}

Config current() {
  std::lock_guard<std::mutex> lk(g_cfg_mu);
  return g_cfg;
}

// Device staging for the host-memory entry points, with a pinned host
// mirror: the caller's arrays are copied into it on the host and cross PCIe
// in two DMA transfers each way instead of six pageable ones.  The
// reference's unsplit is not reentrant (its `save` scratch arrays,
// hydro/umuscl.f90:41-53); the host-memory form is likewise serialised on
// one library-owned stream.
struct Staging {
  std::mutex mu;
  hipStream_t stream = nullptr;
  double* buf = nullptr;   // device
  double* host = nullptr;  // pinned host memory of the same size
  size_t cap = 0;          // doubles
  int ensure(size_t n) {
    if (!stream) {
      hipError_t e = hipStreamCreateWithFlags(&stream, hipStreamNonBlocking);
      if (e != hipSuccess) return hip_fail(e, "hipStreamCreate");
    }
    if (n <= cap) return 0;
    // every call synchronises its stream before returning, so nothing is in
    // flight on the old buffers
    if (buf) (void)hipFree(buf);
    if (host) (void)hipHostFree(host);
    buf = nullptr;
    host = nullptr;
    cap = 0;
    hipError_t e = hipMalloc(&buf, n * sizeof(double));
    if (e != hipSuccess) return hip_fail(e, "hipMalloc");
    e = hipHostMalloc(reinterpret_cast<void**>(&host), n * sizeof(double), hipHostMallocDefault);
    if (e != hipSuccess) {
      (void)hipFree(buf);
      buf = nullptr;
      return hip_fail(e, "hipHostMalloc");
    }
    cap = n;
    return 0;
  }
};
Staging g_stage;

}  // namespace

extern "C" {

int ramses_hydro_config(const ramses_hydro_params* p) {
  if (!p) return fail(-1, "ramses_hydro_config: null parameters");
  if (!(p->gamma > 1.0)) return fail(-2, "gamma must be > 1");
  if (!(p->smallr > 0.0) || !(p->smallc > 0.0)) return fail(-3, "smallr and smallc must be > 0");
  const int st = p->slope_type;
  if (!(st == 0 || st == 1 || st == 2 || st == 3 || st == 7 || st == 8))
    return fail(-4, "unsupported slope_type for NDIM=3 (supported: 0,1,2,3,7,8)");
  if (p->riemann < RAMSES_RIEMANN_LLF || p->riemann > RAMSES_RIEMANN_EXACT)
    return fail(-5, "unsupported riemann solver (llf, hllc, hll, acoustic, exact)");
  if (p->niter_riemann < 0 || p->niter_riemann > 1000) return fail(-10, "niter_riemann out of [0, 1000]");
  if (p->difmag != 0.0) return fail(-6, "difmag > 0 (artificial diffusion) is not supported");
  if (p->precision != RAMSES_FP64 && p->precision != RAMSES_FP32_MIXED) return fail(-7, "bad precision");
  if (p->nvector <= 0 || p->nvector > 4096) return fail(-8, "nvector out of range");
  if (!(p->courant_factor > 0.0)) return fail(-9, "courant_factor must be > 0");
  Config c;
  c.hp.gamma = p->gamma;
  c.hp.smallr = p->smallr;
  c.hp.smallc = p->smallc;
  c.hp.slope_theta = p->slope_theta;
  c.hp.courant_factor = p->courant_factor;
  c.hp.slope_type = p->slope_type;
  c.hp.riemann = p->riemann;
  c.hp.niter_riemann = p->niter_riemann;
  c.precision = p->precision;
  c.nvector = p->nvector;
  c.set = true;
  std::lock_guard<std::mutex> lk(g_cfg_mu);
  c.interpol_type = g_cfg.interpol_type;
  c.interpol_var = g_cfg.interpol_var;
  c.rp = g_cfg.rp;
  g_cfg = c;
  ++g_config_gen;
  return 0;
}

const char* ramses_last_error(void) { return g_err.c_str(); }

int ramses_unsplit_device(const double* uin, const double* gravin, double* flux, double* tmp, double dx,
                          double dy, double dz, double dt, long ngroups, int ngrid_last, void* stream) {
  const Config c = current();
  if (ngroups < 0) return fail(-1, "ngroups < 0");
  if (ngroups == 0) return 0;
  if (ngrid_last < 0 || ngrid_last > c.nvector) return fail(-1, "ngrid_last out of [0, nvector]");
  if (!uin || !flux || !tmp) return fail(-1, "null array");
  if (!(dx > 0) || !(dy > 0) || !(dz > 0)) return fail(-1, "dx, dy, dz must be > 0");
  if ((double)ngroups * 12.0 * c.nvector > 2147483647.0 * 256.0) return fail(-1, "batch too large");
  ramses::UnsplitArgs a;
  a.uin = uin;
  a.grav = gravin;
  a.flux = flux;
  a.tmp = tmp;
  a.dx[0] = dx;
  a.dx[1] = dy;
  a.dx[2] = dz;
  a.dt = dt;
  a.d_dt = nullptr;
  a.ngroups = ngroups;
  a.ngrid_last = ngrid_last;
  a.nv = c.nvector;
  a.hp = c.hp;
  hipError_t e = ramses::launch_unsplit(a, c.precision == RAMSES_FP32_MIXED, (hipStream_t)stream);
  if (e != hipSuccess) return hip_fail(e, "unsplit launch");
  return 0;
}

int ramses_unsplit(const double* uin, const double* gravin, double* flux, double* tmp, double dx, double dy,
                   double dz, double dt, int ngrid) {
  const Config c = current();
  if (!uin || !flux || !tmp) return fail(-1, "null array");
  if (ngrid < 0 || ngrid > c.nvector) return fail(-1, "ngrid out of [0, nvector]");
  if (ngrid == 0) return 0;
  const size_t nv = (size_t)c.nvector;
  const size_t nu = nv * 216 * 5, ng = nv * 216 * 3, nf = nv * 27 * 5 * 3, nt = nv * 27 * 2 * 3;
  std::lock_guard<std::mutex> lk(g_stage.mu);
  int rc = g_stage.ensure(nu + ng + nf + nt);
  if (rc) return rc;
  // uin, flux, tmp, gravin: one transfer in (gravin last, left out when
  // null) and one out (flux and tmp)
  double* d_u = g_stage.buf;
  double* d_f = d_u + nu;
  double* d_t = d_f + nf;
  double* d_g = d_t + nt;
  double* h_u = g_stage.host;  // the same layout in pinned host memory
  double* h_f = h_u + nu;
  double* h_t = h_f + nf;
  double* h_g = h_t + nt;
  hipStream_t s = g_stage.stream;
  hipError_t e;
  // flux/tmp travel both ways so that entries the reference never writes keep the caller's values
  std::memcpy(h_u, uin, nu * 8);
  std::memcpy(h_f, flux, nf * 8);
  std::memcpy(h_t, tmp, nt * 8);
  if (gravin) std::memcpy(h_g, gravin, ng * 8);
  if ((e = hipMemcpyAsync(d_u, h_u, (nu + nf + nt + (gravin ? ng : 0)) * 8, hipMemcpyHostToDevice, s)) != hipSuccess)
    return hip_fail(e, "H2D uin, flux, tmp, gravin");
  rc = ramses_unsplit_device(d_u, gravin ? d_g : nullptr, d_f, d_t, dx, dy, dz, dt, 1, ngrid, s);
  if (rc) {
    (void)hipStreamSynchronize(s);  // the copies above still read the pinned buffer
    return rc;
  }
  if ((e = hipMemcpyAsync(h_f, d_f, (nf + nt) * 8, hipMemcpyDeviceToHost, s)) != hipSuccess) {
    (void)hipStreamSynchronize(s);
    return hip_fail(e, "D2H flux, tmp");
  }
  if ((e = hipStreamSynchronize(s)) != hipSuccess) return hip_fail(e, "unsplit sync");
  std::memcpy(flux, h_f, nf * 8);
  std::memcpy(tmp, h_t, nt * 8);
  return 0;
}

}  // extern "C"

// Staging of ramses_unsplit_noalloc: device memory the library never
// allocates.  One call stages uin, gravin, flux and tmp of one batch, 2295
// doubles per lane of nvector, so the 2 MiB array serves nvector <= 114.
// Each GPU the library runs on holds its own copy until the process ends.
constexpr size_t kNoAllocDoubles = 262144;
static __device__ __align__(256) double scratch[kNoAllocDoubles];

namespace {
constexpr int kNoAllocMaxDevices = 64;
struct NoAllocStaging {
  std::mutex mu;
  // Device address of `scratch` on each device, taken once per device by
  // hipGetSymbolAddress with that device current.  Host code uses only
  // these addresses: the name `scratch` in host code is a host-side stand-in.
  double* ptr[kNoAllocMaxDevices] = {};
};
NoAllocStaging g_noalloc;
}  // namespace

extern "C" {

int ramses_unsplit_noalloc(const double* uin, const double* gravin, double* flux, double* tmp, double dx, double dy,
                           double dz, double dt, int ngrid, void* stream) {
  const Config c = current();
  if (!uin || !flux || !tmp) return fail(-1, "null array");
  if (ngrid < 0 || ngrid > c.nvector) return fail(-1, "ngrid out of [0, nvector]");
  const size_t nv = (size_t)c.nvector;
  const size_t nu = nv * 216 * 5, ng = nv * 216 * 3, nf = nv * 27 * 5 * 3, nt = nv * 27 * 2 * 3;
  if (nu + ng + nf + nt > kNoAllocDoubles)
    return fail(-11, "ramses_unsplit_noalloc: nvector " + std::to_string(nv) + " needs " +
                         std::to_string((nu + ng + nf + nt) * 8) + " bytes of staging, more than its fixed " +
                         std::to_string(kNoAllocDoubles * 8) + " (nvector <= 114)");
  if (ngrid == 0) return 0;
  int dev = 0;
  hipError_t e;
  if ((e = hipGetDevice(&dev)) != hipSuccess) return hip_fail(e, "hipGetDevice");
  if (dev < 0 || dev >= kNoAllocMaxDevices) return fail(-1, "ramses_unsplit_noalloc: device index beyond 63");
  // One call at a time per process: every call on a device shares its
  // staging, and the call returns only after its results reached the host.
  std::lock_guard<std::mutex> lk(g_noalloc.mu);
  if (!g_noalloc.ptr[dev]) {
    void* p = nullptr;
    if ((e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch))) != hipSuccess) return hip_fail(e, "staging address");
    g_noalloc.ptr[dev] = static_cast<double*>(p);
  }
  double* d_u = g_noalloc.ptr[dev];
  double* d_g = d_u + nu;
  double* d_f = d_g + ng;
  double* d_t = d_f + nf;
  hipStream_t s = (hipStream_t)stream;
  // flux/tmp travel both ways so that entries the reference never writes
  // keep the caller's values; the staging keeps the previous call's data,
  // and every word the kernel reads is copied in first
  if ((e = hipMemcpyAsync(d_u, uin, nu * 8, hipMemcpyHostToDevice, s)) != hipSuccess) return hip_fail(e, "H2D uin");
  if (gravin && (e = hipMemcpyAsync(d_g, gravin, ng * 8, hipMemcpyHostToDevice, s)) != hipSuccess)
    return hip_fail(e, "H2D gravin");
  if ((e = hipMemcpyAsync(d_f, flux, nf * 8, hipMemcpyHostToDevice, s)) != hipSuccess) return hip_fail(e, "H2D flux");
  if ((e = hipMemcpyAsync(d_t, tmp, nt * 8, hipMemcpyHostToDevice, s)) != hipSuccess) return hip_fail(e, "H2D tmp");
  int rc = ramses_unsplit_device(d_u, gravin ? d_g : nullptr, d_f, d_t, dx, dy, dz, dt, 1, ngrid, s);
  if (rc) return rc;
  if ((e = hipMemcpyAsync(flux, d_f, nf * 8, hipMemcpyDeviceToHost, s)) != hipSuccess) return hip_fail(e, "D2H flux");
  if ((e = hipMemcpyAsync(tmp, d_t, nt * 8, hipMemcpyDeviceToHost, s)) != hipSuccess) return hip_fail(e, "D2H tmp");
  if ((e = hipStreamSynchronize(s)) != hipSuccess) return hip_fail(e, "unsplit sync");
  return 0;
}

void unsplit_(const double* uin, const double* gravin, const double* pin, double* flux, double* tmp,
              const double* dx, const double* dy, const double* dz, const double* dt, const int* ngrid) {
  (void)pin;  // declared but never read by the reference body (hydro/umuscl.f90:32)
  // godfine1 passes zeros when poisson is off (its gloc starts at 0.0 and
  // is filled only under `if(poisson)`, hydro/godunov_fine.f90:512,637).
  // An array of +0.0 words gives
  // the same bits as no array (zero acceleration enters ctoprim either way),
  // so it is not copied and the gravity-free kernels run; a -0.0 or any
  // other word keeps the array.
  const double* g = gravin;
  if (g) {
    const size_t n = (size_t)current().nvector * 216 * 3;
    size_t i = 0;
    for (; i < n; ++i) {
      uint64_t w;
      std::memcpy(&w, g + i, sizeof w);
      if (w != 0) break;
    }
    if (i == n) g = nullptr;
  }
  const int rc = ramses_unsplit(uin, g, flux, tmp, *dx, *dy, *dz, *dt, *ngrid);
  if (rc) {
    std::fprintf(stderr, "unsplit: %s\n", ramses_last_error());
    std::abort();
  }
}

int ramses_cmpdt(const double* uu, const double* gg, double dx, int ncell, double* dt) {
  const Config c = current();
  if (!uu || !dt) return fail(-1, "null argument");
  if (ncell < 0 || ncell > c.nvector) return fail(-1, "ncell out of [0, nvector]");
  const size_t nv = (size_t)c.nvector;
  std::lock_guard<std::mutex> lk(g_stage.mu);
  int rc = g_stage.ensure(nv * 5 + nv * 3 + 1);
  if (rc) return rc;
  double* d_u = g_stage.buf;
  double* d_g = d_u + nv * 5;
  double* d_dt = d_g + nv * 3;
  hipStream_t s = g_stage.stream;
  hipError_t e;
  if ((e = hipMemcpyAsync(d_u, uu, nv * 5 * 8, hipMemcpyHostToDevice, s)) != hipSuccess) return hip_fail(e, "H2D uu");
  if (gg && (e = hipMemcpyAsync(d_g, gg, nv * 3 * 8, hipMemcpyHostToDevice, s)) != hipSuccess)
    return hip_fail(e, "H2D gg");
  if ((e = ramses::launch_cmpdt(d_u, gg ? d_g : nullptr, (int)nv, ncell, dx, c.hp, d_dt, s)) != hipSuccess)
    return hip_fail(e, "cmpdt launch");
  if ((e = hipMemcpyAsync(dt, d_dt, 8, hipMemcpyDeviceToHost, s)) != hipSuccess) return hip_fail(e, "D2H dt");
  if ((e = hipStreamSynchronize(s)) != hipSuccess) return hip_fail(e, "cmpdt sync");
  return 0;
}

namespace {
int amr_view(const ramses_amr_tree* t, ramses::AmrTreeView& v) {
  if (!t || !t->son || !t->father || !t->nbor) return fail(-1, "null tree array");
  if (t->ncoarse < 1 || t->ngridmax < 1 || t->nx < 1 || t->ny < 1 || t->nz < 1) return fail(-1, "bad tree sizes");
  if (t->ncoarse != t->nx * t->ny * t->nz) return fail(-1, "ncoarse must be nx*ny*nz");
  // cell indices are 32-bit, as the reference's integer arrays are
  if ((long)t->ncoarse + 8L * t->ngridmax > 2147483647L) return fail(-1, "ncoarse + 8*ngridmax exceeds int32");
  v.son = t->son;
  v.father = t->father;
  v.nbor = t->nbor;
  v.ncoarse = t->ncoarse;
  v.ngridmax = t->ngridmax;
  v.nx = t->nx;
  v.ny = t->ny;
  v.nz = t->nz;
  v.ncell = (long)t->ncoarse + 8L * t->ngridmax;
  return 0;
}

size_t amr_work(int noct, int nv) {
  const size_t nb = (size_t)((noct + nv - 1) / nv);
  const size_t per = (size_t)nv * (216 * 5 + 27 * 15 + 27 * 6) * sizeof(double) + (size_t)nv * 216;
  return nb * per + 256;
}
}  // namespace

int ramses_amr_config(int interpol_type, int interpol_var) {
  if (interpol_type < 0 || interpol_type > 4) return fail(-2, "interpol_type must be 0..4");
  if (interpol_var < 0 || interpol_var > 2) return fail(-2, "interpol_var must be 0..2");
  // the reference stops on this pair (hydro/interpol_hydro.f90:369-372)
  if (interpol_type == 4 && interpol_var != 2) return fail(-2, "interpol_type=4 is designed for interpol_var=2");
  std::lock_guard<std::mutex> lk(g_cfg_mu);
  g_cfg.interpol_type = interpol_type;
  g_cfg.interpol_var = interpol_var;
  ++g_config_gen;
  return 0;
}

unsigned long ramses_amr_work_bytes(int noct) { return (unsigned long)amr_work(noct < 0 ? 0 : noct, current().nvector); }

int ramses_amr_godunov(const ramses_amr_tree* t, const ramses_amr_level* lv, const double* uold, double* unew,
                       const double* d_dt, void* work, unsigned long work_bytes, int* d_error, void* stream) {
  const Config c = current();
  ramses::AmrTreeView v;
  int rc = amr_view(t, v);
  if (rc) return rc;
  if (!lv || !uold || !unew || !d_dt || !d_error) return fail(-1, "null argument");
  if (lv->noct < 0 || lv->noct > t->ngridmax) return fail(-1, "noct out of range");
  if (lv->noct > 0 && (!lv->active || !lv->pos)) return fail(-1, "null active/pos");
  if (lv->ilevel < 1) return fail(-1, "ilevel must be >= 1");
  if (!(lv->dx > 0)) return fail(-1, "dx must be > 0");
  if (c.nvector * 27 > 1024) return fail(-8, "nvector too large for the AMR gather (max 37)");
  const size_t need = amr_work(lv->noct, c.nvector);
  if (!work || work_bytes < need) return fail(-1, "work buffer too small (see ramses_amr_work_bytes)");
  ramses::AmrLevelArgs a;
  a.t = v;
  a.uold = uold;
  a.unew = unew;
  a.ustride = v.ncell + 1;
  a.active = lv->active;
  a.noct = lv->noct;
  a.dnoct = nullptr;
  a.dnoct_coarse = nullptr;
  a.ilevel = lv->ilevel;
  a.pos = lv->pos;
  const int nv = c.nvector;
  const size_t nb = (size_t)((lv->noct + nv - 1) / nv);
  char* w = (char*)(((uintptr_t)work + 255) & ~(uintptr_t)255);
  a.uin = (double*)w;
  a.flux = a.uin + nb * nv * 216 * 5;
  double* tmp = a.flux + nb * nv * 27 * 15;
  a.ok = (unsigned char*)(tmp + nb * nv * 27 * 6);
  a.nv = nv;
  a.interpol_type = c.interpol_type;
  a.interpol_var = c.interpol_var;
  a.hp = c.hp;
  a.active_coarse = lv->active_coarse;
  a.noct_coarse = lv->active_coarse ? lv->noct_coarse : 0;
  a.error = d_error;
  hipError_t e = ramses::launch_amr_godunov(a, lv->dx, d_dt, tmp, c.precision == RAMSES_FP32_MIXED,
                                            (hipStream_t)stream);
  if (e != hipSuccess) return hip_fail(e, "amr godunov launch");
  return 0;
}

int ramses_amr_copy(const ramses_amr_tree* t, const ramses_amr_level* lv, const double* src, double* dst,
                    void* stream) {
  ramses::AmrTreeView v;
  int rc = amr_view(t, v);
  if (rc) return rc;
  if (!lv || !src || !dst || (lv->noct > 0 && !lv->active)) return fail(-1, "null argument");
  if (lv->noct < 0 || lv->noct > t->ngridmax) return fail(-1, "noct out of range");
  hipError_t e = ramses::launch_amr_copy(lv->active, lv->noct, v, src, dst, v.ncell + 1, (hipStream_t)stream);
  if (e != hipSuccess) return hip_fail(e, "amr copy launch");
  return 0;
}

int ramses_amr_upload(const ramses_amr_tree* t, const ramses_amr_level* lv, double* u, void* stream) {
  const Config c = current();
  ramses::AmrTreeView v;
  int rc = amr_view(t, v);
  if (rc) return rc;
  if (!lv || !u || (lv->noct > 0 && !lv->active)) return fail(-1, "null argument");
  if (lv->noct < 0 || lv->noct > t->ngridmax) return fail(-1, "noct out of range");
  hipError_t e = ramses::launch_amr_upload(lv->active, lv->noct, v, u, v.ncell + 1, c.hp.smallr, c.interpol_var,
                                           (hipStream_t)stream);
  if (e != hipSuccess) return hip_fail(e, "amr upload launch");
  return 0;
}

int ramses_amr_courant(const ramses_amr_tree* t, const ramses_amr_level* lv, const double* u, double dt_start,
                       double* d_dt, void* stream) {
  const Config c = current();
  ramses::AmrTreeView v;
  int rc = amr_view(t, v);
  if (rc) return rc;
  if (!lv || !u || !d_dt || (lv->noct > 0 && !lv->active)) return fail(-1, "null argument");
  if (lv->noct < 0 || lv->noct > t->ngridmax) return fail(-1, "noct out of range");
  if (!(lv->dx > 0)) return fail(-1, "dx must be > 0");
  double start = c.hp.courant_factor * lv->dx / c.hp.smallc;
  if (dt_start < start) start = dt_start;
  hipError_t e = ramses::launch_set_double(d_dt, start, (hipStream_t)stream);
  if (e != hipSuccess) return hip_fail(e, "dt init launch");
  e = ramses::launch_amr_courant(lv->active, lv->noct, v, u, v.ncell + 1, lv->dx, c.hp, d_dt, (hipStream_t)stream);
  if (e != hipSuccess) return hip_fail(e, "amr courant launch");
  return 0;
}

int ramses_refine_config(const ramses_refine_params* p) {
  ramses::RefineParams rp = Config().rp;  // defaults: every criterion off, nexpand 1
  if (p) {
    if (p->nexpand < 0 || p->nexpand > 16) return fail(-2, "nexpand must be 0..16");
    if (!(p->floor_d >= 0) || !(p->floor_u >= 0) || !(p->floor_p >= 0)) return fail(-2, "floors must be >= 0");
    rp = ramses::RefineParams{p->err_grad_d, p->err_grad_u, p->err_grad_p,
                              p->floor_d,    p->floor_u,    p->floor_p,    p->nexpand};
  }
  std::lock_guard<std::mutex> lk(g_cfg_mu);
  g_cfg.rp = rp;
  ++g_config_gen;
  return 0;
}

int ramses_amr_flag_fine(const ramses_amr_tree* t, const ramses_amr_level* lv, int icount, int levelmin,
                         int nlevelmax, const double* uold, int* flag1, int* flag2, void* stream) {
  const Config c = current();
  ramses::AmrTreeView v;
  int rc = amr_view(t, v);
  if (rc) return rc;
  if (!lv || !uold || !flag1 || !flag2 || (lv->noct > 0 && !lv->active)) return fail(-1, "null argument");
  if (lv->noct < 0 || lv->noct > t->ngridmax) return fail(-1, "noct out of range");
  if (lv->ilevel < 1 || levelmin < 1 || nlevelmax < levelmin) return fail(-1, "bad levels");
  hipError_t e = ramses::launch_flag_fine(lv->active, lv->noct, v, lv->ilevel, icount, levelmin, nlevelmax, uold,
                                          v.ncell + 1, flag1, flag2, c.rp, c.hp, (hipStream_t)stream);
  if (e != hipSuccess) return hip_fail(e, "flag_fine launch");
  return 0;
}

unsigned long ramses_amr_refine_work_bytes(int noct, int noct_fine) {
  return (unsigned long)ramses::refine_work_bytes(noct < 0 ? 0 : noct, noct_fine < 0 ? 0 : noct_fine);
}

int ramses_amr_refine_fine(const ramses_amr_tree* t, const ramses_amr_level* lv, int* active_fine, int* noct_fine,
                           int* pos, ramses_amr_free* fl, int* flag1, int* flag2, double* xg, double* uold,
                           double* unew, void* work, unsigned long work_bytes, int* d_error, void* stream) {
  const Config c = current();
  ramses::AmrTreeView v;
  int rc = amr_view(t, v);
  if (rc) return rc;
  if (!lv || !active_fine || !noct_fine || !pos || !fl || !fl->ring || !flag1 || !flag2 || !uold || !unew ||
      !d_error || (lv->noct > 0 && !lv->active))
    return fail(-1, "null argument");
  if (lv->noct < 0 || lv->noct > t->ngridmax || *noct_fine < 0 || *noct_fine > t->ngridmax)
    return fail(-1, "noct out of range");
  // The creation and kill order (hence grid numbering, the free list and the
  // next level's list order) follows refine_fine's loop over nvector-sized
  // batches (amr/refine_utils.f90:395-415, 492-511); the refine kernels reproduce it
  // for nvector = 32, the reference build's value (bin/Makefile:11).
  if (c.nvector != 32) return fail(-8, "refine_fine requires nvector = 32 (the batch order of its grid numbering)");
  if (fl->head < 0 || fl->head >= t->ngridmax || fl->numbf < 0 || fl->numbf > t->ngridmax)
    return fail(-1, "bad free list");
  if ((unsigned long)ramses::refine_work_bytes(lv->noct, *noct_fine) > work_bytes || !work)
    return fail(-1, "work buffer too small (see ramses_amr_refine_work_bytes)");
  ramses::TreeMut m{t->son, t->father, t->nbor};
  int nofree = 0;
  if (lv->ilevel < 1) return fail(-1, "ilevel must be >= 1");
  hipError_t e = ramses::launch_refine_fine(lv->active, lv->noct, lv->ilevel, v, m, active_fine, noct_fine, pos,
                                            fl->ring, &fl->head, &fl->numbf, &fl->used_mem, flag1, flag2, xg,
                                            uold, unew,
                                            c.interpol_type, c.interpol_var, c.hp.smallr, work, d_error, &nofree,
                                            (hipStream_t)stream);
  if (e != hipSuccess) return hip_fail(e, "refine_fine");
  if (nofree) return fail(-9, "No more free memory: increase ngridmax");
  return 0;
}

// ---------------------------------------------------------------------------
// amr_step (amr/amr_step.f90) over a device-resident mesh, host recursion.
// ---------------------------------------------------------------------------
namespace {

struct StepCtx {
  ramses_amr_mesh* m;
  hipStream_t s;
  Config cfg;
  bool cons_armed;  // nstep_coarse changed since the last update_time
  bool cons_done;   // the step's first update_time has been passed
};

// The grid lists and host counts of every level.
ramses::ConsLevels cons_levels(const ramses_amr_mesh* m) {
  ramses::ConsLevels lv{};
  for (int l = 1; l <= m->nlevelmax && l < ramses::kConsMaxLevels; ++l) {
    lv.active[l] = m->active[l];
    lv.noct[l] = m->noct[l];
  }
  return lv;
}

ramses_amr_level level_of(const ramses_amr_mesh* m, int l) {
  ramses_amr_level lv;
  lv.ilevel = l;
  lv.dx = __builtin_ldexp(1.0, -l) * m->boxlen;
  lv.active = m->active[l];
  lv.noct = m->noct[l];
  lv.pos = m->pos;
  lv.active_coarse = (l > 1 && m->noct[l - 1] > 0) ? m->active[l - 1] : nullptr;
  lv.noct_coarse = (l > 1) ? m->noct[l - 1] : 0;
  return lv;
}

int ensure_work(ramses_amr_mesh* m, size_t need, hipStream_t s) {
  if (m->work && m->work_bytes >= need) return 0;
  hipError_t e;
  if (m->work) {
    if ((e = hipStreamSynchronize(s)) != hipSuccess) return hip_fail(e, "coarse step sync");
    (void)hipFree(m->work);
    m->work = nullptr;
    m->work_bytes = 0;
  }
  const size_t bytes = need + need / 4 + 4096;
  if ((e = hipMalloc(&m->work, bytes)) != hipSuccess) return hip_fail(e, "coarse step workspace");
  m->work_bytes = bytes;
  return 0;
}

int amr_step_rec(StepCtx& c, int l, int icount) {
  ramses_amr_mesh* m = c.m;
  if (m->noct[l] == 0) return 0;
  int rc;
  if (m->adaptive && m->levelmin < m->nlevelmax && (l == m->levelmin || icount > 1)) {
    for (int i = l; i < m->nlevelmax; ++i) {
      if (m->noct[i] == 0) continue;
      const size_t need = ramses::refine_work_bytes(m->noct[i], m->noct[i + 1]);
      if ((rc = ensure_work(m, need, c.s))) return rc;
      ramses_amr_level lv = level_of(m, i);
      rc = ramses_amr_refine_fine(&m->tree, &lv, m->active[i + 1], &m->noct[i + 1], m->pos, &m->free, m->flag1,
                                  m->flag2, m->xg, m->uold, m->unew, m->work, m->work_bytes, m->error, c.s);
      if (rc) return rc;
    }
  }
  ramses_amr_level lv = level_of(m, l);
  ramses::AmrTreeView v;
  if ((rc = amr_view(&m->tree, v))) return rc;
  // newdt_fine: dtold = dtnew, dtnew = min(boxlen/smallc, courant_factor*dx/smallc,
  // dtnew(l-1)/nsubcycle(l-1)), then the courant reduction over the leaves
  double start = c.cfg.hp.courant_factor * lv.dx / c.cfg.hp.smallc;
  const double cap = m->boxlen / c.cfg.hp.smallc;
  if (cap < start) start = cap;
  const int coarser = l > m->levelmin;
  hipError_t e = ramses::launch_amr_dt(0, m->dt, m->dtold, m->time, l, coarser, coarser ? m->nsubcycle[l - 1] : 1,
                                       0, start, c.s);
  if (e != hipSuccess) return hip_fail(e, "dt launch");
  // courant_fine and set_unew in one pass (with the conservation sums:
  // mass_tot = mass_tot + mass_all etc., hydro/courant_fine.f90:152-154)
  if (m->sums && (rc = ensure_work(m, ramses::amr_begin_partials_bytes(lv.noct, false), c.s))) return rc;
  e = ramses::launch_amr_begin(lv.active, lv.noct, v, m->uold, m->unew, v.ncell + 1, lv.dx, c.cfg.hp, m->dt + l, c.s,
                               nullptr, m->sums ? (double*)m->work : nullptr, m->sums);
  if (e != hipSuccess) return hip_fail(e, "courant / set_unew launch");
  if (l < m->nlevelmax && m->noct[l + 1] > 0) {
    if ((rc = amr_step_rec(c, l + 1, 1))) return rc;
    if (m->nsubcycle[l] == 2 && (rc = amr_step_rec(c, l + 1, 2))) return rc;
  } else {
    e = ramses::launch_amr_dt(1, m->dt, m->dtold, m->time, l, l < m->nlevelmax, m->nsubcycle[l], 0, 0.0, c.s);
    if (e != hipSuccess) return hip_fail(e, "dt launch");
    m->nstep += 1;
    // update_time's conservation baselines (amr/update_time.f90:229-246)
    if (m->cons0 && !c.cons_done && c.cons_armed) {
      e = ramses::launch_amr_cons0(cons_levels(m), nullptr, v, m->uold, v.ncell + 1, m->levelmin, m->nlevelmax,
                                   m->boxlen, c.cfg.nvector, nullptr, m->cons0, c.s);
      if (e != hipSuccess) return hip_fail(e, "conservation baselines launch");
    }
    c.cons_done = true;
  }
  lv = level_of(m, l);  // the finer levels' refinements changed noct(l+1) only
  if ((rc = ensure_work(m, amr_work(lv.noct, current().nvector), c.s))) return rc;
  rc = ramses_amr_godunov(&m->tree, &lv, m->uold, m->unew, m->dt + l, m->work, m->work_bytes, m->error, c.s);
  if (rc) return rc;
  m->cell_updates += 8LL * lv.noct;
  e = ramses::launch_amr_finish(lv.active, lv.noct, v, m->unew, m->uold, v.ncell + 1, c.cfg.hp.smallr, c.cfg.interpol_var,
                                c.s);
  if (e != hipSuccess) return hip_fail(e, "set_uold/upload launch");
  if (m->adaptive) {
    // ensure_ref_rules runs for icount < nsubcycle(l-1) (amr/flag_utils.f90:94-98)
    const int ic = (l <= m->levelmin || m->nsubcycle[l - 1] == 2) ? icount : 2;
    rc = ramses_amr_flag_fine(&m->tree, &lv, ic, m->levelmin, m->nlevelmax, m->uold, m->flag1, m->flag2, c.s);
    if (rc) return rc;
  }
  if (l > m->levelmin) {
    e = ramses::launch_amr_dt(2, m->dt, m->dtold, m->time, l, 1, m->nsubcycle[l - 1], icount, 0.0, c.s);
    if (e != hipSuccess) return hip_fail(e, "dt launch");
  }
  return 0;
}

// --- the captured coarse step -------------------------------------------

// What a captured step reads and writes besides the mesh arrays: the level
// counts, the free list, the refinement decision and the counters.  One
// copy in before the replay and one copy out after it.
struct AmrDevState {
  int noct[RAMSES_MAX_LEVELS + 2];
  int free[4];       // head, numbf, used_mem
  int ctrl[4];       // refine decision: ncreate, nkill, ok
  int error;         // *m->error after the step
  int cons_armed;    // nstep_coarse changed since the last update_time
  long long cnt[2];  // nstep, cell_updates
};

struct GraphState {
  hipStream_t cap = nullptr;  // capture stream (default priority)
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  AmrDevState* dev = nullptr;
  AmrDevState* host = nullptr;  // pinned
  void* work = nullptr;
  size_t work_bytes = 0;
  int capacity = 0;
  std::vector<long> key;
  void drop_graph() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    exec = nullptr;
    graph = nullptr;
  }
  void release() {
    drop_graph();
    if (work) (void)hipFree(work);
    if (dev) (void)hipFree(dev);
    if (host) (void)hipHostFree(host);
    if (cap) (void)hipStreamDestroy(cap);
    work = nullptr;
    dev = nullptr;
    host = nullptr;
    cap = nullptr;
  }
};

void graph_key(const ramses_amr_mesh* m, int capacity, std::vector<long>& k) {
  k.clear();
  auto P = [&](const void* q) { k.push_back((long)(uintptr_t)q); };
  P(m->tree.son);
  P(m->tree.father);
  P(m->tree.nbor);
  for (int v : {m->tree.ncoarse, m->tree.ngridmax, m->tree.nx, m->tree.ny, m->tree.nz}) k.push_back(v);
  for (int l = 1; l <= m->nlevelmax; ++l) {
    P(m->active[l]);
    k.push_back(m->nsubcycle[l]);
  }
  for (const void* q : {(const void*)m->pos, (const void*)m->xg, (const void*)m->flag1, (const void*)m->flag2,
                        (const void*)m->free.ring, (const void*)m->uold, (const void*)m->unew, (const void*)m->dt,
                        (const void*)m->dtold, (const void*)m->time, (const void*)m->error, (const void*)m->sums,
                        (const void*)m->cons0})
    P(q);
  long bl;
  std::memcpy(&bl, &m->boxlen, sizeof(bl));
  for (long v : {(long)m->levelmin, (long)m->nlevelmax, (long)m->adaptive, bl, (long)capacity,
                 (long)g_config_gen.load()})
    k.push_back(v);
}

struct GCtx {
  ramses_amr_mesh* m;
  hipStream_t s;
  Config c;
  AmrDevState* dev;
  void* work;
  int capacity;
  ramses::AmrTreeView v;
  ramses::TreeMut tm;
  bool cons_done = false;  // the baselines' launch has been emitted
};

// amr_step(l, icount) as a fixed launch sequence: every level step down to
// nlevelmax is emitted, and each kernel of a level step does nothing when
// that level has no octs at run time, which is when the reference does not
// call it (amr/amr_step.f90:345-365).  The counts a step's kernels read do
// not change during that step: refine_fine(i) changes only level i+1, and a
// level's step refines only its own and finer levels.
hipError_t gstep(GCtx& g, int l, int icount) {
  ramses_amr_mesh* m = g.m;
  int* dn = g.dev->noct;
  const long us = g.v.ncell + 1;
  hipError_t e;
  if (m->adaptive && m->levelmin < m->nlevelmax && (l == m->levelmin || icount > 1)) {
    for (int i = l; i < m->nlevelmax; ++i) {
      e = ramses::launch_refine_fine_dev(m->active[i], dn + i, i, g.v, g.tm, m->active[i + 1], dn + i + 1, m->pos,
                                         m->free.ring, g.dev->free, g.dev->ctrl, m->flag1, m->flag2, m->xg, m->uold,
                                         m->unew, g.c.interpol_type, g.c.interpol_var, g.c.hp.smallr, g.work, m->error,
                                         g.capacity, g.s);
      if (e != hipSuccess) return e;
    }
  }
  const double dx = __builtin_ldexp(1.0, -l) * m->boxlen;
  double start = g.c.hp.courant_factor * dx / g.c.hp.smallc;
  const double capdt = m->boxlen / g.c.hp.smallc;
  if (capdt < start) start = capdt;
  const int coarser = l > m->levelmin;
  e = ramses::launch_amr_dt(0, m->dt, m->dtold, m->time, l, coarser, coarser ? m->nsubcycle[l - 1] : 1, 0, start, g.s,
                            dn, g.dev->cnt);
  if (e != hipSuccess) return e;
  double* partials = m->sums ? (double*)(((uintptr_t)g.work + 255) & ~(uintptr_t)255) : nullptr;
  e = ramses::launch_amr_begin(m->active[l], 0, g.v, m->uold, m->unew, us, dx, g.c.hp, m->dt + l, g.s, dn + l,
                               partials, m->sums);
  if (e != hipSuccess) return e;  // courant_fine + set_unew
  if (l < m->nlevelmax) {
    if ((e = gstep(g, l + 1, 1)) != hipSuccess) return e;
    if (m->nsubcycle[l] == 2 && (e = gstep(g, l + 1, 2)) != hipSuccess) return e;
  }
  e = ramses::launch_amr_dt(1, m->dt, m->dtold, m->time, l, l < m->nlevelmax, m->nsubcycle[l], 0, 0.0, g.s, dn,
                            g.dev->cnt);
  if (e != hipSuccess) return e;
  // update_time's conservation baselines: emitted after the first update_time
  // of the capture order (the deepest level's).  At run time the step's
  // first update_time may come at a shallower level, but every launch
  // between the two belongs to an empty level and does nothing, and the
  // kernel finds the levels of the first descent from the device counts.
  if (m->cons0 && !g.cons_done) {
    g.cons_done = true;
    e = ramses::launch_amr_cons0(cons_levels(m), dn, g.v, m->uold, us, m->levelmin, m->nlevelmax, m->boxlen,
                                 g.c.nvector, &g.dev->cons_armed, m->cons0, g.s);
    if (e != hipSuccess) return e;
  }
  ramses::AmrLevelArgs a;
  a.t = g.v;
  a.uold = m->uold;
  a.unew = m->unew;
  a.ustride = us;
  a.active = m->active[l];
  a.noct = 0;
  a.dnoct = dn + l;
  a.dnoct_coarse = l > 1 ? dn + l - 1 : nullptr;
  a.ilevel = l;
  a.pos = m->pos;
  const int nv = g.c.nvector;
  const size_t nb = (size_t)((g.capacity + nv - 1) / nv);
  char* w = (char*)(((uintptr_t)g.work + 255) & ~(uintptr_t)255);
  a.uin = (double*)w;
  a.flux = a.uin + nb * nv * 216 * 5;
  double* tmp = a.flux + nb * nv * 27 * 15;
  a.ok = (unsigned char*)(tmp + nb * nv * 27 * 6);
  a.nv = nv;
  a.interpol_type = g.c.interpol_type;
  a.interpol_var = g.c.interpol_var;
  a.hp = g.c.hp;
  a.active_coarse = l > 1 ? m->active[l - 1] : nullptr;
  a.noct_coarse = 0;
  a.error = m->error;
  e = ramses::launch_amr_godunov(a, dx, m->dt + l, tmp, g.c.precision == RAMSES_FP32_MIXED, g.s);
  if (e != hipSuccess) return e;
  e = ramses::launch_amr_finish(m->active[l], 0, g.v, m->unew, m->uold, us, g.c.hp.smallr, g.c.interpol_var, g.s,
                                dn + l);
  if (e != hipSuccess) return e;
  if (m->adaptive) {
    const int ic = (l <= m->levelmin || m->nsubcycle[l - 1] == 2) ? icount : 2;
    e = ramses::launch_flag_fine(m->active[l], 0, g.v, l, ic, m->levelmin, m->nlevelmax, m->uold, us, m->flag1,
                                 m->flag2, g.c.rp, g.c.hp, g.s, dn + l);
    if (e != hipSuccess) return e;
  }
  if (l > m->levelmin) {
    e = ramses::launch_amr_dt(2, m->dt, m->dtold, m->time, l, 1, m->nsubcycle[l - 1], icount, 0.0, g.s, dn,
                              g.dev->cnt);
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

// The whole coarse step (amr_step(levelmin, 1) and the end of adaptive_loop's
// step) as launches on g.s, for capture.
hipError_t gcoarse(GCtx& g) {
  ramses_amr_mesh* m = g.m;
  // the refusal bits (4: no free grid, 8: capacity) describe this step only
  hipError_t e = ramses::launch_clear_bits(m->error, 4 | 8, g.s);
  if (e != hipSuccess) return e;
  // adaptive_loop resets the totals at every coarse step (amr/adaptive_loop.f90:89-91)
  if (m->sums && (e = ramses::launch_set_doubles(m->sums, 3, 0.0, g.s)) != hipSuccess) return e;
  e = gstep(g, m->levelmin, 1);
  if (e != hipSuccess) return e;
  if (m->adaptive && m->levelmin < m->nlevelmax) {
    const long us = g.v.ncell + 1;
    for (int l = m->levelmin - 1; l >= 1; --l)
      if ((e = ramses::launch_amr_upload(m->active[l], 0, g.v, m->uold, us, g.c.hp.smallr, g.c.interpol_var, g.s,
                                         g.dev->noct + l)) != hipSuccess)
        return e;
    for (int l = m->levelmin - 1; l >= 1; --l)
      if ((e = ramses::launch_flag_fine(m->active[l], 0, g.v, l, 2, m->levelmin, m->nlevelmax, m->uold, us, m->flag1,
                                        m->flag2, g.c.rp, g.c.hp, g.s, g.dev->noct + l)) != hipSuccess)
        return e;
    if ((e = ramses::launch_set_ints(m->flag1 + 1, m->tree.ncoarse, 1, g.s)) != hipSuccess) return e;
  }
  return hipSuccess;
}

// Level steps a captured coarse step emits: sum over levels of the product
// of the coarser levels' nsubcycle.
long graph_level_steps(const ramses_amr_mesh* m) {
  long tot = 0, mult = 1;
  for (int l = m->levelmin; l <= m->nlevelmax; ++l) {
    tot += mult;
    mult *= m->nsubcycle[l];
  }
  return tot;
}

constexpr long kMaxGraphLevelSteps = 1023;

int coarse_step_graph(ramses_amr_mesh* m, hipStream_t s) {
  const Config c = current();
  if (c.nvector * 27 > 1024) return fail(-8, "nvector too large for the AMR gather (max 37)");
  if (graph_level_steps(m) > kMaxGraphLevelSteps)
    return fail(-1, "captured mode: too many level steps per coarse step (reduce nlevelmax-levelmin)");
  const int capacity = (m->graph_capacity > 0 && m->graph_capacity < m->tree.ngridmax) ? m->graph_capacity
                                                                                      : m->tree.ngridmax;
  for (int l = 1; l <= m->nlevelmax; ++l)
    if (m->noct[l] > capacity) return fail(-10, "captured mode: a level holds more octs than graph_capacity");
  ramses::AmrTreeView v;
  int rc = amr_view(&m->tree, v);
  if (rc) return rc;
  GraphState* gs = static_cast<GraphState*>(m->graph_state);
  hipError_t e;
  if (!gs) {
    gs = new GraphState();
    m->graph_state = gs;
    if ((e = hipStreamCreate(&gs->cap)) != hipSuccess) return hip_fail(e, "capture stream");
    if ((e = hipMalloc(&gs->dev, sizeof(AmrDevState))) != hipSuccess) return hip_fail(e, "graph state");
    if ((e = hipHostMalloc(&gs->host, sizeof(AmrDevState))) != hipSuccess) return hip_fail(e, "graph staging");
  }
  if (gs->capacity != capacity || !gs->work) {
    if (gs->work) {
      if ((e = hipStreamSynchronize(s)) != hipSuccess) return hip_fail(e, "graph workspace sync");
      (void)hipFree(gs->work);
      gs->work = nullptr;
    }
    size_t need = amr_work(capacity, c.nvector);
    const size_t rneed = ramses::refine_work_bytes(capacity, capacity);
    if (rneed > need) need = rneed;
    const size_t pneed = ramses::amr_begin_partials_bytes(0, true) + 256;
    if (pneed > need) need = pneed;
    if ((e = hipMalloc(&gs->work, need)) != hipSuccess) return hip_fail(e, "graph workspace");
    gs->work_bytes = need;
    gs->capacity = capacity;
    gs->drop_graph();
  }
  std::vector<long> key;
  graph_key(m, capacity, key);
  if (!gs->exec || key != gs->key) {
    gs->drop_graph();
    GCtx g{m, gs->cap, c, gs->dev, gs->work, capacity, v, ramses::TreeMut{m->tree.son, m->tree.father, m->tree.nbor}};
    if ((e = hipStreamBeginCapture(gs->cap, hipStreamCaptureModeThreadLocal)) != hipSuccess)
      return hip_fail(e, "begin capture");
    hipError_t ge = gcoarse(g);
    e = hipStreamEndCapture(gs->cap, &gs->graph);
    if (ge != hipSuccess) return hip_fail(ge, "captured coarse step launch");
    if (e != hipSuccess) return hip_fail(e, "end capture");
    if ((e = hipGraphInstantiate(&gs->exec, gs->graph, nullptr, nullptr, 0)) != hipSuccess)
      return hip_fail(e, "graph instantiate");
    gs->key = key;
  }
  if (m->graph == 2) return 0;  // capture only
  AmrDevState* h = gs->host;
  std::memset(h, 0, sizeof(AmrDevState));
  for (int l = 1; l <= m->nlevelmax; ++l) h->noct[l] = m->noct[l];
  h->free[0] = m->free.head;
  h->free[1] = m->free.numbf;
  h->free[2] = m->free.used_mem;
  h->cnt[0] = m->nstep;
  h->cnt[1] = m->cell_updates;
  h->cons_armed = m->nstep_coarse != m->nstep_coarse_old;
  if ((e = hipMemcpyAsync(gs->dev, h, sizeof(AmrDevState), hipMemcpyHostToDevice, s)) != hipSuccess)
    return hip_fail(e, "graph state in");
  if ((e = hipGraphLaunch(gs->exec, s)) != hipSuccess) return hip_fail(e, "graph launch");
  if ((e = hipMemcpyAsync(&gs->dev->error, m->error, sizeof(int), hipMemcpyDeviceToDevice, s)) != hipSuccess)
    return hip_fail(e, "graph error copy");
  if ((e = hipMemcpyAsync(h, gs->dev, sizeof(AmrDevState), hipMemcpyDeviceToHost, s)) != hipSuccess)
    return hip_fail(e, "graph state out");
  if ((e = hipStreamSynchronize(s)) != hipSuccess) return hip_fail(e, "graph step sync");
  for (int l = 1; l <= m->nlevelmax; ++l) m->noct[l] = h->noct[l];
  m->free.head = h->free[0];
  m->free.numbf = h->free[1];
  m->free.used_mem = h->free[2];
  m->nstep = h->cnt[0];
  m->cell_updates = h->cnt[1];
  m->nstep_coarse_old = m->nstep_coarse;  // update_time's bookkeeping (amr/update_time.f90:310)
  m->nstep_coarse += 1;
  if (h->error & 4) return fail(-9, "No more free memory: increase ngridmax");
  if (h->error & 8) return fail(-10, "captured mode: a refinement exceeded graph_capacity");
  return 0;
}

}  // namespace

int ramses_amr_coarse_step(ramses_amr_mesh* m, void* stream) {
  if (!m) return fail(-1, "null mesh");
  if (m->levelmin < 1 || m->nlevelmax < m->levelmin || m->nlevelmax > RAMSES_MAX_LEVELS - 1)
    return fail(-1, "levels out of range");
  if (!m->uold || !m->unew || !m->dt || !m->dtold || !m->time || !m->error || !m->pos || !m->flag1 || !m->flag2)
    return fail(-1, "null mesh array");
  for (int l = 1; l <= m->nlevelmax; ++l) {
    if (m->noct[l] < 0 || m->noct[l] > m->tree.ngridmax) return fail(-1, "noct out of range");
    if (m->noct[l] > 0 && !m->active[l]) return fail(-1, "null grid list");
    if (m->nsubcycle[l] != 1 && m->nsubcycle[l] != 2) return fail(-2, "nsubcycle must be 1 or 2");
  }
  if (!(m->boxlen > 0)) return fail(-1, "boxlen must be > 0");
  if (m->adaptive && m->levelmin < m->nlevelmax && current().nvector != 32)
    return fail(-8, "adaptive coarse steps require nvector = 32 (the batch order of refine_fine's grid numbering)");
  if (m->adaptive && m->levelmin < m->nlevelmax) {
    const ramses_amr_free& f = m->free;
    if (!f.ring) return fail(-1, "null free list");
    if (f.head < 0 || f.head >= m->tree.ngridmax || f.numbf < 0 || f.numbf > m->tree.ngridmax)
      return fail(-1, "free list head/numbf out of range");
  }
  if (m->graph) return coarse_step_graph(m, (hipStream_t)stream);
  StepCtx c{m, (hipStream_t)stream, current(), m->nstep_coarse != m->nstep_coarse_old, false};
  // adaptive_loop resets the totals at every coarse step (amr/adaptive_loop.f90:89-91)
  if (m->sums) {
    hipError_t e = ramses::launch_set_doubles(m->sums, 3, 0.0, c.s);
    if (e != hipSuccess) return hip_fail(e, "sums reset launch");
  }
  int rc = amr_step_rec(c, m->levelmin, 1);
  if (rc) return rc;
  m->nstep_coarse_old = m->nstep_coarse;  // update_time's bookkeeping (amr/update_time.f90:310)
  m->nstep_coarse += 1;
  if (m->adaptive && m->levelmin < m->nlevelmax) {
    for (int l = m->levelmin - 1; l >= 1; --l) {
      ramses_amr_level lv = level_of(m, l);
      if ((rc = ramses_amr_upload(&m->tree, &lv, m->uold, c.s))) return rc;
    }
    for (int l = m->levelmin - 1; l >= 1; --l) {
      ramses_amr_level lv = level_of(m, l);
      rc = ramses_amr_flag_fine(&m->tree, &lv, 2, m->levelmin, m->nlevelmax, m->uold, m->flag1, m->flag2, c.s);
      if (rc) return rc;
    }
    hipError_t e = ramses::launch_set_ints(m->flag1 + 1, m->tree.ncoarse, 1, c.s);  // flag_coarse
    if (e != hipSuccess) return hip_fail(e, "flag_coarse launch");
  }
  return 0;
}

int ramses_amr_mesh_release(ramses_amr_mesh* m) {
  if (!m) return 0;
  if (m->graph_state) {
    GraphState* gs = static_cast<GraphState*>(m->graph_state);
    gs->release();
    delete gs;
    m->graph_state = nullptr;
  }
  if (m->work) {
    hipError_t e = hipFree(m->work);
    m->work = nullptr;
    m->work_bytes = 0;
    if (e != hipSuccess) return hip_fail(e, "hipFree");
  }
  return 0;
}

int ramses_tune(int knob, int value) {
  ramses::Tuning& t = ramses::tuning();
  ++g_config_gen;
  switch (knob) {
    case RAMSES_TUNE_UNSPLIT_BLOCK:
      if (value < 0 || value > 768 || value % 64) return fail(-2, "unsplit block must be a multiple of 64 in [64, 768]");
      t.unsplit_block = value;
      return 0;
    case RAMSES_TUNE_LEVEL_TILE:
      if (value < 0 || value > 5) return fail(-2, "level tile must be 0..5");
      t.level_tile = value;
      return 0;
    case RAMSES_TUNE_LEVEL_TZ:
      if (value != 0 && value < 4) return fail(-2, "level tz must be 0 or >= 4");
      t.level_tz = value;
      return 0;
    case RAMSES_TUNE_PERSIST_BLOCKS:
      if (value != 0 && (value < 32 || value > 8192)) return fail(-2, "persist blocks must be 0 or in [32, 8192]");
      t.persist_blocks = value;
      return 0;
    case RAMSES_TUNE_AMR_FACE_OCTS:
      if (value < -1) return fail(-2, "amr face octs must be -1, 0 or a positive oct count");
      t.amr_face_octs = value;
      return 0;
    default:
      return fail(-1, "unknown tuning knob");
  }
}

int ramses_tune_get(int knob) {
  const ramses::Tuning& t = ramses::tuning();
  switch (knob) {
    case RAMSES_TUNE_UNSPLIT_BLOCK: return t.unsplit_block;
    case RAMSES_TUNE_LEVEL_TILE: return t.level_tile;
    case RAMSES_TUNE_LEVEL_TZ: return t.level_tz;
    case RAMSES_TUNE_PERSIST_BLOCKS: return (int)ramses::persist_blocks();
    case RAMSES_TUNE_AMR_FACE_OCTS: return ramses::amr_face_octs();
    default: return fail(-1, "unknown tuning knob");
  }
}

int ramses_level_slot_init(const ramses_level_slot* slot, double t_end, void* stream) {
  const Config c = current();
  if (!slot || !slot->buf[0] || !slot->buf[1] || !slot->state) return fail(-1, "null slot pointer");
  if (slot->n < 16 || slot->n % 16 != 0) return fail(-1, "slot n must be a multiple of 16");
  if (!(slot->dx > 0)) return fail(-1, "dx must be > 0");
  const double boxlen = slot->n * slot->dx;
  double start = boxlen / c.hp.smallc;
  const double cap = c.hp.courant_factor * slot->dx / c.hp.smallc;
  if (cap < start) start = cap;
  hipError_t e = ramses::launch_init_slot(slot->state, t_end, start, (hipStream_t)stream);
  if (e != hipSuccess) return hip_fail(e, "slot init launch");
  e = ramses::launch_courant(slot->buf[0], slot->n, slot->dx, c.hp, slot->state + 3, (hipStream_t)stream);
  if (e != hipSuccess) return hip_fail(e, "slot courant launch");
  return 0;
}

int ramses_level_batch_step(const ramses_level_slot* slots, int nslots, void* stream) {
  const Config c = current();
  if (nslots < 0 || nslots > RAMSES_MAX_SLOTS) return fail(-1, "nslots out of [0, RAMSES_MAX_SLOTS]");
  if (nslots == 0) return 0;
  if (!slots) return fail(-1, "null slots");
  if (c.hp.slope_type == 3) return fail(-4, "slope_type 3 is not supported on the resident level path");
  double* b0[RAMSES_MAX_SLOTS];
  double* b1[RAMSES_MAX_SLOTS];
  double* st[RAMSES_MAX_SLOTS];
  int n[RAMSES_MAX_SLOTS];
  double dx[RAMSES_MAX_SLOTS];
  long blocks = 0;
  for (int i = 0; i < nslots; ++i) {
    const ramses_level_slot& s = slots[i];
    if (!s.buf[0] || !s.buf[1] || !s.state) return fail(-1, "null slot pointer");
    if (s.n < 16 || s.n % 16 != 0) return fail(-1, "slot n must be a multiple of 16");
    b0[i] = s.buf[0];
    b1[i] = s.buf[1];
    st[i] = s.state;
    n[i] = s.n;
    dx[i] = s.dx;
    const int tz = ramses::batch_tz(s.n);
    blocks += (long)(s.n / 16) * (s.n / 8) * (s.n / tz);
  }
  if (blocks > 2147483647L) return fail(-1, "batch too large");
  hipError_t e = ramses::launch_level_batch(b0, b1, st, n, dx, nslots, c.hp, c.precision == RAMSES_FP32_MIXED,
                                           (hipStream_t)stream);
  if (e != hipSuccess) return hip_fail(e, "level batch launch");
  return 0;
}

int ramses_uniform_dt_init(double* d_dt, double value, void* stream) {
  if (!d_dt) return fail(-1, "null dt pointer");
  hipError_t e = ramses::launch_set_double(d_dt, value, (hipStream_t)stream);
  if (e != hipSuccess) return hip_fail(e, "dt init launch");
  return 0;
}

unsigned long ramses_uniform_work_bytes(int n) {
  return n < 2 ? 0 : (unsigned long)ramses::uniform_sums_work_bytes(n);
}

int ramses_uniform_courant_sums(const double* u, int n, double dx, double dt_start, double* d_dt, double* d_sums,
                                void* work, unsigned long work_bytes, void* stream) {
  const Config c = current();
  if (!u || !d_dt) return fail(-1, "null pointer");
  if (n < 2) return fail(-1, "n must be >= 2");
  if (!(dx > 0)) return fail(-1, "dx must be > 0");
  if (d_sums && (!work || work_bytes < ramses::uniform_sums_work_bytes(n)))
    return fail(-1, "work buffer too small (see ramses_uniform_work_bytes)");
  // courant_fine starts from dtnew (dt_start); every cmpdt batch starts from
  // courant_factor*dx/smallc (hydro/courant_fine.f90:35, godunov_utils.f90:114)
  double start = c.hp.courant_factor * dx / c.hp.smallc;
  if (dt_start < start) start = dt_start;
  hipError_t e = ramses::launch_set_double(d_dt, start, (hipStream_t)stream);
  if (e != hipSuccess) return hip_fail(e, "dt init launch");
  e = ramses::launch_courant(u, n, dx, c.hp, d_dt, (hipStream_t)stream, d_sums ? (double*)work : nullptr, d_sums);
  if (e != hipSuccess) return hip_fail(e, "courant launch");
  return 0;
}

int ramses_uniform_courant(const double* u, int n, double dx, double dt_start, double* d_dt, void* stream) {
  return ramses_uniform_courant_sums(u, n, dx, dt_start, d_dt, nullptr, nullptr, 0, stream);
}

int ramses_uniform_godunov_sums(const double* u, double* unew, int n, double dx, const double* d_dt,
                                double* d_dt_next, double* d_sums, void* work, unsigned long work_bytes,
                                void* stream) {
  const Config c = current();
  if (!u || !unew || !d_dt) return fail(-1, "null pointer");
  if (u == unew) return fail(-1, "unew must not alias u");
  if (n < 8 || n % 8 != 0) return fail(-1, "n must be a multiple of 8 (>= 8)");
  if (!(dx > 0)) return fail(-1, "dx must be > 0");
  if (c.hp.slope_type == 3) return fail(-4, "slope_type 3 is not supported on the resident level path");
  if (d_sums && (!work || work_bytes < ramses::uniform_sums_work_bytes(n)))
    return fail(-1, "work buffer too small (see ramses_uniform_work_bytes)");
  hipError_t e = ramses::launch_uniform_godunov(u, unew, n, dx, d_dt, d_dt_next, c.hp,
                                                c.precision == RAMSES_FP32_MIXED, (hipStream_t)stream,
                                                d_sums ? (double*)work : nullptr, d_sums);
  if (e != hipSuccess) return hip_fail(e, "uniform godunov launch");
  return 0;
}

int ramses_uniform_godunov(const double* u, double* unew, int n, double dx, const double* d_dt, double* d_dt_next,
                           void* stream) {
  return ramses_uniform_godunov_sums(u, unew, n, dx, d_dt, d_dt_next, nullptr, nullptr, 0, stream);
}

}  // extern "C"
