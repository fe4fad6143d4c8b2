// Mesh refinement on the GPU: the refinement map of a level (flag_fine,
// amr/flag_utils.f90:57-106) and the creation / destruction of the octs of
// the next level (refine_fine, amr/refine_utils.f90:332-585), with the tree
// (son / father / nbor), the flags and the free-grid list resident in HBM.
//
// flag_fine      flag_init   test_flag: a refined cell stays flagged while a
//                            child is refined or flagged (:168-208)
//                flag_smooth one Jacobi pass of smooth_fine (:555-677),
//                            flag1 and flag2 used as ping-pong buffers
//                flag_hydro  hydro_flag / hydro_refine gradient criteria
//                            (hydro/hydro_flag.f90:1-180,
//                            hydro/godunov_utils.f90:125-263)
//                flag_ensure ensure_ref_rules (:213-276)
// refine_fine    MarkSrc     one mark per (batch, ind, lane) slot, the
//                            order in which the reference visits cells, so
//                            an exclusive scan of the marks gives every new
//                            or killed oct its rank in the reference's order
//                            (the marks are computed inside their scan)
//                refine_make make_grid_fine (:590-948): the k-th new oct takes
//                            the k-th free grid from the head of the free
//                            list, links to its father and neighbour cells,
//                            and its cells are prolongated (interpol_hydro)
//                refine_kill kill_grid (:953-1158): the k-th killed oct is
//                            cleared and appended to the tail of the free list
//                list_*      the level's grid list rebuilt as build_comm does
//                            from the linked list (amr/virtual_boundaries.f90:
//                            1286-1420): survivors in order, then new octs in
//                            creation order.
// The free list of the reference is a FIFO (taken at headf, returned at
// tailf); here it is a ring of ngridmax grid indices.
#include "hydro_device.hpp"
#include "amr_args.hpp"
#include "amr_device.hpp"

namespace ramses {

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

namespace {

// --- flag_fine ----------------------------------------------------------

// Cells of the level in cell-major order: i -> (ind = i / noct + 1, a = i %
// noct).  The flag kernels loop over them grid-stride; the oct count is the
// host value or, with dn set, the one in device memory.
__device__ __forceinline__ void level_cell(const int* active, int noct, const AmrTreeView& t, long i, int& g,
                                           int& ind, long& c) {
  ind = (int)(i / noct) + 1;
  g = active[i % noct];
  c = cell_of(t, ind, g);
}

#define FOR_LEVEL_CELLS(noct_h, dn)                                                             \
  const int noct = level_count(noct_h, dn);                                                     \
  const long ncells_ = 8L * noct, stride_ = (long)gridDim.x * blockDim.x;                       \
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < ncells_; i += stride_)

__global__ void flag_init(const int* active, int noct_h, const int* dn, AmrTreeView t, int* flag1, int all) {
  FOR_LEVEL_CELLS(noct_h, dn) {
    int g, ind;
    long c;
    level_cell(active, noct, t, i, g, ind, c);
    if (all) {  // below levelmin every cell is flagged (init_flag :150-157)
      flag1[c] = 1;
      continue;
    }
    const int s = t.son[c];
    int f = 0;
    if (s > 0)
      for (int k = 1; k <= 8; ++k) {
        const long cc = cell_of(t, k, s);
        if (t.son[cc] > 0 || flag1[cc] == 1) f = 1;
      }
    flag1[c] = f;
  }
}

// One pass of smooth_fine's dilation (amr/flag_utils.f90:555-677): a cell
// is flagged when it was, or when at least nthr of its six neighbours at
// the level were (flag1[0] == 0 stands for a missing neighbour).  The
// reference counts into flag2 and merges into flag1; reading `in` and
// writing `out` gives the same Jacobi pass in one launch.
__global__ void flag_smooth(const int* active, int noct_h, const int* dn, AmrTreeView t, const int* in, int* out,
                            int nthr) {
  FOR_LEVEL_CELLS(noct_h, dn) {
    int g, ind;
    long c;
    level_cell(active, noct, t, i, g, ind, c);
    int n = 0;
#pragma unroll
    for (int nb = 1; nb <= 6; ++nb) n += in[nbor_cell(t, g, ind, nb)];
    out[c] = (in[c] == 1 || n >= nthr) ? 1 : 0;
  }
}

// Copy of the level's flags from one buffer to the other.
__global__ void flag_copy(const int* active, int noct_h, const int* dn, AmrTreeView t, const int* in, int* out) {
  FOR_LEVEL_CELLS(noct_h, dn) {
    int g, ind;
    long c;
    level_cell(active, noct, t, i, g, ind, c);
    out[c] = in[c];
  }
}

struct Prim3 {
  double d, v[3], p;
};

// hydro_refine's conversion to primitives (godunov_utils.f90:149-180).
__device__ __forceinline__ Prim3 refine_prim(const double* u, long us, long c, const HydroParams& hp) {
  Prim3 q;
  q.d = fmax_f(u[c], hp.smallr);
  for (int d = 0; d < 3; ++d) q.v[d] = u[(long)(d + 1) * us + c] / q.d;
  double ek = 0.0;
  for (int d = 0; d < 3; ++d) ek = ek + (0.5 * q.d) * (q.v[d] * q.v[d]);
  q.p = (hp.gamma - 1.0) * (u[4 * us + c] - ek);
  return q;
}

__device__ __forceinline__ double grad_err(double g, double m, double d, double floor_x) {
  return 2.0 * fmax_f(__builtin_fabs((d - m) / ((d + m) + floor_x)), __builtin_fabs((m - g) / ((m + g) + floor_x)));
}

__device__ void flag_hydro_cell(const int* active, int noct, const AmrTreeView& t, const double* u, long us,
                                int* flag1, const RefineParams& rp, const HydroParams& hp, long i) {
  int g, ind;
  long c;
  level_cell(active, noct, t, i, g, ind, c);
  // The six neighbours' states are loaded before any test (independent
  // loads, issued together); the tests have no side effects, so OR-ing all
  // of them gives the flag of the reference's short-circuit order.
  const Prim3 m = refine_prim(u, us, c, hp);
  long cn[6];
#pragma unroll
  for (int k = 0; k < 6; ++k) {
    const long cc = nbor_cell(t, g, ind, k + 1);
    cn[k] = cc != 0 ? cc : nbor_of(t, g, k + 1);  // missing neighbour: the father's neighbour
  }
  Prim3 nq[6];
#pragma unroll
  for (int k = 0; k < 6; ++k) nq[k] = refine_prim(u, us, cn[k], hp);
  bool ok = false;
#pragma unroll
  for (int idim = 0; idim < 3; ++idim) {
    const Prim3& l = nq[2 * idim];
    const Prim3& r = nq[2 * idim + 1];
    if (rp.err_grad_d >= 0.0) ok = ok | (grad_err(l.d, m.d, r.d, rp.floor_d) > rp.err_grad_d);
    if (rp.err_grad_p >= 0.0) ok = ok | (grad_err(l.p, m.p, r.p, rp.floor_p) > rp.err_grad_p);
    if (rp.err_grad_u >= 0.0) {
      const double fu2 = rp.floor_u * rp.floor_u;
      const double cg = __builtin_sqrt(fmax_f(hp.gamma * l.p / l.d, fu2));
      const double cm = __builtin_sqrt(fmax_f(hp.gamma * m.p / m.d, fu2));
      const double cd = __builtin_sqrt(fmax_f(hp.gamma * r.p / r.d, fu2));
      for (int d = 0; d < 3; ++d) {
        const double vg = l.v[d], vm = m.v[d], vd = r.v[d];
        const double e1 = __builtin_fabs((vd - vm) / ((((cd + cm) + __builtin_fabs(vd)) + __builtin_fabs(vm)) + rp.floor_u));
        const double e2 = __builtin_fabs((vm - vg) / ((((cm + cg) + __builtin_fabs(vm)) + __builtin_fabs(vg)) + rp.floor_u));
        ok = ok | (2.0 * fmax_f(e1, e2) > rp.err_grad_u);
      }
    }
  }
  if (ok) flag1[c] = 1;
}

__global__ void flag_hydro(const int* active, int noct_h, const int* dn, AmrTreeView t, const double* u, long us,
                           int* flag1, RefineParams rp, HydroParams hp) {
  FOR_LEVEL_CELLS(noct_h, dn) flag_hydro_cell(active, noct, t, u, us, flag1, rp, hp, i);
}

// Thread = (grid, father neighbour j of 27): a missing or unrefined
// neighbour unflags all cells of the grid (every such thread writes the
// same zeros), as ensure_ref_rules does for a grid with any such neighbour.
__global__ void flag_ensure(const int* active, int noct_h, const int* dn, AmrTreeView t, int ilevel, int* flag1) {
  const long n = 27L * level_count(noct_h, dn), stride = (long)gridDim.x * blockDim.x;
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    const int g = active[i / 27], j = (int)(i % 27);
    const long nf = nbor_father_cell(t, t.father[g], j % 3, (j / 3) % 3, j / 9, ilevel);
    if (nf == 0 || t.son[nf] == 0)
      for (int ind = 1; ind <= 8; ++ind) flag1[cell_of(t, ind, g)] = 0;
  }
}

// --- exclusive scan of 64-bit counters ----------------------------------

constexpr int kScanThreads = 256;
constexpr int kScanItems = 8;
constexpr int kScanTile = kScanThreads * kScanItems;

__device__ long long block_excl_scan(long long x, long long* sh, long long& total) {
  // wave-level inclusive scan, then across the block's waves
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  long long v = x;
  for (int o = 1; o < 64; o <<= 1) {
    const long long y = __shfl_up(v, o, 64);
    if (lane >= o) v += y;
  }
  if (lane == 63) sh[wid] = v;
  __syncthreads();
  long long base = 0, tot = 0;
  const int nw = blockDim.x >> 6;
  for (int w = 0; w < nw; ++w) {
    if (w < wid) base += sh[w];
    tot += sh[w];
  }
  __syncthreads();
  total = tot;
  return base + v - x;
}

// Length of a scanned array: a host value, or derived from a device oct count
// (kind 1: the count itself; kind 2: the refine slots of that many octs).
struct DevLen {
  long n;
  const int* d;
  int kind;
  __device__ __forceinline__ long get() const {
    if (!d) return n;
    const long c = *d;
    return kind == 2 ? ((c + 31) / 32) * 256 : c;
  }
};

// The scanned values: an array, or computed by the kernel that scans them
// (the producer stores each value too, for the scan's consumers).
struct ArraySrc {
  const long long* in;
  __device__ __forceinline__ bool skip() const { return false; }
  __device__ __forceinline__ long long operator()(long i) const { return in[i]; }
};

// Exclusive scan in two launches: scan_tiles leaves out[i] = the sum of the
// values before i within its tile of kScanTile and sums[tile] = the tile's
// total; scan_sums turns sums[] into the tiles' exclusive offsets (and
// sums[ntiles] = the grand total).  The consumers read out[i] + sums[i /
// kScanTile] (scan_offset), so no third pass adds the offsets.  Block b
// scans tiles b, b + gridDim.x, ...
template <class Src>
__global__ __launch_bounds__(kScanThreads) void scan_tiles(Src src, long long* out, long long* sums, DevLen len) {
  __shared__ long long sh[kScanThreads / 64];
  if (src.skip()) return;
  const long n = len.get();
  const long ntiles = (n + kScanTile - 1) / kScanTile;
  for (long tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
    const long base = tile * kScanTile + (long)threadIdx.x * kScanItems;
    long long v[kScanItems], s = 0;
    for (int k = 0; k < kScanItems; ++k) {
      v[k] = base + k < n ? src(base + k) : 0;
      s += v[k];
    }
    long long total;
    long long run = block_excl_scan(s, sh, total);
    for (int k = 0; k < kScanItems; ++k) {
      if (base + k < n) out[base + k] = run;
      run += v[k];
    }
    if (threadIdx.x == 0) {
      sums[tile] = total;
      if (gridDim.x == 1 && ntiles == 1) {  // one tile: its exclusive offsets are final
        sums[0] = 0;
        sums[1] = total;
      }
    }
  }
}

// One block: exclusive scan of the tile totals; sums[ntiles] = grand total.
__global__ __launch_bounds__(1024) void scan_sums(long long* sums, DevLen len) {
  __shared__ long long sh[1024 / 64];
  const long ntiles = (len.get() + kScanTile - 1) / kScanTile;
  long long carry = 0;
  for (long b0 = 0; b0 < ntiles; b0 += 1024) {
    const long i = b0 + threadIdx.x;
    const long long x = i < ntiles ? sums[i] : 0;
    long long total;
    const long long e = block_excl_scan(x, sh, total);
    if (i < ntiles) sums[i] = carry + e;
    carry += total;
  }
  if (threadIdx.x == 0) sums[ntiles] = carry;
}

__device__ __forceinline__ long long scan_offset(const long long* out, const long long* sums, long i) {
  return out[i] + sums[i / kScanTile];
}

// --- refine_fine ----------------------------------------------------------

// Slot p = batch*256 + (ind-1)*32 + lane over the level's active list in
// nvector=32 batches.  mark = create (low 32 bits) | kill << 32, computed
// and stored by the scan of the marks.
struct MarkSrc {
  const int* active;
  int noct_h;
  const int* dn;
  AmrTreeView t;
  const int* flag1;
  long long* mark;
  __device__ __forceinline__ bool skip() const { return false; }
  __device__ __forceinline__ long long operator()(long p) const {
    const int noct = level_count(noct_h, dn);
    const int b = (int)(p >> 8), r = (int)(p & 255);
    const int ind = (r >> 5) + 1, lane = r & 31;
    const int a = b * 32 + lane;
    long long m = 0;
    if (a < noct) {
      const long c = cell_of(t, ind, active[a]);
      const int s = t.son[c], f = flag1[c];
      if (f == 1 && s == 0) m = 1;
      if (f == 0 && s > 0) m = 1LL << 32;
    }
    mark[p] = m;
    return m;
  }
};

struct RefineRun {
  AmrTreeView t;      // read view of the same arrays as m
  TreeMut m;
  const int* active;  // level ilevel
  int noct;
  const long long* mark;
  const long long* off;   // exclusive scan of mark within tiles
  const long long* sums;  // the tiles' offsets (scan_offset)
  long nslot;
  int* ring;
  int cap;          // ngridmax
  int head;         // ring index of headf
  int numbf;        // free grids before this refine
  int* newg;        // created grids in creation order
  double* uold;
  double* unew;
  long us;
  int* flag1;
  int* flag2;
  double* xg;       // grid centres xg[3*igrid + d], or null
  double dx;        // cell size of the level being refined (0.5^ilevel)
  int interpol_type, interpol_var;
  double smallr;
  int* error;
  // device-decided refinement (no host synchronisation): the level's oct
  // count, the free list {head, numbf} and the decision {ncreate, nkill, ok}
  // are read from device memory when these are set
  const int* dnoct;
  const int* dfree;
  const int* ctrl;
  __device__ __forceinline__ long slots() const {
    return dnoct ? (long)((*dnoct + 31) / 32) * 256 : nslot;
  }
  __device__ __forceinline__ int head_() const { return dfree ? dfree[0] : head; }
  __device__ __forceinline__ int numbf_() const { return dfree ? dfree[1] : numbf; }
};

__device__ void make_slot(const RefineRun& r, long p) {
  if ((r.mark[p] & 0xffffffffLL) == 0) return;
  const int b = (int)(p >> 8), q = (int)(p & 255);
  const int ind = (q >> 5) + 1, lane = q & 31;
  const int igrid = r.active[b * 32 + lane];
  const long c = cell_of(r.t, ind, igrid);
  const int k = (int)(scan_offset(r.off, r.sums, p) & 0xffffffffLL);
  const int g = r.ring[(int)(((long)r.head_() + k) % r.cap)];
  r.newg[k] = g;
  r.m.son[c] = g;
  r.m.father[g] = (int)c;
  if (r.xg) {  // centre of the new oct = centre of its father cell (make_grid_fine :663-673)
    const int p[3] = {(ind - 1) & 1, ((ind - 1) >> 1) & 1, (ind - 1) >> 2};
    for (int d = 0; d < 3; ++d) r.xg[3L * g + d] = r.xg[3L * igrid + d] + ((double)p[d] - 0.5) * r.dx;
  }
  long nb[7];
  nb[0] = c;
  for (int in = 1; in <= 6; ++in) {
    const long n = nbor_cell(r.t, igrid, ind, in);
    if (n == 0) atomicOr(r.error, 2);  // "Fatal error in make_grid_fine"
    nb[in] = n;
    r.m.nbor[(long)(in - 1) * r.cap + g] = (int)n;
  }
  double u1[7][5], u2[8][5];
  for (int j = 0; j < 7; ++j)
    for (int v = 0; v < 5; ++v) u1[j][v] = r.uold[(long)v * r.us + nb[j]];
  interpol_oct(u1, u2, r.interpol_type, r.interpol_var, r.smallr);
  for (int j = 1; j <= 8; ++j) {
    const long cc = cell_of(r.t, j, g);
    for (int v = 0; v < 5; ++v) r.uold[(long)v * r.us + cc] = u2[j - 1][v];
    r.m.son[cc] = 0;
    r.flag1[cc] = 0;
  }
}

__global__ void refine_make(RefineRun r) {
  if (r.ctrl && !r.ctrl[2]) return;
  const long n = r.slots(), stride = (long)gridDim.x * blockDim.x;
  for (long p = (long)blockIdx.x * blockDim.x + threadIdx.x; p < n; p += stride) make_slot(r, p);
}

__device__ void kill_slot(const RefineRun& r, long p) {
  if ((r.mark[p] >> 32) == 0) return;
  const int b = (int)(p >> 8), q = (int)(p & 255);
  const int ind = (q >> 5) + 1, lane = q & 31;
  const long c = cell_of(r.t, ind, r.active[b * 32 + lane]);
  const int k = (int)(scan_offset(r.off, r.sums, p) >> 32);
  const int g = r.m.son[c];
  r.m.son[c] = 0;
  r.m.father[g] = 0;
  if (r.xg)
    for (int d = 0; d < 3; ++d) r.xg[3L * g + d] = 0.0;
  for (int in = 1; in <= 6; ++in) r.m.nbor[(long)(in - 1) * r.cap + g] = 0;
  for (int j = 1; j <= 8; ++j) {
    const long cc = cell_of(r.t, j, g);
    r.m.son[cc] = 0;
    r.flag1[cc] = 0;
    r.flag2[cc] = 0;
    for (int v = 0; v < 5; ++v) {
      r.uold[(long)v * r.us + cc] = 0.0;
      r.unew[(long)v * r.us + cc] = 0.0;
    }
  }
  r.ring[(int)(((long)r.head_() + r.numbf_() + k) % r.cap)] = g;
}

__global__ void refine_kill(RefineRun r) {
  if (r.ctrl && !r.ctrl[2]) return;
  const long n = r.slots(), stride = (long)gridDim.x * blockDim.x;
  for (long p = (long)blockIdx.x * blockDim.x + threadIdx.x; p < n; p += stride) kill_slot(r, p);
}

// The refinement decision on the device, from the scanned marks: ctrl =
// {ncreate, nkill, ok}; ok is 0 when there is nothing to do or when the
// free list cannot hold the new octs ("No more free memory",
// amr/refine_utils.f90:415-423: error bit 4, nothing is modified).
// A refinement that would take the next level beyond `capacity` octs (the
// workspace of a captured step) is refused the same way (error bit 8).
__global__ void refine_decide(const long long* sums, DevLen len, const int* dfree, const int* dn_f, int capacity,
                              int* ctrl, int* error) {
  const long ntiles = (len.get() + kScanTile - 1) / kScanTile;
  const long long tot = sums[ntiles];
  const int nc = (int)(tot & 0xffffffffLL), nk = (int)(tot >> 32);
  int ok = (nc > 0 || nk > 0) ? 1 : 0;
  if (ok && nc >= dfree[1]) {
    ok = 0;
    atomicOr(error, 4);
  }
  if (ok && (long)(nk > 0 ? *dn_f - nk : *dn_f) + nc > capacity) {
    ok = 0;
    atomicOr(error, 8);
  }
  ctrl[0] = nc;
  ctrl[1] = nk;
  ctrl[2] = ok;
}

// Survivors of the next level's list (father still set) in list order,
// computed and stored by their scan.  With ctrl (the device-decided
// refinement), nothing is scanned unless octs were killed: the list then
// keeps its order.
struct AliveSrc {
  const int* list;
  const int* father;
  long long* mark;
  const int* ctrl;
  __device__ __forceinline__ bool skip() const { return ctrl && (!ctrl[2] || ctrl[1] == 0); }
  __device__ __forceinline__ long long operator()(long i) const {
    const long long m = father[list[i]] != 0 ? 1 : 0;
    mark[i] = m;
    return m;
  }
};

__global__ void list_compact(const int* list, int n, const long long* mark, const long long* off,
                             const long long* sums, int* tmp) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n && mark[i]) tmp[scan_offset(off, sums, i)] = list[i];
}

// list[0..nalive) = tmp[0..nalive) (if tmp), list[nalive..) = newg; pos follows.
__global__ void list_assemble(int* list, const int* tmp, int nalive, const int* newg, int ncreate, int* pos) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nalive + ncreate) return;
  if (i < nalive && !tmp) return;  // survivors already in place
  const int g = i < nalive ? tmp[i] : newg[i - nalive];
  list[i] = g;
  pos[g] = i;
}

// The same list maintenance with the counts and the decision in device
// memory: dnf = the next level's count before this refinement.
__global__ void list_compact_d(const int* list, const int* dnf, const long long* mark, const long long* off,
                               const long long* sums, int* tmp, const int* ctrl) {
  if (!ctrl[2] || ctrl[1] == 0) return;  // nothing killed: the list keeps its order
  const long n = *dnf, stride = (long)gridDim.x * blockDim.x;
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
    if (mark[i]) tmp[scan_offset(off, sums, i)] = list[i];
}

__global__ void list_assemble_d(int* list, const int* tmp, const int* dnf, const int* newg, int* pos,
                                const int* ctrl) {
  if (!ctrl[2]) return;
  const int nf = *dnf, nc = ctrl[0], nk = ctrl[1];
  const int nalive = nk > 0 ? nf - nk : nf;
  const long n = (long)nalive + nc, stride = (long)gridDim.x * blockDim.x;
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    if (i < nalive && nk == 0) continue;  // survivors already in place
    const int g = i < nalive ? tmp[i] : newg[i - nalive];
    list[i] = g;
    pos[g] = (int)i;
  }
}

// The new counts: next level, free-list head / numbf, used_mem.
__global__ void refine_commit(int* dnf, int* dfree, const int* ctrl, int cap) {
  if (!ctrl[2]) return;
  const int nf = *dnf, nc = ctrl[0], nk = ctrl[1];
  *dnf = (nk > 0 ? nf - nk : nf) + nc;
  const int numbf = dfree[1];
  dfree[0] = (int)(((long)dfree[0] + nc) % cap);
  if (nc > 0) dfree[2] = cap - (numbf - nc);  // make_grid_fine's last used_mem
  dfree[1] = numbf - nc + nk;
}

inline unsigned nblk(long n, int b) { return (unsigned)((n + b - 1) / b); }

// The scan of n values from src (host length): one tile scans into its
// final offsets (sums = {0, total}); more tiles also scan the tile totals.
template <class Src>
hipError_t scan64(Src src, long long* out, long long* sums, long n, hipStream_t s) {
  const long nt = (n + kScanTile - 1) / kScanTile;
  if (nt == 0) return hipSuccess;
  const DevLen len{n, nullptr, 0};
  hipLaunchKernelGGL(scan_tiles<Src>, dim3((unsigned)nt), dim3(kScanThreads), 0, s, src, out, sums, len);
  if (nt > 1) hipLaunchKernelGGL(scan_sums, dim3(1), dim3(1024), 0, s, sums, len);
  return hipGetLastError();
}

// The scan with its length in device memory: a persistent grid, and the
// tile totals always scanned so that sums[ntiles] is the grand total.
template <class Src>
hipError_t scan64_dev(Src src, long long* out, long long* sums, DevLen len, hipStream_t s) {
  hipLaunchKernelGGL(scan_tiles<Src>, dim3(persist_blocks() / 2), dim3(kScanThreads), 0, s, src, out, sums, len);
  hipLaunchKernelGGL(scan_sums, dim3(1), dim3(1024), 0, s, sums, len);
  return hipGetLastError();
}

}  // namespace

// flag_fine(ilevel, icount) for one level (amr/flag_utils.f90:57-106).  With
// dn the oct count is read from device memory (persistent grids).
hipError_t launch_flag_fine(const int* active, int noct, const AmrTreeView& t, int ilevel, int icount, int levelmin,
                            int nlevelmax, const double* u, long us, int* flag1, int* flag2, const RefineParams& rp,
                            const HydroParams& hp, hipStream_t s, const int* dn) {
  if (ilevel == nlevelmax || (!dn && noct <= 0)) return hipSuccess;
  const unsigned nb = dn ? persist_blocks() : nblk(8L * noct, 256);
  hipLaunchKernelGGL(flag_init, dim3(nb), dim3(256), 0, s, active, noct, dn, t, flag1, ilevel < levelmin ? 1 : 0);
  if (ilevel < levelmin) return hipGetLastError();
  // the flags ping-pong between flag1 and flag2, one launch per pass
  int* cur = flag1;
  int* other = flag2;
  auto smooth = [&]() {
    const int nthr[3] = {1, 2, 2};
    for (int k = 0; k < 3; ++k) {
      hipLaunchKernelGGL(flag_smooth, dim3(nb), dim3(256), 0, s, active, noct, dn, t, (const int*)cur, other,
                         nthr[k]);
      int* tmp = cur;
      cur = other;
      other = tmp;
    }
  };
  smooth();
  if (rp.err_grad_d >= 0.0 || rp.err_grad_p >= 0.0 || rp.err_grad_u >= 0.0)
    hipLaunchKernelGGL(flag_hydro, dim3(nb), dim3(256), 0, s, active, noct, dn, t, u, us, cur, rp, hp);
  for (int e = 0; e < rp.nexpand; ++e) smooth();
  if (cur != flag1)
    hipLaunchKernelGGL(flag_copy, dim3(nb), dim3(256), 0, s, active, noct, dn, t, (const int*)cur, flag1);
  if (ilevel > levelmin && icount < 2)
    hipLaunchKernelGGL(flag_ensure, dim3(dn ? persist_blocks() : nblk(27L * noct, 256)), dim3(256), 0, s, active, noct, dn,
                       t, ilevel, flag1);
  return hipGetLastError();
}

// Workspace of refine_fine for a level of noct octs whose next level has
// noct_fine octs.
size_t refine_work_bytes(int noct, int noct_fine) {
  const long nslot = (long)((noct + 31) / 32) * 256;
  const long nmax = nslot > noct_fine ? nslot : noct_fine;
  const long nt = (nmax + kScanTile - 1) / kScanTile + 1;  // tile totals + grand total
  return (size_t)(2 * nmax + nt) * sizeof(long long) + (size_t)(nslot + noct_fine) * sizeof(int) + 256;
}

// refine_fine(ilevel) (amr/refine_utils.f90:332-585).  The one host
// synchronisation reads the number of octs to create and to kill.
// On return *noct_fine, *head, *numbf describe the rebuilt next level and
// the free list.  *nofree = 1 (and nothing is modified) when the free list
// cannot hold the new octs.
hipError_t launch_refine_fine(const int* active, int noct, int ilevel, const AmrTreeView& t, const TreeMut& m,
                              int* active_fine, int* noct_fine, int* pos, int* ring, int* head, int* numbf,
                              int* used_mem, int* flag1, int* flag2, double* xg, double* uold, double* unew,
                              int interpol_type, int interpol_var, double smallr, void* work, int* d_error,
                              int* nofree, hipStream_t s) {
  *nofree = 0;
  if (noct <= 0) return hipSuccess;
  const long nslot = (long)((noct + 31) / 32) * 256;
  const int nf = *noct_fine;
  const long nmax = nslot > nf ? nslot : nf;
  const long nt = (nmax + kScanTile - 1) / kScanTile + 1;
  char* w = (char*)(((uintptr_t)work + 255) & ~(uintptr_t)255);
  long long* mark = (long long*)w;
  long long* off = mark + nmax;
  long long* sums = off + nmax;  // nt entries
  int* newg = (int*)(sums + nt);
  int* tmp = newg + nslot;
  hipError_t e = scan64(MarkSrc{active, noct, nullptr, t, flag1, mark}, off, sums, nslot, s);
  if (e != hipSuccess) return e;
  const long ntiles = (nslot + kScanTile - 1) / kScanTile;
  long long tot = 0;
  if ((e = hipMemcpyAsync(&tot, sums + ntiles, sizeof(tot), hipMemcpyDeviceToHost, s)) != hipSuccess) return e;
  if ((e = hipStreamSynchronize(s)) != hipSuccess) return e;
  const int ncreate = (int)(tot & 0xffffffffLL), nkill = (int)(tot >> 32);
  if (ncreate == 0 && nkill == 0) return hipSuccess;
  if (ncreate >= *numbf) {  // "No more free memory" (:415-423)
    *nofree = 1;
    return hipSuccess;
  }
  RefineRun r;
  r.t = t;
  r.m = m;
  r.active = active;
  r.noct = noct;
  r.mark = mark;
  r.off = off;
  r.sums = sums;
  r.nslot = nslot;
  r.ring = ring;
  r.cap = t.ngridmax;
  r.head = *head;
  r.numbf = *numbf;
  r.newg = newg;
  r.uold = uold;
  r.unew = unew;
  r.us = t.ncell + 1;
  r.flag1 = flag1;
  r.flag2 = flag2;
  r.xg = xg;
  r.dx = __builtin_ldexp(1.0, -ilevel);
  r.interpol_type = interpol_type;
  r.interpol_var = interpol_var;
  r.smallr = smallr;
  r.error = d_error;
  r.dnoct = nullptr;
  r.dfree = nullptr;
  r.ctrl = nullptr;
  if (ncreate > 0) hipLaunchKernelGGL(refine_make, dim3(nblk(nslot, 128)), dim3(128), 0, s, r);
  if (nkill > 0) hipLaunchKernelGGL(refine_kill, dim3(nblk(nslot, 256)), dim3(256), 0, s, r);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  int nalive = nf;
  if (nkill > 0 && nf > 0) {
    if ((e = scan64(AliveSrc{active_fine, m.father, mark, nullptr}, off, sums, nf, s)) != hipSuccess) return e;
    hipLaunchKernelGGL(list_compact, dim3(nblk(nf, 256)), dim3(256), 0, s, (const int*)active_fine, nf,
                       (const long long*)mark, (const long long*)off, (const long long*)sums, tmp);
    nalive = nf - nkill;
    if (nalive + ncreate > 0)
      hipLaunchKernelGGL(list_assemble, dim3(nblk(nalive + ncreate, 256)), dim3(256), 0, s, active_fine,
                         (const int*)tmp, nalive, (const int*)newg, ncreate, pos);
  } else if (ncreate > 0) {
    // nothing killed: survivors stay in place, the new octs are appended
    hipLaunchKernelGGL(list_assemble, dim3(nblk(nf + ncreate, 256)), dim3(256), 0, s, active_fine,
                       (const int*)nullptr, nf, (const int*)newg, ncreate, pos);
  }
  if ((e = hipGetLastError()) != hipSuccess) return e;
  *noct_fine = nalive + ncreate;
  *head = (int)(((long)*head + ncreate) % t.ngridmax);
  if (ncreate > 0) *used_mem = t.ngridmax - (*numbf - ncreate);  // make_grid_fine's last value
  *numbf = *numbf - ncreate + nkill;
  return hipSuccess;
}

// refine_fine(ilevel) with no host synchronisation, for a captured launch
// sequence: the level counts dn_l (level ilevel) and dn_f (level ilevel+1),
// the free list dfree = {head, numbf, used_mem} and the decision ctrl[3] live
// in device memory; work holds refine_work_bytes(capacity, capacity), every
// level holding at most `capacity` octs.  A refinement that would exhaust
// the free list sets error bit 4, one that would exceed the capacity error
// bit 8; either changes nothing.
hipError_t launch_refine_fine_dev(const int* active, const int* dn_l, int ilevel, const AmrTreeView& t,
                                  const TreeMut& m, int* active_fine, int* dn_f, int* pos, int* ring, int* dfree,
                                  int* ctrl, int* flag1, int* flag2, double* xg, double* uold, double* unew,
                                  int interpol_type, int interpol_var, double smallr, void* work, int* d_error,
                                  int capacity, hipStream_t s) {
  const int cap = t.ngridmax;
  const long nslot = (long)((capacity + 31) / 32) * 256;
  const long nmax = nslot > capacity ? nslot : capacity;
  const long nt = (nmax + kScanTile - 1) / kScanTile + 1;
  char* w = (char*)(((uintptr_t)work + 255) & ~(uintptr_t)255);
  long long* mark = (long long*)w;
  long long* off = mark + nmax;
  long long* sums = off + nmax;
  int* newg = (int*)(sums + nt);
  int* tmp = newg + nslot;
  const DevLen slots{0, dn_l, 2};
  hipError_t e = scan64_dev(MarkSrc{active, 0, dn_l, t, flag1, mark}, off, sums, slots, s);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(refine_decide, dim3(1), dim3(1), 0, s, (const long long*)sums, slots, (const int*)dfree,
                     (const int*)dn_f, capacity, ctrl, d_error);
  RefineRun r;
  r.t = t;
  r.m = m;
  r.active = active;
  r.noct = 0;
  r.mark = mark;
  r.off = off;
  r.sums = sums;
  r.nslot = 0;
  r.ring = ring;
  r.cap = cap;
  r.head = 0;
  r.numbf = 0;
  r.newg = newg;
  r.uold = uold;
  r.unew = unew;
  r.us = t.ncell + 1;
  r.flag1 = flag1;
  r.flag2 = flag2;
  r.xg = xg;
  r.dx = __builtin_ldexp(1.0, -ilevel);
  r.interpol_type = interpol_type;
  r.interpol_var = interpol_var;
  r.smallr = smallr;
  r.error = d_error;
  r.dnoct = dn_l;
  r.dfree = dfree;
  r.ctrl = ctrl;
  hipLaunchKernelGGL(refine_make, dim3(persist_blocks()), dim3(128), 0, s, r);
  hipLaunchKernelGGL(refine_kill, dim3(persist_blocks()), dim3(256), 0, s, r);
  if ((e = scan64_dev(AliveSrc{active_fine, m.father, mark, ctrl}, off, sums, DevLen{0, dn_f, 1}, s)) != hipSuccess)
    return e;
  hipLaunchKernelGGL(list_compact_d, dim3(persist_blocks()), dim3(256), 0, s, (const int*)active_fine, (const int*)dn_f,
                     (const long long*)mark, (const long long*)off, (const long long*)sums, tmp, (const int*)ctrl);
  hipLaunchKernelGGL(list_assemble_d, dim3(persist_blocks()), dim3(256), 0, s, active_fine, (const int*)tmp,
                     (const int*)dn_f, (const int*)newg, pos, (const int*)ctrl);
  hipLaunchKernelGGL(refine_commit, dim3(1), dim3(1), 0, s, dn_f, dfree, (const int*)ctrl, cap);
  return hipGetLastError();
}

}  // namespace ramses
