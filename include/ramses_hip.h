/*
 * ramses_hip -- MI355X (gfx950) implementation of the RAMSES unsplit
 * MUSCL-Hancock hydro step behind the reference's own call boundary.
 *
 * C ABI only: plain pointers, sizes and scalars; no C++ or torch types.
 * Every entry point cites the reference interface it replaces.  Array
 * arguments use the reference's Fortran layouts (column-major, the oct lane
 * `l` fastest with leading dimension nvector):
 *
 *   uin   (nvector, -1:4, -1:4, -1:4, nvar)        hydro/umuscl.f90:33
 *   gravin(nvector, -1:4, -1:4, -1:4, ndim)        hydro/umuscl.f90:34
 *   pin   (nvector, -1:4, -1:4, -1:4)   (unused)   hydro/umuscl.f90:32
 *   flux  (nvector, 1:3, 1:3, 1:3, nvar, ndim)     hydro/umuscl.f90:37
 *   tmp   (nvector, 1:3, 1:3, 1:3, 2, ndim)        hydro/umuscl.f90:38
 *
 * NDIM=3, NVAR=5, NENER=0.  Only the 36 defined faces per oct and only lanes
 * l < ngrid of flux/tmp are written, as in the reference; every other entry
 * keeps the caller's contents.
 *
 * Status codes: 0 success, < 0 invalid argument or configuration,
 * > 0 HIP runtime error (the hipError_t value).  ramses_last_error() returns
 * a message for the last failure on the calling thread.
 */
#ifndef RAMSES_HIP_H
#define RAMSES_HIP_H

#ifdef __cplusplus
extern "C" {
#endif

#define RAMSES_RIEMANN_LLF 0
#define RAMSES_RIEMANN_HLLC 1
#define RAMSES_RIEMANN_HLL 2
#define RAMSES_RIEMANN_ACOUSTIC 3 /* riemann_acoustic, hydro/godunov_utils.f90:500 */
#define RAMSES_RIEMANN_EXACT 4    /* riemann_approx ('exact'), hydro/godunov_utils.f90:268 */

#define RAMSES_FP64 0        /* every operation in FP64, bitwise the reference */
#define RAMSES_FP32_MIXED 1  /* FP32 primitives/slopes/trace/Riemann; FP64 state,
                                E-kinetic recovery, flux accumulation and dt */

/* Solver configuration.  Replaces the module globals the reference reads in
 * every routine of the path: gamma, smallr, smallc, slope_type, slope_theta,
 * riemann, scheme, difmag, courant_factor (hydro/hydro_parameters.f90:74-85)
 * and the compile-time nvector (amr/amr_parameters.f90:49-53). */
typedef struct ramses_hydro_params {
  double gamma;          /* ratio of specific heats, > 1 */
  double smallr;         /* density floor */
  double smallc;         /* sound-speed floor */
  double slope_theta;    /* slope_type 8 parameter */
  double courant_factor; /* CFL number used by the dt reduction */
  double difmag;         /* must be 0: artificial diffusion (cmpdivu/consup) is not built */
  int slope_type;        /* 0,1,2,3,7,8 (the NDIM=3 branches of uslope) */
  int riemann;           /* RAMSES_RIEMANN_* */
  int precision;         /* RAMSES_FP64 or RAMSES_FP32_MIXED */
  int nvector;           /* leading dimension of the Fortran arrays (32 in the reference build) */
  int niter_riemann;     /* Newton iterations of the exact solver (reference default 10) */
} ramses_hydro_params;

/* Validate and install the configuration used by every later call.
 * Replaces the namelist-read module globals (hydro/read_hydro_params.f90:43-53);
 * an unsupported value is rejected here, never mid-kernel (the reference
 * stops inside the kernel: hydro/umuscl.f90:801-803,1474-1477). */
int ramses_hydro_config(const ramses_hydro_params *p);

/* Message for the last non-zero status returned on this thread. */
const char *ramses_last_error(void);

/* Link-level drop-in for the reference's external subroutine
 *   subroutine unsplit(uin,gravin,pin,flux,tmp,dx,dy,dz,dt,ngrid)
 * (hydro/umuscl.f90:22, called from hydro/godunov_fine.f90:681).  All
 * arguments by reference, host memory.  On failure it prints the message and
 * aborts, as the reference's `stop` does.  A gravity array of +0.0 words
 * only (godfine1's without Poisson) is treated as none: the same bits,
 * without its copy. */
void unsplit_(const double *uin, const double *gravin, const double *pin, double *flux, double *tmp,
              const double *dx, const double *dy, const double *dz, const double *dt, const int *ngrid);

/* ISO_C_BINDING form of the same call without `pin` (the 9-argument form of
 * rhd/umuscl.f90:31 and patch/6eq/umuscl.f90:22).  Host memory; gravin may
 * be NULL (gravity zero). */
int ramses_unsplit(const double *uin, const double *gravin, double *flux, double *tmp, double dx,
                   double dy, double dz, double dt, int ngrid);

/* ramses_unsplit without any device allocation; it has
 * no counterpart in the reference.  For callers that cannot allocate device memory (tight
 * per-batch loops such as godfine1's, harnesses that forbid hipMalloc).
 * Same arguments and results as ramses_unsplit, enqueued on `stream`
 * (hipStream_t, NULL = default stream) and synchronised before it returns.
 * The batch is staged in a fixed 2 MiB device array of the library (one per
 * GPU, held until the process ends), so nvector <= 114; a larger configured
 * nvector returns -11.  Calls on one GPU share that array: the library
 * serialises them within a process, and every value the kernel reads is
 * copied in by the call itself. */
int ramses_unsplit_noalloc(const double *uin, const double *gravin, double *flux, double *tmp, double dx,
                           double dy, double dz, double dt, int ngrid, void *stream);

/* Device-resident batched form: `ngroups` consecutive unsplit array sets in
 * device memory (each laid out exactly like one call above, group g at
 * offset g*nvector*216*5 for uin, etc.), every group full except the last
 * which has `ngrid_last` lanes.  Equivalent to the godunov_fine sweep loop
 * (hydro/godunov_fine.f90:24-31) calling unsplit once per batch.  Enqueued
 * on `stream` (a hipStream_t; NULL = default stream), asynchronous.
 * gravin may be NULL (gravity zero, the same bits as a zero array): that
 * selects kernels with no gravity loads, 1.51 ms against 1.97 ms with a
 * zero array for 2^20 octs in FP64. */
int ramses_unsplit_device(const double *uin, const double *gravin, double *flux, double *tmp, double dx,
                          double dy, double dz, double dt, long ngroups, int ngrid_last, void *stream);

/* cmpdt (hydro/godunov_utils.f90:5-120) over `ncell` cells of a (nvector,
 * nvar) host array `uu` and (nvector, ndim) gravity `gg` (NULL = zero).
 * Unlike the reference `uu` is not overwritten.  Returns the status and
 * writes the time step to *dt. */
int ramses_cmpdt(const double *uu, const double *gg, double dx, int ncell, double *dt);

/* ---------------------------------------------------------------------
 * Resident uniform level (a fully refined periodic level, levelmin ==
 * levelmax): the godunov_fine -> godfine1 sweep (hydro/godunov_fine.f90:5-35,
 * 486-911), set_unew/set_uold (:40-232) and courant_fine
 * (hydro/courant_fine.f90:1-159) over conserved state kept in device memory
 * as u[var][z][y][x], var-major SoA, n^3 cells per variable.
 * --------------------------------------------------------------------- */

/* courant_fine over the level: writes min(dt_start, min_cell cmpdt) to
 * *d_dt (device memory, one double).  u: device state. */
int ramses_uniform_courant(const double *u, int n, double dx, double dt_start, double *d_dt, void *stream);

/* godfine1 over every oct of the level with the time step read from device
 * memory *d_dt: unew = u + sum_dim (flux_left - flux_right).  When
 * d_dt_next is not NULL the courant reduction of the new state is fused into
 * the same kernel and min-combined into *d_dt_next, which the caller must
 * have set to the starting value (ramses_uniform_dt_init). */
int ramses_uniform_godunov(const double *u, double *unew, int n, double dx, const double *d_dt,
                           double *d_dt_next, void *stream);

/* courant_fine's conservation sums (hydro/courant_fine.f90:95-119) beside
 * the dt: the *_sums forms of the two calls above also write, for the state
 * they reduce (the level's state, or the new state of the step), d_sums[0] =
 * total mass sum rho*vol, d_sums[1] = total energy sum E*vol (the
 * reference's ekin_tot), d_sums[2] = internal energy sum (E - |m|^2/(2
 * max(rho,smallr)))*vol, vol = dx^3, into 3 device doubles (overwritten).
 * The sums are reduced per workgroup and then over workgroups in a fixed
 * order, so a given launch gives the same bits on every run (the reference's
 * serial order rounds differently).  `work` is device memory of at least
 * ramses_uniform_work_bytes(n) bytes; d_sums NULL computes no sums and needs
 * no work buffer. */
unsigned long ramses_uniform_work_bytes(int n);
int ramses_uniform_courant_sums(const double *u, int n, double dx, double dt_start, double *d_dt, double *d_sums,
                                void *work, unsigned long work_bytes, void *stream);
/* godfine1 over the level with the fused courant_fine of the new state and
 * its sums (hydro/godunov_fine.f90:5-35, hydro/courant_fine.f90:95-119). */
int ramses_uniform_godunov_sums(const double *u, double *unew, int n, double dx, const double *d_dt,
                                double *d_dt_next, double *d_sums, void *work, unsigned long work_bytes,
                                void *stream);

/* Writes `value` into one device double with a kernel (graph-capture safe):
 * the per-step reset of the dt accumulator, dtnew(ilevel)=boxlen/smallc in
 * newdt_fine (pm/newdt_fine.f90:51). */
int ramses_uniform_dt_init(double *d_dt, double value, void *stream);

/* ---------------------------------------------------------------------
 * Continuous batching of independent levels (serving).  Each slot is one
 * periodic uniform level advanced by the per-level step of amr_step
 * (amr/amr_step.f90:327-424: newdt_fine/courant_fine, godunov_fine,
 * set_uold) with its own courant dt.  The last step of a slot is shortened
 * to end exactly at the slot's end time; the reference has no such cap
 * (pm/newdt_fine.f90:51-173 ignores the output times, amr/amr_step.f90:151
 * writes an output after the step that crosses one), so this is the
 * server's own contract.  Slots of different sizes share one launch.
 * --------------------------------------------------------------------- */
#define RAMSES_MAX_SLOTS 32
/* Slot state: 8 device doubles, read by the host to follow progress. */
#define RAMSES_SLOT_TIME 0    /* t */
#define RAMSES_SLOT_TEND 1    /* end time */
#define RAMSES_SLOT_DT 2      /* dt of the last step */
#define RAMSES_SLOT_DTNEXT 3  /* courant dt of the current state */
#define RAMSES_SLOT_DONE 4    /* 1 once t >= end time */
#define RAMSES_SLOT_STEPS 5   /* steps taken */
#define RAMSES_SLOT_PARITY 6  /* buf[parity] holds the current state */
#define RAMSES_SLOT_STEPPED 7 /* internal */

typedef struct ramses_level_slot {
  double *buf[2]; /* two device buffers u[var][z][y][x], n^3 per variable */
  double *state;  /* 8 device doubles (RAMSES_SLOT_*) */
  int n;          /* cells per side, a multiple of 16 */
  double dx;      /* cell size */
} ramses_level_slot;

/* Start a slot at t=0 from buf[0]: courant_fine of the initial state
 * (hydro/courant_fine.f90:1-159) into its state. */
int ramses_level_slot_init(const ramses_level_slot *slot, double t_end, void *stream);

/* One step of every slot that is not done (the godunov_fine sweep,
 * hydro/godunov_fine.f90:5-35, of each level, fused courant reduction of
 * the new state).  Results depend only on each slot's own data: a slot
 * advanced alone or in any mix gives the same bits. */
int ramses_level_batch_step(const ramses_level_slot *slots, int nslots, void *stream);

/* ---------------------------------------------------------------------
 * Oct-tree (AMR) levels.  The tree and the state stay in device memory in
 * the reference's own memory model (amr/amr_commons.f90:54-75,
 * hydro/hydro_commons.f90:4): 1-based cell index
 * icell = ncoarse + (ind-1)*ngridmax + igrid, arrays addressed so that
 * element i of the reference is p[i] (element 0 unused), state
 * u[var*(ncell+1) + icell].  nx*ny*nz coarse cells, periodic.
 * --------------------------------------------------------------------- */
typedef struct ramses_amr_tree {
  int *son;    /* son[icell], icell = 1..ncoarse+8*ngridmax */
  int *father; /* father[igrid], igrid = 1..ngridmax */
  int *nbor;   /* nbor[(j-1)*ngridmax + igrid], j = 1..6 (-x,+x,-y,+y,-z,+z) */
  int ncoarse, ngridmax, nx, ny, nz;
} ramses_amr_tree; /* only ramses_amr_refine_fine writes the arrays */

typedef struct ramses_amr_level {
  int ilevel;
  double dx;                /* 0.5^ilevel * boxlen/nx */
  const int *active;        /* active(ilevel)%igrid, device, noct entries */
  int noct;
  const int *pos;           /* pos[igrid] = index of igrid in `active` (read for this level's grids) */
  const int *active_coarse; /* grids of level ilevel-1 (reflux targets); NULL at the coarsest level */
  int noct_coarse;
} ramses_amr_level;

/* interpol_type (0-4) and interpol_var (0-2) of the buffer-oct
 * prolongation, hydro/hydro_parameters.f90:88-89 (defaults 1 and 0).
 * Type 4 requires interpol_var 2, as the reference does
 * (hydro/interpol_hydro.f90:367-377); other pairs with type 4 return -2.
 * interpol_var also selects what the restriction averages (upl,
 * hydro/interpol_hydro.f90:226-261): total energy for 0, internal energy
 * for 1 and 2. */
int ramses_amr_config(int interpol_type, int interpol_var);

/* Device workspace the godunov sweep of a level with `noct` octs needs
 * (its nvector batches of stencils, flags and fluxes). */
unsigned long ramses_amr_work_bytes(int noct);

/* godunov_fine(ilevel) (hydro/godunov_fine.f90:5-35): godfine1 over every
 * oct of the level -- 6x6x6 gather through son/nbor with interpolated
 * buffer octs, unsplit, flux reset at refined faces, unew update at the
 * level and conservative reflux into coarser leaf neighbours -- with dt
 * read from device memory *d_dt.  *d_error becomes nonzero if an oct's
 * 27 father neighbours do not all exist (the refinement rule is broken). */
int ramses_amr_godunov(const ramses_amr_tree *t, const ramses_amr_level *lv, const double *uold, double *unew,
                       const double *d_dt, void *work, unsigned long work_bytes, int *d_error, void *stream);

/* set_unew / set_uold of a level (hydro/godunov_fine.f90:40,135): copy the
 * state of the level's active cells from src to dst. */
int ramses_amr_copy(const ramses_amr_tree *t, const ramses_amr_level *lv, const double *src, double *dst,
                    void *stream);

/* upload_fine(ilevel) (hydro/interpol_hydro.f90:5-263): refined cells of
 * the level take the average of their children (energy per the configured
 * interpol_var). */
int ramses_amr_upload(const ramses_amr_tree *t, const ramses_amr_level *lv, double *u, void *stream);

/* courant_fine(ilevel) (hydro/courant_fine.f90:1-159) over the level's leaf
 * cells: *d_dt = min(dt_start, courant_factor*dx/smallc, min cmpdt). */
int ramses_amr_courant(const ramses_amr_tree *t, const ramses_amr_level *lv, const double *u, double dt_start,
                       double *d_dt, void *stream);

/* Refinement criteria: err_grad_d/u/p and floor_d/u/p
 * (hydro/hydro_parameters.f90:47-52; a criterion is off when its err_grad
 * is negative, the default -1) and nexpand (amr/amr_parameters.f90:85,
 * default 1, the same for every level). */
typedef struct ramses_refine_params {
  double err_grad_d, err_grad_u, err_grad_p;
  double floor_d, floor_u, floor_p;
  int nexpand;
} ramses_refine_params;

/* Set the refinement criteria (hydro/hydro_parameters.f90:47-52,
 * amr/amr_parameters.f90:85); NULL restores the defaults. */
int ramses_refine_config(const ramses_refine_params *p);

/* flag_fine(ilevel, icount) (amr/flag_utils.f90:57-106) for one level: the
 * refinement map flag1 of the level's cells -- refined cells with refined
 * or flagged children (test_flag), one smooth_fine buffer, the gradient
 * criteria (hydro_flag, hydro/hydro_flag.f90:1-180; hydro_refine,
 * hydro/godunov_utils.f90:125-263), nexpand more buffers, and for
 * ilevel > levelmin with icount < 2 the refinement rule (ensure_ref_rules).
 * flag1 and flag2 have ncoarse+8*ngridmax+1 entries, element 0 zero;
 * flag2 is workspace.  No-op at ilevel == nlevelmax.  ensure_ref_rules
 * runs for icount < 2, i.e. icount < nsubcycle(ilevel-1) with nsubcycle 2;
 * a caller whose level ilevel-1 does not subcycle passes icount = 2. */
int ramses_amr_flag_fine(const ramses_amr_tree *t, const ramses_amr_level *lv, int icount, int levelmin,
                         int nlevelmax, const double *uold, int *flag1, int *flag2, void *stream);

/* The free-grid list (headf/tailf/numbf/used_mem of amr/amr_commons.f90):
 * a device ring of ngridmax grid indices holding the numbf free grids from
 * index `head` on, in the reference's free-list order; used_mem is
 * ngridmax-numbf as of the last oct creation, as make_grid_fine leaves it
 * (amr/refine_utils.f90:655-661). */
typedef struct ramses_amr_free {
  int *ring;
  int head;
  int numbf;
  int used_mem;
} ramses_amr_free;

/* Device workspace of ramses_amr_refine_fine for a level of noct octs whose
 * next level has noct_fine octs. */
unsigned long ramses_amr_refine_work_bytes(int noct, int noct_fine);

/* refine_fine(ilevel) (amr/refine_utils.f90:332-585): every leaf cell of the
 * level with flag1 == 1 gets a new oct (make_grid_fine, :590-948: the next
 * free grid, its son/father/nbor links, its cells prolongated with
 * interpol_hydro from the father cell and its six neighbours), then every
 * refined cell with flag1 == 0 loses its oct (kill_grid, :953-1158: links,
 * flags and state cleared, grid returned to the tail of the free list).
 * active_fine[0..*noct_fine) is the grid list of level ilevel+1 (capacity
 * ngridmax); it is rebuilt as build_comm rebuilds it from the level's
 * linked list (amr/virtual_boundaries.f90:1286-1420) -- survivors in order,
 * then the new octs in creation order -- and pos[] follows.  xg (optional,
 * 3 doubles per grid: xg[3*igrid+d], the grid centres of amr/amr_commons.f90
 * in units of the coarse cell) is set for new octs (:663-673) and cleared
 * for killed ones.  Writes the tree arrays of t, the state, flag1/flag2 and *fl.  Synchronises `stream` once
 * (to learn how many octs are made and killed).  Returns -9 and changes
 * nothing if the free list cannot hold the new octs ("No more free memory",
 * :415-423); *d_error gets bit 2 if a new oct lacks a neighbour cell. */
int ramses_amr_refine_fine(const ramses_amr_tree *t, const ramses_amr_level *lv, int *active_fine, int *noct_fine,
                           int *pos, ramses_amr_free *fl, int *flag1, int *flag2, double *xg, double *uold,
                           double *unew, void *work, unsigned long work_bytes, int *d_error, void *stream);

/* ---------------------------------------------------------------------------
 * One coarse step of the main loop on a device-resident mesh.
 * ------------------------------------------------------------------------- */
#define RAMSES_MAX_LEVELS 32

/* The whole mesh of amr/amr_commons.f90 + hydro/hydro_commons.f90 for one
 * domain.  Device arrays as in ramses_amr_tree / ramses_amr_level (element 0
 * unused); active[l] has room for ngridmax grids; dt, dtold hold dtnew(l)
 * and dtold(l) at index l (RAMSES_MAX_LEVELS+2 entries); time is t.
 * noct, free, the counters and work are updated by the library; work is a
 * device workspace the library grows as needed (start with NULL / 0 and
 * release it with ramses_amr_mesh_release). */
typedef struct ramses_amr_mesh {
  ramses_amr_tree tree;
  int *active[RAMSES_MAX_LEVELS + 1];
  int noct[RAMSES_MAX_LEVELS + 1];
  int *pos;
  double *xg;
  int *flag1, *flag2;
  ramses_amr_free free;
  double *uold, *unew;
  double *dt, *dtold, *time;
  int *error;
  int levelmin, nlevelmax;
  int nsubcycle[RAMSES_MAX_LEVELS + 1]; /* 1 or 2 for level l at index l */
  int adaptive;                         /* refine and flag during the step */
  double boxlen;
  long long nstep, nstep_coarse, cell_updates;
  void *work;
  unsigned long work_bytes;
  /* Captured mode (graph != 0): the whole coarse step is captured once as a
   * hipGraph and replayed.  Every level kernel then reads its oct count from
   * device memory and loops over a fixed grid, refine_fine decides on the
   * device with no host synchronisation, and one copy in and one copy out
   * per step carry noct, the free list and the counters.  The graph is
   * captured again when a pointer, a size, nsubcycle or a configuration
   * call changes.  Its workspace holds graph_capacity octs per level (0:
   * ngridmax; about 13.4 KB per oct); a refinement that would take a level
   * beyond it is refused and the step returns -10.  Results are identical to
   * the direct mode.  graph == 2 captures (if needed) without running a step,
   * so that a timed loop replays only. */
  int graph;
  int graph_capacity;
  void *graph_state; /* library-owned: NULL before the first step */
  /* Optional (NULL: not computed): 3 device doubles that each coarse step
   * resets and every courant_fine of the step adds its level's leaf sums to,
   * as mass_tot, ekin_tot and eint_tot are (amr/adaptive_loop.f90:89-91,
   * hydro/courant_fine.f90:95-119,152-154): total mass, sum of E*vol and
   * internal energy, vol = dx^3 of each level.  Deterministic: workgroup
   * partial sums added in block order. */
  double *sums;
  /* Optional (NULL: not kept): update_time's conservation baselines
   * (amr/update_time.f90:229-246), 2 device doubles {einit, mass_tot_0}
   * that the snapshot files carry.  At the first update_time of a coarse
   * step with nstep_coarse != nstep_coarse_old (every step but a run's
   * first, when nstep_coarse_old starts equal to nstep_coarse as
   * adaptive_loop sets it), each one still zero is set from mass_tot and
   * epot_tot + ekin_tot as they stand there: the courant_fine sums of the
   * levels the step's first descent visits, each level summed in the
   * reference's serial order, so the values round as the reference's do.
   * The step then sets nstep_coarse_old = nstep_coarse. */
  double *cons0;
  long long nstep_coarse_old;
} ramses_amr_mesh;

/* amr_step(levelmin, 1) (amr/amr_step.f90) followed by the end of a coarse
 * step of adaptive_loop (amr/adaptive_loop.f90:143-191: upload_fine and
 * flag_fine below levelmin, flag_coarse): per level, refine_fine of the
 * finer levels (at levelmin and on second substeps), newdt_fine /
 * courant_fine, set_unew, the finer levels' substeps, godunov_fine,
 * set_uold, upload_fine and flag_fine, with every dt kept on the device.
 * nstep counts finest-level steps (amr/update_time.f90:332). */
int ramses_amr_coarse_step(ramses_amr_mesh *m, void *stream);

/* Frees the workspace (and the captured graph) a mesh's coarse steps
 * allocated. */
int ramses_amr_mesh_release(ramses_amr_mesh *m);

/* ---------------------------------------------------------------------
 * Launch-shape tuning.  The reference's only performance knob is the
 * compile-time vector length nvector (amr/amr_parameters.f90:49-53, set by
 * bin/Makefile:11); on the GPU the equivalent choices are the workgroup and
 * tile shapes below, picked per device and workload by the autotuner
 * (ramses_amd/autotune.py).  value 0 restores the built-in default.
 * --------------------------------------------------------------------- */
#define RAMSES_TUNE_UNSPLIT_BLOCK 1 /* threads per unsplit workgroup: multiple of 64, <= 768 */
#define RAMSES_TUNE_LEVEL_TILE 2    /* level tile: 1 = 32x8, 2 = 16x8, 3 = 8x8, 4 = 32x16, 5 = 16x12 columns */
#define RAMSES_TUNE_LEVEL_TZ 3      /* z planes per level workgroup: a divisor of n, >= 4 */
#define RAMSES_TUNE_PERSIST_BLOCKS 4 /* workgroups of a captured AMR step's fixed grids (0: 1024) */
#define RAMSES_TUNE_AMR_FACE_OCTS 5  /* levels of at most this many octs run the AMR unsplit one face per
                                        thread instead of one row per thread (0: 8192, -1: never) */

/* Set one tuning knob (amr/amr_parameters.f90:49-53 is the reference's
 * counterpart); returns 0, or < 0 for an unknown knob or invalid value.
 * ramses_tune_get returns the current value (0 = default). */
int ramses_tune(int knob, int value);
int ramses_tune_get(int knob);

#ifdef __cplusplus
}
#endif
#endif /* RAMSES_HIP_H */
