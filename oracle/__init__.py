"""TEST INFRASTRUCTURE ONLY: the CPU oracle for the unsplit hydro step.

Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may import this package, and only as the checker.  The product
package ``ramses_amd`` never imports it.

Two CPU implementations are exposed through ctypes:

* ``Oracle``  -- ``libhydro_oracle.so``, the C restatement in ``hydro_ref.c``
  (built by ``make -C oracle``), and
* ``Reference`` -- ``_ref/libramses_ref.so``, the reference's own Fortran
  hot-path routines (hydro/umuscl.f90, hydro/godunov_utils.f90) compiled from
  the reference sources by ``make -C oracle ref``.  It is present only where
  it was built.

All arrays are numpy float64 in the reference's Fortran layouts, passed as
C-ordered numpy arrays whose axes are the Fortran axes reversed:
``uin[nvar][k][j][i][l]`` (6x6x6 cells, ``l`` < nvector fastest),
``flux[ndim][nvar][k][j][i][l]`` (3x3x3 face slots), ``tmp[ndim][2][k][j][i][l]``.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_LIB = os.path.join(_HERE, "libhydro_oracle.so")
REF_LIB = os.path.join(_HERE, "_ref", "libramses_ref.so")

NVAR, NDIM = 5, 3
RIEMANN = {"llf": 0, "hllc": 1, "hll": 2, "acoustic": 3, "exact": 4}

_dp = ctypes.POINTER(ctypes.c_double)


def _ptr(a):
    if a is None:
        return None
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(_dp)


class _Params(ctypes.Structure):
    _fields_ = [("gamma", ctypes.c_double), ("smallr", ctypes.c_double),
                ("smallc", ctypes.c_double), ("slope_theta", ctypes.c_double),
                ("courant_factor", ctypes.c_double), ("difmag", ctypes.c_double),
                ("slope_type", ctypes.c_int), ("riemann", ctypes.c_int), ("niter_riemann", ctypes.c_int)]


def stencil_shape(nv):
    return (NVAR, 6, 6, 6, nv)


def grav_shape(nv):
    return (NDIM, 6, 6, 6, nv)


def flux_shape(nv):
    return (NDIM, NVAR, 3, 3, 3, nv)


def tmp_shape(nv):
    return (NDIM, 2, 3, 3, 3, nv)


class Oracle:
    """ctypes front end of the C restatement (hydro_ref.c)."""

    def __init__(self, path=ORACLE_LIB):
        if not os.path.exists(path):
            raise FileNotFoundError(f"{path} is not built (make -C oracle)")
        lib = ctypes.CDLL(path)
        lib.hr_set_params.argtypes = [ctypes.POINTER(_Params)]
        lib.hr_set_params.restype = ctypes.c_int
        lib.hr_unsplit.argtypes = [_dp, _dp, _dp, _dp] + [ctypes.c_double] * 4 + [ctypes.c_int] * 2
        lib.hr_unsplit.restype = ctypes.c_int
        lib.hr_unsplit_batched.argtypes = [_dp, _dp, _dp, _dp, ctypes.c_double, ctypes.c_double,
                                           ctypes.c_long, ctypes.c_int, ctypes.c_int, ctypes.c_int]
        lib.hr_unsplit_batched.restype = ctypes.c_int
        lib.hr_cmpdt.argtypes = [_dp, _dp, ctypes.c_double, ctypes.c_int, ctypes.c_int]
        lib.hr_cmpdt.restype = ctypes.c_double
        lib.hr_uniform_godunov.argtypes = [_dp, _dp, ctypes.c_int, ctypes.c_double, ctypes.c_double,
                                           ctypes.c_int, ctypes.c_int]
        lib.hr_uniform_godunov.restype = ctypes.c_int
        lib.hr_uniform_courant.argtypes = [_dp, ctypes.c_int, ctypes.c_double, ctypes.c_double,
                                           ctypes.c_int, _dp, _dp, _dp]
        lib.hr_uniform_courant.restype = ctypes.c_double
        lib.hr_uniform_step.argtypes = [_dp, _dp, ctypes.c_int, ctypes.c_double, ctypes.c_int,
                                        ctypes.c_int]
        lib.hr_uniform_step.restype = ctypes.c_double
        self.lib = lib

    def set_params(self, gamma=1.4, smallr=1e-10, smallc=1e-10, slope_type=1, slope_theta=1.5,
                   riemann="llf", courant_factor=0.5, difmag=0.0, niter_riemann=10):
        p = _Params(gamma, smallr, smallc, slope_theta, courant_factor, difmag, slope_type,
                    RIEMANN[riemann], niter_riemann)
        rc = self.lib.hr_set_params(ctypes.byref(p))
        if rc != 0:
            raise ValueError(f"oracle rejected the configuration (code {rc})")

    def unsplit(self, uin, gravin, flux, tmp, dx, dt, ngrid, dy=None, dz=None):
        nv = uin.shape[-1]
        rc = self.lib.hr_unsplit(_ptr(uin), _ptr(gravin), _ptr(flux), _ptr(tmp), dx,
                                 dx if dy is None else dy, dx if dz is None else dz, dt, ngrid, nv)
        if rc:
            raise RuntimeError(f"hr_unsplit failed ({rc})")

    def unsplit_batched(self, uin, gravin, flux, tmp, dx, dt, ngrid_last=None, nthreads=0):
        """uin[g][nvar][6][6][6][nv]; all groups full except the last."""
        ng, nv = uin.shape[0], uin.shape[-1]
        rc = self.lib.hr_unsplit_batched(_ptr(uin), _ptr(gravin), _ptr(flux), _ptr(tmp), dx, dt,
                                         ng, nv if ngrid_last is None else ngrid_last, nv, nthreads)
        if rc:
            raise RuntimeError(f"hr_unsplit_batched failed ({rc})")

    def cmpdt(self, uu, gg, dx, ncell):
        """uu[nvar][nv] is overwritten like the reference does; returns dt."""
        return self.lib.hr_cmpdt(_ptr(uu), _ptr(gg), dx, ncell, uu.shape[-1])

    def uniform_godunov(self, u, dx, dt, nv=32, nthreads=0):
        n = u.shape[-1]
        unew = np.empty_like(u)
        rc = self.lib.hr_uniform_godunov(_ptr(u), _ptr(unew), n, dx, dt, nv, nthreads)
        if rc:
            raise RuntimeError(f"hr_uniform_godunov failed ({rc})")
        return unew

    def uniform_courant(self, u, dx, dt_start, nv=32):
        n = u.shape[-1]
        m, e, ei = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
        dt = self.lib.hr_uniform_courant(_ptr(u), n, dx, dt_start, nv, ctypes.byref(m),
                                         ctypes.byref(e), ctypes.byref(ei))
        return dt, m.value, e.value, ei.value

    def uniform_step(self, u, dx, nv=32, nthreads=0):
        n = u.shape[-1]
        unew = np.empty_like(u)
        dt = self.lib.hr_uniform_step(_ptr(u), _ptr(unew), n, dx, nv, nthreads)
        if dt < 0:
            raise RuntimeError("hr_uniform_step failed")
        return unew, dt


class Reference:
    """ctypes front end of the reference's own compiled routines (oracle/_ref)."""

    def __init__(self, path=REF_LIB):
        if not os.path.exists(path):
            raise FileNotFoundError(f"{path} is not built (make -C oracle ref)")
        # lazy binding: the AMR routines reference clean_stop (amr/end.f90),
        # called only on corrupt trees and not part of this build
        lib = ctypes.CDLL(path, mode=os.RTLD_LAZY)
        lib.ref_set_hydro_params.argtypes = [ctypes.c_double] * 3 + [ctypes.c_int, ctypes.c_double,
                                                                       ctypes.c_int, ctypes.c_double,
                                                                       ctypes.c_double]
        lib.ref_unsplit.argtypes = [_dp, _dp, _dp, _dp, _dp] + [ctypes.c_double] * 4 + [ctypes.c_int]
        lib.ref_cmpdt.argtypes = [_dp, _dp, ctypes.c_double, _dp, ctypes.c_int]
        nv, nvar, nd = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        lib.ref_shape(ctypes.byref(nv), ctypes.byref(nvar), ctypes.byref(nd))
        self.nvector, self.nvar, self.ndim = nv.value, nvar.value, nd.value
        self.lib = lib
        self._pin = np.zeros((6, 6, 6, self.nvector))
        self._grav0 = np.zeros(grav_shape(self.nvector))

    def set_params(self, gamma=1.4, smallr=1e-10, smallc=1e-10, slope_type=1, slope_theta=1.5,
                   riemann="llf", courant_factor=0.5, difmag=0.0, niter_riemann=10):
        self.lib.ref_set_hydro_params(gamma, smallr, smallc, slope_type, slope_theta,
                                      RIEMANN[riemann], courant_factor, difmag)
        self.lib.ref_set_niter_riemann.argtypes = [ctypes.c_int]
        self.lib.ref_set_niter_riemann(niter_riemann)

    def unsplit(self, uin, gravin, flux, tmp, dx, dt, ngrid, dy=None, dz=None):
        assert uin.shape == stencil_shape(self.nvector)
        g = self._grav0 if gravin is None else gravin
        self.lib.ref_unsplit(_ptr(uin), _ptr(g), _ptr(self._pin), _ptr(flux), _ptr(tmp), dx,
                             dx if dy is None else dy, dx if dz is None else dz, dt, ngrid)

    def cmpdt(self, uu, gg, dx, ncell):
        dt = ctypes.c_double()
        g = np.zeros((NDIM, self.nvector)) if gg is None else gg
        self.lib.ref_cmpdt(_ptr(uu), _ptr(g), dx, ctypes.byref(dt), ncell)
        return dt.value


class ReferenceAMR:
    """The reference's own level routines (godunov_fine, set_unew, set_uold,
    courant_fine, upload_fine, flag_fine, refine_fine) on an oct tree
    installed from numpy arrays (ref_amr_glue.f90)."""

    def __init__(self, ref):
        self.ref = ref
        lib = ref.lib
        ip = ctypes.POINTER(ctypes.c_int)
        lib.ref_amr_init.argtypes = [ctypes.c_int, ctypes.c_int]
        lib.ref_amr_set_tree.argtypes = [ip, ip, ip]
        lib.ref_amr_set_level.argtypes = [ctypes.c_int, ctypes.c_int, ip]
        lib.ref_amr_set_interpol.argtypes = [ctypes.c_int, ctypes.c_int]
        lib.ref_amr_put_u.argtypes = [ctypes.c_int, _dp]
        lib.ref_amr_get_u.argtypes = [ctypes.c_int, _dp]
        lib.ref_amr_set_unew.argtypes = [ctypes.c_int]
        lib.ref_amr_set_uold.argtypes = [ctypes.c_int]
        lib.ref_amr_upload.argtypes = [ctypes.c_int]
        lib.ref_amr_godunov.argtypes = [ctypes.c_int, ctypes.c_double]
        lib.ref_amr_courant.argtypes = [ctypes.c_int, ctypes.c_double, _dp, _dp, _dp, _dp]
        lib.ref_amr_set_free.argtypes = [ctypes.c_int, ip]
        lib.ref_amr_set_refine.argtypes = [ctypes.c_double] * 6 + [ctypes.c_int, ctypes.c_int]
        lib.ref_amr_flag_fine.argtypes = [ctypes.c_int, ctypes.c_int]
        lib.ref_amr_refine_fine.argtypes = [ctypes.c_int]
        lib.ref_amr_get_active.argtypes = [ctypes.c_int, ip]
        lib.ref_amr_get_active.restype = ctypes.c_int
        lib.ref_amr_get_free.argtypes = [ip]
        lib.ref_amr_get_free.restype = ctypes.c_int
        lib.ref_amr_get_tree.argtypes = [ip, ip, ip, ip]
        lib.ref_amr_set_flag1.argtypes = [ip]
        lib.ref_amr_set_xg.argtypes = [_dp]
        lib.ref_amr_get_xg.argtypes = [_dp]
        lib.ref_amr_set_sedov.argtypes = [ctypes.c_double] * 6
        lib.ref_amr_set_init.argtypes = [ctypes.c_int]
        lib.ref_amr_init_flow_fine.argtypes = [ctypes.c_int]
        lib.ref_amr_set_dt.argtypes = [_dp, _dp]
        lib.ref_amr_get_used_mem.restype = ctypes.c_int
        lib.ref_set_regions.argtypes = [ctypes.c_int, ip] + [_dp] * 12
        lib.ref_condinit.argtypes = [ctypes.c_int, _dp, _dp, ctypes.c_double]
        lib.ref_amr_set_nsubcycle.argtypes = [ctypes.c_int, ip]
        lib.ref_amr_set_boxlen.argtypes = [ctypes.c_double]
        lib.ref_amr_count_create.argtypes = [ctypes.c_int]
        lib.ref_amr_count_create.restype = ctypes.c_int
        lib.ref_amr_numbf.restype = ctypes.c_int
        lib.ref_amr_set_outputs.argtypes = [ctypes.c_int, _dp]
        lib.ref_amr_dump.argtypes = [ctypes.c_char_p, ctypes.c_double, ctypes.c_int, ctypes.c_int]
        lib.ref_amr_coarse_step.argtypes = [_dp]
        lib.ref_amr_set_clock.argtypes = [ctypes.c_double, ctypes.c_int, ctypes.c_int]
        lib.ref_amr_get_run.argtypes = [ctypes.c_int, _dp, ip, _dp, _dp]
        lib.ref_amr_get_baselines.argtypes = [_dp]
        self.lib = lib

    def install(self, tree, nlevelmax=None):
        ip = ctypes.POINTER(ctypes.c_int)
        self.tree = tree
        self.lib.ref_amr_init(nlevelmax or tree.nlevel, tree.ngridmax)
        self._keep = [np.ascontiguousarray(a, dtype=np.int32) for a in (tree.son, tree.father, tree.nbor)]
        self.lib.ref_amr_set_tree(*[a.ctypes.data_as(ip) for a in self._keep])
        for l, ids in tree.active.items():
            a = np.ascontiguousarray(ids, dtype=np.int32)
            self.lib.ref_amr_set_level(l, len(a), a.ctypes.data_as(ip))
        # free list: the unused grids in increasing order
        used = np.zeros(tree.ngridmax + 1, dtype=bool)
        for ids in tree.active.values():
            used[np.asarray(ids)] = True
        free = np.ascontiguousarray(np.nonzero(~used[1:])[0] + 1, dtype=np.int32)
        self.lib.ref_amr_set_free(len(free), free.ctypes.data_as(ip))
        self.nlevelmax = nlevelmax or tree.nlevel

    def set_interpol(self, interpol_type=1, interpol_var=0):
        self.lib.ref_amr_set_interpol(interpol_type, interpol_var)

    def put(self, which, u):
        self.lib.ref_amr_put_u(which, _ptr(np.ascontiguousarray(u, dtype=np.float64)))

    def get(self, which):
        out = np.zeros((NVAR, self.tree.ncell))
        self.lib.ref_amr_get_u(which, _ptr(out))
        return out

    def set_unew(self, l):
        self.lib.ref_amr_set_unew(l)

    def set_uold(self, l):
        self.lib.ref_amr_set_uold(l)

    def upload(self, l):
        self.lib.ref_amr_upload(l)

    def godunov(self, l, dt):
        self.lib.ref_amr_godunov(l, dt)

    def courant(self, l, dt_start):
        v = [ctypes.c_double() for _ in range(4)]
        self.lib.ref_amr_courant(l, dt_start, *[ctypes.byref(x) for x in v])
        return tuple(x.value for x in v)

    def set_refine(self, levelmin, err_grad_d=-1.0, err_grad_u=-1.0, err_grad_p=-1.0, floor_d=1e-10,
                   floor_u=1e-10, floor_p=1e-10, nexpand=1):
        self.lib.ref_amr_set_refine(err_grad_d, err_grad_u, err_grad_p, floor_d, floor_u, floor_p, nexpand,
                                    levelmin)

    def flag_fine(self, l, icount):
        self.lib.ref_amr_flag_fine(l, icount)

    def flag_coarse(self):
        self.lib.ref_amr_flag_coarse()

    def refine_fine(self, l):
        """refine_fine(l) + build_comm(l+1).  The reference stops the whole
        process (exit status 0) when the free list is too short; that case
        is refused here with an exception instead."""
        if l < self.nlevelmax:
            need, free = self.lib.ref_amr_count_create(l), self.lib.ref_amr_numbf()
            if need > 0 and need >= free:
                raise MemoryError(f"refine_fine({l}) would need {need} grids with {free} free "
                                  "(the reference would stop: No more free memory)")
        self.lib.ref_amr_refine_fine(l)

    def set_flag1(self, flag1):
        a = np.ascontiguousarray(flag1, dtype=np.int32)
        self.lib.ref_amr_set_flag1(a.ctypes.data_as(ctypes.POINTER(ctypes.c_int)))

    def grids(self, l):
        out = np.zeros(self.tree.ngridmax, dtype=np.int32)
        n = self.lib.ref_amr_get_active(l, out.ctypes.data_as(ctypes.POINTER(ctypes.c_int)))
        return out[:n].copy()

    def free_list(self):
        out = np.zeros(self.tree.ngridmax, dtype=np.int32)
        n = self.lib.ref_amr_get_free(out.ctypes.data_as(ctypes.POINTER(ctypes.c_int)))
        return out[:n].copy()

    def set_xg(self, xg):
        """Grid centres, xg[igrid-1, d]."""
        self.lib.ref_amr_set_xg(_ptr(np.ascontiguousarray(xg, dtype=np.float64)))

    def get_xg(self):
        out = np.zeros((self.tree.ngridmax, 3))
        self.lib.ref_amr_get_xg(_ptr(out))
        return out

    def set_sedov(self, d_amb=1.0, p_amb=1e-5, p_point=0.4, center=(0.5, 0.5, 0.5)):
        self.lib.ref_amr_set_sedov(d_amb, p_amb, p_point, *center)

    def set_init(self, flag):
        self.lib.ref_amr_set_init(1 if flag else 0)

    def init_flow(self):
        self.lib.ref_amr_init_flow()

    def init_flow_fine(self, l):
        self.lib.ref_amr_init_flow_fine(l)

    def set_dt(self, dtold, dtnew):
        n = self.nlevelmax
        a = np.zeros(n)
        b = np.zeros(n)
        a[:min(n, len(dtold))] = dtold[:n]
        b[:min(n, len(dtnew))] = dtnew[:n]
        self.lib.ref_amr_set_dt(_ptr(a), _ptr(b))

    def defrag(self):
        self.lib.ref_amr_defrag()

    def used_mem(self):
        return self.lib.ref_amr_get_used_mem()

    def set_boxlen(self, b):
        self.lib.ref_amr_set_boxlen(b)

    def set_outputs(self, tout):
        a = np.ascontiguousarray(tout, dtype=np.float64)
        self.lib.ref_amr_set_outputs(len(a), _ptr(a))

    def set_nsubcycle(self, ns):
        a = np.ascontiguousarray(ns, dtype=np.int32)
        self.lib.ref_amr_set_nsubcycle(len(a), a.ctypes.data_as(ctypes.POINTER(ctypes.c_int)))

    def set_regions(self, regions):
        """namelist.Region-like objects (type, center, length, exp, d, u, p)."""
        n = len(regions)
        ip = ctypes.POINTER(ctypes.c_int)
        rt = np.array([1 if r.type == "square" else 2 for r in regions], dtype=np.int32)
        cols = [np.array([r.center[i] for r in regions]) for i in range(3)] + \
               [np.array([r.length[i] for r in regions]) for i in range(3)] + \
               [np.array([r.exp for r in regions]), np.array([r.d for r in regions])] + \
               [np.array([r.u[i] for r in regions]) for i in range(3)] + [np.array([r.p for r in regions])]
        self._reg_keep = [np.ascontiguousarray(c, dtype=np.float64) for c in cols]
        self.lib.ref_set_regions(n, rt.ctypes.data_as(ip), *[_ptr(c) for c in self._reg_keep])

    def condinit(self, x, dx):
        """condinit on points x (N, 3) -> conserved (5, N)."""
        n = x.shape[0]
        out = np.zeros((NVAR, n))
        for i0 in range(0, n, 32):
            m = min(32, n - i0)
            xb = np.zeros((3, 32))
            xb[:, :m] = x[i0:i0 + m].T
            ub = np.zeros((NVAR, 32))
            self.lib.ref_condinit(m, _ptr(xb), _ptr(ub), dx)
            out[:, i0:i0 + m] = ub[:, :m]
        return out

    def restart(self, outdir, nrestart, nlevelmax, ngridmax):
        """The reference's restart (init_amr + init_hydro, ref_amr_restart)
        from outdir/output_NNNNN, NNNNN = nrestart.  The reference stops the
        process on a missing or incompatible file, so the file header is
        checked here first and a mismatch raises ValueError."""
        import os
        import struct
        import types
        tag = f"{nrestart:05d}"
        amr = os.path.join(outdir, f"output_{tag}", f"amr_{tag}.out00001")
        hyd = os.path.join(outdir, f"output_{tag}", f"hydro_{tag}.out00001")
        for f in (amr, hyd):
            if not os.path.exists(f):
                raise ValueError(f"{f} not found")
        with open(amr, "rb") as f:
            data = f.read(4096)
        vals, p = [], 0
        for _ in range(7):      # ncpu, ndim, (nx,ny,nz), nlevelmax, ngridmax, nboundary, ngrid
            (n,) = struct.unpack_from("<i", data, p)
            vals.append(struct.unpack_from(f"<{n // 4}i", data, p + 4))
            p += n + 8
        ncpu, ndim, nxyz, lmax, _, nbound, ngrid = vals
        if (ncpu[0], ndim[0], tuple(nxyz), nbound[0]) != (1, 3, (1, 1, 1), 0) or lmax[0] > nlevelmax \
                or ngrid[0] > ngridmax:
            raise ValueError(f"{amr} is not compatible with a one-domain restart at these sizes")
        ip = ctypes.POINTER(ctypes.c_int)
        self.lib.ref_amr_restart.argtypes = [ctypes.c_int] * 3
        self.lib.ref_amr_get_run.argtypes = [ctypes.c_int, _dp, ip, _dp, _dp]
        cwd = os.getcwd()
        os.chdir(outdir)
        try:
            self.lib.ref_amr_restart(nlevelmax, ngridmax, nrestart)
        finally:
            os.chdir(cwd)
        self.tree = types.SimpleNamespace(ngridmax=ngridmax, ncell=1 + 8 * ngridmax, nlevel=nlevelmax)
        self.nlevelmax = nlevelmax

    def run_state(self):
        """(t, nstep, nstep_coarse, ifout, iout, used_mem, dtold[1:L], dtnew[1:L])."""
        ip = ctypes.POINTER(ctypes.c_int)
        L = self.nlevelmax
        t = np.zeros(1)
        ints = np.zeros(5, dtype=np.int32)
        dto, dtn = np.zeros(L), np.zeros(L)
        self.lib.ref_amr_get_run(L, _ptr(t), ints.ctypes.data_as(ip), _ptr(dto), _ptr(dtn))
        return (float(t[0]),) + tuple(int(v) for v in ints) + (dto, dtn)

    def coarse_step(self):
        """One coarse step of the reference's own driver: adaptive_loop's
        body around amr_step(levelmin,1) (amr/adaptive_loop.f90:84-194,
        amr/amr_step.f90, pm/newdt_fine.f90, amr/update_time.f90), with the
        level time steps, t and nstep kept in the reference's module
        globals.  The coarse cell must be flagged (flag_coarse) before the
        first call.  Returns the step's conservation sums (mass, sum E*vol,
        internal energy) as courant_fine accumulates them."""
        sums = np.zeros(3)
        self.lib.ref_amr_coarse_step(_ptr(sums))
        return tuple(float(v) for v in sums)

    def set_clock(self, t=0.0, nstep=0, nstep_coarse=0):
        """The run clock of the reference's driver: t, nstep, nstep_coarse."""
        self.lib.ref_amr_set_clock(t, nstep, nstep_coarse)

    def clock(self):
        """(t, nstep, nstep_coarse, dtold[1:L], dtnew[1:L]) of the reference's driver."""
        st = self.run_state()
        return st[0], st[1], st[2], st[6], st[7]

    def baselines(self):
        """update_time's conservation baselines (einit, mass_tot_0)."""
        out = np.zeros(2)
        self.lib.ref_amr_get_baselines(_ptr(out))
        return out

    def dump(self, directory, t, nstep=0, nstep_coarse=0):
        """dump_all's snapshot files into an existing directory."""
        self.lib.ref_amr_dump(directory.encode(), t, nstep, nstep_coarse)

    def tree_arrays(self):
        """(son, father, nbor, flag1) as 0-based copies of the 1-based arrays."""
        t = self.tree
        ip = ctypes.POINTER(ctypes.c_int)
        son = np.zeros(t.ncell, dtype=np.int32)
        father = np.zeros(t.ngridmax, dtype=np.int32)
        nbor = np.zeros(6 * t.ngridmax, dtype=np.int32)
        flag1 = np.zeros(t.ncell, dtype=np.int32)
        self.lib.ref_amr_get_tree(son.ctypes.data_as(ip), father.ctypes.data_as(ip), nbor.ctypes.data_as(ip),
                                  flag1.ctypes.data_as(ip))
        return son, father, nbor, flag1


def reference_available():
    return os.path.exists(REF_LIB)
