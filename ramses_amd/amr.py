"""Oct-tree (AMR) levels resident on the GPU.

``AmrLevels`` keeps an oct tree (``tree.OctTree``: son / father / nbor and
the per-level active lists of amr/amr_commons.f90) and the conserved state
(uold / unew of hydro/hydro_commons.f90) in device memory, in the
reference's own index conventions, and exposes the reference's level
routines on it:

  godunov(l, dt)   godunov_fine(l)   hydro/godunov_fine.f90:5-35 (godfine1)
  set_unew(l)      set_unew(l)       hydro/godunov_fine.f90:40-130
  set_uold(l)      set_uold(l)       hydro/godunov_fine.f90:135-232
  upload(l)        upload_fine(l)    hydro/interpol_hydro.f90:5-263
  courant(l, dt0)  courant_fine(l)   hydro/courant_fine.f90:1-159

All work is enqueued on one stream without host synchronisation; ``dt`` may
be a host float or a device scalar produced by ``courant``.
"""
import ctypes

import numpy as np

from . import _check, _stream_handle, load_library


class ramses_amr_tree(ctypes.Structure):
    _fields_ = [("son", ctypes.c_void_p), ("father", ctypes.c_void_p), ("nbor", ctypes.c_void_p),
                ("ncoarse", ctypes.c_int), ("ngridmax", ctypes.c_int), ("nx", ctypes.c_int),
                ("ny", ctypes.c_int), ("nz", ctypes.c_int)]


class ramses_amr_level(ctypes.Structure):
    _fields_ = [("ilevel", ctypes.c_int), ("dx", ctypes.c_double), ("active", ctypes.c_void_p),
                ("noct", ctypes.c_int), ("pos", ctypes.c_void_p), ("active_coarse", ctypes.c_void_p),
                ("noct_coarse", ctypes.c_int)]


def _bind(lib):
    P = ctypes.POINTER
    lib.ramses_amr_config.argtypes = [ctypes.c_int, ctypes.c_int]
    lib.ramses_amr_config.restype = ctypes.c_int
    lib.ramses_amr_work_bytes.argtypes = [ctypes.c_int]
    lib.ramses_amr_work_bytes.restype = ctypes.c_ulong
    lib.ramses_amr_godunov.argtypes = [P(ramses_amr_tree), P(ramses_amr_level), ctypes.c_void_p, ctypes.c_void_p,
                                       ctypes.c_void_p, ctypes.c_void_p, ctypes.c_ulong, ctypes.c_void_p,
                                       ctypes.c_void_p]
    lib.ramses_amr_godunov.restype = ctypes.c_int
    lib.ramses_amr_copy.argtypes = [P(ramses_amr_tree), P(ramses_amr_level), ctypes.c_void_p, ctypes.c_void_p,
                                    ctypes.c_void_p]
    lib.ramses_amr_copy.restype = ctypes.c_int
    lib.ramses_amr_upload.argtypes = [P(ramses_amr_tree), P(ramses_amr_level), ctypes.c_void_p, ctypes.c_void_p]
    lib.ramses_amr_upload.restype = ctypes.c_int
    lib.ramses_amr_courant.argtypes = [P(ramses_amr_tree), P(ramses_amr_level), ctypes.c_void_p, ctypes.c_double,
                                       ctypes.c_void_p, ctypes.c_void_p]
    lib.ramses_amr_courant.restype = ctypes.c_int


def configure_interpolation(interpol_type=1, interpol_var=0):
    """hydro/hydro_parameters.f90:88-89 (defaults: minmod, conservative)."""
    lib = load_library()
    _bind(lib)
    _check(lib.ramses_amr_config(interpol_type, interpol_var), "ramses_amr_config")


class AmrLevels:
    def __init__(self, tree, u, device="cuda", stream=None, boxlen=1.0):
        """tree: tree.OctTree; u: conserved state (5, ncell) over every cell
        of every level, icell-1 indexed (OctTree.fill)."""
        import torch
        self.lib = load_library()
        _bind(self.lib)
        self.tree = tree
        self.stream = stream
        self.boxlen = boxlen
        self.device = device

        def dev_int(a):  # element 0 unused so that p[i] is the reference's element i
            return torch.as_tensor(np.concatenate([[0], np.asarray(a, dtype=np.int32)]), dtype=torch.int32,
                                   device=device)

        self.son = dev_int(tree.son)
        self.father = dev_int(tree.father)
        # nbor[(j-1)*ngridmax + igrid]: shift by one so that igrid indexes directly
        self.nbor = torch.as_tensor(np.concatenate([[0], tree.nbor]).astype(np.int32), device=device)
        self.t = ramses_amr_tree(self.son.data_ptr(), self.father.data_ptr(), self.nbor.data_ptr(), 1,
                                 tree.ngridmax, 1, 1, 1)
        self.active = {l: torch.as_tensor(ids, dtype=torch.int32, device=device) for l, ids in tree.active.items()}
        pos = np.zeros(tree.ngridmax + 1, dtype=np.int32)
        for l, ids in tree.active.items():
            pos[ids] = np.arange(len(ids), dtype=np.int32)
        self.pos = torch.as_tensor(pos, device=device)
        u = np.asarray(u, dtype=np.float64)
        full = np.zeros((5, tree.ncell + 1))
        full[:, 1:] = u
        self.uold = torch.as_tensor(full, device=device).contiguous()
        self.unew = self.uold.clone()
        maxoct = max(len(v) for v in tree.active.values())
        self.work = torch.empty(int(self.lib.ramses_amr_work_bytes(maxoct)), dtype=torch.uint8, device=device)
        self.error = torch.zeros(1, dtype=torch.int32, device=device)
        self.dt = torch.zeros(tree.nlevel + 2, dtype=torch.float64, device=device)

    def dx(self, l):
        return 0.5 ** l * self.boxlen

    def _level(self, l):
        ids = self.active.get(l)
        coarse = self.active.get(l - 1)
        return ramses_amr_level(l, self.dx(l), 0 if ids is None else ids.data_ptr(),
                                0 if ids is None else len(ids), self.pos.data_ptr(),
                                0 if coarse is None else coarse.data_ptr(), 0 if coarse is None else len(coarse))

    def _s(self):
        return _stream_handle(self.stream)

    def set_unew(self, l):
        _check(self.lib.ramses_amr_copy(ctypes.byref(self.t), ctypes.byref(self._level(l)),
                                        ctypes.c_void_p(self.uold.data_ptr()), ctypes.c_void_p(self.unew.data_ptr()),
                                        self._s()), "ramses_amr_copy")

    def set_uold(self, l):
        _check(self.lib.ramses_amr_copy(ctypes.byref(self.t), ctypes.byref(self._level(l)),
                                        ctypes.c_void_p(self.unew.data_ptr()), ctypes.c_void_p(self.uold.data_ptr()),
                                        self._s()), "ramses_amr_copy")

    def upload(self, l):
        _check(self.lib.ramses_amr_upload(ctypes.byref(self.t), ctypes.byref(self._level(l)),
                                          ctypes.c_void_p(self.uold.data_ptr()), self._s()), "ramses_amr_upload")

    def courant(self, l, dt_start=None):
        """courant_fine(l) into the device scalar self.dt[l]; returns it."""
        start = self.boxlen / 1e-10 if dt_start is None else dt_start
        _check(self.lib.ramses_amr_courant(ctypes.byref(self.t), ctypes.byref(self._level(l)),
                                           ctypes.c_void_p(self.uold.data_ptr()), start,
                                           ctypes.c_void_p(self.dt.data_ptr() + 8 * l), self._s()),
               "ramses_amr_courant")
        return self.dt[l]

    def godunov(self, l, dt=None):
        """godunov_fine(l) with dt (host float, or None for self.dt[l])."""
        if dt is not None:
            self.dt[l] = float(dt)
        _check(self.lib.ramses_amr_godunov(ctypes.byref(self.t), ctypes.byref(self._level(l)),
                                           ctypes.c_void_p(self.uold.data_ptr()), ctypes.c_void_p(self.unew.data_ptr()),
                                           ctypes.c_void_p(self.dt.data_ptr() + 8 * l),
                                           ctypes.c_void_p(self.work.data_ptr()), self.work.numel(),
                                           ctypes.c_void_p(self.error.data_ptr()), self._s()), "ramses_amr_godunov")

    def state(self, which="new"):
        t = self.unew if which == "new" else self.uold
        return t[:, 1:].cpu().numpy()

    def check(self):
        if int(self.error.item()) != 0:
            raise RuntimeError("oct tree violates the refinement rule (missing father neighbours)")


class AmrSolver(AmrLevels):
    """amr_step (amr/amr_step.f90) on a static oct tree: per coarse step the
    levels are advanced recursively with nsubcycle=2 time steps per level,
    the fine levels first, exactly in the reference's order of operations:

      newdt_fine(l) -> dt_l = min(courant_fine(l), dt_{l-1}/nsubcycle)  :327-330
      set_unew(l)                                                       :334
      amr_step(l+1, 1); amr_step(l+1, 2)                                :345-352
      godunov_fine(l) with dt_l                                         :389
      set_uold(l); upload_fine(l)                                       :424-442
      dt_{l-1} = dtold_l + dt_l after the second substep               :573-576

    The time steps stay on the device (min/sum of device scalars), so a
    coarse step needs no host synchronisation.  Refinement is out of scope:
    the tree is the one it was built with.
    """

    def __init__(self, tree, u, nsubcycle=2, **kw):
        super().__init__(tree, u, **kw)
        import torch
        self.levelmin = tree.levelmin
        self.nlevelmax = tree.nlevel
        self.nsubcycle = nsubcycle
        self.dtold = torch.zeros_like(self.dt)
        self.time = torch.zeros(1, dtype=torch.float64, device=self.device)
        self.fine_steps = 0

    def noct(self, l):
        ids = self.active.get(l)
        return 0 if ids is None else len(ids)

    def coarse_step(self):
        self.amr_step(self.levelmin, 1)

    def amr_step(self, l, icount):
        import torch
        if self.noct(l) == 0:
            return
        # newdt_fine (hydro/courant_fine.f90 via pm/newdt_fine.f90):
        # dtold = dtnew, then dtnew = courant_fine starting from boxlen/smallc
        from . import _config_cache
        self.dtold[l] = self.dt[l]
        self.courant(l, self.boxlen / _config_cache.get("smallc", 1e-10))
        if l > self.levelmin:
            self.dt[l] = torch.minimum(self.dt[l - 1] / float(self.nsubcycle), self.dt[l])
        self.set_unew(l)
        if l < self.nlevelmax and self.noct(l + 1) > 0:
            self.amr_step(l + 1, 1)
            if self.nsubcycle == 2:
                self.amr_step(l + 1, 2)
        else:
            if l < self.nlevelmax:
                self.dtold[l + 1] = self.dt[l] / float(self.nsubcycle)
                self.dt[l + 1] = self.dt[l] / float(self.nsubcycle)
            self.time += self.dt[l]        # update_time (amr/update_time.f90:217,331)
            self.fine_steps += 1
        self.godunov(l)
        self.set_uold(l)
        self.upload(l)
        if l > self.levelmin:
            if self.nsubcycle == 1:
                self.dt[l - 1] = self.dt[l]
            if icount == 2:
                self.dt[l - 1] = self.dtold[l] + self.dt[l]

    def cell_updates_per_coarse_step(self):
        """sum over levels of 8*noct_l*nsubcycle^(l-levelmin)."""
        return sum(8 * self.noct(l) * self.nsubcycle ** (l - self.levelmin)
                   for l in range(self.levelmin, self.nlevelmax + 1))
