"""Oct-tree (AMR) levels resident on the GPU.

``AmrLevels`` keeps an oct tree (``tree.OctTree``: son / father / nbor and
the per-level grid lists of amr/amr_commons.f90) and the conserved state
(uold / unew of hydro/hydro_commons.f90) in device memory, in the
reference's own index conventions, and exposes the reference's level
routines on it:

  godunov(l, dt)   godunov_fine(l)   hydro/godunov_fine.f90:5-35 (godfine1)
  set_unew(l)      set_unew(l)       hydro/godunov_fine.f90:40-130
  set_uold(l)      set_uold(l)       hydro/godunov_fine.f90:135-232
  upload(l)        upload_fine(l)    hydro/interpol_hydro.f90:5-263
  courant(l, dt0)  courant_fine(l)   hydro/courant_fine.f90:1-159
  flag(l, icount)  flag_fine(l)      amr/flag_utils.f90:57-106
  refine(l)        refine_fine(l)    amr/refine_utils.f90:332-585

The grid list of every level (active(l)%igrid) has room for ngridmax grids
so that refinement can grow it in place; the free grids sit in a device
ring in the reference's free-list order (headf .. tailf).  All work is
enqueued on one stream; only ``refine`` synchronises (to learn how many
octs it made and killed).
"""
import ctypes

import numpy as np

from . import _check, _config_cache, _stream_handle, load_library


class ramses_amr_tree(ctypes.Structure):
    _fields_ = [("son", ctypes.c_void_p), ("father", ctypes.c_void_p), ("nbor", ctypes.c_void_p),
                ("ncoarse", ctypes.c_int), ("ngridmax", ctypes.c_int), ("nx", ctypes.c_int),
                ("ny", ctypes.c_int), ("nz", ctypes.c_int)]


class ramses_amr_level(ctypes.Structure):
    _fields_ = [("ilevel", ctypes.c_int), ("dx", ctypes.c_double), ("active", ctypes.c_void_p),
                ("noct", ctypes.c_int), ("pos", ctypes.c_void_p), ("active_coarse", ctypes.c_void_p),
                ("noct_coarse", ctypes.c_int)]


class ramses_refine_params(ctypes.Structure):
    _fields_ = [("err_grad_d", ctypes.c_double), ("err_grad_u", ctypes.c_double), ("err_grad_p", ctypes.c_double),
                ("floor_d", ctypes.c_double), ("floor_u", ctypes.c_double), ("floor_p", ctypes.c_double),
                ("nexpand", ctypes.c_int)]


class ramses_amr_free(ctypes.Structure):
    _fields_ = [("ring", ctypes.c_void_p), ("head", ctypes.c_int), ("numbf", ctypes.c_int),
                ("used_mem", ctypes.c_int)]


MAX_LEVELS = 32  # RAMSES_MAX_LEVELS


class ramses_amr_mesh(ctypes.Structure):
    _fields_ = [("tree", ramses_amr_tree), ("active", ctypes.c_void_p * (MAX_LEVELS + 1)),
                ("noct", ctypes.c_int * (MAX_LEVELS + 1)), ("pos", ctypes.c_void_p), ("xg", ctypes.c_void_p),
                ("flag1", ctypes.c_void_p), ("flag2", ctypes.c_void_p), ("free", ramses_amr_free),
                ("uold", ctypes.c_void_p), ("unew", ctypes.c_void_p), ("dt", ctypes.c_void_p),
                ("dtold", ctypes.c_void_p), ("time", ctypes.c_void_p), ("error", ctypes.c_void_p),
                ("levelmin", ctypes.c_int), ("nlevelmax", ctypes.c_int),
                ("nsubcycle", ctypes.c_int * (MAX_LEVELS + 1)), ("adaptive", ctypes.c_int),
                ("boxlen", ctypes.c_double), ("nstep", ctypes.c_longlong), ("nstep_coarse", ctypes.c_longlong),
                ("cell_updates", ctypes.c_longlong), ("work", ctypes.c_void_p), ("work_bytes", ctypes.c_ulong),
                ("graph", ctypes.c_int), ("graph_capacity", ctypes.c_int), ("graph_state", ctypes.c_void_p),
                ("sums", ctypes.c_void_p), ("cons0", ctypes.c_void_p), ("nstep_coarse_old", ctypes.c_longlong)]


def _bind(lib):
    if getattr(lib, "_amr_bound", False):
        return
    P = ctypes.POINTER
    vp, ci, cd, ul = ctypes.c_void_p, ctypes.c_int, ctypes.c_double, ctypes.c_ulong
    T, L = P(ramses_amr_tree), P(ramses_amr_level)
    sig = {
        "ramses_amr_config": [ci, ci],
        "ramses_amr_work_bytes": [ci],
        "ramses_amr_godunov": [T, L, vp, vp, vp, vp, ul, vp, vp],
        "ramses_amr_copy": [T, L, vp, vp, vp],
        "ramses_amr_upload": [T, L, vp, vp],
        "ramses_amr_courant": [T, L, vp, cd, vp, vp],
        "ramses_refine_config": [P(ramses_refine_params)],
        "ramses_amr_flag_fine": [T, L, ci, ci, ci, vp, vp, vp, vp],
        "ramses_amr_refine_work_bytes": [ci, ci],
        "ramses_amr_refine_fine": [T, L, vp, P(ci), vp, P(ramses_amr_free), vp, vp, vp, vp, vp, vp, ul, vp, vp],
        "ramses_amr_coarse_step": [P(ramses_amr_mesh), vp],
        "ramses_amr_mesh_release": [P(ramses_amr_mesh)],
    }
    for name, args in sig.items():
        f = getattr(lib, name)
        f.argtypes = args
        f.restype = ul if name.endswith("work_bytes") else ci
    lib._amr_bound = True


def configure_interpolation(interpol_type=1, interpol_var=0):
    """hydro/hydro_parameters.f90:88-89 (defaults: minmod, conservative)."""
    lib = load_library()
    _bind(lib)
    _check(lib.ramses_amr_config(interpol_type, interpol_var), "ramses_amr_config")


def configure_refinement(err_grad_d=-1.0, err_grad_u=-1.0, err_grad_p=-1.0, floor_d=1e-10, floor_u=1e-10,
                         floor_p=1e-10, nexpand=1):
    """Refinement criteria, hydro/hydro_parameters.f90:47-52 and
    amr/amr_parameters.f90:85 (defaults: every criterion off, nexpand 1)."""
    lib = load_library()
    _bind(lib)
    p = ramses_refine_params(err_grad_d, err_grad_u, err_grad_p, floor_d, floor_u, floor_p, nexpand)
    _check(lib.ramses_refine_config(ctypes.byref(p)), "ramses_refine_config")


class SnapshotTree:
    """A restarted mesh (snapshot.read_output) in the form AmrLevels reads
    an OctTree: 0-based son/father/nbor/flag1 arrays, the level lists, grid
    centres, the free list and used_mem."""

    def __init__(self, mesh, levelmin):
        self.levelmin = levelmin
        self.levelmax = mesh.nlevelmax
        self.nlevel = max([l for l, g in mesh.lists.items() if len(g)] + [1])
        self.ngridmax = mesh.ngridmax
        self.ncell = 1 + 8 * mesh.ngridmax
        self.son, self.father, self.nbor = mesh.son, mesh.father, mesh.nbor
        self.flag1 = mesh.flag1
        self.active = {l: np.asarray(g, dtype=np.int32) for l, g in mesh.lists.items()}
        self.xg = mesh.xg
        self.free = mesh.free
        self.used_mem = mesh.used_mem if mesh.used_mem >= 0 else mesh.ngridmax - len(mesh.free)


class AmrLevels:
    def __init__(self, tree, u, device="cuda", stream=None, boxlen=1.0, nlevelmax=None):
        """tree: tree.OctTree; u: conserved state (5, ncell) over every cell
        of every level, icell-1 indexed (OctTree.fill), or None for zeros.
        nlevelmax: deepest level refinement may create (default: the tree's
        levelmax)."""
        import torch
        self.lib = load_library()
        _bind(self.lib)
        self.tree = tree
        self.stream = stream
        self.boxlen = boxlen
        self.device = device
        self.nlevelmax = max(tree.nlevel, tree.levelmax if nlevelmax is None else nlevelmax)
        if self.nlevelmax > MAX_LEVELS - 1:
            raise ValueError(f"at most {MAX_LEVELS - 1} levels")
        ngm = tree.ngridmax

        def dev_int(a):  # element 0 unused so that p[i] is the reference's element i
            return torch.as_tensor(np.concatenate([[0], np.asarray(a, dtype=np.int32)]), dtype=torch.int32,
                                   device=device)

        self.son = dev_int(tree.son)
        self.father = dev_int(tree.father)
        self.nbor = dev_int(tree.nbor)
        self.t = ramses_amr_tree(self.son.data_ptr(), self.father.data_ptr(), self.nbor.data_ptr(), 1, ngm, 1, 1, 1)
        # grid lists with room for ngridmax grids each, and their lengths
        self.lists = {}
        self.nocts = {}
        pos = np.zeros(ngm + 1, dtype=np.int32)
        for l in range(1, self.nlevelmax + 1):
            ids = np.asarray(tree.active.get(l, []), dtype=np.int32)
            buf = np.zeros(ngm, dtype=np.int32)
            buf[:len(ids)] = ids
            pos[ids] = np.arange(len(ids), dtype=np.int32)
            self.lists[l] = torch.as_tensor(buf, device=device)
            self.nocts[l] = len(ids)
        self.pos = torch.as_tensor(pos, device=device)
        # grid centres xg (amr/amr_commons.f90) in units of the coarse cell
        xg = np.zeros((ngm + 1, 3))
        if getattr(tree, "xg", None) is not None:
            xg[1:] = tree.xg
        else:
            for (l, gx, gy, gz), i in tree.grid_at.items():
                xg[i] = (np.array([gx, gy, gz]) + 0.5) * 0.5 ** (l - 1)
        self.xg = torch.as_tensor(xg, device=device)
        if getattr(tree, "free", None) is not None:
            free = np.asarray(tree.free, dtype=np.int32)
            used_mem = tree.used_mem
        else:
            # free grids in increasing order, as the reference's free list
            # after allocating grids 1..n in order (amr/init_amr.f90:245-256)
            used = np.zeros(ngm + 1, dtype=bool)
            for ids in tree.active.values():
                used[np.asarray(ids)] = True
            free = np.nonzero(~used[1:])[0].astype(np.int32) + 1
            used_mem = ngm - len(free)
        ring = np.zeros(ngm, dtype=np.int32)
        ring[:len(free)] = free
        self.ring = torch.as_tensor(ring, device=device)
        self.free = ramses_amr_free(self.ring.data_ptr(), 0, len(free), used_mem)
        if u is None:
            self.uold = torch.zeros((5, tree.ncell + 1), dtype=torch.float64, device=device)
        else:
            full = np.zeros((5, tree.ncell + 1))
            full[:, 1:] = np.asarray(u, dtype=np.float64)
            self.uold = torch.as_tensor(full, device=device).contiguous()
        self.unew = self.uold.clone()
        self.flag1 = torch.zeros(tree.ncell + 1, dtype=torch.int32, device=device)
        if getattr(tree, "flag1", None) is not None:
            self.flag1[1:] = torch.as_tensor(np.asarray(tree.flag1, dtype=np.int32), device=device)
        self.flag2 = torch.zeros_like(self.flag1)
        self.work = torch.empty(0, dtype=torch.uint8, device=device)
        self.error = torch.zeros(1, dtype=torch.int32, device=device)
        self.dt = torch.zeros(MAX_LEVELS + 2, dtype=torch.float64, device=device)

    # -- bookkeeping ---------------------------------------------------------
    def noct(self, l):
        return self.nocts.get(l, 0)

    @property
    def active(self):
        """Current grid list of every non-empty level (device views)."""
        return {l: self.lists[l][:n] for l, n in self.nocts.items() if n > 0}

    def grids(self, l):
        return self.lists[l][:self.noct(l)].cpu().numpy()

    def dx(self, l):
        return 0.5 ** l * self.boxlen

    def _level(self, l):
        lst, n = self.lists.get(l), self.noct(l)
        coarse, nc = self.lists.get(l - 1), self.noct(l - 1)
        return ramses_amr_level(l, self.dx(l), 0 if lst is None else lst.data_ptr(), n, self.pos.data_ptr(),
                                0 if (coarse is None or nc == 0) else coarse.data_ptr(), nc)

    def _s(self):
        return _stream_handle(self.stream)

    def _work(self, nbytes):
        import torch
        if self.work.numel() < nbytes:
            self.work = torch.empty(int(nbytes * 1.25) + 4096, dtype=torch.uint8, device=self.device)
        return self.work

    # -- level routines ------------------------------------------------------
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
        w = self._work(self.lib.ramses_amr_work_bytes(self.noct(l)))
        _check(self.lib.ramses_amr_godunov(ctypes.byref(self.t), ctypes.byref(self._level(l)),
                                           ctypes.c_void_p(self.uold.data_ptr()), ctypes.c_void_p(self.unew.data_ptr()),
                                           ctypes.c_void_p(self.dt.data_ptr() + 8 * l),
                                           ctypes.c_void_p(w.data_ptr()), w.numel(),
                                           ctypes.c_void_p(self.error.data_ptr()), self._s()), "ramses_amr_godunov")

    def flag(self, l, icount, levelmin):
        """flag_fine(l, icount): the refinement map of level l in flag1."""
        _check(self.lib.ramses_amr_flag_fine(ctypes.byref(self.t), ctypes.byref(self._level(l)), icount, levelmin,
                                             self.nlevelmax, ctypes.c_void_p(self.uold.data_ptr()),
                                             ctypes.c_void_p(self.flag1.data_ptr()),
                                             ctypes.c_void_p(self.flag2.data_ptr()), self._s()),
               "ramses_amr_flag_fine")

    def refine(self, l):
        """refine_fine(l): make and kill the octs of level l+1 from flag1."""
        if l >= self.nlevelmax or self.noct(l) == 0:
            return
        nf = ctypes.c_int(self.noct(l + 1))
        w = self._work(self.lib.ramses_amr_refine_work_bytes(self.noct(l), nf.value))
        _check(self.lib.ramses_amr_refine_fine(ctypes.byref(self.t), ctypes.byref(self._level(l)),
                                               ctypes.c_void_p(self.lists[l + 1].data_ptr()), ctypes.byref(nf),
                                               ctypes.c_void_p(self.pos.data_ptr()), ctypes.byref(self.free),
                                               ctypes.c_void_p(self.flag1.data_ptr()),
                                               ctypes.c_void_p(self.flag2.data_ptr()),
                                               ctypes.c_void_p(self.xg.data_ptr()),
                                               ctypes.c_void_p(self.uold.data_ptr()),
                                               ctypes.c_void_p(self.unew.data_ptr()), ctypes.c_void_p(w.data_ptr()),
                                               w.numel(), ctypes.c_void_p(self.error.data_ptr()), self._s()),
               "ramses_amr_refine_fine")
        self.nocts[l + 1] = nf.value

    def cells(self, l):
        """(icell, centres) of every cell of level l on the device, cell-major
        (ind 1..8 outer, the level's grid order inner); centres in units of
        the coarse cell (amr/amr_commons.f90 xg plus the cell offsets)."""
        import torch
        g = self.lists[l][:self.noct(l)].long()
        ngm = self.tree.ngridmax
        ind = torch.arange(8, device=self.device)
        off = torch.stack([(ind & 1), (ind >> 1) & 1, ind >> 2], 1).double() - 0.5
        x = self.xg[g][None, :, :] + off[:, None, :] * 0.5 ** l
        icell = 1 + ind[:, None] * ngm + g[None, :]
        return icell.reshape(-1), x.reshape(-1, 3)

    def init_flow_fine(self, l, ic):
        """init_flow_fine(l) (hydro/init_flow_fine.f90:28-470, the condinit
        branch): every cell of level l takes ic(centres (N,3), dx) -> (5, N)."""
        if self.noct(l) == 0:
            return
        icell, x = self.cells(l)
        self.uold[:, icell] = ic(x * self.boxlen, 0.5 ** l * self.boxlen)

    def defrag(self):
        """defrag (amr/load_balance.f90:993-1608): renumber the grids 1..n in
        level-list order (level 1 first), remap father / nbor / son, move xg,
        flag1 and uold with their grids (positions past n up to the old
        highest grid become zero), leave unew where it is, and make the free
        list n+1..ngridmax.  The reference runs it before every snapshot."""
        import torch
        ngm, nc, dev = self.tree.ngridmax, 1, self.device
        order = torch.cat([self.lists[l][:self.noct(l)] for l in range(1, self.nlevelmax + 1)]).long()
        n2 = order.numel()
        gmax = int(order.max().item())
        new = torch.zeros(ngm + 1, dtype=torch.long, device=dev)
        new[order] = torch.arange(1, n2 + 1, device=dev)

        def remap(c):
            g = torch.remainder(c - nc - 1, ngm) + 1
            return torch.where(c > nc, nc + torch.div(c - nc - 1, ngm, rounding_mode="floor") * ngm + new[g], c)

        def regrid(dst, base, vals):
            buf = torch.zeros(gmax, dtype=dst.dtype, device=dev)
            buf[:n2] = vals
            dst[base + 1:base + gmax + 1] = buf

        regrid(self.father, 0, remap(self.father[order].long()).to(self.father.dtype))
        for j in range(6):
            regrid(self.nbor, j * ngm, remap(self.nbor[j * ngm + order].long()).to(self.nbor.dtype))
        xg = torch.zeros((gmax, 3), dtype=self.xg.dtype, device=dev)
        xg[:n2] = self.xg[order]
        self.xg[1:gmax + 1] = xg
        for ind in range(8):
            base = nc + ind * ngm
            regrid(self.son, base, new[self.son[base + order].long()].to(self.son.dtype))
            regrid(self.flag1, base, self.flag1[base + order])
            for v in range(5):
                regrid(self.uold[v], base, self.uold[v, base + order])
        start = 0
        for l in range(1, self.nlevelmax + 1):
            n = self.noct(l)
            ids = torch.arange(start + 1, start + n + 1, dtype=torch.int32, device=dev)
            self.lists[l][:n] = ids
            self.pos[start + 1:start + n + 1] = torch.arange(n, dtype=torch.int32, device=dev)
            start += n
        self.ring[:ngm - n2] = torch.arange(n2 + 1, ngm + 1, dtype=torch.int32, device=dev)
        self.free.head = 0
        self.free.numbf = ngm - n2

    # -- host views ------------------------------------------------------------
    def state(self, which="new"):
        t = self.unew if which == "new" else self.uold
        return t[:, 1:].cpu().numpy()

    def tree_arrays(self):
        """(son, father, nbor, flag1) as the reference's 1-based arrays, 0-based here."""
        return (self.son[1:].cpu().numpy(), self.father[1:].cpu().numpy(), self.nbor[1:].cpu().numpy(),
                self.flag1[1:].cpu().numpy())

    def host_mesh(self):
        """The mesh as a snapshot.Mesh (host copies)."""
        from .snapshot import Mesh
        son, father, nbor, flag1 = self.tree_arrays()
        return Mesh(self.tree.ngridmax, self.nlevelmax, {l: self.grids(l) for l in range(1, self.nlevelmax + 1)},
                    son, father, nbor, flag1, self.xg[1:].cpu().numpy(), self.state("old"), self.free_list(),
                    self.free.used_mem)

    def free_list(self):
        r = self.ring.cpu().numpy()
        ngm = self.tree.ngridmax
        return r[(self.free.head + np.arange(self.free.numbf)) % ngm]

    def check(self):
        e = int(self.error.item())
        if e & 1:
            raise RuntimeError("oct tree violates the refinement rule (missing father neighbours)")
        if e & 2:
            raise RuntimeError("a new oct has no neighbour cell (refinement rule broken)")


class AmrSolver(AmrLevels):
    """amr_step (amr/amr_step.f90) with nsubcycle 1 or 2 per level: per
    coarse step the levels are advanced recursively, the fine levels first,
    in the reference's order of operations:

      refine_fine(i), i = l..nlevelmax, at l == levelmin or icount > 1  :40-93
      newdt_fine(l) -> dt_l = min(courant_fine(l), dt_{l-1}/nsubcycle)  :327-330
      set_unew(l)                                                       :334
      amr_step(l+1, 1); amr_step(l+1, 2)                                :345-352
      godunov_fine(l) with dt_l                                         :389
      set_uold(l); upload_fine(l)                                       :424-442
      flag_fine(l, icount)                                              :532
      dt_{l-1} = dtold_l + dt_l after the second substep               :573-576

    The time steps stay on the device (min/sum of device scalars).  With
    refine=False the tree is static (the reference's `static` mode).
    """

    def __init__(self, tree, u, nsubcycle=2, refine=False, graph=False, graph_capacity=0, sums=False, **kw):
        """nsubcycle: 1 or 2 for every level, or a list per level (index
        0 is level 1), as &RUN_PARAMS nsubcycle (amr/amr_commons.f90:51).
        graph: run each coarse step as the library's captured hipGraph (one
        launch and one synchronisation per step; identical results), with a
        workspace for graph_capacity octs per level (0: ngridmax).
        sums: keep courant_fine's conservation totals of each coarse step
        in ``self.sums`` (mass_tot, ekin_tot = sum E*vol, eint_tot;
        hydro/courant_fine.f90:95-119,152-154), on the device."""
        super().__init__(tree, u, **kw)
        self.graph = bool(graph)
        self.graph_capacity = int(graph_capacity)
        import torch
        nsub = [nsubcycle] * (self.nlevelmax + 1) if isinstance(nsubcycle, int) else list(nsubcycle)
        nsub = (nsub + [2] * (self.nlevelmax + 1))[:self.nlevelmax + 1]
        if any(n not in (1, 2) for n in nsub):
            raise ValueError("nsubcycle must be 1 or 2 per level")
        self.levelmin = tree.levelmin
        self.nsub = {l: nsub[l - 1] for l in range(1, self.nlevelmax + 1)}
        self.nsubcycle = nsubcycle
        self.adaptive = refine
        self.dtold = torch.zeros_like(self.dt)
        self.time = torch.zeros(1, dtype=torch.float64, device=self.device)
        self.fine_steps = 0        # nstep: finest-level update_time calls
        self.nstep_coarse = 0
        self.cell_updates = 0
        self.sums = torch.zeros(3, dtype=torch.float64, device=self.device) if sums else None
        # update_time's conservation baselines {einit, mass_tot_0}
        # (amr/update_time.f90:229-246) and its nstep_coarse_old, which
        # adaptive_loop starts equal to nstep_coarse (amr/adaptive_loop.f90:75)
        self.cons0 = torch.zeros(2, dtype=torch.float64, device=self.device)
        self.nstep_coarse_old = 0
        self._mesh = ramses_amr_mesh()

    @classmethod
    def from_snapshot(cls, outdir, nrestart, levelmin, nlevelmax=None, ngridmax=None, smallr=1e-10, gamma=None,
                      ordering=None, tout=None, **kw):
        """The restart of amr/adaptive_loop.f90:32-34: the solver and run
        state (snapshot.RunInfo) of output_NNNNN under outdir, with
        NNNNN = nrestart.  The mesh, free list, refinement map, state, level
        dts, t, nstep and nstep_coarse are those init_amr and init_hydro
        rebuild (snapshot.read_output); kw as for AmrSolver."""
        from . import snapshot
        mesh, run = snapshot.read_output(outdir, nrestart, ngridmax=ngridmax, nlevelmax=nlevelmax, smallr=smallr,
                                         gamma=gamma, ordering=ordering, tout=tout)
        run.levelmin = levelmin
        tree = SnapshotTree(mesh, levelmin)
        s = cls(tree, mesh.uold, nlevelmax=mesh.nlevelmax, boxlen=run.boxlen, **kw)
        n = len(run.dtnew)
        dt = s.dt.cpu().numpy()
        dto = s.dtold.cpu().numpy()
        dt[1:n + 1], dto[1:n + 1] = run.dtnew, run.dtold
        s.dt.copy_(s.dt.new_tensor(dt))
        s.dtold.copy_(s.dtold.new_tensor(dto))
        s.time.fill_(run.t)
        s.fine_steps, s.nstep_coarse = run.nstep, run.nstep_coarse
        s.nstep_coarse_old = run.nstep_coarse
        s.cons0.copy_(s.cons0.new_tensor([run.einit, run.mass_tot_0]))
        return s, run

    def __del__(self):
        try:
            self.lib.ramses_amr_mesh_release(ctypes.byref(self._mesh))
        except Exception:
            pass

    def flag_all(self):
        """The initial refinement map: flag_fine(l, 2) from the finest level
        to the coarsest (amr/adaptive_loop.f90:187-190)."""
        for l in range(self.nlevelmax, 0, -1):
            if self.noct(l):
                self.flag(l, 2, self.levelmin)

    def init_flow(self, ic):
        """init_flow (hydro/init_flow_fine.f90:5-23): condinit on every level
        >= levelmin from the finest, each level followed by upload_fine."""
        for l in range(self.nlevelmax, 0, -1):
            if l >= self.levelmin:
                self.init_flow_fine(l, ic)
            self.upload(l)

    def init_refine(self, ic):
        """Build the initial mesh and state from the initial conditions ic as
        init_refine and init_refine_2 do (amr/init_refine.f90:1-150):
        repeated init_flow / flag / refine passes, then passes that refine,
        re-impose the initial conditions level by level, upload and flag."""
        L, lmin = self.nlevelmax, self.levelmin
        for il in range(lmin + 1, L + 1):
            self.init_flow(ic)
            for l in range(L - 1, 0, -1):
                self.flag(l, 2, lmin)
            for l in range(lmin, L):
                self.refine(l)
            if self.noct(il) == 0:
                break
        self.init_flow(ic)
        for _ in range(lmin, L + 2):
            for l in range(1, L + 1):
                if l >= lmin:
                    self.refine(l)
                self.init_flow_fine(l, ic)
            for l in range(L, lmin - 1, -1):
                self.upload(l)
            for l in range(L, 0, -1):
                self.flag(l, 2, lmin)

    def capture(self):
        """Capture the coarse step's hipGraph now (graph mode), without
        running a step; the next coarse_step only replays it."""
        if not self.graph:
            return
        self.graph = 2
        try:
            self.coarse_step()
        finally:
            self.graph = True

    def coarse_step(self):
        """One coarse step of adaptive_loop (amr/adaptive_loop.f90:86-191):
        amr_step(levelmin, 1) and, with refinement on, upload_fine and
        flag_fine below levelmin and flag_coarse -- run by the library's
        native driver (ramses_amr_coarse_step) in one call."""
        m = self._mesh
        m.tree = self.t
        for l in range(1, self.nlevelmax + 1):
            m.active[l] = self.lists[l].data_ptr()
            m.noct[l] = self.noct(l)
            m.nsubcycle[l] = self.nsub[l]
        m.pos, m.xg = self.pos.data_ptr(), self.xg.data_ptr()
        m.flag1, m.flag2 = self.flag1.data_ptr(), self.flag2.data_ptr()
        m.free = self.free
        m.uold, m.unew = self.uold.data_ptr(), self.unew.data_ptr()
        m.dt, m.dtold, m.time = self.dt.data_ptr(), self.dtold.data_ptr(), self.time.data_ptr()
        m.error = self.error.data_ptr()
        m.levelmin, m.nlevelmax = self.levelmin, self.nlevelmax
        m.adaptive = 1 if self.adaptive else 0
        m.boxlen = self.boxlen
        m.nstep, m.nstep_coarse, m.cell_updates = self.fine_steps, self.nstep_coarse, self.cell_updates
        m.graph, m.graph_capacity = int(self.graph), self.graph_capacity
        m.sums = None if self.sums is None else self.sums.data_ptr()
        m.cons0, m.nstep_coarse_old = self.cons0.data_ptr(), self.nstep_coarse_old
        rc = self.lib.ramses_amr_coarse_step(ctypes.byref(m), self._s())
        for l in range(1, self.nlevelmax + 1):
            self.nocts[l] = m.noct[l]
        self.free.head, self.free.numbf, self.free.used_mem = m.free.head, m.free.numbf, m.free.used_mem
        self.fine_steps, self.nstep_coarse, self.cell_updates = m.nstep, m.nstep_coarse, m.cell_updates
        self.nstep_coarse_old = m.nstep_coarse_old
        _check(rc, "ramses_amr_coarse_step")

    def write_output(self, outdir, run):
        """dump_all for the hydro mesh: defrag, then the snapshot files
        (snapshot.write_output) with this solver's time and level dts."""
        from . import snapshot
        self.defrag()
        run.t = float(self.time.item())
        run.nstep, run.nstep_coarse = self.fine_steps, self.nstep_coarse
        run.einit, run.mass_tot_0 = (float(v) for v in self.cons0.cpu().numpy())
        run.dtold = tuple(self.dtold[1:self.nlevelmax + 1].cpu().numpy())
        run.dtnew = tuple(self.dt[1:self.nlevelmax + 1].cpu().numpy())
        return snapshot.write_output(self.host_mesh(), run, outdir)

    def cell_updates_per_coarse_step(self):
        """sum over levels of 8*noct_l*prod(nsubcycle of the coarser levels)."""
        tot, mult = 0, 1
        for l in range(self.levelmin, self.nlevelmax + 1):
            tot += 8 * self.noct(l) * mult
            mult *= self.nsub[l]
        return tot
