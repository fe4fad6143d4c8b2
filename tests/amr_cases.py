"""Oct trees and states shared by the AMR tests and the golden generator."""
import numpy as np

from helpers import conserved
from ramses_amd.tree import OctTree


def sphere_tree(levelmin=3, levelmax=5, r0=0.22, extra_grids=0):
    """A refined ball around the box centre (radius shrinking with level),
    plus an off-centre blob, so every level has coarse/fine interfaces in
    every direction and around corners."""
    def refine(l, x):
        r = np.sqrt(((x - 0.5) ** 2).sum(1))
        r2 = np.sqrt(((x - np.array([0.2, 0.7, 0.35])) ** 2).sum(1))
        return (r < r0 * 0.75 ** (l - levelmin)) | (r2 < 0.08)
    return OctTree(levelmin, levelmax, refine, extra_grids=extra_grids)


def smooth_state(tree, gamma=5.0 / 3.0, seed=0, amp=0.3):
    """Smooth periodic fields evaluated at every cell centre of every level."""
    rng = np.random.default_rng(seed)
    k = rng.integers(1, 3, size=(5, 3))
    ph = rng.uniform(0, 2 * np.pi, size=(5, 3))

    def fn(x, l):
        s = [np.prod(np.sin(2 * np.pi * k[i] * x + ph[i]), axis=1) for i in range(5)]
        rho = 1.0 + amp * s[0] + 0.1 * (x[:, 0] > 0.5)
        vel = np.stack([0.8 * s[1], 0.8 * s[2], 0.8 * s[3]])
        p = 1.0 + amp * s[4] + 2.0 * np.exp(-((x - 0.5) ** 2).sum(1) / 0.01)
        return conserved(rho, vel, p, gamma)
    return tree.fill(fn)


def reference_coarse_step(A, tree, levelmin, dt, dtold, nsub=2):
    """The same amr_step sequence driving the reference's compiled routines
    (host floats for the level time steps)."""
    nlevelmax = tree.nlevel

    def step(l, icount):
        if len(tree.active.get(l, [])) == 0:
            return
        dtold[l] = dt[l]
        dt[l] = A.courant(l, 1.0 / 1e-10)[0]
        if l > levelmin:
            dt[l] = min(dt[l - 1] / float(nsub), dt[l])
        A.set_unew(l)
        if l < nlevelmax and len(tree.active.get(l + 1, [])) > 0:
            step(l + 1, 1)
            step(l + 1, 2)
        elif l < nlevelmax:
            dtold[l + 1] = dt[l] / float(nsub)
            dt[l + 1] = dt[l] / float(nsub)
        A.godunov(l, dt[l])
        A.set_uold(l)
        A.upload(l)
        if l > levelmin and icount == 2:
            dt[l - 1] = dtold[l] + dt[l]
    step(levelmin, 1)


def reference_amr_step(A, levelmin, nlevelmax, dt, dtold, adaptive=True, smallc=1e-10, boxlen=1.0, nsub=None,
                       stats=None, clock=None):
    """One coarse step of amr_step (amr/amr_step.f90) driving the
    reference's compiled routines, with refinement (refine_fine at
    levelmin / second substeps, flag_fine after each level step).
    nsub: {level: 1 or 2} (default 2 everywhere); stats["cells"] counts
    8 cells per oct per godunov_fine call and stats["sums"] adds up the
    three conservation sums each courant_fine call returns (mass, sum
    E*vol, eint), as mass_tot etc. do; clock["t"] / clock["nstep"]
    advance as update_time advances t and nstep (amr/update_time.f90:330-332)
    where amr_step calls it (amr/amr_step.f90:357,363)."""
    ns = nsub or {}

    def update_time(l):
        if clock is not None:
            clock["t"] = clock["t"] + dt[l]
            clock["nstep"] = clock["nstep"] + 1

    def n(l):
        return len(A.grids(l)) if 1 <= l <= nlevelmax else 0

    def step(l, icount):
        if n(l) == 0:
            return
        if adaptive and levelmin < nlevelmax and (l == levelmin or icount > 1):
            for i in range(l, nlevelmax + 1):
                A.refine_fine(i)
        dtold[l] = dt[l]
        cf = A.courant(l, boxlen / smallc)
        dt[l] = cf[0]
        if stats is not None:
            acc = stats.setdefault("sums", [0.0, 0.0, 0.0])
            for k in range(3):
                acc[k] = acc[k] + cf[1 + k]
        if l > levelmin:
            dt[l] = min(dt[l - 1] / float(ns.get(l - 1, 2)), dt[l])
        A.set_unew(l)
        if l < nlevelmax and n(l + 1) > 0:
            step(l + 1, 1)
            if ns.get(l, 2) == 2:
                step(l + 1, 2)
        elif l < nlevelmax:
            dtold[l + 1] = dt[l] / float(ns.get(l, 2))
            dt[l + 1] = dt[l] / float(ns.get(l, 2))
            update_time(l)
        else:
            update_time(l)
        A.godunov(l, dt[l])
        if stats is not None:
            stats["cells"] = stats.get("cells", 0) + 8 * n(l)
        A.set_uold(l)
        A.upload(l)
        if adaptive:
            A.flag_fine(l, icount)
        if l > levelmin:
            if ns.get(l - 1, 2) == 1:
                dt[l - 1] = dt[l]
            if icount == 2:
                dt[l - 1] = dtold[l] + dt[l]
    step(levelmin, 1)
    if adaptive and levelmin < nlevelmax:  # amr/adaptive_loop.f90:143-191
        for l in range(levelmin - 1, 0, -1):
            A.upload(l)
        for l in range(levelmin - 1, 0, -1):
            A.flag_fine(l, 2)
        A.flag_coarse()


def reference_driver_step(A, dt=None, dtold=None, stats=None, clock=None):
    """One coarse step of the reference's OWN driver (oracle.ReferenceAMR.
    coarse_step: adaptive_loop's body around amr_step(levelmin,1),
    amr/adaptive_loop.f90:84-194 with amr/amr_step.f90, pm/newdt_fine.f90
    and amr/update_time.f90), which keeps the level time steps, t and nstep
    in its own module globals.  Copies them out for the caller: dt[l] /
    dtold[l] (dtnew / dtold of level l), stats["sums"] the step's
    conservation sums, clock["t"] / clock["nstep"].  The coarse cell must
    have been flagged (A.flag_coarse()) before the first step."""
    sums = A.coarse_step()
    t, nstep, _, dto, dtn = A.clock()
    for l in range(1, len(dtn) + 1):
        if dt is not None and l < len(dt):
            dt[l] = float(dtn[l - 1])
        if dtold is not None and l < len(dtold):
            dtold[l] = float(dto[l - 1])
    if stats is not None:
        stats["sums"] = list(sums)
    if clock is not None:
        clock["t"], clock["nstep"] = t, nstep


def reference_init_refine(A, levelmin, nlevelmax):
    """init_refine + init_refine_2 (amr/init_refine.f90) on a tree installed
    up to levelmin, with the reference's own init_flow / flag_fine /
    refine_fine; init is set during the first phase as the reference sets
    it (make_grid_fine then leaves new cells to init_flow)."""
    def n(l):
        return len(A.grids(l))
    A.set_init(True)
    for il in range(levelmin + 1, nlevelmax + 1):
        A.init_flow()
        for l in range(nlevelmax - 1, 0, -1):
            A.flag_fine(l, 2)
        for l in range(1, nlevelmax):
            A.refine_fine(l)
        if n(il) == 0:
            break
    A.set_init(False)
    A.init_flow()
    for _ in range(levelmin, nlevelmax + 2):
        for l in range(1, nlevelmax + 1):
            A.refine_fine(l)
            A.init_flow_fine(l)
        for l in range(nlevelmax, levelmin - 1, -1):
            A.upload(l)
        for l in range(nlevelmax, 0, -1):
            A.flag_fine(l, 2)


def digest(a):
    """sha256 of an array's little-endian bytes (int32 or float64)."""
    import hashlib
    a = np.asarray(a)
    dt = "<i4" if a.dtype.kind in "iu" else "<f8"
    return hashlib.sha256(np.ascontiguousarray(a, dtype=dt).tobytes()).hexdigest()


def mesh_digests(son, father, nbor, flag1, grids, free, uold, unew):
    """Digests of a whole adaptive mesh: tree arrays, flags, grid lists per
    level (dict), free list and both state arrays."""
    return dict(son=digest(son), father=digest(father), nbor=digest(nbor), flag1=digest(flag1),
                grids={str(l): digest(g) for l, g in grids.items()}, nocts={str(l): len(g) for l, g in grids.items()},
                free=digest(free), nfree=len(free), uold=digest(uold), unew=digest(unew))


REFINE_CASE = dict(levelmin=3, levelmax=5, nlevelmax=6, extra_grids=20000, seed=3, err_grad_d=0.12,
                   err_grad_p=0.25, gamma=5.0 / 3.0, riemann="hllc", slope_type=1, courant_factor=0.5)
