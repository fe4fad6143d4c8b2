"""TEST INFRASTRUCTURE ONLY: the reference's own level routines on a uniform
periodic level, for the parity tests and the CPU baseline of bench.py.

A fully refined periodic level ``L`` (n = 2^L cells per side, levelmin =
levelmax = L) is installed as an oct tree into the reference's module
globals (``ReferenceAMR``, oracle/_ref), and one level step runs the
reference's own routines in amr_step's order for a level without finer
levels (amr/amr_step.f90:327-424): courant_fine (hydro/courant_fine.f90:1-159,
started from dtnew = boxlen/smallc as newdt_fine does, pm/newdt_fine.f90:51),
set_unew, godunov_fine (hydro/godunov_fine.f90:5-35) and set_uold.  The
state crosses the boundary as u[var][z][y][x] (x fastest), the layout of the
resident-level entry points, mapped cell by cell to the reference's cell
index ncoarse + (ind-1)*ngridmax + igrid (hydro/godunov_fine.f90:601).
"""
import numpy as np

from . import Reference, ReferenceAMR


def uniform_tree_index(tree, level):
    """idx[z, y, x] = icell - 1 of every cell of a fully refined level."""
    from ramses_amd.tree import NCOARSE
    n = 2 ** level
    ids = np.asarray(tree.active[level], dtype=np.int64)
    gp = tree.grid_pos[ids]
    idx = np.empty((n, n, n), dtype=np.int64)
    for ind in range(8):
        ix, iy, iz = ind & 1, (ind >> 1) & 1, ind >> 2
        idx[2 * gp[:, 2] + iz, 2 * gp[:, 1] + iy, 2 * gp[:, 0] + ix] = NCOARSE + ind * tree.ngridmax + ids - 1
    return idx


class ReferenceUniformLevel:
    """The reference's godunov_fine/courant_fine on an n^3 periodic level."""

    def __init__(self, n, gamma, riemann="hllc", slope_type=1, courant_factor=0.5, smallr=1e-10, smallc=1e-10,
                 ref=None):
        from ramses_amd.tree import OctTree
        level = int(round(np.log2(n)))
        if 2 ** level != n or level < 2:
            raise ValueError("n must be a power of two >= 4")
        self.n, self.level, self.smallc = n, level, smallc
        self.R = ref or Reference()
        self.R.set_params(gamma=gamma, riemann=riemann, slope_type=slope_type, courant_factor=courant_factor,
                          smallr=smallr, smallc=smallc)
        self.A = ReferenceAMR(self.R)
        self.tree = OctTree(level, level)
        self.A.install(self.tree, nlevelmax=level)
        self.idx = uniform_tree_index(self.tree, level)

    def load(self, u):
        """Install u[var][z][y][x] as uold (and unew) of the level."""
        full = np.zeros((5, self.tree.ncell))
        for v in range(5):
            full[v, self.idx] = u[v]
        self.A.put(0, full)
        self.A.put(1, full)

    def state(self):
        full = self.A.get(0)
        return np.ascontiguousarray(np.stack([full[v][self.idx] for v in range(5)]))

    def courant(self, boxlen=1.0):
        """courant_fine of the current state: (dt, mass, sum E*vol, eint)."""
        return self.A.courant(self.level, boxlen / self.smallc)

    def step(self, dt=None):
        """One level step; dt None takes courant_fine's.  Returns the dt used."""
        if dt is None:
            dt = self.courant()[0]
        self.A.set_unew(self.level)
        self.A.godunov(self.level, dt)
        self.A.set_uold(self.level)
        return dt

    def release(self):
        from ramses_amd.tree import OctTree
        self.A.install(OctTree(1, 1), nlevelmax=1)
