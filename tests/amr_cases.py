"""Oct trees and states shared by the AMR tests and the golden generator."""
import numpy as np

from helpers import conserved
from ramses_amd.tree import OctTree


def sphere_tree(levelmin=3, levelmax=5, r0=0.22):
    """A refined ball around the box centre (radius shrinking with level),
    plus an off-centre blob, so every level has coarse/fine interfaces in
    every direction and around corners."""
    def refine(l, x):
        r = np.sqrt(((x - 0.5) ** 2).sum(1))
        r2 = np.sqrt(((x - np.array([0.2, 0.7, 0.35])) ** 2).sum(1))
        return (r < r0 * 0.75 ** (l - levelmin)) | (r2 < 0.08)
    return OctTree(levelmin, levelmax, refine)


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
