"""GPU parity of the oct-tree level path (gather with buffer-oct
interpolation, flux reset, update, reflux, courant, upload) against the
reference's own compiled level routines and the committed AMR fixtures."""
import json
import os

import numpy as np
import pytest

import oracle
from amr_cases import smooth_state, sphere_tree

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(__file__), "golden", "amr_sphere.npz")
needs_ref = pytest.mark.skipif(not oracle.reference_available(), reason="reference build (oracle/_ref) not present")


def _gpu_sweep(tree, u, l, dt, levelmin=3):
    from ramses_amd.amr import AmrLevels
    g = AmrLevels(tree, u)
    if l > levelmin:
        g.set_unew(l - 1)
    g.set_unew(l)
    g.godunov(l, dt)
    g.check()
    return g.state("new")


def test_amr_sweeps_golden(gpu):
    """Every level sweep of the committed fixture, bitwise."""
    import ramses_amd as ra
    from ramses_amd.amr import AmrLevels, configure_interpolation
    d = np.load(GOLD)
    meta = json.loads(str(d["meta"]))
    tree = sphere_tree()
    u = smooth_state(tree)
    assert np.array_equal(u, d["u0"])
    ra.configure(gamma=meta["gamma"], riemann=meta["riemann"], slope_type=1, courant_factor=0.5)
    configure_interpolation(1, 0)
    for l in meta["levels"]:
        g = AmrLevels(tree, u)
        dt = float(g.courant(l))
        assert dt == float(d[f"dt{l}"]), l
        out = _gpu_sweep(tree, u, l, dt)
        idx = d[f"idx{l}"]
        assert np.array_equal(np.nonzero((out != u).any(0))[0], idx), l
        assert np.array_equal(out[:, idx], d[f"val{l}"]), l
        # upload_fine of the level
        g = AmrLevels(tree, u)
        g.upload(l)
        uo = g.state("old")
        uidx = d[f"upidx{l}"]
        assert np.array_equal(np.nonzero((uo != u).any(0))[0], uidx), l
        assert np.array_equal(uo[:, uidx], d[f"upval{l}"]), l
    ra.configure()


@needs_ref
@pytest.mark.parametrize("riemann,slope", [("hllc", 1), ("llf", 2), ("hll", 8)])
def test_amr_sweeps_vs_reference(gpu, riemann, slope):
    import ramses_amd as ra
    from ramses_amd.amr import configure_interpolation
    R = oracle.Reference()
    A = oracle.ReferenceAMR(R)
    tree = sphere_tree()
    u = smooth_state(tree, seed=11)
    A.install(tree)
    A.set_interpol(1, 0)
    R.set_params(gamma=5 / 3, riemann=riemann, slope_type=slope, courant_factor=0.5)
    ra.configure(gamma=5 / 3, riemann=riemann, slope_type=slope, courant_factor=0.5)
    configure_interpolation(1, 0)
    for l in (3, 4, 5):
        A.put(0, u)
        A.put(1, u)
        dt = A.courant(l, 1e10)[0]
        if l > 3:
            A.set_unew(l - 1)
        A.set_unew(l)
        A.godunov(l, dt)
        ref = A.get(1)
        out = _gpu_sweep(tree, u, l, dt)
        assert np.array_equal(out, ref), (riemann, slope, l)
    ra.configure()


@needs_ref
@pytest.mark.parametrize("itype,ivar", [(0, 0), (2, 0), (3, 0), (1, 1), (2, 2), (3, 2)])
def test_amr_interpolation_vs_reference(gpu, itype, ivar):
    import ramses_amd as ra
    from ramses_amd.amr import configure_interpolation
    R = oracle.Reference()
    A = oracle.ReferenceAMR(R)
    tree = sphere_tree()
    u = smooth_state(tree, seed=5, amp=0.4)
    A.install(tree)
    A.set_interpol(itype, ivar)
    R.set_params(gamma=1.4, riemann="hllc", slope_type=1, courant_factor=0.5)
    ra.configure(gamma=1.4, riemann="hllc", slope_type=1, courant_factor=0.5)
    configure_interpolation(itype, ivar)
    l = 5
    A.put(0, u)
    A.put(1, u)
    dt = A.courant(l, 1e10)[0]
    A.set_unew(l - 1)
    A.set_unew(l)
    A.godunov(l, dt)
    out = _gpu_sweep(tree, u, l, dt)
    assert np.array_equal(out, A.get(1)), (itype, ivar)
    configure_interpolation(1, 0)
    ra.configure()


def test_amr_uniform_tree_equals_uniform_level(gpu):
    """On a fully refined tree the oct path and the resident uniform level
    give the same bits."""
    import ramses_amd as ra
    from ramses_amd.amr import AmrLevels
    from ramses_amd.tree import OctTree
    from helpers import random_grid
    rng = np.random.default_rng(4)
    n = 16
    ug = random_grid(rng, n)
    tree = OctTree(4, 4)

    def fn(x, l):
        if l < 4:
            return np.ones((5, len(x)))
        c = np.floor(x * n).astype(int)
        return ug[:, c[:, 2], c[:, 1], c[:, 0]]
    u = tree.fill(fn)
    ra.configure(gamma=5 / 3, riemann="hllc", slope_type=1, courant_factor=0.5)
    g = AmrLevels(tree, u)
    dt = float(g.courant(4))
    g.set_unew(4)
    g.godunov(4)
    lvl = ra.UniformLevel(ug)
    dtu = float(lvl.step())
    assert dt == dtu
    ic, x = tree.cells_of_level(4)
    c = np.floor(x * n).astype(int)
    out = g.state("new")
    assert np.array_equal(out[:, ic - 1], lvl.u.cpu().numpy()[:, c[:, 2], c[:, 1], c[:, 0]])


@needs_ref
def test_amr_subcycled_steps_vs_reference(gpu):
    """Three coarse steps of the subcycled multi-level step (levels 3-5,
    4 fine steps per coarse step) against the reference's own routines
    driven in amr_step's order: every cell and every level dt, bitwise."""
    import ramses_amd as ra
    from amr_cases import reference_coarse_step
    from ramses_amd.amr import AmrSolver, configure_interpolation
    R = oracle.Reference()
    A = oracle.ReferenceAMR(R)
    tree = sphere_tree()
    u = smooth_state(tree, seed=3)
    R.set_params(gamma=5 / 3, riemann="hllc", slope_type=1, courant_factor=0.5)
    ra.configure(gamma=5 / 3, riemann="hllc", slope_type=1, courant_factor=0.5)
    A.install(tree)
    A.set_interpol(1, 0)
    configure_interpolation(1, 0)
    A.put(0, u)
    A.put(1, u)
    s = AmrSolver(tree, u)
    dt = [0.0] * (tree.nlevel + 2)
    dtold = [0.0] * (tree.nlevel + 2)
    for k in range(3):
        reference_coarse_step(A, tree, 3, dt, dtold)
        s.coarse_step()
        s.check()
        assert np.array_equal(s.state("old"), A.get(0)), k
        got = s.dt.cpu().numpy()
        for l in (3, 4, 5):
            assert got[l] == dt[l], (k, l)
    ra.configure()
