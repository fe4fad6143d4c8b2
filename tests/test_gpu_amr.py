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
needs_ref = pytest.mark.needs_ref


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
@pytest.mark.parametrize("riemann,slope", [("hllc", 1), ("llf", 2), ("hll", 8), ("hllc", 3), ("llf", 7),
                                           ("acoustic", 1), ("exact", 2)])
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
        if riemann == "exact":  # pow in the rarefaction branches: last-bit differences allowed
            np.testing.assert_allclose(out, ref, rtol=1e-12, atol=1e-300)
        else:
            assert np.array_equal(out, ref), (riemann, slope, l)
    ra.configure()


@needs_ref
@pytest.mark.parametrize("which,riemann,slope", [("dropin", "llf", 2), ("dropin", "hllc", 1),
                                                 ("patched", "hllc", 1)])
def test_reference_godunov_through_library_unsplit(gpu, which, riemann, slope):
    """The drop-in end to end: the reference's own godunov_fine
    (hydro/godunov_fine.f90), built without umuscl.o so that godfine1's
    `call unsplit` (:681) reaches this library -- through its link-level
    symbol unsplit_ (dropin), or through the integration guide's
    ISO_C_BINDING patch file forwarding to ramses_unsplit (patched, compiled
    from the guide by oracle/Makefile) -- gives the all-reference
    build's unew bit for bit on levels 3-5 of the sphere tree, coarse/fine
    buffer octs included.  The patch configures the library once per
    process from the reference's module globals, so it runs one case."""
    import ramses_amd as ra
    path = os.path.join(os.path.dirname(oracle.REF_LIB), f"libramses_ref_{which}.so")
    assert os.path.exists(path), f"{path} is missing (make -C oracle dropin)"
    ra.load_library()  # the HIP library (and torch's runtime) before the drop-in build that needs it
    # the host-memory unsplit the drop-ins reach must work here first: on a
    # failure the link-level symbol aborts and the patch stops the process
    ra.configure(gamma=5 / 3, riemann=riemann, slope_type=slope, courant_factor=0.5)
    uin = np.zeros(ra.stencil_shape(32))
    uin[0] = 1.0
    uin[4] = 1.0
    ra.unsplit(uin, None, np.zeros(ra.flux_shape(32)), np.zeros(ra.tmp_shape(32)), 0.1, 0.1, 0.1, 1e-3, 32)
    tree = sphere_tree()
    u = smooth_state(tree, seed=11)
    out = {}
    for name, lib in (("reference", oracle.REF_LIB), (which, path)):
        R = oracle.Reference(lib)
        A = oracle.ReferenceAMR(R)
        A.install(tree)
        A.set_interpol(1, 0)
        R.set_params(gamma=5 / 3, riemann=riemann, slope_type=slope, courant_factor=0.5)
        ra.configure(gamma=5 / 3, riemann=riemann, slope_type=slope, courant_factor=0.5)
        res = []
        for l in (3, 4, 5):
            A.put(0, u)
            A.put(1, u)
            dt = A.courant(l, 1e10)[0]
            if l > 3:
                A.set_unew(l - 1)
            A.set_unew(l)
            A.godunov(l, dt)
            res.append((dt, A.get(1)))
        out[name] = res
    for l, (a, b) in zip((3, 4, 5), zip(out["reference"], out[which])):
        assert a[0] == b[0], (which, l)
        assert np.array_equal(a[1], b[1]), (which, l)
    assert not np.array_equal(out["reference"][0][1], u)
    ra.configure()


@needs_ref
@pytest.mark.parametrize("itype,ivar", [(0, 0), (2, 0), (3, 0), (1, 1), (2, 2), (3, 2), (4, 2)])
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
@pytest.mark.parametrize("graph", [False, True])
def test_amr_subcycled_steps_vs_reference(gpu, graph):
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
    s = AmrSolver(tree, u, graph=graph)
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


def _refine_setup(egd, egp, nlevelmax=6, seed=3, extra_grids=20000, **solver_kw):
    import ramses_amd as ra
    from ramses_amd.amr import AmrSolver, configure_interpolation, configure_refinement
    R = oracle.Reference()
    A = oracle.ReferenceAMR(R)
    tree = sphere_tree(extra_grids=extra_grids)
    u = smooth_state(tree, seed=seed)
    R.set_params(gamma=5 / 3, riemann="hllc", slope_type=1, courant_factor=0.5)
    ra.configure(gamma=5 / 3, riemann="hllc", slope_type=1, courant_factor=0.5)
    A.install(tree, nlevelmax=nlevelmax)
    A.set_interpol(1, 0)
    A.put(0, u)
    A.put(1, u)
    A.set_refine(3, err_grad_d=egd, err_grad_p=egp)
    configure_interpolation(1, 0)
    configure_refinement(err_grad_d=egd, err_grad_p=egp)
    s = AmrSolver(tree, u, refine=True, nlevelmax=nlevelmax, **solver_kw)
    return A, s, tree


def _same_tree(A, s, nlevelmax, tag):
    rs, rf, rn, rfl = A.tree_arrays()
    gs, gf, gn, gfl = s.tree_arrays()
    assert np.array_equal(gs, rs), tag
    assert np.array_equal(gf, rf), tag
    assert np.array_equal(gn, rn), tag
    for l in range(1, nlevelmax + 1):
        assert np.array_equal(s.grids(l), A.grids(l)), (tag, l)
    assert np.array_equal(s.free_list(), A.free_list()), tag
    return gfl, rfl


@needs_ref
@pytest.mark.parametrize("egd,egp", [(0.12, 0.25), (0.35, 0.8)])
def test_flag_and_refine_vs_reference(gpu, egd, egp):
    """flag_fine on every level (icount 2 and 1), then refine_fine of
    levels 3..5: flags, son/father/nbor, grid lists, free list and the
    prolongated / cleared state, all bitwise."""
    import ramses_amd as ra
    A, s, tree = _refine_setup(egd, egp)
    for icount in (2, 1):
        for l in range(6, 0, -1):
            if s.noct(l):
                A.flag_fine(l, icount)
                s.flag(l, icount, 3)
        gfl, rfl = _same_tree(A, s, 6, f"flags icount={icount}")
        assert np.array_equal(gfl, rfl), icount
    made = killed = 0
    for l in range(3, 6):
        before = {k: s.noct(k) for k in range(1, 7)}
        A.refine_fine(l)
        s.refine(l)
        s.check()
        gfl, rfl = _same_tree(A, s, 6, f"refine {l}")
        assert np.array_equal(gfl, rfl), l
        assert np.array_equal(s.state("old"), A.get(0)), l
        assert np.array_equal(s.state("new"), A.get(1)), l
        made += max(0, s.noct(l + 1) - before[l + 1])
        killed += max(0, before[l + 1] - s.noct(l + 1))
    assert made + killed > 0
    ra.configure()


@needs_ref
@pytest.mark.parametrize("graph,face_octs", [(False, 0), (True, 0), (False, -1), (True, -1)])
def test_amr_adaptive_steps_vs_reference(gpu, graph, face_octs):
    """Three coarse steps of the adaptive multi-level step (refine at the
    start of the levelmin step and of every second substep, flag after
    every level step) against the reference's own coarse step
    (amr/amr_step.f90 under adaptive_loop's body): tree, grid lists, free
    list, state and dts, bitwise.
    face_octs 0 keeps the default (these small levels run the level unsplit
    one face per thread); -1 runs it one row of faces per thread."""
    import ramses_amd as ra
    lib = ra.load_library()
    assert lib.ramses_tune(5, face_octs) == 0
    try:
        _adaptive_steps(graph)
    finally:
        lib.ramses_tune(5, 0)
    ra.configure()


def _adaptive_steps(graph):
    from amr_cases import reference_driver_step
    A, s, tree = _refine_setup(0.12, 0.25)
    s.graph = graph
    for l in range(6, 0, -1):
        if s.noct(l):
            A.flag_fine(l, 2)
    A.flag_coarse()
    s.flag_all()
    dt = [0.0] * 8
    dtold = [0.0] * 8
    counts = []
    for k in range(3):
        reference_driver_step(A, dt, dtold)
        s.coarse_step()
        s.check()
        counts.append({l: s.noct(l) for l in range(3, 7)})
        gfl, rfl = _same_tree(A, s, 6, f"step {k}")
        assert np.array_equal(gfl, rfl), k
        assert np.array_equal(s.state("old"), A.get(0)), k
        got = s.dt.cpu().numpy()
        for l in range(3, 7):
            if s.noct(l):
                assert got[l] == dt[l], (k, l)
    assert counts[0] != counts[-1] or counts[0][6] > 0


def test_adaptive_mesh_golden(gpu):
    """flag_fine on every level, refine_fine of levels 3-5 and two adaptive
    coarse steps against the committed digests of the reference's mesh
    (tree, flags, grid lists, free list, uold, unew) and its level dts."""
    import ramses_amd as ra
    from amr_cases import REFINE_CASE as C, mesh_digests, smooth_state, sphere_tree
    from ramses_amd.amr import AmrSolver, configure_interpolation, configure_refinement
    with open(os.path.join(os.path.dirname(__file__), "golden", "amr_refine.json")) as f:
        gold = json.load(f)
    assert gold["case"] == C
    tree = sphere_tree(C["levelmin"], C["levelmax"], extra_grids=C["extra_grids"])
    u = smooth_state(tree, seed=C["seed"])
    ra.configure(gamma=C["gamma"], riemann=C["riemann"], slope_type=C["slope_type"],
                 courant_factor=C["courant_factor"])
    configure_interpolation(1, 0)
    configure_refinement(err_grad_d=C["err_grad_d"], err_grad_p=C["err_grad_p"])
    L = C["nlevelmax"]
    s = AmrSolver(tree, u, refine=True, nlevelmax=L)

    def snap():
        son, father, nbor, flag1 = s.tree_arrays()
        return mesh_digests(son, father, nbor, flag1, {l: s.grids(l) for l in range(1, L + 1)}, s.free_list(),
                            s.state("old"), s.state("new"))
    s.flag_all()
    assert snap() == gold["stages"]["flagged"]
    for l in range(C["levelmin"], L):
        s.refine(l)
    assert snap() == gold["stages"]["refined"]
    for k in range(2):
        s.coarse_step()
        s.check()
        assert snap() == gold["stages"][f"step{k + 1}"], k
        got = s.dt.cpu().numpy()[:L + 2]
        assert [float(x).hex() for x in got] == gold["dts"][k], k
    ra.configure()


@needs_ref
@pytest.mark.parametrize("graph", [False, True])
def test_sedov_amr_init_and_steps_vs_reference(gpu, graph):
    """The adaptive Sedov set-up (namelist/sedov3d.nml regions, gamma 1.4,
    hllc, minmod, CFL 0.8, err_grad_d = err_grad_p = 0.1, interpol_type 1)
    at levelmin 3, levelmax 6: the initial mesh built by init_refine /
    init_refine_2 on the GPU against the reference's own init_flow /
    flag_fine / refine_fine, then three adaptive coarse steps of the
    reference's own driver -- mesh, grid centres, state and dts bitwise."""
    import ramses_amd as ra
    from amr_cases import mesh_digests, reference_driver_step, reference_init_refine
    from ramses_amd.amr import AmrSolver, configure_interpolation, configure_refinement
    from ramses_amd.ics import sedov_regions
    from ramses_amd.tree import OctTree
    lmin, L = 3, 6
    tree = OctTree(lmin, lmin, extra_grids=20000)
    R = oracle.Reference()
    A = oracle.ReferenceAMR(R)
    R.set_params(gamma=1.4, riemann="hllc", slope_type=1, courant_factor=0.8)
    ra.configure(gamma=1.4, riemann="hllc", slope_type=1, courant_factor=0.8)
    A.install(tree, nlevelmax=L)
    A.set_interpol(1, 0)
    A.set_refine(lmin, err_grad_d=0.1, err_grad_p=0.1)
    configure_interpolation(1, 0)
    configure_refinement(err_grad_d=0.1, err_grad_p=0.1)
    s = AmrSolver(tree, None, refine=True, nlevelmax=L, graph=graph)
    A.set_xg(s.xg[1:].cpu().numpy())
    A.set_sedov(1.0, 1e-5, 0.4, (0.5, 0.5, 0.5))
    reference_init_refine(A, lmin, L)
    s.init_refine(lambda x, dx: sedov_regions(x, dx, gamma=1.4))

    def snap(src):
        son, father, nbor, flag1 = src.tree_arrays()
        if src is s:
            return mesh_digests(son, father, nbor, flag1, {l: s.grids(l) for l in range(1, L + 1)}, s.free_list(),
                                s.state("old"), s.state("new"))
        return mesh_digests(son, father, nbor, flag1, {l: A.grids(l) for l in range(1, L + 1)}, A.free_list(),
                            A.get(0), A.get(1))
    assert np.array_equal(s.xg[1:].cpu().numpy(), A.get_xg())
    go, ro = s.state("old"), A.get(0)
    bad = np.nonzero((go != ro).any(0))[0]
    if len(bad):
        ngm = tree.ngridmax
        lev = s.xg.new_zeros(1)  # noqa: F841
        grid = (bad - 1) % ngm + 1
        levels = {l: int(np.isin(grid, s.grids(l)).sum()) for l in range(1, L + 1)}
        k = bad[:4]
        raise AssertionError(f"uold differs in {len(bad)} cells, per level {levels}; "
                             f"cells {k.tolist()}: gpu {go[:, k].T.tolist()} ref {ro[:, k].T.tolist()}")
    assert snap(s) == snap(A)
    assert s.noct(L) > 0
    A.flag_coarse()   # init_refine_2's last pass ends with it (amr/init_refine.f90:121)
    dt, dtold = [0.0] * (L + 2), [0.0] * (L + 2)
    for k in range(3):
        reference_driver_step(A, dt, dtold)
        s.coarse_step()
        s.check()
        assert snap(s) == snap(A), k
        got = s.dt.cpu().numpy()
        for l in range(lmin, L + 1):
            if s.noct(l):
                assert got[l] == dt[l], (k, l)
    assert np.array_equal(s.xg[1:].cpu().numpy(), A.get_xg())
    ra.configure()


@needs_ref
def test_sedov_amr_levels_6_10_vs_reference(gpu):
    """The adaptive Sedov workload at its own size: levelmin 6, levelmax 10,
    err_grad_d = err_grad_p = 0.1, interpol_type 1, gamma 1.4, hllc, minmod,
    CFL 0.8, nsubcycle 2 -- init_refine, then four coarse steps (64 finest
    steps) in direct and captured mode against the reference's own
    init_refine sequence and coarse step: tree, grid lists, free list,
    refinement map, state, grid centres, level dts, t and nstep, bitwise
    after every step."""
    import ramses_amd as ra
    from amr_cases import mesh_digests, reference_driver_step, reference_init_refine
    from ramses_amd.amr import AmrSolver, configure_interpolation, configure_refinement
    from ramses_amd.ics import sedov_regions
    from ramses_amd.tree import OctTree
    lmin, L, ngm = 6, 10, 120_000
    tree = OctTree(lmin, lmin, extra_grids=ngm - sum(8 ** k for k in range(lmin)))
    ra.configure(gamma=1.4, riemann="hllc", slope_type=1, courant_factor=0.8)
    configure_interpolation(1, 0)
    configure_refinement(err_grad_d=0.1, err_grad_p=0.1)

    def snap_native(s):
        son, father, nbor, flag1 = s.tree_arrays()
        return mesh_digests(son, father, nbor, flag1, {l: s.grids(l) for l in range(1, L + 1)}, s.free_list(),
                            s.state("old"), s.state("new"))

    native = {}
    for graph in (False, True):
        s = AmrSolver(tree, None, refine=True, nlevelmax=L, graph=graph)
        s.init_refine(lambda x, dx: sedov_regions(x, dx, gamma=1.4))
        seq = [(snap_native(s), s.xg[1:].cpu().numpy().copy())]
        for _ in range(4):
            s.coarse_step()
            s.check()
            seq.append((snap_native(s), None, s.dt.cpu().numpy()[:L + 1].copy(), float(s.time.item()),
                        s.fine_steps))
        assert s.fine_steps == 64 and s.noct(L) > 0
        native[graph] = seq
        del s
    R = oracle.Reference()
    A = oracle.ReferenceAMR(R)
    R.set_params(gamma=1.4, riemann="hllc", slope_type=1, courant_factor=0.8)
    A.install(tree, nlevelmax=L)
    A.set_interpol(1, 0)
    A.set_refine(lmin, err_grad_d=0.1, err_grad_p=0.1)
    A.set_xg(native[False][0][1])
    A.set_sedov(1.0, 1e-5, 0.4, (0.5, 0.5, 0.5))
    reference_init_refine(A, lmin, L)

    def snap_ref():
        son, father, nbor, flag1 = A.tree_arrays()
        return mesh_digests(son, father, nbor, flag1, {l: A.grids(l) for l in range(1, L + 1)}, A.free_list(),
                            A.get(0), A.get(1))
    for graph in (False, True):
        assert native[graph][0][0] == snap_ref(), ("init", graph)
    A.flag_coarse()
    dt, dtold, clock = [0.0] * (L + 2), [0.0] * (L + 2), {}
    for k in range(1, 5):
        reference_driver_step(A, dt, dtold, clock=clock)
        ref = snap_ref()
        for graph in (False, True):
            snap, _, got_dt, t, nstep = native[graph][k]
            assert snap == ref, (k, graph)
            assert (t, nstep) == (clock["t"], clock["nstep"]), (k, graph)
            for l in range(lmin, L + 1):
                assert got_dt[l] == dt[l], (k, l, graph)
    A.install(OctTree(1, 1), nlevelmax=1)  # release the reference's arrays
    configure_refinement()
    ra.configure()


@needs_ref
def test_snapshot_after_adaptive_steps_vs_reference(gpu):
    """Two adaptive coarse steps, then dump_all on both sides: the GPU
    defrag + snapshot writer give the reference's files byte for byte, and
    a further step after the defrag stays bitwise."""
    import shutil
    import tempfile
    import ramses_amd as ra
    from amr_cases import reference_driver_step
    from ramses_amd import snapshot
    A, s, tree = _refine_setup(0.12, 0.25)
    xg = s.xg[1:].cpu().numpy()
    A.set_xg(xg)
    for l in range(6, 0, -1):
        if s.noct(l):
            A.flag_fine(l, 2)
    A.flag_coarse()
    s.flag_all()
    L = 6
    clock = {}
    for _ in range(2):
        reference_driver_step(A, clock=clock)
        s.coarse_step()
    s.check()
    assert (float(s.time.item()), s.fine_steps) == (clock["t"], clock["nstep"])
    base = tempfile.mkdtemp(prefix="snap")
    try:
        os.makedirs(os.path.join(base, "ref"))
        A.dump(os.path.join(base, "ref"), clock["t"], clock["nstep"], 2)
        run = snapshot.RunInfo(levelmin=3, gamma=5 / 3, ordering="planar", bound_key=(-1.0, 1e30))
        out = s.write_output(os.path.join(base, "ours"), run)
        for name in ("info_00001.txt", "amr_00001.out00001", "hydro_00001.out00001", "hydro_file_descriptor.txt"):
            assert open(os.path.join(out, name), "rb").read() == \
                open(os.path.join(base, "ref", name), "rb").read(), name
    finally:
        shutil.rmtree(base)
    reference_driver_step(A)
    s.coarse_step()
    s.check()
    son, father, nbor, flag1 = s.tree_arrays()
    rs, rf, rn, rfl = A.tree_arrays()
    assert np.array_equal(son, rs) and np.array_equal(father, rf) and np.array_equal(nbor, rn)
    assert np.array_equal(flag1, rfl)
    for l in range(1, L + 1):
        assert np.array_equal(s.grids(l), A.grids(l)), l
    assert np.array_equal(s.free_list(), A.free_list())
    assert np.array_equal(s.state("old"), A.get(0))
    ra.configure()


@needs_ref
@pytest.mark.parametrize("graph", [False, True])
def test_sedov_amr_no_subcycling_vs_reference(gpu, graph):
    """nsubcycle=1 on every level (as namelist/sedov3d.nml sets it): one
    step per level per coarse step, the coarse dt taken from the finest
    level; init_refine and three coarse steps bitwise."""
    import ramses_amd as ra
    from amr_cases import reference_driver_step, reference_init_refine
    from ramses_amd.amr import AmrSolver, configure_interpolation, configure_refinement
    from ramses_amd.ics import sedov_regions
    from ramses_amd.tree import OctTree
    lmin, L = 3, 6
    tree = OctTree(lmin, lmin, extra_grids=20000)
    R = oracle.Reference()
    A = oracle.ReferenceAMR(R)
    R.set_params(gamma=1.4, riemann="hllc", slope_type=1, courant_factor=0.8)
    ra.configure(gamma=1.4, riemann="hllc", slope_type=1, courant_factor=0.8)
    A.install(tree, nlevelmax=L)
    A.set_interpol(1, 0)
    A.set_refine(lmin, err_grad_d=0.1, err_grad_p=0.1)
    A.set_nsubcycle([1] * 10)
    configure_interpolation(1, 0)
    configure_refinement(err_grad_d=0.1, err_grad_p=0.1)
    s = AmrSolver(tree, None, refine=True, nlevelmax=L, nsubcycle=1, graph=graph)
    A.set_xg(s.xg[1:].cpu().numpy())
    A.set_sedov(1.0, 1e-5, 0.4, (0.5, 0.5, 0.5))
    reference_init_refine(A, lmin, L)
    s.init_refine(lambda x, dx: sedov_regions(x, dx, gamma=1.4))
    A.flag_coarse()
    dt, dtold = [0.0] * (L + 2), [0.0] * (L + 2)
    for k in range(3):
        reference_driver_step(A, dt, dtold)
        s.coarse_step()
        s.check()
        son, father, nbor, flag1 = s.tree_arrays()
        rs, rf, rn, rfl = A.tree_arrays()
        assert np.array_equal(son, rs) and np.array_equal(flag1, rfl), k
        for l in range(1, L + 1):
            assert np.array_equal(s.grids(l), A.grids(l)), (k, l)
        assert np.array_equal(s.state("old"), A.get(0)), k
        got = s.dt.cpu().numpy()
        for l in range(lmin, L + 1):
            if s.noct(l):
                assert got[l] == dt[l], (k, l)
    A.set_nsubcycle([2] * 10)
    ra.configure()


SMALL_SEDOV_NML = """&RUN_PARAMS
hydro=.true.
nsubcycle=2,2,2,1,1,1
nstepmax=2
/
&AMR_PARAMS
levelmin=3
levelmax=6
ngridtot=20000
nexpand=1
boxlen=0.5
/
&INIT_PARAMS
nregion=2
region_type(1)='square'
region_type(2)='point'
x_center=0.25,0.25
y_center=0.25,0.25
z_center=0.25,0.25
length_x=10.0,1.0
length_y=10.0,1.0
length_z=10.0,1.0
exp_region=10.0,10.0
d_region=1.0,0.0
p_region=1e-5,0.4
/
&OUTPUT_PARAMS
noutput=1
tout=1.0
/
&HYDRO_PARAMS
gamma=1.4
courant_factor=0.8
riemann='hllc'
slope_type=1
/
&REFINE_PARAMS
err_grad_d=0.1
err_grad_p=0.1
interpol_type=1
/
"""


@needs_ref
@pytest.mark.parametrize("graph", [False, True])
def test_namelist_run_matches_reference(gpu, graph):
    """A parameter file run end to end (ramses_amd.run: namelist, region
    initial conditions in a box of length 0.5, init_refine, two adaptive
    coarse steps with mixed subcycling, snapshot) against the reference's
    own routines (init_refine's sequence, then the reference's own coarse
    step): the four snapshot files are byte-identical."""
    import shutil
    import tempfile
    import ramses_amd as ra
    from amr_cases import reference_driver_step, reference_init_refine
    from ramses_amd import namelist, run
    from ramses_amd.tree import OctTree
    base = tempfile.mkdtemp(prefix="nml")
    try:
        path = os.path.join(base, "p.nml")
        with open(path, "w") as f:
            f.write(SMALL_SEDOV_NML)
        s, outs = run.run(path, outdir=os.path.join(base, "o"), log=lambda *a: None, ordering="planar",
                          bound_key=(-1.0, 1e30), graph=graph)   # the reference build's one-domain ordering
        p = namelist.Params.from_namelist(path)
        L, lmin = p.levelmax, p.levelmin
        R = oracle.Reference()
        A = oracle.ReferenceAMR(R)
        R.set_params(gamma=p.gamma, riemann=p.riemann, slope_type=p.slope_type, courant_factor=p.courant_factor)
        tree = OctTree(lmin, lmin, extra_grids=p.ngridmax - sum(8 ** k for k in range(lmin)))
        A.install(tree, nlevelmax=L)
        A.set_boxlen(p.boxlen)
        A.set_interpol(p.interpol_type, p.interpol_var)
        A.set_refine(lmin, err_grad_d=p.err_grad_d, err_grad_p=p.err_grad_p)
        A.set_nsubcycle(p.nsubcycle[:10])
        A.set_outputs(p.tout)
        xg = np.zeros((tree.ngridmax, 3))
        for (l, gx, gy, gz), i in tree.grid_at.items():
            xg[i - 1] = (np.array([gx, gy, gz]) + 0.5) * 0.5 ** (l - 1)
        A.set_xg(xg)
        A.set_regions(p.regions)
        reference_init_refine(A, lmin, L)
        A.flag_coarse()
        clock = {}
        for _ in range(p.nstepmax):
            reference_driver_step(A, clock=clock)
        os.makedirs(os.path.join(base, "ref"))
        A.dump(os.path.join(base, "ref"), clock["t"], clock["nstep"], p.nstepmax)
        out = outs[-1]
        for name in ("info_00001.txt", "amr_00001.out00001", "hydro_00001.out00001", "hydro_file_descriptor.txt"):
            assert open(os.path.join(out, name), "rb").read() == \
                open(os.path.join(base, "ref", name), "rb").read(), name
        A.set_nsubcycle([2] * 10)
        A.set_outputs([])
    finally:
        shutil.rmtree(base)
        ra.configure()


@needs_ref
@pytest.mark.parametrize("itype,ivar", [(0, 0), (2, 1), (3, 2), (4, 2)])
def test_refine_interpolation_variants_vs_reference(gpu, itype, ivar):
    """make_grid_fine prolongates new octs with interpol_hydro: every
    interpolation type / variable choice gives the reference's new cells,
    and one adaptive coarse step stays bitwise."""
    import ramses_amd as ra
    from amr_cases import reference_driver_step
    from ramses_amd.amr import configure_interpolation
    A, s, tree = _refine_setup(0.12, 0.25, extra_grids=60000)
    A.set_interpol(itype, ivar)
    configure_interpolation(itype, ivar)
    for l in range(6, 0, -1):
        if s.noct(l):
            A.flag_fine(l, 2)
    A.flag_coarse()
    s.flag_all()
    for l in range(3, 6):
        A.refine_fine(l)
        s.refine(l)
    s.check()
    _same_tree(A, s, 6, "refine")
    assert np.array_equal(s.state("old"), A.get(0))
    reference_driver_step(A)
    s.coarse_step()
    s.check()
    _same_tree(A, s, 6, "step")
    assert np.array_equal(s.state("old"), A.get(0))
    A.set_interpol(1, 0)
    configure_interpolation(1, 0)
    ra.configure()


@needs_ref
def test_velocity_criterion_and_nexpand_vs_reference(gpu):
    """err_grad_u (hydro_refine's velocity test with sound-speed floor) and
    nexpand = 2 buffers: flags on every level and one refinement bitwise."""
    import ramses_amd as ra
    from ramses_amd.amr import configure_refinement
    A, s, tree = _refine_setup(-1.0, -1.0)
    A.set_refine(3, err_grad_u=0.3, nexpand=2)
    configure_refinement(err_grad_u=0.3, nexpand=2)
    for l in range(6, 0, -1):
        if s.noct(l):
            A.flag_fine(l, 2)
            s.flag(l, 2, 3)
    gfl, rfl = _same_tree(A, s, 6, "flags")
    assert np.array_equal(gfl, rfl)
    assert 0 < int(gfl.sum()) < len(gfl)
    for l in range(3, 6):
        A.refine_fine(l)
        s.refine(l)
    s.check()
    _same_tree(A, s, 6, "refine")
    assert np.array_equal(s.state("old"), A.get(0))
    configure_refinement()
    ra.configure()


def test_refine_out_of_grids_is_an_error_and_changes_nothing(gpu):
    """With too few free grids refine_fine returns the reference's "No more
    free memory" condition as an error (status -9) and leaves the mesh as it
    was: level 3 wants 57 new octs, 50 grids are free."""
    import ramses_amd as ra
    from ramses_amd.amr import AmrSolver, configure_refinement
    tree = sphere_tree(extra_grids=50)
    u = smooth_state(tree, seed=3)
    ra.configure(gamma=5 / 3, riemann="hllc", slope_type=1, courant_factor=0.5)
    configure_refinement(err_grad_d=0.12, err_grad_p=0.25)
    s = AmrSolver(tree, u, refine=True, nlevelmax=6)
    s.flag_all()
    before = (s.tree_arrays(), {l: s.grids(l).copy() for l in range(1, 7)}, s.free_list().copy(), s.state("old"))
    with pytest.raises(ra.RamsesError, match="free memory"):
        s.refine(3)
    after = (s.tree_arrays(), {l: s.grids(l) for l in range(1, 7)}, s.free_list(), s.state("old"))
    assert all(np.array_equal(a, b) for a, b in zip(before[0], after[0]))
    assert all(np.array_equal(before[1][l], after[1][l]) for l in range(1, 7))
    assert np.array_equal(before[2], after[2]) and np.array_equal(before[3], after[3])
    assert int(s.error.item()) == 0
    configure_refinement()
    ra.configure()


@pytest.mark.parametrize("graph", [False, True])
def test_adaptive_fp32_mixed_close_to_fp64(gpu, graph):
    """FP32-mixed adaptive coarse step: same refinement decisions here and
    a relative L1 difference on the leaves below 1e-5 per variable."""
    import ramses_amd as ra
    from ramses_amd.amr import AmrSolver, configure_interpolation, configure_refinement
    tree = sphere_tree(extra_grids=20000)
    u = smooth_state(tree, seed=3)
    out = {}
    for prec in ("fp64", "fp32"):
        ra.configure(gamma=5 / 3, riemann="hllc", slope_type=1, courant_factor=0.5, precision=prec)
        configure_interpolation(1, 0)
        configure_refinement(err_grad_d=0.12, err_grad_p=0.25)
        s = AmrSolver(tree, u, refine=True, nlevelmax=6, graph=graph)
        s.flag_all()
        s.coarse_step()
        s.check()
        out[prec] = (s.tree_arrays()[0], s.state("old"))
    son64, u64 = out["fp64"]
    son32, u32 = out["fp32"]
    assert np.array_equal(son64, son32)
    leaf = son64 == 0
    for v in range(5):
        a, b = u32[v][leaf], u64[v][leaf]
        assert np.abs(a - b).sum() / np.abs(b).sum() < 1e-5, v
    configure_refinement()
    ra.configure()


def test_captured_step_recaptures_after_configuration_change(gpu):
    """A captured coarse step holds the configuration it was captured with:
    after a configuration call it is captured again, so three adaptive
    steps with a change of courant_factor and of the refinement criteria
    between them give the direct mode's mesh, state and dts bit for bit."""
    import ramses_amd as ra
    from ramses_amd.amr import AmrSolver, configure_interpolation, configure_refinement
    tree = sphere_tree(extra_grids=60000)
    u = smooth_state(tree, seed=3)
    out = {}
    for graph in (False, True):
        ra.configure(gamma=5 / 3, riemann="hllc", slope_type=1, courant_factor=0.5)
        configure_interpolation(1, 0)
        configure_refinement(err_grad_d=0.12, err_grad_p=0.25)
        s = AmrSolver(tree, u, refine=True, nlevelmax=6, graph=graph)
        s.flag_all()
        s.coarse_step()
        ra.configure(gamma=5 / 3, riemann="hllc", slope_type=1, courant_factor=0.4)
        s.coarse_step()
        configure_refinement(err_grad_d=0.3, err_grad_p=0.5)
        s.coarse_step()
        s.check()
        out[graph] = (s.tree_arrays(), {l: s.grids(l) for l in range(1, 7)}, s.free_list(), s.state("old"),
                      s.dt.cpu().numpy(), float(s.time.item()), s.fine_steps, s.cell_updates, s.free.used_mem)
        del s
    a, b = out[False], out[True]
    assert all(np.array_equal(x, y) for x, y in zip(a[0], b[0]))
    assert all(np.array_equal(a[1][l], b[1][l]) for l in range(1, 7))
    for x, y in zip(a[2:], b[2:]):
        assert np.array_equal(x, y)
    configure_refinement()
    ra.configure()


def test_captured_step_errors(gpu):
    """Captured mode refuses a level larger than its workspace before it
    runs, and a step whose refinement would exhaust the free list raises
    the direct mode's error."""
    import ramses_amd as ra
    from ramses_amd.amr import AmrSolver, configure_interpolation, configure_refinement
    tree = sphere_tree(extra_grids=20000)
    u = smooth_state(tree, seed=3)
    ra.configure(gamma=5 / 3, riemann="hllc", slope_type=1, courant_factor=0.5)
    configure_interpolation(1, 0)
    configure_refinement(err_grad_d=0.12, err_grad_p=0.25)
    s = AmrSolver(tree, u, refine=True, nlevelmax=6, graph=True, graph_capacity=8)
    s.flag_all()
    with pytest.raises(ra.RamsesError, match="graph_capacity"):
        s.coarse_step()
    small = sphere_tree(extra_grids=50)
    s = AmrSolver(small, smooth_state(small, seed=3), refine=True, nlevelmax=6, graph=True)
    s.flag_all()
    with pytest.raises(ra.RamsesError, match="free memory"):
        s.coarse_step()
    configure_refinement()
    ra.configure()


@needs_ref
@pytest.mark.parametrize("graph", [False, True])
def test_coarse_step_conservation_sums_vs_reference(gpu, graph):
    """courant_fine's mass, sum E*vol and internal-energy totals of each
    adaptive coarse step (hydro/courant_fine.f90:95-119, summed over every
    courant_fine call of the step and reset per coarse step as
    amr/adaptive_loop.f90:89-91 does) agree with those of the reference's
    own coarse step to 1e-12;
    computing them leaves the mesh and state bitwise equal to the
    reference's, and the totals are the same bits on a second run.
    update_time's baselines einit and mass_tot_0, which the snapshots
    carry, are the reference's bits from the second step on."""
    import ramses_amd as ra
    from amr_cases import reference_driver_step
    runs = []
    for rep in range(2):
        A, s, tree = _refine_setup(0.12, 0.25, graph=graph, sums=True)
        A.set_xg(s.xg[1:].cpu().numpy())
        for l in range(6, 0, -1):
            if s.noct(l):
                A.flag_fine(l, 2)
        A.flag_coarse()
        s.flag_all()
        dt, dtold = [0.0] * 8, [0.0] * 8
        got = []
        for step in range(3):
            stats = {}
            reference_driver_step(A, dt, dtold, stats=stats)
            s.coarse_step()
            s.check()
            mine = s.sums.cpu().numpy()
            np.testing.assert_allclose(mine, stats["sums"], rtol=1e-12, atol=0)
            got.append(mine.copy())
            # update_time's baselines: set at the second step's first
            # update_time, summed in the reference's order -- the same bits
            b = A.baselines()
            assert np.array_equal(s.cons0.cpu().numpy(), b), step
            assert (b[0] != 0.0) == (step >= 1), step
        assert np.array_equal(s.state("old"), A.get(0))
        runs.append(got)
        del s
    assert all(np.array_equal(a, b) for a, b in zip(*runs))
    ra.configure()
