"""Restart from snapshot files (output_NNNNN/): snapshot.read_output
against the reference's own restart (init_amr + init_hydro, compiled in
oracle/_ref) on the same files, bit for bit, and a restarted GPU run
against the reference continuing from the same files."""
import os
import shutil
import tempfile

import numpy as np
import pytest

import oracle
from amr_cases import reference_driver_step, smooth_state, sphere_tree
from ramses_amd import snapshot

needs_ref = pytest.mark.needs_ref

L = 6
GAMMA = 5 / 3


def _reference_snapshot(base, seed=3):
    """A refined, stepped mesh (levels 3-6) written by the reference's
    dump_all into base/output_00001; returns the reference driver and the
    level dts it recorded."""
    R = oracle.Reference()
    A = oracle.ReferenceAMR(R)
    tree = sphere_tree(extra_grids=20000)
    u = smooth_state(tree, seed=seed)
    R.set_params(gamma=GAMMA, riemann="hllc", slope_type=1, courant_factor=0.5)
    A.install(tree, nlevelmax=L)
    A.set_interpol(1, 0)
    A.put(0, u)
    A.put(1, u)
    A.set_refine(3, err_grad_d=0.12, err_grad_p=0.25)
    xg = np.zeros((tree.ngridmax, 3))
    for (l, gx, gy, gz), i in tree.grid_at.items():
        xg[i - 1] = (np.array([gx, gy, gz]) + 0.5) * 0.5 ** (l - 1)
    A.set_xg(xg)
    for l in range(L, 0, -1):
        if len(A.grids(l)):
            A.flag_fine(l, 2)
    A.flag_coarse()
    clock = {}
    reference_driver_step(A, clock=clock)
    d = os.path.join(base, "output_00001")
    os.makedirs(d)
    A.dump(d, clock["t"], clock["nstep"], 1)
    return R, A, tree


def _reference_restart_state(A, ngridmax):
    son, father, nbor, flag1 = A.tree_arrays()
    lists = {l: A.grids(l) for l in range(1, L + 1)}
    return dict(son=son, father=father, nbor=nbor, flag1=flag1, lists=lists, u=A.get(0), xg=A.get_xg(),
                free=A.free_list(), run=A.run_state())


@needs_ref
@pytest.mark.parametrize("extra", [0, 7001])
def test_read_output_matches_reference_restart(extra):
    """The mesh, free list, refinement map, conserved state, times and
    level dts that read_output rebuilds equal the reference's restart of
    the same files; extra > 0 restarts with a larger ngridmax, which
    renumbers every father and neighbour cell index."""
    base = tempfile.mkdtemp(prefix="rst")         # the reference's file names hold 80 characters
    try:
        R, A, tree = _reference_snapshot(base)
        ngm = tree.ngridmax + extra
        mesh, run = snapshot.read_output(base, 1, ngridmax=ngm, nlevelmax=L, smallr=1e-10, gamma=GAMMA,
                                         ordering="planar")
        A.restart(base, 1, L, ngm)
        ref = _reference_restart_state(A, ngm)
        assert np.array_equal(mesh.son, ref["son"])
        assert np.array_equal(mesh.father, ref["father"])
        assert np.array_equal(mesh.nbor, ref["nbor"])
        assert np.array_equal(mesh.flag1, ref["flag1"])
        for l in range(1, L + 1):
            assert np.array_equal(mesh.lists.get(l, np.zeros(0)), ref["lists"][l]), l
        assert np.array_equal(mesh.xg, ref["xg"])
        assert np.array_equal(mesh.free, ref["free"])
        assert np.array_equal(mesh.uold, ref["u"])
        t, nstep, nstep_coarse, ifout, iout, used_mem, dto, dtn = ref["run"]
        assert (run.t, run.nstep, run.nstep_coarse, run.ifout, run.iout, mesh.used_mem) == \
            (t, nstep, nstep_coarse, ifout, iout, used_mem)
        assert np.array_equal(np.array(run.dtold), dto) and np.array_equal(np.array(run.dtnew), dtn)
        assert len(mesh.lists[6]) > 0 and len(mesh.free) == ngm - sum(len(g) for g in mesh.lists.values())
    finally:
        shutil.rmtree(base)


@needs_ref
def test_read_output_round_trip_and_errors():
    """Writing a restarted mesh again gives the same amr file (the hydro
    file's primitives pass through a conserved -> primitive round trip, so
    only its layout is compared), and incompatible files raise."""
    base = tempfile.mkdtemp(prefix="rst")
    try:
        _reference_snapshot(base)
        mesh, run = snapshot.read_output(base, 1, smallr=1e-10, gamma=GAMMA, ordering="planar")
        run.levelmin = 3
        run.ifout = 1
        out = snapshot.write_output(mesh, run, os.path.join(base, "again"))
        a = open(os.path.join(out, "amr_00001.out00001"), "rb").read()
        b = open(os.path.join(base, "output_00001", "amr_00001.out00001"), "rb").read()
        assert a == b
        ha = snapshot._records(os.path.join(out, "hydro_00001.out00001"))
        hb = snapshot._records(os.path.join(base, "output_00001", "hydro_00001.out00001"))
        assert [len(x) for x in ha] == [len(x) for x in hb]
        with pytest.raises(ValueError, match="not found"):
            snapshot.read_output(base, 2)
        with pytest.raises(ValueError, match="not compatible"):
            snapshot.read_output(base, 1, nlevelmax=L - 1)
        with pytest.raises(ValueError, match="not compatible"):
            snapshot.read_output(base, 1, ngridmax=10)
        with pytest.raises(ValueError, match="Ordering"):
            snapshot.read_output(base, 1, ordering="hilbert")
    finally:
        shutil.rmtree(base)


@pytest.mark.gpu
@needs_ref
def test_restarted_run_matches_reference(gpu):
    """AmrSolver.from_snapshot and the reference's restart of the same
    files, then two adaptive coarse steps on each: tree, grid lists, free
    list, refinement map, state and level dts equal after every step."""
    import ramses_amd as ra
    from ramses_amd.amr import AmrSolver, configure_interpolation, configure_refinement
    base = tempfile.mkdtemp(prefix="rst")
    try:
        R, A, tree = _reference_snapshot(base)
        ngm = tree.ngridmax + 3000
        ra.configure(gamma=GAMMA, riemann="hllc", slope_type=1, courant_factor=0.5)
        configure_interpolation(1, 0)
        configure_refinement(err_grad_d=0.12, err_grad_p=0.25)
        s, run = AmrSolver.from_snapshot(base, 1, 3, nlevelmax=L, ngridmax=ngm, gamma=GAMMA, ordering="planar",
                                         refine=True)
        A.restart(base, 1, L, ngm)
        t, nstep, nstep_coarse, _, _, _, dto, dtn = A.run_state()
        assert float(s.time.item()) == t and (s.fine_steps, s.nstep_coarse) == (nstep, nstep_coarse)
        dt, dtold = [0.0] + list(dtn) + [0.0], [0.0] + list(dto) + [0.0]
        for k in range(2):
            reference_driver_step(A, dt, dtold)
            s.coarse_step()
            s.check()
            rs, rf, rn, rfl = A.tree_arrays()
            gs, gf, gn, gfl = s.tree_arrays()
            assert np.array_equal(gs, rs) and np.array_equal(gf, rf) and np.array_equal(gn, rn), k
            assert np.array_equal(gfl, rfl), k
            for l in range(1, L + 1):
                assert np.array_equal(s.grids(l), A.grids(l)), (k, l)
            assert np.array_equal(s.free_list(), A.free_list()), k
            assert np.array_equal(s.state("old"), A.get(0)), k
            got = s.dt.cpu().numpy()
            for l in range(3, L + 1):
                if s.noct(l):
                    assert got[l] == dt[l], (k, l)
        assert s.nstep_coarse == nstep_coarse + 2
    finally:
        shutil.rmtree(base)
        ra.configure()
        configure_refinement()


@pytest.mark.gpu
@needs_ref
def test_namelist_restart_matches_reference(gpu):
    """A parameter file run (two coarse steps, output_00001), then the same
    file with nrestart=1 and nstepmax=4 continuing it through
    ramses_amd.run; the reference restarts from the same output_00001 and
    takes the same two coarse steps: output_00002's four files are
    byte-identical."""
    import ramses_amd as ra
    from ramses_amd import namelist, run
    from test_gpu_amr import SMALL_SEDOV_NML
    base = tempfile.mkdtemp(prefix="nml")
    try:
        path = os.path.join(base, "p.nml")
        with open(path, "w") as f:
            f.write(SMALL_SEDOV_NML)
        o = os.path.join(base, "o")
        run.run(path, outdir=o, log=lambda *a: None, ordering="planar", bound_key=(-1.0, 1e30))
        path2 = os.path.join(base, "p2.nml")
        with open(path2, "w") as f:
            f.write(SMALL_SEDOV_NML.replace("nstepmax=2", "nstepmax=4\nnrestart=1"))
        s, outs = run.run(path2, outdir=o, log=lambda *a: None)
        assert s.nstep_coarse == 4 and outs[-1].endswith("output_00002")
        p = namelist.Params.from_namelist(path2)
        L, lmin = p.levelmax, p.levelmin
        R = oracle.Reference()
        A = oracle.ReferenceAMR(R)
        R.set_params(gamma=p.gamma, riemann=p.riemann, slope_type=p.slope_type, courant_factor=p.courant_factor)
        A.restart(o, 1, L, p.ngridmax)
        A.set_interpol(p.interpol_type, p.interpol_var)
        A.set_refine(lmin, err_grad_d=p.err_grad_d, err_grad_p=p.err_grad_p)
        A.set_nsubcycle(p.nsubcycle[:10])
        A.set_outputs(p.tout)
        t, nstep, nstep_coarse, ifout, _, _, dto, dtn = A.run_state()
        assert (ifout, nstep_coarse) == (2, 2)
        clock = {}
        for _ in range(2):
            reference_driver_step(A, clock=clock)
        os.makedirs(os.path.join(base, "ref"))
        A.dump(os.path.join(base, "ref"), clock["t"], clock["nstep"], 4)
        for name in ("info_00002.txt", "amr_00002.out00001", "hydro_00002.out00001", "hydro_file_descriptor.txt"):
            assert open(os.path.join(outs[-1], name), "rb").read() == \
                open(os.path.join(base, "ref", name), "rb").read(), name
        A.set_nsubcycle([2] * 10)
        A.set_outputs([])
    finally:
        shutil.rmtree(base)
        ra.configure()
