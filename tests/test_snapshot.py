"""Snapshot files (output_NNNNN/): the writer against the reference's own
dump_all writers (output_info, backup_amr, backup_hydro, compiled in
oracle/_ref) on the same defragmented mesh, byte for byte."""
import os
import shutil
import tempfile

import numpy as np
import pytest

import oracle
from amr_cases import reference_driver_step, smooth_state, sphere_tree
from ramses_amd import snapshot

needs_ref = pytest.mark.needs_ref


def test_fortran_e_format():
    """Ew.d as the Fortran standard defines it (0.d1..dd E+xx)."""
    assert snapshot.fortran_e(0.5) == "  0.500000000000000E+00"
    assert snapshot.fortran_e(1.0) == "  0.100000000000000E+01"
    assert snapshot.fortran_e(0.0) == "  0.000000000000000E+00"
    assert snapshot.fortran_e(-2.5e-7) == " -0.250000000000000E-06"
    assert snapshot.fortran_e(2.0 ** 21) == "  0.209715200000000E+07"


def reference_mesh(A, L):
    son, father, nbor, flag1 = A.tree_arrays()
    return snapshot.Mesh(A.tree.ngridmax, L, {l: A.grids(l) for l in range(1, L + 1)}, son, father, nbor, flag1,
                         A.get_xg(), A.get(0), A.free_list(), A.used_mem())


@needs_ref
def test_snapshot_matches_reference_writers():
    """A refined, stepped mesh (levels 3-6, gradient refinement) written by
    the reference's dump_all writers and by snapshot.write_output after the
    reference's defrag: info, amr, hydro and descriptor files identical."""
    R = oracle.Reference()
    A = oracle.ReferenceAMR(R)
    tree = sphere_tree(extra_grids=20000)
    u = smooth_state(tree, seed=3)
    R.set_params(gamma=5 / 3, riemann="hllc", slope_type=1, courant_factor=0.5)
    L = 6
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
    dt, dtold, clock = [0.0] * (L + 2), [0.0] * (L + 2), {}
    reference_driver_step(A, dt, dtold, clock=clock)
    base = tempfile.mkdtemp(prefix="snap")         # the reference's file names hold 80 characters
    ref_dir = os.path.join(base, "ref")
    os.makedirs(ref_dir)
    A.dump(ref_dir, clock["t"], clock["nstep"], 1)  # defrag + dump_all's writers
    mesh = reference_mesh(A, L)                    # the defragmented mesh
    assert np.array_equal(mesh.lists[1], [1])
    run = snapshot.RunInfo(levelmin=3, gamma=5 / 3, t=clock["t"], nstep=clock["nstep"], nstep_coarse=1,
                           dtold=tuple(dtold[1:L + 1]),
                           dtnew=tuple(dt[1:L + 1]), ordering="planar", bound_key=(-1.0, 1e30))
    out = snapshot.write_output(mesh, run, os.path.join(base, "ours"))
    assert run.ifout == 2
    try:
        for name in ("info_00001.txt", "amr_00001.out00001", "hydro_00001.out00001", "hydro_file_descriptor.txt"):
            a = open(os.path.join(out, name), "rb").read()
            b = open(os.path.join(ref_dir, name), "rb").read()
            assert a == b, name
    finally:
        shutil.rmtree(base)
