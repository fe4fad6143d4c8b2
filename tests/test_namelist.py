"""Parameter files: the namelist reader on the reference's own
namelist/sedov3d.nml (kept as an input fixture under tests/golden/), and the
region initial conditions against the reference's compiled condinit."""
import os

import numpy as np
import pytest

import oracle
from ramses_amd import namelist

HERE = os.path.dirname(__file__)
needs_ref = pytest.mark.needs_ref


def test_parse_sedov3d():
    g = namelist.read(os.path.join(HERE, "golden", "sedov3d.nml"))
    assert g["run_params"]["nsubcycle"] == [1] * 10
    assert g["amr_params"]["levelmin"] == 8 and g["amr_params"]["boxlen"] == 0.5
    assert g["init_params"]["region_type"] == ["square", "point"]
    assert g["init_params"]["p_region"] == [1e-5, 0.4]
    p = namelist.Params.from_namelist(os.path.join(HERE, "golden", "sedov3d.nml"))
    assert (p.levelmin, p.levelmax, p.ngridmax, p.boxlen) == (8, 8, 3000000, 0.5)
    assert p.nsubcycle[:10] == [1] * 10 and p.nsubcycle[10] == 2
    assert (p.gamma, p.courant_factor, p.slope_type, p.riemann) == (1.4, 0.8, 1, "llf")
    assert [r.type for r in p.regions] == ["square", "point"]
    assert p.regions[0].length == (10.0, 10.0, 10.0) and p.regions[1].center == (0.0, 0.0, 0.0)
    assert p.regions[1].p == 0.4 and p.tout == [1.0]


def test_parse_forms():
    g = namelist.parse("""
&HYDRO_PARAMS
 gamma=1.6666667d0, riemann='hllc'  ! comment, with = sign
 slope_type=2
/
&REFINE_PARAMS
err_grad_d=0.1
err_grad_p=0.1 interpol_type=0
/
&AMR_PARAMS
nexpand=3*2,1
/
""")
    assert g["hydro_params"] == {"gamma": 1.6666667, "riemann": "hllc", "slope_type": 2}
    assert g["refine_params"] == {"err_grad_d": 0.1, "err_grad_p": 0.1, "interpol_type": 0}
    assert g["amr_params"]["nexpand"] == [2, 2, 2, 1]
    with pytest.raises(ValueError):
        namelist.Params.from_namelist(namelist.parse("&HYDRO_PARAMS\nscheme='plmde'\n/\n"))


@needs_ref
@pytest.mark.parametrize("case", ["sedov", "mixed"])
def test_region_condinit_vs_reference(case):
    """region_condinit + condinit at random cell centres: bitwise with the
    reference for max-norm squares and points; exp-norm squares (a pow()
    in both) agree wherever the inside/outside decision does."""
    import torch
    from ramses_amd.ics import region_condinit
    R = oracle.Reference()
    A = oracle.ReferenceAMR(R)
    R.set_params(gamma=1.4, riemann="hllc", slope_type=1, courant_factor=0.8)
    if case == "sedov":
        regs = namelist.Params.from_namelist(os.path.join(HERE, "golden", "sedov3d.nml")).regions
    else:
        R_ = namelist.Region
        regs = [R_("square", (0.5, 0.5, 0.5), (0.8, 0.6, 1.0), 2.0, 2.0, (0.3, -0.2, 0.1), 3.0),
                R_("square", (0.25, 0.7, 0.5), (0.3, 0.3, 0.3), 10.0, 0.5, (0.0, 1.0, 0.0), 0.2),
                R_("point", (0.5, 0.5, 0.5), (1.0, 1.0, 1.0), 10.0, 1.5, (0.5, 0.0, -0.5), 0.7)]
    A.install(__import__("ramses_amd.tree", fromlist=["OctTree"]).OctTree(1, 1), nlevelmax=1)
    A.set_regions(regs)
    rng = np.random.default_rng(5)
    for l in (3, 5):
        dx = 0.5 ** l
        idx = rng.integers(0, 2 ** l, size=(300, 3))
        x = (idx + 0.5) * dx
        want = A.condinit(x, dx)
        got = region_condinit(torch.as_tensor(x), dx, regs, gamma=1.4).numpy()
        assert np.array_equal(got, want), (case, l)
