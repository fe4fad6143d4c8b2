"""CPU checks of the C ABI: the library loads, exports every function the
header declares, and validates configurations (no GPU compute calls)."""
import ctypes
import os
import re

import pytest

import ramses_amd


def test_library_exports_header():
    lib = ramses_amd.load_library()
    names = ramses_amd.header_functions()
    assert "unsplit_" in names and "ramses_unsplit" in names
    for n in names:
        assert hasattr(lib, n), n


def test_header_cites_reference():
    """Every entry point's comment block names the reference file it
    replaces, or says that it has none."""
    text = open(ramses_amd.HEADER).read()
    for n in ramses_amd.header_functions():
        i = text.index(n + "(")
        block = text[max(0, i - 900):i]
        assert ".f90:" in block or "no counterpart in the reference" in block, n


@pytest.mark.parametrize("bad", [dict(slope_type=4), dict(slope_type=5), dict(riemann_code=7), dict(difmag=0.1),
                                 dict(gamma=1.0), dict(nvector=0), dict(precision_code=9),
                                 dict(niter_riemann=-1)])
def test_config_rejects(bad):
    lib = ramses_amd.load_library()
    p = ramses_amd.ramses_hydro_params(1.4, 1e-10, 1e-10, 1.5, 0.5, 0.0, 1, 0, 0, 32, 10)
    for k, v in bad.items():
        setattr(p, {"riemann_code": "riemann", "precision_code": "precision"}.get(k, k), v)
    assert lib.ramses_hydro_config(ctypes.byref(p)) < 0
    assert len(lib.ramses_last_error()) > 0


def test_config_accepts_and_python_mirror():
    ramses_amd.configure(gamma=5 / 3, riemann="hllc", slope_type=1, precision="fp32")
    ramses_amd.configure(riemann="acoustic")
    ramses_amd.configure(riemann="exact", niter_riemann=3)
    ramses_amd.configure()  # defaults of hydro/hydro_parameters.f90:74-85
    with pytest.raises(ramses_amd.RamsesError):
        ramses_amd.configure(slope_type=6)


def test_product_does_not_import_oracle():
    """The product package never imports, links or runs the oracle."""
    pkg = os.path.dirname(ramses_amd.__file__)
    for root, _, files in os.walk(pkg):
        for f in files:
            if f.endswith((".py", ".hip", ".hpp", ".cpp", ".h")):
                text = open(os.path.join(root, f)).read()
                assert "import oracle" not in text and not re.search(r"\bhydro_ref\b|libramses_ref", text), f


def test_mesh_struct_layout_matches_header(tmp_path):
    """The ctypes mirror of ramses_amr_mesh has the C layout of the header."""
    import shutil
    import subprocess
    from ramses_amd import amr
    if not shutil.which("gcc"):
        pytest.skip("no host C compiler")
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    src = tmp_path / "m.c"
    src.write_text('#include <stdio.h>\n#include <stddef.h>\n#include "include/ramses_hip.h"\n'
                   'int main(){printf("%zu %zu %zu %zu %zu %zu %zu\\n", sizeof(ramses_amr_mesh),'
                   ' offsetof(ramses_amr_mesh, free), offsetof(ramses_amr_mesh, levelmin),'
                   ' offsetof(ramses_amr_mesh, boxlen), offsetof(ramses_amr_mesh, work),'
                   ' offsetof(ramses_amr_mesh, graph), offsetof(ramses_amr_mesh, graph_state));return 0;}\n')
    exe = tmp_path / "m"
    subprocess.run(["gcc", "-I", root, str(src), "-o", str(exe)], check=True)
    got = [int(x) for x in subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split()]
    M = amr.ramses_amr_mesh
    assert got == [ctypes.sizeof(M), M.free.offset, M.levelmin.offset, M.boxlen.offset, M.work.offset,
                   M.graph.offset, M.graph_state.offset]


def test_amr_entry_points_reject_bad_trees_before_any_device_work():
    """Argument checks of the oct-tree entry points run on the host before
    any launch: a tree whose cell indices would not fit the reference's
    32-bit integers, inconsistent coarse sizes, and a level count beyond
    ngridmax are refused with a message (no device memory is touched: the
    pointers here are placeholders)."""
    from ramses_amd import amr
    lib = ramses_amd.load_library()
    amr._bind(lib)
    fake = ctypes.c_void_p(0x1000)
    lv = amr.ramses_amr_level(3, 0.125, fake, 8, fake, None, 0)
    cases = [
        (amr.ramses_amr_tree(fake, fake, fake, 1, 300_000_000, 1, 1, 1), "int32"),
        (amr.ramses_amr_tree(fake, fake, fake, 2, 1000, 1, 1, 1), "ncoarse"),
        (amr.ramses_amr_tree(None, fake, fake, 1, 1000, 1, 1, 1), "null"),
    ]
    for tree, msg in cases:
        rc = lib.ramses_amr_upload(ctypes.byref(tree), ctypes.byref(lv), fake, None)
        assert rc < 0, msg
        assert msg in lib.ramses_last_error().decode(), lib.ramses_last_error()
    tree = amr.ramses_amr_tree(fake, fake, fake, 1, 1000, 1, 1, 1)
    big = amr.ramses_amr_level(3, 0.125, fake, 1001, fake, None, 0)
    assert lib.ramses_amr_courant(ctypes.byref(tree), ctypes.byref(big), fake, 1.0, fake, None) < 0
    assert "noct" in lib.ramses_last_error().decode()


def test_library_binds_after_torch_runtime():
    """Loading the library in a fresh process imports torch first, so that
    the library binds to the HIP runtime torch bundles rather than bringing
    in a second one (a process with two HIP runtimes finds no device in the
    second)."""
    import subprocess
    import sys
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    code = ("import sys, ramses_amd\n"
            "assert 'torch' not in sys.modules\n"
            "ramses_amd.load_library()\n"
            "assert 'torch' in sys.modules\n")
    subprocess.run([sys.executable, "-c", code], cwd=root, check=True, timeout=300)


def test_tuning_knobs_validate_and_read_back():
    """ramses_tune refuses out-of-range launch shapes with a message and
    ramses_tune_get reads back what was set (0 restores the default)."""
    lib = ramses_amd.load_library()
    lib.ramses_tune.argtypes = [ctypes.c_int, ctypes.c_int]
    lib.ramses_tune_get.argtypes = [ctypes.c_int]
    bad = [(1, 100), (1, 832), (2, 6), (2, -1), (3, 2), (4, 16), (4, 9000), (5, -2), (99, 0)]
    for knob, value in bad:
        assert lib.ramses_tune(knob, value) < 0, (knob, value)
        assert len(lib.ramses_last_error()) > 0
    assert lib.ramses_tune(5, 4096) == 0 and lib.ramses_tune_get(5) == 4096
    assert lib.ramses_tune(5, -1) == 0 and lib.ramses_tune_get(5) == 0  # never
    assert lib.ramses_tune(5, 0) == 0 and lib.ramses_tune_get(5) == 8192  # default
    assert lib.ramses_tune(4, 512) == 0 and lib.ramses_tune_get(4) == 512
    assert lib.ramses_tune(4, 0) == 0 and lib.ramses_tune_get(4) == 1024


def test_unsplit_wrappers_reject_mismatched_shapes():
    """The C entry points take no array sizes: they address uin, gravin,
    flux and tmp with the configured nvector.  The Python wrappers therefore
    refuse, before any call into the library, arrays whose nvector differs
    from the configuration, flux/tmp/gravin whose batch count or shape
    differs from uin, and non-float64 arrays."""
    import numpy as np
    ra = ramses_amd
    ra.configure(gamma=5 / 3, riemann="hllc", slope_type=1)  # nvector 32
    uin = np.zeros(ra.stencil_shape(32))
    flux = np.zeros(ra.flux_shape(32))
    tmp = np.zeros(ra.tmp_shape(32))
    with pytest.raises(ValueError, match="nvector"):
        ra.unsplit(np.zeros(ra.stencil_shape(16)), None, np.zeros(ra.flux_shape(16)),
                   np.zeros(ra.tmp_shape(16)), 0.1, 0.1, 0.1, 1e-3, 16)
    with pytest.raises(ValueError, match="flux"):
        ra.unsplit(uin, None, np.zeros(ra.flux_shape(32))[:2], tmp, 0.1, 0.1, 0.1, 1e-3, 32)
    with pytest.raises(ValueError, match="tmp"):
        ra.unsplit(uin, None, flux, np.zeros((3, 2, 3, 3, 2, 32)), 0.1, 0.1, 0.1, 1e-3, 32)
    with pytest.raises(ValueError, match="gravin"):
        ra.unsplit(uin, np.zeros((2, 6, 6, 6, 32)), flux, tmp, 0.1, 0.1, 0.1, 1e-3, 32)
    with pytest.raises(TypeError, match="float64"):
        ra.unsplit(uin.astype(np.float32), None, flux, tmp, 0.1, 0.1, 0.1, 1e-3, 32)
    with pytest.raises(ValueError, match="pin"):
        ra.unsplit_fortran(uin, None, np.zeros((6, 6, 6, 16)), flux, tmp, 0.1, 0.1, 0.1, 1e-3, 32)
    # batched device form: the same checks on the batch axis (numpy stand-ins
    # are enough, the check runs before any device pointer is taken)
    g_u = np.zeros((4,) + ra.stencil_shape(32))
    with pytest.raises(ValueError, match="flux"):
        ra.unsplit_device(g_u, None, np.zeros((3,) + ra.flux_shape(32)), np.zeros((4,) + ra.tmp_shape(32)),
                          0.1, 1e-3)
    with pytest.raises(ValueError, match="tmp"):
        ra.unsplit_device(g_u, None, np.zeros((4,) + ra.flux_shape(32)), np.zeros((2,) + ra.tmp_shape(32)),
                          0.1, 1e-3)
    with pytest.raises(ValueError, match="nvector"):
        ra.unsplit_device(np.zeros((4,) + ra.stencil_shape(16)), None, np.zeros((4,) + ra.flux_shape(16)),
                          np.zeros((4,) + ra.tmp_shape(16)), 0.1, 1e-3)
    ra.configure()


def test_refinement_refuses_nvector_other_than_32():
    """refine_fine numbers new grids in the order of its nvector-sized
    batches; the device refinement reproduces the reference build's
    nvector = 32 only, so another nvector is refused up front (host-side
    check: the pointers are placeholders and nothing is launched)."""
    from ramses_amd import amr
    lib = ramses_amd.load_library()
    amr._bind(lib)
    fake = ctypes.c_void_p(0x1000)
    tree = amr.ramses_amr_tree(fake, fake, fake, 1, 1000, 1, 1, 1)
    lv = amr.ramses_amr_level(3, 0.125, fake, 8, fake, None, 0)
    fl = amr.ramses_amr_free(fake, 0, 10, 990)
    nfine = ctypes.c_int(0)
    try:
        ramses_amd.configure(nvector=16)
        rc = lib.ramses_amr_refine_fine(ctypes.byref(tree), ctypes.byref(lv), fake, ctypes.byref(nfine), fake,
                                        ctypes.byref(fl), fake, fake, None, fake, fake, fake, 1 << 20, fake, None)
        assert rc == -8 and "nvector = 32" in lib.ramses_last_error().decode()
    finally:
        ramses_amd.configure()


def test_interpolation_config_validates():
    """ramses_amr_config accepts interpol_type 0..4 and interpol_var 0..2,
    and refuses type 4 with any interpol_var but 2, where the reference
    stops (hydro/interpol_hydro.f90:367-372)."""
    from ramses_amd import amr
    lib = ramses_amd.load_library()
    amr._bind(lib)
    for t, v in [(0, 0), (1, 1), (2, 2), (3, 0), (4, 2)]:
        assert lib.ramses_amr_config(t, v) == 0, (t, v)
    for t, v in [(4, 0), (4, 1), (5, 2), (-1, 0), (1, 3)]:
        assert lib.ramses_amr_config(t, v) < 0, (t, v)
        assert len(lib.ramses_last_error()) > 0
    assert lib.ramses_amr_config(1, 0) == 0
