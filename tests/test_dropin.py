"""CPU checks of the two drop-in builds of the reference (oracle/Makefile
`dropin`): the reference's objects without umuscl.o, with godfine1's
`call unsplit` (hydro/godunov_fine.f90:681) resolved by this repository's
library.  No GPU call is made here; the GPU test
test_gpu_amr.py::test_reference_godunov_through_library_unsplit runs them."""
import ctypes
import os
import re
import subprocess

import pytest

import oracle
import ramses_amd

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF_DIR = os.path.dirname(oracle.REF_LIB)
needs_ref = pytest.mark.needs_ref


def _addr(lib, name):
    return ctypes.cast(getattr(lib, name), ctypes.c_void_p).value


def _lib(name):
    path = os.path.join(REF_DIR, name)
    if not os.path.exists(path):
        pytest.skip(f"{path} is not built (make -C oracle dropin)")
    return path


@needs_ref
def test_link_level_dropin_resolves_unsplit_to_library():
    """Without umuscl.o the reference's `unsplit_` comes from libramses_hip.so."""
    hip = ramses_amd.load_library()
    path = _lib("libramses_ref_dropin.so")
    out = subprocess.run(["nm", "-D", path], capture_output=True, text=True, check=True).stdout
    assert re.search(r"^\s+U unsplit_$", out, re.M), "unsplit_ must be undefined in the drop-in build"
    assert re.search(r" T godfine1_$", out, re.M)
    lib = ctypes.CDLL(path, mode=os.RTLD_LAZY)
    assert _addr(lib, "unsplit_") == _addr(hip, "unsplit_")


@needs_ref
def test_integration_patch_compiles_and_forwards():
    """The integration guide's patch file (its first fortran block), compiled
    by flang against the reference's modules, is what the patched build links: it defines
    `unsplit_` itself and calls the library's ramses_unsplit and
    ramses_hydro_config."""
    text = open(os.path.join(ROOT, "INTEGRATION.md")).read()
    block = re.search(r"```fortran\n(.*?)```", text, re.S).group(1)
    src = os.path.join(REF_DIR, "umuscl_patch.f90")
    if not os.path.exists(src):
        pytest.skip("patch not extracted (make -C oracle dropin)")
    assert open(src).read() == block, "oracle/_ref/umuscl_patch.f90 is stale: rebuild with make -C oracle dropin"
    hip = ramses_amd.load_library()
    path = _lib("libramses_ref_patched.so")
    out = subprocess.run(["nm", "-D", path], capture_output=True, text=True, check=True).stdout
    assert re.search(r" T unsplit_$", out, re.M)
    assert re.search(r"^\s+U ramses_unsplit$", out, re.M)
    assert re.search(r"^\s+U ramses_hydro_config$", out, re.M)
    lib = ctypes.CDLL(path, mode=os.RTLD_LAZY)
    assert _addr(lib, "ramses_unsplit") == _addr(hip, "ramses_unsplit")
