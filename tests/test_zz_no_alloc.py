"""ramses_unsplit_noalloc (staging in the library's fixed device array)
against ramses_unsplit (staging the library allocates): same bits, on two
calls in a row, the second with a smaller batch."""
import ctypes

import numpy as np
import pytest

import ramses_amd as ra
from helpers import NV, random_stencils

_dp = ctypes.POINTER(ctypes.c_double)


def _noalloc_fn():
    lib = ra.load_library()
    f = lib.ramses_unsplit_noalloc
    f.argtypes = [_dp, _dp, _dp, _dp] + [ctypes.c_double] * 4 + [ctypes.c_int, ctypes.c_void_p]
    f.restype = ctypes.c_int
    return f


def _p(a):
    return None if a is None else a.ctypes.data_as(_dp)


def _both(uin, grav, ngrid, dx, dt, fill):
    """(flux, tmp) of ramses_unsplit and of ramses_unsplit_noalloc, both
    starting from flux/tmp filled with `fill`."""
    f_ref = np.full(ra.flux_shape(NV), fill)
    t_ref = np.full(ra.tmp_shape(NV), fill)
    ra.unsplit(uin, grav, f_ref, t_ref, dx, dx, dx, dt, ngrid)
    f_na = np.full(ra.flux_shape(NV), fill)
    t_na = np.full(ra.tmp_shape(NV), fill)
    rc = _noalloc_fn()(_p(uin), _p(grav), _p(f_na), _p(t_na), dx, dx, dx, dt, ngrid, None)
    ra._check(rc, "ramses_unsplit_noalloc")
    return (f_ref, t_ref), (f_na, t_na)


@pytest.mark.gpu
@pytest.mark.parametrize("riemann,slope,precision", [("hllc", 1, "fp64"), ("llf", 2, "fp64"), ("hllc", 1, "fp32")])
def test_noalloc_matches_allocating_form(gpu, riemann, slope, precision):
    rng = np.random.default_rng(11)
    ra.configure(gamma=5 / 3, riemann=riemann, slope_type=slope, precision=precision)
    try:
        # call 1: a full batch with gravity; call 2: 7 lanes, no gravity, so
        # the staging still holds call 1's data beyond what call 2 copies in
        uin1 = random_stencils(rng, 1)[0]
        grav1 = rng.normal(0.0, 0.3, ra.grav_shape(NV))
        uin2 = random_stencils(rng, 1)[0]
        for uin, grav, ngrid, fill in ((uin1, grav1, NV, 0.0), (uin2, None, 7, -3.5)):
            (f_ref, t_ref), (f_na, t_na) = _both(uin, grav, ngrid, 1 / 128, 2e-4, fill)
            assert np.array_equal(f_na, f_ref), (riemann, slope, precision, ngrid)
            assert np.array_equal(t_na, t_ref), (riemann, slope, precision, ngrid)
            # lanes beyond ngrid keep the caller's values, as the reference leaves them
            assert np.all(f_na[..., ngrid:] == fill) and np.all(t_na[..., ngrid:] == fill)
            assert np.any(f_na[..., :ngrid] != fill)
    finally:
        ra.configure()


def test_noalloc_rejects_nvector_beyond_staging():
    """A configured nvector whose batch does not fit the fixed staging is
    refused before any device work (runs without a GPU)."""
    f = _noalloc_fn()
    ra.configure(nvector=128)
    try:
        uin = np.zeros(ra.stencil_shape(128))
        flux = np.zeros(ra.flux_shape(128))
        tmp = np.zeros(ra.tmp_shape(128))
        rc = f(_p(uin), None, _p(flux), _p(tmp), 0.1, 0.1, 0.1, 1e-3, 1, None)
        assert rc == -11
        assert "nvector <= 114" in ra.load_library().ramses_last_error().decode()
    finally:
        ra.configure()
