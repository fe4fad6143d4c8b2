"""ramses_amd -- the RAMSES unsplit MUSCL-Hancock hydro step on MI355X.

Host-side mirror of the reference's operator interface for this path.  The
numerics run only in the gfx950 library ``libramses_hip.so`` (HIP kernels
behind the C ABI of ``include/ramses_hip.h``); this module loads it with
ctypes and raises if it is missing -- there is no CPU fallback.

Reference interface mirrored here:
  * ``unsplit(uin, gravin, flux, tmp, dx, dy, dz, dt, ngrid)`` --
    hydro/umuscl.f90:22 (the ``pin`` argument is never read by the reference
    and is accepted and ignored by ``unsplit_``).
  * ``cmpdt(uu, gg, dx, ncell)`` -- hydro/godunov_utils.f90:5.
  * ``configure(...)`` -- the hydro module globals,
    hydro/hydro_parameters.f90:74-85.
  * ``UniformLevel`` -- godunov_fine/godfine1, set_unew/set_uold and
    courant_fine over a fully refined periodic level kept in device memory.
Arrays use the reference's Fortran layouts; as C-ordered numpy/torch arrays
their axes are the Fortran axes reversed (see ``stencil_shape`` etc.).
"""
import ctypes
import os
import re

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libramses_hip.so")
HEADER = os.path.join(os.path.dirname(HERE), "include", "ramses_hip.h")

NVAR, NDIM = 5, 3
RIEMANN = {"llf": 0, "hllc": 1, "hll": 2, "acoustic": 3, "exact": 4}
PRECISION = {"fp64": 0, "fp32": 1, "fp32_mixed": 1}

_dp = ctypes.POINTER(ctypes.c_double)


class ramses_hydro_params(ctypes.Structure):
    _fields_ = [("gamma", ctypes.c_double), ("smallr", ctypes.c_double), ("smallc", ctypes.c_double),
                ("slope_theta", ctypes.c_double), ("courant_factor", ctypes.c_double),
                ("difmag", ctypes.c_double), ("slope_type", ctypes.c_int), ("riemann", ctypes.c_int),
                ("precision", ctypes.c_int), ("nvector", ctypes.c_int), ("niter_riemann", ctypes.c_int)]


class RamsesError(RuntimeError):
    pass


_lib = None


def load_library(path=LIB_PATH):
    """Load libramses_hip.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise ImportError(f"{path} is missing: build it with `python -m ramses_amd.build` "
                          "(the HIP library is the only implementation; there is no CPU fallback)")
    # PyTorch ships its own HIP and HSA runtimes. Loading the library first
    # would bring in the system copies, and a process holding two runtimes
    # finds no device in the second one; importing torch first makes the
    # library bind to the runtime torch already holds.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    lib = ctypes.CDLL(path)
    lib.ramses_hydro_config.argtypes = [ctypes.POINTER(ramses_hydro_params)]
    lib.ramses_hydro_config.restype = ctypes.c_int
    lib.ramses_last_error.restype = ctypes.c_char_p
    lib.ramses_unsplit.argtypes = [_dp, _dp, _dp, _dp] + [ctypes.c_double] * 4 + [ctypes.c_int]
    lib.ramses_unsplit.restype = ctypes.c_int
    lib.unsplit_.argtypes = [_dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp, ctypes.POINTER(ctypes.c_int)]
    lib.unsplit_.restype = None
    lib.ramses_unsplit_device.argtypes = [ctypes.c_void_p] * 4 + [ctypes.c_double] * 4 + \
        [ctypes.c_long, ctypes.c_int, ctypes.c_void_p]
    lib.ramses_unsplit_device.restype = ctypes.c_int
    lib.ramses_cmpdt.argtypes = [_dp, _dp, ctypes.c_double, ctypes.c_int, _dp]
    lib.ramses_cmpdt.restype = ctypes.c_int
    lib.ramses_uniform_courant.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_double, ctypes.c_double,
                                           ctypes.c_void_p, ctypes.c_void_p]
    lib.ramses_uniform_courant.restype = ctypes.c_int
    lib.ramses_uniform_godunov.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_double,
                                           ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
    lib.ramses_uniform_godunov.restype = ctypes.c_int
    lib.ramses_uniform_dt_init.argtypes = [ctypes.c_void_p, ctypes.c_double, ctypes.c_void_p]
    lib.ramses_uniform_dt_init.restype = ctypes.c_int
    _lib = lib
    return lib


def header_functions(path=HEADER):
    """Names of the functions declared in include/ramses_hip.h."""
    text = open(path).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b([A-Za-z_][A-Za-z0-9_]*)\s*\(", text)) -
                  {"if", "defined", "sizeof"})


def _check(rc, what):
    if rc != 0:
        msg = load_library().ramses_last_error().decode()
        raise RamsesError(f"{what} failed (status {rc}): {msg}")


def _ptr(a):
    if a is None:
        return None
    if a.dtype != np.float64 or not a.flags.c_contiguous:
        raise TypeError("expected a C-contiguous float64 array")
    return a.ctypes.data_as(_dp)


def _dev(t):
    """Device address of a torch tensor (or None)."""
    if t is None:
        return None
    if not t.is_cuda or not t.is_contiguous():
        raise TypeError("expected a contiguous device tensor")
    return ctypes.c_void_p(t.data_ptr())


def _stream_handle(stream):
    if stream is None:
        import torch
        return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    if isinstance(stream, int):
        return ctypes.c_void_p(stream)
    return ctypes.c_void_p(stream.cuda_stream)


_config_cache = {}


def configure(gamma=1.4, smallr=1e-10, smallc=1e-10, slope_type=1, slope_theta=1.5, riemann="llf",
              courant_factor=0.5, difmag=0.0, precision="fp64", nvector=32, niter_riemann=10):
    """Install the solver configuration (the reference's hydro module globals,
    hydro/hydro_parameters.f90:74-85, with the reference's defaults)."""
    lib = load_library()
    p = ramses_hydro_params(gamma, smallr, smallc, slope_theta, courant_factor, difmag, slope_type,
                            RIEMANN[riemann], PRECISION[precision], nvector, niter_riemann)
    _check(lib.ramses_hydro_config(ctypes.byref(p)), "ramses_hydro_config")
    _config_cache.clear()
    _config_cache.update(gamma=gamma, smallr=smallr, smallc=smallc, slope_type=slope_type,
                         slope_theta=slope_theta, riemann=riemann, courant_factor=courant_factor,
                         precision=precision, nvector=nvector, niter_riemann=niter_riemann)


def stencil_shape(nv=32):
    return (NVAR, 6, 6, 6, nv)


def grav_shape(nv=32):
    return (NDIM, 6, 6, 6, nv)


def flux_shape(nv=32):
    return (NDIM, NVAR, 3, 3, 3, nv)


def tmp_shape(nv=32):
    return (NDIM, 2, 3, 3, 3, nv)


def _configured_nvector():
    return _config_cache.get("nvector", 32)


def _check_shapes(arrays, ng, nv):
    """Raise ValueError unless every array has the unsplit layout for `ng`
    batches (ng None: one call, no batch axis) of the configured nvector.
    The C entry points take no sizes beyond ngrid: they address every array
    with the configured nvector, so a smaller array would be overrun."""
    if nv != _configured_nvector():
        raise ValueError(f"arrays have nvector={nv} but the library is configured with "
                         f"nvector={_configured_nvector()}")
    lead = () if ng is None else (ng,)
    for name, a, shape in arrays:
        if a is None:
            continue
        if tuple(a.shape) != lead + shape(nv):
            raise ValueError(f"{name} has shape {tuple(a.shape)}, expected {lead + shape(nv)}")
        dt = getattr(a, "dtype", None)
        if str(dt) not in ("float64", "torch.float64"):
            raise TypeError(f"{name} must be float64, got {dt}")


def unsplit(uin, gravin, flux, tmp, dx, dy, dz, dt, ngrid):
    """hydro/umuscl.f90:22 on host numpy arrays (flux/tmp updated in place)."""
    lib = load_library()
    nv = uin.shape[-1]
    _check_shapes([("uin", uin, stencil_shape), ("gravin", gravin, grav_shape), ("flux", flux, flux_shape),
                   ("tmp", tmp, tmp_shape)], None, nv)
    _check(lib.ramses_unsplit(_ptr(uin), _ptr(gravin), _ptr(flux), _ptr(tmp), dx, dy, dz, dt, ngrid),
           "ramses_unsplit")


def unsplit_fortran(uin, gravin, pin, flux, tmp, dx, dy, dz, dt, ngrid):
    """Calls the link-level symbol ``unsplit_`` exactly as Fortran would."""
    lib = load_library()
    nv = uin.shape[-1]
    _check_shapes([("uin", uin, stencil_shape), ("gravin", gravin, grav_shape), ("flux", flux, flux_shape),
                   ("tmp", tmp, tmp_shape)], None, nv)
    if pin is not None and tuple(pin.shape) != (6, 6, 6, nv):
        raise ValueError(f"pin has shape {tuple(pin.shape)}, expected {(6, 6, 6, nv)}")
    d = [ctypes.c_double(v) for v in (dx, dy, dz, dt)]
    lib.unsplit_(_ptr(uin), _ptr(gravin), _ptr(pin), _ptr(flux), _ptr(tmp), ctypes.byref(d[0]),
                 ctypes.byref(d[1]), ctypes.byref(d[2]), ctypes.byref(d[3]), ctypes.byref(ctypes.c_int(ngrid)))


def unsplit_device(uin, gravin, flux, tmp, dx, dt, ngrid_last=None, dy=None, dz=None, stream=None):
    """Batched unsplit over device tensors uin[g][5][6][6][6][nv] etc."""
    lib = load_library()
    ng, nv = uin.shape[0], uin.shape[-1]
    _check_shapes([("uin", uin, stencil_shape), ("gravin", gravin, grav_shape), ("flux", flux, flux_shape),
                   ("tmp", tmp, tmp_shape)], ng, nv)
    rc = lib.ramses_unsplit_device(_dev(uin), _dev(gravin), _dev(flux), _dev(tmp), dx,
                                   dx if dy is None else dy, dx if dz is None else dz, dt, ng,
                                   nv if ngrid_last is None else ngrid_last, _stream_handle(stream))
    _check(rc, "ramses_unsplit_device")


def cmpdt(uu, gg, dx, ncell):
    """hydro/godunov_utils.f90:5 on a host (nvar, nvector) array; returns dt."""
    lib = load_library()
    dt = ctypes.c_double()
    _check(lib.ramses_cmpdt(_ptr(uu), _ptr(gg), dx, ncell, ctypes.byref(dt)), "ramses_cmpdt")
    return dt.value


from .uniform import UniformLevel  # noqa: E402,F401
