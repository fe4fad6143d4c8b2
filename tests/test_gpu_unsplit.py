"""GPU parity of the batched unsplit kernel through the C ABI against the
reference's golden vectors and the CPU oracle (bitwise in FP64)."""
import numpy as np
import pytest

import oracle
import ramses_amd as ra
from helpers import NV, defined_face_mask, random_stencils

pytestmark = pytest.mark.gpu


def _golden_inputs(meta, d, ci):
    uin = np.zeros(ra.stencil_shape(NV))
    uin[..., :2] = d[f"c{ci}_uin"]
    g = None
    if meta[ci]["grav"]:
        g = np.zeros(ra.grav_shape(NV))
        g[..., :2] = d[f"c{ci}_grav"]
    return uin, g


def test_unsplit_golden_host_abi(gpu, golden_unsplit):
    meta, d = golden_unsplit
    for ci, c in enumerate(meta):
        ra.configure(gamma=c["gamma"], riemann=c["riemann"], slope_type=c["slope_type"])
        uin, g = _golden_inputs(meta, d, ci)
        flux = np.full(ra.flux_shape(NV), np.nan)
        tmp = np.full(ra.tmp_shape(NV), np.nan)
        ra.unsplit(uin, g, flux, tmp, c["dx"], c["dx"], c["dx"], c["dt"], 2)
        assert np.array_equal(flux[..., :2], d[f"c{ci}_flux"], equal_nan=True), c
        assert np.array_equal(tmp[..., :2], d[f"c{ci}_tmp"], equal_nan=True), c
        assert np.all(np.isnan(flux[..., 2:])) and np.all(np.isnan(tmp[..., 2:]))


def test_unsplit_golden_fortran_symbol(gpu, golden_unsplit):
    meta, d = golden_unsplit
    pin = np.zeros((6, 6, 6, NV))
    for ci, c in enumerate(meta):
        if c["kind"] != "random":
            continue
        ra.configure(gamma=c["gamma"], riemann=c["riemann"], slope_type=c["slope_type"])
        uin, g = _golden_inputs(meta, d, ci)
        gz = np.zeros(ra.grav_shape(NV)) if g is None else g
        flux = np.full(ra.flux_shape(NV), np.nan)
        tmp = np.full(ra.tmp_shape(NV), np.nan)
        ra.unsplit_fortran(uin, gz, pin, flux, tmp, c["dx"], c["dx"], c["dx"], c["dt"], 2)
        assert np.array_equal(flux[..., :2], d[f"c{ci}_flux"], equal_nan=True), c


def test_fortran_symbol_zero_gravity_forms(gpu):
    """unsplit_ hands an all-(+0.0) gravity array over as no array, which
    gives the same bits as the array itself; a -0.0 array stays an array
    (with a momentum of -0.0, -0.0 + (-0.0 * dt/2) keeps the sign of the
    primitive velocity that -0.0 + 0.0 drops).  Compared as bit patterns, so
    the sign of zero counts."""
    rng = np.random.default_rng(91)
    uin = np.ascontiguousarray(random_stencils(rng, 1)[0])
    uin[1:4] = -0.0
    ra.configure(gamma=5 / 3, riemann="hllc", slope_type=1)
    pin = np.zeros((6, 6, 6, NV))
    dx, dt = 1 / 128, 2e-4

    def run(fn, g):
        flux = np.zeros(ra.flux_shape(NV))
        tmp = np.zeros(ra.tmp_shape(NV))
        if fn == "fortran":
            ra.unsplit_fortran(uin, g, pin, flux, tmp, dx, dx, dx, dt, NV)
        else:
            ra.unsplit(uin, g, flux, tmp, dx, dx, dx, dt, NV)
        return flux.view(np.uint64), tmp.view(np.uint64)

    neg = np.full(ra.grav_shape(NV), -0.0)
    pos = np.zeros(ra.grav_shape(NV))
    for a, b in zip(run("fortran", neg), run("c", neg)):
        np.testing.assert_array_equal(a, b)
    for a, b in zip(run("fortran", pos), run("c", None)):
        np.testing.assert_array_equal(a, b)
    for a, b in zip(run("fortran", pos), run("c", pos)):
        np.testing.assert_array_equal(a, b)
    ra.configure()


@pytest.mark.parametrize("gravity", [False, True])
@pytest.mark.parametrize("riemann", ["llf", "hllc", "hll", "acoustic", "exact"])
@pytest.mark.parametrize("slope", [0, 1, 2, 3, 7, 8])
def test_unsplit_device_batched_vs_oracle(gpu, oracle_lib, riemann, slope, gravity):
    """Bitwise against the oracle; the exact solver's rarefaction branches
    call pow, whose device and host libraries may round the last bit apart,
    so it is held to the FP64 bar of 1e-12 relative instead.  Calls with and
    without a gravity array run different kernel instances (the ones without
    carry no gravity loads), so every solver and slope type runs both."""
    import torch
    rng = np.random.default_rng(1000 + 10 * slope + len(riemann))
    G = 40
    uin = random_stencils(rng, G)
    grav = rng.normal(0, 1, (G,) + ra.grav_shape(NV)) if gravity else None
    dx, dt, last = 1 / 128, 2e-4, 7
    ra.configure(gamma=5 / 3, riemann=riemann, slope_type=slope)
    oracle_lib.set_params(gamma=5 / 3, riemann=riemann, slope_type=slope)
    f_ref = np.full((G,) + ra.flux_shape(NV), np.nan)
    t_ref = np.full((G,) + ra.tmp_shape(NV), np.nan)
    oracle_lib.unsplit_batched(uin, grav, f_ref, t_ref, dx, dt, ngrid_last=last)
    d_u = torch.from_numpy(uin).cuda()
    d_g = None if grav is None else torch.from_numpy(grav).cuda()
    d_f = torch.full(f_ref.shape, float("nan"), dtype=torch.float64, device="cuda")
    d_t = torch.full(t_ref.shape, float("nan"), dtype=torch.float64, device="cuda")
    ra.unsplit_device(d_u, d_g, d_f, d_t, dx, dt, ngrid_last=last)
    torch.cuda.synchronize()
    f, t = d_f.cpu().numpy(), d_t.cpu().numpy()
    if riemann == "exact":
        np.testing.assert_allclose(f, f_ref, rtol=1e-12, atol=1e-300, equal_nan=True)
        np.testing.assert_allclose(t, t_ref, rtol=1e-12, atol=1e-300, equal_nan=True)
    else:
        assert np.array_equal(f, f_ref, equal_nan=True)
        assert np.array_equal(t, t_ref, equal_nan=True)


def test_unsplit_fp32_mixed_error(gpu, oracle_lib):
    """FP32-mixed fluxes against the FP64 oracle: relative L1 per variable,
    without and with a gravity array."""
    import torch
    rng = np.random.default_rng(55)
    G = 64
    uin = random_stencils(rng, G)
    grav = rng.normal(0, 1, (G,) + ra.grav_shape(NV))
    dx, dt = 1 / 128, 2e-4
    m = defined_face_mask(NV)
    for riemann in ("hllc", "llf"):
        for g in (None, grav):
            ra.configure(gamma=5 / 3, riemann=riemann, slope_type=1, precision="fp32")
            oracle_lib.set_params(gamma=5 / 3, riemann=riemann, slope_type=1)
            f_ref = np.zeros((G,) + ra.flux_shape(NV))
            t_ref = np.zeros((G,) + ra.tmp_shape(NV))
            oracle_lib.unsplit_batched(uin, g, f_ref, t_ref, dx, dt)
            d_f = torch.zeros(f_ref.shape, dtype=torch.float64, device="cuda")
            d_t = torch.zeros(t_ref.shape, dtype=torch.float64, device="cuda")
            d_g = None if g is None else torch.from_numpy(g).cuda()
            ra.unsplit_device(torch.from_numpy(uin).cuda(), d_g, d_f, d_t, dx, dt)
            torch.cuda.synchronize()
            f = d_f.cpu().numpy()
            for v in range(5):
                a, b = f[:, :, v][:, m[:, v]], f_ref[:, :, v][:, m[:, v]]
                rel_l1 = np.abs(a - b).sum() / np.abs(b).sum()
                assert rel_l1 < 1e-5, (riemann, g is None, v, rel_l1)
    ra.configure()


def test_cmpdt_golden(gpu, golden_cmpdt):
    meta, d = golden_cmpdt
    for ci, c in enumerate(meta):
        ra.configure(gamma=c["gamma"], courant_factor=c["courant_factor"])
        uu = d[f"c{ci}_uu_in"].copy()
        dt = ra.cmpdt(uu, d[f"c{ci}_gg"].copy(), c["dx"], c["ncell"])
        assert dt == float(d[f"c{ci}_dt"]), c
    ra.configure()


@pytest.mark.parametrize("gravity", [False, True])
@pytest.mark.parametrize("precision", ["fp32", "fp64"])
def test_unsplit_nvector_layouts_agree(gpu, precision, gravity):
    """The same octs laid out in batches of 32 and of 16 give the same bits.
    With 32 the FP32-mixed kernel stages each cell's primitives once in LDS;
    with 16 the row kernel evaluates them per row, so this pins the two
    forms against each other, without and with a gravity array."""
    import torch
    rng = np.random.default_rng(77)
    G = 8
    u32 = random_stencils(rng, G, nv=32)
    u16 = u32.reshape(G, 5, 6, 6, 6, 2, 16).transpose(0, 5, 1, 2, 3, 4, 6).reshape(2 * G, 5, 6, 6, 6, 16)
    g32 = g16 = None
    if gravity:
        g32 = rng.normal(0, 1, (G,) + ra.grav_shape(32))
        g16 = g32.reshape(G, 3, 6, 6, 6, 2, 16).transpose(0, 5, 1, 2, 3, 4, 6).reshape(2 * G, 3, 6, 6, 6, 16)
    dx, dt = 1 / 128, 2e-4
    out = {}
    for nv, u, g in ((32, u32, g32), (16, u16, g16)):
        ra.configure(gamma=5 / 3, riemann="hllc", slope_type=1, precision=precision, nvector=nv)
        f = torch.zeros((u.shape[0],) + ra.flux_shape(nv), dtype=torch.float64, device="cuda")
        t = torch.zeros((u.shape[0],) + ra.tmp_shape(nv), dtype=torch.float64, device="cuda")
        d_g = None if g is None else torch.from_numpy(np.ascontiguousarray(g)).cuda()
        ra.unsplit_device(torch.from_numpy(np.ascontiguousarray(u)).cuda(), d_g, f, t, dx, dt)
        torch.cuda.synchronize()
        out[nv] = (f.cpu().numpy(), t.cpu().numpy())
    ra.configure()
    for k in range(2):  # flux, tmp (undefined slots stay zero in both)
        a = out[32][k]
        b = out[16][k]
        b = np.moveaxis(b.reshape((G, 2) + b.shape[1:]), 1, -2).reshape(a.shape)
        assert np.abs(a).sum() > 0
        np.testing.assert_array_equal(a, b)
