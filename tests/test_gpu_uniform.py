"""GPU parity of the resident uniform level (godfine1 sweep + fused courant)
against the reference-composed golden run and the CPU oracle."""
import numpy as np
import pytest

import ramses_amd as ra
from helpers import random_grid

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("fused", [True, False])
def test_sedov16_golden(gpu, golden_sedov16, fused):
    g = golden_sedov16
    ra.configure(gamma=1.4, riemann="hllc", slope_type=1, courant_factor=0.8)
    lvl = ra.UniformLevel(g["u0"])
    for s in range(1, 11):
        dt = lvl.step(fused=fused)
        assert float(dt) == g["dt"][s - 1], s
        if f"u_step{s}" in g:
            assert np.array_equal(lvl.u.cpu().numpy(), g[f"u_step{s}"]), s


@pytest.mark.parametrize("riemann,slope", [("hllc", 1), ("llf", 1), ("hll", 2), ("hllc", 7), ("llf", 8),
                                           ("hllc", 0), ("acoustic", 1), ("exact", 1)])
def test_random_grid_vs_oracle(gpu, oracle_lib, riemann, slope):
    rng = np.random.default_rng(77)
    n = 32
    u = random_grid(rng, n)
    ra.configure(gamma=5 / 3, riemann=riemann, slope_type=slope, courant_factor=0.5)
    oracle_lib.set_params(gamma=5 / 3, riemann=riemann, slope_type=slope, courant_factor=0.5)
    lvl = ra.UniformLevel(u)
    ref = u
    for s in range(2):
        ref, dt_ref = oracle_lib.uniform_step(ref, 1.0 / n)
        dt = float(lvl.step())
        if riemann == "exact":  # pow in the rarefaction branches: last-bit differences allowed
            assert abs(dt - dt_ref) <= 1e-12 * dt_ref
            np.testing.assert_allclose(lvl.u.cpu().numpy(), ref, rtol=1e-12, atol=1e-300)
            ref = lvl.u.cpu().numpy()
            continue
        assert dt == dt_ref
        assert np.array_equal(lvl.u.cpu().numpy(), ref), (riemann, slope, s)


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_every_tile_shape_same_bits(gpu, oracle_lib, precision):
    """Every launch shape the autotuner may pick (tiles 32x8, 16x8, 8x8,
    32x16 and 16x12 with z chunks of 4..64 planes) gives the same bits: equal
    to the oracle in FP64, and to each other in FP32-mixed.  At n = 64 the
    16x12 tiles cover rows 0..47 and a second launch of 16x8 tiles rows
    48..63."""
    from ramses_amd import autotune
    rng = np.random.default_rng(9)
    n = 64
    u = random_grid(rng, n, mach=2.0)
    ra.configure(gamma=5 / 3, riemann="hllc", slope_type=1, courant_factor=0.5, precision=precision)
    oracle_lib.set_params(gamma=5 / 3, riemann="hllc", slope_type=1, courant_factor=0.5)
    ref = None
    if precision == "fp64":
        ref, dt_ref = oracle_lib.uniform_step(u, 1.0 / n)
    try:
        for tile in (1, 2, 3, 4, 5):
            for tz in (4, 16, 64):
                autotune._set(autotune.LEVEL_TILE, tile)
                autotune._set(autotune.LEVEL_TZ, tz)
                lvl = ra.UniformLevel(u)
                dt = float(lvl.step())
                out = lvl.u.cpu().numpy()
                if ref is None:
                    ref, dt_ref = out, dt
                assert dt == dt_ref, (tile, tz)
                assert np.array_equal(out, ref), (precision, tile, tz)
    finally:
        autotune._set(autotune.LEVEL_TILE, 0)
        autotune._set(autotune.LEVEL_TZ, 0)
        ra.configure()


def test_conservation_and_determinism(gpu):
    rng = np.random.default_rng(5)
    n = 64
    u = random_grid(rng, n, mach=2.0)
    ra.configure(gamma=5 / 3, riemann="hllc", slope_type=1, courant_factor=0.5)
    runs = []
    for _ in range(2):
        lvl = ra.UniformLevel(u)
        for _ in range(3):
            lvl.step()
        runs.append(lvl.u.cpu().numpy())
    assert np.array_equal(runs[0], runs[1])
    for v in range(5):
        assert abs(runs[0][v].sum() - u[v].sum()) <= 1e-12 * np.abs(u[v]).sum()


def test_fp32_mixed_sedov_error(gpu, golden_sedov16):
    g = golden_sedov16
    ra.configure(gamma=1.4, riemann="hllc", slope_type=1, courant_factor=0.8, precision="fp32")
    lvl = ra.UniformLevel(g["u0"])
    for s in range(10):
        lvl.step()
    u = lvl.u.cpu().numpy()
    ref = g["u_step10"]
    for v in (0, 4):
        rel = np.abs(u[v] - ref[v]).sum() / np.abs(ref[v]).sum()
        assert rel < 1e-5, (v, rel)
    ra.configure()


def test_graph_replay_equals_eager(gpu):
    """Two fused steps captured in a hipGraph and replayed several times,
    with eager work and synchronisations in between, give the eager bits."""
    import torch
    rng = np.random.default_rng(21)
    u = random_grid(rng, 32, mach=2.0)
    ra.configure(gamma=5 / 3, riemann="hllc", slope_type=1, courant_factor=0.5)
    eager = ra.UniformLevel(u)
    graphed = ra.UniformLevel(u)
    for rep in range(4):
        for _ in range(2):
            eager.step()
        graphed.advance(2, graph=True)
        torch.cuda.synchronize()
        _ = (graphed.u * 2).sum()  # eager work between replays
        assert torch.equal(eager.u, graphed.u), rep
        assert torch.equal(eager.dt, graphed.dt), rep
    graphed.advance(3, graph=True)
    for _ in range(3):
        eager.step()
    assert torch.equal(eager.u, graphed.u)


def test_uniform_conservation_sums_vs_reference(gpu):
    """courant_fine's sums (hydro/courant_fine.f90:95-119) of a uniform
    level: the initial state's and, after each fused step, the new state's,
    against the reference's own courant_fine on the same states to 1e-12;
    the sums change neither dt nor the state, and repeat bit for bit."""
    import torch
    from oracle.ref_uniform import ReferenceUniformLevel
    from ramses_amd import ics
    n = 32
    u0 = ics.turbulent_host(n, 5.0, rho_sigma=0.5, seed=4)
    ra.configure(gamma=5 / 3, riemann="hllc", slope_type=1, courant_factor=0.5)
    R = ReferenceUniformLevel(n, 5 / 3, riemann="hllc", slope_type=1, courant_factor=0.5)
    R.load(u0)
    plain = ra.UniformLevel(u0)
    runs = []
    for rep in range(2):
        lvl = ra.UniformLevel(u0, sums=True)
        lvl.courant()
        got = [lvl.sums.cpu().numpy().copy()]
        for s in range(3):
            lvl.step()
            got.append(lvl.sums.cpu().numpy().copy())
        runs.append((got, lvl.u.clone()))
    for s in range(3):
        plain.step()
    assert torch.equal(runs[0][1], plain.u) and torch.equal(runs[1][1], plain.u)
    assert all(np.array_equal(a, b) for a, b in zip(runs[0][0], runs[1][0]))
    for s in range(4):
        ref = R.courant()
        np.testing.assert_allclose(runs[0][0][s], ref[1:], rtol=1e-12, atol=0)
        if s < 3:
            R.step(ref[0])
    R.release()
    ra.configure()
