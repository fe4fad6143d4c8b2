"""GPU parity on the workloads the metric is quoted on, at their own sizes.

* Sedov 64^3 (levelmin = levelmax = 6, gamma 1.4, hllc, minmod, CFL 0.8):
  20 level steps bitwise against the fixture written by the reference's own
  courant_fine / godunov_fine (tests/golden/sedov64.npz).
* Synthetic oct batches, 2^20 octs: the full batched launch the benchmark
  times, a strided sample of its batches bitwise against the oracle; the
  FP32-mixed LDS-staged kernel with a ragged last batch.
* 512^3 Mach-5 k^-4 field (rho lognormal 0.5, P = rho/gamma, gamma 5/3,
  hllc, minmod, CFL 0.5): the same field at 64^3 bitwise against the
  reference's own level routines and at 128^3 against the oracle; at 512^3
  the conservation of mass, momentum and energy, run-to-run identity, the
  first dt against the oracle's courant_fine, a 128^3 sub-box bitwise
  against the oracle for three steps, and FP32-mixed within 1e-5 of FP64.
* Serving: requests of the serving benchmark's stream (sides from
  default_rng(3), Mach-2 k^-4 fields), one 128^3 to t = 0.02 and a 64^3 to
  t = 0.1, served with continuous batching; every FP64 result bitwise equal
  to the oracle stepped with the same end-time cap, FP32-mixed within 1e-5.
The reference routines these follow: hydro/godunov_fine.f90:486-911,
hydro/umuscl.f90:22-171, hydro/courant_fine.f90:44-131.
"""
import os

import numpy as np
import pytest

import oracle
import ramses_amd as ra
from conftest import GOLDEN
from helpers import defined_face_mask, random_stencils

pytestmark = pytest.mark.gpu
needs_ref = pytest.mark.needs_ref

# hydro parameters of the Mach-5 k^-4 turbulence level
TURB_HYDRO = dict(gamma=5.0 / 3.0, riemann="hllc", slope_type=1, courant_factor=0.5)
REL_L1_FP32 = 1e-5  # FP32-mixed vs FP64, relative L1 per conserved variable


def _rel_l1(a, b):
    return float(np.abs(a - b).sum() / np.abs(b).sum())


# ---------------------------------------------------------------- Sedov 64^3
@pytest.mark.parametrize("fused", [True, False])
def test_sedov64_golden(gpu, fused):
    from ramses_amd import ics
    g = np.load(os.path.join(GOLDEN, "sedov64.npz"))
    ra.configure(gamma=1.4, riemann="hllc", slope_type=1, courant_factor=0.8)
    # the device generator of the benchmark builds the fixture's initial state
    assert np.array_equal(ics.sedov(64).cpu().numpy(), g["u0"])
    lvl = ra.UniformLevel(g["u0"])
    for s in range(1, 21):
        dt = float(lvl.step(fused=fused))
        assert dt == g["dt"][s - 1], s
        if f"u_step{s}" in g:
            assert np.array_equal(lvl.u.cpu().numpy(), g[f"u_step{s}"]), s
    ra.configure()


# ------------------------------------------------- 2^20 synthetic oct batches
@pytest.mark.parametrize("riemann", ["hllc", "llf"])
def test_unsplit_full_launch_strided_vs_oracle(gpu, oracle_lib, riemann):
    """One launch over all 32768 batches of the benchmark's input; every 64th
    batch and the last compared bitwise with the oracle (FP64), and the
    FP32-mixed launch's fluxes within 1e-5 relative L1 of FP64."""
    import torch
    from ramses_amd import ics
    dx = 1.0 / 128
    uin, dt = ics.oct_batches(1 << 20)
    G = uin.shape[0]
    sample = torch.tensor(list(range(0, G, 64)) + [G - 1], device=uin.device)
    print(f"full launch: {G} batches, dt={dt!r}, {riemann}", flush=True)
    out = {}
    for prec in ("fp64", "fp32"):
        ra.configure(**dict(TURB_HYDRO, riemann=riemann), precision=prec)
        flux = torch.zeros((G,) + ra.flux_shape(32), dtype=torch.float64, device=uin.device)
        tmp = torch.zeros((G,) + ra.tmp_shape(32), dtype=torch.float64, device=uin.device)
        ra.unsplit_device(uin, None, flux, tmp, dx, dt)
        torch.cuda.synchronize()
        out[prec] = (flux[sample].cpu().numpy(), tmp[sample].cpu().numpy())
        del flux, tmp
    u_s = np.ascontiguousarray(uin[sample].cpu().numpy())
    del uin
    oracle_lib.set_params(**dict(TURB_HYDRO, riemann=riemann))
    f_ref = np.zeros((len(sample),) + ra.flux_shape(32))
    t_ref = np.zeros((len(sample),) + ra.tmp_shape(32))
    oracle_lib.unsplit_batched(u_s, None, f_ref, t_ref, dx, dt)
    assert np.array_equal(out["fp64"][0], f_ref)
    assert np.array_equal(out["fp64"][1], t_ref)
    for v in range(5):
        assert _rel_l1(out["fp32"][0][:, :, v], f_ref[:, :, v]) < REL_L1_FP32, v
    ra.configure()


@pytest.mark.parametrize("last", [1, 17, 31])
def test_fp32_staged_ragged_last_batch(gpu, oracle_lib, last):
    """The FP32-mixed kernel (nvector 32) stages the primitives of half a
    batch per workgroup in LDS, so a last batch of 1, 17 or 31 octs leaves a
    workgroup with no live lane or a partly live one.  Its live lanes give
    the same bits as the same octs in a full batch, its dead lanes keep the
    caller's values, and the result is within 1e-5 of the FP64 oracle."""
    import torch
    rng = np.random.default_rng(11 + last)
    uin = random_stencils(rng, 5)
    dx, dt = 1.0 / 128, 2e-4
    ra.configure(**TURB_HYDRO, precision="fp32")
    d_u = torch.from_numpy(uin).cuda()
    res = {}
    for ngl in (32, last):
        f = torch.full((5,) + ra.flux_shape(32), 7.0, dtype=torch.float64, device="cuda")
        t = torch.full((5,) + ra.tmp_shape(32), 7.0, dtype=torch.float64, device="cuda")
        ra.unsplit_device(d_u, None, f, t, dx, dt, ngrid_last=ngl)
        res[ngl] = (f.cpu().numpy(), t.cpu().numpy())
    full, rag = res[32], res[last]
    for a, b in zip(full, rag):
        assert np.array_equal(a[:4], b[:4])
        assert np.array_equal(a[4, ..., :last], b[4, ..., :last])
        assert np.all(b[4, ..., last:] == 7.0)
    oracle_lib.set_params(**TURB_HYDRO)
    f_ref = np.full((5,) + ra.flux_shape(32), 7.0)
    t_ref = np.full((5,) + ra.tmp_shape(32), 7.0)
    oracle_lib.unsplit_batched(uin, None, f_ref, t_ref, dx, dt, ngrid_last=last)
    live = np.ones(f_ref.shape, dtype=bool)
    live[4, ..., last:] = False
    live &= defined_face_mask()[None]
    for v in range(5):
        assert _rel_l1(rag[0][:, :, v][live[:, :, v]], f_ref[:, :, v][live[:, :, v]]) < REL_L1_FP32, v
    ra.configure()


# ------------------------------------------------------- Mach-5 512^3 field
@needs_ref
def test_c4_field_vs_reference_routines_64(gpu):
    """The 512^3 workload's field generator at 64^3: three level steps
    bitwise against the reference's own courant_fine and godunov_fine."""
    from oracle.ref_uniform import ReferenceUniformLevel
    from ramses_amd import ics
    u0 = ics.turbulent_host(64, 5.0, gamma=TURB_HYDRO["gamma"], rho_sigma=0.5, seed=4)
    R = ReferenceUniformLevel(64, TURB_HYDRO["gamma"], riemann="hllc", slope_type=1, courant_factor=0.5)
    R.load(u0)
    ra.configure(**TURB_HYDRO)
    lvl = ra.UniformLevel(u0)
    for s in range(3):
        dt_ref = R.step()
        assert float(lvl.step()) == dt_ref, s
        assert np.array_equal(lvl.u.cpu().numpy(), R.state()), s
    R.release()
    ra.configure()


def test_c4_field_vs_oracle_128(gpu, oracle_lib):
    from ramses_amd import ics
    u0 = ics.turbulent_host(128, 5.0, gamma=TURB_HYDRO["gamma"], rho_sigma=0.5, seed=4)
    ra.configure(**TURB_HYDRO)
    oracle_lib.set_params(**TURB_HYDRO)
    lvl = ra.UniformLevel(u0)
    ref = u0
    for s in range(3):
        ref, dt_ref = oracle_lib.uniform_step(ref, 1.0 / 128)
        assert float(lvl.step()) == dt_ref, s
        assert np.array_equal(lvl.u.cpu().numpy(), ref), s
    ra.configure()


@pytest.fixture(scope="module")
def c4_512():
    from ramses_amd import ics
    return ics.turbulent_host(512, 5.0, gamma=TURB_HYDRO["gamma"], rho_sigma=0.5, seed=4)


def test_c4_512_full_size(gpu, oracle_lib, c4_512):
    """10 FP64 steps of the 512^3 workload: mass, momentum and energy
    conserved to 1e-12 of sum |u|, two runs bitwise identical, the first dt
    equal to the oracle's courant_fine over all 134M cells, the interior of
    a 128^3 sub-box bitwise equal to the oracle stepped with the same dts
    (3 steps; the sub-box is periodic for the oracle, so cells within 4 per
    step of its faces are left out), and FP32-mixed within 1e-5 of FP64."""
    import torch
    n, dx, steps = 512, 1.0 / 512, 10
    u0 = c4_512
    ra.configure(**TURB_HYDRO)
    oracle_lib.set_params(**TURB_HYDRO)
    start = min(1.0 / 1e-10, TURB_HYDRO["courant_factor"] * dx / 1e-10)
    dt0_ref = oracle_lib.uniform_courant(u0, dx, start)[0]
    sb = slice(192, 320)
    lvl = ra.UniformLevel(u0)
    before = lvl.u.sum(dim=(1, 2, 3))
    scale = lvl.u.abs().sum(dim=(1, 2, 3))
    dts, subs = [], []
    for k in range(steps):
        dts.append(float(lvl.step()))
        if k < 3:
            subs.append(lvl.u[:, sb, sb, sb].cpu().numpy())
    print(f"512^3 dts: {dts[:3]}", flush=True)
    assert dts[0] == dt0_ref
    after = lvl.u.sum(dim=(1, 2, 3))
    drift = ((after - before).abs() / scale).cpu().numpy()
    assert np.all(drift <= 1e-12), drift
    final = lvl.u.clone()
    del lvl
    again = ra.UniformLevel(u0)
    for _ in range(steps):
        again.step()
    assert torch.equal(again.u, final)
    del again
    ref = np.ascontiguousarray(u0[:, sb, sb, sb])
    for k in range(3):
        ref = oracle_lib.uniform_godunov(ref, dx, dts[k])
        m = 4 * (k + 1)
        inner = slice(m, 128 - m)
        assert np.array_equal(ref[:, inner, inner, inner], subs[k][:, inner, inner, inner]), k
    ra.configure(**TURB_HYDRO, precision="fp32")
    lvl32 = ra.UniformLevel(u0)
    for _ in range(steps):
        lvl32.step()
    for v in range(5):
        rel = float((lvl32.u[v] - final[v]).abs().sum() / final[v].abs().sum())
        assert rel < REL_L1_FP32, (v, rel)
    del lvl32, final
    ra.configure()


# ------------------------------------------------------------------- serving
SERVE_HYDRO = dict(gamma=5.0 / 3.0, riemann="hllc", slope_type=1, courant_factor=0.5)
C3_REQUESTS = [0, 1, 2, 3, 4, 7, 10, 11]  # indices into the benchmark's request stream


def _c3_sides(count=12):
    """Request sides as the serving benchmark draws them (default_rng(3))."""
    rng = np.random.default_rng(3)
    return [int(16 * 2 ** rng.integers(0, 4)) for _ in range(count)]


def _oracle_serve(O, u0, n, t_end, boxlen=1.0, smallc=1e-10):
    """The oracle's level step repeated until t >= t_end with each dt capped
    at t_end - t, as the server steps a request."""
    dx = boxlen / n
    start = min(boxlen / smallc, SERVE_HYDRO["courant_factor"] * dx / smallc)
    dtc = O.uniform_courant(u0, dx, start)[0]
    u, t, steps = u0, 0.0, 0
    while not t >= t_end:
        dt = min(dtc, t_end - t)
        u = O.uniform_godunov(u, dx, dt)
        dtc = O.uniform_courant(u, dx, start)[0]
        t = t + dt
        steps += 1
    return u, steps


@pytest.fixture(scope="module")
def c3_cases(gpu):
    from ramses_amd import ics
    sides = _c3_sides()
    O = oracle.Oracle()
    O.set_params(**SERVE_HYDRO)
    cases = []
    for i in C3_REQUESTS:
        n = sides[i]
        t_end = 0.02 if n == 128 else 0.1
        u0 = ics.turbulent(n, 2.0, gamma=SERVE_HYDRO["gamma"], rho_sigma=0.0, seed=1000 + i, generator="device")
        host = u0.cpu().numpy()
        ref, steps = _oracle_serve(O, host, n, t_end)
        print(f"request {i}: {n}^3 to t={t_end}, {steps} oracle steps", flush=True)
        cases.append(dict(i=i, n=n, t_end=t_end, u0=u0, ref=ref, steps=steps))
    assert {c["n"] for c in cases} == {16, 32, 64, 128}
    return cases


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_c3_served_vs_oracle(gpu, c3_cases, precision):
    from ramses_amd.serve import HydroServer, Request
    ra.configure(**SERVE_HYDRO, precision=precision)
    reqs = [Request(c["i"], c["n"], c["t_end"], arrival=0.001 * k, ic=lambda n, u=c["u0"]: u.clone(),
                    keep_result=True) for k, c in enumerate(c3_cases)]
    srv = HydroServer(hbm_budget=1 << 30, max_slots=4)
    srv.run(reqs)
    assert srv.max_active >= 2, "requests did not overlap"
    for r, c in zip(reqs, c3_cases):
        got = r.result.cpu().numpy()
        if precision == "fp64":
            assert r.steps == c["steps"], (c["i"], r.steps, c["steps"])
            assert np.array_equal(got, c["ref"]), c["i"]
        else:
            for v in range(5):
                rel = _rel_l1(got[v], c["ref"][v])
                assert rel < REL_L1_FP32, (c["i"], v, rel)
    ra.configure()
