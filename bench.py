"""Benchmark: the RAMSES unsplit MUSCL-Hancock level step on one MI355X.

Default workload (one "step" = one level step of the resident uniform
level: the dt reset, then one launch doing godunov_fine over every oct, the
conservative update and the courant_fine reduction of the new state):
  512^3 periodic box (134M cells, NVAR=5), a seeded Gaussian random
  velocity field with P(k) ~ k^-4 at rms Mach 5, rho lognormal (sigma 0.5)
  with P = rho/gamma, gamma = 5/3, riemann = hllc, minmod slopes, CFL 0.5.
  The initial state comes from numpy default_rng(4) noise filtered with
  scipy's FFT (ramses_amd.ics.turbulent_host) and is resident in HBM before
  the timed region.

Reported: cell-updates/s of the FP64 path (value) and of the FP32-mixed
path, per-step latency p50/p99, the relative L1 error of FP32-mixed against
FP64 after the timed steps, the level kernel's roofline against MI355X HBM
(8 TB/s) with algorithmic bytes of 80 B per cell-update (read U^n, write
U^n+1, 5 doubles each), the HBM traffic per launch measured with rocprofv3
counters (read from the profile summary of the same kernel and launch
shape, null when none matches), and the reference's own CPU level step
(oracle/_ref: courant_fine + godunov_fine, the Fortran compiled from the
reference sources) run as one process per host core on 128^3 levels of the
same generator, its per-cell rate extrapolated to the 512^3 workload.

Other workloads (--workload): unsplit (2^20 synthetic oct batches through the
batched unsplit boundary), sedov64, turb256, serve, sedov_amr.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload W]
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

NOCT = 1 << 20
NV = 32
GAMMA = 5.0 / 3.0
DX = 1.0 / 128
CFL = 0.5
BYTES_PER_OCT = 216 * 5 * 8 + 36 * 5 * 8 + 36 * 2 * 8  # 10,656
HBM_PEAK = 8.0e12


NEEDED_BYTES_PER_OCT = 80 * 5 * 8 + 36 * 5 * 8 + 36 * 2 * 8  # 5,216: the 80 stencil cells a face needs


WARM_S = 0.5  # seconds of untimed launches of each measured kernel before its timed steps


def warm_device(torch, seconds=0.3):
    """Untimed: keep the GPU busy for a moment before any measurement.  The
    first tens of milliseconds of work after an idle period run slower (one
    FP64 unsplit launch took 2.2 ms at first and 1.75 ms after 25 ms of
    work), so without this the first timed variant would depend on whether
    an autotuning pass happened to run before it.  A plain torch kernel is
    used so that the warm-up does not show under the measured kernels' names
    in a kernel trace."""
    x = torch.zeros(1 << 26, device="cuda")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    while time.perf_counter() - t0 < seconds:
        for _ in range(16):
            x.add_(1.0)
        torch.cuda.synchronize()
    del x


def time_variant(torch, ra, uin, flux, tmp, dt, riemann, precision, steps, warmup, tune=True, block=0):
    from ramses_amd import autotune
    ra.configure(gamma=GAMMA, riemann=riemann, slope_type=1, courant_factor=CFL, precision=precision)
    stream = torch.cuda.current_stream()
    tuned = None
    if precision == "fp32":  # nvector 32 runs the LDS-staged kernel: fixed 192-thread workgroups
        tuned = (192, None)
    elif not tune:
        autotune._set(autotune.UNSPLIT_BLOCK, block)
        tuned = (block or 256, None)
    else:  # untimed: pick the workgroup size on these inputs
        tuned = autotune.tune_unsplit(lambda: ra.unsplit_device(uin, None, flux, tmp, DX, dt, stream=stream),
                                      key=("unsplit", uin.shape[0], precision, riemann))
    # Untimed warm-up: at least `warmup` launches and at least WARM_S seconds.
    # The FP64 launches take tens of launches to settle (2.0 ms down to
    # 1.75-1.8 ms on the default workload; FP32 is flat), so a fixed short
    # warm-up would time the settling.
    k, t0 = 0, time.perf_counter()
    while k < warmup or time.perf_counter() - t0 < WARM_S:
        ra.unsplit_device(uin, None, flux, tmp, DX, dt, stream=stream)
        k += 1
        if k % 16 == 0:
            torch.cuda.synchronize()
    torch.cuda.synchronize()
    starts = [torch.cuda.Event(enable_timing=True) for _ in range(steps)]
    ends = [torch.cuda.Event(enable_timing=True) for _ in range(steps)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(steps):
        starts[k].record(stream)
        ra.unsplit_device(uin, None, flux, tmp, DX, dt, stream=stream)
        ends[k].record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ms = sorted(s.elapsed_time(e) for s, e in zip(starts, ends))
    avg = sum(ms) / len(ms)
    pct = lambda q: ms[min(len(ms) - 1, int(round(q * (len(ms) - 1))))]  # noqa: E731
    return dict(wall_s=wall, kernel_ms_avg=avg, p50_ms=pct(0.5), p99_ms=pct(0.99),
                ms_per_step=1e3 * wall / steps, block=tuned[0] if tuned else None)


def unsplit_cpu_baseline(dt, budget_s=12.0):
    """The reference's own unsplit (oracle/_ref) on one core over the first
    32 batches of the benchmark's own input (the first chunk of
    ics.oct_batches_host with seed 1) with its dt, a bounded sample."""
    import numpy as np
    import oracle
    from ramses_amd import ics
    groups = ics.oct_batches_host_chunk(np.random.default_rng(1), 1024, NV, GAMMA, DX, CFL)[0][:32]
    if oracle.reference_available():
        impl, kind = oracle.Reference(), "reference"
    else:
        impl, kind = oracle.Oracle(), "port"
    impl.set_params(gamma=GAMMA, riemann="hllc", slope_type=1, courant_factor=CFL)
    flux = np.zeros((3, 5, 3, 3, 3, NV))
    tmp = np.zeros((3, 2, 3, 3, 3, NV))
    n = 0
    t0 = time.perf_counter()
    while True:
        impl.unsplit(np.ascontiguousarray(groups[n % 32]), None, flux, tmp, DX, dt, NV)
        n += 1
        el = time.perf_counter() - t0
        if el > budget_s and n >= 32:
            break
    cells = 8 * NV * n
    return {"value": cells / el, "unit": "cell-updates/s", "cores": 1, "kind": kind,
            "sample": f"{n} unsplit calls of {NV} octs ({cells} cell-updates) in {el:.1f} s, hllc minmod, "
                      f"the first 32 batches of the benchmark's input, cycled; "
                      f"{'flang -O2 build of hydro/umuscl.f90' if kind == 'reference' else 'C restatement'}"}


def progress(msg):
    """One line on stderr per finished stage (set-up stages take tens of
    seconds at 512^3)."""
    print(f"bench: {msg}", file=sys.stderr, flush=True)


def host_cores():
    """Host cores this run may use: the scheduler's affinity set, at most 16
    (the host share of one GPU on the machines this runs on)."""
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(16, n))


def level_cpu_baseline(workload, n_sub, budget_s=12.0, seed=4, target=""):
    """The reference's own level step (courant_fine, set_unew, godunov_fine,
    set_uold; oracle/_ref) on n_sub^3 levels of the workload's generator, one
    process per host core started together (oracle/cpu_bench.py); the
    reference has no threads of its own and this image has no MPI, so
    concurrent independent processes are how it uses several cores.  The
    value is the sum of the processes' cell-update rates."""
    import subprocess
    cores = host_cores()
    env = dict(os.environ, OMP_NUM_THREADS="1")
    cmd = [sys.executable, "-m", "oracle.cpu_bench", "--workload", workload, "--n", str(n_sub),
           "--budget", str(budget_s), "--seed", str(seed)]
    procs = [subprocess.Popen(cmd, cwd=ROOT, env=env, stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True)
             for _ in range(cores)]
    res = []
    for p in procs:
        out, err = p.communicate(timeout=600)
        if p.returncode != 0:
            raise RuntimeError(f"cpu baseline worker failed ({p.returncode}): {err[-2000:]}")
        res.append(json.loads(out.strip().splitlines()[-1]))
    rate = sum(r["cells"] / r["seconds"] for r in res)
    kind = res[0]["kind"]
    steps = sum(r["steps"] for r in res)
    cells = sum(r["cells"] for r in res)
    what = ("the reference's courant_fine + godunov_fine (flang -O2 build of the reference sources)"
            if kind == "reference" else "the C restatement of the same level step")
    return {"value": rate, "unit": "cell-updates/s", "cores": cores, "kind": kind,
            "sample": f"{cores} concurrent single-core processes, each {what} on a {n_sub}^3 periodic level of "
                      f"this workload's generator for >= {budget_s:.0f} s: {steps} level steps, {cells} "
                      f"cell-updates; rate per cell extrapolated to {target or 'the workload'}"}


def profiled_traffic(workload, precision, tile, tz):
    """HBM bytes per launch of the level kernel from the committed rocprofv3
    counter summary (profiles/uniform_godunov_<workload>.md): lines
    'traffic <workload> <precision> tile=<t> tz=<z> fetch_kib=<F>
    write_kib=<W> bytes=<B>', where B = 2.0 * F KiB + 1.0 * W KiB (the
    scale factors a device copy of known size gives these counters).  None
    unless a line matches this kernel's precision and launch shape."""
    import re
    path = os.path.join(ROOT, "profiles", f"uniform_godunov_{workload}.md")
    if not os.path.exists(path):
        return None
    pat = re.compile(r"^traffic (\S+) (fp64|fp32) tile=(\d+) tz=(\d+) fetch_kib=(\d+) write_kib=(\d+) bytes=(\d+)\s*$")
    for line in open(path):
        m = pat.match(line.strip())
        if m and m.group(1) == workload and m.group(2) == precision and int(m.group(3)) == tile \
                and int(m.group(4)) == tz:
            return int(m.group(7))
    return None


def profiled_unsplit_traffic(variant, block):
    """HBM bytes per launch of the batched unsplit kernel from the committed
    counter summary (profiles/unsplit_rows_2p20_octs.md): lines 'traffic
    unsplit <variant> block=<b> fetch_kib=<F> write_kib=<W> bytes=<B>', with
    B as in profiled_traffic.  None unless a line matches the variant and
    its workgroup size."""
    import re
    path = os.path.join(ROOT, "profiles", "unsplit_rows_2p20_octs.md")
    if not os.path.exists(path):
        return None
    pat = re.compile(r"^traffic unsplit (\S+) block=(\d+) fetch_kib=(\d+) write_kib=(\d+) bytes=(\d+)\s*$")
    for line in open(path):
        m = pat.match(line.strip())
        if m and m.group(1) == variant and int(m.group(2)) == block:
            return int(m.group(5))
    return None


LEVEL_WORKLOADS = {
    # 3D Sedov, uniform 64^3 periodic, gamma=1.4, hllc, minmod, CFL 0.8
    "sedov64": dict(n=64, gamma=1.4, courant=0.8, ic="sedov", cpu="sedov", cpu_n=64,
                    name="3D Sedov blast, uniform 64^3 periodic (levelmin=levelmax=6), gamma=1.4, hllc, minmod, "
                         "CFL 0.8"),
    # 512^3 periodic Mach-5 k^-4 velocity field, lognormal rho (sigma 0.5), P=rho/gamma
    "turb512": dict(n=512, gamma=5.0 / 3.0, courant=0.5, ic="turb", mach=5.0, rho_sigma=0.5, seed=4, cpu="turb",
                    cpu_n=128,
                    name="512^3 periodic (134M cells, NVAR=5), Mach-5 k^-4 velocity field, rho lognormal(0.5), "
                         "P=rho/gamma, gamma=5/3, hllc, minmod, CFL 0.5"),
    "turb256": dict(n=256, gamma=5.0 / 3.0, courant=0.5, ic="turb", mach=5.0, rho_sigma=0.5, seed=4, cpu="turb",
                    cpu_n=128,
                    name="256^3 periodic, Mach-5 k^-4 velocity field, rho lognormal(0.5), P=rho/gamma, "
                         "gamma=5/3, hllc, minmod, CFL 0.5"),
}


def level_bench(torch, ra, args):
    """Resident uniform level: untimed autotuning of the tile shape and
    warm-up steps, then K timed steps from the initial state in FP64 and in
    FP32-mixed (one step = the dt reset and the fused godunov + courant
    launch).  HIP events on the level's stream bracket each update launch;
    the wall clock brackets the K steps."""
    from ramses_amd import autotune, ics
    w = LEVEL_WORKLOADS[args.workload]
    n = w["n"]
    if w["ic"] == "sedov":
        u0 = ics.sedov(n, gamma=w["gamma"])
        data = "synthetic: Sedov point blast (region_condinit of namelist/sedov3d.nml)"
    else:
        u0 = ics.turbulent(n, w["mach"], gamma=w["gamma"], rho_sigma=w["rho_sigma"], seed=w["seed"])
        data = f"synthetic: numpy default_rng({w['seed']}) noise, k^-4 filtered (scipy FFT), resident in HBM"
    torch.cuda.synchronize()
    progress(f"initial state {n}^3 resident")
    cells = n ** 3
    out, states = {}, {}
    for prec in ("fp64", "fp32"):
        ra.configure(gamma=w["gamma"], riemann="hllc", slope_type=1, courant_factor=w["courant"],
                     precision=prec)
        lvl = ra.UniformLevel(u0)
        if args.no_tune:
            autotune._set(autotune.LEVEL_TILE, args.level_tile)
            autotune._set(autotune.LEVEL_TZ, args.level_tz)
            tuned = ((args.level_tile, args.level_tz), None)
        else:
            tuned = autotune.tune_level(lvl.update_only, n, key=("level", n, prec))  # untimed
        k, t0 = 0, time.perf_counter()
        while k < args.warmup or time.perf_counter() - t0 < WARM_S:  # untimed
            lvl.step()
            k += 1
        torch.cuda.synchronize()
        lvl = ra.UniformLevel(u0)  # the timed run starts from the initial state
        lvl.courant()
        ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
              for _ in range(args.steps)]
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for k in range(args.steps):
            lvl.step(events=ev[k])
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        kms = sorted(a.elapsed_time(b) for a, b in ev)
        kavg = sum(kms) / len(kms)
        pct = lambda q: kms[min(len(kms) - 1, int(round(q * (len(kms) - 1))))]  # noqa: E731
        ach = 80.0 * cells / (kavg * 1e-3)
        tile, tz = tuned[0]
        traffic = profiled_traffic(args.workload, prec, tile, tz)
        out[prec] = {"cell_updates_per_s": cells * args.steps / wall, "ms_per_step": 1e3 * wall / args.steps,
                     "tile": tile, "tz": tz, "kernel_ms_avg": kavg, "kernel_ms_p50": pct(0.5),
                     "kernel_ms_p99": pct(0.99),
                     "roofline": {"bound": "hbm", "achieved": ach / 1e9, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                                  "frac": ach / HBM_PEAK, "traffic": traffic}}
        states[prec] = lvl.u.clone()
        del lvl
        progress(f"{prec}: {kavg:.3f} ms per launch, tile {tile} tz {tz}")
    err = {}
    names = ["rho", "mx", "my", "mz", "E"]
    for v in range(5):
        a, b = states["fp32"][v], states["fp64"][v]
        err[names[v]] = float((a - b).abs().sum() / b.abs().sum())
    del states
    ra.configure()
    head = out["fp64"]
    line = {
        "metric": "cell-updates/s FP64 vs FP32-mixed, % HBM roofline; p50/p99 solve latency",
        "value": head["cell_updates_per_s"], "unit": "cell-updates/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": head["ms_per_step"], "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": data,
        "config": {"workload": w["name"], "cells": cells, "steps_timed": args.steps, "parallelism": "1 GPU"},
        "latency_ms": {"p50": head["kernel_ms_p50"], "p99": head["kernel_ms_p99"]},
        "roofline": head["roofline"],
        "fp32_mixed_value": out["fp32"]["cell_updates_per_s"],
        "fp32_over_fp64": out["fp32"]["cell_updates_per_s"] / head["cell_updates_per_s"],
        "fp32_rel_l1_vs_fp64": err,
        "variants": out,
    }
    if not args.no_cpu:
        line["cpu_baseline"] = level_cpu_baseline(w["cpu"], w["cpu_n"], seed=w.get("seed", 4),
                                                  target=f"{n}^3" if w["cpu_n"] != n else "")
    return line


AMR_WORKLOAD = dict(levelmin=6, levelmax=10, gamma=1.4, courant=0.8, err_grad_d=0.1, err_grad_p=0.1,
                    fine_steps=50, ngridmax=1 << 21,
                    name="3D Sedov blast, levelmin=6 levelmax=10, gradient refinement err_grad_d=err_grad_p=0.1, "
                         "interpol_type=1, gamma=1.4, hllc, minmod, CFL 0.8, nsubcycle=2, 50 fine steps")


def amr_bench(torch, ra, args):
    """Adaptive oct-tree Sedov run: initial mesh from init_refine on the GPU
    (untimed), then coarse steps of amr_step (refine, courant, godunov,
    upload, flag on every level, subcycled) until `fine_steps` finest-level
    steps; cell-updates = sum over level steps of 8 * octs."""
    from ramses_amd.amr import AmrSolver, configure_interpolation, configure_refinement
    from ramses_amd.ics import sedov_regions
    from ramses_amd.tree import OctTree
    w = AMR_WORKLOAD
    nfine = args.fine_steps or w["fine_steps"]
    lmin, lmax = w["levelmin"], args.levelmax or w["levelmax"]
    tree = OctTree(lmin, lmin, extra_grids=w["ngridmax"] - sum(8 ** k for k in range(lmin)))
    ic = lambda x, dx: sedov_regions(x, dx, gamma=w["gamma"])  # noqa: E731
    out, finals, tuned = {}, {}, {}
    for prec in ("fp64", "fp32"):
        ra.configure(gamma=w["gamma"], riemann="hllc", slope_type=1, courant_factor=w["courant"], precision=prec)
        configure_interpolation(1, 0)
        configure_refinement(err_grad_d=w["err_grad_d"], err_grad_p=w["err_grad_p"])
        for warm in (True, False):
            s = AmrSolver(tree, None, refine=True, nlevelmax=lmax, graph=args.amr_mode == "graph")
            s.init_refine(ic)
            init_octs = {l: s.noct(l) for l in range(lmin, lmax + 1)}
            s.capture()  # graph mode: the one-time capture stays outside the timed region
            torch.cuda.synchronize()
            if warm:
                if s.graph and not args.no_tune:
                    from ramses_amd.autotune import tune_amr_graph
                    tuned[prec] = tune_amr_graph(s, ("amr_graph", lmin, lmax, prec))[0]
                for _ in range(max(1, args.warmup // 3)):
                    s.coarse_step()
                torch.cuda.synchronize()
                del s
                continue
            t0 = time.perf_counter()
            ncoarse = 0
            per_step = []
            while s.fine_steps < nfine:
                a = time.perf_counter()
                s.coarse_step()
                torch.cuda.synchronize()
                per_step.append(time.perf_counter() - a)
                ncoarse += 1
            wall = time.perf_counter() - t0
            s.check()
        cu = s.cell_updates
        ach = 80.0 * cu / wall
        out[prec] = {"cell_updates_per_s": cu / wall, "wall_s": wall, "coarse_steps": ncoarse,
                     "fine_steps": s.fine_steps, "cell_updates": cu,
                     "ms_per_coarse_step": 1e3 * wall / ncoarse,
                     "ms_each_coarse_step": [round(1e3 * x, 3) for x in per_step],
                     "octs_initial": init_octs, "octs_final": {l: s.noct(l) for l in range(lmin, lmax + 1)},
                     "roofline": {"bound": "hbm", "achieved": ach / 1e9, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                                  "frac": ach / HBM_PEAK, "traffic": None}}
        finals[prec] = (s.son.clone(), s.uold.clone(), float(s.time.item()))
        del s
    cpu = None if args.no_cpu else amr_cpu_baseline(w, lmin, lmax)
    err = None
    same_mesh = bool(torch.equal(finals["fp64"][0], finals["fp32"][0]))
    if same_mesh:
        a, b = finals["fp32"][1], finals["fp64"][1]
        leaf = (finals["fp64"][0] == 0)
        err = {n: float(((a[v] - b[v]).abs() * leaf).sum() / ((b[v]).abs() * leaf).sum().clamp_min(1e-300))
               for v, n in enumerate(["rho", "mx", "my", "mz", "E"])}
    ra.configure()
    configure_refinement()
    head = out["fp64"]
    return {
        "metric": "cell-updates/s FP64 vs FP32-mixed, % HBM roofline; p50/p99 solve latency",
        "value": head["cell_updates_per_s"], "unit": "cell-updates/s", "n_gpus": 1, "steps": head["coarse_steps"],
        "warmup": args.warmup, "ms_per_step": head["ms_per_coarse_step"], "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic (Sedov initial conditions)",
        "config": {"workload": w["name"], "levelmin": lmin, "levelmax": lmax, "fine_steps": nfine,
                   "parallelism": "1 GPU",
                   "mode": "captured hipGraph per coarse step" if args.amr_mode == "graph" else "direct launches"},
        "roofline": head["roofline"], "variants": out, "fp32_same_mesh": same_mesh,
        "autotuned_graph_blocks": tuned or None,
        "fp32_rel_l1_vs_fp64": err, "t_end": {k: v[2] for k, v in finals.items()},
        "fp32_over_fp64": out["fp32"]["cell_updates_per_s"] / head["cell_updates_per_s"],
        **({"cpu_baseline": cpu} if cpu else {}),
    }


def amr_cpu_baseline(w, lmin, lmax, budget_s=15.0):
    """The reference's own coarse step (oracle/_ref: adaptive_loop's body
    around amr_step(levelmin,1), with newdt_fine, courant_fine,
    godunov_fine, flag_fine, refine_fine and upload_fine as the reference
    calls them) on the same adaptive Sedov workload after the reference's
    init_refine sequence, one host core.  Cell-updates are counted as on
    the GPU (8 per oct per godunov_fine call), by a second, untimed pass of
    the same steps through the level routines one by one, which gives the
    same mesh bit for bit."""
    import numpy as np
    import oracle
    if not oracle.reference_available():
        return None
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from amr_cases import reference_amr_step, reference_init_refine
    from ramses_amd.tree import OctTree

    def setup():
        R = oracle.Reference()
        A = oracle.ReferenceAMR(R)
        R.set_params(gamma=w["gamma"], riemann="hllc", slope_type=1, courant_factor=w["courant"])
        tree = OctTree(lmin, lmin, extra_grids=w["ngridmax"] - sum(8 ** k for k in range(lmin)))
        A.install(tree, nlevelmax=lmax)
        A.set_interpol(1, 0)
        A.set_refine(lmin, err_grad_d=w["err_grad_d"], err_grad_p=w["err_grad_p"])
        xg = np.zeros((tree.ngridmax, 3))
        for (l, gx, gy, gz), i in tree.grid_at.items():
            xg[i - 1] = (np.array([gx, gy, gz]) + 0.5) * 0.5 ** (l - 1)
        A.set_xg(xg)
        A.set_sedov(1.0, 1e-5, 0.4, (0.5, 0.5, 0.5))
        reference_init_refine(A, lmin, lmax)
        A.flag_coarse()
        return A
    A = setup()
    steps = 0
    t0 = time.perf_counter()
    while time.perf_counter() - t0 < budget_s and steps < 8:
        A.coarse_step()
        steps += 1
    el = time.perf_counter() - t0
    A = setup()
    dt, dtold = [0.0] * (lmax + 2), [0.0] * (lmax + 2)
    stats = {"cells": 0}
    for _ in range(steps):
        reference_amr_step(A, lmin, lmax, dt, dtold, stats=stats)
    cells = stats["cells"]
    A.install(OctTree(1, 1), nlevelmax=1)  # release the reference's arrays
    return {"value": cells / el, "unit": "cell-updates/s", "cores": 1, "kind": "reference",
            "sample": f"{steps} coarse steps of the same adaptive Sedov run ({cells} cell-updates) in {el:.1f} s "
                      f"after init_refine; flang -O2 build of the reference's own amr_step driver and level "
                      f"routines"}


def serve_bench(torch, ra, args):
    """Continuous batching of independent periodic solves: sides log-uniform in
    {16,32,64,128} (seed 3), request i's velocity field a k^-4 Gaussian field at
    rms Mach 2 (seed 1000+i), rho=1, P=1/gamma, gamma=5/3, to t=0.1 sound
    crossing times, FP32-mixed, hllc, minmod.  Capacity is measured first with
    a closed batch; arrivals are then Poisson at `load` x capacity."""
    import numpy as np
    from ramses_amd import ics
    from ramses_amd.serve import HydroServer, Request
    ra.configure(gamma=5.0 / 3.0, riemann="hllc", slope_type=1, courant_factor=0.5, precision="fp32")
    rng = np.random.default_rng(3)
    sides = [int(16 * 2 ** rng.integers(0, 4)) for _ in range(args.requests)]

    def ic(i):
        # the device generator: a request's field is made inside the serving
        # loop, where host FFTs would stall admissions
        return lambda n: ics.turbulent(n, 2.0, gamma=5.0 / 3.0, rho_sigma=0.0, seed=1000 + i, generator="device")

    budget = 64 << 30
    # warm-up (untimed): FFT plans of the IC generator and the kernels
    warm = [Request(-1 - j, n, 0.01, 0.0, ic(j)) for j, n in enumerate((16, 32, 64, 128) * 2)]
    HydroServer(hbm_budget=budget, steps_per_tick=args.steps_per_tick).run(warm)
    torch.cuda.synchronize()
    # capacity: a closed batch of the first 128 solves, all present at t=0
    cap_reqs = [Request(i, sides[i], 0.1, 0.0, ic(i)) for i in range(min(128, args.requests))]
    srv = HydroServer(hbm_budget=budget, steps_per_tick=args.steps_per_tick)
    t0 = time.perf_counter()
    srv.run(cap_reqs)
    cap_wall = time.perf_counter() - t0
    capacity = len(cap_reqs) / cap_wall
    cap_cells = sum(r.n ** 3 * r.steps for r in cap_reqs) / cap_wall
    runs = {}
    for load in [float(x) for x in str(args.load).split(",")]:
        lam = load * capacity
        arr = np.cumsum(np.random.default_rng(30).exponential(1.0 / lam, args.requests))
        reqs = [Request(i, sides[i], 0.1, float(arr[i]), ic(i)) for i in range(args.requests)]
        srv = HydroServer(hbm_budget=budget, steps_per_tick=args.steps_per_tick)
        t0 = time.perf_counter()
        srv.run(reqs)
        wall = time.perf_counter() - t0
        lat = np.array(sorted(r.latency for r in reqs))
        cells = sum(r.n ** 3 * r.steps for r in reqs)
        runs[load] = {"arrival_rate_per_s": lam, "wall_s": wall, "cell_updates_per_s": cells / wall,
                      "latency_s": {"p50": float(np.percentile(lat, 50)), "p99": float(np.percentile(lat, 99)),
                                    "max": float(lat[-1])},
                      "ticks": srv.ticks, "max_in_flight": srv.max_active, "max_hbm_bytes": srv.max_used,
                      "total_steps": int(sum(r.steps for r in reqs))}
    head = runs[max(runs)]
    wall, cells = head["wall_s"], head["cell_updates_per_s"] * head["wall_s"]
    ra.configure()
    ach = 80.0 * cells / wall
    cpu = None if args.no_cpu else level_cpu_baseline("turb2", 64, seed=1000, target="the served mix of sizes")
    return {
        "metric": "cell-updates/s FP64 vs FP32-mixed, % HBM roofline; p50/p99 solve latency",
        "value": cells / wall, "unit": "cell-updates/s", "n_gpus": 1, "steps": head["ticks"], "warmup": 8,
        "ms_per_step": 1e3 * wall / max(1, head["ticks"]), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32-mixed", "data": "synthetic",
        "config": {"workload": f"serving: {args.requests} periodic solves, side log-uniform in {{16,32,64,128}}, "
                               "Mach-2 k^-4 field, to t=0.1, FP32-mixed, continuous batching",
                   "requests": args.requests, "parallelism": "1 GPU"},
        "latency_s": head["latency_s"],
        "capacity_solves_per_s": capacity, "capacity_cell_updates_per_s": cap_cells,
        "offered_load": max(runs), "runs_by_load": {str(k): v for k, v in runs.items()},
        "hbm_budget_bytes": budget,
        "steps_per_tick": args.steps_per_tick,
        # whole-job rate over the run's wall clock (admission, IC generation
        # and host bookkeeping included), 80 B per cell-update
        "roofline": {"bound": "hbm", "achieved": ach / 1e9, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                     "frac": ach / HBM_PEAK, "traffic": None},
        **({"cpu_baseline": cpu} if cpu else {}),
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--octs", type=int, default=NOCT)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--variants", default="fp64-hllc,fp32-hllc,fp64-llf,fp32-llf")
    ap.add_argument("--workload", default="turb512",
                    choices=["unsplit", "sedov64", "turb512", "turb256", "serve", "sedov_amr"],
                    help="turb512 (default: the 512^3 resident level), the batched unsplit boundary on 2^20 "
                         "octs, another resident-level workload, the serving loop or the adaptive Sedov run")
    ap.add_argument("--fine-steps", type=int, default=0, help="sedov_amr: finest-level steps (default 50)")
    ap.add_argument("--levelmax", type=int, default=0, help="sedov_amr: deepest level (default 10)")
    ap.add_argument("--amr-mode", default="graph", choices=["graph", "direct"],
                    help="sedov_amr: coarse steps replayed from a captured hipGraph, or launched directly")
    ap.add_argument("--requests", type=int, default=2000, help="serve: number of solves")
    ap.add_argument("--load", default="1.2,0.8", help="serve: offered loads / measured capacity (comma list)")
    ap.add_argument("--steps-per-tick", type=int, default=8, help="serve: level steps per scheduler tick")
    ap.add_argument("--no-tune", action="store_true", help="use the shapes given below instead of autotuning")
    ap.add_argument("--unsplit-block", type=int, default=0, help="with --no-tune: unsplit workgroup size")
    ap.add_argument("--level-tile", type=int, default=0, help="with --no-tune: 1=32x8, 2=16x8, 3=8x8, 4=32x16, 5=16x12")
    ap.add_argument("--level-tz", type=int, default=0, help="with --no-tune: z planes per workgroup")
    ap.add_argument("--calibrate-copy", type=int, default=0,
                    help="run only a device copy of this many MiB (for HBM counter scaling)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build measures one GPU only; sharding the path over several GPUs is "
              "designed but not built. Run with --gpus 1.", file=sys.stderr)
        sys.exit(2)

    import torch
    import ramses_amd as ra
    ra.load_library()
    if args.calibrate_copy:
        n = args.calibrate_copy * (1 << 20) // 8
        a = torch.ones(n, dtype=torch.float64, device="cuda")
        b = torch.empty_like(a)
        for _ in range(5):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"calibrate_copy_bytes": n * 8}))
        return

    warm_device(torch)
    if args.workload == "serve":
        print(json.dumps(serve_bench(torch, ra, args)), flush=True)
        return
    if args.workload == "sedov_amr":
        print(json.dumps(amr_bench(torch, ra, args)), flush=True)
        return
    if args.workload != "unsplit":
        print(json.dumps(level_bench(torch, ra, args)), flush=True)
        return

    noct = args.octs
    from ramses_amd import ics
    uin, dt = ics.oct_batches_host(noct, NV, seed=1, gamma=GAMMA, dx=DX, cfl=CFL, progress=progress)
    progress("oct batches resident")
    G = noct // NV
    flux = torch.zeros((G, 3, 5, 3, 3, 3, NV), dtype=torch.float64, device="cuda")
    tmp = torch.zeros((G, 3, 2, 3, 3, 3, NV), dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()

    results = {}
    for v in args.variants.split(","):
        prec, riemann = v.split("-")
        results[v] = time_variant(torch, ra, uin, flux, tmp, dt, riemann, prec, args.steps, args.warmup,
                                  tune=not args.no_tune, block=args.unsplit_block)
    ra.configure()

    cells = 8 * noct
    head_key = "fp64-hllc" if "fp64-hllc" in results else next(iter(results))
    head = results[head_key]
    alg_bytes = BYTES_PER_OCT * noct

    def roof(r, variant):
        ach = alg_bytes / (r["kernel_ms_avg"] * 1e-3)
        traffic = profiled_unsplit_traffic(variant, r["block"]) if noct == NOCT else None
        return {"bound": "hbm", "achieved": ach / 1e9, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                "frac": ach / HBM_PEAK, "traffic": traffic,
                # the same time against the bytes the kernel needs: the 80 of
                # 216 stencil cells that some face depends on, and the 36 faces
                "frac_needed_bytes": NEEDED_BYTES_PER_OCT * noct / (r["kernel_ms_avg"] * 1e-3) / HBM_PEAK}

    variants = {}
    for k, r in results.items():
        variants[k] = {"cell_updates_per_s": cells / (r["kernel_ms_avg"] * 1e-3),
                       "kernel_ms_avg": r["kernel_ms_avg"], "p50_ms": r["p50_ms"], "p99_ms": r["p99_ms"],
                       "roofline_frac": roof(r, k)["frac"], "roofline_frac_needed_bytes": roof(r, k)["frac_needed_bytes"],
                       "traffic": roof(r, k)["traffic"], "workgroup": r["block"]}
    line = {
        "metric": "cell-updates/s FP64 vs FP32-mixed, % HBM roofline; p50/p99 solve latency",
        "value": cells / (head["ms_per_step"] * 1e-3),
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": head["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic: numpy default_rng(1) stencils (ics.oct_batches_host), resident in HBM",
        "config": {"workload": "unsplit oct batches: 2^20 octs (8.4M cells) as 6x6x6 stencils, nvector=32, "
                               "NVAR=5, hllc, minmod, gamma=5/3, dx=1/128, dt at CFL 0.5",
                   "octs": noct, "nvector": NV, "riemann": "hllc", "slope_type": 1, "parallelism": "1 GPU"},
        "latency_ms": {"p50": head["p50_ms"], "p99": head["p99_ms"]},
        "roofline": roof(head, head_key),
        "variants": variants,
    }
    if "fp32-hllc" in variants:
        line["fp32_mixed_value"] = variants["fp32-hllc"]["cell_updates_per_s"]
        line["fp32_over_fp64"] = variants["fp32-hllc"]["cell_updates_per_s"] / variants["fp64-hllc"]["cell_updates_per_s"] \
            if "fp64-hllc" in variants else None
    if not args.no_cpu:
        cb = unsplit_cpu_baseline(dt)
        line["cpu_baseline"] = cb
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
