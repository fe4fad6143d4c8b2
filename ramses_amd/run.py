"""Run a RAMSES parameter file on the GPU: the hydro path of the reference's
main loop (amr/adaptive_loop.f90) driven by its own namelist.

    python -m ramses_amd.run params.nml [--outdir DIR] [--precision fp32]

Reads &RUN_PARAMS, &AMR_PARAMS, &INIT_PARAMS, &HYDRO_PARAMS, &REFINE_PARAMS
and &OUTPUT_PARAMS (ramses_amd.namelist), builds the initial mesh from the
regions as init_refine does, advances coarse steps (amr_step with
refinement and subcycling) until t reaches the last output time or
nstepmax coarse steps have run, and writes output_NNNNN/ snapshots when t
passes each tout (dump_all's trigger, amr/amr_step.f90) and at the end.
"""
import argparse
import sys
import time


def _configure(p, precision):
    import ramses_amd as ra
    from .amr import configure_interpolation, configure_refinement
    ra.configure(gamma=p.gamma, smallr=p.smallr, smallc=p.smallc, slope_type=p.slope_type,
                 slope_theta=p.slope_theta, riemann=p.riemann, courant_factor=p.courant_factor, precision=precision,
                 niter_riemann=p.niter_riemann)
    configure_interpolation(p.interpol_type, p.interpol_var)
    configure_refinement(err_grad_d=p.err_grad_d, err_grad_u=p.err_grad_u, err_grad_p=p.err_grad_p,
                         floor_d=p.floor_d, floor_u=p.floor_u, floor_p=p.floor_p, nexpand=p.nexpand)


def _ngridmax(p):
    return max(p.ngridmax, sum(8 ** k for k in range(p.levelmin)) + 1)


def build(params, precision="fp64", device="cuda", graph=False):
    """An AmrSolver with the mesh and state of the parameter file's
    initial conditions (init_refine + init_refine_2); graph: run its coarse
    steps as the captured hipGraph."""
    from .amr import AmrSolver
    from .ics import region_condinit
    from .tree import OctTree
    p = params
    _configure(p, precision)
    base = sum(8 ** k for k in range(p.levelmin))
    tree = OctTree(p.levelmin, p.levelmin, extra_grids=_ngridmax(p) - base)
    s = AmrSolver(tree, None, refine=p.levelmax > p.levelmin, nlevelmax=p.levelmax, nsubcycle=p.nsubcycle,
                  boxlen=p.boxlen, device=device, graph=graph)
    ic = lambda x, dx: region_condinit(x, dx, p.regions, gamma=p.gamma, smallr=p.smallr, smallc=p.smallc)  # noqa
    if s.adaptive:
        s.init_refine(ic)
    else:
        s.init_flow(ic)
    return s


def restart(params, outdir, precision="fp64", device="cuda", ordering=None, graph=False):
    """An AmrSolver and snapshot.RunInfo restarted from
    outdir/output_NNNNN with NNNNN = params.nrestart (init_amr and
    init_hydro's restart, AmrSolver.from_snapshot)."""
    from .amr import AmrSolver
    p = params
    _configure(p, precision)
    return AmrSolver.from_snapshot(outdir, p.nrestart, p.levelmin, nlevelmax=p.levelmax, ngridmax=_ngridmax(p),
                                   smallr=p.smallr, gamma=p.gamma, ordering=ordering, tout=p.tout,
                                   refine=p.levelmax > p.levelmin, nsubcycle=p.nsubcycle, device=device,
                                   graph=graph)


def run(path, outdir=".", precision="fp64", max_coarse=None, log=print, ordering="hilbert", bound_key=None,
        graph=False):
    """ordering / bound_key: the domain decomposition the snapshot records
    (one domain; the Hilbert key range by default).  With nrestart > 0 in
    &RUN_PARAMS the run continues from outdir/output_NNNNN instead, with
    the decomposition that snapshot records.  graph: each coarse step is
    replayed from a captured hipGraph (same results)."""
    from . import snapshot
    from .namelist import Params
    p = Params.from_namelist(path)
    if p.nrestart > 0:
        s, info = restart(p, outdir, precision, ordering=None, graph=graph)
    else:
        s = build(p, precision, graph=graph)
        info = snapshot.RunInfo(levelmin=p.levelmin, gamma=p.gamma, smallr=p.smallr, boxlen=p.boxlen,
                                noutput=len(p.tout), tout=tuple(p.tout), ordering=ordering, bound_key=bound_key)
    info.levelmin, info.gamma, info.smallr = p.levelmin, p.gamma, p.smallr
    t_end = p.tout[-1] if p.tout else float("inf")
    nmax = p.nstepmax
    t0 = time.perf_counter()
    outputs = []
    steps = 0
    while s.nstep_coarse < nmax and (max_coarse is None or steps < max_coarse):
        t = float(s.time.item())
        if t >= t_end:
            break
        s.coarse_step()
        s.check()
        steps += 1
        t = float(s.time.item())
        octs = {l: s.noct(l) for l in range(p.levelmin, p.levelmax + 1) if s.noct(l)}
        log(f"coarse step {s.nstep_coarse:6d}  t={t:.6e}  fine steps {s.fine_steps}  octs {octs}")
        while info.iout <= len(p.tout) and t >= p.tout[info.iout - 1]:
            outputs.append(s.write_output(outdir, info))
    if not outputs:
        outputs.append(s.write_output(outdir, info))
    log(f"done: {s.nstep_coarse} coarse steps, {s.cell_updates} cell updates in {time.perf_counter() - t0:.2f} s")
    return s, outputs


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("namelist")
    ap.add_argument("--outdir", default=".")
    ap.add_argument("--precision", default="fp64", choices=["fp64", "fp32"])
    ap.add_argument("--max-coarse", type=int, default=None, help="stop after this many coarse steps (of this run)")
    ap.add_argument("--graph", action="store_true", help="replay each coarse step from a captured hipGraph")
    a = ap.parse_args(argv)
    run(a.namelist, a.outdir, a.precision, a.max_coarse, graph=a.graph)


if __name__ == "__main__":
    sys.exit(main())
