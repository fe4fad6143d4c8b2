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


def build(params, precision="fp64", device="cuda"):
    """An AmrSolver with the mesh and state of the parameter file's
    initial conditions (init_refine + init_refine_2)."""
    import ramses_amd as ra
    from .amr import AmrSolver, configure_interpolation, configure_refinement
    from .ics import region_condinit
    from .tree import OctTree
    p = params
    ra.configure(gamma=p.gamma, smallr=p.smallr, smallc=p.smallc, slope_type=p.slope_type,
                 slope_theta=p.slope_theta, riemann=p.riemann, courant_factor=p.courant_factor, precision=precision)
    configure_interpolation(p.interpol_type, p.interpol_var)
    configure_refinement(err_grad_d=p.err_grad_d, err_grad_u=p.err_grad_u, err_grad_p=p.err_grad_p,
                         floor_d=p.floor_d, floor_u=p.floor_u, floor_p=p.floor_p, nexpand=p.nexpand)
    base = sum(8 ** k for k in range(p.levelmin))
    ngridmax = max(p.ngridmax, base + 1)
    tree = OctTree(p.levelmin, p.levelmin, extra_grids=ngridmax - base)
    s = AmrSolver(tree, None, refine=p.levelmax > p.levelmin, nlevelmax=p.levelmax, nsubcycle=p.nsubcycle,
                  boxlen=p.boxlen, device=device)
    ic = lambda x, dx: region_condinit(x, dx, p.regions, gamma=p.gamma, smallr=p.smallr, smallc=p.smallc)  # noqa
    if s.adaptive:
        s.init_refine(ic)
    else:
        s.init_flow(ic)
    return s


def run(path, outdir=".", precision="fp64", max_coarse=None, log=print, ordering="hilbert", bound_key=None):
    """ordering / bound_key: the domain decomposition the snapshot records
    (one domain; the Hilbert key range by default)."""
    from . import snapshot
    from .namelist import Params
    p = Params.from_namelist(path)
    s = build(p, precision)
    info = snapshot.RunInfo(levelmin=p.levelmin, gamma=p.gamma, smallr=p.smallr, boxlen=p.boxlen,
                            noutput=len(p.tout), tout=tuple(p.tout), ordering=ordering, bound_key=bound_key)
    t_end = p.tout[-1] if p.tout else float("inf")
    nmax = p.nstepmax if max_coarse is None else min(p.nstepmax, max_coarse)
    t0 = time.perf_counter()
    outputs = []
    while s.nstep_coarse < nmax:
        t = float(s.time.item())
        if t >= t_end:
            break
        s.coarse_step()
        s.check()
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
    ap.add_argument("--max-coarse", type=int, default=None, help="stop after this many coarse steps")
    a = ap.parse_args(argv)
    run(a.namelist, a.outdir, a.precision, a.max_coarse)


if __name__ == "__main__":
    sys.exit(main())
