"""Autotuner for the launch shapes of the library's kernels.

The reference tunes one thing, the compile-time vector length nvector
(amr/amr_parameters.f90:49-53).  Here the equivalent choices are the unsplit
workgroup size, the resident-level tile / z-chunk and the fixed grid of a
captured oct-tree coarse step, and they depend on the device and on the
problem size, so they are measured: every candidate is timed with HIP
events on the caller's stream (median of a few launches) on the real
inputs, the fastest is installed with ``ramses_tune`` and the choice is
remembered per (kernel, size, precision, solver) for the process.  (The
oct-group size nvector itself stays 32 on the oct tree: it fixes the order
of the coarse-cell refluxes, so changing it would change the results.)
"""
import ctypes

from . import _check, load_library

UNSPLIT_BLOCK, LEVEL_TILE, LEVEL_TZ, PERSIST_BLOCKS, AMR_FACE_OCTS = 1, 2, 3, 4, 5
_cache = {}


def _set(knob, value):
    lib = load_library()
    lib.ramses_tune.argtypes = [ctypes.c_int, ctypes.c_int]
    _check(lib.ramses_tune(knob, value), "ramses_tune")


def _time(fn, reps=5):
    import torch
    ts = []
    for _ in range(reps + 1):
        a = torch.cuda.Event(enable_timing=True)
        b = torch.cuda.Event(enable_timing=True)
        a.record()
        fn()
        b.record()
        torch.cuda.synchronize()
        ts.append(a.elapsed_time(b))
    ts = sorted(ts[1:]) if reps > 0 else ts
    return ts[len(ts) // 2]


def tune_unsplit(run, key, blocks=(128, 192, 256, 320, 384, 512, 768), verbose=False):
    """run(): one unsplit launch on the real inputs.  Returns (block, ms)."""
    if key in _cache:
        _set(UNSPLIT_BLOCK, _cache[key][0])
        return _cache[key]
    best = None
    for b in blocks:
        _set(UNSPLIT_BLOCK, b)
        ms = _time(run)
        if verbose:
            print(f"autotune unsplit block={b}: {ms:.3f} ms")
        if best is None or ms < best[1]:
            best = (b, ms)
    _set(UNSPLIT_BLOCK, best[0])
    _cache[key] = best
    return best


def tile_16x12_fits(n):
    """16x12 tiles cover rows [0, m) for the largest multiple m of 12 that
    leaves a multiple of 8 rows, which 16x8 tiles cover (n a multiple of 16)."""
    return n % 16 == 0 and any((n - m) % 8 == 0 for m in range(n // 12 * 12, 11, -12))


def tune_level(run, n, key, verbose=False):
    """run(): one level update on the real state.  Returns ((tile, tz), ms)."""
    if key in _cache:
        (tile, tz), _ = _cache[key]
        _set(LEVEL_TILE, tile)
        _set(LEVEL_TZ, tz)
        return _cache[key]
    tiles = [t for t, w in ((1, 32), (2, 16), (3, 8), (4, 32)) if n % w == 0] + ([5] if tile_16x12_fits(n) else [])
    tzs = [0] + [z for z in (8, 16, 32, 64, 128) if n % z == 0 and z < n]
    best = None
    for t in tiles:
        for z in tzs:
            _set(LEVEL_TILE, t)
            _set(LEVEL_TZ, z)
            ms = _time(run, reps=5)  # the median of 5: shapes within a few percent of each other
            if verbose:
                print(f"autotune level tile={t} tz={z}: {ms:.3f} ms")
            if best is None or ms < best[1]:
                best = ((t, z), ms)
    _set(LEVEL_TILE, best[0][0])
    _set(LEVEL_TZ, best[0][1])
    _cache[key] = best
    return best


def tune_amr_graph(solver, key, blocks=(256, 512, 1024, 2048), rounds=2, verbose=False):
    """The fixed-grid size of a captured coarse step (AmrSolver with
    graph=True), timed on the solver's own coarse steps: each candidate is
    captured (untimed) and one step replayed, `rounds` times over the
    candidates, keeping each one's fastest step.  The solver advances by
    len(blocks) * rounds coarse steps.  Returns (blocks, ms)."""
    import torch
    if key in _cache:
        _set(PERSIST_BLOCKS, _cache[key][0])
        return _cache[key]
    best = {}
    for _ in range(rounds):
        for b in blocks:
            _set(PERSIST_BLOCKS, b)
            solver.capture()
            torch.cuda.synchronize()
            ms = _time(solver.coarse_step, reps=0)
            if verbose:
                print(f"autotune amr graph blocks={b}: {ms:.3f} ms")
            best[b] = min(best.get(b, ms), ms)
    b = min(best, key=best.get)
    _set(PERSIST_BLOCKS, b)
    _cache[key] = (b, best[b])
    return _cache[key]


def reset():
    for k in (UNSPLIT_BLOCK, LEVEL_TILE, LEVEL_TZ, PERSIST_BLOCKS, AMR_FACE_OCTS):
        _set(k, 0)
