"""RAMSES snapshot files (output_NNNNN/) from an oct-tree mesh.

The files are those the reference writes in dump_all
(amr/output_amr.f90:5-206) for one domain, record for record:

  info_NNNNN.txt             output_info   (amr/output_amr.f90:411-491)
  amr_NNNNN.out00001         backup_amr    (amr/output_amr.f90:211-406)
  hydro_NNNNN.out00001       backup_hydro  (hydro/output_hydro.f90:1-179)
  hydro_file_descriptor.txt  dump_header_info / dump_var_info
                             (io/dump_utils.f90:122-139)

so the reference's readers (its restart in amr/init_amr.f90, and
tests/visu/visu_ramses.py) load them.  The binary files are Fortran
sequential unformatted records: a 4-byte little-endian length, the payload,
the length again.  The reference defragments the mesh before it writes
(dump_all calls defrag); ``AmrSolver.write_output`` runs the same defrag on
the GPU (``AmrLevels.defrag``) and then writes the host copy of the mesh.
"""
import os
import struct
from dataclasses import dataclass, field

import numpy as np

NCOARSE = 1


@dataclass
class Mesh:
    """Host copy of a one-domain mesh, 0-based arrays holding the
    reference's 1-based elements (son[i-1] is son(i))."""
    ngridmax: int
    nlevelmax: int
    lists: dict            # level -> grid ids in list order
    son: np.ndarray        # (ncell,)
    father: np.ndarray     # (ngridmax,)
    nbor: np.ndarray       # (6*ngridmax,), nbor[(j-1)*ngridmax + igrid-1]
    flag1: np.ndarray      # (ncell,)
    xg: np.ndarray         # (ngridmax, 3)
    uold: np.ndarray       # (5, ncell)
    free: np.ndarray       # free grids from headf to tailf
    used_mem: int = -1     # make_grid_fine's last ngridmax-numbf (-1: ngridmax-len(free))


@dataclass
class RunInfo:
    """The run state and parameters a snapshot records.  Defaults are the
    reference's for a non-cosmological run (amr/amr_parameters.f90,
    amr/amr_commons.f90, hydro/hydro_commons.f90)."""
    levelmin: int
    gamma: float = 1.4
    smallr: float = 1e-10
    t: float = 0.0
    nstep: int = 0
    nstep_coarse: int = 0
    dtold: tuple = ()
    dtnew: tuple = ()
    boxlen: float = 1.0
    noutput: int = 0
    iout: int = 1
    ifout: int = 1
    tout: tuple = ()
    aout: tuple = ()
    ordering: str = "hilbert"
    bound_key: tuple = None       # default: the one-domain Hilbert key range
    einit: float = 0.0
    mass_tot_0: float = 0.0
    rho_tot: float = 0.0
    omega_m: float = 1.0
    omega_l: float = 0.0
    omega_k: float = 0.0
    omega_b: float = 0.045
    h0: float = 1.0
    aexp_ini: float = 10.0
    boxlen_ini: float = 0.0
    aexp: float = 1.0
    hexp: float = 0.0
    aexp_old: float = 1.0
    epot_tot_int: float = 0.0
    epot_tot_old: float = 0.0
    mass_sph: float = 0.0
    units: tuple = (1.0, 1.0, 1.0)  # unit_l, unit_d, unit_t
    extra: dict = field(default_factory=dict)


class _Records:
    def __init__(self, path):
        self.f = open(path, "wb")

    def rec(self, *parts):
        payload = b"".join(np.ascontiguousarray(p).tobytes() for p in parts)
        n = struct.pack("<i", len(payload))
        self.f.write(n + payload + n)

    def close(self):
        self.f.close()


def _i4(*v):
    return np.array(v, dtype="<i4")


def _f8(*v):
    return np.array(v, dtype="<f8")


def fortran_e(x, w=23, d=15):
    """Fortran Ew.d edit descriptor: 0.ddd...dE+xx, right-justified."""
    x = float(x)
    if x == 0.0:
        s = "0." + "0" * d + "E+00"
    else:
        m, e = f"{abs(x):.{d - 1}e}".split("e")
        digits = m.replace(".", "")
        ex = int(e) + 1
        s = "0." + digits + (f"E{ex:+03d}" if abs(ex) <= 99 else f"{ex:+04d}")
        if x < 0:
            s = "-" + s
    return s.rjust(w)


def _cells(mesh, ind, grids):
    return NCOARSE + (ind - 1) * mesh.ngridmax + grids - 1  # 0-based positions


def _level_grids(mesh, l):
    return np.asarray(mesh.lists.get(l, []), dtype=np.int64)


def write_info(mesh, run, path):
    L = mesh.nlevelmax
    bk = run.bound_key if run.bound_key is not None else (0.0, float(2 ** (3 * (L + 1))))
    lines = [f"ncpu        ={1:11d}", f"ndim        ={3:11d}", f"levelmin    ={run.levelmin:11d}",
             f"levelmax    ={L:11d}", f"ngridmax    ={mesh.ngridmax:11d}", f"nstep_coarse={run.nstep_coarse:11d}",
             "", f"boxlen      ={fortran_e(run.boxlen)}", f"time        ={fortran_e(run.t)}",
             f"aexp        ={fortran_e(run.aexp)}", f"H0          ={fortran_e(run.h0)}",
             f"omega_m     ={fortran_e(run.omega_m)}", f"omega_l     ={fortran_e(run.omega_l)}",
             f"omega_k     ={fortran_e(run.omega_k)}", f"omega_b     ={fortran_e(run.omega_b)}",
             f"unit_l      ={fortran_e(run.units[0])}", f"unit_d      ={fortran_e(run.units[1])}",
             f"unit_t      ={fortran_e(run.units[2])}", "",
             "ordering type=" + run.ordering.ljust(80)[:80],
             "   DOMAIN   ind_min                 ind_max",
             f"{1:8d} {fortran_e(bk[0])} {fortran_e(bk[1])}"]
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")


def write_amr(mesh, run, path):
    L, ngm = mesh.nlevelmax, mesh.ngridmax
    counts = [len(_level_grids(mesh, l)) for l in range(1, L + 1)]
    heads = [int(_level_grids(mesh, l)[0]) if counts[l - 1] else 0 for l in range(1, L + 1)]
    tails = [int(_level_grids(mesh, l)[-1]) if counts[l - 1] else 0 for l in range(1, L + 1)]
    ngrid = sum(counts)
    numbtot = np.zeros((L, 10), dtype="<i4")
    numbtot[:, 0:4] = np.array(counts)[:, None]
    free = np.asarray(mesh.free)
    headf = int(free[0]) if len(free) else 0
    tailf = int(free[-1]) if len(free) else 0
    used = mesh.used_mem if mesh.used_mem >= 0 else ngm - len(free)
    bk = run.bound_key if run.bound_key is not None else (0.0, float(2 ** (3 * (L + 1))))
    dtold = np.zeros(L) if len(run.dtold) == 0 else np.asarray(run.dtold[:L], dtype=float)
    dtnew = np.zeros(L) if len(run.dtnew) == 0 else np.asarray(run.dtnew[:L], dtype=float)
    r = _Records(path)
    r.rec(_i4(1))
    r.rec(_i4(3))
    r.rec(_i4(1, 1, 1))
    r.rec(_i4(L))
    r.rec(_i4(ngm))
    r.rec(_i4(0))
    r.rec(_i4(ngrid))
    r.rec(_f8(run.boxlen))
    r.rec(_i4(run.noutput, run.iout, run.ifout))
    huge = np.finfo(np.float64).max   # default of tout / aout (amr/amr_parameters.f90:118-119)
    r.rec(_f8(*(list(run.tout) + [huge] * run.noutput)[:run.noutput]))
    r.rec(_f8(*(list(run.aout) + [huge] * run.noutput)[:run.noutput]))
    r.rec(_f8(run.t))
    r.rec(_f8(*dtold))
    r.rec(_f8(*dtnew))
    r.rec(_i4(run.nstep, run.nstep_coarse))
    r.rec(_f8(run.einit, run.mass_tot_0, run.rho_tot))
    r.rec(_f8(run.omega_m, run.omega_l, run.omega_k, run.omega_b, run.h0, run.aexp_ini, run.boxlen_ini))
    r.rec(_f8(run.aexp, run.hexp, run.aexp_old, run.epot_tot_int, run.epot_tot_old))
    r.rec(_f8(run.mass_sph))
    r.rec(_i4(*heads))
    r.rec(_i4(*tails))
    r.rec(_i4(*counts))
    r.rec(numbtot)
    r.rec(_i4(headf, tailf, len(free), used, used))
    r.rec(np.frombuffer(run.ordering.ljust(128)[:128].encode(), dtype=np.uint8))
    r.rec(_f8(*bk))
    r.rec(_i4(*mesh.son[:NCOARSE]))
    r.rec(_i4(*([1] * NCOARSE)))      # flag_coarse flags the coarse cells
    r.rec(_i4(*([1] * NCOARSE)))      # cpu_map of the coarse cells
    for l in range(1, L + 1):
        g = _level_grids(mesh, l)
        if len(g) == 0:
            continue
        nxt = np.concatenate([g[1:], [0]]).astype("<i4")
        prv = np.concatenate([[0], g[:-1]]).astype("<i4")
        r.rec(g.astype("<i4"))
        r.rec(nxt)
        r.rec(prv)
        for d in range(3):
            r.rec(mesh.xg[g - 1, d].astype("<f8"))
        r.rec(mesh.father[g - 1].astype("<i4"))
        for j in range(6):
            r.rec(mesh.nbor[j * ngm + g - 1].astype("<i4"))
        for ind in range(1, 9):
            r.rec(mesh.son[_cells(mesh, ind, g)].astype("<i4"))
        for ind in range(1, 9):
            r.rec(np.ones(len(g), dtype="<i4"))   # cpu_map: one domain
        for ind in range(1, 9):
            r.rec(mesh.flag1[_cells(mesh, ind, g)].astype("<i4"))
    r.close()


HYDRO_FIELDS = [("density", "d"), ("velocity_x", "d"), ("velocity_y", "d"), ("velocity_z", "d"),
                ("pressure", "d")]


def write_hydro(mesh, run, path, desc_path):
    L = mesh.nlevelmax
    r = _Records(path)
    r.rec(_i4(1))
    r.rec(_i4(5))
    r.rec(_i4(3))
    r.rec(_i4(L))
    r.rec(_i4(0))
    r.rec(_f8(run.gamma))
    u = mesh.uold
    for l in range(1, L + 1):
        g = _level_grids(mesh, l)
        r.rec(_i4(l))
        r.rec(_i4(len(g)))
        if len(g) == 0:
            continue
        for ind in range(1, 9):
            c = _cells(mesh, ind, g)
            rho = u[0, c]
            rmax = np.maximum(rho, run.smallr)
            r.rec(rho.astype("<f8"))
            for d in range(1, 4):
                r.rec((u[d, c] / rmax).astype("<f8"))
            p = u[4, c]
            for d in range(1, 4):
                p = p - (0.5 * (u[d, c] * u[d, c])) / rmax
            r.rec(((run.gamma - 1.0) * p).astype("<f8"))
    r.close()
    with open(desc_path, "w") as f:
        f.write("# version:  1\n# ivar, variable_name, variable_type\n")
        for i, (name, kind) in enumerate(HYDRO_FIELDS, 1):
            f.write(f"{i:2d}, {name}, {kind}\n")


def write_output(mesh, run, outdir):
    """dump_all's file set for one domain (amr/output_amr.f90:5-206):
    output_NNNNN/ with NNNNN = run.ifout; the counters are advanced first
    as dump_all advances them (ifout+1, iout+1 once t reaches tout(iout)),
    and the advanced values are what the files record.  Returns the
    directory; `run` is updated in place."""
    tag = f"{run.ifout:05d}"
    run.ifout += 1
    huge = np.finfo(np.float64).max
    tout = run.tout[run.iout - 1] if run.iout <= len(run.tout) else huge
    aout = run.aout[run.iout - 1] if run.iout <= len(run.aout) else huge
    if run.t >= tout or run.aexp >= aout:
        run.iout += 1
    d = os.path.join(outdir, f"output_{tag}")
    os.makedirs(d, exist_ok=True)
    write_info(mesh, run, os.path.join(d, f"info_{tag}.txt"))
    write_amr(mesh, run, os.path.join(d, f"amr_{tag}.out00001"))
    write_hydro(mesh, run, os.path.join(d, f"hydro_{tag}.out00001"), os.path.join(d, "hydro_file_descriptor.txt"))
    return d


# -- reading (restart) ------------------------------------------------------

def _records(path):
    """The payloads of a Fortran sequential unformatted file, in order."""
    with open(path, "rb") as f:
        data = f.read()
    out, p = [], 0
    while p < len(data):
        if p + 4 > len(data):
            raise ValueError(f"{path}: truncated record marker at byte {p}")
        (n,) = struct.unpack_from("<i", data, p)
        end = p + 4 + n
        if n < 0 or end + 4 > len(data) or struct.unpack_from("<i", data, end)[0] != n:
            raise ValueError(f"{path}: bad record at byte {p}")
        out.append(data[p + 4:end])
        p = end + 4
    return out


class _Reader:
    def __init__(self, path):
        self.path = path
        self.recs = _records(path)
        self.k = 0

    def raw(self):
        if self.k >= len(self.recs):
            raise ValueError(f"{self.path}: fewer records than the header announces")
        r = self.recs[self.k]
        self.k += 1
        return r

    def i4(self):
        return np.frombuffer(self.raw(), dtype="<i4").astype(np.int64)

    def f8(self):
        return np.frombuffer(self.raw(), dtype="<f8").copy()


def output_files(outdir, nrestart):
    """The amr and hydro file paths of output_NNNNN under outdir."""
    tag = f"{nrestart:05d}"
    d = os.path.join(outdir, f"output_{tag}")
    return os.path.join(d, f"amr_{tag}.out00001"), os.path.join(d, f"hydro_{tag}.out00001")


def read_output(outdir, nrestart, ngridmax=None, nlevelmax=None, smallr=1e-10, gamma=None, ordering=None,
                tout=None):
    """The restart of amr/adaptive_loop.f90:32-34 for one domain: the mesh
    and run state of output_NNNNN (NNNNN = nrestart) under outdir, as
    init_amr (amr/init_amr.f90:261-551) and init_hydro
    (hydro/init_hydro.f90:56-250) rebuild them.

    ngridmax / nlevelmax: the restarted run's sizes (default: the file's);
    cell indices are renumbered when ngridmax differs (:440-449).  The free
    list is every grid after the file's grids in increasing order (:386-391)
    and used_mem is the recorded value.  ifout is the recorded one, at most
    nrestart+1 (:331); iout is the first output time in `tout` (the
    restarted run's) after t (:353-363).  The hydro file's primitives become
    conserved variables as init_hydro converts them (density; velocity times
    max(density, smallr); pressure/(gamma-1) plus the kinetic energy, or 0
    where the density is not positive); the coarse cells are 0.  Raises
    ValueError where the reference would stop (incompatible file)."""
    amr_path, hydro_path = output_files(outdir, nrestart)
    for p in (amr_path, hydro_path):
        if not os.path.exists(p):
            raise ValueError(f"Restart failed: file {p} not found")
    r = _Reader(amr_path)
    ncpu2, ndim2 = int(r.i4()[0]), int(r.i4()[0])
    nx2, ny2, nz2 = (int(v) for v in r.i4())
    L2, ngm2, nbound2, ngrid_current = (int(r.i4()[0]) for _ in range(4))
    boxlen = float(r.f8()[0])
    ngm = ngm2 if ngridmax is None else int(ngridmax)
    L = L2 if nlevelmax is None else int(nlevelmax)
    if ncpu2 != 1:
        raise ValueError(f"Number of processes not compatible: the file has ncpu={ncpu2}, this build runs one domain")
    if ndim2 != 3 or (nx2, ny2, nz2) != (1, 1, 1) or nbound2 != 0 or L2 > L or ngrid_current > ngm:
        raise ValueError("file is not compatible with the run: ndim nx ny nz nlevelmax ngrid nboundary = "
                         f"{ndim2} {nx2} {ny2} {nz2} {L2} {ngrid_current} {nbound2} (run: 3 1 1 1 {L} {ngm} 0)")
    noutput2, iout2, ifout2 = (int(v) for v in r.i4())
    tout2, aout2 = r.f8(), r.f8()
    run = RunInfo(levelmin=1, boxlen=boxlen)
    run.ifout = min(ifout2, nrestart + 1)
    run.t = float(r.f8()[0])
    run.dtold, run.dtnew = tuple(r.f8()), tuple(r.f8())
    run.nstep, run.nstep_coarse = (int(v) for v in r.i4())
    run.einit, run.mass_tot_0, run.rho_tot = (float(v) for v in r.f8())
    (run.omega_m, run.omega_l, run.omega_k, run.omega_b, run.h0, run.aexp_ini,
     run.boxlen_ini) = (float(v) for v in r.f8())
    run.aexp, run.hexp, run.aexp_old, run.epot_tot_int, run.epot_tot_old = (float(v) for v in r.f8())
    run.mass_sph = float(r.f8()[0])
    huge = np.finfo(np.float64).max
    run.tout = tuple(tout if tout is not None else tout2)
    run.noutput = len(run.tout)
    run.iout = 1
    if run.tout and not all(x == huge for x in run.tout):
        while run.iout <= len(run.tout) and run.tout[run.iout - 1] <= run.t:
            run.iout += 1
    headl, taill, numbl = r.i4(), r.i4(), r.i4()
    r.raw()                                  # numbtot
    _headf, _tailf, _numbf, used_mem, _used_tot = (int(v) for v in r.i4())
    ordering2 = r.raw().decode("ascii", "replace").strip()
    if ordering is not None and ordering2 != ordering:
        raise ValueError(f"Ordering is incompatible: the file has {ordering2!r}, the run {ordering!r}")
    run.ordering = ordering2
    run.bound_key = tuple(r.f8())
    ncell = NCOARSE + 8 * ngm
    son = np.zeros(ncell, dtype=np.int32)
    flag1 = np.zeros(ncell, dtype=np.int32)
    father = np.zeros(ngm, dtype=np.int32)
    nbor = np.zeros(6 * ngm, dtype=np.int32)
    xg = np.zeros((ngm, 3))
    son[:NCOARSE] = r.i4()
    flag1[:NCOARSE] = r.i4()
    r.raw()                                  # cpu_map of the coarse cells

    def remap(cells, l):
        if ngm == ngm2 or l <= 1:
            return cells
        pos = (cells - NCOARSE - 1) // ngm2
        return NCOARSE + pos * ngm + (cells - NCOARSE - pos * ngm2)

    lists = {}
    for l in range(1, L2 + 1):
        n = int(numbl[l - 1])
        if n == 0:
            continue
        g = r.i4()
        nxt, prv = r.i4(), r.i4()
        if len(g) != n or int(g[0]) != int(headl[l - 1]) or int(g[-1]) != int(taill[l - 1]):
            raise ValueError(f"{amr_path}: level {l} list does not match its head/tail/count")
        if np.any(nxt[:-1] != g[1:]) or nxt[-1] != 0 or prv[0] != 0 or np.any(prv[1:] != g[:-1]):
            raise ValueError(f"{amr_path}: level {l} next/prev links are not the list order")
        if g.max() > ngm or g.min() < 1:
            raise ValueError(f"{amr_path}: grid index beyond ngridmax at level {l}")
        for d in range(3):
            xg[g - 1, d] = r.f8()
        father[g - 1] = remap(r.i4(), l)
        for j in range(6):
            nbor[j * ngm + g - 1] = remap(r.i4(), l)
        for ind in range(1, 9):
            son[NCOARSE + (ind - 1) * ngm + g - 1] = r.i4()
        for ind in range(1, 9):
            r.raw()                          # cpu_map
        for ind in range(1, 9):
            flag1[NCOARSE + (ind - 1) * ngm + g - 1] = r.i4()
        lists[l] = g.astype(np.int32)
    free = np.arange(ngrid_current + 1, ngm + 1, dtype=np.int32)

    h = _Reader(hydro_path)
    ncpu2, nvar2, ndim2, L2h, nbound2 = (int(h.i4()[0]) for _ in range(5))
    gamma2 = float(h.f8()[0])
    g_ = gamma2 if gamma is None else float(gamma)
    if ncpu2 != 1 or nvar2 != 5 or ndim2 != 3 or nbound2 != 0 or L2h != L2:
        raise ValueError(f"{hydro_path}: not a one-domain 3D hydro file with 5 variables and {L2} levels")
    uold = np.zeros((5, ncell))
    for l in range(1, L2 + 1):
        l2, n2 = int(h.i4()[0]), int(h.i4()[0])
        g = lists.get(l, np.zeros(0, dtype=np.int32)).astype(np.int64)
        if l2 != l or n2 != len(g):
            raise ValueError(f"{hydro_path}: level {l2} has {n2} grids, the amr file {len(g)}")
        if n2 == 0:
            continue
        for ind in range(1, 9):
            c = NCOARSE + (ind - 1) * ngm + g - 1
            rho = h.f8()
            uold[0, c] = rho
            rmax = np.maximum(rho, smallr)
            for d in range(1, 4):
                uold[d, c] = h.f8() * rmax
            e = h.f8() / (g_ - 1.0)
            for d in range(1, 4):
                e = e + (0.5 * (uold[d, c] * uold[d, c])) / rmax
            uold[4, c] = np.where(rho > 0.0, e, 0.0)
    mesh = Mesh(ngm, L, lists, son, father, nbor, flag1, xg, uold, free, used_mem)
    run.gamma, run.smallr = g_, smallr
    return mesh, run
