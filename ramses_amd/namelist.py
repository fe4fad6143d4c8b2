"""RAMSES parameter files (Fortran namelists) for this hydro path.

``read(path)`` parses the groups the reference reads with
``read(1, NML=...)`` (amr/read_params.f90, hydro/read_hydro_params.f90):
``&RUN_PARAMS``, ``&AMR_PARAMS``, ``&INIT_PARAMS``, ``&HYDRO_PARAMS``,
``&REFINE_PARAMS``, ``&OUTPUT_PARAMS``.  Values follow Fortran list-directed
input: integers, reals (``1e-5``, ``1.0d0``), logicals (``.true.``), quoted
strings, comma lists, repeat counts (``10*1``) and element assignments
(``region_type(2)='point'``).  Names are case-insensitive.

``Params.from_namelist`` collects what this path uses, with the reference's
defaults (hydro/hydro_parameters.f90, amr/amr_parameters.f90), and rejects
what it does not implement (``scheme`` other than 'muscl', ``difmag`` > 0,
boundaries, gravity, cooling, particles).
"""
import re
from dataclasses import dataclass, field


def _value(tok):
    t = tok.strip()
    if not t:
        return None
    if (t[0] == t[-1]) and t[0] in "'\"":
        return t[1:-1]
    lo = t.lower()
    if lo in (".true.", "t", ".t."):
        return True
    if lo in (".false.", "f", ".f."):
        return False
    try:
        return int(t)
    except ValueError:
        pass
    try:
        return float(lo.replace("d", "e"))
    except ValueError:
        return t


def _split_values(text):
    """Comma/space separated items, quotes kept together, r*v expanded."""
    out, cur, q = [], "", None
    for ch in text:
        if q:
            cur += ch
            if ch == q:
                q = None
        elif ch in "'\"":
            q = ch
            cur += ch
        elif ch in ", \t\n":
            if cur:
                out.append(cur)
            cur = ""
        else:
            cur += ch
    if cur:
        out.append(cur)
    vals = []
    for item in out:
        m = re.fullmatch(r"(\d+)\*(.+)", item)
        if m:
            vals.extend([_value(m.group(2))] * int(m.group(1)))
        else:
            vals.append(_value(item))
    return vals


def parse(text):
    """{group: {name: value or list}} of every &GROUP ... / block."""
    groups = {}
    lines = []
    for line in text.splitlines():
        # strip comments outside quotes
        q, cut = None, len(line)
        for i, ch in enumerate(line):
            if q:
                if ch == q:
                    q = None
            elif ch in "'\"":
                q = ch
            elif ch == "!":
                cut = i
                break
        lines.append(line[:cut])
    body = "\n".join(lines)
    for m in re.finditer(r"&(\w+)(.*?)(?:^\s*/\s*$|/\s*$)", body, re.S | re.M):
        name = m.group(1).lower()
        g = groups.setdefault(name, {})
        content = m.group(2)
        # split into assignments at  key=  or  key(i)=
        parts = re.split(r"(?:^|[\s,])([A-Za-z_]\w*(?:\(\s*\d+\s*\))?)\s*=", content)
        for key, val in zip(parts[1::2], parts[2::2]):
            vals = _split_values(val)
            km = re.fullmatch(r"(\w+)\(\s*(\d+)\s*\)", key)
            if km:
                k, idx = km.group(1).lower(), int(km.group(2))
                arr = g.get(k)
                if not isinstance(arr, list):
                    arr = [] if arr is None else [arr]
                while len(arr) < idx + len(vals) - 1:
                    arr.append(None)
                arr[idx - 1:idx - 1 + len(vals)] = vals
                g[k] = arr
            else:
                g[key.lower()] = vals[0] if len(vals) == 1 else vals
    return groups


def read(path):
    with open(path) as f:
        return parse(f.read())


def _list(v, n=None, fill=None):
    if v is None:
        v = []
    if not isinstance(v, list):
        v = [v]
    if n is not None:
        v = (v + [fill] * n)[:n]
    return v


@dataclass
class Region:
    """One &INIT_PARAMS region (hydro/init_flow_fine.f90:475-596)."""
    type: str
    center: tuple
    length: tuple
    exp: float
    d: float
    u: tuple
    p: float


@dataclass
class Params:
    levelmin: int = 1
    levelmax: int = 1
    ngridmax: int = 0
    boxlen: float = 1.0
    nexpand: int = 1
    nsubcycle: list = field(default_factory=lambda: [2] * 100)
    nstepmax: int = 1000000
    nrestart: int = 0
    gamma: float = 1.4
    courant_factor: float = 0.5
    smallr: float = 1e-10
    smallc: float = 1e-10
    slope_type: int = 1
    slope_theta: float = 1.5
    riemann: str = "llf"
    niter_riemann: int = 10
    interpol_type: int = 1
    interpol_var: int = 0
    err_grad_d: float = -1.0
    err_grad_u: float = -1.0
    err_grad_p: float = -1.0
    floor_d: float = 1e-10
    floor_u: float = 1e-10
    floor_p: float = 1e-10
    regions: list = field(default_factory=list)
    tout: list = field(default_factory=list)

    @classmethod
    def from_namelist(cls, groups_or_path):
        g = read(groups_or_path) if isinstance(groups_or_path, str) else groups_or_path
        run, amr = g.get("run_params", {}), g.get("amr_params", {})
        hyd, ref = g.get("hydro_params", {}), g.get("refine_params", {})
        ini, out = g.get("init_params", {}), g.get("output_params", {})
        p = cls()
        if not run.get("hydro", True):
            raise ValueError("hydro=.false.: nothing for this path to run")
        for k in ("poisson", "pic", "cooling", "sink", "rt"):
            if run.get(k, False):
                raise ValueError(f"{k}=.true. is outside this hydro path")
        p.levelmin = int(amr.get("levelmin", 1))
        p.levelmax = int(amr.get("levelmax", p.levelmin))
        p.ngridmax = int(amr.get("ngridmax", amr.get("ngridtot", 0)))
        p.boxlen = float(amr.get("boxlen", 1.0))
        for k in ("nx", "ny", "nz"):
            if int(amr.get(k, 1)) != 1:
                raise ValueError("only nx=ny=nz=1 coarse grids are supported")
        p.nexpand = int(_list(amr.get("nexpand", 1))[0])
        nsub = _list(run.get("nsubcycle"))
        p.nsubcycle = [int(x) for x in nsub] + [2] * (100 - len(nsub))
        p.nstepmax = int(run.get("nstepmax", p.nstepmax))
        p.nrestart = int(run.get("nrestart", 0))
        if str(hyd.get("scheme", "muscl")).strip().lower() != "muscl":
            raise ValueError("only scheme='muscl' is implemented")
        if float(hyd.get("difmag", 0.0)) > 0:
            raise ValueError("difmag > 0 is not implemented")
        p.gamma = float(hyd.get("gamma", p.gamma))
        p.courant_factor = float(hyd.get("courant_factor", p.courant_factor))
        p.smallr = float(hyd.get("smallr", p.smallr))
        p.smallc = float(hyd.get("smallc", p.smallc))
        p.slope_type = int(hyd.get("slope_type", p.slope_type))
        p.slope_theta = float(hyd.get("slope_theta", p.slope_theta))
        p.riemann = str(hyd.get("riemann", p.riemann)).strip().lower()
        if p.riemann not in ("llf", "hllc", "hll", "acoustic", "exact"):
            raise ValueError(f"unknown riemann solver {p.riemann!r}")
        p.niter_riemann = int(hyd.get("niter_riemann", p.niter_riemann))
        for k in ("interpol_type", "interpol_var"):
            setattr(p, k, int(ref.get(k, getattr(p, k))))
        for k in ("err_grad_d", "err_grad_u", "err_grad_p", "floor_d", "floor_u", "floor_p"):
            setattr(p, k, float(ref.get(k, getattr(p, k))))
        n = int(ini.get("nregion", 0))
        types = _list(ini.get("region_type"), n, "square")
        cols = {k: _list(ini.get(k), n, dflt) for k, dflt in
                (("x_center", 0.0), ("y_center", 0.0), ("z_center", 0.0), ("length_x", 1e10), ("length_y", 1e10),
                 ("length_z", 1e10), ("exp_region", 2.0), ("d_region", 0.0), ("u_region", 0.0), ("v_region", 0.0),
                 ("w_region", 0.0), ("p_region", 0.0))}
        for i in range(n):
            p.regions.append(Region(str(types[i]).strip(), (cols["x_center"][i], cols["y_center"][i],
                                                           cols["z_center"][i]),
                                    (cols["length_x"][i], cols["length_y"][i], cols["length_z"][i]),
                                    float(cols["exp_region"][i]), float(cols["d_region"][i]),
                                    (cols["u_region"][i], cols["v_region"][i], cols["w_region"][i]),
                                    float(cols["p_region"][i])))
        nout = int(out.get("noutput", 0))
        p.tout = [float(x) for x in _list(out.get("tout"), nout, None) if x is not None]
        return p
