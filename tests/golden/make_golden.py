"""Write the golden vectors under tests/golden/ from the reference's own code.

Runs only in the container that holds the reference sources: it loads
``oracle/_ref/libramses_ref.so`` -- the reference's hydro/umuscl.f90 and
hydro/godunov_utils.f90 compiled by ``make -C oracle ref`` -- and records
inputs and outputs as small .npz fixtures.  The fixtures are data; the GPU
tests and the oracle tests read them without the reference present.

    python tests/golden/make_golden.py

Families
  unsplit_cases.npz   F1: unsplit() on 2 octs per case, solvers x slopes x
                      gravity, plus hand-built edge stencils (uniform state,
                      stationary contact, supersonic left/right flow,
                      density/pressure floors, exact-zero ties).
  cmpdt_cases.npz     F2: cmpdt() inputs, in-place outputs and dt.
  sedov16.npz         F3: 16^3 periodic Sedov (gamma=1.4, hllc, minmod,
                      CFL 0.8): conserved state after steps 1, 2 and 10 and
                      the dt sequence, composed from the reference unsplit
                      and cmpdt by the uniform-level driver below.
  sedov64.npz         F3 at the Sedov configuration's own size: 64^3
                      periodic Sedov (gamma=1.4, hllc, minmod, CFL 0.8),
                      the initial state, the dt of each of 20 level steps
                      and the conserved state after steps 1, 2, 3 and 20,
                      from the reference's own courant_fine, set_unew,
                      godunov_fine and set_uold on a fully refined level-6
                      oct tree (oracle/ref_uniform.py).
  amr_sphere.npz      F4: an oct tree of levels 3-5 (tests/amr_cases.py)
                      driven through the reference's own courant_fine,
                      set_unew, godunov_fine and upload_fine
                      (oracle/ref_amr_glue.f90): dt per level and the cells
                      each routine changed.
  amr_refine.json     F5: the same kind of tree with room for 20000 more
                      grids, refined by the reference's own flag_fine and
                      refine_fine (gradient criteria on density and
                      pressure), then advanced two adaptive coarse steps in
                      amr_step's order: sha256 digests of son / father /
                      nbor / flag1, every level's grid list, the free list
                      and uold / unew after each stage, and the level dts.
"""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

import oracle  # noqa: E402

NLANE = 2  # octs stored per unsplit case


def conserved(rho, vel, p, gamma):
    """primitive -> conservative, as hydro/condinit.f90:36-71."""
    u = np.empty((5,) + rho.shape)
    u[0] = rho
    u[1:4] = rho * vel
    e = np.zeros_like(rho)
    for d in range(3):
        e = e + 0.5 * rho * vel[d] ** 2
    u[4] = e + p / (gamma - 1.0)
    return u


def random_stencil(rng, nv, gamma):
    rho = rng.lognormal(0.0, 1.0, (6, 6, 6, nv))
    vel = rng.normal(0.0, 1.0, (3, 6, 6, 6, nv))
    p = rng.lognormal(-0.5, 0.5, (6, 6, 6, nv))
    return conserved(rho, vel, p, gamma)


def edge_stencils(nv, gamma):
    """Hand-built stencils that hit the branches of the limiter and solvers."""
    out = {}
    ones = np.ones((6, 6, 6, nv))
    z3 = np.zeros((3, 6, 6, 6, nv))
    out["uniform"] = conserved(ones * 1.3, z3 + np.array([0.2, -0.1, 0.05])[:, None, None, None, None],
                               ones * 0.7, gamma)
    x = np.arange(-1, 5)[:, None, None, None]  # axis i is the last spatial axis (x)
    xs = np.broadcast_to(np.arange(-1, 5)[None, None, :, None], (6, 6, 6, nv))
    rho = np.where(xs < 2, 1.0, 0.125)
    out["stationary_contact"] = conserved(rho, z3, ones * 1.0, gamma)
    vsup = z3.copy()
    vsup[0] = 8.0
    out["supersonic_right"] = conserved(ones * (1.0 + 0.1 * (xs > 1)), vsup, ones, gamma)
    vsub = z3.copy()
    vsub[0] = -8.0
    out["supersonic_left"] = conserved(ones * (1.0 + 0.1 * (xs > 1)), vsub, ones, gamma)
    # density below smallr and negative internal energy trigger the floors
    rfl = ones.copy()
    rfl[:, :, 2] = 1e-14
    u = conserved(rfl, z3, ones * 1e-3, gamma)
    u[4][:, 3] = -1.0
    out["floors"] = u
    # Sedov-like: symmetric blast around the stencil centre, exact ties
    r2 = ((xs - 1.5) ** 2 + (np.broadcast_to(np.arange(-1, 5)[None, :, None, None], xs.shape) - 1.5) ** 2
          + (np.broadcast_to(np.arange(-1, 5)[:, None, None, None], xs.shape) - 1.5) ** 2)
    p = np.where(r2 < 1.0, 1e4, 1e-5)
    out["sedov_centre"] = conserved(ones, z3, p, gamma)
    del x
    return out


def unsplit_cases(R):
    rng = np.random.default_rng(101)
    nv = R.nvector
    cases = []
    dx = 1.0 / 128
    for riemann in ("llf", "hllc", "hll"):
        for slope in (0, 1, 2, 3, 7, 8):
            for grav in (False, True):
                if slope in (0, 3, 7, 8) and (grav or riemann == "hll"):
                    continue
                cases.append(dict(riemann=riemann, slope_type=slope, grav=grav, gamma=5.0 / 3.0,
                                  kind="random", dx=dx, dt=0.5 * dx / 6.0))
    for name in ("uniform", "stationary_contact", "supersonic_right", "supersonic_left", "floors",
                 "sedov_centre"):
        for riemann in ("llf", "hllc", "hll"):
            cases.append(dict(riemann=riemann, slope_type=1, grav=False, gamma=1.4, kind=name,
                              dx=1.0 / 64, dt=1e-3))
    arrays = {}
    meta = []
    for ci, c in enumerate(cases):
        R.set_params(gamma=c["gamma"], riemann=c["riemann"], slope_type=c["slope_type"])
        if c["kind"] == "random":
            uin = random_stencil(rng, nv, c["gamma"])
        else:
            uin = np.ascontiguousarray(edge_stencils(nv, c["gamma"])[c["kind"]])
        g = rng.normal(0.0, 1.0, oracle.grav_shape(nv)) if c["grav"] else None
        flux = np.full(oracle.flux_shape(nv), np.nan)
        tmp = np.full(oracle.tmp_shape(nv), np.nan)
        R.unsplit(uin, g, flux, tmp, c["dx"], c["dt"], NLANE)
        arrays[f"c{ci}_uin"] = uin[..., :NLANE].copy()
        if g is not None:
            arrays[f"c{ci}_grav"] = g[..., :NLANE].copy()
        arrays[f"c{ci}_flux"] = flux[..., :NLANE].copy()
        arrays[f"c{ci}_tmp"] = tmp[..., :NLANE].copy()
        meta.append(c)
    np.savez_compressed(os.path.join(HERE, "unsplit_cases.npz"), meta=json.dumps(meta), **arrays)
    print(f"unsplit_cases.npz: {len(cases)} cases")


def cmpdt_cases(R):
    rng = np.random.default_rng(202)
    nv = R.nvector
    arrays, meta = {}, []
    ci = 0
    for gamma in (1.4, 5.0 / 3.0):
        for C in (0.5, 0.8):
            for grav in (False, True):
                R.set_params(gamma=gamma, courant_factor=C)
                rho = rng.lognormal(0, 1, nv)
                v = rng.normal(0, 1, (3, nv))
                p = rng.lognormal(-0.5, 0.5, nv)
                uu = conserved(rho, v, p, gamma)
                uu[:, 5] = uu[:, 5] * 0 + np.array([1e-14, 0, 0, 0, -1.0])  # floors
                gg = rng.normal(0, 1, (3, nv)) if grav else np.zeros((3, nv))
                ncell = nv - 3
                arrays[f"c{ci}_uu_in"] = uu.copy()
                arrays[f"c{ci}_gg"] = gg
                dt = R.cmpdt(uu, gg, 1.0 / 64, ncell)
                arrays[f"c{ci}_uu_out"] = uu
                arrays[f"c{ci}_dt"] = np.array(dt)
                meta.append(dict(gamma=gamma, courant_factor=C, dx=1.0 / 64, ncell=ncell))
                ci += 1
    np.savez_compressed(os.path.join(HERE, "cmpdt_cases.npz"), meta=json.dumps(meta), **arrays)
    print(f"cmpdt_cases.npz: {ci} cases")


def sedov_ic(n, gamma=1.4, e_total=1.0):
    """3D Sedov point blast on a periodic unit box as region_condinit builds it
    (hydro/init_flow_fine.f90:475-596, namelist/sedov3d.nml): ambient
    rho=1, P=1e-5, and a 'point' region at the box centre whose CIC weights
    put 1/8 of p_region/vol in each of the 8 central cells."""
    dx = 1.0 / n
    vol = dx ** 3
    p_region = e_total * (gamma - 1.0)
    q = np.zeros((5, n, n, n))
    q[0] = 1.0
    q[4] = 1e-5
    xc = (np.arange(n) + 0.5) * dx
    w = np.maximum(1.0 - np.abs(xc - 0.5) / dx, 0.0)
    r = w[:, None, None] * w[None, :, None] * w[None, None, :]
    q[0] = q[0] + 0.0 * r / vol
    q[4] = q[4] + p_region * r / vol
    u = np.zeros_like(q)
    u[0] = q[0]
    u[1:4] = q[0] * q[1:4]
    e = np.zeros((n, n, n))
    for d in range(3):
        e = e + 0.5 * q[0] * q[1 + d] ** 2
    u[4] = e + q[4] / (gamma - 1.0)
    return u


def uniform_step_with(R, u, dx, nv):
    """One courant + godunov step of a fully refined periodic level, driving
    the reference's cmpdt and unsplit (hydro/courant_fine.f90:44-131,
    hydro/godunov_fine.f90:553-681,751-792)."""
    n = u.shape[-1]
    ncell = n ** 3
    flat = u.reshape(5, ncell)
    dt = 1.0 / 1e-10  # newdt_fine: boxlen/smallc
    for c0 in range(0, ncell, nv):
        m = min(nv, ncell - c0)
        uu = np.zeros((5, nv))
        uu[:, :m] = flat[:, c0:c0 + m]
        dt = min(dt, R.cmpdt(uu, None, dx, m))
    no = n // 2
    noct = no ** 3
    unew = u.copy()
    offs = np.arange(-1, 5)
    for b0 in range(0, noct, nv):
        ng = min(nv, noct - b0)
        io = np.arange(b0, b0 + ng)
        ox, oy, oz = io % no, (io // no) % no, io // (no * no)
        X = (2 * ox[None, :] + offs[:, None] - 1) % n  # [i3][l]
        Y = (2 * oy[None, :] + offs[:, None] - 1) % n
        Z = (2 * oz[None, :] + offs[:, None] - 1) % n
        uin = np.zeros((5, 6, 6, 6, nv))
        # uin[v][k][j][i][l] = u[v][Z][Y][X]
        uin[..., :ng] = u[:, Z[:, None, None, :], Y[None, :, None, :], X[None, None, :, :]]
        flux = np.zeros(oracle.flux_shape(nv))
        tmp = np.zeros(oracle.tmp_shape(nv))
        R.unsplit(uin, None, flux, tmp, dx, dt, ng)
        for d in range(3):
            e = [0, 0, 0]
            e[d] = 1
            for k2 in range(2):
                for j2 in range(2):
                    for i2 in range(2):
                        fl = flux[d, :, k2, j2, i2, :ng]
                        fr = flux[d, :, k2 + e[2], j2 + e[1], i2 + e[0], :ng]
                        zc, yc, xc = 2 * oz + k2, 2 * oy + j2, 2 * ox + i2
                        unew[:, zc, yc, xc] = unew[:, zc, yc, xc] + (fl - fr)
    return unew, dt


def sedov16(R):
    n, nv = 16, R.nvector
    R.set_params(gamma=1.4, riemann="hllc", slope_type=1, courant_factor=0.8)
    u = sedov_ic(n)
    dts, snaps = [], {}
    for step in range(1, 11):
        u, dt = uniform_step_with(R, u, 1.0 / n, nv)
        dts.append(dt)
        if step in (1, 2, 10):
            snaps[f"u_step{step}"] = u.copy()
    np.savez_compressed(os.path.join(HERE, "sedov16.npz"), u0=sedov_ic(n), dt=np.array(dts),
                        meta=json.dumps(dict(n=n, gamma=1.4, riemann="hllc", slope_type=1,
                                             courant_factor=0.8, steps=10)), **snaps)
    print("sedov16.npz: dt =", dts[:3], "...")


def sedov64(R):
    from oracle.ref_uniform import ReferenceUniformLevel
    n, steps = 64, 20
    lvl = ReferenceUniformLevel(n, 1.4, riemann="hllc", slope_type=1, courant_factor=0.8, ref=R)
    u0 = sedov_ic(n)
    lvl.load(u0)
    dts, snaps = [], {}
    for step in range(1, steps + 1):
        dts.append(lvl.step())
        if step in (1, 2, 3, steps):
            snaps[f"u_step{step}"] = lvl.state()
    lvl.release()
    np.savez_compressed(os.path.join(HERE, "sedov64.npz"), u0=u0, dt=np.array(dts),
                        meta=json.dumps(dict(n=n, gamma=1.4, riemann="hllc", slope_type=1,
                                             courant_factor=0.8, steps=steps)), **snaps)
    print("sedov64.npz: dt =", dts[:3], "...")


def amr_sphere(R):
    """F4: one oct tree (levels 3-5, refined ball + blob), smooth state; for
    each level: the reference's courant_fine dt, then set_unew(l-1),
    set_unew(l), godunov_fine(l) (changed cells only) and upload_fine(l)."""
    sys.path.insert(0, os.path.dirname(HERE))
    from amr_cases import smooth_state, sphere_tree
    tree = sphere_tree()
    u = smooth_state(tree)
    A = oracle.ReferenceAMR(R)
    A.install(tree)
    A.set_interpol(1, 0)
    R.set_params(gamma=5.0 / 3.0, riemann="hllc", slope_type=1, courant_factor=0.5)
    out = {"u0": u}
    for l in (3, 4, 5):
        A.put(0, u)
        A.put(1, u)
        dt = A.courant(l, 1e10)[0]
        if l > 3:
            A.set_unew(l - 1)
        A.set_unew(l)
        A.godunov(l, dt)
        un = A.get(1)
        idx = np.nonzero((un != u).any(0))[0]
        out[f"dt{l}"] = np.array(dt)
        out[f"idx{l}"] = idx
        out[f"val{l}"] = un[:, idx]
        A.put(0, u)
        A.upload(l)
        uo = A.get(0)
        uidx = np.nonzero((uo != u).any(0))[0]
        out[f"upidx{l}"] = uidx
        out[f"upval{l}"] = uo[:, uidx]
    np.savez_compressed(os.path.join(HERE, "amr_sphere.npz"),
                        meta=json.dumps(dict(gamma=5.0 / 3.0, riemann="hllc", levels=[3, 4, 5],
                                             interpol_type=1, interpol_var=0)), **out)
    print("amr_sphere.npz:", {l: len(out[f"idx{l}"]) for l in (3, 4, 5)})


def amr_refine(R):
    """F5: flag_fine on every level, refine_fine of levels 3-5, then two
    adaptive coarse steps of the reference's own driver, on the
    REFINE_CASE tree; digests of the mesh after each stage."""
    sys.path.insert(0, os.path.dirname(HERE))
    from amr_cases import REFINE_CASE as C, mesh_digests, reference_driver_step, smooth_state, sphere_tree
    tree = sphere_tree(C["levelmin"], C["levelmax"], extra_grids=C["extra_grids"])
    u = smooth_state(tree, seed=C["seed"])
    A = oracle.ReferenceAMR(R)
    R.set_params(gamma=C["gamma"], riemann=C["riemann"], slope_type=C["slope_type"],
                 courant_factor=C["courant_factor"])
    A.install(tree, nlevelmax=C["nlevelmax"])
    A.set_interpol(1, 0)
    A.put(0, u)
    A.put(1, u)
    A.set_refine(C["levelmin"], err_grad_d=C["err_grad_d"], err_grad_p=C["err_grad_p"])
    L = C["nlevelmax"]

    def snap():
        son, father, nbor, flag1 = A.tree_arrays()
        return mesh_digests(son, father, nbor, flag1, {l: A.grids(l) for l in range(1, L + 1)}, A.free_list(),
                            A.get(0), A.get(1))
    stages = {}
    for l in range(L, 0, -1):
        if len(A.grids(l)):
            A.flag_fine(l, 2)
    stages["flagged"] = snap()
    for l in range(C["levelmin"], L):
        A.refine_fine(l)
    stages["refined"] = snap()
    A.flag_coarse()
    dt, dtold = [0.0] * (L + 2), [0.0] * (L + 2)
    dts = []
    for k in range(2):
        reference_driver_step(A, dt, dtold)
        stages[f"step{k + 1}"] = snap()
        dts.append([float(x).hex() for x in dt])
    with open(os.path.join(HERE, "amr_refine.json"), "w") as f:
        json.dump(dict(case=C, stages=stages, dts=dts), f, indent=1, sort_keys=True)
    print("amr_refine.json:", {k: v["nocts"] for k, v in stages.items()})


def main():
    R = oracle.Reference()
    only = sys.argv[1:]
    for name, fn in [("unsplit", unsplit_cases), ("cmpdt", cmpdt_cases), ("sedov16", sedov16), ("sedov64", sedov64),
                     ("amr_sphere", amr_sphere), ("amr_refine", amr_refine)]:
        if not only or name in only:
            fn(R)


if __name__ == "__main__":
    main()
