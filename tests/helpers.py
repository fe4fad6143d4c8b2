"""Input generators shared by the tests (numpy, seeded)."""
import numpy as np

NV = 32


def conserved(rho, vel, p, gamma):
    u = np.empty((5,) + rho.shape)
    u[0] = rho
    u[1:4] = rho * vel
    e = np.zeros_like(rho)
    for d in range(3):
        e = e + 0.5 * rho * vel[d] ** 2
    u[4] = e + p / (gamma - 1.0)
    return u


def random_stencils(rng, ngroups, nv=NV, gamma=5.0 / 3.0):
    """Synthetic gathered oct batches: rho~logN(0,1), v~N(0,1), P~logN(-0.5,0.5)."""
    rho = rng.lognormal(0.0, 1.0, (ngroups, 6, 6, 6, nv))
    vel = rng.normal(0.0, 1.0, (3, ngroups, 6, 6, 6, nv))
    p = rng.lognormal(-0.5, 0.5, (ngroups, 6, 6, 6, nv))
    u = conserved(rho, vel, p, gamma)            # (5, g, 6,6,6, nv)
    return np.ascontiguousarray(np.moveaxis(u, 0, 1))  # (g, 5, 6,6,6, nv)


def random_grid(rng, n, gamma=5.0 / 3.0, mach=1.0):
    rho = rng.lognormal(0.0, 0.5, (n, n, n))
    vel = rng.normal(0.0, mach, (3, n, n, n))
    p = rng.lognormal(-0.5, 0.5, (n, n, n))
    return np.ascontiguousarray(conserved(rho, vel, p, gamma))


def sedov_ic(n, gamma=1.4, e_total=1.0):
    """3D Sedov point blast (hydro/init_flow_fine.f90:475-596 'square' ambient
    plus 'point' region with CIC weights at the box centre)."""
    dx = 1.0 / n
    vol = dx ** 3
    q = np.zeros((5, n, n, n))
    q[0] = 1.0
    q[4] = 1e-5
    xc = (np.arange(n) + 0.5) * dx
    w = np.maximum(1.0 - np.abs(xc - 0.5) / dx, 0.0)
    r = w[:, None, None] * w[None, :, None] * w[None, None, :]
    q[4] = q[4] + e_total * (gamma - 1.0) * r / vol
    u = np.zeros_like(q)
    u[0] = q[0]
    u[4] = q[4] / (gamma - 1.0)
    return u


def defined_face_mask(nv=NV):
    """True on the 36 flux slots per oct the reference defines (x: i=1..3,
    j,k=1..2; y, z alike), in the C-ordered (3, 5, 3,3,3, nv) view."""
    m = np.zeros((3, 5, 3, 3, 3, nv), dtype=bool)
    m[0, :, 0:2, 0:2, 0:3] = True   # [d][v][k][j][i]
    m[1, :, 0:2, 0:3, 0:2] = True
    m[2, :, 0:3, 0:2, 0:2] = True
    return m
