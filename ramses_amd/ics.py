"""Initial conditions for the level workloads, generated on the device.

* ``sedov(n)`` -- the 3D Sedov point blast the reference builds from
  namelist/sedov3d.nml through region_condinit
  (hydro/init_flow_fine.f90:475-596) and condinit's primitive->conservative
  conversion (hydro/condinit.f90:36-71): a 'square' region covering the box
  with rho=1, u=0, P=1e-5, plus a 'point' region at the box centre carrying
  p_region = E*(gamma-1) spread over the 8 central cells with CIC weights.
* ``turbulent(n, mach, ...)`` -- a periodic Gaussian random velocity field
  with power spectrum P(k) ~ k^-4 (|v_hat| ~ k^-2, k=0 removed), scaled to
  the requested 3D rms Mach number for c_s=1, optionally a lognormal density
  and an isothermal pressure P = rho/gamma.  The default generator draws its
  white noise from numpy's ``default_rng(seed)`` (PCG64) on the host --
  velocity x, y, z, then the density -- and filters it with scipy's FFT, so
  the field can be regenerated with numpy alone; ``generator="device"``
  draws torch Philox noise on the GPU instead (fast enough to create a
  request's field inside a serving loop, but only reproducible with torch).
Both return conserved state u[var][z][y][x] as a float64 tensor.
* ``oct_batches(noct)`` / ``oct_batches_host(noct)`` -- synthetic gathered
  oct stencils for the batched unsplit boundary: torch Philox noise drawn on
  the device, or numpy ``default_rng(seed)`` noise drawn on the host in
  chunks and copied over (the benchmark's input, regenerable without torch).
* ``region_condinit(x, dx, regions)`` -- the &INIT_PARAMS regions evaluated
  at arbitrary cell centres x (N, 3) with the cell size dx of their level,
  as init_flow_fine calls condinit on every level of an oct tree; returns
  the conserved state (5, N).  ``sedov_regions`` is the sedov3d.nml case.
"""
import math
import os


def sedov(n, gamma=1.4, e_total=1.0, device="cuda"):
    import torch
    dx = 1.0 / n
    vol = dx ** 3
    xc = (torch.arange(n, dtype=torch.float64, device=device) + 0.5) * dx
    w = torch.clamp(1.0 - _div((xc - 0.5).abs(), dx), min=0.0)
    r = w[:, None, None] * w[None, :, None] * w[None, None, :]
    p = 1e-5 + _div(e_total * (gamma - 1.0) * r, vol)
    u = torch.zeros((5, n, n, n), dtype=torch.float64, device=device)
    u[0] = 1.0
    u[4] = _div(p, gamma - 1.0)
    return u


def turbulent(n, mach, gamma=5.0 / 3.0, rho_sigma=0.0, seed=0, device="cuda", slope=-2.0, generator="host"):
    import torch
    if generator == "host":
        u = turbulent_host(n, mach, gamma=gamma, rho_sigma=rho_sigma, seed=seed, slope=slope)
        return torch.from_numpy(u).to(device)
    if generator != "device":
        raise ValueError("generator must be 'host' or 'device'")
    gen = torch.Generator(device=device)
    gen.manual_seed(seed)
    k1 = torch.fft.fftfreq(n, d=1.0 / n, device=device, dtype=torch.float64)
    kr = torch.fft.rfftfreq(n, d=1.0 / n, device=device, dtype=torch.float64)
    kk = torch.sqrt(k1[:, None, None] ** 2 + k1[None, :, None] ** 2 + kr[None, None, :] ** 2)
    amp = torch.where(kk > 0, kk.clamp(min=1e-30) ** slope, torch.zeros_like(kk))
    vel = torch.empty((3, n, n, n), dtype=torch.float64, device=device)
    for d in range(3):
        white = torch.randn((n, n, n), generator=gen, dtype=torch.float64, device=device)
        vel[d] = torch.fft.irfftn(torch.fft.rfftn(white) * amp, s=(n, n, n))
        del white
    vel -= vel.mean(dim=(1, 2, 3), keepdim=True)
    rms = torch.sqrt((vel * vel).sum(0).mean())
    vel *= mach / rms  # c_s = 1
    if rho_sigma > 0:
        rho = torch.exp(rho_sigma * torch.randn((n, n, n), generator=gen, dtype=torch.float64, device=device))
    else:
        rho = torch.ones((n, n, n), dtype=torch.float64, device=device)
    p = rho / gamma
    u = torch.empty((5, n, n, n), dtype=torch.float64, device=device)
    u[0] = rho
    u[1:4] = rho * vel
    u[4] = p / (gamma - 1.0) + 0.5 * rho * (vel * vel).sum(0)
    del vel
    return u


def turbulent_host(n, mach, gamma=5.0 / 3.0, rho_sigma=0.0, seed=0, slope=-2.0, workers=None):
    """The k^-4 velocity field of ``turbulent`` as a host numpy array (5, n,
    n, n), from ``numpy.random.default_rng(seed)``: three standard-normal
    cubes (vx, vy, vz), each filtered in Fourier space by |k|^slope with the
    k = 0 mode removed, the mean removed, scaled to 3D rms Mach `mach` for
    c_s = 1; then, when rho_sigma > 0, rho = exp(rho_sigma * N(0,1)) per
    cell from the same stream, else rho = 1; P = rho / gamma."""
    import numpy as np
    import scipy.fft as sfft
    if workers is None:
        workers = min(16, os.cpu_count() or 1)
    rng = np.random.default_rng(seed)
    k1 = np.fft.fftfreq(n, d=1.0 / n)
    kr = np.fft.rfftfreq(n, d=1.0 / n)
    kk = np.sqrt(k1[:, None, None] ** 2 + k1[None, :, None] ** 2 + kr[None, None, :] ** 2)
    amp = np.zeros_like(kk)
    np.power(kk, slope, out=amp, where=kk > 0)
    vel = np.empty((3, n, n, n))
    for d in range(3):
        white = rng.standard_normal((n, n, n))
        spec = sfft.rfftn(white, workers=workers)
        del white
        spec *= amp
        vel[d] = sfft.irfftn(spec, s=(n, n, n), workers=workers)
        del spec
    vel -= vel.mean(axis=(1, 2, 3), keepdims=True)
    rms = np.sqrt((vel * vel).sum(0).mean())
    vel *= mach / rms  # c_s = 1
    if rho_sigma > 0:
        rho = np.exp(rho_sigma * rng.standard_normal((n, n, n)))
    else:
        rho = np.ones((n, n, n))
    u = np.empty((5, n, n, n))
    u[0] = rho
    for d in range(3):
        u[1 + d] = rho * vel[d]
    ek = vel[0] * vel[0]
    ek += vel[1] * vel[1]
    ek += vel[2] * vel[2]
    del vel
    ek *= 0.5 * rho
    u[4] = rho / gamma / (gamma - 1.0) + ek
    return u


def oct_batches(noct, nv=32, seed=1, gamma=5.0 / 3.0, dx=1.0 / 128, cfl=0.5, device="cuda", chunk=2048):
    """Synthetic gathered oct batches for the batched unsplit entry point:
    uin[g][5][6][6][6][nv] (the reference's uin layout per batch, axes
    reversed) for noct // nv batches, every stencil cell drawn i.i.d. as
    rho ~ logN(0, 1), u, v, w ~ N(0, 1), P ~ logN(-0.5, 0.5) and converted
    to conserved variables; and dt = the smallest cmpdt time step at CFL
    `cfl` without gravity (hydro/godunov_utils.f90:5-120) over the 8 centre
    cells of every oct.  The noise is torch Philox drawn on the device in
    chunks of `chunk` batches (2^20 octs, 9 GB of stencils, in well under a
    second; oct_batches_host draws the same distribution with numpy in about
    30 s)."""
    import torch
    G = noct // nv
    uin = torch.empty((G, 5, 6, 6, 6, nv), dtype=torch.float64, device=device)
    gen = torch.Generator(device=device)
    gen.manual_seed(seed)
    dtmin = float("inf")
    for g0 in range(0, G, chunk):
        g1 = min(G, g0 + chunk)
        shp = (g1 - g0, 6, 6, 6, nv)
        rho = torch.exp(torch.randn(shp, generator=gen, device=device, dtype=torch.float64))
        v = torch.randn((3,) + shp, generator=gen, device=device, dtype=torch.float64)
        p = torch.exp(-0.5 + 0.5 * torch.randn(shp, generator=gen, device=device, dtype=torch.float64))
        u = uin[g0:g1]
        u[:, 0] = rho
        u[:, 1:4] = (rho.unsqueeze(0) * v).permute(1, 0, 2, 3, 4, 5)
        u[:, 4] = p / (gamma - 1.0) + 0.5 * rho * (v * v).sum(0)
        c = torch.sqrt(gamma * p[:, 1:3, 1:3, 1:3] / rho[:, 1:3, 1:3, 1:3])
        w = 3.0 * c + v[:, :, 1:3, 1:3, 1:3].abs().sum(0)
        r = 1e-4
        dtc = dx / w * ((1.0 + 2.0 * cfl * r) ** 0.5 - 1.0) / r
        dtmin = min(dtmin, float(dtc.min()))
        del rho, v, p, c, w, dtc
    return uin, dtmin


def oct_batches_host_chunk(rng, ngroups, nv=32, gamma=5.0 / 3.0, dx=1.0 / 128, cfl=0.5):
    """ngroups batches of the distribution of oct_batches drawn with numpy:
    rho = exp(N(0,1)), then the three velocity components, then P =
    exp(-0.5 + 0.5 N(0,1)), each as one standard_normal call of shape
    (ngroups, 6, 6, 6, nv) in that order from `rng`.  Returns the conserved
    batches uin[g][5][6][6][6][nv] and the chunk's smallest cmpdt time step
    (see oct_batches)."""
    import numpy as np
    shp = (ngroups, 6, 6, 6, nv)
    rho = np.exp(rng.standard_normal(shp))
    v = rng.standard_normal((3,) + shp)
    p = np.exp(-0.5 + 0.5 * rng.standard_normal(shp))
    u = np.empty((ngroups, 5) + shp[1:])
    u[:, 0] = rho
    u[:, 1:4] = np.moveaxis(rho[None] * v, 0, 1)
    u[:, 4] = p / (gamma - 1.0) + 0.5 * rho * (v * v).sum(0)
    c = np.sqrt(gamma * p[:, 1:3, 1:3, 1:3] / rho[:, 1:3, 1:3, 1:3])
    w = 3.0 * c + np.abs(v[:, :, 1:3, 1:3, 1:3]).sum(0)
    r = 1e-4
    dtc = dx / w * ((1.0 + 2.0 * cfl * r) ** 0.5 - 1.0) / r
    return u, float(dtc.min())


def oct_batches_host(noct, nv=32, seed=1, gamma=5.0 / 3.0, dx=1.0 / 128, cfl=0.5, device="cuda", chunk=1024,
                     progress=None):
    """oct_batches with numpy default_rng(seed) noise, drawn on the host in
    chunks of `chunk` batches (oct_batches_host_chunk, chunk after chunk from
    one generator) and copied to `device`, so the inputs can be regenerated
    without torch.  The first chunk alone gives the first `chunk` batches."""
    import numpy as np
    import torch
    G = noct // nv
    uin = torch.empty((G, 5, 6, 6, 6, nv), dtype=torch.float64, device=device)
    rng = np.random.default_rng(seed)
    dtmin = float("inf")
    for g0 in range(0, G, chunk):
        g1 = min(G, g0 + chunk)
        u, dt = oct_batches_host_chunk(rng, g1 - g0, nv, gamma, dx, cfl)
        uin[g0:g1].copy_(torch.from_numpy(u))
        dtmin = min(dtmin, dt)
        if progress and (g0 // chunk) % 8 == 7:
            progress(f"{g1} of {G} batches drawn")
    return uin, dtmin


def sound_crossing_time(boxlen=1.0, cs=1.0):
    return boxlen / cs


__all__ = ["sedov", "turbulent", "sound_crossing_time", "math"]


def _div(a, b):
    """a / b rounded as IEEE division.  torch's GPU kernels turn division
    by a host scalar into a multiplication by its reciprocal, which is not
    the reference's rounding unless b is a power of two."""
    import torch
    return a / torch.full_like(a, b)


def region_condinit(x, dx, regions, gamma=1.4, smallr=1e-10, smallc=1e-10):
    """condinit (hydro/condinit.f90) with the region initial conditions of
    &INIT_PARAMS (region_condinit, hydro/init_flow_fine.f90:475-596) at cell
    centres x (N, 3) of cell size dx: start from the floor state (smallr, 0,
    smallr*smallc^2/gamma); a 'square' region (exp_region < 10: the
    exp-norm of 2|x-c|/length, else the max norm, < 1) sets its state, a
    'point' region adds its CIC share d*r/dx^3, u*r, p*r/dx^3 with
    r = prod max(1 - |x-c|/dx, 0).  Regions are namelist.Region-like
    (type, center, length, exp, d, u, p).  Returns conserved (5, N)."""
    import torch
    n = x.shape[0]
    q1 = torch.full((n,), smallr, dtype=x.dtype, device=x.device)
    qv = [torch.zeros(n, dtype=x.dtype, device=x.device) for _ in range(3)]
    q5 = torch.full((n,), smallr * smallc ** 2 / gamma, dtype=x.dtype, device=x.device)
    for reg in regions:
        c = torch.tensor(reg.center, dtype=x.dtype, device=x.device)
        if reg.type == "square":
            ln = torch.tensor(reg.length, dtype=x.dtype, device=x.device)
            xn = 2.0 * (x - c).abs() / ln.expand_as(x)
            if reg.exp < 10:
                r = (xn[:, 0] ** reg.exp + xn[:, 1] ** reg.exp + xn[:, 2] ** reg.exp) ** (1.0 / reg.exp)
            else:
                r = torch.maximum(torch.maximum(xn[:, 0], xn[:, 1]), xn[:, 2])
            inside = r < 1.0
            q1 = torch.where(inside, torch.full_like(q1, reg.d), q1)
            qv = [torch.where(inside, torch.full_like(q1, reg.u[d]), qv[d]) for d in range(3)]
            q5 = torch.where(inside, torch.full_like(q1, reg.p), q5)
        elif reg.type == "point":
            vol = dx * dx * dx
            w = torch.clamp(1.0 - _div((x - c).abs(), dx), min=0.0)
            r = w[:, 0] * w[:, 1] * w[:, 2]
            q1 = q1 + _div(reg.d * r, vol)
            qv = [qv[d] + reg.u[d] * r for d in range(3)]
            q5 = q5 + _div(reg.p * r, vol)
        else:
            raise ValueError(f"region_type {reg.type!r} is not implemented")
    u = torch.empty((5, n), dtype=x.dtype, device=x.device)
    u[0] = q1
    e = torch.zeros(n, dtype=x.dtype, device=x.device)
    for d in range(3):
        u[d + 1] = q1 * qv[d]
        e = e + (0.5 * q1) * (qv[d] * qv[d])
    u[4] = e + _div(q5, gamma - 1.0)
    return u


class _Reg:
    def __init__(self, type, center, length, exp, d, u, p):
        self.type, self.center, self.length, self.exp, self.d, self.u, self.p = type, center, length, exp, d, u, p


def sedov_regions(x, dx, gamma=1.4, d_amb=1.0, p_amb=1e-5, p_point=0.4, center=(0.5, 0.5, 0.5), boxlen=1.0):
    """The two regions of namelist/sedov3d.nml -- a 'square' covering the
    box (d_amb, p_amb) and a 'point' at `center` depositing p_point -- through
    region_condinit; zero velocity."""
    regs = [_Reg("square", (0.5 * boxlen,) * 3, (10.0 * boxlen,) * 3, 10.0, d_amb, (0.0, 0.0, 0.0), p_amb),
            _Reg("point", center, (1.0, 1.0, 1.0), 10.0, 0.0, (0.0, 0.0, 0.0), p_point)]
    return region_condinit(x, dx, regs, gamma=gamma)
