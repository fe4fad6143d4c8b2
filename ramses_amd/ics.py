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
  and an isothermal pressure P = rho/gamma.
Both return conserved state u[var][z][y][x] as a float64 tensor.
* ``region_condinit(x, dx, regions)`` -- the &INIT_PARAMS regions evaluated
  at arbitrary cell centres x (N, 3) with the cell size dx of their level,
  as init_flow_fine calls condinit on every level of an oct tree; returns
  the conserved state (5, N).  ``sedov_regions`` is the sedov3d.nml case.
"""
import math


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


def turbulent(n, mach, gamma=5.0 / 3.0, rho_sigma=0.0, seed=0, device="cuda", slope=-2.0):
    import torch
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
