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
* ``sedov_regions(x, dx)`` -- the same Sedov set-up evaluated at arbitrary
  cell centres x (N, 3) with the cell size dx of their level, as
  init_flow_fine calls condinit on every level of an oct tree; returns the
  conserved state (5, N).
"""
import math


def sedov(n, gamma=1.4, e_total=1.0, device="cuda"):
    import torch
    dx = 1.0 / n
    vol = dx ** 3
    xc = (torch.arange(n, dtype=torch.float64, device=device) + 0.5) * dx
    w = torch.clamp(1.0 - (xc - 0.5).abs() / dx, min=0.0)
    r = w[:, None, None] * w[None, :, None] * w[None, None, :]
    p = 1e-5 + e_total * (gamma - 1.0) * r / vol
    u = torch.zeros((5, n, n, n), dtype=torch.float64, device=device)
    u[0] = 1.0
    u[4] = p / (gamma - 1.0)
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


def sedov_regions(x, dx, gamma=1.4, d_amb=1.0, p_amb=1e-5, p_point=0.4, center=(0.5, 0.5, 0.5)):
    """region_condinit (hydro/init_flow_fine.f90:475-596) for the two regions
    of namelist/sedov3d.nml -- a 'square' covering the box (d_amb, p_amb)
    and a 'point' at `center` adding p_point*r/dx^3 with the CIC weight
    r = prod max(1 - |x - c|/dx, 0) -- then condinit's conversion to
    conserved variables (hydro/condinit.f90:36-71); zero velocity."""
    import torch
    w = torch.clamp(1.0 - (x - torch.tensor(center, dtype=x.dtype, device=x.device)).abs() / dx, min=0.0)
    r = w[:, 0] * w[:, 1] * w[:, 2]
    vol = dx * dx * dx
    rho = torch.full_like(r, d_amb) + 0.0 * r / vol
    p = torch.full_like(r, p_amb) + p_point * r / vol
    u = torch.zeros((5, x.shape[0]), dtype=x.dtype, device=x.device)
    u[0] = rho
    u[4] = p / (gamma - 1.0)
    return u
