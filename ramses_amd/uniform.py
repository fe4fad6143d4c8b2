"""Resident uniform periodic level (levelmin == levelmax).

Mirrors the reference's per-level step on a fully refined periodic grid
(amr/amr_step.f90:327-424): ``newdt_fine -> courant_fine`` gives dt
(pm/newdt_fine.f90:51,177; hydro/courant_fine.f90:1-159), ``set_unew``,
``godunov_fine`` sweeps ``godfine1`` over every oct
(hydro/godunov_fine.f90:5-35,486-911) and ``set_uold`` copies the result
back.  Here the conserved state stays in device memory as
``u[var][z][y][x]``; set_unew/set_uold become a swap of two buffers and the
courant reduction of the new state is fused into the update kernel, so a
step is one kernel launch and no host synchronisation.

torch is used only to own device memory and streams.
"""
import ctypes

from . import _check, _stream_handle, load_library


def _bind_sums(lib):
    if getattr(lib, "_sums_bound", False):
        return
    vp = ctypes.c_void_p
    lib.ramses_uniform_work_bytes.argtypes = [ctypes.c_int]
    lib.ramses_uniform_work_bytes.restype = ctypes.c_ulong
    lib.ramses_uniform_courant_sums.argtypes = [vp, ctypes.c_int, ctypes.c_double, ctypes.c_double, vp, vp, vp,
                                                ctypes.c_ulong, vp]
    lib.ramses_uniform_courant_sums.restype = ctypes.c_int
    lib.ramses_uniform_godunov_sums.argtypes = [vp, vp, ctypes.c_int, ctypes.c_double, vp, vp, vp, vp,
                                                ctypes.c_ulong, vp]
    lib.ramses_uniform_godunov_sums.restype = ctypes.c_int
    lib._sums_bound = True


class UniformLevel:
    """Conserved state of an n^3 periodic level resident on the GPU."""

    def __init__(self, u0, boxlen=1.0, smallc=1e-10, device="cuda", stream=None, sums=False):
        """sums=True also computes courant_fine's conservation sums
        (hydro/courant_fine.f90:95-119) of the current state with every dt
        reduction: ``self.sums`` = (mass, sum E*vol, internal energy) on the
        device, after each step those of the new state."""
        import torch
        n = u0.shape[-1]
        if tuple(u0.shape) != (5, n, n, n):
            raise ValueError("u0 must have shape (5, n, n, n)")
        self.n = n
        self.dx = boxlen / n
        self.boxlen = boxlen
        self.smallc = smallc
        self.stream = stream
        t = torch.as_tensor(u0, dtype=torch.float64)
        self.u = t.to(device).contiguous()
        if self.u.data_ptr() == getattr(u0, "data_ptr", lambda: None)():
            self.u = self.u.clone()
        self.unew = torch.empty_like(self.u)
        # dt slots: [0] current step, [1] fused reduction for the next step
        self.dt = torch.zeros(2, dtype=torch.float64, device=device)
        self.lib = load_library()
        _bind_sums(self.lib)
        self.sums = None
        self._work = None
        if sums:
            self.sums = torch.zeros(3, dtype=torch.float64, device=device)
            nbytes = int(self.lib.ramses_uniform_work_bytes(n))
            self._work = torch.empty((nbytes + 7) // 8, dtype=torch.float64, device=device)
        self._dt_ready = False
        self.steps = 0
        self.t = 0.0

    def _ptr(self, t, off=0):
        return ctypes.c_void_p(t.data_ptr() + 8 * off)

    def courant(self, dt_start=None):
        """courant_fine on the current state; returns dt (host float)."""
        start = self.boxlen / self.smallc if dt_start is None else dt_start
        s = _stream_handle(self.stream)
        _check(self.lib.ramses_uniform_courant_sums(self._ptr(self.u), self.n, self.dx, start,
                                                    self._ptr(self.dt, 0), *self._sums_args(), s),
               "ramses_uniform_courant")
        self._dt_ready = True
        return self.dt[0]

    def step(self, fused=True, events=None):
        """One level step: dt = courant(u); u <- godunov(u, dt).  Returns the
        device tensor holding the dt used (no host synchronisation).
        ``events`` = (start, end) torch events recorded around the update
        kernel alone."""
        s = _stream_handle(self.stream)
        if not self._dt_ready:
            self.courant()
        d_dt = self._ptr(self.dt, 0)
        d_next = self._ptr(self.dt, 1)
        if fused:
            # newdt_fine restarts every step from boxlen/smallc (pm/newdt_fine.f90:51);
            # the C*dx/smallc cap of every cmpdt batch is applied at init
            import torch  # noqa: F401
            start = min(self.boxlen / self.smallc, self._courant_cap())
            _check(self.lib.ramses_uniform_dt_init(d_next, start, s), "ramses_uniform_dt_init")
        if events is not None:
            events[0].record(self._torch_stream())
        _check(self.lib.ramses_uniform_godunov_sums(self._ptr(self.u), self._ptr(self.unew), self.n, self.dx,
                                                    d_dt, d_next if fused else None,
                                                    *(self._sums_args() if fused else (None, None, 0)), s),
               "ramses_uniform_godunov")
        if events is not None:
            events[1].record(self._torch_stream())
        used = self.dt[0].clone()
        self.u, self.unew = self.unew, self.u
        if fused:
            self.dt[0:1].copy_(self.dt[1:2])
        else:
            self._dt_ready = False
        self.steps += 1
        return used

    def advance(self, nsteps, graph=True):
        """nsteps fused steps.  With graph=True a pair of steps (the buffer
        swap returns to the same assignment after two) is captured once in a
        hipGraph and replayed, so a launch-bound small level pays one graph
        launch per two steps instead of six launches.  Bitwise the same as
        calling step() nsteps times."""
        import torch
        if not self._dt_ready:
            self.courant()
        if not graph or nsteps < 2:
            for _ in range(nsteps):
                self.step()
            return
        if getattr(self, "_graph", None) is None:
            u_before = self.u
            g = torch.cuda.CUDAGraph()
            torch.cuda.synchronize()
            with torch.cuda.graph(g):
                self.stream = torch.cuda.current_stream()
                self.step()
                self.step()
            self.stream = None
            assert self.u is u_before
            self.steps -= 2  # capture does not execute the steps
            self._graph = g
        for _ in range(nsteps // 2):
            self._graph.replay()
            self.steps += 2
        if nsteps % 2:
            self.step()

    def update_only(self):
        """Launch the update kernel into the spare buffer without advancing
        the level (used by the autotuner to time launch shapes)."""
        s = _stream_handle(self.stream)
        if not self._dt_ready:
            self.courant()
        _check(self.lib.ramses_uniform_godunov_sums(self._ptr(self.u), self._ptr(self.unew), self.n, self.dx,
                                                    self._ptr(self.dt, 0), self._ptr(self.dt, 1),
                                                    *self._sums_args(), s),
               "ramses_uniform_godunov")

    def _sums_args(self):
        if self.sums is None:
            return (None, None, 0)
        return (self._ptr(self.sums), self._ptr(self._work), self._work.numel() * 8)

    def _torch_stream(self):
        import torch
        return torch.cuda.current_stream() if self.stream is None else self.stream

    def _courant_cap(self):
        from . import _config_cache
        return _config_cache["courant_factor"] * self.dx / _config_cache["smallc"]
