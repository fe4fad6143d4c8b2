"""Continuous-batching server for independent hydro solves on one GPU.

Each request is a periodic uniform level (side n, a multiple of 16) advanced
from t=0 to its end time by the reference's per-level step
(amr/amr_step.f90:327-424), each with its own courant dt; the last step is
shortened to end exactly at the request's end time, a cap the reference does
not have (its newdt_fine ignores output times).  Requests that
arrive while others run are admitted as soon as a slot and enough HBM are
free; every scheduler tick then advances *all* in-flight requests with one
bookkeeping launch and one update launch (``ramses_level_batch_step``), so
their work shares the device inside the same kernel.  A finished request
frees its slot and its buffers at once for the next one.

The host never waits for the tick it just launched: a copy of the slot states
into pinned memory is queued behind each tick and read one tick later, so
completions are seen with a delay of one tick while the GPU keeps running.

Results depend only on each request's own data (per-slot dt, no reduction
across slots), so a request gives the same bits alone and in any mix.
"""
import ctypes
import collections
import time

from . import _check, _stream_handle, load_library

SLOT_DOUBLES = 8
T_TIME, T_TEND, T_DT, T_DTNEXT, T_DONE, T_STEPS, T_PARITY = 0, 1, 2, 3, 4, 5, 6
MAX_SLOTS = 32


class ramses_level_slot(ctypes.Structure):
    _fields_ = [("buf", ctypes.c_void_p * 2), ("state", ctypes.c_void_p), ("n", ctypes.c_int),
                ("dx", ctypes.c_double)]


class Request:
    """One solve: side n, IC generator, end time, arrival time (s)."""

    def __init__(self, rid, n, t_end, arrival=0.0, ic=None, keep_result=False):
        self.id = rid
        self.n = n
        self.t_end = t_end
        self.arrival = arrival
        self.ic = ic                  # callable(n) -> device tensor (5, n, n, n)
        self.keep_result = keep_result
        self.admitted = None
        self.finished = None
        self.steps = None
        self.result = None
        self.checksum = None

    @property
    def latency(self):
        return None if self.finished is None else self.finished - self.arrival

    @property
    def bytes(self):
        return 2 * 5 * self.n ** 3 * 8


def admit(queue, free_slots, used_bytes, budget):
    """Admission policy: first come first served, as long as a slot is free
    and the request's two state buffers fit in the HBM budget.  A request
    that does not fit waits (head-of-line) so large solves are not starved
    by a stream of small ones."""
    out = []
    while queue and free_slots > len(out):
        r = queue[0]
        if used_bytes + r.bytes > budget:
            break
        out.append(queue.popleft())
        used_bytes += r.bytes
    return out


class _Pool:
    """Buffers per size class, reused across requests (no allocation in the
    steady state)."""

    def __init__(self, device):
        self.free = collections.defaultdict(list)
        self.device = device

    def get(self, n):
        import torch
        if self.free[n]:
            return self.free[n].pop()
        return (torch.empty((5, n, n, n), dtype=torch.float64, device=self.device),
                torch.empty((5, n, n, n), dtype=torch.float64, device=self.device))

    def put(self, n, bufs):
        self.free[n].append(bufs)


class HydroServer:
    def __init__(self, hbm_budget=64 << 30, max_slots=MAX_SLOTS, device="cuda", stream=None, steps_per_tick=8,
                 use_graph=False, graph_after=3):
        import torch
        if not 1 <= max_slots <= MAX_SLOTS:
            raise ValueError("max_slots out of range")
        self.lib = load_library()
        self.lib.ramses_level_slot_init.argtypes = [ctypes.POINTER(ramses_level_slot), ctypes.c_double,
                                                    ctypes.c_void_p]
        self.lib.ramses_level_slot_init.restype = ctypes.c_int
        self.lib.ramses_level_batch_step.argtypes = [ctypes.POINTER(ramses_level_slot), ctypes.c_int,
                                                     ctypes.c_void_p]
        self.lib.ramses_level_batch_step.restype = ctypes.c_int
        self.budget = hbm_budget
        self.max_slots = max_slots
        self.device = device
        self.stream = stream
        self.state = torch.zeros((max_slots, SLOT_DOUBLES), dtype=torch.float64, device=device)
        self.host_state = torch.zeros((max_slots, SLOT_DOUBLES), dtype=torch.float64)
        self.pool = _Pool(device)
        self.ticks = 0
        self.steps_per_tick = steps_per_tick
        self.use_graph = use_graph
        self.graph_after = graph_after
        self.graph_replays = 0
        self.graph_captures = 0
        self.max_steps = 1_000_000
        self.max_used = 0
        self.max_active = 0

    def _slot_struct(self, row, bufs, n):
        s = ramses_level_slot()
        s.buf[0] = bufs[0].data_ptr()
        s.buf[1] = bufs[1].data_ptr()
        s.state = self.state.data_ptr() + 8 * SLOT_DOUBLES * row
        s.n = n
        s.dx = 1.0 / n
        return s

    def run(self, requests, clock=time.perf_counter):
        """Serve `requests` (sorted by arrival) to completion; returns them."""
        import torch
        stream = self.stream if self.stream is not None else torch.cuda.current_stream()
        sh = _stream_handle(stream)
        pending = collections.deque(sorted(requests, key=lambda r: r.arrival))
        queue = collections.deque()
        free_rows = list(range(self.max_slots))[::-1]
        active = {}          # row -> (request, bufs, slot struct, admission tick)
        used = 0
        snaps = [torch.zeros_like(self.host_state, pin_memory=True) for _ in range(2)]
        events = [None, None]
        graph, graph_key, same = None, None, 0
        t0 = clock()
        while pending or queue or active:
            now = clock() - t0
            while pending and pending[0].arrival <= now:
                queue.append(pending.popleft())
            for r in admit(queue, len(free_rows), used, self.budget):
                row = free_rows.pop()
                bufs = self.pool.get(r.n)
                bufs[0].copy_(r.ic(r.n))
                st = self._slot_struct(row, bufs, r.n)
                _check(self.lib.ramses_level_slot_init(ctypes.byref(st), r.t_end, sh), "ramses_level_slot_init")
                r.admitted = clock() - t0
                active[row] = (r, bufs, st, self.ticks)
                used += r.bytes
            self.max_used = max(self.max_used, used)
            self.max_active = max(self.max_active, len(active))
            if not active:
                if pending and not queue:
                    wait = pending[0].arrival - (clock() - t0)
                    if wait > 0:
                        time.sleep(min(wait, 0.005))
                continue
            # tick k: advance every in-flight solve by steps_per_tick steps
            # (solves that reach their end time stop inside the tick), then
            # queue a pinned copy of the slot states behind it.  While the set
            # of in-flight solves does not change the tick is replayed from a
            # captured hipGraph (one launch instead of 2*steps_per_tick).
            k = self.ticks
            rows = tuple(sorted(active))
            key = tuple((i, active[i][0].id) for i in rows)
            if self.use_graph and graph_key == key and graph is not None:
                graph.replay()
                self.graph_replays += 1
            else:
                arr = (ramses_level_slot * len(rows))(*[active[i][2] for i in rows])
                same += 1 if graph_key == key else -same
                if self.use_graph and same >= self.graph_after:
                    # the same set for several ticks: capture a tick (on
                    # torch's capture stream) and replay it from now on
                    graph = torch.cuda.CUDAGraph()
                    with torch.cuda.graph(graph):
                        gsh = _stream_handle(None)
                        for _ in range(self.steps_per_tick):
                            _check(self.lib.ramses_level_batch_step(arr, len(rows), gsh), "ramses_level_batch_step")
                    self.graph_captures += 1
                    graph.replay()
                    self.graph_replays += 1
                else:
                    graph = None
                    for _ in range(self.steps_per_tick):
                        _check(self.lib.ramses_level_batch_step(arr, len(rows), sh), "ramses_level_batch_step")
                graph_key = key
            snaps[k % 2].copy_(self.state, non_blocking=True)
            ev = torch.cuda.Event()
            ev.record(stream)
            events[k % 2] = ev
            self.ticks += 1
            # read tick k-1's snapshot while tick k runs (at the very first
            # tick there is none: wait for this one instead)
            j = k - 1 if k > 0 and events[(k - 1) % 2] is not None else k
            events[j % 2].synchronize()
            snap = snaps[j % 2].numpy()
            for row in list(active):
                r, bufs, st, adm = active[row]
                if adm > j:
                    continue  # snapshot older than this occupant
                if snap[row, T_DONE] == 0.0:
                    if snap[row, T_STEPS] > self.max_steps:
                        raise RuntimeError(f"request {r.id} exceeded {self.max_steps} steps (t={snap[row, T_TIME]})")
                    continue
                active.pop(row)
                r.finished = clock() - t0
                r.steps = int(snap[row, T_STEPS])
                cur = bufs[int(snap[row, T_PARITY])]
                r.checksum = torch.stack([cur[0].sum(), cur[4].sum()])  # read later, no sync here
                if r.keep_result:
                    r.result = cur.clone()
                self.pool.put(r.n, bufs)
                free_rows.append(row)
                used -= r.bytes
        torch.cuda.synchronize()
        for r in requests:
            if r.checksum is not None:
                r.checksum = tuple(float(x) for x in r.checksum.cpu())
        return requests
