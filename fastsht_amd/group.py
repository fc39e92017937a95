"""Uneven batches of jobs run as one schedule, recorded in a HIP graph.

A job is one transform of `nsim` maps of one shape (nside, lmax): forward
(t2alm refined with the group's `niter`, plus qu2eb for a polarised job) or
backward (alm2t, plus eb2qu), in FP64 or mixed precision,
on device tensors in the layouts of the `SHT` object: maps (npix, nsim) and
alm (nsim, lmax+1, lmax+1), Fortran order.

The reference runs one shape per process and one job at a time, each
transform as per-order launches on per-order streams (SHT.py:17-62,
src/sht_core.f90:48,123, src/sht_data_init.f90:259-267).  A `Group` runs a
list of jobs of mixed shapes, sizes and polarisations together:

- one library handle per shape, created for the largest job of that shape
  (spin-2 capable when a polarised job has it); every job runs on its own
  tensors through `fsht_set_batch`, so no maps or alms are copied;
- the shapes' job lists are placed on up to `streams` HIP streams, longest
  estimated work first, each on the stream with the least work so far;
  jobs of one shape run one after another on one stream (they share the
  handle's workspace), jobs of different shapes run concurrently;
- the whole schedule is recorded once as a HIP graph (forked streams joined
  back into the capture stream) and replayed: one launch for the group.
  Replays read the input tensors and write the output tensors given at
  capture, as CUDA/HIP graphs do;
- `tune()` times the graph of each candidate schedule (stream count, order
  of the shapes) and keeps the fastest.

Every job's result is the result of its own `SHT` object up to the order in
which the FP64 analysis sums a map's rings in the two column-block kinds
(about 1e-16 relative).
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass

from . import _lib
from .sht import SHT, _fortran_strides


def _torch():
    import torch

    return torch


@dataclass
class Job:
    """inputs: forward (T,) or (T, Q, U) maps; backward (T,) or (T, E, B)
    alm.  outputs: the same count of alm (forward) or maps (backward)."""
    nside: int
    lmax: int
    direction: str  # "forward" | "backward"
    inputs: tuple
    outputs: tuple
    precision: str = "fp64"

    @property
    def pol(self):
        return len(self.inputs) == 3

    @property
    def nsim(self):
        x = self.inputs[0]
        return x.shape[1] if self.direction == "forward" else x.shape[0]

    def work(self, niter):
        """Relative cost: Legendre work (rings x orders x l per map) times the
        passes the transform makes; a Q/U or E/B pair costs four temperature
        maps."""
        passes = 2 * niter + 1 if self.direction == "forward" else 1
        per_map = (2 * self.nside) * (self.lmax + 1) ** 2
        return per_map * self.nsim * passes * (5 if self.pol else 1)


def estimate_handle_bytes(nside, lmax, cap, pol, niter):
    """Upper estimate of a handle's device bytes (workspace with every ring
    kept for every order, plus tables), before it exists."""
    npix, L, nsp = 12 * nside ** 2, lmax + 1, (cap + 15) // 16 * 16
    ncomp = 2 if pol else 1
    g = 2 * L * (2 * nside) * 4 * nsp // (1 if pol else 2)
    ws = ncomp * cap * npix + npix * ncomp * nsp + g * (2 if niter > 0 else 1)
    return 8 * ws + 64 * L * L + 8 * npix + 0.3e9


def job_bytes(nside, lmax, nsim, pol):
    """Device bytes of a job's inputs and outputs."""
    return 8 * (3 if pol else 1) * nsim * (12 * nside ** 2 + (lmax + 1) ** 2)


class Group:
    def __init__(self, jobs, niter=0, streams=4):
        torch = _torch()
        if not jobs:
            raise ValueError("a group needs at least one job")
        for j in jobs:
            if j.direction not in ("forward", "backward"):
                raise ValueError(f"direction must be 'forward' or 'backward', got {j.direction!r}")
            if j.precision not in ("fp64", "mixed"):
                raise ValueError(f"precision must be 'fp64' or 'mixed', got {j.precision!r}")
            if len(j.inputs) not in (1, 3) or len(j.outputs) != len(j.inputs):
                raise ValueError("a job has one (T) or three (T, Q/E, U/B) inputs and as many outputs")
        self.jobs, self.niter = list(jobs), niter
        self.lib = _lib.load()
        self.shapes = {}  # (nside, lmax) -> [job indices in order]
        for i, j in enumerate(self.jobs):
            self.shapes.setdefault((j.nside, j.lmax), []).append(i)
        self.handles = {}
        for key, idx in self.shapes.items():
            cap = max(self.jobs[i].nsim for i in idx)
            pol = any(self.jobs[i].pol for i in idx)
            sht = SHT(key[0], key[1], cap, niter=niter, pol=pol)
            if any(self.jobs[i].precision == "mixed" for i in idx):
                sht.set_precision("mixed")  # builds the int8 tables now, outside any capture
            self.handles[key] = sht
        self.streams = [torch.cuda.Stream() for _ in range(max(1, streams))]
        self.plan = self._plan(len(self.streams), order="work")
        self.graph = None
        self._bound = {}  # handle -> stream it was set to
        torch.cuda.synchronize()

    # ---- schedule ---------------------------------------------------------
    def _plan(self, nstreams, order="work"):
        """Shapes per stream: longest-processing-time first ("work") or in
        the order of the jobs ("given")."""
        keys = list(self.shapes)
        cost = {k: sum(self.jobs[i].work(self.niter) for i in self.shapes[k]) for k in keys}
        if order == "work":
            keys.sort(key=lambda k: -cost[k])
        load = [0.0] * nstreams
        plan = [[] for _ in range(nstreams)]
        for k in keys:
            s = min(range(nstreams), key=lambda t: load[t])
            plan[s].append(k)
            load[s] += cost[k]
        return [p for p in plan if p]

    def _bind(self, sht, stream):
        """Point the handle at the stream (outside any capture: switching
        streams waits for the handle's previous stream)."""
        s = stream.cuda_stream
        if self._bound.get(id(sht)) != s:
            _lib.check(self.lib.fsht_set_stream(sht._h, ctypes.c_void_p(s)), sht._h)
            self._bound[id(sht)] = s

    def _run_job(self, sht, j):
        lib, h, D = self.lib, sht._h, _lib.FSHT_DEVICE
        p = lambda t: ctypes.c_void_p(t.data_ptr())
        _lib.check(lib.fsht_set_precision(h, _lib.FSHT_PREC_MIXED if j.precision == "mixed" else _lib.FSHT_PREC_FP64), h)
        _lib.check(lib.fsht_set_batch(h, j.nsim), h)
        x, y = j.inputs, j.outputs
        if j.direction == "forward":
            _lib.check(lib.fsht_t2alm(h, p(x[0]), p(y[0]), self.niter, D), h)
            if j.pol:
                _lib.check(lib.fsht_qu2eb(h, p(x[1]), p(x[2]), p(y[1]), p(y[2]), self.niter, D), h)
        else:
            _lib.check(lib.fsht_alm2t(h, p(x[0]), p(y[0]), D), h)
            if j.pol:
                _lib.check(lib.fsht_eb2qu(h, p(x[1]), p(x[2]), p(y[1]), p(y[2]), D), h)

    def _check_tensors(self):
        torch = _torch()
        for j in self.jobs:
            L, npix, n = j.lmax + 1, 12 * j.nside ** 2, j.nsim
            mshape, ashape = (npix, n), (n, L, L)
            ins, outs = (mshape, ashape) if j.direction == "forward" else (ashape, mshape)
            for x, shape in [(x, ins) for x in j.inputs] + [(y, outs) for y in j.outputs]:
                if (not x.is_cuda or x.dtype != torch.float64 or tuple(x.shape) != shape
                        or tuple(x.stride()) != _fortran_strides(shape)):
                    raise ValueError(f"job ({j.nside}, {j.lmax}, {j.direction}): tensors must be float64 on the "
                                     f"GPU in Fortran order, shape {shape}")

    def _enqueue(self, plan):
        """Fork the plan's streams from the current stream, run each stream's
        shapes' jobs in order, join them back."""
        torch = _torch()
        cur = torch.cuda.current_stream()
        for s, keys in enumerate(plan):
            st = self.streams[s]
            st.wait_stream(cur)
            with torch.cuda.stream(st):
                for k in keys:
                    for i in self.shapes[k]:
                        self._run_job(self.handles[k], self.jobs[i])
        for s in range(len(plan)):
            cur.wait_stream(self.streams[s])

    def _bind_plan(self, plan):
        for s, keys in enumerate(plan):
            for k in keys:
                self._bind(self.handles[k], self.streams[s])

    # ---- public -------------------------------------------------------------
    def run(self, graph=False):
        """Run every job once.  graph=True replays the recorded schedule
        (recording it on first use: one eager run, then the capture)."""
        if not graph:
            self._check_tensors()
            self._bind_plan(self.plan)
            self._enqueue(self.plan)
            return
        if self.graph is None:
            self.graph = self._capture(self.plan)
        self.graph.replay()

    def _capture(self, plan):
        torch = _torch()
        self._check_tensors()
        self._bind_plan(plan)
        self._enqueue(plan)  # eager pass: any first-use allocation happens here
        torch.cuda.synchronize()
        g = torch.cuda.CUDAGraph()
        side = torch.cuda.Stream()
        side.wait_stream(torch.cuda.current_stream())
        with torch.cuda.graph(g, stream=side):
            self._enqueue(plan)
        torch.cuda.current_stream().wait_stream(side)
        torch.cuda.synchronize()
        return g

    def time_graph(self, g, reps=3):
        """Median ms of `reps` replays of a recorded schedule."""
        torch = _torch()
        times = []
        for _ in range(reps):
            a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            a.record()
            g.replay()
            b.record()
            b.synchronize()
            times.append(a.elapsed_time(b))
        times.sort()
        return times[len(times) // 2]

    def tune(self, reps=3, max_streams=None, orders=("work", "given"), log=None):
        """Record and time the schedule for 1..len(streams) streams, shapes
        ordered by work or as given; keep the fastest.  Returns
        [(streams, order, ms)] in the order tried."""
        n = min(len(self.streams), len(self.shapes), max_streams or len(self.streams))
        tried, best = [], None
        seen = set()
        for s in range(1, n + 1):
            for order in orders:
                plan = self._plan(s, order)
                key = tuple(tuple(p) for p in plan)
                if key in seen:
                    continue
                seen.add(key)
                g = self._capture(plan)
                ms = self.time_graph(g, reps)
                tried.append((s, order, ms))
                if log:
                    log(f"tune: {s} streams, order {order}: {ms:.1f} ms")
                if best is None or ms < best[0]:
                    best = (ms, plan, g)
                else:
                    del g
        self.plan, self.graph = best[1], best[2]
        return tried

    def synchronize(self):
        _torch().cuda.synchronize()

    def close(self):
        """Drop the graph and the handles (their device memory is freed)."""
        self.synchronize()
        self.graph = None
        self.handles = {}
        _torch().cuda.empty_cache()
