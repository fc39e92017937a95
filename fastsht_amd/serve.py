"""Continuous-batching request loop for SHT transforms of mixed shapes.

A request is one transform of `nsim` maps of one shape: forward (t2alm,
refined with `niter` Jacobi iterations, plus qu2eb for polarised requests) or
backward (alm2t, plus eb2qu), in FP64 or mixed precision,
with its inputs and outputs as device tensors in the layouts of the `SHT`
object: maps (npix, n) and alm (n, lmax+1, lmax+1), Fortran order.

The reference configures one shape per process and recomputes nothing
between calls (SHT.py:17-62, src/sht_data_init.f90:51-59); here one process
serves every shape:

- A *lane* per (nside, lmax) holds one library handle (spin-2 capable)
  created for a capacity of maps: its ring tables, Legendre seeds and (once a
  mixed request arrives) int8 Legendre tables stay resident for every later
  request of that shape, temperature-only or polarised.  Each lane works on
  its own HIP stream, so lanes of different shapes run on the device at the
  same time.
- Requests of a lane are packed in arrival order into batches of up to the
  capacity (fsht_set_batch sets the batch's map count; column tiles the batch
  leaves empty are skipped), a request larger than the room left is split
  over batches, and a batch of one direction and precision is launched as
  soon as the lane's previous batch is done.  Requests that arrive while a
  batch runs join the next one.
- Admission is by an HBM budget: a request is admitted when its inputs and
  outputs fit; a lane is created when its measured footprint fits, after
  idle lanes are closed, oldest first.

Every map of a batch is transformed independently of the others (the
contractions are per column, the mixed-precision exponents per column), so a
request's result does not depend on what it is batched with, up to rounding:
the FP64 analysis sums a map's rings in another order when the map lands in
another kind of column block (about 1e-16 relative).
"""
from __future__ import annotations

import ctypes
import math
import itertools
import time
from collections import deque
from dataclasses import dataclass, field

from . import _lib
from .sht import SHT

GB = 1e9


def _torch():
    import torch

    return torch


@dataclass
class Request:
    """One transform.  inputs: forward (T,) or (T, Q, U) maps; backward (T,)
    or (T, E, B) alm.  outputs are filled in the same order: alm for forward,
    maps for backward."""
    nside: int
    lmax: int
    direction: str  # "forward" | "backward"
    inputs: tuple
    pol: bool = False
    precision: str = "fp64"
    rid: int = -1
    arrival: float = 0.0
    done_time: float = None
    outputs: tuple = None
    precision_used: str = None
    maps_done: int = 0
    ready: object = None  # event after which the inputs are valid on the device
    n: int = 0            # maps (set at submission from the inputs)

    @property
    def nsim(self):
        if not self.n:
            x = self.inputs[0]
            self.n = int(x.shape[1] if self.direction == "forward" else x.shape[0])
        return self.n

    @property
    def done(self):
        return self.done_time is not None

    @property
    def latency(self):
        return None if self.done_time is None else self.done_time - self.arrival

    def nbytes(self):
        L = self.lmax + 1
        n = self.nsim * (3 if self.pol else 1)
        return 8 * n * (12 * self.nside ** 2 + L * L)


def estimate_lane_bytes(nside, lmax, cap, mixed=False, mixed_pol=False):
    """Device bytes of a lane (an upper estimate: spin-2 workspace, batch
    buffers of three components, tables; with `mixed`, the int8 P tables in
    both operand orders and the sliced operands of two components; with
    `mixed_pol` also the two buffers the spin-2 F1/F2 tables of a range of
    orders are generated into, at most 6 GiB each).  Admission of a request
    that needs a new lane reserves this estimate; the lane, once created, is
    charged what it holds: the library's allocation at creation (measured)
    and its batch buffers, and, when its first mixed (and first polarised
    mixed) request is admitted, the bytes the library plans for that state
    (fsht_info), which it builds at its first mixed transform of each spin."""
    npix, L, nsp = 12 * nside ** 2, lmax + 1, (cap + 15) // 16 * 16
    ws = 8 * npix * (2 * cap + 8 * nsp)
    io = 8 * 3 * cap * (npix + L * L)
    tables = 64 * L * L + 8 * npix + 0.3e9
    mix = 0
    ptab = 12 * 0.5 * L * L * 2 * nside  # spin 0: six slices, two operand orders
    if mixed or mixed_pol:
        # sliced operands of two components: AQ, and FQ unless it fits in the
        # map-major spectrum buffer (2 cap maps), where the library keeps it
        aq = 2 * 2 * (nsp // 16) * 6144 * L * (L / 64 + 1)
        fq = 2 * 2 * (nsp // 16) * 6144 * L * math.ceil(2 * nside / 32)
        mix += ptab + aq + (0 if fq <= 16 * cap * npix else fq)
    if mixed_pol:
        # spin 2: F1 and F2 buffers of one range of orders, each at most 6 GiB,
        # 1/16 of the FP64 workspace and the function's whole table
        mix += 2 * min(6 * 2 ** 30, max(2 ** 28, ws / 16), ptab / 2)
    return ws + io + tables + mix


class Lane:
    def __init__(self, server, nside, lmax, cap):
        torch = _torch()
        self.nside, self.lmax, self.cap = nside, lmax, cap
        self.npix, self.L = 12 * nside ** 2, lmax + 1
        self.stream = torch.cuda.Stream()
        torch.cuda.synchronize()
        free0 = torch.cuda.mem_get_info()[0]
        self.sht = SHT(nside, lmax, cap, niter=server.niter, pol=True)
        self.lib = _lib.load()
        _lib.check(self.lib.fsht_set_stream(self.sht._h, ctypes.c_void_p(self.stream.cuda_stream)), self.sht._h)
        torch.cuda.synchronize()
        lib_bytes = free0 - torch.cuda.mem_get_info()[0]  # the library allocates outside torch's cache
        with torch.cuda.stream(self.stream):  # batch buffers: T, then Q/U or E/B
            self.maps = torch.empty((3, self.npix * cap), dtype=torch.float64, device="cuda")
            self.alms = torch.empty((3, cap * self.L * self.L), dtype=torch.float64, device="cuda")
        self.footprint = lib_bytes + 8 * (self.maps.numel() + self.alms.numel())
        self.mixed_ready = False
        self.mixed_pol_ready = False  # charged for the spin-2 mixed-precision buffers
        self.pending = deque()  # requests (or their remaining maps) in arrival order
        self.inflight = None    # (event, chunks, precision)
        self.last_used = time.perf_counter()

    def map_view(self, c, n):
        return self.maps[c, :self.npix * n].view(n, self.npix).t()

    def alm_view(self, c, n):
        L = self.L
        return self.alms[c, :n * L * L].view(L, L, n).permute(2, 1, 0)

    def mixed_bytes(self, mixed, pol):
        """Device bytes of the lane's mixed-precision state as the library
        plans it (fsht_info): the sliced operands and temperature tables with
        `mixed`, and also the spin-2 exponents and range buffers with `pol`."""
        info = (ctypes.c_longlong * 39)()
        _lib.check(self.lib.fsht_info(self.sht._h, info, 39), self.sht._h)
        return (info[36] + info[37] if mixed or pol else 0) + (info[38] if pol else 0)

    def _switch_mixed(self):
        self.sht.set_precision("mixed")

    def enable_mixed(self, pol=False):
        """Switch the lane's handle to mixed precision (once) and charge the
        lane for the state its requests will build: the temperature tables and
        sliced operands, and with `pol` the spin-2 buffers (the library builds
        each part at the first transform that needs it)."""
        if not self.mixed_ready:
            self._switch_mixed()
        pol = pol or self.mixed_pol_ready
        self.footprint += self.mixed_bytes(True, pol) - self.mixed_bytes(self.mixed_ready, self.mixed_pol_ready)
        self.mixed_ready, self.mixed_pol_ready = True, pol

    def take_batch(self):
        """The next batch: the oldest pending requests of the oldest request's
        direction and precision, in arrival order, up to the capacity (the
        last one possibly in part).  Returns the key, the chunks (request,
        first map, maps, T offset, pair offset) and the T and pair counts."""
        head = self.pending[0]
        key = (head.direction, head.precision_used)
        chunks, n, npol = [], 0, 0
        for r in list(self.pending):
            if (r.direction, r.precision_used) != key or n == self.cap:
                continue
            k = min(r.nsim - r.maps_done, self.cap - n)
            chunks.append((r, r.maps_done, k, n, npol))
            n += k
            npol += k if r.pol else 0
        for r, start, k, _, _ in chunks:  # requests fully packed leave the queue
            if start + k == r.nsim:
                self.pending.remove(r)
        return key, chunks, n, npol

    def admit(self, r):
        """Allocate the request's outputs on the lane's stream."""
        torch = _torch()
        L, n, npix = r.lmax + 1, r.nsim, 12 * r.nside ** 2
        with torch.cuda.stream(self.stream):
            if r.direction == "forward":
                r.outputs = tuple(torch.empty((L, L, n), dtype=torch.float64, device="cuda").permute(2, 1, 0)
                                  for _ in r.inputs)
            else:
                r.outputs = tuple(torch.empty((n, npix), dtype=torch.float64, device="cuda").t() for _ in r.inputs)
        self.pending.append(r)

    def launch(self, niter):
        """Pack the next batch (temperature maps of every request, then the
        Q/U or E/B pairs of the polarised ones) and enqueue it on the lane's
        stream."""
        torch = _torch()
        key, chunks, n, npol = self.take_batch()
        direction, precision = key
        fwd = direction == "forward"
        p = lambda t: ctypes.c_void_p(t.data_ptr())
        h, lib, D = self.sht._h, self.lib, _lib.FSHT_DEVICE
        with torch.cuda.stream(self.stream):
            for r, _, _, _, _ in chunks:
                if r.ready is not None:
                    self.stream.wait_event(r.ready)
            # views of the batch buffers: component c holds n (T) or npol (pairs) maps
            cnt = (n, npol, npol)
            mv = [self.map_view(c, cnt[c]) for c in range(3)]
            av = [self.alm_view(c, cnt[c]) for c in range(3)]
            src = mv if fwd else av
            for r, start, k, off, offp in chunks:  # pack
                for c, x in enumerate(r.inputs):
                    o = off if c == 0 else offp
                    if fwd:
                        src[c][:, o:o + k].copy_(x[:, start:start + k])
                    else:
                        src[c][o:o + k].copy_(x[start:start + k])
            _lib.check(lib.fsht_set_precision(h, _lib.FSHT_PREC_MIXED if precision == "mixed"
                                              else _lib.FSHT_PREC_FP64), h)
            _lib.check(lib.fsht_set_batch(h, n), h)
            if fwd:
                _lib.check(lib.fsht_t2alm(h, p(mv[0]), p(av[0]), niter, D), h)
            else:
                _lib.check(lib.fsht_alm2t(h, p(av[0]), p(mv[0]), D), h)
            if npol:
                _lib.check(lib.fsht_set_batch(h, npol), h)
                if fwd:
                    _lib.check(lib.fsht_qu2eb(h, p(mv[1]), p(mv[2]), p(av[1]), p(av[2]), niter, D), h)
                else:
                    _lib.check(lib.fsht_eb2qu(h, p(av[1]), p(av[2]), p(mv[1]), p(mv[2]), D), h)
            dst = av if fwd else mv
            for r, start, k, off, offp in chunks:  # unpack into the requests' outputs
                for c, y in enumerate(r.outputs):
                    o = off if c == 0 else offp
                    if fwd:
                        y[start:start + k].copy_(dst[c][o:o + k])
                    else:
                        y[:, start:start + k].copy_(dst[c][:, o:o + k])
            ev = torch.cuda.Event()
            ev.record(self.stream)
        self.inflight = (ev, chunks)
        self.last_used = time.perf_counter()
        return n

    def close(self):
        torch = _torch()
        self.stream.synchronize()
        self.sht = None
        self.maps = self.alms = None
        torch.cuda.empty_cache()


@dataclass
class Server:
    """hbm_budget: device bytes the loop may hold (lanes, batch buffers,
    request inputs and outputs).  capacity(nside, lmax) -> maps per
    batch (default: 64, down to 16 where a lane's estimate would pass a
    quarter of the budget).  Mixed requests run in mixed precision for
    mixed_min_nside <= nside <= mixed_max_nside and in FP64 otherwise
    (precision_used says which): above mixed_max_nside the int8 tables do not
    fit the budget, and below mixed_min_nside the transforms are too short
    for the int8 contractions to pay (t2alm(3) + alm2t on one MI355X: mixed
    0.88x FP64 at nside 128, 1.06x at 256, 1.10x at 512, 1.32x at 1024 for
    32-64 maps), so FP64 is as fast and exact."""
    hbm_budget: float = 200 * GB
    niter: int = 3
    mixed_min_nside: int = 256
    mixed_max_nside: int = 1024
    capacity: object = None
    lanes: dict = field(default_factory=dict)
    waiting: deque = field(default_factory=deque)   # submitted, not yet admitted
    held: float = 0.0                               # bytes of admitted request buffers
    _ids: object = field(default_factory=itertools.count)
    _t0: float = field(default_factory=time.perf_counter)
    batches: int = 0
    evictions: int = 0
    lane_factory: object = Lane  # (server, nside, lmax, capacity) -> lane
    device_events: bool = True   # record an input-ready event at submission
    known: dict = field(default_factory=dict)  # (nside, lmax) -> bytes of a lane of that shape, held before

    def clock(self):
        return time.perf_counter() - self._t0

    def cap_for(self, nside, lmax):
        if self.capacity is not None:
            return int(self.capacity(nside, lmax))
        cap = 64  # halved while a lane would pass a quarter of the budget, down to 16
        while cap > 16 and estimate_lane_bytes(nside, lmax, cap) > self.hbm_budget / 4:
            cap //= 2  # (below 16 maps the padded workspace no longer shrinks)
        return cap

    def reserve(self, key, mixed, mpol):
        """Bytes a new lane of shape key = (nside, lmax) is admitted against:
        what a lane of that shape held before (its measured allocation and
        the library's planned mixed-precision bytes), else the upper
        estimate."""
        if key not in self.known:
            return estimate_lane_bytes(*key, self.cap_for(*key), mixed, mpol)
        base, mix, pol = self.known[key]
        return base + (mix if mixed or mpol else 0) + (pol if mpol else 0)

    def lane_bytes(self):
        return sum(l.footprint for l in self.lanes.values())

    def submit(self, req: Request):
        """Queue a request; its inputs must be device tensors of its shape."""
        req.rid = next(self._ids)
        req.n = 0
        _ = req.nsim  # fixed from the inputs now: the client may drop them once done
        req.arrival = self.clock()
        if self.device_events:
            req.ready = _torch().cuda.Event()
            req.ready.record()
        mixed_ok = req.precision == "mixed" and self.mixed_min_nside <= req.nside <= self.mixed_max_nside
        req.precision_used = "mixed" if mixed_ok else "fp64"
        self.waiting.append(req)
        return req

    def _admit(self):
        while self.waiting:
            r = self.waiting[0]
            key = (r.nside, r.lmax)
            lane = self.lanes.get(key)
            need = r.nbytes()
            mixed = r.precision_used == "mixed"
            mpol = mixed and r.pol
            if lane is None:
                need += self.reserve(key, mixed, mpol)
            elif (mixed and not lane.mixed_ready) or (mpol and not lane.mixed_pol_ready):
                need += (lane.mixed_bytes(mixed or lane.mixed_ready, mpol or lane.mixed_pol_ready)
                         - lane.mixed_bytes(lane.mixed_ready, lane.mixed_pol_ready))
            if self.lane_bytes() + self.held + need > self.hbm_budget and not self._evict(need, keep=key):
                if not any(l.pending or l.inflight is not None for l in self.lanes.values()) and self.held == 0:
                    raise RuntimeError(f"request {r.rid} needs {need / GB:.1f} GB, more than the "
                                       f"{self.hbm_budget / GB:.1f} GB budget holds")
                return  # wait for memory to free up (arrival order is kept)
            self.waiting.popleft()
            if lane is None:
                lane = self.lane_factory(self, r.nside, r.lmax, self.cap_for(*key))
                self.lanes[key] = lane
                mix = lane.mixed_bytes(True, False)
                self.known[key] = (lane.footprint, mix, lane.mixed_bytes(True, True) - mix)
            if r.precision_used == "mixed":
                lane.enable_mixed(pol=r.pol)
            self.held += r.nbytes()
            lane.admit(r)

    def _evict(self, need, keep=None):
        """Close idle lanes other than `keep`, least recently used first, until
        `need` fits."""
        idle = sorted((l for k, l in self.lanes.items() if k != keep and l.inflight is None and not l.pending),
                      key=lambda l: l.last_used)
        for l in idle:
            if self.lane_bytes() + self.held + need <= self.hbm_budget:
                break
            del self.lanes[(l.nside, l.lmax)]
            l.close()
            self.evictions += 1
        return self.lane_bytes() + self.held + need <= self.hbm_budget

    def poll(self):
        """One scheduling round: admit, retire finished batches, launch the
        next batch of every idle lane.  Returns the requests completed."""
        done = []
        self._admit()
        for lane in list(self.lanes.values()):
            if lane.inflight is not None and lane.inflight[0].query():
                _, chunks = lane.inflight
                lane.inflight = None
                now = self.clock()
                for r, _, k, _, _ in chunks:
                    r.maps_done += k
                    if r.maps_done == r.nsim:
                        r.done_time = now
                        self.held -= r.nbytes()
                        done.append(r)
            if lane.inflight is None and lane.pending:
                lane.launch(self.niter)
                self.batches += 1
        return done

    def busy(self):
        return bool(self.waiting) or any(l.pending or l.inflight is not None for l in self.lanes.values())

    def run_until_idle(self, sleep=0.0002):
        out = []
        while self.busy():
            got = self.poll()
            out += got
            if not got:
                time.sleep(sleep)
        return out

    def close(self):
        for l in self.lanes.values():
            l.close()
        self.lanes.clear()
