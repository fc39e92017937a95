"""Host logic of the serving loop, without a GPU: a stand-in lane records the
batches the scheduler forms and finishes each one after one poll."""
from collections import deque

import pytest

from fastsht_amd.serve import Lane, Request, Server, estimate_lane_bytes


class Done:
    def query(self):
        return True


class FakeLane(Lane):
    made = []

    def __init__(self, server, nside, lmax, cap):
        self.nside, self.lmax, self.cap = nside, lmax, cap
        self.footprint = estimate_lane_bytes(nside, lmax, cap)
        self.mixed_ready = False
        self.mixed_pol_ready = False
        self.pending = deque()
        self.inflight = None
        self.last_used = 0.0
        self.batches = []
        self.closed = False
        FakeLane.made.append(self)

    def mixed_bytes(self, mixed, pol):
        base = estimate_lane_bytes(self.nside, self.lmax, self.cap)
        return estimate_lane_bytes(self.nside, self.lmax, self.cap, mixed, pol) - base

    def _switch_mixed(self):
        pass

    def admit(self, r):
        r.outputs = ()
        self.pending.append(r)

    def launch(self, niter):
        key, chunks, n, npol = self.take_batch()
        self.batches.append((key, [(c[0].rid, c[1], c[2], c[3], c[4]) for c in chunks], n, npol))
        self.inflight = (Done(), chunks)

    def close(self):
        self.closed = True


class FakeMaps:
    def __init__(self, n):
        self.shape = (1, n)


def req(nside, n, direction="forward", pol=False, precision="fp64"):
    ins = (FakeMaps(n),) * (3 if pol else 1)
    if direction == "backward":
        ins = tuple(type("A", (), {"shape": (n, 1, 1)})() for _ in ins)
    return Request(nside=nside, lmax=2 * nside, direction=direction, inputs=ins, pol=pol, precision=precision)


def server(**kw):
    FakeLane.made = []
    return Server(lane_factory=FakeLane, device_events=False, **kw)


def test_packing_order_split_and_pair_offsets():
    srv = server(hbm_budget=1e12, capacity=lambda ns, lm: 8)
    a = srv.submit(req(16, 5))
    b = srv.submit(req(16, 6, pol=True))
    c = srv.submit(req(16, 2, direction="backward"))
    d = srv.submit(req(16, 4))
    done = srv.run_until_idle()
    assert {r.rid for r in done} == {a.rid, b.rid, c.rid, d.rid}
    (lane,) = FakeLane.made
    # first batch: a (5 maps) and the first 3 maps of b; pairs only for b
    assert lane.batches[0] == (("forward", "fp64"), [(a.rid, 0, 5, 0, 0), (b.rid, 0, 3, 5, 0)], 8, 3)
    # then the rest of b with d (same direction), c waits for its own batch
    assert lane.batches[1] == (("forward", "fp64"), [(b.rid, 3, 3, 0, 0), (d.rid, 0, 4, 3, 3)], 7, 3)
    assert lane.batches[2] == (("backward", "fp64"), [(c.rid, 0, 2, 0, 0)], 2, 0)
    assert all(r.maps_done == r.nsim for r in done)


def test_precisions_batch_separately_and_mixed_cap():
    srv = server(hbm_budget=1e12, capacity=lambda ns, lm: 64, mixed_min_nside=0, mixed_max_nside=32)
    a = srv.submit(req(16, 3, precision="mixed"))
    b = srv.submit(req(16, 3))
    c = srv.submit(req(64, 3, precision="mixed"))  # above mixed_max_nside: served in FP64
    srv.run_until_idle()
    assert a.precision_used == "mixed" and b.precision_used == "fp64" and c.precision_used == "fp64"
    lane16 = [l for l in FakeLane.made if l.nside == 16][0]
    assert [k for k, _, _, _ in lane16.batches] == [("forward", "mixed"), ("forward", "fp64")]
    assert lane16.mixed_ready
    assert not [l for l in FakeLane.made if l.nside == 64][0].mixed_ready


def test_budget_evicts_idle_lanes_and_never_the_target():
    # room for the nside-32 lane, not for both lanes: every change of shape
    # closes the other (idle) lane first
    budget = 1.1 * estimate_lane_bytes(32, 64, 4)
    srv = server(hbm_budget=budget, capacity=lambda ns, lm: 4)
    reqs = [srv.submit(req(ns, 3)) for ns in (16, 32, 16, 32, 32)]
    done = srv.run_until_idle()
    assert len(done) == len(reqs) and all(r.done for r in reqs)
    assert srv.evictions >= 2
    assert all(l.closed for l in FakeLane.made[:-1])
    assert srv.held == pytest.approx(0.0)


def test_request_above_budget_raises():
    srv = server(hbm_budget=1.0)
    srv.submit(req(16, 3))
    with pytest.raises(RuntimeError, match="budget"):
        srv.run_until_idle()


def test_spin2_buffers_charged_on_first_polarised_mixed_request():
    """A lane is charged the spin-2 chunk buffers only when its first
    polarised mixed request is admitted, once."""
    srv = server(hbm_budget=1e13, capacity=lambda ns, lm: 16, mixed_min_nside=0)
    t = srv.submit(req(32, 3, precision="mixed"))
    srv.run_until_idle()
    lane = [l for l in FakeLane.made if l.nside == 32][-1]
    assert lane.mixed_ready and not lane.mixed_pol_ready and t.done
    q = srv.submit(req(32, 2, pol=True, precision="mixed"))
    srv.run_until_idle()
    assert lane.mixed_pol_ready and q.done
    base = estimate_lane_bytes(32, 64, 16, True)
    full = estimate_lane_bytes(32, 64, 16, True, True)
    assert full > base > estimate_lane_bytes(32, 64, 16)
    assert lane.footprint == pytest.approx(full)
    srv.submit(req(32, 2, pol=True, precision="mixed"))
    srv.run_until_idle()
    assert lane.footprint == pytest.approx(full)  # charged once


def test_recreated_lane_reserves_what_it_held():
    """A lane closed for memory is re-admitted against the bytes it held
    (here half its estimate), not against the estimate."""
    class HalfLane(FakeLane):
        def __init__(self, server, nside, lmax, cap):
            super().__init__(server, nside, lmax, cap)
            self.footprint = 0.5 * estimate_lane_bytes(nside, lmax, cap)

    est16, est32 = estimate_lane_bytes(16, 32, 4), estimate_lane_bytes(32, 64, 4)
    srv = Server(lane_factory=HalfLane, device_events=False, capacity=lambda ns, lm: 4, hbm_budget=1e13)
    srv.submit(req(16, 3))
    srv.submit(req(32, 3))
    srv.run_until_idle()
    assert srv.reserve((32, 64), False, False) == pytest.approx(0.5 * est32)
    assert srv.reserve((64, 128), False, False) == pytest.approx(estimate_lane_bytes(64, 128, 4))
    # room for either lane as held, not for both, and less than the nside-32 estimate
    srv.hbm_budget = 0.5 * est32 + 0.5 * est16 - 1.0
    assert srv.hbm_budget < est32
    srv.submit(req(32, 2))
    srv.run_until_idle()
    assert srv.evictions == 1 and list(srv.lanes) == [(32, 64)]
    srv.submit(req(16, 2))
    srv.submit(req(32, 2))
    done = srv.run_until_idle()
    assert len(done) == 2 and srv.evictions == 3


def test_small_mixed_requests_run_in_fp64():
    """Below mixed_min_nside (default 256) mixed requests run in FP64, and
    their lane never builds mixed-precision state."""
    srv = server(hbm_budget=1e13, capacity=lambda ns, lm: 16)
    a = srv.submit(req(128, 3, precision="mixed"))
    b = srv.submit(req(256, 3, precision="mixed"))
    srv.run_until_idle()
    assert a.precision_used == "fp64" and b.precision_used == "mixed"
    lane128 = [l for l in FakeLane.made if l.nside == 128][0]
    assert not lane128.mixed_ready and lane128.footprint == pytest.approx(estimate_lane_bytes(128, 256, 16))
