"""The serving loop (fastsht_amd/serve.py) and batches below a handle's
capacity (fsht_set_batch).

Mix invariance: every request's result, computed in a mix of shapes,
directions, polarisations and precisions packed into shared batches on
concurrent lanes, equals the same request transformed alone by its own `SHT`
object.  Tolerance: 1e-14 of the largest output (FP64 column blocks of
different kinds sum the contraction in a different order; mixed precision
accumulates exactly in integers and is compared at the same bound)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu
TOL = 1e-14


@pytest.fixture(scope="module")
def torch_cuda():
    torch = pytest.importorskip("torch")
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch


def rel(a, b):
    return float((a - b).abs().max() / b.abs().max())


def make_request(torch, rng, nside, direction, pol, nsim, precision, lmax=None):
    from fastsht_amd.serve import Request
    lmax = 2 * nside if lmax is None else lmax
    L, npix = lmax + 1, 12 * nside * nside
    ncomp = 3 if pol else 1
    if direction == "forward":
        ins = tuple(torch.from_numpy(rng.standard_normal((nsim, npix))).cuda().t() for _ in range(ncomp))
    else:
        ins = tuple(torch.from_numpy(rng.standard_normal((L, L, nsim))).cuda().permute(2, 1, 0)
                    for _ in range(ncomp))
    return Request(nside=nside, lmax=lmax, direction=direction, inputs=ins, pol=pol, precision=precision)


def alone(torch, r, niter):
    """The request through its own SHT object (no batching, no lanes)."""
    from fastsht_amd import SHT
    S = SHT(r.nside, r.lmax, r.nsim, niter=niter, pol=r.pol, precision=r.precision_used)
    if r.direction == "forward":
        out = [S.t2alm(r.inputs[0])]
        if r.pol:
            out += list(S.qu2eb(r.inputs[1], r.inputs[2]))
    else:
        out = [S.alm2t(r.inputs[0])]
        if r.pol:
            out += list(S.eb2qu(r.inputs[1], r.inputs[2]))
    torch.cuda.synchronize()
    return out


def check(torch, reqs, niter):
    for r in reqs:
        assert r.done
        ref = alone(torch, r, niter)
        for got, want in zip(r.outputs, ref):
            assert rel(got, want) < TOL, (r.rid, r.nside, r.direction, r.pol, r.precision_used, rel(got, want))


def test_set_batch_matches_own_handle(torch_cuda):
    """A handle created for 40 maps running 5 (then 23) maps gives what
    handles created for 5 (and 23) maps give."""
    torch = torch_cuda
    from fastsht_amd import SHT
    ns, lmax = 16, 40
    rng = np.random.default_rng(2)
    big = SHT(ns, lmax, 40, niter=2, pol=True)
    for n in (5, 23):
        print(f"case set_batch nside={ns} n={n}", flush=True)
        maps = torch.from_numpy(rng.standard_normal((n, 12 * ns * ns))).cuda().t()
        Q = torch.from_numpy(rng.standard_normal((n, 12 * ns * ns))).cuda().t()
        big.set_batch(n)
        own = SHT(ns, lmax, n, niter=2, pol=True)
        assert rel(big.t2alm(maps), own.t2alm(maps)) < TOL
        a = big.t2alm(maps)
        assert rel(big.alm2t(a), own.alm2t(a)) < TOL
        E, B = big.qu2eb(maps, Q)
        Eo, Bo = own.qu2eb(maps, Q)
        assert rel(E, Eo) < TOL and rel(B, Bo) < TOL
    with pytest.raises(Exception, match="batch must hold"):
        big.set_batch(41)


def test_mix_invariance(torch_cuda):
    torch = torch_cuda
    from fastsht_amd.serve import Server
    rng = np.random.default_rng(7)
    srv = Server(hbm_budget=60e9, niter=3, mixed_min_nside=0)  # mixed lanes at every nside here
    shapes = [(16, "forward", False, 5, "fp64"), (32, "backward", True, 9, "fp64"), (16, "forward", True, 3, "mixed"),
              (64, "forward", False, 17, "mixed"), (32, "forward", False, 2, "fp64"), (64, "backward", False, 7, "mixed"),
              (16, "backward", True, 11, "fp64"), (32, "forward", True, 6, "mixed"), (64, "forward", True, 4, "fp64"),
              (16, "forward", False, 1, "mixed"), (32, "backward", False, 12, "mixed"), (64, "backward", True, 3, "fp64")]
    reqs = [srv.submit(make_request(torch, rng, *s)) for s in shapes]
    print(f"case serve mix: {len(reqs)} requests", flush=True)
    done = srv.run_until_idle()
    assert len(done) == len(reqs)
    assert srv.batches < len(reqs)  # requests were packed together
    check(torch, reqs, 3)
    srv.close()


def test_split_requests_and_eviction(torch_cuda):
    """Capacity 8 splits larger requests over batches; a budget that holds
    one lane at a time closes idle lanes for new shapes.  Two shapes of one
    nside (lmax 128 and 95) make lanes of similar size, each larger than any
    request; a closed lane is re-admitted against the bytes it held."""
    torch = torch_cuda
    from fastsht_amd.serve import Server
    rng = np.random.default_rng(9)
    srv = Server(hbm_budget=100e9, niter=2, capacity=lambda nside, lmax: 8, mixed_min_nside=0)
    first = srv.submit(make_request(torch, rng, 64, "forward", True, 10, "mixed"))
    warm = srv.submit(make_request(torch, rng, 64, "forward", True, 2, "mixed", lmax=95))
    srv.run_until_idle()
    shapes = [(64, "forward", False, 19, "fp64", 95), (64, "forward", True, 10, "mixed"),
              (64, "backward", True, 13, "fp64", 95), (64, "backward", False, 9, "fp64")]
    reqs = [make_request(torch, rng, *s) for s in shapes]
    # room for either lane as held and the largest request, not for both lanes
    big = max(r.nbytes() for r in reqs)
    srv.hbm_budget = max(l.footprint for l in srv.lanes.values()) + big + 1e6
    assert srv.lane_bytes() > srv.hbm_budget, (srv.lane_bytes(), srv.hbm_budget)
    for r in reqs:
        srv.submit(r)
    print(f"case serve split/evict: budget {srv.hbm_budget / 1e9:.3f} GB", flush=True)
    srv.run_until_idle()
    assert srv.evictions >= 2, srv.evictions
    check(torch, [first, warm] + reqs, 2)
    srv.close()
