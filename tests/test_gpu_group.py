"""Grouped launches across jobs, their HIP-graph replay and an occupancy
tuner: every job's result equals its own handle's result.

Skipped: the grouped launches and the tuner were taken out of the library
after two GPU memory faults.  Both runs faulted in spin-2 transforms through
the per-handle FP64 Legendre kernels after those kernels' bodies had been
moved into device functions shared with the grouped kernels (spin-0 cases
passed); the library is back to the kernels as they were before."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu
TOL = 1e-14

JOBS = [(16, 32, 5, False), (32, 95, 17, True), (16, 47, 33, True), (64, 128, 3, False), (32, 64, 40, False),
        (64, 191, 9, True)]


@pytest.fixture(scope="module")
def torch_cuda():
    torch = pytest.importorskip("torch")
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch


def rel(a, b):
    return float((a - b).abs().max() / b.abs().max())


def maps(torch, rng, ns, n):
    return torch.from_numpy(rng.standard_normal((n, 12 * ns * ns))).cuda().t()


def alms(torch, rng, lmax, n):
    L = lmax + 1
    return torch.from_numpy(rng.standard_normal((L, L, n))).cuda().permute(2, 1, 0)


def empty_like_alm(torch, lmax, n):
    L = lmax + 1
    return torch.empty((L, L, n), dtype=torch.float64, device="cuda").permute(2, 1, 0)


@pytest.mark.skip(reason="grouped launches were taken out: the first GPU run of this test ended in a device "
                         "memory fault (in the spin-2 transforms of the refactored per-handle kernels it ran first)")
def test_group_matches_solo_and_graph_replays(torch_cuda):
    torch = torch_cuda
    from fastsht_amd import SHT
    from fastsht_amd.group import Group
    rng = np.random.default_rng(3)
    niter = 3
    shts = [SHT(ns, lmax, n, niter=niter, pol=pol) for ns, lmax, n, pol in JOBS]
    T = [maps(torch, rng, ns, n) for ns, _, n, _ in JOBS]
    Q = [maps(torch, rng, ns, n) if pol else None for ns, _, n, pol in JOBS]
    U = [maps(torch, rng, ns, n) if pol else None for ns, _, n, pol in JOBS]
    solo_t = [s.t2alm(t) for s, t in zip(shts, T)]
    solo_eb = [s.qu2eb(q, u) if q is not None else None for s, q, u in zip(shts, Q, U)]
    torch.cuda.synchronize()
    G = Group(shts)
    out = [empty_like_alm(torch, lmax, n) for _, lmax, n, _ in JOBS]
    E = [empty_like_alm(torch, lmax, n) if pol else None for _, lmax, n, pol in JOBS]
    B = [empty_like_alm(torch, lmax, n) if pol else None for _, lmax, n, pol in JOBS]
    for graph in (False, True, True, True):  # eager, then capture + three replays
        print(f"case group t2alm/qu2eb graph={graph}", flush=True)
        for o in out:
            o.zero_()
        torch.cuda.synchronize()
        G.run("t2alm", T, out, niter=niter, graph=graph)
        G.run("qu2eb", Q, E, in1=U, out1=B, niter=niter, graph=graph)
        G.synchronize()
        for j in range(len(JOBS)):
            assert rel(out[j], solo_t[j]) < TOL, (graph, j)
            if Q[j] is not None:
                assert rel(E[j], solo_eb[j][0]) < TOL and rel(B[j], solo_eb[j][1]) < TOL, (graph, j)
        _ = torch.randn(1000, device="cuda").sum().item()  # eager work between replays
    # backward directions
    A = [alms(torch, rng, lmax, n) for _, lmax, n, _ in JOBS]
    Ea = [alms(torch, rng, lmax, n) if pol else None for _, lmax, n, pol in JOBS]
    Ba = [alms(torch, rng, lmax, n) if pol else None for _, lmax, n, pol in JOBS]
    solo_m = [s.alm2t(a) for s, a in zip(shts, A)]
    solo_qu = [s.eb2qu(e, b) if e is not None else None for s, e, b in zip(shts, Ea, Ba)]
    torch.cuda.synchronize()
    M = [torch.empty((n, 12 * ns * ns), dtype=torch.float64, device="cuda").t() for ns, _, n, _ in JOBS]
    Qo = [torch.empty_like(m) if pol else None for m, (_, _, _, pol) in zip(M, JOBS)]
    Uo = [torch.empty_like(m) if pol else None for m, (_, _, _, pol) in zip(M, JOBS)]
    for graph in (False, True, True):
        print(f"case group alm2t/eb2qu graph={graph}", flush=True)
        G.run("alm2t", A, M, graph=graph)
        G.run("eb2qu", Ea, Qo, in1=Ba, out1=Uo, graph=graph)
        G.synchronize()
        for j in range(len(JOBS)):
            assert rel(M[j], solo_m[j]) < TOL, (graph, j)
            if Ea[j] is not None:
                assert rel(Qo[j], solo_qu[j][0]) < TOL and rel(Uo[j], solo_qu[j][1]) < TOL, (graph, j)


@pytest.mark.skip(reason="the occupancy tuner was taken out with the grouped launches (same kernel refactor)")
def test_handle_tuner_keeps_results(torch_cuda):
    torch = torch_cuda
    import ctypes

    from fastsht_amd import SHT, _lib
    rng = np.random.default_rng(5)
    lib = _lib.load()
    ns, lmax, n = 32, 95, 20
    print("case tune handle nside=32 lmax=95", flush=True)
    S = SHT(ns, lmax, n, niter=2, pol=True)
    x = maps(torch, rng, ns, n)
    y = maps(torch, rng, ns, n)
    ref = S.t2alm(x)
    E, B = S.qu2eb(x, y)
    a = S.alm2t(ref)
    torch.cuda.synchronize()
    ch = (ctypes.c_int * 4)()
    _lib.check(lib.fsht_tune(S._h, ch), S._h)
    assert all(c in (1, 2) for c in ch)
    assert rel(S.t2alm(x), ref) < TOL
    E2, B2 = S.qu2eb(x, y)
    assert rel(E2, E) < TOL and rel(B2, B) < TOL
    assert rel(S.alm2t(ref), a) < TOL
    # a later handle of the same class starts with the tuned variants
    S2 = SHT(ns, lmax, n, niter=2, pol=True)
    assert rel(S2.t2alm(x), ref) < TOL


@pytest.mark.skip(reason="grouped launches were taken out (see test_group_matches_solo_and_graph_replays)")
def test_group_tuner_keeps_results(torch_cuda):
    torch = torch_cuda
    from fastsht_amd import SHT
    from fastsht_amd.group import Group
    rng = np.random.default_rng(6)
    print("case tune group", flush=True)
    shts = [SHT(a, b, c, niter=2, pol=d) for a, b, c, d in JOBS[:3]]
    T = [maps(torch, rng, a, c) for a, _, c, _ in JOBS[:3]]
    solo = [s.t2alm(t) for s, t in zip(shts, T)]
    torch.cuda.synchronize()
    G = Group(shts)
    out = [empty_like_alm(torch, b, c) for _, b, c, _ in JOBS[:3]]
    assert all(c in (1, 2) for c in G.tune("t2alm", T, out, niter=2))
    G.run("t2alm", T, out, niter=2, graph=True)
    G.synchronize()
    for o, r in zip(out, solo):
        assert rel(o, r) < TOL
