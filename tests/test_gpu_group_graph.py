"""Uneven groups of jobs (fastsht_amd/group.py): every job's result, run in
the group eagerly, from the recorded HIP graph (several replays, with other
work between them) and after the schedule tuner, equals the same job run by
its own `SHT` object.

Tolerance 1e-14 of the largest output: a job runs on a handle created for the
largest job of its shape, whose column blocks can be of the other kind than
those of a handle of the job's own size (the FP64 analysis then sums the
rings in another order)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu
TOL = 1e-14

# (nside, lmax, nsim, pol, direction, precision); two jobs share each of the
# shapes (16, 32), (32, 95), (32, 64) and (16, 47), the last two with a
# polarised mixed-precision job (spin-2 tables generated inside the graph)
JOBS = [(16, 32, 5, False, "forward", "fp64"), (32, 95, 17, True, "forward", "fp64"),
        (16, 47, 33, True, "backward", "fp64"), (32, 64, 40, False, "forward", "mixed"),
        (16, 32, 12, True, "forward", "fp64"), (64, 128, 3, False, "backward", "mixed"),
        (32, 95, 9, False, "backward", "fp64"), (32, 64, 7, True, "forward", "mixed"),
        (16, 47, 6, True, "backward", "mixed")]
NITER = 2


@pytest.fixture(scope="module")
def torch_cuda():
    torch = pytest.importorskip("torch")
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch


def rel(a, b):
    return float((a - b).abs().max() / b.abs().max())


def make_jobs(torch, rng):
    from fastsht_amd.group import Job
    jobs = []
    for ns, lmax, n, pol, direction, prec in JOBS:
        L, npix, c = lmax + 1, 12 * ns * ns, 3 if pol else 1
        maps = lambda: torch.from_numpy(rng.standard_normal((n, npix))).cuda().t()
        alms = lambda: torch.from_numpy(rng.standard_normal((L, L, n))).cuda().permute(2, 1, 0)
        emaps = lambda: torch.empty((n, npix), dtype=torch.float64, device="cuda").t()
        ealms = lambda: torch.empty((L, L, n), dtype=torch.float64, device="cuda").permute(2, 1, 0)
        if direction == "forward":
            ins, outs = tuple(maps() for _ in range(c)), tuple(ealms() for _ in range(c))
        else:
            ins, outs = tuple(alms() for _ in range(c)), tuple(emaps() for _ in range(c))
        jobs.append(Job(ns, lmax, direction, ins, outs, prec))
    return jobs


def solo(torch, j):
    from fastsht_amd import SHT
    S = SHT(j.nside, j.lmax, j.nsim, niter=NITER, pol=j.pol, precision=j.precision)
    if j.direction == "forward":
        out = [S.t2alm(j.inputs[0])]
        if j.pol:
            out += list(S.qu2eb(j.inputs[1], j.inputs[2]))
    else:
        out = [S.alm2t(j.inputs[0])]
        if j.pol:
            out += list(S.eb2qu(j.inputs[1], j.inputs[2]))
    torch.cuda.synchronize()
    return out


def check(jobs, ref, what):
    for i, (j, r) in enumerate(zip(jobs, ref)):
        for got, want in zip(j.outputs, r):
            e = rel(got, want)
            assert e < TOL, (what, i, JOBS[i], e)


def clear(torch, jobs):
    for j in jobs:
        for y in j.outputs:
            y.fill_(float("nan"))
    torch.cuda.synchronize()


def test_group_matches_solo_eager_graph_and_tuned(torch_cuda):
    torch = torch_cuda
    from fastsht_amd.group import Group
    rng = np.random.default_rng(11)
    jobs = make_jobs(torch, rng)
    ref = [solo(torch, j) for j in jobs]
    G = Group(jobs, niter=NITER, streams=3)
    print(f"case group eager: {len(jobs)} jobs, {len(G.shapes)} shapes, plan {G.plan}", flush=True)
    clear(torch, jobs)
    G.run()
    G.synchronize()
    check(jobs, ref, "eager")
    for rep in range(3):
        print(f"case group graph replay {rep}", flush=True)
        clear(torch, jobs)
        G.run(graph=True)
        G.synchronize()
        check(jobs, ref, f"graph {rep}")
        _ = torch.randn(1000, device="cuda").sum().item()  # other work between replays
    print("case group tune", flush=True)
    tried = G.tune(reps=2)
    assert tried and all(ms > 0 for _, _, ms in tried)
    clear(torch, jobs)
    G.run(graph=True)
    G.synchronize()
    check(jobs, ref, "tuned graph")
    G.close()


def test_group_rejects_bad_tensors(torch_cuda):
    torch = torch_cuda
    from fastsht_amd.group import Group, Job
    ns, lmax, n = 16, 32, 4
    x = torch.zeros((12 * ns * ns, n), dtype=torch.float64, device="cuda")  # C order: wrong strides
    y = torch.empty((lmax + 1, lmax + 1, n), dtype=torch.float64, device="cuda").permute(2, 1, 0)
    G = Group([Job(ns, lmax, "forward", (x,), (y,))])
    with pytest.raises(ValueError, match="Fortran order"):
        G.run()
    G.close()
