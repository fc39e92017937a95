"""Host logic of uneven groups, without a GPU: the schedule planner and the
memory waves of the uneven-batch benchmark."""
import bench
from fastsht_amd.group import Group, Job, estimate_handle_bytes, job_bytes


class Maps:
    def __init__(self, n):
        self.shape = (1, n)


def planner(specs, niter=3):
    G = object.__new__(Group)  # the planner needs no handles
    G.jobs = [Job(ns, lm, "forward", (Maps(n),) * (3 if pol else 1), ()) for ns, lm, n, pol in specs]
    G.niter = niter
    G.shapes = {}
    for i, j in enumerate(G.jobs):
        G.shapes.setdefault((j.nside, j.lmax), []).append(i)
    return G


def test_plan_longest_first_on_least_loaded_stream():
    G = planner([(16, 32, 10, False), (64, 128, 10, False), (32, 64, 10, True), (16, 32, 30, False)])
    cost = {k: sum(G.jobs[i].work(G.niter) for i in v) for k, v in G.shapes.items()}
    assert cost[(64, 128)] > cost[(32, 64)] > cost[(16, 32)]
    assert G._plan(2) == [[(64, 128)], [(32, 64), (16, 32)]]
    assert G._plan(1, "given") == [[(16, 32), (64, 128), (32, 64)]]
    # never more streams than shapes; every shape on exactly one stream
    p = G._plan(8)
    assert len(p) == 3 and sorted(k for s in p for k in s) == sorted(G.shapes)


def test_job_work_counts_spin2_pairs_and_passes():
    t = Job(16, 32, "forward", (Maps(4),), ())
    tqu = Job(16, 32, "forward", (Maps(4),) * 3, ())
    back = Job(16, 32, "backward", (type("A", (), {"shape": (4, 1, 1)})(),), ())
    assert tqu.work(3) == 5 * t.work(3)
    assert t.work(3) == 7 * back.work(3)  # 4 analyses + 3 residual syntheses against one synthesis


def test_waves_fit_the_budget_and_keep_job_order():
    specs = bench.c3_jobs(0)
    assert len(specs) == 24 and all(ns in (128, 256, 512) and lm in (2 * ns, 3 * ns - 1) and 16 <= n <= 400
                                    for ns, lm, n, _ in specs)
    budget = 200e9
    waves, need = bench.group_waves(specs, 3, budget)
    assert [j for w in waves for j in w] == specs
    assert all(b <= budget for b in need)
    # a job alone always forms a wave, even above the budget
    waves, _ = bench.group_waves(specs[:3], 3, 1.0)
    assert [len(w) for w in waves] == [1, 1, 1]
    assert job_bytes(16, 32, 2, True) == 8 * 3 * 2 * (12 * 256 + 33 * 33)
    assert estimate_handle_bytes(512, 1535, 341, False, 3) > estimate_handle_bytes(512, 1535, 341, False, 0)
