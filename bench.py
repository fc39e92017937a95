"""Benchmark: batched temperature SHT round trip on one MI355X.

One step = t2alm(niter=3) + alm2t over nsim maps at nside=1024, lmax=2047,
inputs resident in HBM, on synthetic maps built as alm2t of Gaussian alm with
C_l = 1/(l(l+1)) (l >= 2) plus white noise of variance 1e-3.

The value is the mixed-precision rate (int8-slice MFMA Legendre contractions,
FP64 FFTs, FP64 Jacobi refinement).  The same handle is also timed in plain
FP64 (FP64 MFMA Legendre); both rates, their ratio, p50/p99 step latencies,
the mixed alm error against the FP64 path for every niter 0..3 and the
alm2t map error are printed on the one JSON line, with the dominant Legendre
kernel's roofline (int8 MFMA for mixed, FP64 MFMA for FP64; HIP events on the
library stream) and the CPU baseline (the CPU restatement of the algorithm:
the full nside=64 configuration at 1 thread and at all threads, median of 5,
and a bounded sample of the nside=1024 workload).

  python bench.py [--steps K] [--warmup W] [--nside 1024 --lmax 2047 --nsim 200 --niter 3]
  python bench.py --only mixed|fp64      # time one precision only (profiling runs)
  python bench.py --pol [--steps K]      # spin-2 qu2eb+eb2qu (separate line, not the value)
  python bench.py --serve [--qps 1 --requests 40 --budget 200]   # serving latency (separate line)
  python bench.py --group [--group-reps 2]   # uneven 24-job batch, grouped graph vs per-job (separate line)
"""
from __future__ import annotations

import argparse
import ctypes
import json
import math
import os
import statistics
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

# MI355X published dense peaks
PEAK_FP64_MFMA_TFLOPS = 78.6
PEAK_INT8_MFMA_TOPS = 5000.0
PEAK_HBM_GBS = 8000.0
KERNEL_CLASSES = ["ring_fft_fwd", "ring_fft_inv", "transpose", "legendre_analysis", "legendre_synthesis",
                  "alias_fold", "spectrum_gather", "operand_slicing", "table_generation"]
NCLS = len(KERNEL_CLASSES)
LEGENDRE = ("legendre_analysis", "legendre_synthesis")


def legendre_flops_per_map(nside, lmax, i0):
    """Algorithmic FP64 flops of one spin-0 Legendre pass for one map with the
    north/south fold: sum_m g_m * 2 * (lmax+1-m) * (2*nside - i0(m)), g_0 = 1,
    g_m = 2 (Re and Im share P)."""
    tot = 0
    for m in range(lmax + 1):
        g = 1 if m == 0 else 2
        tot += g * 2 * (lmax + 1 - m) * (2 * nside - int(i0[m]))
    return float(tot)


def slice_products(k):
    """int8 slice products of one mixed-precision contraction with k slices per
    operand: the digit pairs (i, j) with i + j < k."""
    return k * (k + 1) // 2


def slice_ladder(info, niter, spin=0):
    """Slices per pass of a mixed t2alm(niter) + alm2t (spin 0) or
    qu2eb(niter) + eb2qu (spin 2), from the library's table (fsht_info
    out[16..36)): pass kind k (0 the first analysis, 1 a residual synthesis, 2
    a residual analysis) whose error d = min(2, damping) later Jacobi
    iterations damp uses table[k][d] slices; analysis pass k is damped by
    niter - k iterations, residual synthesis k by niter - k - 1; the output
    synthesis and an unrefined analysis use `out`."""
    b = 16 if spin == 0 else 26
    tab = [[int(info[b + 3 * k + d]) for d in range(3)] for k in range(3)]
    out = int(info[b + 9])
    d = lambda x: min(2, x)
    ana = [out] if niter == 0 else [tab[0][d(niter)]] + [tab[2][d(niter - k - 1)] for k in range(niter)]
    syn = [tab[1][d(niter - k - 1)] for k in range(niter)]
    return {"analysis": ana, "residual_synthesis": syn, ("alm2t" if spin == 0 else "eb2qu"): out,
            "table": {"first_analysis": tab[0], "residual_synthesis": tab[1], "residual_analysis": tab[2]}}


def make_inputs(sht, seed):
    """Seeded Gaussian alm (C_l = 1/(l(l+1))) and maps = alm2t(alm) + N(0, 1e-3), on the GPU."""
    import torch

    L = sht.lmax + 1
    nsim = sht.nsim
    g = torch.Generator(device="cuda")
    g.manual_seed(seed)
    l = torch.arange(L, dtype=torch.float64, device="cuda")
    cl = torch.zeros(L, dtype=torch.float64, device="cuda")
    cl[2:] = 1.0 / (l[2:] * (l[2:] + 1.0))
    # storage order (col, row, sim): col = m for Re (row = l >= m); col c >= 1
    # holds Im of m = L - c at rows l - m.
    c = torch.arange(L, device="cuda")[:, None]
    r = torch.arange(L, device="cuda")[None, :]
    re_part = r >= c
    ell = torch.where(re_part, r, r + L - c).clamp(max=L - 1)
    half = torch.where(re_part & (c == 0), torch.ones_like(cl[0]), torch.full_like(cl[0], 0.5))
    sd = torch.sqrt(cl[ell] * half)
    store = torch.randn((L, L, nsim), dtype=torch.float64, device="cuda", generator=g) * sd[:, :, None]
    alm = store.permute(2, 1, 0)  # (nsim, L, L) Fortran order
    maps = sht.alm2t(alm)
    noise = torch.randn((nsim, sht.npix), dtype=torch.float64, device="cuda", generator=g) * math.sqrt(1e-3)
    maps.add_(noise.t())
    del store, noise
    return alm, maps


def _cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def _all_threads():
    """Threads for the all-cores CPU leg: the host's logical CPUs, capped by
    OMP_NUM_THREADS when the environment sets it (the CPU share given to this
    process)."""
    n = os.cpu_count() or 1
    env = os.environ.get("OMP_NUM_THREADS")
    if env and env.isdigit() and int(env) > 0:
        n = min(n, int(env))
    return n


def cpu_c0(threads, reps=5):
    """The oracle on the full nside=64 configuration: lmax=191, nsim=100 maps
    = alm2t of Gaussian alm (C_l = 1/(l(l+1)), PCG64 seed 0) + N(0, 1e-3)
    noise; t2alm niter=0, t2alm niter=3 and alm2t, each timed whole (wall
    clock, one warm-up call, median of `reps`), per thread count."""
    import numpy as np
    from threadpoolctl import threadpool_limits

    from oracle import oracle as O

    nside, lmax, nsim = 64, 191, 100
    S = O.OracleSHT(nside, lmax, nsim)
    alm = O.random_alm(lmax, nsim, seed=0)
    rng = np.random.default_rng(1)
    maps = np.asfortranarray(S.alm2t(alm) + rng.standard_normal((S.npix, nsim)) * math.sqrt(1e-3))
    out = {}
    for nt in threads:
        with threadpool_limits(nt):
            S.t2alm(maps, niter=0)
            S.alm2t(alm)
            res = {}
            for name, fn in (("t2alm_niter0", lambda: S.t2alm(maps, niter=0)),
                             ("t2alm_niter3", lambda: S.t2alm(maps, niter=3)),
                             ("alm2t", lambda: S.alm2t(alm))):
                ts = []
                for _ in range(reps):
                    t0 = time.perf_counter()
                    fn()
                    ts.append(time.perf_counter() - t0)
                res[name] = statistics.median(ts)
        out[str(nt)] = {
            "maps_per_s": {k: nsim / v for k, v in res.items()} | {
                "t2alm_niter3+alm2t": nsim / (res["t2alm_niter3"] + res["alm2t"])},
            "median_s": res,
        }
    return out


def cpu_sample(nsim=32, mstride=16):
    """The oracle (CPU FP64 restatement of the reference algorithm: numpy ring
    FFTs, OpenBLAS dgemm per m, C Legendre recursion) on a bounded sample of the
    nside=1024, lmax=2047 workload: `nsim` maps, the full forward and inverse
    ring FFTs, and the per-m analysis and synthesis passes for every
    `mstride`-th order m.  The per-m pass times are scaled to all m by the
    per-m work (lmax+1-m) * (4*nside-1 - 2*i0(m)); the round trip is
    fft_fwd + 4 analysis passes + 4 synthesis passes + fft_inv."""
    import numpy as np

    from oracle import oracle as O

    nside, lmax = 1024, 2047
    S = O.OracleSHT(nside, lmax, nsim)
    rng = np.random.default_rng(0)
    maps = np.asfortranarray(rng.standard_normal((12 * nside * nside, nsim)))
    ms = list(range(0, lmax + 1, mstride))
    t = S.timed_sample(maps, ms)
    w = (lmax + 1 - np.arange(lmax + 1)) * (S.nring - 2 * S.rtab)
    scale = float(w.sum() / w[ms].sum())
    total = t["fft_fwd"] + t["fft_inv"] + 4.0 * scale * (t["analysis"] + t["synthesis"])
    return nsim / total, (f"oracle (numpy FFT + OpenBLAS dgemm + C recursion) at nside=1024 lmax=2047, {nsim} maps: "
                          f"ring FFT fwd {t['fft_fwd']:.1f}s + inv {t['fft_inv']:.1f}s, analysis {t['analysis']:.1f}s "
                          f"and synthesis {t['synthesis']:.1f}s over every {mstride}th m scaled x{scale:.2f} by per-m "
                          f"work; round trip = FFTs + 4 analysis + 4 synthesis passes")


def cpu_baseline(c0=True):
    """CPU legs, on rank 0's host: the value is the nside=1024 sample (the
    benchmark's workload) at all threads; `c0` holds the full nside=64
    configuration at 1 thread and at all threads."""
    from threadpoolctl import threadpool_limits

    nt = _all_threads()
    with threadpool_limits(nt):
        v, sample = cpu_sample()
    out = {"value": v, "unit": "maps/s", "cores": nt, "kind": "port", "sample": sample,
           "cpu": _cpu_model(), "host_logical_cpus": os.cpu_count(), "threads_all": nt}
    if c0:
        out["c0"] = {"what": "oracle, full nside=64 lmax=191 nsim=100 (alm2t of seeded Gaussian alm + N(0,1e-3)), "
                             "no extrapolation; wall clock, 1 warm-up, median of 5",
                     "threads": cpu_c0([1, nt]) if nt > 1 else cpu_c0([1])}
    return out


def make_eb(sht, seed):
    """Seeded E/B Gaussian alm (C_l^EE = 1/(l(l+1)), C_l^BB = 0.05 C_l^EE, l >= 2)
    in the packed storage, on the GPU."""
    import torch

    L = sht.lmax + 1
    g = torch.Generator(device="cuda")
    g.manual_seed(seed)
    l = torch.arange(L, dtype=torch.float64, device="cuda")
    cl = torch.zeros(L, dtype=torch.float64, device="cuda")
    cl[2:] = 1.0 / (l[2:] * (l[2:] + 1.0))
    c = torch.arange(L, device="cuda")[:, None]
    r = torch.arange(L, device="cuda")[None, :]
    re_part = r >= c
    ell = torch.where(re_part, r, r + L - c).clamp(max=L - 1)
    half = torch.where(re_part & (c == 0), torch.ones_like(cl[0]), torch.full_like(cl[0], 0.5))
    out = []
    for amp in (1.0, 0.05):
        sd = torch.sqrt(cl[ell] * half * amp)
        out.append((torch.randn((L, L, sht.nsim), dtype=torch.float64, device="cuda", generator=g)
                    * sd[:, :, None]).permute(2, 1, 0))
    return out


def read_profile(sht, lib):
    from fastsht_amd import _lib

    prof = (ctypes.c_double * (2 * NCLS))()
    _lib.check(lib.fsht_profile_read(sht._h, prof, 2 * NCLS), sht._h)
    _lib.check(lib.fsht_profile(sht._h, 0), sht._h)
    return {KERNEL_CLASSES[c]: {"launches": int(prof[2 * c]), "ms": prof[2 * c + 1]} for c in range(NCLS)}


def timed_steps(sht, lib, step, steps, warmup):
    """W untimed steps, then exactly K steps bracketed by synchronisation, with
    one HIP event per step boundary on the library's stream (torch's current
    stream) and the library's per-class kernel timers on."""
    import torch

    from fastsht_amd import _lib

    sht._use_torch_stream()
    for _ in range(warmup):
        step()
    torch.cuda.synchronize()
    _lib.check(lib.fsht_profile(sht._h, 1), sht._h)
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(steps + 1)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ev[0].record()
    for i in range(steps):
        step()
        ev[i + 1].record()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ms_total = ev[0].elapsed_time(ev[-1])
    lat = sorted(ev[i].elapsed_time(ev[i + 1]) for i in range(steps))
    return {"ms_total": ms_total, "lat": lat, "kern": read_profile(sht, lib), "wall": wall}


def pct(lat, p):
    return lat[min(len(lat) - 1, max(0, int(math.ceil(p / 100.0 * len(lat))) - 1))]


def load_executed(kernel, prec, nside, nsim):
    """Executed-MFMA fraction of a Legendre class (MFMA work actually issued
    per second over the peak, from the rocprofv3 counters kept under
    profiles/ for this shape), or None."""
    tf = os.path.join(ROOT, "profiles", f"mfma_executed_nside{nside}_nsim{nsim}.json")
    if os.path.exists(tf):
        return json.load(open(tf)).get(prec, {}).get(kernel)
    return None


def load_traffic(kernel, tag, nside, nsim):
    tf = os.path.join(ROOT, "profiles", f"traffic_{kernel}{tag}_nside{nside}_nsim{nsim}.json")
    if os.path.exists(tf):
        return json.load(open(tf)).get("bytes_per_launch")
    return None


def pol_round_trip(args, lmax):
    """Spin-2 workload (nside 1024, lmax 2047, nsim Q/U pairs): one step =
    qu2eb(niter) + eb2qu on Q/U built as eb2qu of seeded E/B alm
    (C_l^BB = 0.05 C_l^EE).  The same handle runs the step in FP64 and in mixed
    precision (int8-slice F1/F2 contractions, tables generated per pass, FP64
    refinement).  Prints one JSON line: Q/U pairs per second for both, their
    ratio, the E/B error against the input alm per niter for both, and the
    mixed result's error against the FP64 result per niter (qu2eb) and for
    eb2qu (first 16 maps).  The T round trip stays the benchmark value."""
    import torch

    from fastsht_amd import SHT, _lib

    sht = SHT(args.nside, lmax, args.nsim, niter=args.niter, pol=True)
    lib = _lib.load()
    precs = [args.only] if args.only else ["fp64", "mixed"]
    E, B = make_eb(sht, args.seed)
    Q, U = sht.eb2qu(E, B)
    Eo = torch.empty_like(E.permute(2, 1, 0).contiguous()).permute(2, 1, 0)
    Bo = torch.empty_like(Eo)
    p = lambda t: ctypes.c_void_p(t.data_ptr())
    relmax = lambda a, b: float((a - b).abs().max()) / max(float(b.abs().max()), 1e-300)

    def qu2eb(prec, it, e, b):
        sht.set_precision(prec)
        _lib.check(lib.fsht_qu2eb(sht._h, p(Q), p(U), p(e), p(b), it, _lib.FSHT_DEVICE), sht._h)
        torch.cuda.synchronize()

    # per-niter errors, outside the timed region (the first mixed call builds
    # the spin-2 state: exponents, ranges of orders, chunk buffers)
    scale = max(float(E.abs().max()), float(B.abs().max()))
    err_in = {pr: {} for pr in precs}
    for pr in precs:
        for it in range(args.niter + 1):
            qu2eb(pr, it, Eo, Bo)
            err_in[pr][str(it)] = max(float((Eo - E).abs().max()), float((Bo - B).abs().max())) / scale
    del E, B
    torch.cuda.empty_cache()
    err_vs = {}
    eb2qu_err = None
    if len(precs) == 2:
        Ef = torch.empty_like(Eo)
        Bf = torch.empty_like(Eo)
        for it in range(args.niter + 1):
            qu2eb("fp64", it, Ef, Bf)
            qu2eb("mixed", it, Eo, Bo)
            err_vs[str(it)] = max(relmax(Eo, Ef), relmax(Bo, Bf))
        # eb2qu of the first 16 maps' FP64 E/B in both precisions
        n = min(16, args.nsim)
        cmp = lambda t: t[:n].permute(2, 1, 0).contiguous().permute(2, 1, 0)
        En, Bn = cmp(Ef), cmp(Bf)
        del Ef, Bf
        _lib.check(lib.fsht_set_batch(sht._h, n), sht._h)
        outs = {}
        for pr in precs:
            sht.set_precision(pr)
            q = torch.empty((n, sht.npix), dtype=torch.float64, device="cuda").t()
            u = torch.empty_like(q)
            _lib.check(lib.fsht_eb2qu(sht._h, p(En), p(Bn), p(q), p(u), _lib.FSHT_DEVICE), sht._h)
            outs[pr] = (q, u)
        torch.cuda.synchronize()
        eb2qu_err = max(relmax(outs["mixed"][0], outs["fp64"][0]), relmax(outs["mixed"][1], outs["fp64"][1]))
        _lib.check(lib.fsht_set_batch(sht._h, args.nsim), sht._h)
        del outs, En, Bn
        torch.cuda.empty_cache()

    def step():  # Q/U are overwritten by their own round trip (kept in HBM)
        _lib.check(lib.fsht_qu2eb(sht._h, p(Q), p(U), p(Eo), p(Bo), args.niter, _lib.FSHT_DEVICE), sht._h)
        _lib.check(lib.fsht_eb2qu(sht._h, p(Eo), p(Bo), p(Q), p(U), _lib.FSHT_DEVICE), sht._h)

    runs = {}
    for pr in precs:
        sht.set_precision(pr)
        runs[pr] = timed_steps(sht, lib, step, args.steps, args.warmup)
    info = (ctypes.c_longlong * 39)()
    _lib.check(lib.fsht_info(sht._h, info, 39), sht._h)
    ladder = slice_ladder(list(info), args.niter, spin=2)
    fpair = 4 * legendre_flops_per_map(args.nside, lmax, sht.ring_tab())  # F1 and F2 against Q and U
    passes = {"legendre_analysis": args.niter + 1, "legendre_synthesis": args.niter + 1}
    fp64_work = {k: fpair * args.nsim * passes[k] for k in LEGENDRE}
    int8_work = {"legendre_analysis": fpair * args.nsim * sum(slice_products(k) for k in ladder["analysis"]),
                 "legendre_synthesis": fpair * args.nsim * (sum(slice_products(k) for k in ladder["residual_synthesis"])
                                                            + slice_products(ladder["eb2qu"]))}
    rates = {pr: args.nsim * args.steps / (r["ms_total"] / 1e3) for pr, r in runs.items()}
    summ = {}
    for pr, r in runs.items():
        kern = r["kern"]
        if pr == "mixed":
            roof = legendre_roofline(kern, args.steps, int8_work, "TOP/s", PEAK_INT8_MFMA_TOPS, passes)[1]
            roof["fp64_equivalent_tflops"] = legendre_roofline(kern, args.steps, fp64_work, "TFLOP/s",
                                                               PEAK_FP64_MFMA_TFLOPS, passes)[1]["achieved"]
        else:
            roof = legendre_roofline(kern, args.steps, fp64_work, "TFLOP/s", PEAK_FP64_MFMA_TFLOPS, passes)[1]
        summ[pr] = {"qu_pairs_per_s": rates[pr], "ms_per_step": r["ms_total"] / args.steps,
                    "latency_ms_per_step": {"p50": pct(r["lat"], 50), "p99": pct(r["lat"], 99)},
                    "kernels_ms_per_step": {k: v["ms"] / args.steps for k, v in kern.items()},
                    "roofline": roof, "eb_alm_rel_err_vs_input_by_niter": err_in[pr]}
    head = "mixed" if "mixed" in runs else "fp64"
    out = {
        "metric": "Q/U pairs/s qu2eb+eb2qu (spin 2; not the benchmark value)",
        "value": rates[head], "unit": "QU pairs/s", "variant": head,
        "ms_per_step": runs[head]["ms_total"] / args.steps, "steps": args.steps, "warmup": args.warmup,
        "dtype": "i8/f64" if head == "mixed" else "f64",
        "config": {"workload": f"qu2eb(niter={args.niter})+eb2qu nside={args.nside} lmax={lmax} nsim={args.nsim} "
                               "Q/U from seeded E/B (C_l^BB = 0.05 C_l^EE)",
                   "nside": args.nside, "lmax": lmax, "nsim": args.nsim, "niter": args.niter,
                   "slices_per_pass": ladder},
    }
    out.update({pr: summ[pr] for pr in runs})
    if len(runs) == 2:
        out["mixed_vs_fp64"] = rates["mixed"] / rates["fp64"]
        out["mixed_eb_rel_err_vs_fp64_by_niter"] = err_vs
        out["mixed_eb2qu_map_rel_err_vs_fp64"] = eb2qu_err
        out["mixed_err_target"] = 1e-10
    print(json.dumps(out))


def host_round_trip(sht, maps_dev, args):
    """t2alm(niter) + alm2t on host arrays: pageable NumPy maps in (staged
    through pinned buffers, FFTs overlapping the transfer), pinned alms and
    maps out (allocated once).  Prints one JSON line with the PCIe-inclusive
    rate; this is never the benchmark's value."""
    import torch

    from fastsht_amd import empty_alms, empty_maps

    maps_h = maps_dev.cpu().numpy()            # pageable, Fortran order
    del maps_dev
    torch.cuda.empty_cache()
    alm_h = empty_alms(sht.lmax, sht.nsim)     # pinned
    out_h = empty_maps(sht.npix, sht.nsim)     # pinned

    def step():
        sht.t2alm(maps_h, alm_h)
        sht.alm2t(alm_h, out_h)

    for _ in range(args.warmup):
        step()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    wall = time.perf_counter() - t0
    gb = 2 * maps_h.nbytes / 1e9 + 2 * alm_h.nbytes / 1e9
    print(json.dumps({"metric": "PCIe-inclusive maps/s t2alm+alm2t from host arrays (not the benchmark value)",
                      "value": sht.nsim * args.steps / wall, "unit": "maps/s", "ms_per_step": 1e3 * wall / args.steps,
                      "host_bytes_moved_per_step_GB": gb, "steps": args.steps, "warmup": args.warmup,
                      "precision": sht.precision,
                      "config": {"nside": sht.nside, "lmax": sht.lmax, "nsim": sht.nsim, "niter": args.niter,
                                 "maps_in": "pageable NumPy", "alms_and_maps_out": "pinned (fsht_host_alloc)"}}))


def c3_jobs(seed, n=24, nsides=(128, 256, 512), nsim=(16, 400)):
    """The uneven batch: n jobs from a seeded generator, nside in {128, 256,
    512}, lmax in {2 nside, 3 nside - 1}, nsim uniform in 16..400, T or TQU."""
    import numpy as np

    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        ns = int(rng.choice(list(nsides)))
        lmax = int(rng.choice([2 * ns, 3 * ns - 1]))
        out.append((ns, lmax, int(rng.integers(nsim[0], nsim[1] + 1)), bool(rng.integers(0, 2))))
    return out


def group_waves(specs, niter, budget):
    """Consecutive jobs packed into waves whose inputs, outputs and handles
    (one per shape, sized for its largest job) fit the byte budget."""
    from fastsht_amd.group import estimate_handle_bytes, job_bytes

    waves, cur = [], []

    def need(js):
        caps = {}
        for ns, lmax, n, pol in js:
            c, p = caps.get((ns, lmax), (0, False))
            caps[(ns, lmax)] = (max(c, n), p or pol)
        return (sum(job_bytes(*j) for j in js)
                + sum(estimate_handle_bytes(k[0], k[1], c, p, niter) for k, (c, p) in caps.items()))

    for j in specs:
        if cur and need(cur + [j]) > budget:
            waves.append(cur)
            cur = []
        cur.append(j)
    waves.append(cur)
    return waves, [need(w) for w in waves]


def run_group_waves(specs, args, log):
    """Each memory wave of `specs` as per-job launches (one stream, jobs one
    after another, no graph: what a caller looping over SHT objects gets),
    as the group eager on 4 streams, and as the tuned group graph."""
    import torch

    from fastsht_amd.group import Group, Job

    free = torch.cuda.mem_get_info()[0]
    waves, need = group_waves(specs, args.niter, 0.8 * free)
    log(f"group: {len(specs)} jobs in {len(waves)} waves, estimated {[round(b / 1e9, 1) for b in need]} GB "
        f"of {free / 1e9:.1f} GB free")
    res = []
    for w, wj in enumerate(waves):
        g = torch.Generator(device="cuda")
        g.manual_seed(args.seed + w)
        jobs = []
        for ns, lmax, n, pol in wj:
            L, npix, c = lmax + 1, 12 * ns * ns, 3 if pol else 1
            ins = tuple(torch.randn((n, npix), dtype=torch.float64, device="cuda", generator=g).t() for _ in range(c))
            outs = tuple(torch.empty((L, L, n), dtype=torch.float64, device="cuda").permute(2, 1, 0)
                         for _ in range(c))
            jobs.append(Job(ns, lmax, "forward", ins, outs))
        t0 = time.perf_counter()
        G = Group(jobs, niter=args.niter, streams=4)
        setup = time.perf_counter() - t0
        log(f"wave {w}: {len(jobs)} jobs, {len(G.shapes)} shapes, handles set up in {setup:.1f} s")

        def timed_eager(plan, reps):
            G._bind_plan(plan)
            G._enqueue(plan)  # untimed first run
            torch.cuda.synchronize()
            ts = []
            for _ in range(reps):
                a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
                a.record()
                G._enqueue(plan)
                b.record()
                b.synchronize()
                ts.append(a.elapsed_time(b))
            return sorted(ts)[len(ts) // 2]

        per_job = timed_eager([list(G.shapes)], args.group_reps)
        sums = lambda: [(float(y.sum()), float(y.abs().sum())) for j in jobs for y in j.outputs]
        ref = sums()  # checksums of the per-job results
        log(f"wave {w}: per-job launches {per_job:.2f} ms")
        eager4 = timed_eager(G._plan(4), args.group_reps)
        log(f"wave {w}: grouped, 4 streams, eager {eager4:.2f} ms")
        tried = G.tune(reps=args.group_reps, orders=("work",), log=log)
        grouped = G.time_graph(G.graph, args.group_reps)
        same = sums() == ref
        log(f"wave {w}: grouped graph {grouped:.2f} ms, outputs equal to per-job: {same}")
        maps = sum(n * (3 if pol else 1) for _, _, n, pol in wj)
        res.append({"jobs": [list(j) for j in wj], "maps": maps, "setup_s": setup, "per_job_ms": per_job,
                    "grouped_eager_4_streams_ms": eager4, "grouped_graph_ms": grouped,
                    "tuner": [{"streams": s, "order": o, "ms": ms} for s, o, ms in tried],
                    "plan": [[list(k) for k in p] for p in G.plan], "outputs_equal": same})
        G.close()
        del jobs, G
        torch.cuda.empty_cache()
    maps = sum(r["maps"] for r in res)
    t_job = sum(r["per_job_ms"] for r in res) / 1e3
    t_grp = sum(r["grouped_graph_ms"] for r in res) / 1e3
    return {"maps_per_s": maps / t_grp, "per_job_maps_per_s": maps / t_job, "grouped_vs_per_job": t_job / t_grp,
            "waves": res}


def group_bench(args):
    """Uneven batch: the 24 seeded jobs run forward (t2alm(niter) + qu2eb for
    TQU jobs, FP64) on white-noise maps, as one Group per memory wave, against
    per-job launches.  A second set of 24 small jobs (nside 32/64/128, 1-16
    maps) shows the regime where one job does not fill the device."""
    log = lambda m: print(m, file=sys.stderr, flush=True)
    big = run_group_waves(c3_jobs(args.seed), args, log)
    small = run_group_waves(c3_jobs(args.seed, nsides=(32, 64, 128), nsim=(1, 16)), args, log)
    print(json.dumps({
        "metric": "uneven 24-job batch maps/s, grouped HIP graph vs per-job launches (not the benchmark value)",
        "value": big["maps_per_s"], "unit": "maps/s (T, Q and U maps counted)",
        "per_job_maps_per_s": big["per_job_maps_per_s"], "grouped_vs_per_job": big["grouped_vs_per_job"],
        "niter": args.niter, "dtype": "f64", "waves": big["waves"],
        "small_jobs": dict(small, config="24 seeded jobs: nside 32/64/128, lmax 2 nside or 3 nside - 1, "
                                         "nsim 1-16, T or TQU"),
        "data": "white-noise maps N(0,1), seeded", "config": {"workload": "24 seeded jobs: nside 128/256/512, "
                                                              "lmax 2 nside or 3 nside - 1, nsim 16-400, T or TQU, "
                                                              "t2alm(niter) + qu2eb"}}))


def serve_stream(args):
    """Serving benchmark: a seeded Poisson stream of requests at args.qps,
    each with nside 64/256/1024/2048 (weights 0.4/0.3/0.2/0.1), lmax =
    2 nside, nsim uniform in 1..64, T or TQU, forward or backward, FP64 or
    mixed (each half and half), served by fastsht_amd.serve.Server under an
    HBM budget.  Inputs are generated on the GPU at each request's arrival
    (white maps, or alm with C_l = 1/(l(l+1))).  Prints one JSON line with
    p50/p99 latency overall and per nside and the served throughput."""
    import numpy as np
    import torch

    from fastsht_amd.serve import GB, Request, Server

    rng = np.random.default_rng(args.seed)
    n = args.requests
    arrivals = np.cumsum(rng.exponential(1.0 / args.qps, n))
    spec = [dict(nside=int(rng.choice([64, 256, 1024, 2048], p=[0.4, 0.3, 0.2, 0.1])),
                 nsim=int(rng.integers(1, 65)), pol=bool(rng.random() < 0.5),
                 direction="forward" if rng.random() < 0.5 else "backward",
                 precision="mixed" if rng.random() < 0.5 else "fp64") for _ in range(n)]

    def inputs(i):
        sp = spec[i]
        ns, m = sp["nside"], sp["nsim"]
        L, npix = 2 * ns + 1, 12 * ns * ns
        g = torch.Generator(device="cuda")
        g.manual_seed(1000 + i)
        k = 3 if sp["pol"] else 1
        if sp["direction"] == "forward":
            return tuple(torch.randn((m, npix), dtype=torch.float64, device="cuda", generator=g).t() for _ in range(k))
        ell = torch.arange(L, dtype=torch.float64, device="cuda")
        sd = torch.zeros(L, dtype=torch.float64, device="cuda")
        sd[2:] = (1.0 / (ell[2:] * (ell[2:] + 1.0))).sqrt()
        return tuple((torch.randn((L, L, m), dtype=torch.float64, device="cuda", generator=g) * sd[None, :, None])
                     .permute(2, 1, 0) for _ in range(k))

    srv = Server(hbm_budget=args.budget * GB, niter=args.niter)
    # warm-up (untimed): one small request per shape and precision builds the
    # lanes' resident tables, as a serving process does before taking traffic
    for ns in (64, 256, 1024, 2048):
        for prec in ("fp64", "mixed"):
            x = torch.randn((1, 12 * ns * ns), dtype=torch.float64, device="cuda").t()
            srv.submit(Request(nside=ns, lmax=2 * ns, direction="forward", inputs=(x, x, x), pol=True,
                               precision=prec))
            srv.run_until_idle()
    torch.cuda.synchronize()
    srv.batches = 0
    srv._t0 = time.perf_counter()
    reqs, i, done = [], 0, []
    while i < n or srv.busy():
        now = srv.clock()
        while i < n and arrivals[i] <= now:
            sp = spec[i]
            r = srv.submit(Request(nside=sp["nside"], lmax=2 * sp["nside"], direction=sp["direction"],
                                   inputs=inputs(i), pol=sp["pol"], precision=sp["precision"]))
            r.arrival = float(arrivals[i])  # latency counts from the scheduled arrival
            reqs.append(r)
            i += 1
        got = srv.poll()
        for r in got:  # the client takes its result and lets the buffers go
            r.inputs = r.outputs = None
        done += got
        if not got:
            time.sleep(0.0005)
    makespan = srv.clock()
    srv.close()
    lat = sorted(r.latency for r in done)
    maps = sum(r.nsim * (3 if r.pol else 1) for r in done)

    def pcts(v):
        v = sorted(v)
        return {"p50": pct(v, 50) * 1e3, "p99": pct(v, 99) * 1e3, "n": len(v)}
    print(json.dumps({
        "metric": "serving latency (ms) at a fixed request rate (not the benchmark value)",
        "qps": args.qps, "requests": n, "budget_GB": args.budget, "niter": args.niter,
        "latency_ms": pcts(lat), "latency_ms_by_nside": {str(k): pcts([r.latency for r in done if r.nside == k])
                                                          for k in (64, 256, 1024, 2048)},
        "served_maps_per_s": maps / makespan, "makespan_s": makespan, "batches": srv.batches,
        "lane_closures": srv.evictions,
        "mixed_served_as_fp64": sum(1 for r in done if r.precision == "mixed" and r.precision_used == "fp64"),
        "config": {"workload": "Poisson stream: nside 64/256/1024/2048 (0.4/0.3/0.2/0.1), lmax=2 nside, nsim 1-64, "
                               "T or TQU, forward (t2alm/qu2eb niter) or backward (alm2t/eb2qu), FP64 or mixed; "
                               "maps counted per component"},
        "data": "synthetic, generated on the GPU at arrival (white maps; alm with C_l = 1/(l(l+1)))",
    }))


def errors_by_niter(sht, lib, maps, niter):
    """Mixed t2alm against FP64 t2alm on the same handle and maps, for every
    niter 0..niter (max |alm_mixed - alm_fp64| / max |alm_fp64|), and the
    alm2t map error (max |map_mixed - map_fp64| / max |map_fp64|) on the FP64
    alm.  Outside every timed region."""
    import torch

    from fastsht_amd import _lib

    p = lambda t: ctypes.c_void_p(t.data_ptr())
    L = sht.lmax + 1
    ref = torch.empty((L, L, sht.nsim), dtype=torch.float64, device="cuda").permute(2, 1, 0)
    mix = torch.empty_like(ref)
    errs = {}
    for it in range(niter + 1):
        sht.set_precision("fp64")
        _lib.check(lib.fsht_t2alm(sht._h, p(maps), p(ref), it, _lib.FSHT_DEVICE), sht._h)
        sht.set_precision("mixed")
        _lib.check(lib.fsht_t2alm(sht._h, p(maps), p(mix), it, _lib.FSHT_DEVICE), sht._h)
        torch.cuda.synchronize()
        scale = float(ref.abs().max())
        errs[str(it)] = float(mix.sub_(ref).abs_().max()) / scale
    del mix
    out = torch.empty((sht.nsim, sht.npix), dtype=torch.float64, device="cuda").t()
    out_fp = torch.empty_like(out)
    sht.set_precision("fp64")
    _lib.check(lib.fsht_alm2t(sht._h, p(ref), p(out_fp), _lib.FSHT_DEVICE), sht._h)
    sht.set_precision("mixed")
    _lib.check(lib.fsht_alm2t(sht._h, p(ref), p(out), _lib.FSHT_DEVICE), sht._h)
    torch.cuda.synchronize()
    scale = float(out_fp.abs().max())
    map_err = float(out.sub_(out_fp).abs_().max()) / scale
    del ref, out, out_fp
    torch.cuda.empty_cache()
    return errs, map_err


def legendre_roofline(kern, steps, work, bound_unit, peak, passes):
    """Roofline of the dominant Legendre class (the larger total time): the
    algorithmic work of one pass (`work[cls]` per step / `passes[cls]` per
    step) over the class's average pass duration (its time per step over its
    passes per step; a pass may span several recorded launches)."""
    dom = max(LEGENDRE, key=lambda k: kern[k]["ms"])
    avg_ms = kern[dom]["ms"] / steps / passes[dom]
    per_pass = work[dom] / passes[dom]
    achieved = per_pass / (avg_ms / 1e3) / 1e12
    return dom, {"bound": "mfma", "kernel": dom, "achieved": achieved, "peak": peak, "unit": bound_unit,
                 "frac": achieved / peak, "work_per_launch": per_pass, "avg_launch_ms": avg_ms,
                 "launches_per_step": passes[dom]}


def gpu_c0(precision):
    """The nside=64 configuration on the GPU (lmax=191, nsim=100, maps built as
    in the CPU leg): t2alm niter=0, t2alm niter=3 and alm2t, each the median of
    5 HIP-event-timed calls after 2 warm-up calls."""
    import torch

    from fastsht_amd import SHT

    sht = SHT(64, 191, 100, niter=3)
    alm, maps = make_inputs(sht, 1)
    sht.set_precision(precision)
    sht._use_torch_stream()
    alm_out = torch.empty((192, 192, 100), dtype=torch.float64, device="cuda").permute(2, 1, 0)
    maps_out = torch.empty((100, sht.npix), dtype=torch.float64, device="cuda").t()
    from fastsht_amd import _lib
    lib = _lib.load()
    p = lambda t: ctypes.c_void_p(t.data_ptr())
    res = {}
    for name, fn in (("t2alm_niter0", lambda: lib.fsht_t2alm(sht._h, p(maps), p(alm_out), 0, _lib.FSHT_DEVICE)),
                     ("t2alm_niter3", lambda: lib.fsht_t2alm(sht._h, p(maps), p(alm_out), 3, _lib.FSHT_DEVICE)),
                     ("alm2t", lambda: lib.fsht_alm2t(sht._h, p(alm), p(maps_out), _lib.FSHT_DEVICE))):
        for _ in range(2):
            _lib.check(fn(), sht._h)
        ts = []
        for _ in range(5):
            a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            a.record()
            _lib.check(fn(), sht._h)
            b.record()
            torch.cuda.synchronize()
            ts.append(a.elapsed_time(b) / 1e3)
        res[name] = statistics.median(ts)
    return {k: 100 / v for k, v in res.items()} | {"t2alm_niter3+alm2t": 100 / (res["t2alm_niter3"] + res["alm2t"])}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--nside", type=int, default=1024)
    ap.add_argument("--lmax", type=int, default=None)
    ap.add_argument("--nsim", type=int, default=200)
    ap.add_argument("--niter", type=int, default=3)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--only", choices=("mixed", "fp64"), default=None,
                    help="time one precision only and skip the error and nside=64 legs (profiling runs)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--cpu-only", action="store_true", help="only the CPU baseline leg")
    ap.add_argument("--host", action="store_true",
                    help="time the round trip from host arrays instead (pageable maps in, pinned alms/maps out): "
                         "the PCIe-inclusive rate, reported separately from the HBM-resident value")
    ap.add_argument("--pol", action="store_true",
                    help="time the spin-2 round trip qu2eb(niter)+eb2qu instead and report per-niter E/B errors "
                         "(reported separately from the temperature value)")
    ap.add_argument("--serve", action="store_true",
                    help="serve a seeded Poisson stream of mixed requests instead and report p50/p99 latency "
                         "(reported separately from the benchmark value)")
    ap.add_argument("--group", action="store_true",
                    help="uneven 24-job batch: grouped HIP graph vs per-job launches (reported separately)")
    ap.add_argument("--group-reps", type=int, default=3, help="timed runs per schedule of --group")
    ap.add_argument("--qps", type=float, default=1.0, help="request rate of --serve")
    ap.add_argument("--requests", type=int, default=40, help="requests of --serve")
    ap.add_argument("--budget", type=float, default=200.0, help="HBM budget of --serve, GB")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run a device copy of MIB MiB (HBM-counter calibration under rocprofv3)")
    args = ap.parse_args()
    if args.copy_calibration:
        import torch

        n = args.copy_calibration * (1 << 20) // 8
        src = torch.ones(n, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_read": n * 8, "copy_bytes_written": n * 8, "launches": 3}))
        return
    if args.cpu_only:
        print(json.dumps({"cpu_baseline": cpu_baseline()}))
        return
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); "
              "the transform shards over maps with no communication between GPUs", file=sys.stderr)
        sys.exit(2)
    lmax = args.lmax if args.lmax is not None else 2 * args.nside - 1

    import torch

    from fastsht_amd import SHT, _lib

    torch.cuda.set_device(0)
    if args.pol:
        pol_round_trip(args, lmax)
        return
    if args.serve:
        serve_stream(args)
        return
    if args.group:
        group_bench(args)
        return
    sht = SHT(args.nside, lmax, args.nsim, niter=args.niter)
    alm, maps = make_inputs(sht, args.seed)
    L = lmax + 1
    alm_out = torch.empty((L, L, args.nsim), dtype=torch.float64, device="cuda").permute(2, 1, 0)
    maps_out = torch.empty((args.nsim, sht.npix), dtype=torch.float64, device="cuda").t()
    lib = _lib.load()
    if args.host:
        sht.set_precision(args.only or "mixed")
        host_round_trip(sht, maps, args)
        return

    def step():
        sht.t2alm(maps, alm_out)
        sht.alm2t(alm_out, maps_out)

    sht.set_precision("mixed")  # builds the int8 tables before any timing
    info = (ctypes.c_longlong * 39)()
    _lib.check(lib.fsht_info(sht._h, info, 39), sht._h)
    ladder = slice_ladder(list(info), args.niter)
    runs = {}
    for prec in ("fp64", "mixed"):
        if args.only and prec != args.only:
            continue
        sht.set_precision(prec)
        runs[prec] = timed_steps(sht, lib, step, args.steps, args.warmup)
    rates = {k: args.nsim * args.steps / (r["ms_total"] / 1e3) for k, r in runs.items()}
    head = "mixed" if "mixed" in runs else "fp64"

    # round trip against the input alm (noise-limited), with the last timed precision's alm
    rel = float(((alm_out - alm).abs().max() / alm.abs().max()).item())
    del alm
    torch.cuda.empty_cache()

    i0 = sht.ring_tab()
    fpm = legendre_flops_per_map(args.nside, lmax, i0)
    fp64_work = {"legendre_analysis": fpm * args.nsim * (args.niter + 1),
                 "legendre_synthesis": fpm * args.nsim * (args.niter + 1)}
    int8_work = {"legendre_analysis": fpm * args.nsim * sum(slice_products(k) for k in ladder["analysis"]),
                 "legendre_synthesis": fpm * args.nsim * (sum(slice_products(k) for k in ladder["residual_synthesis"])
                                                          + slice_products(ladder["alm2t"]))}
    passes = {"legendre_analysis": args.niter + 1, "legendre_synthesis": args.niter + 1}
    roofs = {}
    for prec, r in runs.items():
        if prec == "mixed":
            dom, roof = legendre_roofline(r["kern"], args.steps, int8_work, "TOP/s", PEAK_INT8_MFMA_TOPS, passes)
            fp = legendre_roofline(r["kern"], args.steps, fp64_work, "TFLOP/s", PEAK_FP64_MFMA_TFLOPS, passes)[1]
            roof["fp64_equivalent_tflops"] = fp["achieved"]
            roof["traffic"] = load_traffic(dom, "_mixed", args.nside, args.nsim)
            roof["executed_frac"] = load_executed(dom, "mixed", args.nside, args.nsim)
        else:
            dom, roof = legendre_roofline(r["kern"], args.steps, fp64_work, "TFLOP/s", PEAK_FP64_MFMA_TFLOPS, passes)
            roof["traffic"] = load_traffic(dom, "", args.nside, args.nsim)
            roof["executed_frac"] = load_executed(dom, "fp64", args.nside, args.nsim)
        roofs[prec] = roof

    def summary(prec):
        r = runs[prec]
        return {"maps_per_s": rates[prec], "ms_per_step": r["ms_total"] / args.steps,
                "latency_ms_per_step": {"p50": pct(r["lat"], 50), "p99": pct(r["lat"], 99), "min": r["lat"][0],
                                        "max": r["lat"][-1]},
                "kernels_ms_per_step": {k: v["ms"] / args.steps for k, v in r["kern"].items()},
                "roofline": roofs[prec], "wall_s": r["wall"]}

    hr = runs[head]
    out = {
        "metric": "maps/s t2alm+alm2t @nside=1024 mixed vs FP64 at 1e-10 err; p50/p99 latency",
        "value": rates[head],
        "unit": "maps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": hr["ms_total"] / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "variant": head,
        "dtype": "i8/f64" if head == "mixed" else "f64",
        "data": "synthetic: seeded Gaussian alm (C_l=1/(l(l+1))) -> alm2t + N(0,1e-3) noise, resident in HBM",
        "config": {"workload": f"t2alm(niter={args.niter})+alm2t nside={args.nside} lmax={lmax} nsim={args.nsim} T maps",
                   "nside": args.nside, "lmax": lmax, "nsim": args.nsim, "niter": args.niter,
                   "precision": ("mixed: int8-slice MFMA Legendre (exact int32 accumulation, FP64 combine), "
                                 "FP64 FFT, FP64 residual refinement" if head == "mixed"
                                 else "plain FP64 path (FP64 MFMA Legendre, FP64 FFT)"),
                   "slices_per_pass": ladder,
                   "parallelism": "1 GPU (maps shard with no communication)"},
        "roofline": roofs[head],
        "latency_ms_per_step": {"p50": pct(hr["lat"], 50), "p99": pct(hr["lat"], 99), "min": hr["lat"][0],
                                "max": hr["lat"][-1],
                                "what": f"one t2alm(niter={args.niter})+alm2t request of {args.nsim} maps"},
        "kernels_ms_per_step": {k: v["ms"] / args.steps for k, v in hr["kern"].items()},
        "wall_s": hr["wall"],
        "roundtrip_alm_rel_err_vs_input": rel,
    }
    if "mixed" in runs:
        out["mixed_maps_per_s"] = rates["mixed"]
        out["mixed"] = summary("mixed")
    if "fp64" in runs:
        out["fp64_maps_per_s"] = rates["fp64"]
        out["fp64"] = summary("fp64")
    if len(runs) == 2:
        out["mixed_vs_fp64"] = rates["mixed"] / rates["fp64"]
    if args.only is None:
        errs, map_err = errors_by_niter(sht, lib, maps, args.niter)
        out["mixed_alm_rel_err_vs_fp64_by_niter"] = errs
        out["mixed_alm2t_map_rel_err_vs_fp64"] = map_err
        out["mixed_err_target"] = 1e-10
        del maps, alm_out, maps_out, step, sht  # frees the handle's device memory
        torch.cuda.empty_cache()
        out["gpu_c0_maps_per_s"] = {p: gpu_c0(p) for p in ("fp64", "mixed")}
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(c0=args.only is None)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
