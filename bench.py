"""Benchmark: batched temperature SHT round trip on one MI355X.

One step = t2alm(niter=3) + alm2t over nsim maps at nside=1024, lmax=2047
(FP64, inputs resident in HBM), on synthetic maps built as alm2t of Gaussian
alm with C_l = 1/(l(l+1)) (l >= 2) plus white noise of variance 1e-3.
Prints one JSON line: throughput, the dominant Legendre kernel's FP64-MFMA
roofline fraction (HIP events on the library stream) and a bounded CPU
baseline from the CPU restatement of the algorithm.

  python bench.py [--steps K] [--warmup W] [--nside 1024 --lmax 2047 --nsim 200 --niter 3]
  python bench.py --pol [--steps K]      # spin-2 qu2eb+eb2qu (separate line, not the value)
"""
from __future__ import annotations

import argparse
import ctypes
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

# MI355X published peaks
PEAK_FP64_MFMA_TFLOPS = 78.6
PEAK_HBM_GBS = 8000.0
KERNEL_CLASSES = ["ring_fft_fwd", "ring_fft_inv", "transpose", "legendre_analysis", "legendre_synthesis",
                  "alias_fold", "spectrum_gather"]


def legendre_flops_per_map(nside, lmax, i0):
    """Algorithmic FP64 flops of one spin-0 Legendre pass for one map with the
    north/south fold: sum_m g_m * 2 * (lmax+1-m) * (2*nside - i0(m)), g_0 = 1,
    g_m = 2 (Re and Im share P)."""
    tot = 0
    for m in range(lmax + 1):
        g = 1 if m == 0 else 2
        tot += g * 2 * (lmax + 1 - m) * (2 * nside - int(i0[m]))
    return float(tot)


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


def cpu_baseline(nsim=32, mstride=16):
    """The oracle (CPU FP64 restatement of the reference algorithm: numpy ring
    FFTs, OpenBLAS dgemm per m, C Legendre recursion) on a bounded sample of the
    same workload (nside=1024, lmax=2047): `nsim` maps, the full forward and
    inverse ring FFTs, and the per-m analysis and synthesis passes for every
    `mstride`-th order m.  The per-m pass times are scaled to all m by the
    per-m work (lmax+1-m) * (4*nside-1 - 2*i0(m)); the round trip is
    fft_fwd + 4 analysis passes + 4 synthesis passes + fft_inv."""
    import numpy as np
    from threadpoolctl import threadpool_info

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
    blas = [p.get("num_threads", 1) for p in threadpool_info() if p.get("user_api") == "blas"]
    return {
        "value": nsim / total,
        "unit": "maps/s",
        "cores": int(max(blas) if blas else 1),
        "kind": "port",
        "sample": (f"oracle (numpy FFT + OpenBLAS dgemm + C recursion) at nside=1024 lmax=2047, {nsim} maps: "
                   f"ring FFT fwd {t['fft_fwd']:.1f}s + inv {t['fft_inv']:.1f}s, analysis {t['analysis']:.1f}s and "
                   f"synthesis {t['synthesis']:.1f}s over every {mstride}th m scaled x{scale:.2f} by per-m work; "
                   f"round trip = FFTs + 4 analysis + 4 synthesis passes"),
        "cpu": _cpu_model(),
    }


def _cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


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


def pol_round_trip(args, lmax):
    """Spin-2 workload (nside 1024, lmax 2047, nsim Q/U pairs): one step =
    qu2eb(niter) + eb2qu on Q/U built as eb2qu of seeded E/B alm.  Prints one
    JSON line (Q/U pairs per second, per-niter E/B error against the input alm);
    the T round trip stays the benchmark value."""
    import torch

    from fastsht_amd import SHT, _lib

    sht = SHT(args.nside, lmax, args.nsim, niter=args.niter, pol=True)
    lib = _lib.load()
    E, B = make_eb(sht, args.seed)
    Q, U = sht.eb2qu(E, B)
    Eo = torch.empty_like(E.permute(2, 1, 0).contiguous()).permute(2, 1, 0)
    Bo = torch.empty_like(Eo)
    p = lambda t: ctypes.c_void_p(t.data_ptr())
    scale = max(float(E.abs().max()), float(B.abs().max()))
    errs = {}
    for it in range(args.niter + 1):  # per-niter error, outside the timed region
        _lib.check(lib.fsht_qu2eb(sht._h, p(Q), p(U), p(Eo), p(Bo), it, _lib.FSHT_DEVICE), sht._h)
        torch.cuda.synchronize()
        errs[str(it)] = max(float((Eo - E).abs().max()), float((Bo - B).abs().max())) / scale
    del E, B
    torch.cuda.empty_cache()

    def step():  # Q/U are overwritten by their own round trip (kept in HBM)
        _lib.check(lib.fsht_qu2eb(sht._h, p(Q), p(U), p(Eo), p(Bo), args.niter, _lib.FSHT_DEVICE), sht._h)
        _lib.check(lib.fsht_eb2qu(sht._h, p(Eo), p(Bo), p(Q), p(U), _lib.FSHT_DEVICE), sht._h)

    sht._use_torch_stream()  # the events below are recorded on torch's current stream
    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    _lib.check(lib.fsht_profile(sht._h, 1), sht._h)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(args.steps):
        step()
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1)
    prof = (ctypes.c_double * 14)()
    _lib.check(lib.fsht_profile_read(sht._h, prof, 14), sht._h)
    kern = {KERNEL_CLASSES[c]: {"launches": int(prof[2 * c]), "ms": prof[2 * c + 1]} for c in range(7)}
    fpair = 4 * legendre_flops_per_map(args.nside, lmax, sht.ring_tab())
    tf = {k: fpair * args.nsim / (kern[k]["ms"] / max(1, kern[k]["launches"]) / 1e3) / 1e12
          for k in ("legendre_analysis", "legendre_synthesis")}
    print(json.dumps({
        "metric": "Q/U pairs/s qu2eb+eb2qu (spin 2; not the benchmark value)",
        "value": args.nsim * args.steps / (ms / 1e3), "unit": "QU pairs/s", "ms_per_step": ms / args.steps,
        "steps": args.steps, "warmup": args.warmup, "dtype": "f64",
        "config": {"workload": f"qu2eb(niter={args.niter})+eb2qu nside={args.nside} lmax={lmax} nsim={args.nsim} "
                               "Q/U from seeded E/B (C_l^BB = 0.05 C_l^EE)",
                   "nside": args.nside, "lmax": lmax, "nsim": args.nsim, "niter": args.niter},
        "legendre_tflops": tf,
        "eb_alm_rel_err_vs_input_by_niter": errs,
        "kernels_ms_per_step": {k: v["ms"] / args.steps for k, v in kern.items()},
    }))


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
                      "config": {"nside": sht.nside, "lmax": sht.lmax, "nsim": sht.nsim, "niter": args.niter,
                                 "maps_in": "pageable NumPy", "alms_and_maps_out": "pinned (fsht_host_alloc)"}}))


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
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--cpu-only", action="store_true", help="only the CPU baseline leg")
    ap.add_argument("--host", action="store_true",
                    help="time the round trip from host arrays instead (pageable maps in, pinned alms/maps out): "
                         "the PCIe-inclusive rate, reported separately from the HBM-resident value")
    ap.add_argument("--pol", action="store_true",
                    help="time the spin-2 round trip qu2eb(niter)+eb2qu instead and report per-niter E/B errors "
                         "(reported separately from the temperature value)")
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
    sht = SHT(args.nside, lmax, args.nsim, niter=args.niter)
    alm, maps = make_inputs(sht, args.seed)
    alm_out = torch.empty_like(alm.permute(2, 1, 0).contiguous()).permute(2, 1, 0)
    maps_out = torch.empty((args.nsim, sht.npix), dtype=torch.float64, device="cuda").t()
    lib = _lib.load()
    if args.host:
        host_round_trip(sht, maps, args)
        return

    def step():
        sht.t2alm(maps, alm_out)
        sht.alm2t(alm_out, maps_out)

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    _lib.check(lib.fsht_profile(sht._h, 1), sht._h)
    # one event per step boundary on the stream the library works on: the
    # whole-run time and the per-step (one request of nsim maps) latencies
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(args.steps + 1)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ev[0].record()
    for i in range(args.steps):
        step()
        ev[i + 1].record()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ms_total = ev[0].elapsed_time(ev[-1])
    lat = sorted(ev[i].elapsed_time(ev[i + 1]) for i in range(args.steps))

    def pct(p):
        return lat[min(len(lat) - 1, int(math.ceil(p / 100.0 * len(lat))) - 1)]
    prof = (ctypes.c_double * 14)()
    _lib.check(lib.fsht_profile_read(sht._h, prof, 14), sht._h)
    _lib.check(lib.fsht_profile(sht._h, 0), sht._h)
    kern = {KERNEL_CLASSES[c]: {"launches": int(prof[2 * c]), "ms": prof[2 * c + 1]} for c in range(7)}
    ms_per_step = ms_total / args.steps
    maps_per_s = args.nsim * args.steps / (ms_total / 1e3)

    # roofline of the dominant kernel (largest total time)
    i0 = sht.ring_tab()
    fpm = legendre_flops_per_map(args.nside, lmax, i0)
    dom = max(("legendre_analysis", "legendre_synthesis"), key=lambda k: kern[k]["ms"])
    avg_ms = kern[dom]["ms"] / max(1, kern[dom]["launches"])
    flops_launch = fpm * args.nsim
    achieved = flops_launch / (avg_ms / 1e3) / 1e12
    traffic = None
    tf = os.path.join(ROOT, "profiles", f"traffic_{dom}_nside{args.nside}_nsim{args.nsim}.json")
    if os.path.exists(tf):
        traffic = json.load(open(tf)).get("bytes_per_launch")

    # sanity: the round trip reproduces the inputs' alm (noise-limited)
    rel = float(((alm_out - alm).abs().max() / alm.abs().max()).item())

    out = {
        "metric": "maps/s t2alm+alm2t @nside=1024 mixed vs FP64 at 1e-10 err; p50/p99 latency",
        "value": maps_per_s,
        "unit": "maps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic: seeded Gaussian alm (C_l=1/(l(l+1))) -> alm2t + N(0,1e-3) noise, resident in HBM",
        "config": {"workload": f"t2alm(niter={args.niter})+alm2t nside={args.nside} lmax={lmax} nsim={args.nsim} T maps",
                   "nside": args.nside, "lmax": lmax, "nsim": args.nsim, "niter": args.niter,
                   "precision": "plain FP64 path (FP64 MFMA Legendre, FP64 FFT)",
                   "parallelism": "1 GPU (maps shard with no communication)"},
        "roofline": {"bound": "mfma", "kernel": dom, "achieved": achieved, "peak": PEAK_FP64_MFMA_TFLOPS,
                     "unit": "TFLOP/s", "frac": achieved / PEAK_FP64_MFMA_TFLOPS, "traffic": traffic,
                     "flops_per_launch": flops_launch, "avg_launch_ms": avg_ms},
        "latency_ms_per_step": {"p50": pct(50), "p99": pct(99), "min": lat[0], "max": lat[-1],
                                "what": f"one t2alm(niter={args.niter})+alm2t request of {args.nsim} maps"},
        "kernels_ms_per_step": {k: v["ms"] / args.steps for k, v in kern.items()},
        "wall_s": wall,
        "roundtrip_alm_rel_err_vs_input": rel,
    }
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline()
    print(json.dumps(out))


if __name__ == "__main__":
    main()
