"""Benchmark: in-place full transpose of a 98304 x 131071 float64 matrix (103 GB) on one MI355X.

Workload (the largest single-GPU configuration): a 98304 x 131071 row-major float64 matrix,
U[0,1) values from a seeded generator, resident in HBM, transposed in place by the library's
planner (hip_transpose_full_es). One step = one full transpose; steps alternate direction
(98304 x 131071 -> 131071 x 98304 and back), each moving the whole 103 GB. Effective bandwidth
= 2 x bytes / time per transpose. A million sampled elements are checked after the timed steps.

Printed: one JSON line with the metric, a roofline object for the route's kernels (every
library launch of the timed steps bracketed by HIP events on the stream it runs on: algorithmic
bytes per step / summed kernel time per step), the HBM traffic per transpose from the committed
rocprofv3 counter summary, and a CPU baseline: the reference test's own CPU golden for a full
transpose (test/cl_unittest.cc:547-551, single-threaded) on the first 768 rows of the same
matrix. Extra measurements of the other configurations ride along in "extra".
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
METRIC = "in-place GB/s (2x bytes/time) and % HBM peak; batched/served p50/p99 latency"
PMC_SUMMARY = os.path.join(ROOT, "profiles", "route_st_98304x131071_f64_pmc.json")
M3A, N3A, ES3A = 98304, 131071, 8


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=20)
    p.add_argument("--warmup", type=int, default=4)
    p.add_argument("--cpu-seconds", type=float, default=10.0,
                   help="approximate CPU work for the baseline sample")
    p.add_argument("--no-extra", action="store_true")
    p.add_argument("--profile-batch", action="store_true",
                   help="only run the 4096-matrix batch (and the same matrices as plain calls)")
    p.add_argument("--profile-st", action="store_true",
                   help="only a 1 GiB calibration copy and one 98304 x 131071 f64 transpose and "
                        "its inverse (for HBM byte counters of the route's kernels)")
    p.add_argument("--profile-st-u8", action="store_true",
                   help="the same for the 250007 x 199999 u8 transpose and its inverse")
    p.add_argument("--profile-large", action="store_true",
                   help="only run the two large prime-dimension transposes (for kernel traces)")
    p.add_argument("--profile-soa", type=int, default=0,
                   help="only run N SoA->ASTA->SoA round trips of the 1 GiB struct array "
                        "(the cycle-leader form), after one untimed round trip")
    p.add_argument("--profile-3stage", type=int, default=0,
                   help="only run N round trips of the reference's 3-stage route on 1800 x 7200 "
                        "int32 (a=18, b=72) and back (a=72, b=18), after one untimed round trip")
    p.add_argument("--profile-pt010", type=int, default=0,
                   help="only run N AoS->ASTA->AoS round trips of the 1 GiB struct array with "
                        "the cycle-following 010 kernel, after one untimed round trip")
    p.add_argument("--no-large", action="store_true",
                   help="skip the 103 GB and 50 GB prime-dimension transposes")
    p.add_argument("--set", action="append", default=[], metavar="KEY=VALUE",
                   help="library tuning knob for a profiling run (repeatable), e.g. "
                        "pt010_lead_order=0 to take the visited-bit form of the "
                        "cycle-following 010")
    p.add_argument("--profile", type=int, default=0,
                   help="only run N AoS->ASTA->AoS round trips (2N LDS-staged launches, as "
                        "in the timed step) and N plain 1 GiB device copies (the calibration "
                        "kernel for HBM byte counters), then exit")
    return p.parse_args()


def time_calls(fn, reps, stream, warm=True):
    """Average device time of fn() over `reps` back-to-back calls, bracketed by one pair of
    HIP events on `stream` (the stream the library launches on). With `warm`, one untimed
    call first (plans the route and sizes the scratch pool outside the timed region)."""
    import torch
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    if warm:
        if fn():
            raise RuntimeError("library call failed")
        stream.synchronize()
    with torch.cuda.stream(stream):
        e0.record(stream)
        for _ in range(reps):
            if fn():
                raise RuntimeError("library call failed")
        e1.record(stream)
    stream.synchronize()
    return e0.elapsed_time(e1) / reps / 1e3  # seconds


def cpu_baseline(rows=768):
    """The reference test's CPU golden for a full transpose (test/cl_unittest.cc:547-551: a 010
    with unit tiles, then a 100), single-threaded as the test runs it, on a bounded sample of
    the headline workload: the first `rows` rows of the 98304 x 131071 f64 matrix (a
    rows x 131071 transpose, same row length, same prime column count). Also the oracle's
    16-thread restatement of the reference's own device route for such a matrix (the
    element-granular cycle-following fallback, src/cl_marshal.cc:769-775) on the same sample."""
    import numpy as np
    import oracle
    x = np.random.default_rng(11).random(rows * N3A)
    want = x.reshape(rows, N3A).T.copy()
    t0 = time.perf_counter()
    oracle.golden_transpose(x, rows, N3A, es=8)
    dt = time.perf_counter() - t0
    if not np.array_equal(x.reshape(N3A, rows), want):
        raise RuntimeError("CPU golden transpose is wrong")
    cores = oracle.num_cpus()
    x = want.reshape(-1).copy()
    t0 = time.perf_counter()
    if oracle.par_0100_pttwac(x, 1, N3A, rows, 1, nthreads=cores):
        raise RuntimeError("CPU fallback restatement failed")
    dtp = time.perf_counter() - t0
    nb = rows * N3A * 8
    base = {"value": round(2 * nb / dt / 1e9, 3), "unit": "GB/s", "cores": 1, "kind": "port",
            "sample": f"{rows} x {N3A} f64 ({nb / 1e9:.2f} GB: the first {rows} of the "
                      f"98304 rows), the reference test's CPU golden (010 with unit tiles, then "
                      f"100; test/cl_unittest.cc:547-551) restated in oracle/oracle.c "
                      f"orc_golden_transpose, single-threaded, {dt:.1f} s"}
    par = {"value": round(2 * nb / dtp / 1e9, 3), "unit": "GB/s", "cores": cores,
           "kind": "port",
           "sample": f"same {rows} x {N3A} f64 back, by the reference's element-granular "
                     f"fallback route (cycle following with atomic bit flags) restated in "
                     f"oracle/oracle.c orc_par_0100_pttwac on {cores} threads, {dtp:.1f} s"}
    return base, par


def headline_3a(lm, steps, warmup):
    """Full in-place transposes of the 98304 x 131071 f64 matrix, alternating direction; the
    timed steps run with per-launch kernel timing on."""
    import torch
    M, N, es = M3A, N3A, ES3A
    n = M * N
    free, _ = torch.cuda.mem_get_info()
    if free < n * es + (16 << 30):
        raise RuntimeError(f"needs {n * es / 1e9:.0f} GB of free device memory")
    x = torch.empty(n, dtype=torch.float64, device="cuda")
    g = torch.Generator(device="cuda").manual_seed(2024)
    step = 1 << 30
    for s in range(0, n, step):
        e = min(n, s + step)
        x[s:e] = torch.rand(e - s, dtype=torch.float64, device="cuda", generator=g)
    idx = torch.randint(0, n, (1 << 20,), device="cuda", generator=g)
    ref = x[idx].clone()
    shape = [M, N]
    routes = {}
    stream = torch.cuda.current_stream()

    def one():
        rows, cols = shape
        if lm.transpose_full(x, rows, cols):
            raise RuntimeError("transpose failed")
        routes[f"{rows}x{cols}"] = lm.last_route()
        shape[0], shape[1] = cols, rows

    for _ in range(warmup):
        one()
    torch.cuda.synchronize()
    lm.set_kernel_timing(True)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    e0.record(stream)
    for _ in range(steps):
        one()
    e1.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    kt = lm.kernel_times()
    lm.set_kernel_timing(False)
    step_s = e0.elapsed_time(e1) / 1e3 / steps
    r, c = idx // N, idx % N
    ok = torch.equal(x[idx] if shape == [M, N] else x[c * M + r], ref)
    del x, idx, ref, r, c
    torch.cuda.empty_cache()
    if not ok:
        raise RuntimeError("sampled elements are not where the transpose puts them")
    return {"step_s": step_s, "wall_s": wall, "kernels": kt, "routes": routes,
            "bytes": 2 * n * es, "checked": 1 << 20}


def extra_batch(lm, reps=3):
    """Mixed batch: 4096 matrices in one graph launch vs the same matrices as per-matrix plain
    calls (asynchronous, on one stream). Both are run once untimed first (the batch captures
    its graph, the plain calls fill the planner's cache), then timed with HIP events. The
    same comparison is repeated per size class on subsets of the same buffers."""
    import torch
    from libmarshal_amd import workloads as wl
    items = wl.batch_mix(4096)
    bufs = [torch.empty(r * c * es, dtype=torch.uint8, device="cuda") for (_, es, r, c, _) in items]
    stream = torch.cuda.current_stream()

    def measure(idx, with_sync=False):
        sel = [items[i] for i in idx]
        sb_ = [bufs[i] for i in idx]
        total = sum(r * c * es for (_, es, r, c, _) in sel)
        batch = lm.Batch([(b, op, es, r, c, t) for b, (op, es, r, c, t) in zip(sb_, sel)])

        def plain():
            for b, (op, es, r, c, t) in zip(sb_, sel):
                if op == wl.OP_TRANSPOSE:
                    e = lm.transpose_full(b, r, c, elem_bytes=es)
                elif op == wl.OP_AOS_ASTA:
                    e = lm.aos_asta(b, r, c, t, elem_bytes=es)
                else:
                    e = lm.soa_asta(b, r, c, t, elem_bytes=es)
                if e:
                    return True
            return False
        if batch.run() or plain():
            raise RuntimeError("batch warm-up failed")
        sb = time_calls(batch.run, reps, stream, warm=False)
        sp = time_calls(plain, 2, stream, warm=False)
        out = {"matrices": len(sel), "gbytes": round(total / 1e9, 2),
               "kernels_per_launch": batch.info("kernels"), "solo": batch.info("solo"),
               "batched_GBs": round(2 * total / sb / 1e9, 1), "batched_ms": round(sb * 1e3, 2),
               "plain_GBs": round(2 * total / sp / 1e9, 1), "plain_ms": round(sp * 1e3, 2),
               "speedup": round(sp / sb, 2)}
        if with_sync:
            # The same plain calls with the library's default timing convention (each call
            # measured and synchronized, as a ported caller of the reference gets by default).
            et = lm.Elapsed()
            lm.set_profiling(True)
            t0 = time.perf_counter()
            for b, (op, es, r, c, t) in zip(sb_, sel):
                if op == wl.OP_TRANSPOSE:
                    e = lm.transpose_full(b, r, c, elem_bytes=es, elapsed=et)
                elif op == wl.OP_AOS_ASTA:
                    e = lm.aos_asta(b, r, c, t, elem_bytes=es, elapsed=et)
                else:
                    e = lm.soa_asta(b, r, c, t, elem_bytes=es, elapsed=et)
                if e:
                    raise RuntimeError("plain call failed")
            ss = time.perf_counter() - t0
            lm.set_profiling(False)
            out["plain_synchronized_GBs"] = round(2 * total / ss / 1e9, 1)
            out["speedup_vs_synchronized"] = round(ss / sb, 2)
        batch.close()
        return out
    out = measure(range(len(items)), with_sync=True)
    classes = {}
    for name, lo, hi in (("lt_1MB", 0, 1 << 20), ("1_16MB", 1 << 20, 16 << 20),
                         ("16_128MB", 16 << 20, 128 << 20), ("ge_128MB", 128 << 20, 1 << 62)):
        idx = [i for i, (_, es, r, c, _) in enumerate(items) if lo <= r * c * es < hi]
        m = measure(idx)
        classes[name] = {k: m[k] for k in ("matrices", "gbytes", "batched_GBs", "plain_GBs",
                                           "speedup")}
    out["by_size"] = classes
    del bufs
    torch.cuda.empty_cache()
    return out


def extra_serve(lm, rate=2000.0, duration=1.0):
    """Poisson request stream through the continuous-batching server vs sequential calls."""
    import numpy as np
    import torch
    from libmarshal_amd import workloads as wl
    trace = wl.request_stream(rate, duration, seed=7)
    arena = torch.empty(64 << 30, dtype=torch.uint8, device="cuda")
    server = lm.Server(256 << 30, streams=4, max_batch=256)
    wl.serve_trace(lm, server, arena, trace[:100], realtime=False)  # grow per-stream buffers
    server.drain()
    server.reset_stats()
    wl.serve_trace(lm, server, arena, trace, realtime=True)
    server.drain(600)
    st = server.stats()
    server.close()
    lat, nb, span = wl.plain_trace(lm, arena, trace, realtime=True)
    lat = np.sort(np.array(lat))
    out = {"requests": len(trace), "rate_per_s": rate,
           "payload_gbytes": round(sum(wl.request_bytes(r) for r in trace) / 1e9, 1),
           "served_p50_ms": round(st["p50_ms"], 2), "served_p99_ms": round(st["p99_ms"], 2),
           "served_GBs": round(st["bytes"] / st["span_s"] / 1e9, 1),
           "batches": st["batches"], "largest_batch": st["max_batch"],
           "sequential_p50_ms": round(1e3 * lat[len(lat) // 2], 2),
           "sequential_p99_ms": round(1e3 * lat[int(0.99 * (len(lat) - 1))], 2),
           "sequential_GBs": round(nb / span / 1e9, 1)}
    del arena
    torch.cuda.empty_cache()
    return out


def extra_large(lm, M, N, es):
    """Full in-place transpose of a large non-square matrix (prime dimensions allowed) and
    back, each timed with HIP events after one untimed round trip; a million sampled
    positions are checked after each direction."""
    import torch
    n = M * N
    free, _ = torch.cuda.mem_get_info()
    if free < n * es + (8 << 30):
        return {"skipped": f"needs {(n * es) / 1e9:.0f} GB of free device memory"}
    dt = torch.float64 if es == 8 else torch.uint8
    x = torch.empty(n, dtype=dt, device="cuda")
    g = torch.Generator(device="cuda").manual_seed(M ^ N)
    step = 1 << 30
    for s in range(0, n, step):
        e = min(n, s + step)
        if es == 8:
            x[s:e].copy_(torch.rand(e - s, dtype=dt, device="cuda", generator=g))
        else:
            x[s:e].copy_(torch.randint(0, 256, (e - s,), dtype=dt, device="cuda", generator=g))
    idx = torch.randint(0, n, (1 << 20,), device="cuda", generator=g)
    ref = x[idx].clone()
    r, c = idx // N, idx % N
    stream = torch.cuda.current_stream()
    for _ in range(1):  # untimed: plans both directions and sizes the scratch pool
        if lm.transpose_full(x, M, N) or lm.transpose_full(x, N, M):
            raise RuntimeError("large transpose failed")
    fwd = time_calls(lambda: lm.transpose_full(x, M, N), 1, stream, warm=False)
    route = lm.last_route()
    ok = torch.equal(x[c * M + r], ref)
    back = time_calls(lambda: lm.transpose_full(x, N, M), 1, stream, warm=False)
    route_back = lm.last_route()
    ok = ok and torch.equal(x[idx], ref)
    del x, idx, ref, r, c
    torch.cuda.empty_cache()
    if not ok:
        raise RuntimeError(f"{M}x{N} transpose gave wrong elements")
    gbs_f = 2 * n * es / fwd / 1e9
    gbs_b = 2 * n * es / back / 1e9
    return {"shape": [M, N], "elem_bytes": es, "route": route, "route_back": route_back,
            "ms": round(fwd * 1e3, 2), "ms_back": round(back * 1e3, 2),
            "GBs": round(gbs_f, 1), "GBs_back": round(gbs_b, 1),
            "pct_hbm_peak": round(100 * gbs_f / HBM_PEAK_GBS, 2),
            "pct_hbm_peak_back": round(100 * gbs_b / HBM_PEAK_GBS, 2)}


def cpu_config1():
    """Config 1 on the host: the reference test's CPU golden (single-threaded) and the oracle's
    restatement of the reference's 3-stage route (100 -> 010 -> 0100) on every host thread,
    1800 x 7200 int32, a=18, b=72."""
    import numpy as np
    import oracle
    M, N = 1800, 7200
    x = oracle.tagged(M * N, 4, seed=3).view(np.uint32)
    want = x.reshape(M, N).T.copy()
    y = x.copy()
    reps = 5
    t0 = time.perf_counter()
    for _ in range(reps):
        oracle.golden_transpose(y, M, N, es=4)
        oracle.golden_transpose(y, N, M, es=4)
    dt1 = (time.perf_counter() - t0) / (2 * reps)
    cores = oracle.num_cpus()
    t0 = time.perf_counter()
    oracle.par_transpose_3(x, M // 18, 18, N // 72, 72, nthreads=cores)
    dt = time.perf_counter() - t0
    if not np.array_equal(x.reshape(N, M), want):
        raise RuntimeError("CPU 3-stage restatement is wrong")
    return {"golden_1thread": {"value": round(2 * x.nbytes / dt1 / 1e9, 3), "unit": "GB/s",
                               "cores": 1, "kind": "port",
                               "sample": "1800 x 7200 int32 and back, 5 round trips, the "
                                         "reference test's CPU golden (oracle/oracle.c "
                                         "orc_golden_transpose)"},
            "route3": {"value": round(2 * x.nbytes / dt / 1e9, 3), "unit": "GB/s",
                       "cores": cores, "kind": "port",
                       "sample": "1800 x 7200 int32, a=18 b=72, stages 100 -> 010 -> 0100 "
                                 "(oracle/oracle.c orc_par_transpose_3)"}}


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this benchmark runs on one GPU only; --gpus must be 1 "
                          "(a transpose larger than one GPU shards by rows with one all-to-all exchange)"}))
        sys.exit(2)
    import torch
    import libmarshal_amd as lm

    torch.cuda.set_device(0)
    lm.set_profiling(False)  # asynchronous calls: no per-call synchronization
    for kv in args.set:
        key, _, value = kv.partition("=")
        lm.set_tuning(key, int(value))
    if args.profile_st or args.profile_st_u8:
        c1 = torch.rand(1 << 28, device="cuda")  # 1 GiB
        c2 = torch.empty_like(c1)
        c2.copy_(c1)  # the calibration kernel: reads and writes exactly 1 GiB
        del c1, c2
        torch.cuda.empty_cache()
        if args.profile_st:
            M, N, dt = 98304, 131071, torch.float64
        else:
            M, N, dt = 250007, 199999, torch.uint8
        x = torch.empty(M * N, dtype=dt, device="cuda")
        step = 1 << 30
        for s0 in range(0, M * N, step):
            s1 = min(M * N, s0 + step)
            if dt == torch.uint8:
                x[s0:s1].random_(0, 256)
            else:
                torch.rand(s1 - s0, dtype=dt, out=x[s0:s1])
        if lm.transpose_full(x, M, N):
            raise RuntimeError("transpose failed")
        r1 = lm.last_route()
        if lm.transpose_full(x, N, M):
            raise RuntimeError("transpose failed")
        torch.cuda.synchronize()
        print(f"profile run: 1 GiB copy, {M} x {N} {str(dt).split('.')[-1]} transpose ({r1}) "
              f"and back ({lm.last_route()})", flush=True)
        return
    if args.profile_batch:
        print(json.dumps({"profile_batch": extra_batch(lm)}), flush=True)
        return
    if args.profile_large:
        for (M, N, es) in ((98304, 131071, 8), (250007, 199999, 1)):
            print(json.dumps({"profile_large": extra_large(lm, M, N, es)}), flush=True)
        return
    n, w, t = 1 << 24, 16, 64
    nbytes = n * w * 4
    stream = torch.cuda.current_stream()
    g = torch.Generator(device="cuda").manual_seed(1234)
    d = torch.rand(n * w, device="cuda", generator=g) * 2 - 1  # U[-1, 1)
    ref = d.clone()

    if args.profile_3stage:
        M, N = 1800, 7200
        x = torch.randint(-2**31, 2**31 - 1, (M * N,), dtype=torch.int32, device="cuda")
        x0 = x.clone()

        def tri_round():
            return lm.transpose(x, M // 18, 18, N // 72, 72, 1, 3) or \
                lm.transpose(x, N // 72, 72, M // 18, 18, 1, 3)
        if tri_round():
            raise RuntimeError("3-stage transpose failed")
        s = time_calls(tri_round, args.profile_3stage, stream, warm=False) / 2
        print(json.dumps({"profile_3stage": {"route": lm.last_route(),
                                             "GBs": round(2 * M * N * 4 / s / 1e9, 1),
                                             "round_trips": args.profile_3stage,
                                             "restored": bool(torch.equal(x, x0))}}), flush=True)
        return
    if args.profile_pt010:
        def pt_round():
            return lm.aos_asta(d, n, w, t, method=2, elem_bytes=4) or \
                lm.asta_aos(d, n, w, t, method=2, elem_bytes=4)
        if pt_round():
            raise RuntimeError("conversion failed")
        s = time_calls(pt_round, args.profile_pt010, stream, warm=False) / 2
        print(json.dumps({"profile_pt010": {"route": lm.last_route(),
                                            "GBs": round(2 * nbytes / s / 1e9, 1),
                                            "round_trips": args.profile_pt010,
                                            "restored": bool(torch.equal(d, ref))}}), flush=True)
        return
    if args.profile_soa:
        def soa_round():
            return lm.soa_asta(d, n, w, t, elem_bytes=4) or lm.asta_soa(d, n, w, t, elem_bytes=4)
        if soa_round():
            raise RuntimeError("conversion failed")
        s = time_calls(soa_round, args.profile_soa, stream, warm=False) / 2
        ok = torch.equal(d, ref)
        print(json.dumps({"profile_soa": {"route": lm.last_route(), "GBs": round(2 * nbytes / s / 1e9, 1),
                                          "round_trips": args.profile_soa, "restored": ok}}), flush=True)
        return
    if args.profile:
        d2 = torch.empty_like(d)
        for _ in range(60):  # GPU ramp-up on plain copies, so the profiled launches are steady
            d2.copy_(d)
        for _ in range(args.profile):
            if lm.aos_asta(d, n, w, t, method=1, elem_bytes=4) or \
                    lm.asta_aos(d, n, w, t, method=1, elem_bytes=4):
                raise RuntimeError("conversion failed")
            d2.copy_(d)
        torch.cuda.synchronize()
        print(f"profile run: {2 * args.profile} x k_010_bs and {args.profile + 60} x 1 GiB copy",
              flush=True)
        return

    # The headline: the 103 GB matrix, with the device to itself.
    head = headline_3a(lm, args.steps, args.warmup)
    step_s = head["step_s"]
    value = head["bytes"] / step_s / 1e9
    kern_ms = sum(k["ms"] for k in head["kernels"]) / args.steps
    achieved = head["bytes"] / (kern_ms / 1e3) / 1e9
    traffic = None
    if os.path.exists(PMC_SUMMARY):
        try:
            traffic = json.load(open(PMC_SUMMARY)).get("hbm_bytes_per_transpose")
        except (OSError, ValueError):
            traffic = None
    kernels = [{"name": k["name"], "launches_per_step": round(k["launches"] / args.steps, 2),
                "us_per_launch": round(1e3 * k["ms"] / max(k["launches"], 1), 2),
                "ms_per_step": round(k["ms"] / args.steps, 3),
                "GBs": round(k["bytes"] / (k["ms"] / 1e3) / 1e9, 1) if k["ms"] > 0 else None}
               for k in head["kernels"]]

    extra = {}
    if not args.no_extra:
        # Config 2: AoS -> ASTA -> AoS of 2^24 structs x 16 f32, LDS-staged and cycle
        # following, and SoA -> ASTA -> SoA. The first ~20 round trips after start-up run up to
        # 30 % slower while the GPU ramps up, so 30 untimed round trips come first.
        def bs_round():
            return lm.aos_asta(d, n, w, t, method=1, elem_bytes=4) or \
                lm.asta_aos(d, n, w, t, method=1, elem_bytes=4)
        for _ in range(30):
            if bs_round():
                raise RuntimeError("conversion failed")
        c2 = {}
        s = time_calls(bs_round, 50, stream, warm=False) / 2
        c2["aos_asta_bs_GBs"] = round(2 * nbytes / s / 1e9, 1)
        c2["aos_asta_bs_pct_hbm_peak"] = round(100 * 2 * nbytes / s / 1e9 / HBM_PEAK_GBS, 2)
        s = time_calls(lambda: lm.aos_asta(d, n, w, t, method=2, elem_bytes=4) or
                       lm.asta_aos(d, n, w, t, method=2, elem_bytes=4), 5, stream) / 2
        c2["aos_asta_cycle_following_GBs"] = round(2 * nbytes / s / 1e9, 1)
        s = time_calls(lambda: lm.soa_asta(d, n, w, t, elem_bytes=4) or
                       lm.asta_soa(d, n, w, t, elem_bytes=4), 5, stream) / 2
        c2["soa_asta_GBs"] = round(2 * nbytes / s / 1e9, 1)
        c2["soa_route"] = lm.last_route()
        if not torch.equal(d.view(torch.int32), ref.view(torch.int32)):
            raise RuntimeError("config-2 round trips are not bit-exact")
        extra["aos_soa_asta_2e24x16_f32"] = c2
        # Config 1: 1800 x 7200 int32, a = 18, b = 72, the reference's 3-stage route, and the
        # planner's route (~40 us per call: 50 calls so that launch gaps average out).
        M, N = 1800, 7200
        x = torch.randint(-2**31, 2**31 - 1, (M * N,), dtype=torch.int32, device="cuda")
        x0 = x.clone()
        tri = lambda: lm.transpose(x, M // 18, 18, N // 72, 72, 1, 3) or \
            lm.transpose(x, N // 72, 72, M // 18, 18, 1, 3)
        s = time_calls(tri, 25, stream) / 2
        extra["transpose_1800x7200_i32_3stage_GBs"] = round(2 * M * N * 4 / s / 1e9, 1)
        s = time_calls(lambda: lm.transpose_full(x, M, N) or lm.transpose_full(x, N, M), 25,
                       stream) / 2
        extra["transpose_1800x7200_i32_planner_GBs"] = round(2 * M * N * 4 / s / 1e9, 1)
        extra["planner_route"] = lm.last_route()
        if not torch.equal(x, x0):
            raise RuntimeError("config-1 round trips are not bit-exact")
        del x, x0
        # the other sizes of the reference's paper run (test/cl_unittest.cc:386-387), float
        sizes = {}
        for (M, N) in ((2500, 5100), (3200, 4000), (3900, 3300), (5100, 2500), (7200, 1800)):
            x = torch.rand(M * N, device="cuda")
            s = time_calls(lambda: lm.transpose_full(x, M, N), 50, stream)
            sizes[f"{M}x{N}"] = {"GBs": round(2 * M * N * 4 / s / 1e9, 1), "route": lm.last_route()}
            del x
        extra["transpose_paper_sizes_f32_planner"] = sizes
        torch.cuda.empty_cache()
        extra["cpu_transpose_1800x7200_i32"] = cpu_config1()
        if not args.no_large:
            extra["transpose_250007x199999_u8"] = extra_large(lm, 250007, 199999, 1)
        extra["batch_4096_mixed"] = extra_batch(lm)
        extra["serve_poisson_2000"] = extra_serve(lm)
        # three times the load for half a second: requests queue, so admission matters
        extra["serve_poisson_6000"] = extra_serve(lm, rate=6000.0, duration=0.5)
    del ref, d
    torch.cuda.empty_cache()

    base, par = cpu_baseline()
    extra["cpu_3a_sample_fallback_route"] = par
    out = {
        "metric": METRIC,
        "value": round(value, 2),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(step_s * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic: U[0,1) f64 from a seeded generator, resident in HBM",
        "config": {"workload": "98304 x 131071 f64 full in-place transpose (103 GB), steps "
                               "alternating direction",
                   "rows": M3A, "cols": N3A, "elem_bytes": ES3A,
                   "routes": head["routes"], "parallelism": "single GPU"},
        "pct_hbm_peak": round(100 * value / HBM_PEAK_GBS, 2),
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": traffic,
                     "kernel": "the route's kernels together: " +
                               " + ".join(k["name"] for k in head["kernels"]),
                     "algorithmic_bytes_per_step": head["bytes"],
                     "kernel_ms_per_step": round(kern_ms, 3),
                     "kernels": kernels},
        "wall_s": round(head["wall_s"], 3),
        "checked_elements": head["checked"],
        "extra": extra,
        "cpu_baseline": base,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
