#!/usr/bin/env python3
"""Benchmark: exact touch-configuration search on one MI355X.

Workload (one "step"): 65,536 segments that start at exact-verified free
states of the torus scene (env torus R=0.3, r=0.08, 2,048 triangles; robot
torus R=0.1, r=0.02, 1,024 triangles), head in uniform random SE(3)
directions for a length ~ U[0.05, 0.5], are checked at discrete steps of
1/1024 with early exit at the first contact (transitStateToWithContact), and
refine each contact with 20 bisection levels.  A query is one state validity
decision made under the reference's semantics: every evaluated discrete step
plus every bisection probe.  Inputs sit in HBM before the timed region.

Prints one JSON line (metric, value in queries/s, roofline of the segment
kernel, CPU-oracle baseline on a bounded sample).  The default run then
measures the other configurations the same way and adds them under
"workloads": 1M state validity (torus), 4M uneven segments and a 4,096-
milestone visibility matrix (shell scene), and the learned prescreen (shell
scene), each with its kernel time, roofline, oracle baseline and mismatch
count (--headline-only skips them; --workload runs one of them alone).
"""
import argparse
import copy
import json
import os
import subprocess
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "exact-verified SE(3) validity queries/sec (0 mismatches) + free samples/sec w/ surrogate"
N_SEG = 65536
VD = 1.0 / 1024
LEVELS = 20
# Straight-line binary64 operation counts of the oracle's node-pair box test
# and 17-axis triangle test (oracle/oracle.c obb_disjoint / or_tri_intersect;
# adds, subtracts, multiplies and square roots).
C_BV = 286
C_TRI = 753
PEAK_FP32_TFLOPS = 157.3
# The bound field takes "hbm" or "mfma"; these kernels are compute-bound on
# the vector ALUs and issue no matrix instructions, so the roof they are
# measured against is the fp32 vector peak, named here.
VALU_ROOF = "fp32 VALU (vector ALUs; the kernel issues no matrix instructions)"
# HBM bytes per launch of the touch kernel on this workload, from rocprofv3
# --pmc FETCH_SIZE and WRITE_SIZE (KiB per dispatch), each scaled by what the
# same counter reads for a 256 MiB device copy (FETCH_SIZE 131,094 KiB, i.e.
# x2.000; WRITE_SIZE 262,144 KiB, x1.000); see
# profiles/k_segments_touch_torus65536.md.  Measured, not computed here.
TRAFFIC_FETCH_KIB = 166188
TRAFFIC_WRITE_KIB = 490440
# the same for one launch of the uneven-segment kernel (4M segments, shell
# scene); see profiles/k_segments_transit_shell_uneven.md
TRAFFIC_UNEVEN_FETCH_KIB = 64489375
TRAFFIC_UNEVEN_WRITE_KIB = 122651438
PEAK_HBM_GBS = 8000.0
FETCH_SCALE = 262144 / 131094
WRITE_SCALE = 1.0


def build_workload(uw, W, n_seg, seed=0):
    starts = []
    have = 0
    k = 0
    while have < n_seg:
        qs = W.random_states(2 * n_seg, seed=1000 + seed + k)
        v = uw.is_valid_states(qs)
        starts.append(qs[v])
        have += int(v.sum())
        k += 1
    starts = np.concatenate(starts)[:n_seg]
    ends = W.segment_ends(starts, seed=2000 + seed)
    return np.ascontiguousarray(starts), np.ascontiguousarray(ends)


def cpu_baseline(ow, starts, ends, gpu_res, target_s, single_s):
    """Time the oracle on a bounded sample of the same segments (all threads
    OMP allows, then one thread) and check the GPU's answers for that sample."""
    from oracle import oracle as O
    threads = _sample_threads()
    # timed with the host's C library, as the reference runs
    O.set_libm(1)
    n = 256
    t0 = time.perf_counter()
    o = ow.touch_batch(starts[:n], ends[:n], VD, LEVELS, nthreads=threads)
    dt = time.perf_counter() - t0
    if dt < target_s:
        n = int(min(len(starts), max(n, n * target_s / max(dt, 1e-3))))
        t0 = time.perf_counter()
        o = ow.touch_batch(starts[:n], ends[:n], VD, LEVELS, nthreads=threads)
        dt = time.perf_counter() - t0
    q = int(o["queries"].sum())
    # a workload shorter than the budget is timed again until the budget is
    # spent, so the rate is not a fraction-of-a-second reading
    passes, q_all, dt_all = 1, q, dt
    while dt_all < target_s:
        t0 = time.perf_counter()
        o = ow.touch_batch(starts[:n], ends[:n], VD, LEVELS, nthreads=threads)
        dt_all += time.perf_counter() - t0
        q_all += int(o["queries"].sum())
        passes += 1
    what = f"{passes} passes over " if passes > 1 else ""

    def one(k):
        sl = slice((k % 256) * 16, (k % 256 + 1) * 16)
        return int(ow.touch_batch(starts[sl], ends[sl], VD, LEVELS, nthreads=1)["queries"].sum())
    st, k = _single_thread(one, single_s)
    # parity: the correctly rounded restatement on the same segments
    O.set_libm(0)
    o = ow.touch_batch(starts[:n], ends[:n], VD, LEVELS, nthreads=threads)
    mism = 0
    for key, gk in (("done", "done"), ("lo", "lo"), ("hi", "hi"), ("queries", "steps")):
        a = o[key]
        b = gpu_res[gk][:n]
        if key == "done":
            a = a.astype(bool)
        mism += int((a != b).sum())
    return _cpu_baseline(q_all / dt_all, "queries/s",
                         f"{what}first {n} of the {len(starts)} touch segments ({q} queries per pass, "
                         f"{dt_all:.1f} s; the host C library's sin/acos/atan2, as the reference runs)", st,
                         f"{k} chunks of 16 of the same segments, one thread"), mism, n


def oracle_flops_per_query(ow, starts, ends, n, vd=VD):
    """Mean algorithmic flops per query from the oracle's traversal counters
    over the states the first n segments evaluate."""
    from oracle import oracle as O
    qs = []
    for i in range(n):
        lf, st, done, lt, t, ns = ow.transit(starts[i], ends[i], vd)
        dist = O.distance(starts[i], ends[i])
        d = vd
        for k in range(ns):
            qs.append(O.interpolate(starts[i], ends[i], d * (1.0 / dist)))
            if d < dist:
                d = min(d + vd, dist)
            else:
                d += vd
    qs = np.array(qs)
    _, nbv, ntri = ow.valid_states(qs, counters=True)
    return float(C_BV * nbv.mean() + C_TRI * ntri.mean()), float(nbv.mean()), float(ntri.mean())


def _compute_roofline(flops_q, nbv, ntri, queries, kern_ms, kernel, sample):
    """The bench line's roofline object for a VALU-bound traversal kernel:
    algorithmic binary64 flops per query (the oracle's node and triangle
    tests on `sample`) over the kernel's time, against the fp32 peak.  HBM
    traffic is counted for the headline kernel only (null here)."""
    achieved = queries * flops_q / (kern_ms * 1e-3) / 1e12
    return dict(bound="mfma", roof=VALU_ROOF, achieved=round(achieved, 3), peak=PEAK_FP32_TFLOPS, unit="TFLOP/s",
                frac=round(achieved / PEAK_FP32_TFLOPS, 5), traffic=None,
                note="compute-bound on vector ALUs (no matrix instructions); the compute roof is MI355X's fp32 "
                     "peak; algorithmic flops = oracle node/triangle test counts x their straight-line binary64 "
                     "op counts",
                flops_per_query=round(flops_q, 1), oracle_node_tests_per_query=round(nbv, 2),
                oracle_tri_tests_per_query=round(ntri, 3), kernel=kernel, kernel_ms=round(kern_ms, 4),
                flops_sample=sample)


def _world(args, pyosr, env, rob):
    """The workload's world, with the traversal knobs the command line set."""
    uw = pyosr.create_unit_world_from_arrays(env, rob)
    knobs = {k: v for k, v in (("leaf_size", args.leaf_size), ("pool_chunk", args.pool_chunk),
                               ("queue_depth", args.queue_depth)) if v}
    if knobs:
        uw.tuning = knobs
    return uw


class _Timer:
    """HIP events on the stream the engine's kernels run on."""

    def __init__(self, torch, stream):
        self.torch, self.stream = torch, stream

    def run(self, fn, steps, warmup):
        torch = self.torch
        for _ in range(warmup):
            fn()
        torch.cuda.synchronize()
        ev0 = torch.cuda.Event(enable_timing=True)
        ev1 = torch.cuda.Event(enable_timing=True)
        t0 = time.perf_counter()
        ev0.record(self.stream)
        for _ in range(steps):
            fn()
        ev1.record(self.stream)
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        return wall, ev0.elapsed_time(ev1) / steps


def _line(workload, value, wall, args, extra):
    line = {
        "metric": METRIC, "value": round(value, 1), "unit": "queries/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(wall / args.steps * 1e3, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64 (exact leaf tests) + f32 (conservative culling)",
        "data": "synthetic (seeded scenes, Shoemake states)",
        "config": {"workload": workload},
    }
    line.update(extra)
    return line


_HOST = None


def host_cpu():
    """The host CPU the oracle's baseline runs on: lscpu's model and topology
    (of the whole host), the CPUs this process may use (a container or
    cgroup slice may hold fewer) and the OpenMP settings in force."""
    global _HOST
    if _HOST is not None:
        return _HOST
    info = {}
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=20).stdout
        for line in out.splitlines():
            k, _, v = line.partition(":")
            v = v.strip()
            if k == "Model name":
                info["model"] = v
            elif k == "Socket(s)":
                info["sockets"] = int(v)
            elif k == "Core(s) per socket":
                info["cores_per_socket"] = int(v)
            elif k == "Thread(s) per core":
                info["threads_per_core"] = int(v)
            elif k == "CPU(s)":
                info["cpus"] = int(v)
    except (OSError, ValueError, subprocess.SubprocessError):
        pass
    if "model" not in info:
        try:
            with open("/proc/cpuinfo") as f:
                for line in f:
                    if line.startswith("model name"):
                        info["model"] = line.split(":", 1)[1].strip()
                        break
        except OSError:
            pass
    tpc = info.get("threads_per_core", 1)
    info["smt"] = "on" if tpc > 1 else "off"
    info["cpus_usable"] = len(os.sched_getaffinity(0))
    info["physical_cores_usable"] = max(1, info["cpus_usable"] // max(tpc, 1))
    info["omp"] = {k: os.environ.get(k) for k in ("OMP_NUM_THREADS", "OMP_PROC_BIND", "OMP_PLACES")}
    _HOST = info
    return info


def _sample_threads():
    """Threads of the oracle's multi-threaded baseline: OMP_NUM_THREADS when
    set, else the physical cores this process may use."""
    if os.environ.get("OMP_NUM_THREADS"):
        return int(os.environ["OMP_NUM_THREADS"])
    return host_cpu()["physical_cores_usable"]


def _single_thread(fn, budget_s):
    """fn(k) runs chunk k of a workload on one thread and returns its units;
    chunks run until budget_s seconds are spent.  Returns (rate, chunks)."""
    done, dt, k = 0, 0.0, 0
    while dt < budget_s:
        t0 = time.perf_counter()
        done += fn(k)
        dt += time.perf_counter() - t0
        k += 1
    return done / dt, k


def _cpu_baseline(value, unit, sample, single, single_sample):
    return dict(value=value, unit=unit, cores=_sample_threads(), kind="port", sample=sample,
                single_thread=dict(value=single, unit=unit, sample=single_sample), host=host_cpu())


def run_states(args, torch, pyosr, W):
    """1M single-state validity queries, torus scene."""
    env, rob = W.torus_scene(0)
    uw = _world(args, pyosr, env, rob)
    n = args.segments or 1_000_000
    qs = W.random_states(n, seed=1)
    stream = torch.cuda.Stream()
    torch.cuda.set_stream(stream)
    d_q = torch.from_numpy(qs).cuda()
    d_out = torch.zeros(n, dtype=torch.uint8, device="cuda")
    wall, kern_ms = _Timer(torch, stream).run(
        lambda: uw.valid_states_device(d_q.data_ptr(), n, d_out.data_ptr(), stream.cuda_stream), args.steps, args.warmup)
    uw.check_device()
    out = d_out.cpu().numpy().astype(bool)
    extra = {"mismatches": None, "cpu_baseline": None, "kernel": "k_valid_states", "kernel_ms": round(kern_ms, 4)}
    if not args.no_cpu:
        from oracle import oracle as O
        O.set_libm(False)
        ow = O.OracleWorld().setup(*env, *rob)
        # the states in order, 200,000 at a time, until the time budget is spent
        m, dt, mism, cs = 0, 0.0, 0, []
        while dt < args.cpu_seconds and m < n:
            sl = slice(m, min(n, m + 200_000))
            t0 = time.perf_counter()
            c, nbv, ntri = ow.valid_states(qs[sl], nthreads=_sample_threads(), counters=True)
            dt += time.perf_counter() - t0
            mism += int((out[sl] != c.astype(bool)).sum())
            cs.append((nbv, ntri))
            m = sl.stop
        nbv = np.concatenate([a for a, _ in cs])
        ntri = np.concatenate([b for _, b in cs])
        # the whole workload takes well under the budget: time further passes
        passes, done_q = 1, m
        while dt < args.cpu_seconds:
            t0 = time.perf_counter()
            ow.valid_states(qs[:m], nthreads=_sample_threads())
            dt += time.perf_counter() - t0
            done_q += m
            passes += 1
        extra["mismatches"] = mism
        what = f"{passes} passes over " if passes > 1 else ""
        nch = max(1, min(m, n) // 2000)
        st, k = _single_thread(lambda k: int(ow.valid_states(qs[(k % nch) * 2000:(k % nch + 1) * 2000],
                                                               nthreads=1).size), args.single_seconds)
        extra["cpu_baseline"] = _cpu_baseline(done_q / dt, "queries/s",
                                              f"{what}first {m} of the {n} states ({dt:.2f} s)", st,
                                              f"{k} chunks of 2,000 of the same states, one thread")
        fq = C_BV * nbv.mean() + C_TRI * ntri.mean()
        extra["roofline"] = _compute_roofline(float(fq), float(nbv.mean()), float(ntri.mean()), n, kern_ms,
                                              "k_valid_states", f"the first {m} states")
    line = _line(f"state validity: torus env 2,048 tris / robot 1,024 tris, {n} states", n * args.steps / wall,
                 wall, args, extra)
    line["config"]["states"] = n
    return line


def _free_states(uw, W, n, seed):
    out, have, k = [], 0, 0
    while have < n:
        qs = W.random_states(2 * n, seed=seed + k)
        v = uw.is_valid_states(qs)
        out.append(qs[v])
        have += int(v.sum())
        k += 1
    return np.ascontiguousarray(np.concatenate(out)[:n])


def run_uneven(args, torch, pyosr, W):
    """Advance-until-contact over uneven segments (lognormal step counts),
    shell scene, in the persistent segment kernel."""
    env, rob = W.shell_scene(0)
    uw = _world(args, pyosr, env, rob)
    n = args.segments or 4 * 1024 * 1024
    vd = VD
    starts = _free_states(uw, W, n, 5000)
    steps_req = W.lognormal_steps(n, seed=5001)
    ends = W.segment_ends(starts, seed=5002, lengths=steps_req * vd)
    stream = torch.cuda.Stream()
    torch.cuda.set_stream(stream)
    d_f = torch.from_numpy(starts).cuda()
    d_t = torch.from_numpy(ends).cuda()
    d_done = torch.zeros(n, dtype=torch.uint8, device="cuda")
    d_lt = torch.zeros(n, dtype=torch.float64, device="cuda")
    d_ta = torch.zeros(n, dtype=torch.float64, device="cuda")
    d_st = torch.zeros(n, dtype=torch.int32, device="cuda")
    timer = _Timer(torch, stream)
    wall, kern_ms = timer.run(
        lambda: uw.transit_device(d_f.data_ptr(), d_t.data_ptr(), n, vd, d_done.data_ptr(), d_lt.data_ptr(),
                                  d_ta.data_ptr(), d_st.data_ptr(), stream=stream.cuda_stream),
        args.steps, args.warmup)
    uw.check_device()
    steps = d_st.cpu().numpy().astype(np.int64)
    done = d_done.cpu().numpy().astype(bool)
    q = int(steps.sum())
    extra = {"mismatches": None, "cpu_baseline": None, "kernel": "k_segments<transit> (persistent)",
             "kernel_ms": round(kern_ms, 4)}
    if not args.no_cpu:
        from oracle import oracle as O
        O.set_libm(False)
        ow = O.OracleWorld().setup(*env, *rob)
        lt = d_lt.cpu().numpy()
        done_n, cq, mism, dt = 0, 0, 0, 0.0
        O.set_libm(1)  # timed with the host's C library, as the reference runs
        while dt < args.cpu_seconds and done_n < n:
            sl = slice(done_n, min(n, done_n + 16384))
            t0 = time.perf_counter()
            o = ow.touch_batch(starts[sl], ends[sl], vd, 0, nthreads=_sample_threads())
            dt += time.perf_counter() - t0
            cq += int(o["queries"].sum())
            done_n = sl.stop
        O.set_libm(0)  # parity with the correctly rounded restatement
        par = min(done_n, 65536)
        o = ow.touch_batch(starts[:par], ends[:par], vd, 0, nthreads=_sample_threads())
        mism = int((o["done"].astype(bool) != done[:par]).sum() + (o["queries"] != steps[:par]).sum()
                   + (o["last_tau"] != lt[:par]).sum())
        extra["mismatches"] = mism
        extra["parity_sample"] = f"the first {par} segments"
        O.set_libm(1)

        def one(k):
            sl = slice((k % 512) * 32, (k % 512 + 1) * 32)
            return int(ow.touch_batch(starts[sl], ends[sl], vd, 0, nthreads=1)["queries"].sum())
        st, k = _single_thread(one, args.single_seconds)
        O.set_libm(0)
        extra["cpu_baseline"] = _cpu_baseline(cq / dt, "queries/s",
                                              f"first {done_n} of the {n} segments ({cq} queries, {dt:.2f} s; the "
                                              "host C library's sin/acos/atan2, as the reference runs)", st,
                                              f"{k} chunks of 32 of the same segments, one thread")
        fq, nbv, ntri = oracle_flops_per_query(ow, starts, ends, 1024)
        roof = _compute_roofline(fq, nbv, ntri, q, kern_ms, "k_segments<transit>",
                                 "the states the first 1024 segments evaluate")
        if n == 4 * 1024 * 1024:
            # both bounds, as the shell scene's trees exceed the L2: the
            # compute fraction above and the measured HBM bytes per launch
            traffic = int((TRAFFIC_UNEVEN_FETCH_KIB * FETCH_SCALE + TRAFFIC_UNEVEN_WRITE_KIB * WRITE_SCALE) * 1024)
            roof["traffic"] = traffic
            roof["hbm_gbs"] = round(traffic / (kern_ms * 1e-3) / 1e9, 1)
            roof["hbm_frac"] = round(traffic / (kern_ms * 1e-3) / 1e9 / PEAK_HBM_GBS, 5)
        extra["roofline"] = roof
    line = _line(f"uneven segments: shell env {len(env[1])} tris / bent rod {len(rob[1])} tris, {n} segments, "
                 f"lognormal(5, 1.2) step counts in [8, 4096] at step 1/1024", q * args.steps / wall, wall, args, extra)
    line["config"].update({"segments": n, "queries_per_step": q, "free_segments": int(done.sum())})
    return line


def run_visibility(args, torch, pyosr, W):
    """Bit-packed visibility matrix of 4,096 free milestones, shell scene."""
    env, rob = W.shell_scene(0)
    uw = _world(args, pyosr, env, rob)
    n = args.segments or 4096
    ms = _free_states(uw, W, n, 6000)
    stream = torch.cuda.Stream()
    torch.cuda.set_stream(stream)
    d_q = torch.from_numpy(ms).cuda()
    words = (n * n + 31) // 32
    d_bits = torch.zeros(words, dtype=torch.int32, device="cuda")
    npairs = n * (n - 1) // 2
    d_steps = torch.zeros(npairs, dtype=torch.int32, device="cuda")
    wall, kern_ms = _Timer(torch, stream).run(
        lambda: uw.visibility_bits_device(d_q.data_ptr(), n, VD, d_bits.data_ptr(), d_steps.data_ptr(),
                                          stream.cuda_stream), args.steps, args.warmup)
    uw.check_device()
    steps = d_steps.cpu().numpy().astype(np.int64)
    q = int(steps.sum())
    bits = d_bits.cpu().numpy().view(np.uint32)
    extra = {"mismatches": None, "cpu_baseline": None, "kernel": "k_visibility<matrix>", "kernel_ms": round(kern_ms, 4)}
    if not args.no_cpu:
        from oracle import oracle as O
        O.set_libm(False)
        ow = O.OracleWorld().setup(*env, *rob)
        # 64 x 64 diagonal blocks of the matrix, in order, until the time
        # budget is spent
        flat = np.unpackbits(bits.view(np.uint8), bitorder="little")[: n * n].reshape(n, n)
        blk, k, dt, mism, qs_cpu = 64, 0, 0.0, 0, 0
        row, col = np.triu_indices(blk, 1)
        O.set_libm(1)  # timed with the host's C library, as the reference runs
        while dt < args.cpu_seconds and (k + 1) * blk <= n:
            lo = k * blk
            t0 = time.perf_counter()
            c = ow.visibility_matrix(ms[lo:lo + blk], True, VD, nthreads=_sample_threads())
            dt += time.perf_counter() - t0
            r, cc = row + lo, col + lo
            qs_cpu += int(steps[r * (2 * n - r - 1) // 2 + (cc - r - 1)].sum())
            k += 1
        O.set_libm(0)  # parity with the correctly rounded restatement
        for kb in range(min(k, 8)):
            lo = kb * blk
            c = ow.visibility_matrix(ms[lo:lo + blk], True, VD, nthreads=_sample_threads())
            mism += int((flat[lo:lo + blk, lo:lo + blk].astype(np.int8) != c).sum())
        extra["mismatches"] = mism
        extra["parity_sample"] = f"the first {min(k, 8)} diagonal 64 x 64 blocks"
        O.set_libm(1)
        r16, c16 = np.triu_indices(16, 1)

        def one(kk):
            lo = (kk % (n // 16)) * 16
            ow.visibility_matrix(ms[lo:lo + 16], True, VD, nthreads=1)
            rr, cc = r16 + lo, c16 + lo
            return int(steps[rr * (2 * n - rr - 1) // 2 + (cc - rr - 1)].sum())
        st, kk = _single_thread(one, args.single_seconds)
        O.set_libm(0)
        extra["cpu_baseline"] = _cpu_baseline(qs_cpu / dt, "queries/s",
                                              f"{k} diagonal 64 x 64 blocks of the {n} x {n} matrix "
                                              f"({qs_cpu} queries, {dt:.2f} s)", st,
                                              f"{kk} diagonal 16 x 16 blocks, one thread")
        # states the pairs of the first rows evaluate (all at verify step
        # 1/1024 unless a pair is closer than 2/1024)
        fi, ti = np.triu_indices(n, 1)
        fq, nbv, ntri = oracle_flops_per_query(ow, ms[fi[:512]], ms[ti[:512]], 512)
        extra["roofline"] = _compute_roofline(fq, nbv, ntri, q, kern_ms, "k_visibility<matrix>",
                                              "the states the first 512 pairs evaluate")
    line = _line(f"visibility matrix: shell env {len(env[1])} tris / bent rod {len(rob[1])} tris, {n} milestones, "
                 f"{npairs} pairs at step 1/1024", q * args.steps / wall, wall, args, extra)
    line["config"].update({"milestones": n, "pairs": npairs, "queries_per_step": q,
                           "visible_pairs": int(np.unpackbits(bits.view(np.uint8)).sum()) // 2})
    return line


def run_batch(args, torch, pyosr, W):
    """One persistent launch of uneven segments (lognormal step counts) plus
    single states, shell scene; the launch shape is tuned on this batch first.
    Also times the same work as two separate launches."""
    env, rob = W.shell_scene(0)
    uw = _world(args, pyosr, env, rob)
    n = args.segments or 4 * 1024 * 1024
    m = args.batch_states
    vd = VD
    starts = _free_states(uw, W, n, 5000)
    steps_req = W.lognormal_steps(n, seed=5001)
    ends = W.segment_ends(starts, seed=5002, lengths=steps_req * vd)
    qs = W.random_states(m, seed=5003)
    stream = torch.cuda.Stream()
    torch.cuda.set_stream(stream)
    dev = torch.device("cuda")
    d_f, d_t, d_q = (torch.from_numpy(x).to(dev) for x in (starts, ends, qs))
    d_done = torch.zeros(n, dtype=torch.uint8, device=dev)
    d_lt = torch.zeros(n, dtype=torch.float64, device=dev)
    d_ta = torch.zeros(n, dtype=torch.float64, device=dev)
    d_st = torch.zeros(n, dtype=torch.int32, device=dev)
    d_v = torch.zeros(m, dtype=torch.uint8, device=dev)
    outs = (d_done.data_ptr(), d_lt.data_ptr(), d_ta.data_ptr(), d_st.data_ptr())
    cfg, tuned_ms = uw.batch_tune(d_f.data_ptr(), d_t.data_ptr(), n, vd, *outs, d_q.data_ptr(), m, d_v.data_ptr(),
                                  repeats=1)
    timer = _Timer(torch, stream)
    wall, kern_ms = timer.run(
        lambda: uw.batch_device(d_f.data_ptr(), d_t.data_ptr(), n, vd, *outs, d_q.data_ptr(), m, d_v.data_ptr(),
                                stream=stream.cuda_stream), args.steps, args.warmup)
    uw.check_device()
    steps = d_st.cpu().numpy().astype(np.int64)
    done = d_done.cpu().numpy().astype(bool)
    valid = d_v.cpu().numpy().astype(bool)
    q = int(steps.sum()) + m
    # the same work as two launches
    d_done2 = torch.zeros_like(d_done)
    d_v2 = torch.zeros_like(d_v)
    d_st2 = torch.zeros_like(d_st)

    def separate():
        uw.transit_device(d_f.data_ptr(), d_t.data_ptr(), n, vd, d_done2.data_ptr(), d_lt.data_ptr(),
                          d_ta.data_ptr(), d_st2.data_ptr(), stream=stream.cuda_stream)
        uw.valid_states_device(d_q.data_ptr(), m, d_v2.data_ptr(), stream.cuda_stream)
    _, sep_ms = timer.run(separate, max(1, args.steps // 2), 1)
    uw.check_device()
    same = bool(torch.equal(d_done, d_done2) and torch.equal(d_v, d_v2) and torch.equal(d_st, d_st2))
    extra = {"mismatches": None, "cpu_baseline": None, "kernel": "k_batch", "kernel_ms": round(kern_ms, 4),
             "tuned_config": cfg, "separate_launches_ms": round(sep_ms, 4), "same_as_separate": same}
    if not args.no_cpu:
        from oracle import oracle as O
        O.set_libm(False)
        ow = O.OracleWorld().setup(*env, *rob)
        # chunks of 16,384 segments and 4,096 states (the batch's proportions)
        # until the time budget is spent
        ks = kq = 0
        cq, mism, dt = 0, 0, 0.0
        while dt < args.cpu_seconds and ks < n:
            ss = slice(ks, min(n, ks + 16384))
            sq = slice(kq, min(m, kq + 4096))
            t0 = time.perf_counter()
            o = ow.touch_batch(starts[ss], ends[ss], vd, 0, nthreads=_sample_threads())
            c = ow.valid_states(qs[sq], nthreads=_sample_threads())
            dt += time.perf_counter() - t0
            mism += int((o["done"].astype(bool) != done[ss]).sum() + (o["queries"] != steps[ss]).sum()
                        + (c.astype(bool) != valid[sq]).sum())
            cq += int(o["queries"].sum()) + (sq.stop - sq.start)
            ks, kq = ss.stop, sq.stop
        extra["mismatches"] = mism

        def one(k):
            ss, sq = slice((k % 256) * 32, (k % 256 + 1) * 32), slice((k % 256) * 8, (k % 256 + 1) * 8)
            o = ow.touch_batch(starts[ss], ends[ss], vd, 0, nthreads=1)
            ow.valid_states(qs[sq], nthreads=1)
            return int(o["queries"].sum()) + 8
        st, k1 = _single_thread(one, args.single_seconds)
        extra["cpu_baseline"] = _cpu_baseline(cq / dt, "queries/s",
                                              f"first {ks} segments and {kq} states ({cq} queries, {dt:.2f} s)", st,
                                              f"{k1} chunks of 32 segments and 8 states, one thread")
    line = _line(f"mixed batch, one persistent launch: shell env {len(env[1])} tris / bent rod {len(rob[1])} tris, "
                 f"{n} segments with lognormal(5, 1.2) step counts in [8, 4096] at step 1/1024 plus {m} single "
                 f"states", q * args.steps / wall, wall, args, extra)
    line["config"].update({"segments": n, "states": m, "queries_per_step": q, "free_segments": int(done.sum())})
    return line


ALPHA_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tests", "golden", "ref")
# the alpha environment's and solution path's bounding boxes, widened by 10
ALPHA_LO = np.array([-11.892796, -13.170522, -15.239033])
ALPHA_HI = np.array([68.73492, 75.37971186, 56.18796844])


def run_ompl(args, torch, pyosr, W):
    """Key-configuration screening (src/GP/pipeline/solve2.py:219): OMPL's
    discrete motion validator on every pair of np.tril_indices(N) of N free
    key configurations of the alpha puzzle, at the pipeline's default
    resolution 1e-4."""
    from puzzletunneldiscovery_amd import pyse3ompl
    drv = pyse3ompl.OmplDriver()
    drv.set_model_file(pyse3ompl.MODEL_PART_ENV, os.path.join(ALPHA_DIR, "alpha_env-1.2.org.obj"))
    drv.set_model_file(pyse3ompl.MODEL_PART_ROB, os.path.join(ALPHA_DIR, "alpha-1.2.org.obj"))
    drv.set_bb(ALPHA_LO, ALPHA_HI)
    drv.set_cdres(1e-4)
    n = args.segments or 1024
    path = np.loadtxt(os.path.join(ALPHA_DIR, "alpha-1.2.org.path"))
    rng = np.random.default_rng(7000)
    keys = [path]
    have = len(path)
    while have < n:
        q = np.zeros((4 * n, 7))
        q[:, :3] = rng.uniform(ALPHA_LO + 10, ALPHA_HI - 10, (4 * n, 3))
        r = rng.standard_normal((4 * n, 4))
        q[:, 3:] = r / np.linalg.norm(r, axis=1, keepdims=True)
        q = q[drv.validate_states(q) == 1]
        keys.append(q)
        have += len(q)
    keys = np.ascontiguousarray(np.vstack(keys)[:n])
    fi, ti = np.tril_indices(n)
    npairs = len(fi)
    stream = torch.cuda.Stream()
    torch.cuda.set_stream(stream)
    d0 = torch.from_numpy(np.ascontiguousarray(keys[fi])).cuda()
    d1 = torch.from_numpy(np.ascontiguousarray(keys[ti])).cuda()
    d_out = torch.zeros(npairs, dtype=torch.int32, device="cuda")
    d_sam = torch.zeros(npairs, dtype=torch.int32, device="cuda")
    wall, kern_ms = _Timer(torch, stream).run(
        lambda: drv.validate_motion_pairs_device(d0.data_ptr(), d1.data_ptr(), npairs, d_out.data_ptr(),
                                                 d_sam.data_ptr(), stream.cuda_stream), args.steps, args.warmup)
    drv.unit_world.check_device()
    out = d_out.cpu().numpy()
    sam = d_sam.cpu().numpy().astype(np.int64)
    extra = {"mismatches": None, "cpu_baseline": None, "kernel": "k_ompl_motions", "kernel_ms": round(kern_ms, 4)}
    if not args.no_cpu:
        from oracle import oracle as O
        O.set_libm(False)
        ow = O.OracleWorld()
        ow.load_obj(0, os.path.join(ALPHA_DIR, "alpha_env-1.2.org.obj"))
        ow.load_obj(1, os.path.join(ALPHA_DIR, "alpha-1.2.org.obj"))
        ow.scale_to_unit()
        ow.angle_model()
        perm = np.random.default_rng(7001).permutation(npairs)
        done, mism, dt = 0, 0, 0.0
        while dt < args.cpu_seconds and done < npairs:
            idx = perm[done:done + 2048]
            t0 = time.perf_counter()
            c, cs = ow.ompl_validate_motions(keys[fi[idx]], keys[ti[idx]], drv.longest_valid_segment,
                                             nthreads=_sample_threads())
            dt += time.perf_counter() - t0
            mism += int((c != out[idx]).sum() + (cs != sam[idx]).sum())
            done += len(idx)
        extra["mismatches"] = mism

        def one(k):
            idx = perm[(k % 64) * 64:(k % 64 + 1) * 64]
            ow.ompl_validate_motions(keys[fi[idx]], keys[ti[idx]], drv.longest_valid_segment, nthreads=1)
            return len(idx)
        st, k1 = _single_thread(one, args.single_seconds)
        extra["cpu_baseline"] = _cpu_baseline(done / dt, "motions/s",
                                              f"{done} random pairs of the {npairs} ({dt:.2f} s)", st,
                                              f"{k1} chunks of 64 of the same pairs, one thread")
    line = _line(f"OMPL key screening: alpha puzzle (reference meshes), {n} free keys, {npairs} motion pairs "
                 f"(np.tril_indices), resolution 1e-4", npairs * args.steps / wall, wall, args, extra)
    line["unit"] = "motions/s"
    line["metric"] = "OMPL discrete motion checks per second"
    line["config"].update({"keys": n, "pairs": npairs, "valid_pairs": int(out.sum()),
                           "samples_if_valid": int(sam.sum()),
                           "longest_valid_segment": drv.longest_valid_segment})
    return line


def run_clearance(args, torch, pyosr, W):
    """Key-configuration clearance (src/GP/pipeline/keyconf.py:276-317) on
    the alpha puzzle at recommended_cres = scene_scale x 0.001: 128 touch
    samples per key (envconfig.py:141) for 1,024 keys, a tenth of the
    pipeline's 1,024 surface pairs x oversampling ratio 10.  At that step
    (1.5e-5 of the unit world) a touch query averages about 28,000 steps,
    so the full stage is about 4e10 queries.  Keys are free states on the
    reference's solution path (between consecutive path states); the timed
    step is the device launch of all 131,072 touch queries
    (transitStateToWithContact)."""
    from puzzletunneldiscovery_amd import clearance as C
    from puzzletunneldiscovery_amd import pyse3ompl
    uw = pyosr.UnitWorld()
    uw.loadModelFromFile(os.path.join(ALPHA_DIR, "alpha_env-1.2.org.obj"))
    uw.loadRobotFromFile(os.path.join(ALPHA_DIR, "alpha-1.2.org.obj"))
    uw.scaleToUnit()
    uw.angleModel(0.0, 0.0)
    uw.recommended_cres = uw.scene_scale * 0.001
    vd = uw.recommended_cres
    drv = pyse3ompl.OmplDriver.from_unit_world(uw)
    nkey = args.segments or 1024
    nsample = 128
    path = np.loadtxt(os.path.join(ALPHA_DIR, "alpha-1.2.org.path"))
    rng = np.random.default_rng(7100)
    keys, have = [], 0
    while have < nkey:
        # a normalised lerp between consecutive path states, kept where free
        i = rng.integers(0, len(path) - 1, 2 * nkey)
        t = rng.uniform(0, 1, (2 * nkey, 1))
        a, b = path[i], path[i + 1].copy()
        b[:, 3:] *= np.where((a[:, 3:] * b[:, 3:]).sum(axis=1, keepdims=True) < 0, -1.0, 1.0)
        q = a * (1 - t) + b * t
        q[:, 3:] /= np.linalg.norm(q[:, 3:], axis=1, keepdims=True)
        q = q[drv.validate_states(q) == 1]
        keys.append(q)
        have += len(q)
    keys = np.ascontiguousarray(np.vstack(keys)[:nkey])
    print(f"clearance: {nkey} free keys on the solution path", file=sys.stderr, flush=True)
    tr, aa = C.touch_directions(nkey * nsample, np.random.RandomState(7101))
    unit = uw.translate_ompl_to_unit(keys)
    starts = np.repeat(unit, nsample, axis=0)
    ends = pyosr.apply_batch(starts, tr, aa)
    n = len(starts)
    stream = torch.cuda.Stream()
    torch.cuda.set_stream(stream)
    d_f = torch.from_numpy(starts).cuda()
    d_t = torch.from_numpy(ends).cuda()
    d_done = torch.zeros(n, dtype=torch.uint8, device="cuda")
    d_lt = torch.zeros(n, dtype=torch.float64, device="cuda")
    d_ta = torch.zeros(n, dtype=torch.float64, device="cuda")
    d_st = torch.zeros(n, dtype=torch.int32, device="cuda")
    wall, kern_ms = _Timer(torch, stream).run(
        lambda: uw.transit_device(d_f.data_ptr(), d_t.data_ptr(), n, vd, d_done.data_ptr(), d_lt.data_ptr(),
                                  d_ta.data_ptr(), d_st.data_ptr(), stream=stream.cuda_stream),
        args.steps, args.warmup)
    uw.check_device()
    print(f"clearance: timed {args.steps} launches, {kern_ms:.1f} ms each", file=sys.stderr, flush=True)
    steps = d_st.cpu().numpy().astype(np.int64)
    done = d_done.cpu().numpy().astype(bool)
    lt = d_lt.cpu().numpy()
    q = int(steps.sum())
    # the whole stage through the host API (host buffers, PCIe both ways,
    # distances on the host), for comparison: not the timed step
    t0 = time.perf_counter()
    osc = C.touch_distances(uw, unit, tr, aa)
    stage_s = time.perf_counter() - t0
    extra = {"mismatches": None, "cpu_baseline": None, "kernel": "k_segments<transit> (persistent)",
             "kernel_ms": round(kern_ms, 4), "stage_ms_host_api": round(stage_s * 1e3, 2)}
    if not args.no_cpu:
        from oracle import oracle as O
        O.set_libm(False)
        ow = O.OracleWorld()
        ow.load_obj(0, os.path.join(ALPHA_DIR, "alpha_env-1.2.org.obj"))
        ow.load_obj(1, os.path.join(ALPHA_DIR, "alpha-1.2.org.obj"))
        ow.scale_to_unit()
        ow.angle_model()
        perm = np.random.default_rng(7102).permutation(n)
        done_n, cq, mism, dt = 0, 0, 0, 0.0
        while dt < args.cpu_seconds and done_n < n:
            idx = perm[done_n:done_n + 512]
            t0 = time.perf_counter()
            o = ow.touch_batch(starts[idx], ends[idx], vd, 0, nthreads=_sample_threads())
            dt += time.perf_counter() - t0
            mism += int((o["done"].astype(bool) != done[idx]).sum() + (o["queries"] != steps[idx]).sum()
                        + (o["last_tau"] != lt[idx]).sum())
            cq += int(o["queries"].sum())
            done_n += len(idx)
        # DISTANCE_BATCH entries of the first two keys through the oracle's touch queries
        for k in range(2):
            free = [ow.transit(unit[k], ends[k * nsample + j], vd)[0] for j in range(16)]
            mism += int((np.array([O.distance(unit[k], f) for f in free]) != osc[k, :16]).sum())
        print(f"clearance: oracle sample done ({dt:.1f} s)", file=sys.stderr, flush=True)
        extra["mismatches"] = mism

        def one(k):
            idx = perm[(k % 64) * 8:(k % 64 + 1) * 8]
            return int(ow.touch_batch(starts[idx], ends[idx], vd, 0, nthreads=1)["queries"].sum())
        st, k1 = _single_thread(one, args.single_seconds)
        extra["cpu_baseline"] = _cpu_baseline(cq / dt, "queries/s",
                                              f"{done_n} random touch queries of the {n} ({cq} queries, {dt:.2f} s)",
                                              st, f"{k1} chunks of 8 of the same queries, one thread")
        fq, nbv, ntri = oracle_flops_per_query(ow, starts[perm[:32]], ends[perm[:32]], 32, vd)
        extra["roofline"] = _compute_roofline(fq, nbv, ntri, q, kern_ms, "k_segments<transit>",
                                              "the states of 32 random touch queries")
    line = _line(f"key clearance: alpha puzzle (reference meshes), {nkey} keys x {nsample} touch samples at "
                 f"recommended_cres {vd:.6g}", q * args.steps / wall, wall, args, extra)
    line["config"].update({"keys": nkey, "samples_per_key": nsample, "segments": n, "queries_per_step": q,
                           "contact_fraction": round(1.0 - float(done.mean()), 4),
                           "keys_per_s": round(nkey / (kern_ms * 1e-3), 1)})
    return line


def _screen_segments(args, torch, W, uw, net, sq, sy, timer, stream):
    """Segments from exact-verified free states, lognormal(5, 1.2) step counts
    at 1/1024: isValidTransition on every segment, against the product's
    prescreen (surrogate.screen_segments: P(free) at 8 points along each
    segment in the fused inference kernel, then isValidTransition on the
    selected segments, read from the device-side list)."""
    from puzzletunneldiscovery_amd import surrogate as S
    n_seg = args.screen_segments
    free_idx = torch.nonzero(sy.bool()).squeeze(1)[:n_seg]
    starts = sq.index_select(0, free_idx).contiguous()
    n_seg = starts.shape[0]
    steps = W.lognormal_steps(n_seg, seed=7003)
    ends = torch.from_numpy(W.segment_ends(starts.cpu().numpy(), seed=7004, lengths=steps * VD)).cuda()
    sp = stream.cuda_stream
    ok = torch.zeros(n_seg, dtype=torch.uint8, device="cuda")
    nsteps = torch.zeros(n_seg, dtype=torch.int32, device="cuda")
    _, plain_ms = timer.run(lambda: uw.valid_transitions_device(starts.data_ptr(), ends.data_ptr(), n_seg, VD,
                                                                ok.data_ptr(), nsteps.data_ptr(), sp),
                            args.steps, args.warmup)
    free_plain = int(ok.sum().item())
    queries = int(nsteps.to(torch.int64).sum().item())
    state = {}

    def screened():
        state["r"] = S.screen_segments(uw, net, starts, ends, VD, args.threshold, points=8, stream=sp,
                                       out=state.get("r"))

    _, scr_ms = timer.run(screened, args.steps, args.warmup)
    uw.check_device()
    r = state["r"]
    truth = ok.bool()
    sel = S.selected_mask(r, n_seg)
    free_scr = int(r["free"].sum().item())
    recall, precision = S.recall_precision(r, truth)
    return {
        "segments": n_seg, "points_per_segment": 8, "exact_queries": queries, "free_true": free_plain,
        "accepted": int(sel.sum().item()), "free_verified_screened": free_scr,
        "recall": round(recall, 4), "precision": round(precision, 4),
        "exact_ms": round(plain_ms, 3), "screen_total_ms": round(scr_ms, 3),
        "plain_free_per_s": round(free_plain / (plain_ms * 1e-3), 1),
        "screened_free_per_s": round(free_scr / (scr_ms * 1e-3), 1),
        "speedup": round((free_scr / scr_ms) / max(free_plain / plain_ms, 1e-9), 3),
        "mismatches": int((r["free"].bool() != (sel & truth)).sum().item()),
    }


def run_surrogate(args, torch, pyosr, W):
    """Learned prescreen on the shell scene: exact labels, BF16-MFMA training,
    then free samples per second with and without the prescreen (every
    accepted sample exact-verified)."""
    from puzzletunneldiscovery_amd import surrogate as S
    from puzzletunneldiscovery_amd.surrogate import Surrogate
    env, rob = W.shell_scene(0)
    uw = _world(args, pyosr, env, rob)
    n_train = 16 * 1024 * 1024
    n_screen = args.segments or 64 * 1024 * 1024
    stream = torch.cuda.Stream()
    torch.cuda.set_stream(stream)
    sp = stream.cuda_stream
    timer = _Timer(torch, stream)

    def exact(d_q, n, d_out):
        uw.valid_states_device(d_q.data_ptr(), n, d_out.data_ptr(), sp)

    # exact labels
    tq = torch.from_numpy(W.random_states(n_train, seed=7000)).cuda()
    ty = torch.zeros(n_train, dtype=torch.uint8, device="cuda")
    _, label_ms = timer.run(lambda: exact(tq, n_train, ty), 1, 0)
    uw.check_device()
    # untimed warm-up: one epoch of a throwaway network, so that the timed
    # fit does not include loading the kernels or the first allocation of
    # the shuffle's buffers
    warm = Surrogate(max_batch=16384, seed=1)
    warm.fit(tq, ty, epochs=1, batch=16384, lr=2e-3)
    torch.cuda.synchronize()
    del warm
    net = Surrogate(max_batch=16384, seed=7001)
    t0 = time.perf_counter()
    losses = net.fit(tq, ty, epochs=args.train_epochs, batch=16384, lr=2e-3)
    torch.cuda.synchronize()
    train_s = time.perf_counter() - t0
    steps_done = args.train_epochs * (n_train // 16384)
    del tq, ty
    # screen: plain exact kernel vs prescreen + exact re-verification
    sq = torch.from_numpy(W.random_states(n_screen, seed=7002)).cuda()
    sy = torch.zeros(n_screen, dtype=torch.uint8, device="cuda")
    pf = torch.zeros(n_screen, dtype=torch.float32, device="cuda")
    wall_plain, plain_ms = timer.run(lambda: exact(sq, n_screen, sy), args.steps, args.warmup)
    free_plain = int(sy.sum().item())
    wall_mlp, mlp_ms = timer.run(lambda: net.forward_device(sq.data_ptr(), n_screen, pf.data_ptr(), stream=sp),
                                 args.steps, args.warmup)
    thr = args.threshold
    state = {}

    def screened():
        # the product's prescreen: selection fused into the inference kernel,
        # exact re-verification of the device-side list, no host round trip
        state["r"] = S.screen_states(uw, net, sq, thr, stream=sp, out=state.get("r"))

    wall_scr, scr_ms = timer.run(screened, args.steps, args.warmup)
    uw.check_device()
    r = state["r"]
    free_scr = int(r["free"].sum().item())
    seg = _screen_segments(args, torch, W, uw, net, sq, sy, timer, stream) if args.screen_segments else None
    truth = sy.bool()
    sel = S.selected_mask(r, n_screen)
    recall, precision = S.recall_precision(r, truth)
    flops_sample = 2 * (12 * 256 + 3 * 256 * 256 + 256 * 2)
    # one training row: forward, gradient back to layer 0's pre-activation
    # (through the head and W3..W1) and the weight gradients, over the used
    # (unpadded) layer sizes
    flops_train_row = 2 * flops_sample + 2 * (256 * 2 + 3 * 256 * 256)
    rate_plain = free_plain / (plain_ms * 1e-3)
    rate_scr = free_scr / (scr_ms * 1e-3)
    extra = {
        "mismatches": int((r["free"].bool() != (sel & truth)).sum().item()),
        "cpu_baseline": None,
        "surrogate": {
            "labels": n_train, "label_ms": round(label_ms, 2), "train_epochs": args.train_epochs,
            "train_steps": steps_done, "train_s": round(train_s, 3),
            "train_step_us": round(train_s / max(steps_done, 1) * 1e6, 1),
            "train_tflops": round(steps_done * 16384 * flops_train_row / max(train_s, 1e-9) / 1e12, 1),
            "train_mfma_frac_bf16": round(steps_done * 16384 * flops_train_row / max(train_s, 1e-9) / 2.5e15, 4),
            "loss_first_epoch": round(losses[0], 4),
            "loss_last_epoch": round(losses[-1], 4), "threshold": thr, "recall": round(recall, 4),
            "precision": round(precision, 4), "accepted": int(sel.sum().item()), "free_true": free_plain,
            "free_verified_screened": free_scr,
            "plain_free_per_s": round(rate_plain, 1), "screened_free_per_s": round(rate_scr, 1),
            "speedup": round(rate_scr / max(rate_plain, 1e-9), 3),
            "exact_ms": round(plain_ms, 3), "mlp_ms": round(mlp_ms, 3), "screen_total_ms": round(scr_ms, 3),
            "mlp_tflops": round(n_screen * flops_sample / (mlp_ms * 1e-3) / 1e12, 2),
            "mlp_mfma_frac_bf16": round(n_screen * flops_sample / (mlp_ms * 1e-3) / 2.5e15, 4),
        },
        "surrogate_segments": seg,
    }
    if not args.no_cpu:
        # the oracle's exact check of the same random samples: verified free
        # samples per second on the host, and parity with the GPU's labels
        from oracle import oracle as O
        O.set_libm(False)
        ow = O.OracleWorld().setup(*env, *rob)
        hq = sq[: 4 * 1024 * 1024].cpu().numpy()
        hy = sy[: 4 * 1024 * 1024].cpu().numpy().astype(bool)
        m, dt, fr, mism = 0, 0.0, 0, 0
        while dt < args.cpu_seconds and m < len(hq):
            sl = slice(m, min(len(hq), m + 100_000))
            t0 = time.perf_counter()
            c = ow.valid_states(hq[sl], nthreads=_sample_threads()).astype(bool)
            dt += time.perf_counter() - t0
            fr += int(c.sum())
            mism += int((c != hy[sl]).sum())
            m = sl.stop

        def one(k):
            sl = slice((k % 64) * 1000, (k % 64 + 1) * 1000)
            return int(ow.valid_states(hq[sl], nthreads=1).sum())
        st, k1 = _single_thread(one, args.single_seconds)
        extra["mismatches"] += mism
        extra["cpu_baseline"] = _cpu_baseline(fr / dt, "verified free samples/s",
                                              f"exact check of the first {m} screened samples ({dt:.2f} s)", st,
                                              f"{k1} chunks of 1,000 of the same samples, one thread")
    line = _line(f"learned prescreen: shell env {len(env[1])} tris / bent rod {len(rob[1])} tris, "
                 f"{n_train} exact labels, {n_screen} screened samples", rate_scr, wall_scr, args, extra)
    line["unit"] = "verified free samples/s"
    line["value"] = round(rate_scr, 1)
    line["dtype"] = "bf16 MLP (fp32 accumulate) + f64 exact re-verification"
    return line


def run_touch(args, torch, pyosr, W):
    """The headline: touch-configuration search over 65,536 segments of the
    torus scene (one step = one launch)."""
    env, rob = W.torus_scene(0)
    uw = _world(args, pyosr, env, rob)
    starts, ends = build_workload(uw, W, args.segments or N_SEG)
    n = starts.shape[0]
    dev = torch.device("cuda:0")
    d_from = torch.from_numpy(starts).to(dev)
    d_to = torch.from_numpy(ends).to(dev)
    d_done = torch.zeros(n, dtype=torch.uint8, device=dev)
    d_lt = torch.zeros(n, dtype=torch.float64, device=dev)
    d_t = torch.zeros(n, dtype=torch.float64, device=dev)
    d_lo = torch.zeros(n, dtype=torch.float64, device=dev)
    d_hi = torch.zeros(n, dtype=torch.float64, device=dev)
    d_steps = torch.zeros(n, dtype=torch.int32, device=dev)
    # a dedicated (default-priority) stream so the HIP events bracket exactly
    # the engine's kernels
    stream = torch.cuda.Stream()
    torch.cuda.set_stream(stream)
    sp = stream.cuda_stream

    def step():
        uw.touch_search_device(d_from.data_ptr(), d_to.data_ptr(), n, VD, LEVELS, d_done.data_ptr(),
                               d_lt.data_ptr(), d_t.data_ptr(), d_lo.data_ptr(), d_hi.data_ptr(),
                               d_steps.data_ptr(), sp)

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    uw.check_device()
    ev0 = torch.cuda.Event(enable_timing=True)
    ev1 = torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ev0.record(stream)
    for _ in range(args.steps):
        step()
    ev1.record(stream)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    uw.check_device()
    wall = t1 - t0
    kern_ms = ev0.elapsed_time(ev1) / args.steps
    res = dict(done=d_done.cpu().numpy().astype(bool), lo=d_lo.cpu().numpy(), hi=d_hi.cpu().numpy(),
               steps=d_steps.cpu().numpy())
    queries = int(res["steps"].astype(np.int64).sum())
    value = queries * args.steps / wall

    # the same search through the host-buffer entry point: states copied in
    # and results copied out over PCIe on every call (not the headline)
    uw.touch_search(starts, ends, VD, LEVELS)
    reps = 5
    t0 = time.perf_counter()
    for _ in range(reps):
        hres = uw.touch_search(starts, ends, VD, LEVELS)
    host_s = (time.perf_counter() - t0) / reps
    if not (np.array_equal(hres["lo"], res["lo"]) and np.array_equal(hres["hi"], res["hi"])):
        raise RuntimeError("host-buffer touch search differs from the device-buffer one")
    pcie = dict(value=round(queries / host_s, 1), unit="queries/s", ms_per_call=round(host_s * 1e3, 4),
                note="touch_search with host buffers: inputs copied in and results copied out each call")

    cpu = None
    mism = None
    flops_q = None
    if not args.no_cpu:
        from oracle import oracle as O
        O.set_libm(False)
        ow = O.OracleWorld().setup(*env, *rob)
        cpu, mism, n_checked = cpu_baseline(ow, starts, ends, res, args.cpu_seconds, args.single_seconds)
        flops_q, nbv, ntri = oracle_flops_per_query(ow, starts, ends, min(2048, n))
    roof = None
    if flops_q is not None:
        achieved = queries * flops_q / (kern_ms * 1e-3) / 1e12
        traffic = int((TRAFFIC_FETCH_KIB * FETCH_SCALE + TRAFFIC_WRITE_KIB * WRITE_SCALE) * 1024)
        roof = dict(bound="mfma", roof=VALU_ROOF, achieved=round(achieved, 3), peak=PEAK_FP32_TFLOPS, unit="TFLOP/s",
                    frac=round(achieved / PEAK_FP32_TFLOPS, 5), traffic=traffic,
                    note="compute-bound on vector ALUs (no matrix instructions); the compute roof is "
                         "MI355X's fp32 peak; algorithmic flops = oracle node/triangle test counts x "
                         "their straight-line binary64 op counts",
                    flops_per_query=round(flops_q, 1), oracle_node_tests_per_query=round(nbv, 2),
                    oracle_tri_tests_per_query=round(ntri, 3), kernel="k_segments<touch>",
                    kernel_ms=round(kern_ms, 4))
    line = {
        "metric": METRIC,
        "value": round(value, 1),
        "unit": "queries/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(wall / args.steps * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64 (exact leaf tests) + f32 (conservative culling)",
        "data": "synthetic (seeded torus scene, Shoemake states)",
        "config": {"workload": "touch-config search: torus env 2,048 tris / robot 1,024 tris, "
                               f"{n} segments, step 1/1024, {LEVELS} bisection levels",
                   "segments": n, "queries_per_step": queries,
                   "contacts": int((~res["done"]).sum())},
        "mismatches": mism,
        "roofline": roof,
        "cpu_baseline": cpu,
        "pcie_inclusive": pcie,
    }
    return line


SUB_WORKLOADS = (
    # name, function, steps cap, warmup cap: each config timed over a few
    # launches so the whole default run stays within a few minutes
    ("states", "run_states", 20, 3),
    ("uneven", "run_uneven", 3, 1),
    ("visibility", "run_visibility", 2, 1),
    ("surrogate", "run_surrogate", 3, 1),
)
SUB_KEYS = ("value", "unit", "metric", "steps", "warmup", "ms_per_step", "kernel", "kernel_ms", "mismatches", "parity_sample",
            "roofline", "cpu_baseline", "dtype", "config", "surrogate", "surrogate_segments")


def run_all_configs(args, torch, pyosr, W):
    """The other configurations, each measured as its own --workload run
    would, with the steps capped; returned as one object per workload."""
    out = {}
    for name, fn, steps_cap, warm_cap in SUB_WORKLOADS:
        sub = copy.copy(args)
        sub.steps = max(1, min(args.steps, steps_cap))
        sub.warmup = min(args.warmup, warm_cap)
        sub.segments = None
        t0 = time.perf_counter()
        print(f"bench: {name} ...", file=sys.stderr, flush=True)
        r = globals()[fn](sub, torch, pyosr, W)
        torch.cuda.synchronize()
        torch.cuda.empty_cache()
        print(f"bench: {name} done ({time.perf_counter() - t0:.1f} s)", file=sys.stderr, flush=True)
        out[name] = {k: r[k] for k in SUB_KEYS if k in r}
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--segments", type=int, default=None,
                    help="queries per launch (default: the workload's own size)")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--single-seconds", type=float, default=3.0,
                    help="oracle seconds of each single-thread baseline")
    ap.add_argument("--headline-only", action="store_true",
                    help="only the touch-search headline (the default run also measures the states, uneven, "
                         "visibility and surrogate configurations)")
    ap.add_argument("--batch-states", type=int, default=1024 * 1024,
                    help="batch workload: single states in the launch beside the segments")
    ap.add_argument("--workload", choices=["touch", "states", "uneven", "visibility", "surrogate", "ompl", "batch",
                                           "clearance"],
                    default="touch",
                    help="touch (default, the headline): touch-configuration search on the torus scene; "
                         "states: 1M state validity; uneven: 4M lognormal segments on the shell scene; "
                         "visibility: 4,096-milestone matrix on the shell scene; "
                         "surrogate: learned prescreen on the shell scene; "
                         "ompl: OMPL key screening (all pairs of 1,024 keys) on the alpha puzzle; "
                         "batch: 4M uneven segments + 1M states in one persistent launch (shell scene); "
                         "clearance: key clearance (1,024 keys x 128 touch samples) on the alpha puzzle")
    ap.add_argument("--train-epochs", type=int, default=2)
    ap.add_argument("--threshold", type=float, default=0.5)
    ap.add_argument("--screen-segments", type=int, default=8 * 1024 * 1024,
                    help="surrogate workload: segments screened after the samples (0 = none)")
    ap.add_argument("--leaf-size", type=int, default=0, help="triangles per BVH leaf, at most (default: 1)")
    ap.add_argument("--pool-chunk", type=int, default=0, help="states per counter add of the state kernels")
    ap.add_argument("--queue-depth", type=int, default=0, help="segments per counter add of the segment kernels")
    ap.add_argument("--copy-mib", type=int, default=0,
                    help="first copy a buffer of this many MiB three times (a known-size "
                         "kernel that calibrates the FETCH_SIZE / WRITE_SIZE counters)")
    args = ap.parse_args()
    # the oracle's OpenMP threads stay on their cores (read when its library
    # loads, so set before any oracle import)
    os.environ.setdefault("OMP_PROC_BIND", "close")
    os.environ.setdefault("OMP_PLACES", "cores")
    # read the CPU set before OpenMP binds this thread to its first place
    host_cpu()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported "
              "(each query is independent, so several GPUs would each take a contiguous slice of the "
              "batch against their own copy of the scene).", file=sys.stderr)
        sys.exit(2)

    import torch

    from puzzletunneldiscovery_amd import pyosr, workloads as W

    assert torch.cuda.is_available(), "bench.py needs a GPU"
    torch.cuda.set_device(0)
    if args.copy_mib:
        src = torch.ones(args.copy_mib << 18, dtype=torch.float32, device="cuda:0")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
    if args.workload != "touch":
        fn = {"states": run_states, "uneven": run_uneven, "visibility": run_visibility,
              "surrogate": run_surrogate, "ompl": run_ompl, "batch": run_batch,
              "clearance": run_clearance}[args.workload]
        print(json.dumps(fn(args, torch, pyosr, W)), flush=True)
        return
    line = run_touch(args, torch, pyosr, W)
    if not args.headline_only:
        line["workloads"] = run_all_configs(args, torch, pyosr, W)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
