#!/usr/bin/env python3
"""Benchmark on one MI355X.

Workload (default): SSSP on an R-MAT graph, scale 24 (2^24 vertices), edge
factor 16, (a,b,c,d) = (0.57,0.19,0.19,0.05), symmetrized, self-loops and
duplicates removed, u32 weights uniform in [1,64] drawn per unordered pair,
u64 offsets and u32 ids (AoS {dst,w} edge records), source = highest-degree
vertex.  The device edge budget is 40% of the edge+weight bytes, so the
static/on-demand partition streams the other 60% from host memory every
iteration it is needed.

A step is one SSSP from the source over the whole graph (inputs resident in
HBM / pinned host memory before timing starts; the result stays on the
device).  `value` is GTEPS in the Graph500 convention: undirected edges in
the traversed component (sum of reached out-degrees / 2) / seconds per step.

The JSON line also carries:
  roofline     the dominant relax kernel's algorithmic bytes / its summed
               launch time (HIP events on its launch stream), vs 8 TB/s HBM;
  link         streamed on-demand bytes / step time;
  cpu_baseline the oracle's multi-threaded restatement of the reference
               algorithm on the host cores, same graph and source (kind
               "port"), whose result is also compared with the GPU's.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "GTEPS per algo on 1 MI355X; batched BFS/SSSP queries/s + p99 latency"
HBM_PEAK_GBPS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


_gen_ctx = None


def gen_graph(A, **kw):
    """Synthetic graph built on the GPU (asc_generate_device): the host
    generator's CSR bit for bit, in seconds instead of minutes at scale 26."""
    global _gen_ctx
    if _gen_ctx is None:
        _gen_ctx = A.Context()
    return A.generate(device=_gen_ctx, **kw)


def default_cpu_threads():
    """Threads for the CPU baseline: every core this process may run on, up to
    the 16-core share one GPU gets on the benchmark host."""
    try:
        n = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        n = os.cpu_count() or 1
    return max(1, min(16, n))


def oracle_run(label, fn, nt):
    """fn(nthreads) -> (values, stats) of a restated traversal, timed."""
    t0 = time.perf_counter()
    out, st = fn(nt)
    dt = time.perf_counter() - t0
    log(f"cpu {label}: {dt:.2f} s on {nt} threads")
    return out, st, dt


def single_thread_sample(O, label, fn, budget_s=4.0):
    """A bounded single-thread sample of a restated traversal fn(nthreads) on
    the full-size graph: the run stops once n edges are relaxed (n = 2^24,
    2^26, ...) until one sample takes at least a quarter of budget_s or the
    traversal ends first.  Returns the rate in edges relaxed per second
    (the restatement's O(V) set-up included) and what was sampled."""
    n = 1 << 24
    while True:
        with O.edge_limit(n):
            t0 = time.perf_counter()
            _, st = fn(1)
            dt = time.perf_counter() - t0
        whole = st["edges_relaxed"] < n
        if whole or dt >= budget_s / 4 or n >= 1 << 36:
            log(f"cpu {label} single thread: {st['edges_relaxed']} edges, {st['iterations']} iterations, {dt:.2f} s")
            return {"rate": st["edges_relaxed"] / max(dt, 1e-9), "edges": st["edges_relaxed"],
                    "iterations": st["iterations"], "seconds": round(dt, 3), "whole_run": whole}
        n *= 4


def cpu_baseline_block(gteps, nt, st_all, t_all, single, what):
    """cpu_baseline entry: all-core GTEPS of the full restated run, and a
    single-thread figure in edges relaxed per second (with the GTEPS a single
    thread would reach relaxing the same edges at that rate)."""
    rate_all = st_all["edges_relaxed"] / max(t_all, 1e-9)
    span = ("the whole run" if single["whole_run"] else
            f"the first {single['edges']} edges relaxed ({single['iterations']} iterations begun), bounded")
    return {
        "value": round(gteps, 4), "unit": "GTEPS", "cores": nt, "kind": "port",
        "sample": f"{what}, the whole run on {nt} threads ({t_all:.2f} s)",
        "edges_relaxed_per_s": round(rate_all / 1e9, 4),
        "single_thread": {
            "value": round(single["rate"] / 1e9, 4), "unit": "G edges relaxed/s", "cores": 1,
            "gteps_equivalent": round(gteps * single["rate"] / max(rate_all, 1e-9), 4),
            "sample": f"{span}, {single['seconds']} s"},
    }


def gather_baseline(O, csr, ew, vs, V, overload_edges, _unused=None):
    """The reference's 20-thread host gather (GraphMeta.cuh:469-514) timed
    alone: every on-demand vertex with edges, in id order, copied as 8-byte
    {dst, w} records (median of three copies)."""
    deg = np.diff(csr.off)
    lst = np.arange(vs, V, dtype=np.uint32)
    lst = lst[deg[vs:] > 0]
    times = []
    for _ in range(3):
        buf, ms = O.gather_ref(csr, lst, 20, aos=ew)
        times.append(ms)
        del buf
    ms = float(np.median(times))
    nbytes = 8 * int(deg[vs:].sum())
    rate = nbytes / (ms / 1e3)
    log(f"cpu reference gather: {nbytes / 1e9:.2f} GB in {ms:.1f} ms on 20 threads")
    return {"threads": 20, "bytes": nbytes, "ms": round(ms, 2), "GBps": round(rate / 1e9, 2),
            "per_run_ms": round(8 * overload_edges / rate * 1e3, 2),
            "sample": "the whole on-demand region gathered once; per_run_ms = the restated run's gathered edges "
                      "at that rate"}


def algorithmic_bytes(kern, s_e, s_o):
    """Bytes a work-optimal push relax must move: per edge the record and one
    destination value; per frontier vertex an offset pair, its value, a queue
    pop and a push (per list entry of a streamed part: id, u64 prefix, value).
    expand_part_host reads its edge records from mapped host memory (over the
    link, not HBM), so it is reported but never taken as the HBM roofline's
    kernel."""
    out = {
        "expand_local": kern["expand_local"]["edges"] * (s_e + 4) + kern["expand_local"]["vertices"] * (2 * s_o + 12),
        "expand_heavy": kern["expand_heavy"]["edges"] * (s_e + 4),
        "expand_part": kern["expand_part"]["edges"] * (s_e + 4) + kern["expand_part"]["vertices"] * 16,
    }
    if "expand_part_host" in kern:
        out["expand_part_host"] = kern["expand_part_host"]["edges"] * (s_e + 4) + \
            kern["expand_part_host"]["vertices"] * 16
    return out


def serving_bench(A, O, scale, seed, n_queries, rate_factors, cpu_threads, budget_s=60.0, h=None,
                  step_graphs=True):
    """Query serving on a resident weighted R-MAT graph: 70% BFS / 30% SSSP,
    sources uniform over vertices with degree > 0.  Reports one-query-per-
    launch throughput (plain kernels), the server's saturated throughput and
    Poisson-arrival sweeps with p50/p99 submit-to-done latency, and checks a
    sample of served results against the same queries run alone and, for
    four of them, against the oracle; the oracle's restatement answering the
    same kind of queries on the host cores is the CPU baseline."""
    import threading

    t = time.perf_counter()
    if h is None:
        h = gen_graph(A, scale=scale, edge_factor=16, seed=seed, weighted=True, wmin=1, wmax=64)
    V = h.num_vertices
    deg = np.diff(h.offsets())
    cand = np.nonzero(deg > 0)[0]
    log(f"serving graph: R-MAT s{scale} ef16 weighted, V={V} E={h.num_edges} ({time.perf_counter() - t:.1f}s)")
    rng = np.random.default_rng(seed)
    srcs = [int(x) for x in rng.choice(cand, n_queries)]
    kinds = [A.Server.SSSP if x < 0.3 else A.Server.BFS for x in rng.random(n_queries)]
    out = {"workload": f"rmat_s{scale}_ef16_w1-64_resident_70bfs_30sssp", "queries": n_queries}
    # views of the host CSR, which the graph keeps (and frees when closed)
    g_host_arrays = (h.offsets(), h.edge_words())
    with A.Context() as ctx:
        g = ctx.adopt(h, offset_bytes=8)
        # one query per launch
        for k, s in zip(kinds[:2], srcs[:2]):
            (g.sssp if k == A.Server.SSSP else g.bfs)(s, output=False)
        n_plain = min(n_queries, 64)
        ctx.synchronize()
        t0 = time.perf_counter()
        for k, s in zip(kinds[:n_plain], srcs[:n_plain]):
            (g.sssp if k == A.Server.SSSP else g.bfs)(s, output=False)
        ctx.synchronize()
        plain_qps = n_plain / (time.perf_counter() - t0)
        out["plain_qps"] = round(plain_qps, 2)
        with A.Server(g, lanes=64, step_graphs=step_graphs) as srv:
            st = srv.stats()
            out["lanes"] = {"bfs": st["lanes_bfs"], "sssp": st["lanes_sssp"]}
            # exactness on a sample, served in a full mix
            sample = list(range(0, n_queries, max(1, n_queries // 8)))[:8]
            bufs = {i: np.empty(V, np.uint32) for i in sample}
            t0 = time.perf_counter()
            tickets = [srv.submit(k, s, bufs.get(i)) for i, (k, s) in enumerate(zip(kinds, srcs))]
            lat = [srv.wait(tk)["submit_to_done_ms"] for tk in tickets]
            sat = n_queries / (time.perf_counter() - t0)
            exact = all(np.array_equal(bufs[i], (g.sssp if kinds[i] == A.Server.SSSP else g.bfs)(srcs[i])[0])
                        for i in sample)
            # the oracle on two sampled queries of each kind, which also times
            # the CPU answering them (all cores, then one thread, bounded)
            off, ew = g_host_arrays
            csr = O.CSR(off, ew[:, 0], ew[:, 1])
            picks = [i for i in sample if kinds[i] == A.Server.BFS][:2] + \
                    [i for i in sample if kinds[i] == A.Server.SSSP][:2]
            cpu_t = {A.Server.BFS: [], A.Server.SSSP: []}
            oracle_ok = True
            for i in picks:
                fn = (lambda nt, s0=srcs[i]: O.sssp_ref(csr, s0, nthreads=nt)) if kinds[i] == A.Server.SSSP else \
                    (lambda nt, s0=srcs[i]: O.bfs_ref(csr, s0, nthreads=nt))
                ref, _, dt = oracle_run(f"served query {i}", fn, cpu_threads)
                oracle_ok = oracle_ok and bool(np.array_equal(bufs[i], ref))
                cpu_t[kinds[i]].append(dt)
                del ref
            out["saturated_qps"] = round(sat, 2)
            out["saturated_p50_ms"] = round(float(np.percentile(lat, 50)), 3)
            out["saturated_p99_ms"] = round(float(np.percentile(lat, 99)), 3)
            out["speedup_vs_plain"] = round(sat / plain_qps, 2)
            out["exact_sample"] = f"{len(sample)} queries bit-identical to running alone" if exact else "MISMATCH"
            out["oracle_sample"] = (f"{len(picks)} served queries bit-identical to the oracle" if oracle_ok
                                    else "MISMATCH vs oracle")
            tb = float(np.mean(cpu_t[A.Server.BFS])) if cpu_t[A.Server.BFS] else 0.0
            ts = float(np.mean(cpu_t[A.Server.SSSP])) if cpu_t[A.Server.SSSP] else 0.0
            if tb > 0 and ts > 0:
                # single thread: the first BFS query, bounded
                b0 = next(i for i in picks if kinds[i] == A.Server.BFS)
                single = single_thread_sample(O, "served bfs", lambda nt: O.bfs_ref(csr, srcs[b0], nthreads=nt))
                out["cpu_baseline"] = {
                    "value": round(1.0 / (0.7 * tb + 0.3 * ts), 3), "unit": "queries/s", "cores": cpu_threads,
                    "kind": "port",
                    "sample": f"oracle bfs_ref / sssp_ref answering {len(picks)} of the sampled queries one at a time "
                              f"on {cpu_threads} threads (BFS {tb:.2f} s, SSSP {ts:.2f} s each), weighted 70/30",
                    "single_thread": {"value": round(single["rate"] / 1e9, 4), "unit": "G edges relaxed/s",
                                      "cores": 1, "sample": f"the first {single['edges']} edges relaxed by one BFS "
                                                            f"query, {single['seconds']} s"}}
            sweeps = []
            for f in rate_factors:
                rate = f * plain_qps
                n = min(n_queries, max(16, int(rate * min(budget_s / len(rate_factors), 10.0))))
                gaps = np.random.default_rng(seed + int(f * 100)).exponential(1.0 / rate, n)
                tks = []

                def submitter():
                    nxt = time.perf_counter()
                    for i in range(n):
                        nxt += gaps[i]
                        d = nxt - time.perf_counter()
                        if d > 0:
                            time.sleep(d)
                        tks.append(srv.submit(kinds[i], srcs[i]))

                ts = time.perf_counter()
                th = threading.Thread(target=submitter)
                th.start()
                th.join()
                lats = [srv.wait(tk)["submit_to_done_ms"] for tk in tks]
                el = time.perf_counter() - ts
                sweeps.append({"offered_qps": round(rate, 2), "achieved_qps": round(n / el, 2), "queries": n,
                               "p50_ms": round(float(np.percentile(lats, 50)), 3),
                               "p99_ms": round(float(np.percentile(lats, 99)), 3)})
            out["poisson"] = sweeps
        g.close()
    return out


def kernel_bytes(st, name, s_e, s_o, V=0):
    """Algorithmic bytes of one kernel kind over a run.  pr_pull: per in-edge
    the source id and its contribution, per vertex an offset pair, its
    rank and contribution.  bottom_up: per edge read the id and one frontier
    bitmap word, per scanned vertex an offset pair and its value, per launch
    the visited and next bitmaps (V / 8 bytes each).  union_find: per edge
    read the id and two parent probes, per vertex and link pass an offset
    pair and its parent."""
    k = st["kernels"][name]
    if name == "bfs_persistent":   # the whole BFS in one launch: its expansions' bytes, bitmaps once per level
        levels = max(st.get("iterations", 1), 1)
        kk = st["kernels"]
        return (kk["expand_local"]["edges"] * (s_e + 4) + kk["expand_local"]["vertices"] * (2 * s_o + 12) +
                kk["expand_heavy"]["edges"] * (s_e + 4) + kk["bottom_up"]["edges"] * (s_e + 4) +
                kk["bottom_up"]["vertices"] * (2 * s_o + 4) + levels * V // 4)
    if name == "pr_pull":
        return k["edges"] * 8 + k["vertices"] * (2 * s_o + 16)
    if name == "bottom_up":
        return k["edges"] * (s_e + 4) + k["vertices"] * (2 * s_o + 4) + k["launches"] * V // 4
    if name == "union_find":
        return k["edges"] * (s_e + 8) + k["vertices"] * (2 * s_o + 4) * max(k["launches"], 1)
    return algorithmic_bytes(st["kernels"], s_e, s_o)[name]


def _kernel_rate(st, name, s_e, s_o, V=0):
    k = st["kernels"][name]
    if not k["ms"]:
        return None
    return round(kernel_bytes(st, name, s_e, s_o, V) / (k["ms"] / 1e3) / 1e9, 1)


def _kernel_table(st, s_e, s_o, V):
    """{kernel: {ms, launches, GBps, frac}} for the kernels a run used."""
    out = {}
    for name, k in st["kernels"].items():
        if k["ms"] <= 0:
            continue
        gbps = kernel_bytes(st, name, s_e, s_o, V) / (k["ms"] / 1e3) / 1e9
        # expand_part_host reads mapped host memory: no HBM fraction
        out[name] = {"ms": round(k["ms"], 3), "launches": k["launches"], "GBps": round(gbps, 1),
                     "frac": None if name == "expand_part_host" else round(gbps / HBM_PEAK_GBPS, 4)}
    return out


def resident_bench(A, _abi, O, h, steps, cpu_threads):
    """BFS, SSSP and CC on one fully resident weighted R-MAT graph (u64
    offsets, {dst, w} records).  BFS runs direction-optimizing and CC by
    union-find (the graph is symmetric); every result is compared bit for
    bit with the oracle's restatement of the reference at full size, which
    is also timed on all host cores and on one thread (CPU baseline)."""
    off, ew = h.offsets(), h.edge_words()
    V, E = h.num_vertices, h.num_edges
    deg = np.diff(off)
    src = int(np.argmax(deg))
    out = {"V": V, "E_directed": E, "source": src}
    s_e, s_o = 8, 8
    csr = O.CSR(off, ew[:, 0], ew[:, 1])
    oracles = {"bfs": (lambda nt: O.bfs_ref(csr, src, nthreads=nt), "bfs_opt"),
               "sssp": (lambda nt: O.sssp_ref(csr, src, nthreads=nt), "sssp_opt"),
               "cc": (lambda nt: O.cc_ref(csr, nthreads=nt), "cc_opt")}
    # timed runs without per-kernel events (g); the per-kernel tables come from
    # one run of each algorithm on a copy in a profiling context (pg)
    with A.Context() as ctx, ctx.from_arrays(off, ew) as g, A.Context(profile=True) as pctx, \
            pctx.from_arrays(off, ew) as pg:
        out["symmetric"] = g.info()["symmetric"]
        for algo in ("bfs", "sssp", "cc"):
            fn = getattr(g, algo)
            args = () if algo == "cc" else (src,)
            first, st = fn(*args)
            runs = [fn(*args, output=False)[1] for _ in range(steps)]
            prof_run = getattr(pg, algo)(*args, output=False)[1]
            ms = float(np.median([r["total_ms"] for r in runs]))
            num = E / 2 if algo == "cc" else st["reached_degree_sum"] / 2
            out[algo] = {"gteps": round(num / (ms / 1e3) / 1e9, 3), "ms": round(ms, 3), "path": st["path"],
                         "iterations": st["iterations"], "reached": st["reached"],
                         "kernels": _kernel_table(prof_run, s_e, s_o, V),
                         "kernels_from": "one run with per-kernel events after the timed runs"}
            ofn, what = oracles[algo]
            ref, cst, tcpu = oracle_run(f"{algo} resident", ofn, cpu_threads)
            out[algo]["exact"] = bool(np.array_equal(first, ref))
            out[algo]["parity"] = f"oracle {algo}_ref (the reference's {what} restated), full size"
            single = single_thread_sample(O, f"{algo} resident", ofn)
            out[algo]["cpu_baseline"] = cpu_baseline_block(num / tcpu / 1e9, cpu_threads, cst, tcpu, single,
                                                           f"oracle restatement of {what}")
            del ref
            if algo in ("bfs", "cc"):
                # the reference's own schedules (top-down frontier BFS, label
                # propagation) on the same device, for comparison
                with A.Context(bfs_direction=_abi.BFS_TOP_DOWN, cc_method=_abi.CC_LABEL_PROP) as c2, \
                        c2.from_arrays(off, ew) as g2:
                    plain = [getattr(g2, algo)(*args, output=False)[1] for _ in range(max(1, steps // 2))]
                out[algo]["plain_ms"] = round(float(np.median([p["total_ms"] for p in plain])), 3)
                out[algo]["plain_path"] = plain[0]["path"]
            del first
    return out


def per_algo_bench(A, _abi, O, bfs_scale, er_log_v, er_log_e, steps, cpu_threads, pr_method):
    """GTEPS per algorithm: BFS on R-MAT (C1 shape) and CC / PageRank on an
    Erdos-Renyi graph (C3 shape), fully resident.  Every result is compared
    with the oracle at full size (BFS levels and CC labels bit for bit,
    PageRank by the L1 distance to the FP64 power iteration), and the
    oracle's restatement of the reference is timed on all host cores and on
    one thread as the CPU baseline."""
    out = {}
    # BFS: R-MAT scale 20, ef 16, symmetrized, u32 offsets, source 0
    t = time.perf_counter()
    h = gen_graph(A, scale=bfs_scale, edge_factor=16, seed=1)
    off, dst = h.offsets().copy(), h.dst().copy()
    log(f"bfs graph: R-MAT s{bfs_scale} V={h.num_vertices} E={h.num_edges} ({time.perf_counter() - t:.1f}s)")
    V1 = h.num_vertices
    # timed runs without per-kernel events (their records add launch gaps at
    # this size); the per-kernel table comes from one profiled run after them
    with A.Context() as ctx:
        g = ctx.adopt(h, offset_bytes=4)
        lv, st = g.bfs(0)
        runs = [g.bfs(0, output=False)[1] for _ in range(steps)]
        g.close()
    with A.Context(profile=True) as pctx, pctx.from_arrays(off.astype(np.uint32), dst) as pg:
        pg.bfs(0, output=False)
        prof_run = pg.bfs(0, output=False)[1]
    ms = float(np.median([r["total_ms"] for r in runs]))
    num = st["reached_degree_sum"] / 2
    gteps = num / (ms / 1e3) / 1e9
    csr = O.CSR(off, dst)
    ref, cst, tcpu = oracle_run("bfs C1", lambda nt: O.bfs_ref(csr, 0, nthreads=nt), cpu_threads)
    single = single_thread_sample(O, "bfs C1", lambda nt: O.bfs_ref(csr, 0, nthreads=nt))
    out["bfs"] = {"workload": f"bfs_rmat_s{bfs_scale}_ef16_src0_u32off", "gteps": round(gteps, 3),
                  "ms": round(ms, 3), "levels": st["iterations"], "reached": st["reached"],
                  "exact": bool(np.array_equal(lv, ref)), "parity": "oracle bfs_ref, full size",
                  "path": st["path"], "kernels": _kernel_table(prof_run, 4, 4, V1),
                  "kernels_from": "one run with per-kernel events after the timed runs",
                  "cpu_baseline": cpu_baseline_block(num / tcpu / 1e9, cpu_threads, cst, tcpu, single,
                                                     "oracle restatement of bfs_opt")}
    del h, off, dst, ref
    # CC and PageRank: Erdos-Renyi, 2^er_log_v vertices, 2^er_log_e undirected
    # pairs, symmetrized (so the out-CSR is also the in-CSR PageRank pulls over)
    t = time.perf_counter()
    h = gen_graph(A, type=1, num_vertices=1 << er_log_v, num_samples=1 << er_log_e, seed=3)
    V, E = h.num_vertices, h.num_edges
    off, dst = h.offsets(), h.dst()
    log(f"er graph: V={V} E={E} ({time.perf_counter() - t:.1f}s)")
    csr = O.CSR(off, dst)
    # timed runs without per-kernel events; the per-kernel tables come from
    # one run of each algorithm in a profiling context (pctx) after them
    with A.Context() as ctx, A.Context(profile=True) as pctx:
        with ctx.from_arrays(off, dst) as g:
            lab, st = g.cc()
            ccs = [g.cc(output=False)[1] for _ in range(max(1, steps // 2))]
        with pctx.from_arrays(off, dst) as pg:
            pg.cc(output=False)
            cc_prof = pg.cc(output=False)[1]
        ms = float(np.median([c["total_ms"] for c in ccs]))
        out["cc"] = {"workload": f"cc_er_v2^{er_log_v}_e2^{er_log_e}", "gteps": round(E / 2 / (ms / 1e3) / 1e9, 3),
                     "ms": round(ms, 3), "iterations": st["iterations"], "path": st["path"],
                     "components": int(np.count_nonzero(lab == np.arange(V, dtype=np.uint32))),
                     "kernels": _kernel_table(cc_prof, 4, 8, V),
                     "kernels_from": "one run with per-kernel events after the timed runs"}
        # the oracle's label propagation (cc_opt restated), all cores: labels
        # must be identical; it is also the CPU baseline
        ref, cst, tcpu = oracle_run("cc C3", lambda nt: O.cc_ref(csr, nthreads=nt), cpu_threads)
        out["cc"]["exact"] = bool(np.array_equal(lab, ref))
        out["cc"]["parity"] = "oracle cc_ref (label propagation restated), full size"
        single = single_thread_sample(O, "cc C3", lambda nt: O.cc_ref(csr, nthreads=nt))
        out["cc"]["cpu_baseline"] = cpu_baseline_block(E / 2 / tcpu / 1e9, cpu_threads, cst, tcpu, single,
                                                       "oracle restatement of cc_opt")
        with A.Context(cc_method=_abi.CC_LABEL_PROP) as c2, c2.from_arrays(off, dst) as g2:
            _, st2 = g2.cc(output=False)
            out["cc"]["label_propagation_ms"] = round(st2["total_ms"], 3)
        del lab, ref
        outdeg = np.diff(off).astype(np.uint32)
        with ctx.from_arrays(off, dst, out_degree=outdeg) as g:
            r, res, _ = g.pagerank(iterations=20, method=pr_method)
            prs = [g.pagerank(iterations=20, method=pr_method)[2] for _ in range(max(1, steps // 2))]
        with pctx.from_arrays(off, dst, out_degree=outdeg) as pg:
            pr_prof = pg.pagerank(iterations=20, method=pr_method)[2]
        ms = float(np.median([p["total_ms"] for p in prs]))
        out["pagerank"] = {"workload": f"pr_er_v2^{er_log_v}_e2^{er_log_e}_20it_fp32",
                           "gteps": round(E * 20 / (ms / 1e3) / 1e9, 3), "ms": round(ms, 3),
                           "method": {_abi.PR_AUTO: "auto", _abi.PR_PULL: "pull", _abi.PR_TILED: "tiled",
                                      _abi.PR_BLOCKED: "blocked"}[pr_method],
                           "final_residual": float(res[-1]),
                           "kernel_GBps": _kernel_rate(pr_prof, "pr_pull", 4, 8)}
        # FP64 power iteration of the same formula on all cores: L1 distance
        O.set_num_threads(cpu_threads)
        t0 = time.perf_counter()
        ref, _ = O.pr_norm(csr, outdeg, 0.85, 20)
        tcpu = time.perf_counter() - t0
        l1 = float(np.abs(r.astype(np.float64) - ref).sum())
        log(f"cpu pagerank C3: {tcpu:.2f} s on {cpu_threads} threads, L1 {l1:.3g}")
        out["pagerank"]["l1_vs_fp64_oracle"] = l1
        out["pagerank"]["parity"] = ("L1 <= 1e-5 vs oracle pr_norm (FP64), full size" if l1 <= 1e-5
                                     else "MISMATCH: L1 above 1e-5")
        O.set_num_threads(1)
        n = 1 << 24
        while True:
            with O.edge_limit(n):
                t0 = time.perf_counter()
                O.pr_norm(csr, outdeg, 0.85, 1)
                t1 = time.perf_counter() - t0
            pulled = O.last_pulled()
            if pulled < n or t1 >= 1.0 or n >= 1 << 36:
                break
            n *= 4
        O.set_num_threads(cpu_threads)
        log(f"cpu pagerank C3 single thread: {pulled} edges pulled, {t1:.2f} s")
        out["pagerank"]["cpu_baseline"] = {
            "value": round(E * 20 / tcpu / 1e9, 4), "unit": "GTEPS", "cores": cpu_threads, "kind": "port",
            "sample": f"oracle FP64 power iteration, 20 iterations on {cpu_threads} threads ({tcpu:.2f} s)",
            "single_thread": {"value": round(pulled / t1 / 1e9, 4), "unit": "GTEPS", "cores": 1,
                              "sample": f"the first {pulled} edges pulled by one iteration of the same, "
                                        f"the O(V) contribution pass included ({t1:.2f} s)"}}
        del ref, r
    h.free()
    return out


def stream_section(A, _abi, args):
    """The main line: SSSP on the budget-capped R-MAT graph, streamed."""
    t = time.perf_counter()
    h = gen_graph(A, scale=args.scale, edge_factor=args.edge_factor, seed=args.seed, weighted=True, wmin=1,
                  wmax=args.wmax)
    off = h.offsets()
    ew = h.edge_words()
    V, E = h.num_vertices, h.num_edges
    deg = np.diff(off)
    src = int(np.argmax(deg))
    log(f"generated R-MAT s{args.scale} ef{args.edge_factor}: V={V} E={E} max-degree {int(deg[src])} at {src} "
        f"({time.perf_counter() - t:.1f}s)")
    s_e, s_o = 8, 8
    budget = int(args.budget_frac * E * s_e)
    mode = {"gather": _abi.STREAM_GATHER, "zerocopy": _abi.STREAM_ZEROCOPY, "hybrid": _abi.STREAM_HYBRID}[args.stream_mode]
    ctx = A.Context(edge_budget_bytes=budget, part_bytes=args.part_mb << 20, stream_mode=mode, profile=True,
                    sssp_delta=args.delta, stream_pack=args.stream_pack)
    g = ctx.from_arrays(off, ew)
    info = g.info()
    log(f"plan: static edges {info['static_edges']} ({info['static_edges'] / E:.1%}), static vertices "
        f"{info['static_vertices']}, on-demand edges {info['ondemand_edges']}, part {info['part_edges']} edges")

    # warm-up (the first run also returns the distances for the parity check)
    dist, st0 = g.sssp(src)
    for _ in range(max(args.warmup - 1, 0)):
        g.sssp(src, output=False)
    reached_deg = st0["reached_degree_sum"]

    ctx.synchronize()
    t0 = time.perf_counter()
    runs = [g.sssp(src, output=False)[1] for _ in range(args.steps)]
    ctx.synchronize()
    T = time.perf_counter() - t0
    step_s = T / args.steps
    gteps = (reached_deg / 2) / step_s / 1e9

    # per-kernel live timing and algorithmic bytes over the timed steps
    kern = {}
    for name in ("expand_local", "expand_heavy", "expand_part", "expand_part_host"):
        kern[name] = {k: sum(r["kernels"][name][k] for r in runs) for k in ("launches", "ms", "edges", "vertices")}
    abytes = algorithmic_bytes(kern, s_e, s_o)
    dom = max((k for k in abytes if k != "expand_part_host"), key=lambda k: kern[k]["ms"])
    dk = kern[dom]
    achieved = abytes[dom] / (dk["ms"] / 1e3) / 1e9 if dk["ms"] > 0 else 0.0
    streamed = sum(r["streamed_bytes"] for r in runs) / args.steps
    link_h2d, link_d2h = ctx.measure_link(256 << 20, 8)
    if args.copy_probe_mb > 0:
        ctx.copy_probe(args.copy_probe_mb << 20)
    workload = f"sssp_rmat_s{args.scale}_ef{args.edge_factor}_w1-{args.wmax}_budget{int(args.budget_frac * 100)}"
    # HBM bytes per launch of the dominant kernel, from the committed rocprofv3
    # FETCH_SIZE / WRITE_SIZE summary of this workload (calibrated on a copy)
    traffic = None
    tf = os.path.join(ROOT, "profiles", f"{workload}_traffic.json")
    if os.path.exists(tf):
        with open(tf) as f:
            tj = json.load(f)
        if tj.get("kernel") == dom:
            traffic = tj.get("hbm_bytes_per_launch")

    result = {
        "metric": METRIC,
        "value": round(gteps, 4),
        "unit": "GTEPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(step_s * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u32",
        "data": "synthetic",
        "config": {
            "workload": workload,
            "graph": f"R-MAT scale {args.scale}, edge factor {args.edge_factor}, (0.57,0.19,0.19,0.05), "
                     f"symmetrized, weights uniform in [1, {args.wmax}], seed {args.seed}",
            "V": V, "E_directed": E, "source": src, "offsets": "u64", "edge_record": "{u32 dst, u32 w}",
            "edge_budget_bytes": budget, "static_edges": info["static_edges"],
            "ondemand_edges": info["ondemand_edges"], "part_bytes": args.part_mb << 20,
            "stream_mode": args.stream_mode, "sssp_delta": args.delta,
            "iterations": st0["iterations"], "reached": st0["reached"],
        },
        "roofline": {
            "bound": "hbm",
            "kernel": dom,
            "achieved": round(achieved, 2),
            "peak": HBM_PEAK_GBPS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBPS, 4),
            "traffic": traffic,
            "launches_per_step": dk["launches"] / args.steps,
            "algorithmic_bytes_per_launch": abytes[dom] / max(dk["launches"], 1),
            "avg_launch_us": dk["ms"] * 1e3 / max(dk["launches"], 1),
            "kernels": {k: {"ms_per_step": kern[k]["ms"] / args.steps, "launches_per_step": kern[k]["launches"] /
                            args.steps, "GBps": (abytes[k] / (kern[k]["ms"] / 1e3) / 1e9) if kern[k]["ms"] else 0}
                        for k in kern},
        },
        "link": {
            "streamed_bytes_per_step": streamed,
            "achieved_GBps": round(streamed / step_s / 1e9, 3),
            "peak_h2d_GBps": round(link_h2d, 2),
            "peak_d2h_GBps": round(link_d2h, 2),
            "frac": round(streamed / step_s / 1e9 / link_h2d, 4) if link_h2d else None,
            "stream_record_bytes": round(sum(r["streamed_bytes"] for r in runs) /
                                         max(1, sum(r["streamed_edges"] for r in runs))),
            "gather_ms_per_step": sum(r["fill_ms"] for r in runs) / args.steps,
            "prepare_ms_per_step": sum(r["pre_ms"] for r in runs) / args.steps,
            "overload_ms_per_step": sum(r["overload_ms"] for r in runs) / args.steps,
        },
    }

    # comparison points for the on-demand transfer (same graph and source):
    # every chunk holding an active vertex copied whole (partition streaming,
    # the reference's bfs.cu / sssp.cu scheme at 64 MiB partitions), per-vertex
    # gathering only (the CalculateOpt.cu scheme), both at the same budget, and
    # no device edges at all with kernels reading the graph from host memory
    # (the unified-memory "Share" drivers of bfs.cu / sssp.cu in the form they
    # take without page migration)
    if args.stream_baselines:
        result["stream_baselines"] = {}
        for name, thr, bud, md in (("partition_streaming", 1e-9, budget, mode), ("gather_only", 2.0, budget, mode),
                                   ("share_zero_copy", 0.0, 1, _abi.STREAM_ZEROCOPY)):
            with A.Context(edge_budget_bytes=bud, part_bytes=args.part_mb << 20, stream_mode=md,
                           sssp_delta=args.delta, dense_threshold=thr) as c2, c2.from_arrays(off, ew) as g2:
                g2.sssp(src, output=False)
                rs = [g2.sssp(src, output=False)[1] for _ in range(max(2, args.steps // 2))]
                ms = float(np.median([r["total_ms"] for r in rs]))
                sb = float(np.median([r["streamed_bytes"] for r in rs]))
                result["stream_baselines"][name] = {
                    "ms_per_step": round(ms, 3), "gteps": round(reached_deg / 2 / (ms / 1e3) / 1e9, 4),
                    "streamed_GB_per_step": round(sb / 1e9, 3), "link_GBps": round(sb / (ms / 1e3) / 1e9, 2)}
        result["stream_baselines"]["default"] = {
            "ms_per_step": round(step_s * 1e3, 3), "gteps": round(gteps, 4),
            "streamed_GB_per_step": round(streamed / 1e9, 3), "link_GBps": round(streamed / step_s / 1e9, 2)}

    if not args.no_cpu_baseline:
        from oracle import oracle as O
        nt = args.cpu_threads or default_cpu_threads()
        csr = O.CSR(off, ew[:, 0], ew[:, 1])
        S, PE = info["static_edges"], info["part_edges"]
        ref, cst, tcpu = oracle_run("sssp streamed", lambda n: O.sssp_ref(csr, src, S, PE, n), nt)
        ok = bool(np.array_equal(ref, dist))
        del ref
        single = single_thread_sample(O, "sssp streamed", lambda n: O.sssp_ref(csr, src, S, PE, n))
        result["cpu_baseline"] = cpu_baseline_block(
            (reached_deg / 2) / tcpu / 1e9, nt, cst, tcpu, single,
            "one SSSP of the same graph and source by the oracle's restatement of CalculateOpt.cu sssp_opt "
            "(same static prefix and part size)")
        result["cpu_baseline"]["host_cores_visible"] = os.cpu_count()
        # the reference's host gather (fillEdgeArrByMultiThread, 20 threads)
        # timed alone over the whole on-demand region, and the time it would
        # spend on the edges the restated run gathered
        result["cpu_baseline"]["reference_gather"] = gather_baseline(O, csr, ew, int(info["static_vertices"]), V,
                                                                     cst["overload_edges"], args.steps and 1)
        result["parity"] = "bit-exact vs oracle" if ok else "MISMATCH vs oracle"
        if not ok:
            log("GPU distances differ from the oracle")
    g.close()
    ctx.close()
    del h, off, ew
    return result


def c4_shape_section(A, _abi, O, args, cpu_threads):
    """SSSP and BFS at the shape of the out-of-memory configuration (R-MAT
    edge factor 32, weights uniform in [1, 255], about a third of the edge
    bytes resident, the rest streamed from pinned host memory; source = the
    highest-degree vertex) at a scale one benchmark run can generate and
    hold.  Ids then need more bits than fit beside an 8-bit weight in one
    32-bit word, so on-demand edges stream as 5-byte split records (u32 id +
    weight byte), as they would at scale 30.  Reports GTEPS, link GB/s against
    the measured pinned peak, oracle parity and the CPU baseline."""
    t = time.perf_counter()
    sc, ef = args.c4_scale, args.c4_edge_factor
    h = gen_graph(A, scale=sc, edge_factor=ef, seed=args.seed, weighted=True, wmin=1, wmax=255)
    off, ew = h.offsets(), h.edge_words()          # views; the graph keeps the arrays
    V, E = h.num_vertices, h.num_edges
    deg = np.diff(off)
    src = int(np.argmax(deg))
    log(f"c4-shape graph: R-MAT s{sc} ef{ef} w1-255 V={V} E={E} ({time.perf_counter() - t:.1f}s)")
    budget = int(args.c4_budget_frac * 8 * E)
    out = {"workload": f"rmat_s{sc}_ef{2 * ef}directed_w1-255_budget{int(round(args.c4_budget_frac * 100))}",
           "graph": f"R-MAT scale {sc}, {ef} undirected samples per vertex, symmetrized ({2 * ef} directed edges per "
                    f"vertex), (0.57,0.19,0.19,0.05), weights uniform in [1, 255], seed {args.seed} (the scale-30 "
                    f"configuration's shape at a scale whose host copy the oracle can check bit for bit)",
           "V": V, "E_directed": E, "source": src, "edge_budget_bytes": budget}
    # the library copies the arrays and, once the on-demand edges sit in their
    # packed pinned form, frees its {dst, w} copy (release_host_edges): the
    # host footprint is then offsets + 5-byte records, which is what a scale-30
    # graph needs; the generator's arrays stay with this script for the oracle
    with A.Context(edge_budget_bytes=budget, part_bytes=args.part_mb << 20, profile=True,
                   release_host_edges=True) as ctx:
        g = ctx.from_arrays(off, ew)
        info = g.info()
        out.update({"static_edges": info["static_edges"], "ondemand_edges": info["ondemand_edges"],
                    "library_host_bytes": info["host_bytes"], "device_bytes": info["device_bytes"],
                    "edge_record_bytes": 8 * E})
        h2d, _ = ctx.measure_link(256 << 20, 8)
        csr = O.CSR(off, ew[:, 0], ew[:, 1])
        S, PE = info["static_edges"], info["part_edges"]
        n_runs = max(2, args.steps // 4)
        for algo in ("sssp", "bfs"):
            fn = getattr(g, algo)
            first, st = fn(src)
            ctx.synchronize()
            t0 = time.perf_counter()
            runs = [fn(src, output=False)[1] for _ in range(n_runs)]
            ctx.synchronize()
            step_s = (time.perf_counter() - t0) / n_runs
            num = st["reached_degree_sum"] / 2
            sb = float(np.mean([r["streamed_bytes"] for r in runs]))
            sec = {"gteps": round(num / step_s / 1e9, 4), "ms_per_run": round(step_s * 1e3, 3), "runs": n_runs,
                   "iterations": st["iterations"], "reached": st["reached"],
                   "stream_record_bytes": round(sb / max(1.0, float(np.mean([r["streamed_edges"] for r in runs])))),
                   "link": {"streamed_GB_per_run": round(sb / 1e9, 3), "achieved_GBps": round(sb / step_s / 1e9, 2),
                            "peak_h2d_GBps": round(h2d, 2), "frac": round(sb / step_s / 1e9 / h2d, 4) if h2d else None,
                            "gather_ms_per_run": round(float(np.mean([r["fill_ms"] for r in runs])), 2)},
                   "kernels": _kernel_table(runs[len(runs) // 2], 8, 8, V)}
            ofn = (lambda nt: O.sssp_ref(csr, src, S, PE, nt)) if algo == "sssp" else \
                (lambda nt: O.bfs_ref(csr, src, S, PE, nt))
            ref, cst, tcpu = oracle_run(f"{algo} c4-shape", ofn, cpu_threads)
            sec["exact"] = bool(np.array_equal(first, ref))
            sec["parity"] = f"oracle {algo}_ref (same static prefix and part size), full size"
            del ref, first
            single = single_thread_sample(O, f"{algo} c4-shape", ofn)
            sec["cpu_baseline"] = cpu_baseline_block(num / tcpu / 1e9, cpu_threads, cst, tcpu, single,
                                                     f"oracle restatement of {algo}_opt (same static prefix)")
            if algo == "sssp":
                sec["cpu_baseline"]["reference_gather"] = gather_baseline(O, csr, ew, int(info["static_vertices"]),
                                                                          V, cst["overload_edges"])
            out[algo] = sec
            log(f"c4-shape {algo}: {sec['gteps']} GTEPS, {sec['ms_per_run']} ms/run, exact {sec['exact']}")
        g.close()
    h.free()
    return out


def c4_full_section(A, _abi, O, args, shape):
    """BFS and SSSP at the out-of-memory configuration itself: R-MAT at
    args.c4_full_scale (30: 2^30 vertices and 16 undirected samples per
    vertex, symmetrized, i.e. 32 directed edges per vertex, ~34 G directed
    edges and 275 GB of {dst, w} records), weights uniform in [1, 255], a
    96 GB device edge budget (scaled by 2^(scale - 30) below 30), source =
    the highest-degree vertex.  The graph is generated straight into the
    device's static prefix and the pinned packed tail (asc_graph_generate):
    no host copy of its records exists, so parity is checked by properties
    on 2^24 sampled edges drawn by the oracle's restatement of the sampler
    (orc_rmat_pairs): BFS levels differ by at most one across every sampled
    edge, SSSP distances satisfy d[v] <= d[u] + w(u, v) both ways, BFS and
    SSSP reach the same vertices, and the source holds 1.  The scale-26
    section of the same shape checks bit-exact parity with the oracle."""
    sc = args.c4_full_scale
    budget = int(args.c4_budget_gb * 1e9 * 2.0 ** min(0, sc - 30))
    t = time.perf_counter()
    with A.Context(edge_budget_bytes=budget, part_bytes=args.part_mb << 20, profile=True) as ctx:
        g = ctx.generate_graph(type=0, scale=sc, edge_factor=16, seed=args.seed, wmin=1, wmax=255, offset_bytes=8)
        gen_s = time.perf_counter() - t
        info = g.info()
        V, E, src = info["num_vertices"], info["num_edges"], info["max_degree_vertex"]
        log(f"c4 graph: R-MAT s{sc} 32 directed edges per vertex, w1-255: V={V} E={E}, static "
            f"{info['static_edges']} ({gen_s:.1f}s, host {info['host_bytes'] / 1e9:.1f} GB, "
            f"device {info['device_bytes'] / 1e9:.1f} GB)")
        out = {"workload": f"rmat_s{sc}_ef32directed_w1-255_budget{budget / 1e9:.0f}GB",
               "graph": f"R-MAT scale {sc}, 16 undirected samples per vertex, symmetrized (32 directed edges per "
                        f"vertex), (0.57,0.19,0.19,0.05), self-loops and duplicates removed, weights uniform in "
                        f"[1, 255], seed {args.seed}, generated on the GPU straight into the device prefix and the "
                        f"pinned tail",
               "V": V, "E_directed": E, "source": src, "edge_budget_bytes": budget,
               "static_edges": info["static_edges"], "ondemand_edges": info["ondemand_edges"],
               "edge_record_bytes": 8 * E, "library_host_bytes": info["host_bytes"],
               "device_bytes": info["device_bytes"], "generation_s": round(gen_s, 1)}
        h2d, _ = ctx.measure_link(256 << 20, 8)
        vals = {}
        for algo in ("bfs", "sssp"):
            fn = getattr(g, algo)
            first, st = fn(src)
            ctx.synchronize()
            t0 = time.perf_counter()
            runs = [fn(src, output=False)[1] for _ in range(args.c4_full_runs)]
            ctx.synchronize()
            step_s = (time.perf_counter() - t0) / args.c4_full_runs
            num = st["reached_degree_sum"] / 2
            sb = float(np.mean([r["streamed_bytes"] for r in runs]))
            out[algo] = {"gteps": round(num / step_s / 1e9, 4), "ms_per_run": round(step_s * 1e3, 1),
                         "runs": args.c4_full_runs, "iterations": st["iterations"], "reached": st["reached"],
                         "stream_record_bytes": round(sb / max(1.0, float(np.mean([r["streamed_edges"]
                                                                                    for r in runs])))),
                         "link": {"streamed_GB_per_run": round(sb / 1e9, 2),
                                  "achieved_GBps": round(sb / step_s / 1e9, 2), "peak_h2d_GBps": round(h2d, 2),
                                  "frac": round(sb / step_s / 1e9 / h2d, 4) if h2d else None,
                                  "gather_ms_per_run": round(float(np.mean([r["fill_ms"] for r in runs])), 1)},
                         "kernels": _kernel_table(runs[-1], 8, 8, V)}
            vals[algo] = first
            log(f"c4 {algo}: {out[algo]['gteps']} GTEPS, {out[algo]['ms_per_run']} ms/run, "
                f"link {out[algo]['link']['achieved_GBps']} GB/s")
        g.close()
    # properties on sampled edges
    t = time.perf_counter()
    u, v, w = O.rmat_pairs(args.seed, sc, 16 << sc, 1 << 24, wmin=1, wmax=255)
    keep = u != v
    u, v, w = u[keep], v[keep], w[keep]
    lv, d = vals["bfs"], vals["sssp"]
    unreached = np.uint32(V + 1)
    lu, lw = lv[u].astype(np.int64), lv[v].astype(np.int64)
    du, dv = d[u].astype(np.int64), d[v].astype(np.int64)
    checks = {
        "source_holds_1": bool(lv[src] == 1 and d[src] == 1),
        "same_reached_set": bool(np.array_equal(lv == unreached, d == unreached)),
        "reached_count": bool(int((lv != unreached).sum()) == out["bfs"]["reached"] == out["sssp"]["reached"]),
        "bfs_levels_within_one": bool(np.all((lu == V + 1) == (lw == V + 1)) and
                                      np.all(np.abs(lu - lw)[lu != V + 1] <= 1)),
        "sssp_edge_inequality": bool(np.all(((du == V + 1) & (dv == V + 1)) |
                                            ((du != V + 1) & (dv != V + 1) & (dv <= du + w) & (du <= dv + w)))),
    }
    out["parity"] = {"sampled_edges": int(keep.sum()), "checks": checks,
                     "ok": all(checks.values()), "seconds": round(time.perf_counter() - t, 1),
                     "exact_parity": "bit-exact vs the oracle at scale 26 of the same shape (c4_shape)"}
    log(f"c4 properties: {checks}")
    if shape and "sssp" in shape:
        cb = shape["sssp"]["cpu_baseline"]
        out["cpu_baseline"] = {"value": cb["value"], "unit": "GTEPS", "cores": cb["cores"], "kind": "port",
                               "sample": "the oracle's restatement of sssp_opt timed on the scale-26 graph of the same "
                                         "shape (c4_shape); the scale-30 run has no host copy of its records to "
                                         "time it on",
                               "edges_relaxed_per_s": cb["edges_relaxed_per_s"],
                               "single_thread": cb["single_thread"]}
    del vals, lv, d
    return out


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--scale", type=int, default=24, help="R-MAT scale of the main streamed line (0 = skip)")
    ap.add_argument("--edge-factor", type=int, default=16)
    ap.add_argument("--seed", type=int, default=24)
    ap.add_argument("--wmax", type=int, default=64, help="edge weights uniform in [1, wmax]")
    ap.add_argument("--budget-frac", type=float, default=0.4)
    ap.add_argument("--part-mb", type=int, default=64)
    ap.add_argument("--stream-mode", choices=["gather", "zerocopy", "hybrid"], default="gather")
    ap.add_argument("--stream-pack", type=int, default=1, choices=[0, 1, 2],
                    help="streamed record form: 1 packed (4 bytes, else 5 when the weights fit a byte), "
                         "2 always 5 bytes, 0 8-byte {dst, w} records")
    ap.add_argument("--delta", type=int, default=0, help="SSSP band width (0 = library default, <0 = off)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-stream-baselines", dest="stream_baselines", action="store_false",
                    help="skip the partition-streaming and gather-only comparison runs")
    ap.add_argument("--cpu-threads", type=int, default=0,
                    help="CPU baseline threads (0 = the cores this process may use, at most 16)")
    ap.add_argument("--serve-scale", type=int, default=26, help="R-MAT scale of the serving section (0 = skip)")
    ap.add_argument("--algo-bfs-scale", type=int, default=20, help="R-MAT scale of the per-algo BFS line (0 = skip)")
    ap.add_argument("--algo-er", type=str, default="26,30",
                    help="log2 V,log2 undirected pairs of the CC/PageRank graph ('' = skip)")
    ap.add_argument("--serve-queries", type=int, default=512)
    ap.add_argument("--serve-eager", action="store_true",
                    help="serving steps launch their kernels one by one instead of replaying a hipGraph")
    ap.add_argument("--resident-scale", type=int, default=26,
                    help="R-MAT scale of the resident BFS/SSSP/CC section (0 = skip)")
    ap.add_argument("--copy-probe-mb", type=int, default=0,
                    help="also launch one plain device copy of this many MiB (counter calibration)")
    ap.add_argument("--pr-method", choices=["auto", "pull", "tiled", "blocked"], default="auto",
                    help="PageRank schedule of the per-algorithm line")
    ap.add_argument("--c4-scale", type=int, default=26,
                    help="R-MAT scale of the section at the out-of-memory configuration's shape (0 = skip)")
    ap.add_argument("--c4-edge-factor", type=int, default=16,
                    help="undirected samples per vertex (symmetrized: twice as many directed edges)")
    ap.add_argument("--c4-budget-frac", type=float, default=0.355,
                    help="device edge budget of that section as a fraction of its edge+weight bytes "
                         "(96 GB of 270 GB at scale 30)")
    ap.add_argument("--c4-full-scale", type=int, default=30,
                    help="R-MAT scale of the out-of-memory configuration itself, generated on the GPU (0 = skip)")
    ap.add_argument("--c4-budget-gb", type=float, default=96.0,
                    help="device edge budget at scale 30 in GB (scaled by 2^(scale-30) below)")
    ap.add_argument("--c4-full-runs", type=int, default=2, help="timed runs per algorithm at full scale")
    args = ap.parse_args()

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py runs on one MI355X only: the library has no multi-GPU path (a 1-D vertex partition "
              "with an all-to-all of relaxations per step is the intended sharding); use --gpus 1", file=sys.stderr)
        sys.exit(2)

    import ascetic_amd as A
    from ascetic_amd import _abi
    from oracle import oracle as O

    cpu_threads = args.cpu_threads or default_cpu_threads()
    if args.scale > 0:
        result = stream_section(A, _abi, args)
    else:   # sections only (profiling one workload at a time)
        result = {"metric": METRIC, "value": None, "unit": "GTEPS", "n_gpus": 1, "steps": args.steps,
                  "warmup": args.warmup, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
                  "dtype": "u32", "data": "synthetic", "config": {"workload": "sections only"}}
    if args.algo_bfs_scale > 0 and args.algo_er:
        lv, le = (int(x) for x in args.algo_er.split(","))
        pm = {"auto": _abi.PR_AUTO, "pull": _abi.PR_PULL, "tiled": _abi.PR_TILED, "blocked": _abi.PR_BLOCKED}
        result["per_algo"] = per_algo_bench(A, _abi, O, args.algo_bfs_scale, lv, le, args.steps, cpu_threads,
                                            pm[args.pr_method])
        if args.scale > 0:
            result["per_algo"]["sssp"] = {"workload": result["config"]["workload"], "gteps": result["value"]}
    hr = None
    if args.resident_scale > 0:
        t = time.perf_counter()
        hr = gen_graph(A, scale=args.resident_scale, edge_factor=16, seed=args.seed, weighted=True, wmin=1, wmax=64)
        log(f"resident graph: R-MAT s{args.resident_scale} ef16 weighted V={hr.num_vertices} E={hr.num_edges} "
            f"({time.perf_counter() - t:.1f}s)")
        result["resident"] = {"workload": f"rmat_s{args.resident_scale}_ef16_w1-64_resident_u64off",
                              **resident_bench(A, _abi, O, hr, args.steps, cpu_threads)}
        if args.serve_scale != args.resident_scale:
            hr.free()
            hr = None
    if args.serve_scale > 0:
        # steps replayed from hipGraphs ended the process with a host SIGSEGV
        # inside the HIP runtime once under rocprofv3 --kernel-trace (cause not
        # established, no device fault): under a profiler the steps are
        # launched kernel by kernel instead
        profiled = "ROCP_TOOL_LIBRARIES" in os.environ or any(k.startswith("ROCPROF_") for k in os.environ)
        eager = args.serve_eager or profiled
        if profiled and not args.serve_eager:
            log("serving: profiler detected, engine steps launched kernel by kernel (no hipGraph replay)")
        # the serving section adopts (consumes) the resident graph when the scales match
        result["serving"] = serving_bench(A, O, args.serve_scale, args.seed, args.serve_queries,
                                          [0.25, 1.0, 4.0, 8.0], cpu_threads, h=hr, step_graphs=not eager)
        result["serving"]["step_graphs"] = not eager
        hr = None
    if args.c4_scale > 0:
        result["c4_shape"] = c4_shape_section(A, _abi, O, args, cpu_threads)
    if args.c4_full_scale > 0:
        result["c4"] = c4_full_section(A, _abi, O, args, result.get("c4_shape"))
    print(json.dumps(result), flush=True)
    bad = [k for k, v in _parity_flags(result) if not v]
    if bad:
        log("parity failures: " + ", ".join(bad))
        sys.exit(1)


def _parity_flags(result):
    """(section, ok) for every parity check in the JSON line."""
    flags = []
    if "parity" in result:
        flags.append(("main", not result["parity"].startswith("MISMATCH")))
    for k, v in result.get("per_algo", {}).items():
        if "exact" in v:
            flags.append((f"per_algo.{k}", v["exact"]))
        if k == "pagerank":
            flags.append(("per_algo.pagerank", not v.get("parity", "").startswith("MISMATCH")))
    for sec in ("resident", "c4_shape"):
        for k, v in result.get(sec, {}).items():
            if isinstance(v, dict) and "exact" in v:
                flags.append((f"{sec}.{k}", v["exact"]))
    if "c4" in result:
        flags.append(("c4.properties", result["c4"]["parity"]["ok"]))
    sv = result.get("serving", {})
    if sv:
        flags.append(("serving.exact_sample", not sv.get("exact_sample", "").startswith("MISMATCH")))
        flags.append(("serving.oracle_sample", not sv.get("oracle_sample", "").startswith("MISMATCH")))
    return flags


if __name__ == "__main__":
    main()
