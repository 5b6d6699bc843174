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


def algorithmic_bytes(kern, s_e, s_o):
    """Bytes a work-optimal push relax must move: per edge the record and one
    destination value; per frontier vertex an offset pair, its value, a queue
    pop and a push (per list entry of a streamed part: id, u64 prefix, value)."""
    return {
        "expand_local": kern["expand_local"]["edges"] * (s_e + 4) + kern["expand_local"]["vertices"] * (2 * s_o + 12),
        "expand_heavy": kern["expand_heavy"]["edges"] * (s_e + 4),
        "expand_part": kern["expand_part"]["edges"] * (s_e + 4) + kern["expand_part"]["vertices"] * 16,
    }


def serving_bench(A, scale, seed, n_queries, rate_factors, budget_s=60.0, h=None, step_graphs=True):
    """Query serving on a resident weighted R-MAT graph: 70% BFS / 30% SSSP,
    sources uniform over vertices with degree > 0.  Reports one-query-per-
    launch throughput (plain kernels), the server's saturated throughput and
    Poisson-arrival sweeps with p50/p99 submit-to-done latency, and checks a
    sample of served results against the same queries run alone."""
    import threading

    t = time.perf_counter()
    if h is None:
        h = A.rmat(scale, 16, seed=seed, weighted=True, wmin=1, wmax=64)
    V = h.num_vertices
    deg = np.diff(h.offsets())
    cand = np.nonzero(deg > 0)[0]
    log(f"serving graph: R-MAT s{scale} ef16 weighted, V={V} E={h.num_edges} ({time.perf_counter() - t:.1f}s)")
    rng = np.random.default_rng(seed)
    srcs = [int(x) for x in rng.choice(cand, n_queries)]
    kinds = [A.Server.SSSP if x < 0.3 else A.Server.BFS for x in rng.random(n_queries)]
    out = {"workload": f"rmat_s{scale}_ef16_w1-64_resident_70bfs_30sssp", "queries": n_queries}
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
            out["saturated_qps"] = round(sat, 2)
            out["saturated_p50_ms"] = round(float(np.percentile(lat, 50)), 3)
            out["saturated_p99_ms"] = round(float(np.percentile(lat, 99)), 3)
            out["speedup_vs_plain"] = round(sat / plain_qps, 2)
            out["exact_sample"] = f"{len(sample)} queries bit-identical to running alone" if exact else "MISMATCH"
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
        out[name] = {"ms": round(k["ms"], 3), "launches": k["launches"], "GBps": round(gbps, 1),
                     "frac": round(gbps / HBM_PEAK_GBPS, 4)}
    return out


def resident_bench(A, _abi, h, steps):
    """BFS, SSSP and CC on one fully resident weighted R-MAT graph (u64
    offsets, {dst, w} records).  BFS runs direction-optimizing and CC by
    union-find (the graph is symmetric); both are compared bit for bit with
    the library's top-down BFS and label-propagation CC on the same device."""
    off, ew = h.offsets(), h.edge_words()
    V, E = h.num_vertices, h.num_edges
    deg = np.diff(off)
    src = int(np.argmax(deg))
    out = {"V": V, "E_directed": E, "source": src}
    s_e, s_o = 8, 8
    with A.Context(profile=True) as ctx, ctx.from_arrays(off, ew) as g:
        out["symmetric"] = g.info()["symmetric"]
        for algo in ("bfs", "sssp", "cc"):
            fn = getattr(g, algo)
            args = () if algo == "cc" else (src,)
            first, st = fn(*args)
            runs = [fn(*args, output=False)[1] for _ in range(steps)]
            ms = float(np.median([r["total_ms"] for r in runs]))
            num = E / 2 if algo == "cc" else st["reached_degree_sum"] / 2
            out[algo] = {"gteps": round(num / (ms / 1e3) / 1e9, 3), "ms": round(ms, 3), "path": st["path"],
                         "iterations": st["iterations"], "reached": st["reached"],
                         "kernels": _kernel_table(runs[len(runs) // 2], s_e, s_o, V)}
            if algo in ("bfs", "cc"):
                with A.Context(profile=True, bfs_direction=_abi.BFS_TOP_DOWN, cc_method=_abi.CC_LABEL_PROP) as c2, \
                        c2.from_arrays(off, ew) as g2:
                    ref, st2 = getattr(g2, algo)(*args)
                    plain = [getattr(g2, algo)(*args, output=False)[1]["total_ms"] for _ in range(max(1, steps // 2))]
                out[algo]["exact_vs_plain"] = bool(np.array_equal(first, ref))
                out[algo]["plain_ms"] = round(float(np.median(plain)), 3)
                out[algo]["plain_path"] = st2["path"]
                del ref
            del first
    return out


def per_algo_bench(A, _abi, O, bfs_scale, er_log_v, er_log_e, steps, cpu_threads):
    """GTEPS per algorithm: BFS on R-MAT (C1 shape) and CC / PageRank on an
    Erdos-Renyi graph (C3 shape), fully resident."""
    out = {}
    # BFS: R-MAT scale 20, ef 16, symmetrized, u32 offsets, source 0
    t = time.perf_counter()
    h = A.rmat(bfs_scale, 16, seed=1)
    off, dst = h.offsets().copy(), h.dst().copy()
    log(f"bfs graph: R-MAT s{bfs_scale} V={h.num_vertices} E={h.num_edges} ({time.perf_counter() - t:.1f}s)")
    with A.Context(profile=True) as ctx:
        g = ctx.adopt(h, offset_bytes=4)
        lv, st = g.bfs(0)
        runs = [g.bfs(0, output=False)[1] for _ in range(steps)]
        ms = float(np.median([r["total_ms"] for r in runs]))
        gteps = st["reached_degree_sum"] / 2 / (ms / 1e3) / 1e9
        csr = O.CSR(off, dst)
        tc = time.perf_counter()
        ref, _ = O.bfs_ref(csr, 0, nthreads=cpu_threads)
        tcpu = time.perf_counter() - tc
        out["bfs"] = {"workload": f"bfs_rmat_s{bfs_scale}_ef16_src0_u32off", "gteps": round(gteps, 3),
                      "ms": round(ms, 3), "levels": st["iterations"], "reached": st["reached"],
                      "exact": bool(np.array_equal(lv, ref)), "path": st["path"],
                      "kernels": _kernel_table(runs[len(runs) // 2], 4, 4, h.num_vertices),
                      "cpu_port_gteps": round(st["reached_degree_sum"] / 2 / tcpu / 1e9, 3), "cpu_cores": cpu_threads}
        g.close()
    del h, off, dst
    # CC and PageRank: Erdos-Renyi, 2^er_log_v vertices, 2^er_log_e undirected
    # pairs, symmetrized (so the out-CSR is also the in-CSR PageRank pulls over)
    t = time.perf_counter()
    h = A.erdos_renyi(1 << er_log_v, 1 << er_log_e, seed=3)
    V, E = h.num_vertices, h.num_edges
    off, dst = h.offsets(), h.dst()
    log(f"er graph: V={V} E={E} ({time.perf_counter() - t:.1f}s)")
    with A.Context(profile=True) as ctx:
        with ctx.from_arrays(off, dst) as g:
            lab, st = g.cc()
            ccs = [g.cc(output=False)[1] for _ in range(max(1, steps // 2))]
            ms = float(np.median([c["total_ms"] for c in ccs]))
            out["cc"] = {"workload": f"cc_er_v2^{er_log_v}_e2^{er_log_e}", "gteps": round(E / 2 / (ms / 1e3) / 1e9, 3),
                         "ms": round(ms, 3), "iterations": st["iterations"], "path": st["path"],
                         "components": int(np.count_nonzero(lab == np.arange(V, dtype=np.uint32))),
                         "kernels": _kernel_table(ccs[len(ccs) // 2], 4, 8, V)}
        # label propagation on the same graph: the labels must be identical
        with A.Context(profile=True, cc_method=_abi.CC_LABEL_PROP) as c2, c2.from_arrays(off, dst) as g2:
            ref, st2 = g2.cc()
            out["cc"]["exact_vs_plain"] = bool(np.array_equal(lab, ref))
            out["cc"]["plain_ms"] = round(st2["total_ms"], 3)
        del lab, ref
        outdeg = np.diff(off).astype(np.uint32)
        with ctx.from_arrays(off, dst, out_degree=outdeg) as g:
            _, res, _ = g.pagerank(iterations=20)
            prs = [g.pagerank(iterations=20)[2] for _ in range(max(1, steps // 2))]
            ms = float(np.median([p["total_ms"] for p in prs]))
            out["pagerank"] = {"workload": f"pr_er_v2^{er_log_v}_e2^{er_log_e}_20it_fp32",
                               "gteps": round(E * 20 / (ms / 1e3) / 1e9, 3), "ms": round(ms, 3),
                               "final_residual": float(res[-1]), "pr_pull_GBps": _kernel_rate(prs[-1], "pr_pull", 4, 8)}
    h.free()
    return out


def stream_section(A, _abi, args):
    """The main line: SSSP on the budget-capped R-MAT graph, streamed."""
    t = time.perf_counter()
    h = A.rmat(args.scale, args.edge_factor, seed=args.seed, weighted=True, wmin=1, wmax=args.wmax)
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
    for name in ("expand_local", "expand_heavy", "expand_part"):
        kern[name] = {k: sum(r["kernels"][name][k] for r in runs) for k in ("launches", "ms", "edges", "vertices")}
    abytes = algorithmic_bytes(kern, s_e, s_o)
    dom = max(abytes, key=lambda k: kern[k]["ms"])
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
        nt = args.cpu_threads or min(16, os.cpu_count() or 1)
        csr = O.CSR(off, ew[:, 0], ew[:, 1])
        tc = time.perf_counter()
        ref, cst = O.sssp_ref(csr, src, info["static_edges"], info["part_edges"], nt)
        tcpu = time.perf_counter() - tc
        ok = bool(np.array_equal(ref, dist))
        result["cpu_baseline"] = {
            "value": round((reached_deg / 2) / tcpu / 1e9, 4),
            "unit": "GTEPS",
            "cores": nt,
            "kind": "port",
            "sample": f"one full SSSP of the same graph and source by the oracle's restatement of "
                      f"CalculateOpt.cu sssp_opt (same static prefix and part size), {tcpu:.2f} s",
        }
        result["parity"] = "bit-exact vs oracle" if ok else "MISMATCH vs oracle"
        if not ok:
            log("GPU distances differ from the oracle")
    g.close()
    ctx.close()
    del h, off, ew
    return result


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
    ap.add_argument("--cpu-threads", type=int, default=0, help="0 = min(16, cores)")
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
    args = ap.parse_args()

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py runs on one MI355X only: the library has no multi-GPU path (a 1-D vertex partition "
              "with an all-to-all of relaxations per step is the intended sharding); use --gpus 1", file=sys.stderr)
        sys.exit(2)

    import ascetic_amd as A
    from ascetic_amd import _abi

    if args.scale > 0:
        result = stream_section(A, _abi, args)
    else:   # sections only (profiling one workload at a time)
        result = {"metric": METRIC, "value": None, "unit": "GTEPS", "n_gpus": 1, "steps": args.steps,
                  "warmup": args.warmup, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
                  "dtype": "u32", "data": "synthetic", "config": {"workload": "sections only"}}
    if args.algo_bfs_scale > 0 and args.algo_er:
        from oracle import oracle as O
        lv, le = (int(x) for x in args.algo_er.split(","))
        result["per_algo"] = per_algo_bench(A, _abi, O, args.algo_bfs_scale, lv, le, args.steps,
                                            args.cpu_threads or min(16, os.cpu_count() or 1))
        if args.scale > 0:
            result["per_algo"]["sssp"] = {"workload": result["config"]["workload"], "gteps": result["value"]}
    hr = None
    if args.resident_scale > 0:
        t = time.perf_counter()
        hr = A.rmat(args.resident_scale, 16, seed=args.seed, weighted=True, wmin=1, wmax=64)
        log(f"resident graph: R-MAT s{args.resident_scale} ef16 weighted V={hr.num_vertices} E={hr.num_edges} "
            f"({time.perf_counter() - t:.1f}s)")
        result["resident"] = {"workload": f"rmat_s{args.resident_scale}_ef16_w1-64_resident_u64off",
                              **resident_bench(A, _abi, hr, args.steps)}
        if args.serve_scale != args.resident_scale:
            hr.free()
            hr = None
    if args.serve_scale > 0:
        # the serving section adopts (consumes) the resident graph when the scales match
        result["serving"] = serving_bench(A, args.serve_scale, args.seed, args.serve_queries, [0.25, 1.0, 4.0, 8.0],
                                          h=hr, step_graphs=not args.serve_eager)
        result["serving"]["step_graphs"] = not args.serve_eager
    print(json.dumps(result), flush=True)
    if result.get("parity", "").startswith("MISMATCH"):
        sys.exit(1)


if __name__ == "__main__":
    main()
