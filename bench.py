#!/usr/bin/env python3
"""Benchmark of the KD-tree KNN query on one MI355X.

Default workload (``plain_gpu``): d=3 fp32, 4,000,000 reference points U[0,1]^3, 16,000,000
queries (half U[-0.1,1.1]^3, half from a 32-component Gaussian mixture, sigma 0.02), k=16.
A step is one ``TorchKDTree.query`` over all 16M queries (Hilbert-order presort + traversal + outputs:
squared distances, structured int32 indices and original int64 indices), with the tree and the
queries already resident in HBM.  The tree build and upload are outside the timed region, as
in the reference's benchmark (README.md:75).

Prints one JSON line.  ``roofline`` refers to the traversal kernel: its average duration is
measured with HIP events on the launch stream during the timed steps; ``achieved`` divides
the algorithmic bytes of one launch by it: per query d*4 (the query) + k*(4+4) (distance and
int32 index written once) = 140 B at d=3, k=16, plus the tree read once (N*(d*4+4) record bytes
and 5 B per node); ``peak`` is the 8 TB/s HBM3E figure, ``achievable_gbs`` the rate of the
library's own streaming copy on this GPU (``frac_of_achievable``).  The int64 original indices
the kernel also writes (the reference computes them afterwards in torch) are counted only in
``achieved_with_fused_remap``.  ``traffic`` is the HBM traffic of one launch from the rocprofv3
FETCH_SIZE / WRITE_SIZE summary under profiles/, reported when that summary was taken on the same
configuration (tree depth, tie policy, presort).  ``cpu_baseline`` times the oracle (the C++
restatement of the reference CPU search, OpenMP over queries) on bounded samples of the same
queries, on the same tree and on the reference-default-depth tree, with all granted host threads
and with one.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
TRAFFIC_SUMMARY = os.path.join(ROOT, "profiles", "knn_tree_plain_gpu_traffic.txt")


def _measured_counters(levels, tie, presort, variant):
    """Counter figures of one traversal launch from the committed summary (key: value lines), when it
    was taken on the same tree depth, tie policy, presort and launch variant; else {}."""
    try:
        with open(TRAFFIC_SUMMARY) as f:
            kv = dict(line.split(":", 1) for line in f if ":" in line and not line.startswith("#"))
    except OSError:
        return {}
    kv = {k.strip(): v.strip() for k, v in kv.items()}
    if (kv.get("levels") != str(levels) or kv.get("tie_break") != tie or kv.get("presort") != presort
            or kv.get("variant", "0") != str(variant)):
        return {}
    return kv


def _events(torch_knn, stage, fn, reps):
    """Average milliseconds of one call of `fn` in the library's HIP-event bracket `stage`."""
    torch_knn.profile_reset()
    torch_knn.profile_enable(True)
    for _ in range(reps):
        fn()
    torch_knn.profile_enable(False)
    ms, n = torch_knn.profile_read(stage)
    torch_knn.profile_reset()
    return ms / max(n, 1)


def _valu_issue_rate(torch, torch_knn, dev, mode, reps=3):
    """VALU wave instructions per second of the library's issue probe (tk_valu_probe_gpu: 8
    independent chains of 32-bit VALU instructions per lane; mode 0 multiply/add, mode 2
    compare/select/add) over 128 wavefronts per CU."""
    threads = 256 * 128 * 64
    out = torch.empty(threads, dtype=torch.float32, device=dev)
    stream = torch.cuda.current_stream(dev).cuda_stream
    n = torch_knn.valu_probe(out.data_ptr(), threads, 2048, mode, stream)
    torch.cuda.synchronize()
    ms = _events(torch_knn, "valu_probe",
                 lambda: torch_knn.valu_probe(out.data_ptr(), threads, 2048, mode, stream), reps)
    del out
    return n / (ms * 1e-3)


def _achievable_gbs(torch, torch_knn, dev, mib=1024, reps=5):
    """Rate of the library's streaming copy (tk_copy_gpu) over `mib` MiB: bytes read + written per second."""
    src = torch.empty(mib << 18, dtype=torch.float32, device=dev).fill_(1.0)
    dst = torch.empty_like(src)
    stream = torch.cuda.current_stream(dev).cuda_stream
    nbytes = src.numel() * 4
    torch_knn.copy_gpu(dst.data_ptr(), src.data_ptr(), nbytes, stream)
    torch.cuda.synchronize()
    ms = _events(torch_knn, "copy", lambda: torch_knn.copy_gpu(dst.data_ptr(), src.data_ptr(), nbytes, stream), reps)
    del src, dst
    return 2.0 * nbytes / (ms * 1e-3) / 1e9


def _host_facts():
    """nproc and the CPU model (lscpu) of this host."""
    import subprocess
    facts = {"nproc": os.cpu_count()}
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        kv = {a.strip(): b.strip() for a, b in (ln.split(":", 1) for ln in out.splitlines() if ":" in ln)}
        facts.update({"model": kv.get("Model name"), "sockets": kv.get("Socket(s)"),
                      "cores_per_socket": kv.get("Core(s) per socket"), "threads_per_core": kv.get("Thread(s) per core"),
                      "lscpu_cpus": kv.get("CPU(s)")})
    except Exception as e:  # lscpu missing: record why
        facts["lscpu"] = "unavailable: %s" % e
    try:
        facts["sched_affinity"] = len(os.sched_getaffinity(0))
    except Exception:
        pass
    return facts


def _cpu_rate(np, oracle_tree, sample, k, threads, runs):
    """Queries/s of the oracle search on `sample`: (median, mean) over `runs` timed runs after one warm-up."""
    oracle_tree.search(sample[:1024], k, nthreads=threads)
    rates = []
    for _ in range(runs):
        t0 = time.perf_counter()
        oracle_tree.search(sample, k, nthreads=threads)
        rates.append(sample.shape[0] / (time.perf_counter() - t0))
    return float(np.median(rates)), float(np.mean(rates))


def _extra_bruteforce(np, torch, torch_knn, dev, w):
    """Brute-force kernel on the 65,536 x 65,536 cross-check subset of plain_gpu, k=16."""
    n, k = 65536, 16
    half = w["halves"][0]
    ref = torch.from_numpy(np.ascontiguousarray(w["ref"][:n])).to(dev)
    qn = np.ascontiguousarray(np.concatenate([w["queries"][:n // 2], w["queries"][half:half + n // 2]]))
    q = torch.from_numpy(qn).to(dev)
    d = torch.empty((n, k), dtype=torch.float32, device=dev)
    i = torch.empty((n, k), dtype=torch.int32, device=dev)
    stream = torch.cuda.current_stream(dev).cuda_stream

    def run():
        torch_knn.knn_bruteforce(ref.data_ptr(), n, q.data_ptr(), n, 3, "float32", k, d.data_ptr(), i.data_ptr(),
                                 stream)

    run()
    torch.cuda.synchronize()
    ms = _events(torch_knn, "bruteforce", run, 5)
    flops = 8.0 * n * n  # 3 sub + 3 mul + 2 add per (query, reference) pair
    # lists beyond 32 entries (held in LDS): the same subset at k = 128
    k2 = 128
    d2 = torch.empty((n, k2), dtype=torch.float32, device=dev)
    i2 = torch.empty((n, k2), dtype=torch.int32, device=dev)

    def run128():
        torch_knn.knn_bruteforce(ref.data_ptr(), n, q.data_ptr(), n, 3, "float32", k2, d2.data_ptr(), i2.data_ptr(),
                                 stream)

    run128()
    torch.cuda.synchronize()
    ms128 = _events(torch_knn, "bruteforce", run128, 3)
    return {"workload": "plain_gpu refs[:65536] x 65,536 queries (32,768 per half), k=16",
            "ms": ms, "tflops": flops / (ms * 1e-3) / 1e12, "peak_tflops": 157.3,
            "frac": flops / (ms * 1e-3) / 1e12 / 157.3,
            "k128": {"ms": ms128, "tflops": flops / (ms128 * 1e-3) / 1e12,
                     "frac": flops / (ms128 * 1e-3) / 1e12 / 157.3}}


def _extra_fwd_bwd(np, torch, torch_knn, tk, workloads, dev, steps):
    """Differentiation workload: query + loss sum(w * dists) + gradients w.r.t. both clouds."""
    wd = workloads.diff_sphere()
    k = wd["k"]
    ref = torch.from_numpy(wd["ref"]).to(dev).requires_grad_(True)
    q = torch.from_numpy(wd["queries"]).to(dev).requires_grad_(True)
    wts = torch.from_numpy(wd["weights"]).to(dev)
    tree = tk.build_kd_tree(ref, device=dev, levels="auto")

    def step():
        dists, _ = tree.query(q, nr_nns_searches=k)
        loss = (wts * dists).sum()
        return torch.autograd.grad(loss, [q, ref])

    for _ in range(2):
        step()
    torch.cuda.synchronize()
    torch_knn.profile_reset()
    torch_knn.profile_enable(True)
    t0 = time.perf_counter()
    for _ in range(steps):
        step()
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    torch_knn.profile_enable(False)
    bw_ms, bw_n = torch_knn.profile_read("backward")
    tr_ms, tr_n = torch_knn.profile_read("traverse")
    torch_knn.profile_reset()
    nq, n = q.shape[0], ref.shape[0]
    bw_avg = bw_ms / max(bw_n, 1)
    # backward bytes: per query g (4k) + int64 idx (8k) + q (12) + dq (12); per reference point r (12) + dr (12)
    bw_bytes = nq * (4 * k + 8 * k + 24) + n * 24
    return {"workload": "diff_sphere: d=3 fp32, 1,000,000 refs on a perturbed sphere, 1,000,000 queries, k=8, "
                        "loss sum(w*dists), grads w.r.t. queries and refs",
            "levels": tree.levels, "fwd_bwd_per_s": nq * steps / el, "ms_per_step": el * 1e3 / steps,
            "traverse_ms": tr_ms / max(tr_n, 1), "backward_ms": bw_avg,
            "backward_gbs": bw_bytes / (bw_avg * 1e-3) / 1e9, "backward_bytes": bw_bytes}


def _extra_default_levels(torch, tk, dev, ref, q, k, res_d, res_i, steps):
    """The plain_gpu query on a tree of the reference's default depth (nn_distance.py:146-147)."""
    tree = tk.build_kd_tree(torch.from_numpy(ref), device=dev, levels=None)
    tree.query(q, nr_nns_searches=k, result_dists=res_d, result_idx=res_i)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        tree.query(q, nr_nns_searches=k, result_dists=res_d, result_idx=res_i)
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    return {"levels": tree.levels, "queries_per_s": q.shape[0] * steps / el, "ms_per_step": el * 1e3 / steps}


def _extra_lower_index(torch, tree, q, k, res_d, res_i, steps):
    """The plain_gpu query on the same tree with exact distance ties ordered by the lower original
    index (tie_break="lower_index")."""
    policy = tree.tie_break
    tree.tie_break = "lower_index"
    try:
        tree.query(q, nr_nns_searches=k, result_dists=res_d, result_idx=res_i)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            tree.query(q, nr_nns_searches=k, result_dists=res_d, result_idx=res_i)
        torch.cuda.synchronize()
        el = time.perf_counter() - t0
    finally:
        tree.tie_break = policy
    return {"tie_break": "lower_index", "levels": tree.levels, "queries_per_s": q.shape[0] * steps / el,
            "ms_per_step": el * 1e3 / steps}


def _extra_stress(np, torch, torch_knn, tk, workloads, dev, reps):
    """Higher-dimension and fp64 stress: d=8 fwd + bwd at k=32, and the d=3 fp64 forward at k=16."""
    out = {}
    w = workloads.d8_mixture()
    ref = torch.from_numpy(w["ref"]).to(dev).requires_grad_(True)
    q = torch.from_numpy(w["queries"]).to(dev).requires_grad_(True)
    wt = torch.rand((q.shape[0], w["k"]), device=dev, generator=torch.Generator(device=dev).manual_seed(6))
    tree = tk.build_kd_tree(ref, device=dev, levels="auto")

    def step():
        d, _ = tree.query(q, nr_nns_searches=w["k"])
        return torch.autograd.grad((wt * d).sum(), [q, ref])

    step()
    torch.cuda.synchronize()
    torch_knn.profile_reset()
    torch_knn.profile_enable(True)
    t0 = time.perf_counter()
    for _ in range(reps):
        step()
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    torch_knn.profile_enable(False)
    tr, n = torch_knn.profile_read("traverse")
    bw, m = torch_knn.profile_read("backward")
    torch_knn.profile_reset()
    # work of the search: points scanned per query (host search on a sample; the same pruning
    # decisions as the GPU's), and the distance arithmetic it implies, 3d - 1 = 23 flops per point
    sample = np.ascontiguousarray(w["queries"][:1 << 15])
    leaves, points = tree.kdtree.search_stats(sample, w["k"], 16)
    ppq = points / sample.shape[0]
    tr_ms = tr / max(n, 1)
    useful = 23.0 * ppq * q.shape[0] / (tr_ms * 1e-3) / 1e12
    out["d8_fwd_bwd"] = {"workload": "d8_mixture: d=8 fp32, 2,000,000 refs and queries from a 64-component "
                                     "anisotropic mixture, k=32, loss sum(w*dists), grads w.r.t. queries and refs",
                         "levels": tree.levels, "fwd_bwd_per_s": q.shape[0] * reps / el,
                         "ms_per_step": el * 1e3 / reps, "traverse_ms": tr_ms, "backward_ms": bw / max(m, 1),
                         "leaves_per_query": leaves / sample.shape[0], "points_per_query": ppq,
                         "distance_tflops": useful, "distance_frac_of_fp32_peak": useful / 157.3,
                         "work_sample": "first %d queries, host search" % sample.shape[0]}
    del tree, ref, q, wt
    w = workloads.fp64_normal()
    tree = tk.build_kd_tree(torch.from_numpy(w["ref"]), device=dev, levels="auto")
    q = torch.from_numpy(w["queries"]).to(dev)
    tree.query(q, nr_nns_searches=w["k"])
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        tree.query(q, nr_nns_searches=w["k"])
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    out["fp64_fwd"] = {"workload": "fp64_normal: d=3 fp64, 500,000 refs and queries N(0,1)*1e3, k=16",
                       "levels": tree.levels, "queries_per_s": q.shape[0] * reps / el, "ms_per_step": el * 1e3 / reps}
    return out


def _extra_batch(np, torch, torch_knn, tk, workloads, dev, reps):
    """Batching workload: 256 clouds, fwd + bwd, batched graph replay vs a per-cloud loop of the same kernels."""
    clouds = workloads.batch_clouds()
    refs = [c["ref"] for c in clouds]
    qs = [c["queries"] for c in clouds]
    ks = [c["k"] for c in clouds]
    rng = np.random.default_rng(5)
    ws = [rng.random((q.shape[0], k)).astype(np.float32) for q, k in zip(qs, ks)]
    sum_q = sum(q.shape[0] for q in qs)
    sum_n = sum(r.shape[0] for r in refs)
    stream = torch.cuda.current_stream(dev).cuda_stream

    t0 = time.perf_counter()
    batch = tk.build_kd_tree_batch(refs, device=dev)
    step = tk.BatchStep(batch, [torch.from_numpy(q) for q in qs], ks, [torch.from_numpy(w) for w in ws])
    torch.cuda.synchronize()
    plan_s = time.perf_counter() - t0
    step.replay()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        step.replay()
    torch.cuda.synchronize()
    batched_ms = (time.perf_counter() - t0) * 1e3 / reps

    # the same clouds as separate trees (same depths), one plain query + backward call per cloud
    trees = [tk.build_kd_tree(torch.from_numpy(r), device=dev, levels=lv) for r, lv in zip(refs, batch.levels)]
    bufs = []
    for r, q, k, w in zip(refs, qs, ks, ws):
        qd = torch.from_numpy(q).to(dev)
        rd = torch.from_numpy(r).to(dev)
        bufs.append((qd, rd, k, torch.empty((q.shape[0], k), device=dev),
                     torch.empty((q.shape[0], k), dtype=torch.int32, device=dev),
                     torch.empty((q.shape[0], k), dtype=torch.int64, device=dev), torch.from_numpy(w).to(dev),
                     torch.empty_like(qd), torch.empty_like(rd)))

    def loop():
        for tree, (qd, rd, k, d, i, o, w, dq, dr) in zip(trees, bufs):
            tree.kdtree.query(qd.data_ptr(), qd.shape[0], k, d.data_ptr(), i.data_ptr(), orig_idx_ptr=o.data_ptr(),
                              stream=stream)
            torch_knn.knn_backward(qd.data_ptr(), rd.data_ptr(), qd.shape[0], rd.shape[0], 3, "float32", k,
                                   o.data_ptr(), w.data_ptr(), dq.data_ptr(), dr.data_ptr(), stream)

    loop()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        loop()
    torch.cuda.synchronize()
    loop_ms = (time.perf_counter() - t0) * 1e3 / reps
    # the same step with the other traversal launch shape (persistent work queues / one launch per k bucket)
    batch_pb = tk.build_kd_tree_batch(refs, device=dev, levels=batch.levels, persistent=not batch.persistent)
    step_pb = tk.BatchStep(batch_pb, [torch.from_numpy(q) for q in qs], ks, [torch.from_numpy(w) for w in ws])
    step_pb.replay()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        step_pb.replay()
    torch.cuda.synchronize()
    per_bucket_ms = (time.perf_counter() - t0) * 1e3 / reps
    same_pb = all(torch.equal(a, b) for a, b in ((step.dist, step_pb.dist), (step.idx, step_pb.idx),
                                                (step.grad_queries, step_pb.grad_queries),
                                                (step.grad_refs, step_pb.grad_refs)))
    del step_pb, batch_pb
    same = (torch.equal(step.dist, torch.cat([b[3].reshape(-1) for b in bufs]))
            and torch.equal(step.orig, torch.cat([b[5].reshape(-1) for b in bufs]))
            and torch.equal(step.grad_queries, torch.cat([b[7] for b in bufs]))
            and torch.equal(step.grad_refs, torch.cat([b[8] for b in bufs])))
    # algorithmic bytes per fwd+bwd: per query 12 + 8k (fwd: q, dist + idx) and 8k + 24 (bwd: g, idx, q, dq);
    # per reference point 40 (tree records 16, r 12, dr 12)
    alg = sum(q.shape[0] * (12 + 8 * k + 8 * k + 24) for q, k in zip(qs, ks)) + 40 * sum_n
    return {"workload": "batch_clouds: 256 d=3 fp32 clouds (seed 3), N log-uniform [1e3, 5e5], Q = U(0.5,2) N, "
                        "k in {1,4,16}; fwd + bwd, loss sum(w * dists)",
            "clouds": len(clouds), "sum_refs": sum_n, "sum_queries": sum_q, "plan_and_capture_s": plan_s,
            "batched_graph_ms": batched_ms, "batched_fwd_bwd_per_s": sum_q / (batched_ms * 1e-3),
            "traversal_launch": "persistent" if batch.persistent else "per k bucket",
            "other_launch_shape": "per k bucket" if batch.persistent else "persistent",
            "other_launch_shape_graph_ms": per_bucket_ms, "other_launch_shape_same_results": bool(same_pb),
            "per_cloud_loop_ms": loop_ms, "per_cloud_fwd_bwd_per_s": sum_q / (loop_ms * 1e-3),
            "batched_vs_per_cloud": loop_ms / batched_ms, "same_results": bool(same),
            "batched_algorithmic_gbs": alg / (batched_ms * 1e-3) / 1e9}


def _extra_tree_build(np, torch, tk, workloads, dev, w, cpu_build_s, reps):
    """Tree build on the GPU (builder="gpu") against the host build: the 4M-point plain tree and the
    256 trees of the batching workload (one GPU build for all of them)."""
    ref = torch.from_numpy(w["ref"])
    if cpu_build_s is None:
        t0 = time.perf_counter()
        tk.build_kd_tree(ref, device=dev, levels=17)
        cpu_build_s = time.perf_counter() - t0
    tk.build_kd_tree(ref, device=dev, levels=17, builder="gpu")  # first call: allocation warm-up
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        tk.build_kd_tree(ref, device=dev, levels=17, builder="gpu")
    torch.cuda.synchronize()
    gpu_s = (time.perf_counter() - t0) / reps
    refs = [torch.from_numpy(c["ref"]) for c in workloads.batch_clouds()]
    t0 = time.perf_counter()
    cpu_batch = tk.build_kd_tree_batch(refs, device=dev)
    cpu_batch_s = time.perf_counter() - t0
    levels = cpu_batch.levels
    del cpu_batch
    tk.build_kd_tree_batch(refs, device=dev, levels=levels, builder="gpu")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        tk.build_kd_tree_batch(refs, device=dev, levels=levels, builder="gpu")
    torch.cuda.synchronize()
    gpu_batch_s = (time.perf_counter() - t0) / reps
    return {"plain_gpu_tree": {"points": int(ref.shape[0]), "levels": 17, "host_build_s": cpu_build_s,
                               "gpu_build_s": gpu_s, "speedup": cpu_build_s / gpu_s},
            "batch_clouds_trees": {"trees": len(refs), "points": int(sum(r.shape[0] for r in refs)),
                                   "host_build_s": cpu_batch_s, "gpu_build_s": gpu_batch_s,
                                   "speedup": cpu_batch_s / gpu_batch_s},
            "note": "host build = the reference's algorithm, OpenMP over subtrees (one tree) or over trees (256); "
                    "GPU build = one level-by-level radix-sort build, host copies and upload included"}


def _parse():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="plain_gpu", choices=["plain_gpu"])
    ap.add_argument("--levels", default="auto", help="tree depth: 'auto', 'default' or an int")
    ap.add_argument("--tie", default="reference", choices=["reference", "lower_index"])
    ap.add_argument("--presort", default="auto", choices=["auto", "morton", "leaf", "off"])
    ap.add_argument("--cpu-sample", type=int, default=1 << 22, help="queries timed on the CPU (0 = skip)")
    ap.add_argument("--cpu-threads", type=int, default=0, help="0 = min(16, cores)")
    ap.add_argument("--section", default="all",
                    choices=["all", "batching", "fwd_bwd", "stress", "bruteforce", "build"],
                    help="run one secondary measurement alone (for profiling) and print its JSON; 'all' = the "
                         "benchmark line")
    ap.add_argument("--no-extras", action="store_true",
                    help="skip the secondary measurements (brute force, lower-index policy, reference default depth, "
                         "fwd+bwd, d=8 and fp64 stress, batching, GPU tree build)")
    ap.add_argument("--calibrate-copy-mib", type=int, default=0,
                    help="before the timed steps, run one device-to-device copy of this many MiB (a kernel of known "
                         "HBM traffic, for scaling rocprofv3 FETCH_SIZE/WRITE_SIZE counters)")
    return ap.parse_args()


def main():
    args = _parse()
    if args.gpus != 1:
        print("bench.py measures one GPU: there is no multi-GPU path (each GPU would hold the whole tree and a "
              "slice of the queries, with no exchange between GPUs). Run with --gpus 1.", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    import workloads
    import torch_kdtree_amd as tk
    from torch_kdtree_amd import torch_knn

    def log(msg):
        print("[bench] " + msg, file=sys.stderr, flush=True)

    dev = torch.device("cuda", 0)
    if args.section != "all":
        reps = max(1, args.steps)
        if args.section == "batching":
            out = _extra_batch(np, torch, torch_knn, tk, workloads, dev, reps)
        elif args.section == "fwd_bwd":
            out = _extra_fwd_bwd(np, torch, torch_knn, tk, workloads, dev, reps)
        elif args.section == "stress":
            out = _extra_stress(np, torch, torch_knn, tk, workloads, dev, reps)
        elif args.section == "build":
            out = _extra_tree_build(np, torch, tk, workloads, dev, workloads.plain_gpu(), None, reps)
        else:
            out = _extra_bruteforce(np, torch, torch_knn, dev, workloads.plain_gpu())
        print(json.dumps({"section": args.section, "steps": reps, **out}), flush=True)
        return
    log("generating plain_gpu workload")
    w = workloads.plain_gpu()
    ref, qry, k = w["ref"], w["queries"], w["k"]
    n, d = ref.shape
    nq = qry.shape[0]
    levels = {"auto": "auto", "default": None}.get(args.levels, None if args.levels == "default" else args.levels)
    if isinstance(levels, str) and levels != "auto":
        levels = int(levels)
    t0 = time.perf_counter()
    tree = tk.build_kd_tree(torch.from_numpy(ref), device=dev, levels=levels, tie_break=args.tie,
                            presort=args.presort)
    build_s = time.perf_counter() - t0
    log("tree built: levels=%d in %.2f s" % (tree.levels, build_s))

    q = torch.from_numpy(qry).to(dev)
    res_d = torch.empty((nq, k), dtype=torch.float32, device=dev)
    res_i = torch.empty((nq, k), dtype=torch.int32, device=dev)
    for _ in range(args.warmup):
        tree.query(q, nr_nns_searches=k, result_dists=res_d, result_idx=res_i)
    torch.cuda.synchronize()
    if args.calibrate_copy_mib > 0:
        src = torch.empty(args.calibrate_copy_mib << 18, dtype=torch.float32, device=dev).fill_(1.0)
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        torch_knn.copy_gpu(dst.data_ptr(), src.data_ptr(), src.numel() * 4, torch.cuda.current_stream(dev).cuda_stream)
        torch.cuda.synchronize()
        del src, dst
        log("calibration copy of %d MiB done (stream_copy_kernel)" % args.calibrate_copy_mib)
    log("timing %d steps" % args.steps)

    torch_knn.profile_reset()
    torch_knn.profile_enable(True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        tree.query(q, nr_nns_searches=k, result_dists=res_d, result_idx=res_i)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    torch_knn.profile_enable(False)
    trav_ms, trav_n = torch_knn.profile_read("traverse")
    pre_ms, pre_n = torch_knn.profile_read("presort")
    torch_knn.profile_reset()

    ms_per_step = elapsed * 1e3 / args.steps
    value = nq * args.steps / elapsed
    # the traversal kernel that ran: the built-in launch shape, or a tuned variant of it
    variant = tree.kdtree.get_variant()
    if variant > 0:
        vi = torch_knn.tree_variants()[variant]
        kernel_name = "knn_tree_kernel_v<float,3,16,%s> (chunk %d, %d waves per SIMD, block %d)" % (
            "true" if args.tie == "lower_index" else "false", vi["chunk"], vi["waves"], vi["block"])
    else:
        kernel_name = "knn_tree_kernel<float,3,16,%s>" % ("true" if args.tie == "lower_index" else "false")

    # Algorithmic bytes of one traversal launch: every query read once (d*4 B) and its k results
    # written once (4 B distance + 4 B int32 index), plus the tree read once (N*(d*4+4) record
    # bytes, 5 B per node: split value and axis).  The kernel also writes the int64 original index
    # (8 B per result), which the reference computes afterwards in torch: counted only in the
    # fused-remap figure.
    n_nodes = (1 << tree.levels) - 1
    tree_bytes = n * (d * 4 + 4) + n_nodes * 5
    bytes_per_query = d * 4 + k * (4 + 4)
    launch_bytes = nq * bytes_per_query + tree_bytes
    fused_bytes = launch_bytes + nq * k * 8
    trav_avg_ms = trav_ms / max(trav_n, 1)
    achieved = launch_bytes / (trav_avg_ms * 1e-3) / 1e9
    log("measuring the achievable bandwidth (streaming copy, 1 GiB)")
    achievable = _achievable_gbs(torch, torch_knn, dev)
    log("measuring the VALU issue rate (probe kernel)")
    valu_peak = _valu_issue_rate(torch, torch_knn, dev, 0)
    valu_mix = _valu_issue_rate(torch, torch_knn, dev, 2)
    counters = _measured_counters(tree.levels, args.tie, args.presort, variant)
    traffic = int(counters["traffic_bytes_per_launch"]) if "traffic_bytes_per_launch" in counters else None
    # the traversal's own bound: VALU issue.  Its instruction count per launch comes from the committed
    # SQ_INSTS_VALU summary of the same configuration; the ceiling from the probe, measured here.
    valu = {"probe_wave_insts_per_s": valu_peak, "probe_cmp_select_wave_insts_per_s": valu_mix,
            "probe": "tk_valu_probe_gpu, 128 waves per CU: mode 0 = 8 independent v_mul_f32/v_add_f32 chains "
                     "per lane (4-cycle instructions), mode 2 = v_cmp/v_cndmask/v_add chains"}
    if "valu_insts_per_launch" in counters:
        n_valu = float(counters["valu_insts_per_launch"])
        valu.update({"insts_per_launch": n_valu, "wave_insts_per_s": n_valu / (trav_avg_ms * 1e-3),
                     "frac_of_probe": n_valu / (trav_avg_ms * 1e-3) / valu_peak,
                     "frac_of_cmp_select_probe": n_valu / (trav_avg_ms * 1e-3) / valu_mix,
                     "lanes_active_per_inst": float(counters.get("valu_lanes_active", "nan"))})
    # work of the search: leaves and points scanned per query (host search of a balanced sample, the
    # same pruning decisions as the GPU's), and the distance arithmetic it implies (8 flops per point)
    half = w["halves"][0]
    wsample = np.ascontiguousarray(np.concatenate([qry[:1 << 17], qry[half:half + (1 << 17)]]))
    w_leaves, w_points = tree.kdtree.search_stats(wsample, k, 16)
    ppq = w_points / wsample.shape[0]
    work = {"leaves_per_query": w_leaves / wsample.shape[0], "points_per_query": ppq,
            "distance_tflops": 8.0 * ppq * nq / (trav_avg_ms * 1e-3) / 1e12,
            "sample": "first 131,072 queries of each half, host search (same pruning decisions as the GPU)"}
    work["distance_frac_of_fp32_peak"] = work["distance_tflops"] / 157.3

    cpu = None
    if args.cpu_sample > 0:
        log("timing the CPU baseline")
        import oracle

        threads = args.cpu_threads or min(16, os.cpu_count() or 1)
        half = w["halves"][0]

        def balanced(count):
            s = min(count // 2, half)
            return np.ascontiguousarray(np.concatenate([qry[:s], qry[half:half + s]]))

        sample = balanced(args.cpu_sample)
        small = balanced(max(2048, args.cpu_sample // 16))
        ot = oracle.tree_from_product(tree)
        med, mean = _cpu_rate(np, ot, sample, k, threads, 5)
        med1, mean1 = _cpu_rate(np, ot, small, k, 1, 3)
        del ot
        log("timing the CPU baseline on the reference-default-depth tree")
        t13 = tk.build_kd_tree(torch.from_numpy(ref), device="cpu", levels=None)
        ot13 = oracle.tree_from_product(t13)
        s13 = balanced(max(4096, args.cpu_sample // 4))
        med13, mean13 = _cpu_rate(np, ot13, s13, k, threads, 3)
        s13_1 = balanced(max(2048, args.cpu_sample // 64))
        med13_1, mean13_1 = _cpu_rate(np, ot13, s13_1, k, 1, 3)
        del ot13, t13
        cpu = {"value": med, "unit": "queries/s", "cores": threads, "kind": "port",
               "sample": "first %d queries of each half of the 16M-query set (%d total), same tree (levels=%d), k=%d; "
                         "oracle = C++ restatement of KDTreeKNNSearch, OpenMP over queries, %d threads; median of 5 "
                         "runs after a warm-up" % (sample.shape[0] // 2, sample.shape[0], tree.levels, k, threads),
               "host": _host_facts(),
               "runs": {"levels_%d_threads_%d" % (tree.levels, threads): {"queries": sample.shape[0], "median": med,
                                                                        "mean": mean},
                        "levels_%d_threads_1" % tree.levels: {"queries": small.shape[0], "median": med1, "mean": mean1},
                        "levels_13_threads_%d" % threads: {"queries": s13.shape[0], "median": med13, "mean": mean13},
                        "levels_13_threads_1": {"queries": s13_1.shape[0], "median": med13_1, "mean": mean13_1}},
               "note": "levels 13 is the reference's default depth for 4M points (nn_distance.py:146-147); the GPU "
                       "step above runs at levels %d" % tree.levels}

    extras = None
    if not args.no_extras:
        extras = {}
        log("secondary: brute force")
        extras["bruteforce"] = _extra_bruteforce(np, torch, torch_knn, dev, w)
        log("secondary: plain_gpu with the lower-index tie policy")
        extras["lower_index_policy"] = _extra_lower_index(torch, tree, q, k, res_d, res_i, 3)
        log("secondary: plain_gpu at the reference's default depth")
        extras["reference_default_levels"] = _extra_default_levels(torch, tk, dev, ref, q, k, res_d, res_i, 3)
        del res_d, res_i, q
        log("secondary: differentiation fwd+bwd")
        extras["fwd_bwd"] = _extra_fwd_bwd(np, torch, torch_knn, tk, workloads, dev, 10)
        log("secondary: d=8 and fp64 stress")
        extras.update(_extra_stress(np, torch, torch_knn, tk, workloads, dev, 5))
        log("secondary: batching (256 clouds)")
        extras["batching"] = _extra_batch(np, torch, torch_knn, tk, workloads, dev, 3)
        extras["batched_vs_per_cloud"] = extras["batching"]["batched_vs_per_cloud"]
        log("secondary: GPU tree build")
        extras["tree_build"] = _extra_tree_build(np, torch, tk, workloads, dev, w, build_s, 2)

    line = {
        "metric": "KNN queries/s at d=3,k=16 fp32 (+fwd+bwd/s); batched vs per-cloud, same GPU",
        "value": value,
        "unit": "queries/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (numpy PCG64 seed 1): refs U[0,1]^3; queries half U[-0.1,1.1]^3, half 32-GMM sigma 0.02",
        "config": {"workload": "plain_gpu: d=3 fp32, N=4,000,000 refs, Q=16,000,000 queries, k=16",
                   "levels": tree.levels, "tie_break": args.tie, "presort": args.presort,
                   "tree_build_s": round(build_s, 3),
                   "presort_ms_per_step": pre_ms / max(pre_n, 1), "traverse_ms_per_step": trav_avg_ms},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS,
                     "traffic": traffic,
                     "kernel": kernel_name, "launch_variant": variant,
                     "algorithmic_bytes_per_launch": launch_bytes,
                     "bytes_per_query": bytes_per_query, "tree_bytes": tree_bytes,
                     "traverse_ms": trav_avg_ms,
                     "achievable_gbs": achievable, "frac_of_achievable": achieved / achievable,
                     "achieved_with_fused_remap": fused_bytes / (trav_avg_ms * 1e-3) / 1e9,
                     "search_work": work, "valu_issue": valu},
        "cpu_baseline": cpu,
        "secondary": extras,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
