#!/usr/bin/env python
"""Batched ARAP throughput on one MI355X.

Workload: 512 disk-topology islands with log-uniform triangle counts in
[200, 40000] (Delaunay of Poisson-disk points on the unit disk lifted by six
Gaussian bumps, seed 2), LSCM initialisation, then 10 ARAP iterations per
island with every global step solved to relative residual 1e-6.

One step = one replay of the captured graph (prologue + 10 ARAP iterations
over all islands) from the LSCM layout.  value = sum(t_i) * 10 / step time in
Mtri*iter/s.  The roofline figure uses the algorithmic byte model
    B_i = 124 t_i + 72 n_i + 8 nnz_i + k_i (100 n_i + 8 nnz_i)
per island and ARAP iteration (k_i = CG iterations executed, nnz_i with the
diagonal).  The CPU baseline is the float64 CPU restatement of the reference
path (SuperLU) on a bounded sample of the same islands, over all host cores
(worker processes, up to 16) and on one core.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--iters", type=int, default=10)
    ap.add_argument("--islands", type=int, default=512)
    ap.add_argument("--cg-tol", type=float, default=1e-6)
    ap.add_argument("--precision", default="mixed", choices=["mixed", "f32", "f64"])
    ap.add_argument("--cpu-sample", type=int, default=160, help="islands in the CPU baseline sample (0 = skip)")
    ap.add_argument("--verdict-sample", type=int, default=16,
                    help="islands of the CPU sample whose GPU and oracle layouts are validated (0 = skip)")
    ap.add_argument("--no-graph", action="store_true")
    ap.add_argument("--no-tune", dest="tune", action="store_false",
                    help="keep the size-based schedule (no re-planning from measured iterations)")
    ap.add_argument("--no-plain", dest="plain", action="store_false",
                    help="skip the plain per-island GPU baseline")
    ap.add_argument("--workload", default="disks", choices=["disks", "grid", "heightfield", "needles", "validate", "batch", "hier", "small"],
                    help="disks: the throughput workload; grid: the single 7,938-triangle grid island "
                         "(use --iters 20 --cg-tol 1e-8); heightfield: 3,988 islands of 50-1,946 triangles "
                         "cut from a 2M-triangle heightfield (use --iters 15); needles: the precision study on "
                         "256 islands of 5k-60k triangles with 10%% needle triangles (one JSON line with every "
                         "precision variant); validate: validate_uvs of the 512 solved islands packed "
                         "into one atlas, GPU overlap test against the CPU restatement on a sample; batch: "
                         "gpu.Batch.arap(8) of the 512 islands as one batch (one coupled 40-iteration CG "
                         "per global step), against the oracle's Batch on a sample; hier: hier.Solver.relax "
                         "of the 512 islands as the charts of one subset, against the oracle on a sample; "
                         "small: charts.flatten of 40-90-face charts from 8 threads (per-call param, "
                         "FlattenQueue, flatten_many) against the oracle on 1 and 8 cores")
    ap.add_argument("--charts", type=int, default=2000, help="charts of the small workload")
    ap.add_argument("--gen-workers", type=int, default=None, metavar="N",
                    help="worker processes that generate the islands (default: up to 8; use 1 under a "
                         "profiler that initialises the GPU before the program starts, since the workers "
                         "are forked)")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="only run a device-to-device copy of MIB MiB 5 times (HBM counter calibration)")
    return ap.parse_args()


def copy_probe(mib):
    """A plain copy of known size, for scaling FETCH_SIZE / WRITE_SIZE counters."""
    import torch

    n = mib * (1 << 20) // 4
    a = torch.ones(n, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    for _ in range(5):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_probe_bytes_each_way": n * 4, "dispatches": 5}), flush=True)


def kernel_source_digest():
    """SHA-256 (16 hex digits) of the library's kernel and host sources, so a
    counter summary can be matched to the code it measured."""
    import hashlib

    h = hashlib.sha256()
    csrc = ROOT / "ultimate_uv_artist_amd" / "csrc"
    for f in sorted(csrc.iterdir()):
        if f.suffix in (".hip", ".hpp", ".cpp", ".inc"):
            h.update(f.name.encode())
            h.update(f.read_bytes())
    return h.hexdigest()[:16]


TRAFFIC_JSON = "profiles/island_kernel_arap_512islands_traffic.json"


def measured_traffic(digest):
    """HBM bytes per ARAP-iteration launch from the committed rocprofv3 counter
    summary of this workload (FETCH_SIZE + WRITE_SIZE, scaled by a plain copy
    of known size), or None when no summary for this mesh and these kernel
    sources is committed.  Returns (bytes or None, where the figure is from)."""
    f = ROOT / TRAFFIC_JSON
    if not f.exists():
        return None, None
    d = json.loads(f.read_text())
    if d.get("mesh_sha256") != digest[:16]:
        return None, f"{TRAFFIC_JSON}: another mesh"
    src = kernel_source_digest()
    if d.get("kernel_src_sha16") != src:
        return None, f"{TRAFFIC_JSON}: measured on other kernel sources ({d.get('kernel_src_sha16')}, now {src})"
    return d.get("hbm_bytes_per_launch"), (f"{TRAFFIC_JSON} (rocprofv3 FETCH_SIZE + WRITE_SIZE of these kernel "
                                           f"sources, {src}, scaled by a plain copy)")


def unique_edge_count(T):
    a = np.concatenate([T[:, 0], T[:, 1], T[:, 2]])
    b = np.concatenate([T[:, 1], T[:, 2], T[:, 0]])
    n = int(T.max()) + 1
    return len(np.unique(np.minimum(a, b) * n + np.maximum(a, b)))


def orient_majority(U, plan, vo, to, T_local):
    """charts.py:407-410: mirror islands with more negative than positive triangles."""
    import torch

    sa = plan.signed_areas(U).cpu().numpy()
    flip = np.zeros(plan.n, bool)
    for i in range(len(vo) - 1):
        s = sa[to[i]:to[i + 1]]
        if (s < 0).sum() > (s > 0).sum():
            flip[vo[i]:vo[i + 1]] = True
    f = torch.as_tensor(flip, device=U.device)
    U[:, 1] = torch.where(f, -U[:, 1], U[:, 1])
    return U


def _one_thread():
    from threadpoolctl import threadpool_limits

    threadpool_limits(1)


def _oracle_refine(job):
    """One island of the CPU baseline: the oracle's ARAP from the given
    layout, timed inside the worker."""
    from oracle import param_cpu as oracle

    P, T, u0, iters = job
    c0 = time.perf_counter()
    ur, _ = oracle.refine(P, T, u0, iters=iters, tol=0.0)
    return ur, time.perf_counter() - c0


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown CPU"


def precision_study(args):
    """The three precision variants of the ARAP global step on ill-conditioned
    islands: CG iterations to the relative residual, time, and the UV error
    against the float64 oracle (SuperLU) on a sample of islands."""
    import torch

    from oracle import param_cpu as oracle
    from ultimate_uv_artist_amd import _lib, workloads
    from ultimate_uv_artist_amd.plan import IslandPlan

    dev = torch.device("cuda", 0)
    torch.cuda.set_device(dev)
    t0 = time.perf_counter()
    charts = workloads.sliver_islands(seed=3, n_islands=256, workers=args.gen_workers)
    tri = np.array([len(T) for _, T in charts])
    print(f"bench: {len(charts)} needle islands, {tri.sum()} triangles (generated in "
          f"{time.perf_counter() - t0:.1f}s)", flush=True)
    plan = IslandPlan(charts, dev=dev)
    U0 = torch.zeros(plan.n, 2, dtype=torch.float64, device=dev)
    plan.lscm(U0)
    torch.cuda.synchronize()
    vo = plan.vert_off
    order = np.argsort(tri)
    sample = order[np.linspace(0, len(order) - 1, 8).round().astype(int)]
    Uh0 = U0.cpu().numpy()
    ref = {}
    for i in sample:
        P, T = charts[i]
        ref[i], _ = oracle.refine(P, T, Uh0[vo[i]:vo[i + 1]], iters=args.iters, tol=0.0)
    print(f"bench: oracle layouts of {len(sample)} sample islands done", flush=True)
    out = {}
    for name, prec, cg_tol in [(n, p, t) for t in (args.cg_tol, 1e-8) for n, p in
                               (("mixed", _lib.UVA_PREC_MIXED), ("f32", _lib.UVA_PREC_F32),
                                ("f64", _lib.UVA_PREC_F64))]:
        U = U0.clone()
        name = f"{name}@{cg_tol:g}"

        print(f"bench: {name} ...", flush=True)
        # one eager step with HIP events around every launch (the graph
        # replay of the throughput workload adds nothing to the comparison)
        launch = plan.arap_profile(U, args.iters, line_search=True, mode=_lib.UVA_MODE_REFINE, tol=0.0,
                                   cg_tol=cg_tol, precision=prec)
        st = plan.stats()
        ms = float(launch.sum())
        Uh = U.cpu().numpy()
        err = []
        for i in sample:
            g, r = Uh[vo[i]:vo[i + 1]], ref[i]
            lo = r.min(0)
            err.append(float(np.abs(g - r).max() / (r.max(0) - lo).max()))
        kpi = st["cg_iters"] / np.maximum(st["arap_iters"], 1)
        out[name] = {"ms_per_step": round(ms, 2), "mtri_iter_per_s": round(float(tri.sum()) * args.iters / ms / 1e3, 2),
                     "cg_iters_per_global_step_mean": round(float(kpi.mean()), 1),
                     "cg_iters_per_global_step_p95": round(float(np.percentile(kpi, 95)), 1),
                     "converged_islands": int(st["converged"].sum()),
                     "max_rel_resid": float(st["rel_resid"].max()),
                     "uv_err_unit_box_per_sample_island": [round(e, 9) for e in err]}
        print(f"bench: {name}: {ms:.1f} ms/step, CG iters per global step mean {kpi.mean():.1f}, converged "
              f"{int(st['converged'].sum())}/{len(charts)}, max UV err vs FP64 oracle {max(err):.2e}", flush=True)
    print(json.dumps({"workload": f"{len(charts)} needle islands, t uniform [5000, 60000], 10% needles of aspect "
                                  f"U(50, 500), LSCM + {args.iters} ARAP, CG rel resid {args.cg_tol:g} and 1e-8",
                      "triangles": int(tri.sum()), "mesh_sha256": workloads.mesh_digest(charts)[:16],
                      "sample_islands_triangles": [int(tri[i]) for i in sample], "variants": out}), flush=True)


def validate_study(args):
    """validate.validate_uvs (gutter=False) of the 512 disk islands after LSCM
    + ARAP, packed into a k x k grid of the unit square (3.8M triangles):
    time of the whole check with the GPU overlap test, of the overlap kernel
    alone, and of the CPU restatement's overlap test on a sample of cells."""
    import torch

    from oracle import validate_cpu as vc
    from ultimate_uv_artist_amd import _lib, validate, workloads
    from ultimate_uv_artist_amd.plan import IslandPlan

    dev = torch.device("cuda", 0)
    torch.cuda.set_device(dev)
    charts = workloads.disk_islands(seed=2, n_islands=args.islands, workers=args.gen_workers)
    plan = IslandPlan(charts, dev=dev)
    U = torch.zeros(plan.n, 2, dtype=torch.float64, device=dev)
    plan.lscm(U)
    plan.arap(U, args.iters, line_search=True, mode=_lib.UVA_MODE_REFINE, tol=0.0, cg_tol=args.cg_tol)
    Uh = U.cpu().numpy()
    vo = plan.vert_off
    k = int(np.ceil(np.sqrt(len(charts))))
    Ps, Ts, Us, off = [], [], [], 0
    for i, (P, T) in enumerate(charts):
        u = Uh[vo[i]:vo[i + 1]]
        lo = u.min(0)
        u = (u - lo) / float((u.max(0) - lo).max()) * 0.9 / k + np.array([i % k, i // k]) / k
        Ps.append(P + 10.0 * i)
        Ts.append(T + off)
        Us.append(u)
        off += len(P)
    P, T, Ua = np.concatenate(Ps), np.concatenate(Ts), np.concatenate(Us)
    print(f"bench: atlas of {len(charts)} islands, {len(T)} triangles", flush=True)
    validate.validate_uvs(P, T, Ua, gutter=False)           # warm-up
    t0 = time.perf_counter()
    rep, prob = validate.validate_uvs(P, T, Ua, gutter=False)
    t_all = time.perf_counter() - t0
    U32 = np.asarray(Ua, np.float32).astype(np.float64)
    G = int(np.clip(np.sqrt(len(T)) / 2, 64, 4096))
    t0 = time.perf_counter()
    ov = validate.overlap_flags(U32, T, None, 1e-6, False, G)
    t_ov = time.perf_counter() - t0
    # the CPU restatement on the first 16 atlas cells' islands
    nsmp = min(16, len(charts))
    te = int(sum(len(T_) for _, T_ in charts[:nsmp]))
    t0 = time.perf_counter()
    ov_c = vc.overlaps(U32, T[:te], np.zeros(te, bool), 1e-6)
    t_cpu = time.perf_counter() - t0
    same = bool(np.array_equal(ov_c, ov[:te]))
    print(f"bench: validate_uvs {t_all * 1e3:.1f} ms (valid {rep['valid']}, islands {rep['islands']}); overlap test "
          f"{t_ov * 1e3:.1f} ms = {len(T) / t_ov / 1e6:.1f} Mtri/s on the GPU; CPU restatement {te / t_cpu / 1e6:.3f} "
          f"Mtri/s on {te} triangles (one core), same flags {same}", flush=True)
    print(json.dumps({"workload": f"validate_uvs(gutter=False) of {len(charts)} solved disk islands in one atlas",
                      "triangles": int(len(T)), "validate_ms": round(t_all * 1e3, 1),
                      "overlap_ms": round(t_ov * 1e3, 1), "overlap_mtri_per_s": round(len(T) / t_ov / 1e6, 2),
                      "cpu_restatement_mtri_per_s": round(te / t_cpu / 1e6, 4), "cpu_sample_triangles": te,
                      "flags_equal_on_sample": same, "report": rep}), flush=True)


def batch_study(args):
    """gpu.Batch on the 512 disk islands as one batch (3.83M triangles) from
    their LSCM layouts: orient, repair, then the timed Batch.arap(8,
    line_search=True) of the reference's defaults (one Jacobi-PCG of at most
    40 iterations over every chart per global step).  Reports Mtri*iter/s,
    the coupled CG's HBM roofline (uva_csr_pcg, HIP events around 40
    iterations), and the oracle's Batch with the same coupled CG on a sample
    of 16 islands (one core), whose GPU result it also checks."""
    import torch

    from oracle.batch_cpu import BatchOracle
    from ultimate_uv_artist_amd import gpu, workloads
    from ultimate_uv_artist_amd.plan import IslandPlan

    dev = torch.device("cuda", 0)
    torch.cuda.set_device(dev)
    charts = workloads.disk_islands(seed=2, n_islands=args.islands, workers=args.gen_workers)
    plan = IslandPlan(charts, dev=dev)
    U = torch.zeros(plan.n, 2, dtype=torch.float64, device=dev)
    plan.lscm(U)
    Uh = U.cpu().numpy()
    vo = plan.vert_off
    items = [(P, T, Uh[vo[i]:vo[i + 1]]) for i, (P, T) in enumerate(charts)]
    del plan
    nt = int(sum(len(T) for _, T in charts))
    t0 = time.perf_counter()
    b = gpu.Batch(items, dev=dev, dtype=torch.float64)
    b.orient_positive()
    b.repair_flips(80)
    torch.cuda.synchronize()
    t_setup = time.perf_counter() - t0
    U0 = b._U.clone()
    print(f"bench: batch of {len(charts)} charts, {nt} triangles, setup {t_setup:.2f} s", flush=True)
    b.arap(1, line_search=True)                      # warm-up
    times = []
    for k in range(max(args.steps, 1)):
        b._U.copy_(U0)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        b.arap(8, line_search=True)
        torch.cuda.synchronize()
        times.append(time.perf_counter() - t0)
        print(f"bench: Batch.arap(8) step {k}: {times[-1] * 1e3:.1f} ms", flush=True)
    t_step = float(np.median(times))
    # the coupled CG alone: 40 iterations on the global step's rhs, HIP events
    rhs = b._plan.local_rhs(U0, 1)
    x = U0.clone()
    s = torch.cuda.current_stream()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    cg_ms = []
    for _ in range(5):
        x.copy_(U0)
        e0.record(s)
        its = b._plan.pcg(rhs, x, 40, check_every=1, tol2=0.0, strict=True)
        e1.record(s)
        torch.cuda.synchronize()
        cg_ms.append(e0.elapsed_time(e1))
    cg_t = float(np.median(cg_ms)) / 1e3
    info = b._plan.info()
    n, nnz_off = int(info["n_vertices"]), int(info["nnz_true"])
    nnz = nnz_off + n
    # algorithmic bytes of one iteration of uva_plan_pcg (two float64
    # columns): the SpMV launch reads 12 B per stored off-diagonal entry
    # (value hi + lo, column id), the pinned diagonal (8 n), z gathered once
    # (16 n), p and A p read and written (64 n); the update launch reads p,
    # A p, x, r and the pinned diagonal and writes x, r, z (120 n).  The
    # sliced-ELL padding is not counted.
    per_it = 12 * nnz_off + 208 * n
    achieved = per_it * its / cg_t / 1e9
    # CPU: the oracle's Batch with the same coupled CG on 16 islands, one core
    k = min(16, len(items))
    sample = [(P, T, U0.cpu().numpy()[b.offsets[i]:b.offsets[i + 1]]) for i, (P, T, _) in enumerate(items[:k])]
    _one_thread()
    o = BatchOracle(sample)
    c0 = time.perf_counter()
    Uo = o.arap(8, line_search=True, cg_iters=40)
    t_cpu = time.perf_counter() - c0
    nt_s = int(sum(len(T) for _, T, _ in sample))
    bs = gpu.Batch(sample, dev=dev, dtype=torch.float64)
    Ug = bs.arap(8, line_search=True).cpu().numpy()
    err = 0.0
    for i in range(k):
        sl = slice(o.offsets[i], o.offsets[i + 1])
        err = max(err, float(np.abs(Ug[sl] - Uo[sl]).max() / np.ptp(Uo[sl], 0).max()))
    line = {"workload": f"gpu.Batch.arap(8, line_search) of {len(charts)} disk islands as one batch, "
                        f"one coupled CG of 40 iterations per global step",
            "triangles": nt, "ms_per_step": round(t_step * 1e3, 2),
            "mtri_iter_per_s": round(nt * 8 / t_step / 1e6, 1),
            "setup_s": round(t_setup, 3),
            "coupled_cg": {"iterations": its, "ms": round(cg_t * 1e3, 3), "rows": n, "nnz": nnz,
                           "bytes_per_iteration": per_it, "achieved_GBps": round(achieved, 1),
                           "frac": round(achieved / 8000.0, 4)},
            "cpu_baseline": {"value": round(nt_s * 8 / t_cpu / 1e6, 3), "unit": "Mtri*iter/s", "cores": 1,
                             "kind": "port", "sample": f"BatchOracle with the coupled CG on the first {k} islands "
                                                        f"({nt_s} triangles)"},
            "parity": {"sample_islands": k, "max_unit_box_err": err}}
    print(f"bench: Batch.arap(8) {t_step * 1e3:.1f} ms = {line['mtri_iter_per_s']} Mtri*iter/s; coupled CG "
          f"{cg_t * 1e3:.2f} ms / 40 iterations = {achieved:.0f} GB/s; CPU oracle {line['cpu_baseline']['value']} "
          f"Mtri*iter/s on {k} islands; sample error {err:.2e}", flush=True)
    print(json.dumps(line), flush=True)


def _oracle_flatten(job):
    from oracle import param_cpu as oracle

    P, T, iters = job
    return oracle.flatten(P, T, iters=iters)


def small_study(args):
    """The reference's most frequent call: charts.flatten of small charts
    (40-90 faces) from a pool of 8 threads (pipeline.py:608-611,
    charts.py:398-438).  Times, on the same charts:
      * param.lscm + param.refine(8 iterations) per chart, called from 8
        threads (one plan per call, as the reference's threads call it);
      * FlattenQueue: the 8 threads call flatten through one queue that
        gathers their calls into batched solves;
      * flatten_many of all charts in one call;
      * the float64 oracle's flatten on one core and over 8 worker processes.
    Reports charts per second and the per-call latency of each, and the
    largest layout difference of the batched path to the oracle."""
    import multiprocessing as mp
    import threading

    n_threads = 8
    pool = mp.get_context("fork").Pool(n_threads, initializer=_one_thread)
    import torch

    from oracle import param_cpu as oracle
    from ultimate_uv_artist_amd import charts as gch
    from ultimate_uv_artist_amd import param, workloads

    dev = torch.device("cuda", 0)
    torch.cuda.set_device(dev)
    rng = np.random.default_rng(9)
    n = args.charts
    cs = [workloads.disk_island(rng, int(t)) for t in rng.integers(40, 91, n)]
    tri = int(sum(len(T) for _, T in cs))
    iters = 8

    def threaded(fn):
        lat = [0.0] * n
        start = threading.Barrier(n_threads)

        def work(k):
            start.wait()
            for i in range(k, n, n_threads):
                t0 = time.perf_counter()
                fn(*cs[i])
                lat[i] = time.perf_counter() - t0

        th = [threading.Thread(target=work, args=(k,)) for k in range(n_threads)]
        t0 = time.perf_counter()
        for t in th:
            t.start()
        for t in th:
            t.join()
        return time.perf_counter() - t0, float(np.median(lat)) * 1e3

    def plain_call(P, T):
        U = param.lscm(P, T)
        param.refine(P, T, U, iters=iters, tol=1e-4)

    plain_call(*cs[0])                                     # warm-up
    t_plain, lat_plain = threaded(plain_call)
    with gch.FlattenQueue(iters=iters, dev=dev) as q:
        q.flatten(*cs[0])                                  # warm-up
        t_queue, lat_queue = threaded(lambda P, T: q.flatten(P, T, iters))
    gch.flatten_many(cs[:8], iters, dev)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    res = gch.flatten_many(cs, iters, dev)
    t_many = time.perf_counter() - t0
    k = min(args.cpu_sample if args.cpu_sample > 0 else n, n)
    t0 = time.perf_counter()
    ref = [oracle.flatten(P, T, iters=iters) for P, T in cs[:k]]
    t_cpu1 = (time.perf_counter() - t0) / k * n
    t0 = time.perf_counter()
    pool.map(_oracle_flatten, [(P, T, iters) for P, T in cs], chunksize=16)
    t_cpu8 = time.perf_counter() - t0
    pool.close()
    err = 0.0
    for (U, _), Ur in zip(res[:k], ref):          # the oracle leaves out the boundary untangle
        lo = Ur.min(0)
        err = max(err, float(np.abs(U - Ur).max() / (Ur.max(0) - lo).max()))
    line = {"workload": f"charts.flatten of {n} small disk charts (40-90 faces, {tri} triangles), {iters} ARAP "
                        f"iterations, callers on {n_threads} threads",
            "charts": n, "triangles": tri,
            "param_calls_8_threads": {"charts_per_s": round(n / t_plain, 1), "median_latency_ms": round(lat_plain, 3)},
            "flatten_queue_8_threads": {"charts_per_s": round(n / t_queue, 1),
                                        "median_latency_ms": round(lat_queue, 3)},
            "flatten_many": {"charts_per_s": round(n / t_many, 1), "ms": round(t_many * 1e3, 1)},
            "cpu_baseline": {"one_core_charts_per_s": round(n / t_cpu1, 1), "cores": n_threads,
                             "cores_available": len(os.sched_getaffinity(0)),
                             "charts_per_s": round(n / t_cpu8, 1), "kind": "port",
                             "sample": f"oracle flatten (SuperLU) of {k} charts on one core, all {n} over "
                                       f"{n_threads} worker processes"},
            "parity": {"charts": k, "max_unit_box_err": err}}
    print(f"bench: small charts: param calls {line['param_calls_8_threads']}, FlattenQueue "
          f"{line['flatten_queue_8_threads']}, flatten_many {line['flatten_many']}, CPU "
          f"{line['cpu_baseline']['one_core_charts_per_s']} charts/s on one core, "
          f"{line['cpu_baseline']['charts_per_s']} on {n_threads}; max err {err:.1e}", flush=True)
    print(json.dumps(line), flush=True)


def hier_study(args):
    """hier.Solver.relax (defaults: 6 iterations, 20 CG iterations, 24 repair
    rounds) of the 512 disk islands as the charts of one subset of one mesh
    (3.83M triangles) from their LSCM layouts.  The timed call includes what
    the reference's call does besides the iterations: the per-subset set-up
    (chart plan, pinned Laplacian) and the per-chart quality.  Reports
    Mtri*iter/s of the whole call, the oracle's relax on a 16-chart subset
    (one core) and its agreement with the GPU on that subset."""
    import types

    import torch

    from oracle import hier_cpu
    from ultimate_uv_artist_amd import hier, workloads
    from ultimate_uv_artist_amd.plan import IslandPlan

    dev = torch.device("cuda", 0)
    torch.cuda.set_device(dev)
    charts = workloads.disk_islands(seed=2, n_islands=args.islands, workers=args.gen_workers)
    plan = IslandPlan(charts, dev=dev)
    U = torch.zeros(plan.n, 2, dtype=torch.float64, device=dev)
    plan.lscm(U)
    U0 = U.cpu().numpy()
    vo, to = plan.vert_off, plan.tri_off
    del plan
    V = np.concatenate([P for P, _ in charts])
    F = np.concatenate([T + vo[i] for i, (_, T) in enumerate(charts)])
    chart = np.repeat(np.arange(len(charts)), np.diff(to))
    faces = np.arange(len(F))
    nt = len(F)
    S = hier.Solver(types.SimpleNamespace(V=V, F=F), dev=dev)
    S.relax(faces, F, len(V), chart, len(charts), U0)          # warm-up
    times = []
    for k in range(max(args.steps, 1)):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        Uh, flips, scale = S.relax(faces, F, len(V), chart, len(charts), U0)
        S.bad_fraction(2.0, 1.5)
        torch.cuda.synchronize()
        times.append(time.perf_counter() - t0)
        print(f"bench: hier relax step {k}: {times[-1] * 1e3:.1f} ms", flush=True)
    t_step = float(np.median(times))
    # the oracle on the first 16 charts as their own subset, one core
    k = min(16, len(charts))
    nv_s, nt_s = int(vo[k]), int(to[k])
    sub = (np.arange(nt_s), F[:nt_s], nv_s, chart[:nt_s], k, U0[:nv_s])
    _one_thread()
    O = hier_cpu.Solver(V[:nv_s], F[:nt_s])
    c0 = time.perf_counter()
    Uo, _, _ = O.relax(*sub)
    t_cpu = time.perf_counter() - c0
    Sg = hier.Solver(types.SimpleNamespace(V=V[:nv_s], F=F[:nt_s]), dev=dev)
    Ug, flg, _ = Sg.relax(*sub)
    err = 0.0
    for i in range(k):
        sl = slice(vo[i], vo[i + 1])
        err = max(err, float(np.abs(Ug[sl] - Uo[sl]).max() / np.ptp(Uo[sl], 0).max()))
    line = {"workload": f"hier.Solver.relax (6 iterations, 20 CG iterations) of {len(charts)} disk islands as the "
                        f"charts of one subset, set-up and quality included",
            "triangles": nt, "ms_per_call": round(t_step * 1e3, 2),
            "mtri_iter_per_s": round(nt * 6 / t_step / 1e6, 1), "charts_with_flips": int((flips > 0).sum()),
            "cpu_baseline": {"value": round(nt_s * 6 / t_cpu / 1e6, 3), "unit": "Mtri*iter/s", "cores": 1,
                             "kind": "port", "sample": f"hier_cpu.Solver.relax of the first {k} charts "
                                                        f"({nt_s} triangles)"},
            "parity": {"sample_charts": k, "max_unit_box_err": err}}
    print(f"bench: hier relax {t_step * 1e3:.1f} ms = {line['mtri_iter_per_s']} Mtri*iter/s; CPU oracle "
          f"{line['cpu_baseline']['value']} Mtri*iter/s on {k} charts; sample error {err:.2e}", flush=True)
    print(json.dumps(line), flush=True)


def main():
    args = parse()
    if args.copy_probe:
        copy_probe(args.copy_probe)
        return
    if args.gpus != 1:
        print(f"bench.py runs on one GPU; --gpus {args.gpus} is not supported (islands would shard with no "
              "exchange, see DESIGN.md)", file=sys.stderr)
        sys.exit(2)
    if args.workload == "needles":
        precision_study(args)
        return
    if args.workload == "validate":
        validate_study(args)
        return
    if args.workload == "batch":
        batch_study(args)
        return
    if args.workload == "hier":
        hier_study(args)
        return
    if args.workload == "small":
        small_study(args)
        return
    import torch

    from ultimate_uv_artist_amd import _lib, workloads
    from ultimate_uv_artist_amd.plan import IslandPlan

    prec = {"mixed": _lib.UVA_PREC_MIXED, "f32": _lib.UVA_PREC_F32, "f64": _lib.UVA_PREC_F64}[args.precision]
    pool, n_cores = None, 1
    if args.cpu_sample > 0:
        # worker processes for the all-core CPU baseline, forked before the GPU is touched
        import multiprocessing as mp

        n_cores = min(16, len(os.sched_getaffinity(0)))
        pool = mp.get_context("fork").Pool(n_cores, initializer=_one_thread)
    dev = torch.device("cuda", 0)
    torch.cuda.set_device(dev)
    t0 = time.perf_counter()
    if args.workload == "grid":
        charts = [workloads.grid_island(seed=1)]
        wl = (f"one 7,938-triangle grid island (64x64 grid, paraboloid, jitter), LSCM + {args.iters} ARAP, "
              f"CG rel resid {args.cg_tol:g}")
    elif args.workload == "heightfield":
        charts = workloads.heightfield_islands()
        wl = (f"{len(charts)} heightfield islands (1001x1001 fBm grid, Voronoi cells), LSCM + {args.iters} ARAP, "
              f"CG rel resid {args.cg_tol:g}")
    else:
        charts = workloads.disk_islands(seed=2, n_islands=args.islands, workers=args.gen_workers)
        wl = (f"{len(charts)} disk islands, t log-uniform [200,40000], LSCM + {args.iters} ARAP, "
              f"CG rel resid {args.cg_tol:g}")
    t_gen = time.perf_counter() - t0
    tri = np.array([len(T) for _, T in charts])
    nv = np.array([len(P) for P, _ in charts])
    nnz = np.array([len(P) + 2 * unique_edge_count(T) for P, T in charts])
    print(f"bench: {len(charts)} islands, {tri.sum()} triangles, {nv.sum()} vertices (generated in {t_gen:.1f}s)",
          flush=True)

    t0 = time.perf_counter()
    plan = IslandPlan(charts, dev=dev)
    torch.cuda.synchronize()
    t_setup = time.perf_counter() - t0
    U0 = torch.zeros(plan.n, 2, dtype=torch.float64, device=dev)
    t0 = time.perf_counter()
    plan.lscm(U0)
    torch.cuda.synchronize()
    t_lscm = time.perf_counter() - t0
    st_l = plan.stats()
    U0 = orient_majority(U0, plan, plan.vert_off, plan.tri_off, plan.T_local)
    # LSCM byte model (FP64 PCG on the coupled 2n operator, resident form:
    # the vectors stay in registers and LDS, so a CG iteration streams the
    # matrix only): per island and CG iteration 16 B per stored off-diagonal
    # entry (value hi + lo, column, boundary coupling)
    kl = st_l["cg_iters"].astype(np.float64)
    lscm_bytes = float(np.sum(kl * 16.0 * (nnz - nv)))
    lscm = {"ms": round(t_lscm * 1e3, 2), "cg_iters_mean": round(float(kl.mean()), 1), "cg_iters_max": int(kl.max()),
            "algorithmic_bytes": round(lscm_bytes), "achieved_GBps": round(lscm_bytes / t_lscm / 1e9, 1),
            "frac": round(lscm_bytes / t_lscm / 1e9 / HBM_PEAK_GBS, 4)}
    print(f"bench: setup {t_setup:.2f}s, LSCM {t_lscm * 1e3:.1f} ms (CG iters mean {st_l['cg_iters'].mean():.0f}, "
          f"max {st_l['cg_iters'].max()}, converged {int(st_l['converged'].sum())}/{len(charts)}; "
          f"{lscm_bytes / 1e9:.1f} GB algorithmic = {lscm['achieved_GBps']:.0f} GB/s)", flush=True)

    U = torch.empty_like(U0)
    stream = torch.cuda.current_stream(dev)

    def step():
        U.copy_(U0)
        plan.arap(U, args.iters, line_search=True, mode=_lib.UVA_MODE_REFINE, tol=0.0, cg_tol=args.cg_tol,
                  precision=prec, use_graph=not args.no_graph)

    untuned = None
    tuning = None
    if args.tune:
        # the default schedule (team sizes and order from a size-based guess
        # of the CG iterations), timed on a few steps for comparison
        step()
        torch.cuda.synchronize()
        nu = max(1, min(3, args.steps))
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _ in range(nu):
            step()
        e1.record(stream)
        torch.cuda.synchronize()
        ms_u = e0.elapsed_time(e1) / nu
        untuned = {"value": round(float(tri.sum()) * args.iters / (ms_u * 1e-3) / 1e6, 3),
                   "ms_per_step": round(ms_u, 3), "steps": nu}
        print(f"bench: default schedule {ms_u:.2f} ms/step ({untuned['value']} Mtri*iter/s)", flush=True)
        # autotune: the last step measured every island's CG iterations, the
        # plan re-plans team sizes and the workgroup schedule from them
        cost0 = plan.tuning()
        plan.tune()
        tuning = {"default_cost_us": [round(float(c), 4) for c in cost0],
                  "fitted_cost_us": [round(float(c), 4) for c in plan.tuning()]}
        print(f"bench: schedule re-planned from the measured CG iterations per island; cost model (us per CG "
              f"iteration: workgroup, team member, per slice) {tuning['default_cost_us']} -> "
              f"{tuning['fitted_cost_us']}", flush=True)
    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    ev0 = torch.cuda.Event(enable_timing=True)
    ev1 = torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    ev0.record(stream)
    for _ in range(args.steps):
        step()
    ev1.record(stream)
    torch.cuda.synchronize()
    ms = ev0.elapsed_time(ev1) / max(args.steps, 1)
    value = float(tri.sum()) * args.iters / (ms * 1e-3) / 1e6
    # The dominant kernel (one launch = one ARAP iteration of every island):
    # one more step run eagerly with a HIP event pair around every launch on
    # the launch stream.  The solves are deterministic, so this step repeats
    # the timed steps' work (the same CG iterations per island).
    U.copy_(U0)
    launch_ms = plan.arap_profile(U, args.iters, line_search=True, mode=_lib.UVA_MODE_REFINE, tol=0.0,
                                  cg_tol=args.cg_tol, precision=prec)
    arap_ms = launch_ms[1:]
    st = plan.stats()
    k = st["cg_iters"].astype(np.float64)                    # CG iterations over the step, per island
    iters_run = st["arap_iters"]
    bytes_step = float(np.sum(iters_run * (124.0 * tri + 72.0 * nv + 8.0 * nnz) + k * (100.0 * nv + 8.0 * nnz)))
    gbs = bytes_step / (arap_ms.sum() * 1e-3) / 1e9
    print(f"bench: ARAP-iteration kernel {arap_ms.mean():.3f} ms per launch (eager, HIP events; prologue "
          f"{launch_ms[0]:.3f} ms); {bytes_step / args.iters / 1e9:.2f} GB algorithmic per launch", flush=True)
    kpi = k / np.maximum(iters_run, 1)
    busy = st["busy_ms"]
    n_cu = torch.cuda.get_device_properties(dev).multi_processor_count
    # workgroup-slot time the islands used: a team holds m slots, a wavefront
    # task one eighth of one (8 wavefronts per 512-thread workgroup)
    slot_ms = float(np.sum(np.where(st["members"] > 0, st["members"], 1.0 / 8.0) * busy))
    slow = np.argsort(-busy)[:5]
    print(f"bench: island work {busy.sum():.0f} team-ms per step = {busy.sum() / ms:.1f} teams busy on average "
          f"({n_cu} CUs); workgroup slots busy {100.0 * slot_ms / (n_cu * ms):.0f}% of the step; "
          f"slowest islands (n, workgroups, CG iters, ms, us per CG iter): " +
          ", ".join(f"({nv[i]}, {int(st['members'][i])}, {int(k[i])}, {busy[i]:.1f}, "
                    f"{1e3 * busy[i] / max(k[i], 1):.1f})" for i in slow),
          flush=True)
    print(f"bench: {ms:.2f} ms/step, {value:.1f} Mtri*iter/s, {gbs:.0f} GB/s algorithmic; CG iters per global "
          f"step mean {kpi.mean():.1f} p95 {np.percentile(kpi, 95):.1f}; converged {int(st['converged'].sum())}/"
          f"{len(charts)}; max rel resid {st['rel_resid'].max():.2e}", flush=True)

    plain = None
    if args.plain:
        # the same kernels, one island per launch sequence, islands one after another
        plans = [IslandPlan([c], dev=dev) for c in charts]
        U0h = U0.cpu().numpy()
        vo = plan.vert_off
        Us0 = [torch.as_tensor(U0h[vo[i]:vo[i + 1]], device=dev).contiguous() for i in range(len(charts))]
        Us = [u.clone() for u in Us0]

        def plain_step():
            for p, u, u0 in zip(plans, Us, Us0):
                u.copy_(u0)
                p.arap(u, args.iters, line_search=True, mode=_lib.UVA_MODE_REFINE, tol=0.0, cg_tol=args.cg_tol,
                       precision=prec, use_graph=not args.no_graph)

        plain_step()
        torch.cuda.synchronize()
        ev0.record(stream)
        plain_step()
        ev1.record(stream)
        torch.cuda.synchronize()
        ms_plain = ev0.elapsed_time(ev1)
        Ub = U.cpu().numpy()
        diff = 0.0
        for i, u in enumerate(Us):
            a, b = u.cpu().numpy(), Ub[vo[i]:vo[i + 1]]
            lo = b.min(0)
            diff = max(diff, float(np.abs(a - b).max() / max((b.max(0) - lo).max(), 1e-300)))
        plain = {"value": round(float(tri.sum()) * args.iters / (ms_plain * 1e-3) / 1e6, 3),
                 "ms_per_step": round(ms_plain, 3), "speedup_batched": round(ms_plain / ms, 2),
                 "max_layout_diff_unit_box": diff}
        print(f"bench: plain per-island {plain['value']} Mtri*iter/s ({ms_plain:.1f} ms), batched is "
              f"{plain['speedup_batched']}x faster, layouts differ by at most {diff:.1e} (unit box)", flush=True)
        del plans

    cpu = None
    parity = None
    if args.cpu_sample > 0:
        from oracle import param_cpu as oracle

        # islands at evenly spaced quantiles of the triangle mass, so the sample
        # has the batch's mix of small and large islands by work, not by count
        order = np.argsort(tri)
        cum = np.cumsum(tri[order]) / tri.sum()
        qs = (np.arange(min(args.cpu_sample, len(order))) + 0.5) / min(args.cpu_sample, len(order))
        pick = np.unique(order[np.minimum(np.searchsorted(cum, qs), len(order) - 1)])
        Uh0 = U0.cpu().numpy()
        Uh = U.cpu().numpy()
        vo = plan.vert_off
        t_cpu = 0.0
        err, lay = [], []
        for i in pick:
            P, T = charts[i]
            u0 = Uh0[vo[i]:vo[i + 1]]
            c0 = time.perf_counter()
            ur, _ = oracle.refine(P, T, u0, iters=args.iters, tol=0.0)
            t_cpu += time.perf_counter() - c0
            g = Uh[vo[i]:vo[i + 1]]
            lo = ur.min(0)
            s = float((ur.max(0) - lo).max())
            err.append(float(np.abs((g - lo) / s - (ur - lo) / s).max()))
            lay.append((i, g, ur))
        # validate.py verdicts and distortion quantiles of the GPU and oracle
        # layouts on islands spread over the sample's sizes
        from oracle import validate_cpu as vc

        nvs = min(args.verdict_sample, len(lay))
        sel = np.unique(np.linspace(0, len(lay) - 1, nvs).round().astype(int)) if nvs > 0 else []
        nvs, same, dq = len(sel), 0, 0.0
        for j in sel:
            i, g, ur = lay[j]
            P, T = charts[i]
            same += int(vc.island_verdict(P, T, g) == vc.island_verdict(P, T, ur))
            q, qr = oracle.chart_quality(P, T, g), oracle.chart_quality(P, T, ur)
            dq = max(dq, abs(q["aniso_p95"] / qr["aniso_p95"] - 1), abs(q["area_p95"] / qr["area_p95"] - 1))
        cpu_tri_iter = float(tri[pick].sum()) * args.iters
        # the same sample over all cores, islands spread over worker processes
        # (largest first), as the reference spreads charts over a pool
        jobs = sorted(pick, key=lambda i: -tri[i])
        c0 = time.perf_counter()
        res = pool.map(_oracle_refine, [(*charts[i], Uh0[vo[i]:vo[i + 1]], args.iters) for i in jobs], chunksize=1)
        t_all = time.perf_counter() - c0
        pool.close()
        d_all = max(float(np.abs(r[0] - l[2]).max() / np.ptp(l[2], 0).max())
                    for r, l in zip(res, sorted(lay, key=lambda x: -tri[x[0]])))
        cpu = {"value": cpu_tri_iter / t_all / 1e6, "unit": "Mtri*iter/s", "cores": n_cores,
               "cores_available": len(os.sched_getaffinity(0)), "kind": "port",
               "value_1core": cpu_tri_iter / t_cpu / 1e6,
               "sample": f"{len(pick)} islands at evenly spaced quantiles of the triangle mass "
                         f"({int(tri[pick].sum())} triangles), {args.iters} ARAP iterations each from the same "
                         f"LSCM layout, SuperLU factorisation included: {t_all:.1f}s over {n_cores} worker "
                         f"processes, {t_cpu:.1f}s on one core; {cpu_model()}"}
        parity = {"islands": len(pick), "max_uv_err_unit_box": max(err), "verdicts_checked": nvs,
                  "verdicts_identical": same, "max_rel_diff_p95_aniso_area": dq,
                  "cpu_workers_vs_serial_unit_box": d_all}
        print(f"bench: CPU oracle sample {cpu['value']:.3f} Mtri*iter/s on {n_cores} cores, "
              f"{cpu['value_1core']:.3f} on one; parity max err {max(err):.2e}; "
              f"validate verdicts identical on {same}/{nvs} islands; p95 anisotropy/area within {dq:.1e}",
              flush=True)

    traffic, traffic_src = measured_traffic(workloads.mesh_digest(charts))
    rec = {
        "metric": "batched ARAP Mtri*iter/s at relative residual 1e-6",
        "value": round(value, 3),
        "unit": "Mtri*iter/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 SpMV/vectors, f64 dots+residual" if prec == _lib.UVA_PREC_MIXED else args.precision,
        "data": "synthetic",
        "config": {"workload": wl,
                   "islands": len(charts), "triangles": int(tri.sum()), "vertices": int(nv.sum()),
                   "arap_iters": args.iters, "mesh_sha256": workloads.mesh_digest(charts)[:16],
                   "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "achieved": round(gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(gbs / HBM_PEAK_GBS, 4), "traffic": traffic, "traffic_source": traffic_src,
                     "kernel": "island_kernel ARAP iteration", "launch_ms": round(float(arap_ms.mean()), 4),
                     "algorithmic_bytes_per_launch": round(bytes_step / args.iters)},
        "cpu_baseline": cpu,
        "untuned": untuned,
        "tuning": tuning,
        "kernel_src_sha16": kernel_source_digest(),
        "plain_per_island": plain,
        "cg_iters_per_step": {"mean": round(float(kpi.mean()), 2), "p95": round(float(np.percentile(kpi, 95)), 2),
                              "max": int(st["cg_iters_max"].max())},
        "converged_islands": int(st["converged"].sum()),
        "setup_s": round(t_setup, 3),
        "lscm_ms": round(t_lscm * 1e3, 2),
        "lscm": lscm,
        "parity": parity,
    }
    print(json.dumps(rec), flush=True)


if __name__ == "__main__":
    main()
