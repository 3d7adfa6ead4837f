"""Benchmark: D3Q19 BGK + Guo collide-stream on a 256^3 sphere pack (one GPU).

Workload (the single-GPU configuration the metric is quoted on): D3Q19
single-phase BGK, tau = 1, body force g_x = 1e-6, 256^3 box periodic in x
with solid walls on the y and z faces, filled with a seeded sphere pack of
radius 8 at porosity 0.40, solids as half-way bounce-back.

value    = MLUPS of the plain FP64 form (the reference's arithmetic), whole
           run, inputs resident.
fp32s, fp16s = the same for the precision forms (shifted populations in
           float or half storage, float arithmetic, FP64 flux reductions),
           each with its permeability error against FP64.
roofline = the dominant kernel (the FP64 collide-stream step) against 8 TB/s,
           with algorithmic bytes per lattice update = 2 Q s
           (populations read once and written once; s = 8 B FP64, 4 B FP32S,
           2 B FP16S).
permeability = k of the three forms run to |dk|/k <= 1e-7 at two consecutive
           checks 100 steps apart, and the relative errors against FP64.
cpu_baseline = the compiled reference std_case (oracle/_ref, built from the
           reference's sources with MPICH) on the host's cores, one MPI rank
           per core of the process's CPU share (16 on a one-GPU box; nproc
           and the affinity set are recorded beside it), on a 128^3 pack of
           the same generator, from two differenced runs; "port" beside it is
           the oracle's OpenMP restatement of the same loop on the full 256^3
           node list.
host_io  = the FP64 step with the populations handed over from host memory:
           one lb_upload_f and one lb_download_f of every node's populations,
           timed after the headline's timed region, and the rate of the same
           K steps with those two transfers added (never the value).
batched  = the ensemble of independent sphere-pack permeability solves (cubic
           sizes log-uniform in 32..128, porosity U(0.30, 0.50), radius
           U(3, 8), one seed; |dk|/k <= 1e-6 at two consecutive checks 100
           steps apart), all
           resident in one pool and advanced together in FP32S: solves/s over
           the whole ensemble (wall clock from the first admission to the last
           result, host set-up of each solve included), against the plain FP64
           path run one solve at a time (lb_solve_permeability on its own
           context), timed on every solve (--baseline-every m: every m-th by
           fluid count, scaled by m); k error of FP32S and FP16S vs FP64 on
           the solves timed.
channel  = the D2Q9 256x128 channel (solid rows y = 0, 127, periodic x, tau 1,
           F 1e-6, 5000 steps from rest) in FP64: MLUPS and the velocity
           profile against the Poiseuille start-up series.
twophase = the colour-gradient case: 192^3 sphere pack (radius 8, porosity
           0.38), phases drawn per pore node with saturation 0.5, solid
           wettability 1, tau0 = tau1 = 1, sigma = 0.01, beta = 0.7, body force
           1e-6 on the total, 50000 steps in FP64, FP32S and FP16S (the
           reduced forms store floats or halves and compute in FP64): MLUPS,
           the phase field difference and the phase permeabilities
           k_i = nu_i <phi_i sum_q f c_x> / F_x of both.
serving  = a stream of requests, 70 % permeability (single-phase) and 30 %
           relative permeability (colour-gradient, saturation U(0.2, 0.8),
           phases drawn per pore cell), cubic sizes U(48, 160), seeded sphere
           packs; admitted by HBM footprint into one pool, advanced together
           (continuous batching), each resident until it converges (relative
           change of k, or of both phase permeabilities, at most 1e-6 per 100
           steps at two consecutive checks) or reaches its step limit.
           Permeability requests are stored in FP32S (top-level figures) or
           FP16S ("fp16s"), relative-permeability requests in FP32S in both
           (their step computes in FP64 on the stored floats; FP16S storage
           drifts from FP64 by more than 1e-4 over the two-phase
           transient).  FP32S: a burst (all requests at t = 0) gives the
           capacity in solves/s, then Poisson arrivals at 0.7 x that capacity
           give p50/p99 latency; FP16S: the burst.  The same stop rules on the
           plain FP64 path one request at a time (every request;
           --serve-baseline-every m times every m-th permeability request by
           size and scales) give the baseline solves/s and the k errors.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--size 256]
"""
import argparse
import json
import os
import subprocess
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "badchimp-cpp_amd"))

import lbamd  # noqa: E402
from lbamd.workloads import TWO_PHASE, ensemble, request_mix  # noqa: E402

METRIC = "D3Q19 MLUPS & %HBM roofline; batched perm solves/s + p99 at ≤1e-4 k err"
HBM_PEAK_GBS = 8000.0


PROFILE = os.path.join(ROOT, "profiles", "k_bgk_d3q19_256_pack.md")


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def profiled_traffic(form):
    """HBM bytes per collide-stream launch from the committed rocprofv3
    FETCH_SIZE / WRITE_SIZE runs (copy-calibrated) of this workload."""
    try:
        for line in open(PROFILE):
            parts = line.split()
            if len(parts) == 3 and parts[0] == "traffic_bytes_per_launch" and parts[1] == form:
                return int(parts[2])
    except OSError:
        pass
    return None


def make_workload(size, seed=1):
    geo, info = lbamd.sphere_pack((size, size, size), 8.0, 0.40, seed, periodic="x", walls=True)
    t0 = time.perf_counter()
    topo = lbamd.Topology.from_geometry(geo, lbamd.D3Q19, "x")
    info["host_builder_s"] = time.perf_counter() - t0
    return geo, info, topo


def setup_leg(geo, info, topo, text_size=128):
    """Geometry -> node list on the GPU against the host builder (same
    arrays), and the binary .vtklb twin against the text format."""
    import tempfile
    lbamd.Topology.from_geometry_device(np.ones((8, 8, 8), np.int32), lbamd.D3Q19, "x").close()  # HIP init
    t0 = time.perf_counter()
    dev = lbamd.Topology.from_geometry_device(geo, lbamd.D3Q19, "x")
    t_dev = time.perf_counter() - t0
    same = all(np.array_equal(getattr(topo, m)(), getattr(dev, m)())
               for m in ("neighbors", "positions", "node_types", "bulk"))
    dev.close()
    out = {"nodes": int(topo.n_nodes), "host_builder_s": info.get("host_builder_s"), "device_builder_s": t_dev,
           "device_equals_host": bool(same)}
    # the context's layout from the node list: link table, row windows, every
    # link of both checked against the table (lb_set_topology_grid)
    nb, types, bulk, pos = topo.neighbors(), topo.node_types(), topo.bulk(), topo.positions()
    ctx = lbamd.Context(lbamd.D3Q19, lbamd.FP64)
    t0 = time.perf_counter()
    ctx.set_topology_grid(nb, types, bulk, pos)
    ctx.sync()
    out["layout_s"] = time.perf_counter() - t0
    out["layout"] = ctx.layout_note
    ctx.close()
    with tempfile.TemporaryDirectory() as tmp:
        path = os.path.join(tmp, "t.vtklbb")
        t0 = time.perf_counter()
        topo.write_binary(path)
        out["binary_write_s"] = time.perf_counter() - t0
        out["binary_bytes"] = os.path.getsize(path)
        t0 = time.perf_counter()
        b = lbamd.Topology.from_binary(path, lbamd.D3Q19)
        out["binary_read_s"] = time.perf_counter() - t0
        b.close()
        os.remove(path)
        # the text format on a smaller pack (its reader is ~100x slower)
        g2, _ = lbamd.sphere_pack((text_size,) * 3, 8.0, 0.40, 1, periodic="x", walls=True)
        t2 = lbamd.Topology.from_geometry(g2, lbamd.D3Q19, "x")
        tp, bp = os.path.join(tmp, "t.vtklb"), os.path.join(tmp, "t2.vtklbb")
        t2.write_vtklb(tp)
        t2.write_binary(bp)
        t0 = time.perf_counter()
        a = lbamd.Topology.from_vtklb(tp, lbamd.D3Q19)
        t_text = time.perf_counter() - t0
        t0 = time.perf_counter()
        b = lbamd.Topology.from_binary(bp, lbamd.D3Q19)
        t_bin = time.perf_counter() - t0
        out.update({"text_size": text_size, "text_nodes": int(t2.n_nodes), "text_read_s": t_text,
                    "text_bytes": os.path.getsize(tp), "binary_read_same_case_s": t_bin,
                    "text_equals_binary": bool(all(np.array_equal(getattr(a, m)(), getattr(b, m)())
                                                   for m in ("neighbors", "positions", "node_types", "bulk")))})
    log("set-up: host builder %.2f s, device builder %.3f s (same: %s); context layout %.2f s; binary twin "
        "write %.2f s, read %.2f s; text %d^3 read %.2f s vs binary %.3f s" % (
            out["host_builder_s"], t_dev, same, out["layout_s"], out["binary_write_s"], out["binary_read_s"],
            text_size, t_text, t_bin))
    return out


def rest_state(n_nodes, bulk):
    import oracle  # only for the weights table; the GPU legs do not use the oracle
    return oracle.rest_state(1, n_nodes, bulk, np.ones(n_nodes))


def make_ctx(precision, nb, types, bulk, pos, layout):
    ctx = lbamd.Context(lbamd.D3Q19, precision, 0)
    ctx.set_topology_grid(nb, types, bulk, pos, table_only=(layout == "table"))
    return ctx


def gpu_leg(topo, nb, types, bulk, pos, layout, precision, steps, warmup, tau, force, tau_anti=None, table=None,
            host_io=None):
    """Time `steps` steps with the populations resident.  With a dict in
    host_io, also time (after the timed region) one lb_upload_f and one
    lb_download_f of every node's FP64 populations, the host buffers a caller
    hands over and gets back."""
    ctx = make_ctx(precision, nb, types, bulk, pos, layout)
    got_layout = {2: "row windows", 1: "rows"}.get(ctx.layout, "table")
    if ctx.layout_note and ctx.layout_note != got_layout:
        got_layout += " (%s)" % ctx.layout_note
    w = np.array([2.0 / 36.0] * 3 + [1.0 / 36.0] * 6 + [2.0 / 36.0] * 3 + [1.0 / 36.0] * 6 + [12.0 / 36.0])
    f0 = np.zeros((topo.n_nodes, 19))
    f0[bulk] = w[None, :]
    ctx.upload_f(f0)

    if table is not None:
        ctx.set_viscosity_table(table[:, 0], table[:, 1])

    def step(n):
        if table is not None:
            ctx.step_rheology(n, force)
        elif tau_anti is None:
            ctx.step_bgk(n, tau, force)
        else:
            ctx.step_trt(n, tau, tau_anti, force)
    if warmup:
        step(warmup)
    ctx.sync()
    ctx.kernel_time_reset()
    t0 = time.perf_counter()
    step(steps)
    ctx.sync()
    t1 = time.perf_counter()
    kms, launches = ctx.kernel_time()
    if host_io is not None:
        out = np.zeros(f0.size)
        ctx.download_f(out=out)  # first call outside the timing
        ctx.sync()
        t2 = time.perf_counter()
        ctx.upload_f(f0)
        ctx.sync()
        t3 = time.perf_counter()
        ctx.download_f(out=out)
        t4 = time.perf_counter()
        host_io.update(bytes_each_way=int(f0.nbytes), upload_s=t3 - t2, download_s=t4 - t3)
    ctx.close()
    return t1 - t0, kms / max(launches, 1), got_layout


def permeability(nb, types, bulk, pos, layout, n_nodes, precision, tau, force, n_box, tol=1e-7,
                 max_steps=60000, tau_anti=None):
    ctx = make_ctx(precision, nb, types, bulk, pos, layout)
    w = np.array([2.0 / 36.0] * 3 + [1.0 / 36.0] * 6 + [2.0 / 36.0] * 3 + [1.0 / 36.0] * 6 + [12.0 / 36.0])
    f0 = np.zeros((n_nodes, 19))
    f0[bulk] = w[None, :]
    ctx.upload_f(f0)
    t0 = time.perf_counter()
    if tau_anti is None:
        k, steps = ctx.solve_permeability(tau, force, 100, tol, max_steps, n_box)
    else:
        k, steps = ctx.solve_permeability_trt(tau, tau_anti, force, 100, tol, max_steps, n_box)
    dt = time.perf_counter() - t0
    ctx.close()
    return k, steps, dt


def trt_leg(args, geo, topo, nb, types, bulk, pos, force, n_bulk, k_bgk):
    """Two-relaxation-time step on the same pack: tau_sym = 1, tau_anti =
    7/8 ((tau_sym - 1/2)(tau_anti - 1/2) = 3/16), timed in FP32S and FP64, and
    its permeability in FP64 and FP32S beside BGK's at tau = 1; and the
    generalized-Newtonian step timed the same way."""
    ts, ta = 1.0, 0.875
    out = {"tau_sym": ts, "tau_anti": ta}
    for name, prec, bpl in (("fp32s", lbamd.FP32S, 2 * 19 * 4), ("fp64", lbamd.FP64, 2 * 19 * 8)):
        t, kms, _ = gpu_leg(topo, nb, types, bulk, pos, args.layout, prec, args.steps, args.warmup, ts, force,
                            tau_anti=ta)
        gbs = n_bulk * bpl / (kms * 1e-3) / 1e9
        out[name] = {"value": round(n_bulk * args.steps / t / 1e6, 2), "unit": "MLUPS", "avg_kernel_ms": kms,
                     "achieved_GBs": round(gbs, 1), "frac": round(gbs / HBM_PEAK_GBS, 4)}
        log("TRT %s: %.1f MLUPS, kernel %.3f ms" % (name, out[name]["value"], kms))
    # generalized-Newtonian step: per-node relaxation time from a 40-row
    # shear-thinning (Carreau) viscosity table
    sr = np.logspace(-16, -4, 40)
    table = np.stack([sr, 0.02 + (1.0 / 6.0 - 0.02) * (1.0 + 1e10 * sr) ** -0.25], 1)
    out["rheology"] = {"table_rows": 40}
    for name, prec, bpl in (("fp32s", lbamd.FP32S, 2 * 19 * 4), ("fp64", lbamd.FP64, 2 * 19 * 8)):
        t, kms, _ = gpu_leg(topo, nb, types, bulk, pos, args.layout, prec, args.steps, args.warmup, 1.0, force,
                            table=table)
        gbs = n_bulk * bpl / (kms * 1e-3) / 1e9
        out["rheology"][name] = {"value": round(n_bulk * args.steps / t / 1e6, 2), "unit": "MLUPS",
                                 "avg_kernel_ms": kms, "achieved_GBs": round(gbs, 1),
                                 "frac": round(gbs / HBM_PEAK_GBS, 4)}
        log("rheology %s: %.1f MLUPS, kernel %.3f ms" % (name, out["rheology"][name]["value"], kms))
    if not args.no_perm:
        n_box = float(geo.size)
        k64, s64, _ = permeability(nb, types, bulk, pos, args.layout, topo.n_nodes, lbamd.FP64, ts, force, n_box,
                                   tau_anti=ta)
        k32, s32, _ = permeability(nb, types, bulk, pos, args.layout, topo.n_nodes, lbamd.FP32S, ts, force, n_box,
                                   tau_anti=ta)
        out.update({"k_fp64": k64, "steps_fp64": s64, "k_fp32s": k32, "steps_fp32s": s32,
                    "k_rel_err_fp32s": abs(k32 - k64) / abs(k64),
                    "k_bgk_tau1_over_k_trt": (k_bgk / k64) if k_bgk else None})
        log("TRT k FP64 = %.10e (%d steps), FP32S rel err %.2e" % (k64, s64, out["k_rel_err_fp32s"]))
    return out


def host_cores():
    """Cores this process may use: the affinity set, at most 16 (the host
    CPU share of one GPU on the boxes this runs on), with what nproc and
    os.cpu_count() report beside it."""
    try:
        aff = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        aff = os.cpu_count() or 1
    try:
        nproc = int(subprocess.run(["nproc"], capture_output=True, text=True, timeout=10).stdout.strip())
    except (OSError, ValueError, subprocess.SubprocessError):
        nproc = None
    return min(16, aff), {"nproc": nproc, "os_cpu_count": os.cpu_count(), "affinity": aff,
                          "used": min(16, aff)}


def cpu_baseline(nb, types, bulk, n_nodes, tau, force, budget_s=15.0):
    import oracle
    threads, _ = host_cores()
    f0 = rest_state(n_nodes, bulk)
    # one step to size the sample, then as many as fit the budget (at most 50)
    t0 = time.perf_counter()
    oracle.run_std_case(1, nb, types, bulk, f0, tau, force, 1, nthreads=threads)
    one = time.perf_counter() - t0
    steps = int(max(1, min(50, budget_s / max(one, 1e-6))))
    t0 = time.perf_counter()
    oracle.run_std_case(1, nb, types, bulk, f0, tau, force, steps, nthreads=threads)
    dt = time.perf_counter() - t0
    return {"value": bulk.size * steps / dt / 1e6, "unit": "MLUPS", "cores": threads, "kind": "port",
            "sample": "oracle std_case loop (push + bounce-back, FP64) on the same 256^3 node list, "
                      "%d steps, %d OpenMP threads" % (steps, threads)}


def cpu_reference(tau, force, size=128, n_a=20, n_b=520):
    """The compiled reference std_case (oracle/_ref, MPICH) on the host's
    cores, one MPI rank per core, on a bounded sample of the workload: the
    same sphere-pack generator (radius 8, porosity 0.40, seed 1, periodic x,
    walls) at size^3, steps n_b - n_a from two differenced runs."""
    from oracle import ref_run
    if not ref_run.available():
        raise RuntimeError("reference build or MPICH runtime absent")
    ranks, cores = host_cores()
    geo, info = lbamd.sphere_pack((size,) * 3, 8.0, 0.40, 1, periodic="x", walls=True)
    dt, n_bulk, wall = ref_run.time_std_case(geo, "x", tau, force, ranks, n_a, n_b)
    return {"value": n_bulk * (n_b - n_a) / dt / 1e6, "unit": "MLUPS", "cores": ranks, "kind": "reference",
            "host_cores": cores,
            "sample": "reference std_case (src/std_case/main.cpp, D3Q19, -O3, MPICH) on a %d^3 pack of the same "
                      "generator (%d bulk nodes, porosity %.3f), %d MPI ranks (x slabs), %d steps: runs of %d and "
                      "%d steps differenced (%.1f s, %.1f s)" % (size, n_bulk, info["porosity"], ranks, n_b - n_a,
                                                               n_a, n_b, wall[n_a], wall[n_b])}


def batched_leg(n, every, tau, force):
    t0 = time.perf_counter()
    geos, par = ensemble(n)
    log("ensemble: %d samples, %.1f M cells, generated in %.1f s" % (n, sum(g.size for g in geos) / 1e6,
                                                                       time.perf_counter() - t0))
    fps = [lbamd.batch_footprint(lbamd.D3Q19, lbamd.FP32S, g, "x") for g in geos]
    pool = int(sum(f[0] for f in fps) * 1.01) + (64 << 20)
    nfluid = np.array([f[1] for f in fps])
    def run(precision):
        b = lbamd.Batch(lbamd.D3Q19, precision, pool, 100)
        t0 = time.perf_counter()
        assert b.add_many(geos, "x", tau, force, 1e-6, 400000, list(range(n))) == n
        t_adm = time.perf_counter() - t0
        res = {}
        while len(res) < n:
            b.advance(1)
            for r in b.collect():
                res[r["tag"]] = r
        wall = time.perf_counter() - t0
        kms, launches, lups = b.kernel_time()
        shape = b.tuning()
        b.close()
        return res, wall, t_adm, kms, lups, shape

    res, wall, t_adm, kms, lups, shape = run(lbamd.FP32S)
    res16, wall16, _, kms16, lups16, shape16 = run(lbamd.FP16S)
    log("batched FP16S: %d solves in %.2f s, %.1f GLUPS device" % (n, wall16, lups16 / (kms16 * 1e-3) / 1e9))
    log("batched FP32S: %d solves in %.2f s (admission %.2f s), %.1f GLUPS device, tuned %s" %
        (n, wall, t_adm, lups / (kms * 1e-3) / 1e9, shape))
    # plain FP64 one at a time: every solve, or a size-stratified subset
    order = np.argsort(nfluid, kind="stable")
    subset = order[every // 2::every] if every > 1 else order
    t_one = 0.0
    errs, errs16, steps64 = [], [], []
    for i in subset:
        g = geos[i]
        topo = lbamd.Topology.from_geometry(g, lbamd.D3Q19, "x")
        nb, types, bulk, pos = topo.neighbors(), topo.node_types(), topo.bulk(), topo.positions()
        ctx = make_ctx(lbamd.FP64, nb, types, bulk, pos, "rows")
        w = np.array([2.0] * 3 + [1.0] * 6 + [2.0] * 3 + [1.0] * 6 + [12.0]) / 36.0
        f0 = np.zeros((topo.n_nodes, 19))
        f0[bulk] = w[None, :]
        ctx.upload_f(f0)
        ctx.sync()
        t1 = time.perf_counter()
        k64, s64 = ctx.solve_permeability(tau, force, 100, 1e-6, 400000, float(g.size))
        t_one += time.perf_counter() - t1
        ctx.close()
        steps64.append(s64)
        errs.append(abs(res[int(i)]["k"] - k64) / abs(k64))
        errs16.append(abs(res16[int(i)]["k"] - k64) / abs(k64))
    est_one = t_one * n / len(subset)
    steps = np.array([res[i]["steps"] for i in range(n)])
    steps16 = np.array([res16[i]["steps"] for i in range(n)])
    sub = [int(i) for i in subset]
    out = {"n_solves": n, "solves_per_s": n / wall, "wall_s": wall, "admission_s": t_adm,
           "device_GLUPS": lups / (kms * 1e-3) / 1e9, "lups": int(lups), "tuned_launch": shape, "fp16s_tuned_launch": shape16,
           "steps_median": int(np.median(steps)), "steps_max": int(steps.max()),
           "fp16s_steps_median": int(np.median(steps16)), "fp16s_steps_max": int(steps16.max()),
           "steps_total": {"fp32s": int(steps.sum()), "fp16s": int(steps16.sum())},
           "steps_on_fp64_subset": {"fp64": int(np.sum(steps64)), "fp32s": int(steps[sub].sum()),
                                    "fp16s": int(steps16[sub].sum())},
           "fluid_nodes": int(nfluid.sum()), "pool_GB": pool / 1e9,
           "fp64_one_at_a_time_solves_per_s": n / est_one, "fp64_subset": int(len(subset)),
           "speedup_vs_fp64_one_at_a_time": est_one / wall,
           "k_rel_err_fp32s_vs_fp64_max": float(max(errs)), "tol_dk_over_k": 1e-6,
           "fp16s_solves_per_s": n / wall16, "fp16s_device_GLUPS": lups16 / (kms16 * 1e-3) / 1e9,
           "fp16s_speedup_vs_fp64_one_at_a_time": est_one / wall16,
           "k_rel_err_fp16s_vs_fp64_max": float(max(errs16)),
           "converged": int(sum(1 for r in res.values() if r["status"] == 0)),
           "fp16s_converged": int(sum(1 for r in res16.values() if r["status"] == 0))}
    log("FP64 one at a time (%d of %d solves timed%s): %.2f solves/s; batched %.2f solves/s (x%.2f); "
        "max k err %.2e" % (len(subset), n, "" if len(subset) == n else ", scaled",
                            out["fp64_one_at_a_time_solves_per_s"], out["solves_per_s"],
                            out["speedup_vs_fp64_one_at_a_time"], out["k_rel_err_fp32s_vs_fp64_max"]))
    return out


W19 = np.array([2.0] * 3 + [1.0] * 6 + [2.0] * 3 + [1.0] * 6 + [12.0]) / 36.0
TP = TWO_PHASE


def fp64_one(req):
    """The plain FP64 path for one request (own context, same stop rule);
    returns (k or (k0, k1), steps, device seconds)."""
    geo = req.geo
    topo = lbamd.Topology.from_geometry(geo, lbamd.D3Q19, "x")
    nb, types, bulk, pos = topo.neighbors(), topo.node_types(), topo.bulk(), topo.positions()
    n = topo.n_nodes
    nbox = float(geo.size)
    if req.kind == "perm":
        ctx = make_ctx(lbamd.FP64, nb, types, bulk, pos, "rows")
        f0 = np.zeros((n, 19))
        f0[bulk] = W19[None, :]
        ctx.upload_f(f0)
        ctx.sync()
        t0 = time.perf_counter()
        k, steps = ctx.solve_permeability(req.tau, np.asarray(req.force), 100, req.tol, req.max_steps, nbox)
        dt = time.perf_counter() - t0
        ctx.close()
        return k, steps, dt
    r0 = np.zeros(n)
    r0[bulk] = req.rho0[pos[bulk, 0], pos[bulk, 1], pos[bulk, 2]]
    r1 = 1.0 - r0
    sb = types == 1
    r0[sb], r1[sb] = req.wett, 1.0 - req.wett
    f0 = np.zeros((n, 2, 19))
    f0[bulk, 0] = W19[None] * r0[bulk, None]
    f0[bulk, 1] = W19[None] * r1[bulk, None]
    ctx = make_ctx(lbamd.FP64, nb, types, bulk, pos, "rows")
    ctx.upload_f(f0)
    ctx.set_wall_densities(np.stack([r0, r1], 1))
    ctx.sync()
    t0 = time.perf_counter()
    k, steps = ctx.solve_relperm(req.tau0, req.tau1, req.sigma, req.beta, np.asarray(req.force, dtype=float), 100,
                                 req.tol, req.max_steps, nbox)
    dt = time.perf_counter() - t0
    ctx.close()
    return k, steps, dt


def channel_leg(steps=5000):
    from lbamd.geometry import channel_2d
    geo = channel_2d(256, 128)
    topo = lbamd.Topology.from_geometry(geo, lbamd.D2Q9, "x")
    nb, types, bulk, pos = topo.neighbors(), topo.node_types(), topo.bulk(), topo.positions()
    # rest state f = w in the D2Q9 direction order (axis, diagonal alternating, rest last)
    w9 = np.array([4.0, 1.0, 4.0, 1.0, 4.0, 1.0, 4.0, 1.0, 16.0]) / 36.0
    f0 = np.zeros((topo.n_nodes, 9))
    f0[bulk] = w9[None, :]
    ctx = lbamd.Context(lbamd.D2Q9, lbamd.FP64)
    ctx.set_topology_grid(nb, types, bulk, pos)
    ctx.upload_f(f0)
    ctx.sync()
    t0 = time.perf_counter()
    ctx.step_bgk(steps + 1, 1.0, np.array([1e-6, 0.0]))
    ctx.sync()
    dt = time.perf_counter() - t0
    vel = ctx.download_vel()
    ctx.close()
    prof = vel[bulk, 0].reshape(256, 126)[0]
    # start-up series of plane Poiseuille flow between half-way walls (H = 126, nu = 1/6)
    H, nu, F, t = 126.0, 1.0 / 6.0, 1e-6, steps + 0.5
    yh = np.arange(1, 127) - 0.5
    u = F / (2 * nu) * yh * (H - yh)
    for n in range(1, 400, 2):
        u -= 4 * F * H * H / (nu * np.pi ** 3 * n ** 3) * np.sin(n * np.pi * yh / H) * np.exp(-n * n * np.pi ** 2 * nu * t / (H * H))
    return {"workload": "D2Q9 256x128 channel, tau 1, F 1e-6, %d steps, FP64" % steps,
            "MLUPS": bulk.size * (steps + 1) / dt / 1e6, "wall_s": dt,
            "profile_err_vs_startup_series": float(np.abs(prof - u).max() / u.max())}


def twophase_leg(size, steps, seed=3):
    geo, info = lbamd.sphere_pack((size,) * 3, 8.0, 0.38, seed, periodic="x", walls=True)
    topo = lbamd.Topology.from_geometry(geo, lbamd.D3Q19, "x")
    nb, types, bulk, pos = topo.neighbors(), topo.node_types(), topo.bulk(), topo.positions()
    n = topo.n_nodes
    rng = np.random.default_rng(seed + 1)
    r0 = np.zeros(n)
    r0[bulk] = (rng.random(bulk.size) < 0.5).astype(float)
    r1 = np.zeros(n)
    r1[bulk] = 1.0 - r0[bulk]
    sb = types == 1
    r0[sb], r1[sb] = TP["wett"], 1.0 - TP["wett"]
    f0 = np.zeros((n, 2, 19))
    f0[bulk, 0] = W19[None] * r0[bulk, None]
    f0[bulk, 1] = W19[None] * r1[bulk, None]
    rho2 = np.stack([r0, r1], 1)
    F = np.array([1e-6, 0.0, 0.0])
    nbox = float(geo.size)
    res = {}
    for prec, name in ((lbamd.FP64, "fp64"), (lbamd.FP32S, "fp32s"), (lbamd.FP16S, "fp16s")):
        ctx = lbamd.Context(lbamd.D3Q19, prec)
        ctx.set_topology_grid(nb, types, bulk, pos)
        ctx.upload_f(f0)
        ctx.set_wall_densities(rho2)
        ctx.sync()
        ctx.kernel_time_reset()
        t0 = time.perf_counter()
        ctx.step_cg(steps, TP["tau0"], TP["tau1"], TP["sigma"], TP["beta"], lbamd.FORCE_FIXED, F)
        ctx.sync()
        dt = time.perf_counter() - t0
        kms, launches = ctx.kernel_time()
        s0, s1, s2 = ctx.reduce_flux(0)
        phi = ctx.download_scalar(lbamd.PHI)[bulk]
        ctx.close()
        k0 = (TP["tau0"] - 0.5) / 3.0 * (s1 / nbox) / F[0]
        k1 = (TP["tau1"] - 0.5) / 3.0 * (s2 / nbox) / F[0]
        res[name] = {"MLUPS": bulk.size * steps / dt / 1e6, "k0": k0, "k1": k1, "phi": phi,
                     "ms_per_step": kms / steps}
        log("two-phase %s: %.1f MLUPS, k0 %.6e k1 %.6e" % (name, res[name]["MLUPS"], k0, k1))
    a, b, h = res["fp64"], res["fp32s"], res["fp16s"]
    scale = abs(a["k0"]) + abs(a["k1"])
    out = {"workload": "D3Q19 colour gradient, %d^3 sphere pack r=8 porosity %.3f, S_w 0.5, %d steps" %
                       (size, info["porosity"], steps),
           "bulk_nodes": int(bulk.size),
           "fp64_MLUPS": a["MLUPS"], "fp32s_MLUPS": b["MLUPS"],
           "fp64_ms_per_step": a["ms_per_step"], "fp32s_ms_per_step": b["ms_per_step"],
           "k_phase_fp64": [a["k0"], a["k1"]], "k_phase_fp32s": [b["k0"], b["k1"]],
           "k_phase_err_rel_to_sum": [abs(b["k0"] - a["k0"]) / scale, abs(b["k1"] - a["k1"]) / scale],
           "phi_max_abs_diff": float(np.abs(a["phi"] - b["phi"]).max()),
           "phi_mean_abs_diff": float(np.abs(a["phi"] - b["phi"]).mean()),
           "phase_sign_flips": int(np.sum(np.sign(a["phi"]) != np.sign(b["phi"]))),
           "fp16s_MLUPS": h["MLUPS"], "k_phase_fp16s": [h["k0"], h["k1"]],
           "k_phase_err_rel_to_sum_fp16s": [abs(h["k0"] - a["k0"]) / scale, abs(h["k1"] - a["k1"]) / scale],
           "phi_mean_abs_diff_fp16s": float(np.abs(a["phi"] - h["phi"]).mean())}
    log("two-phase FP32S vs FP64: k err %s, phi max %.3e mean %.3e" %
        (out["k_phase_err_rel_to_sum"], out["phi_max_abs_diff"], out["phi_mean_abs_diff"]))
    return out


SERVE_TOL = 1e-6
SERVE_MAX_STEPS = 400000
RELPERM_MAX_STEPS = 20000


def serve_requests(n_req, seed, t_rate=None):
    reqs = request_mix(n_req, seed, t_rate=t_rate, tol=SERVE_TOL, max_steps=SERVE_MAX_STEPS)
    for r in reqs:
        if r.kind == "relperm":
            r.max_steps = RELPERM_MAX_STEPS
    return reqs


def serve_run(pool, precision, n_req, seed, t_rate=None):
    """One serving stream: (results, wall seconds, kernel ms, LUPs)."""
    from lbamd.serve import Server, run_stream
    srv = Server(pool, precision=precision, check_every=100, relperm_precision=lbamd.FP32S)
    done, wall = run_stream(srv, serve_requests(n_req, seed, t_rate=t_rate))
    kms, _, lups = srv.batch.kernel_time()
    srv.close()
    return done, wall, kms, lups


def serving_leg(n_req, pool_gb, every):
    pool = int(pool_gb * 1e9)
    base_reqs = {r.tag: r for r in serve_requests(n_req, 7)}
    # 3) the plain FP64 path one request at a time: every relative-permeability
    #    request, and every permeability request (or, with every > 1, a subset
    #    stratified by size, every m-th by cell count, its time scaled by the
    #    kind's count)
    ref, est, n_sub = {}, 0.0, 0
    for kind in ("perm", "relperm"):
        tags = sorted((t for t in range(n_req) if base_reqs[t].kind == kind), key=lambda t: base_reqs[t].geo.size)
        if not tags:
            continue
        m = max(1, min(every, len(tags))) if kind == "perm" else 1
        sub = tags[m // 2::m] or tags[:1]
        t_kind = 0.0
        for t in sub:
            k, steps, dt = fp64_one(base_reqs[t])
            ref[t] = (kind, k, steps)
            t_kind += dt
            log("  FP64 one at a time: request %d (%s, %d^3) %d steps, %.2f s" % (
                t, kind, base_reqs[t].geo.shape[0], steps, dt))
        est += t_kind * len(tags) / len(sub)
        n_sub += len(sub)
    fp64_rate = n_req / est

    def errors(done):
        """Max k error of the permeability subset; the relative-permeability
        error (max over both phases, of k0 + k1) of every request, split by
        whether the FP64 run converged or stopped at max_steps."""
        byres = {d["tag"]: d for d in done}
        e_perm, e_conv, e_cap = [], [], []
        for t, (kind, k, steps) in ref.items():
            got = byres[t]
            if kind == "perm":
                e_perm.append(abs(got["k"] - k) / abs(k))
            else:
                scale = abs(k[0]) + abs(k[1])
                e = max(abs(got["k_phase"][0] - k[0]), abs(got["k_phase"][1] - k[1])) / scale
                (e_cap if steps >= base_reqs[t].max_steps else e_conv).append(e)
        mx = lambda v: float(max(v)) if v else None  # noqa: E731
        e_all = e_conv + e_cap
        return {"k_rel_err_max_perm": mx(e_perm), "k_rel_err_max_relperm": mx(e_all),
                "k_rel_err_median_relperm": float(np.median(e_all)) if e_all else None,
                "k_rel_err_max_relperm_converged": mx(e_conv), "n_relperm_converged": len(e_conv),
                "k_rel_err_max_relperm_at_max_steps": mx(e_cap), "n_relperm_at_max_steps": len(e_cap)}

    out = {"n_requests": n_req, "mix": "70% permeability / 30% relative permeability, L U(48,160)",
           "pool_GB": pool / 1e9, "fp64_one_at_a_time_solves_per_s": fp64_rate, "fp64_subset": n_sub,
           "stop": "|dk| <= %g |k| (two-phase: |dk_i| <= %g (|k_0|+|k_1|)) per 100 steps at two consecutive "
                   "checks; at most %d steps; two-phase requests run a fixed %d steps (the reference's twophase "
                   "loop runs nIterations with no stop test) unless they settle first" % (
                       SERVE_TOL, SERVE_TOL, SERVE_MAX_STEPS, RELPERM_MAX_STEPS),
           "relperm_storage": "FP32S (FP64 arithmetic)"}
    for precision, name in ((lbamd.FP32S, "fp32s"), (lbamd.FP16S, "fp16s")):
        # 1) burst: capacity
        done, wall, kms, lups = serve_run(pool, precision, n_req, 7)
        cap = len(done) / wall
        log("serving %s burst: %d requests in %.1f s" % (name, len(done), wall))
        # 2) Poisson arrivals at 0.7 x capacity (FP32S)
        if precision == lbamd.FP32S:
            done2, wall2, _, _ = serve_run(pool, precision, n_req, 8, t_rate=0.7 * cap)
            lat = np.array([d["latency"] for d in done2])
        err = errors(done)
        per_kind = {}
        for kind in ("perm", "relperm"):
            ds = [d for d in done if d["kind"] == kind]
            if ds:
                per_kind[kind] = {"n": len(ds), "steps_median": int(np.median([d["steps"] for d in ds])),
                                  "max_steps_reached": int(sum(1 for d in ds if d["status"] == 1)),
                                  "latency_p50_s": float(np.percentile([d["latency"] for d in ds], 50))}
        leg = {"burst_solves_per_s": cap, "burst_wall_s": wall,
               "burst_device_GLUPS": lups / (kms * 1e-3) / 1e9 if kms > 0 else None,
               "burst_device_busy_frac": kms * 1e-3 / wall,
               "speedup_vs_fp64_one_at_a_time": cap / fp64_rate,
               "steps_median": int(np.median([d["steps"] for d in done])), "per_kind_burst": per_kind}
        if precision == lbamd.FP32S:
            leg.update({"poisson_rate_per_s": 0.7 * cap, "poisson_wall_s": wall2,
                        "poisson_solves_per_s": len(done2) / wall2,
                        "p50_latency_s": float(np.percentile(lat, 50)),
                        "p99_latency_s": float(np.percentile(lat, 99))})
        leg.update(err)
        log("serving %s: burst %.2f solves/s (x%.2f vs FP64 one at a time %.3f/s)%s; k err perm %s, relperm "
            "converged %s (%d), at max steps %s (%d)" % (
                name, cap, cap / fp64_rate, fp64_rate,
                ", Poisson %.2f/s p50 %.2f s p99 %.2f s" % (0.7 * cap, leg["p50_latency_s"], leg["p99_latency_s"])
                if "p50_latency_s" in leg else "",
                err["k_rel_err_max_perm"], err["k_rel_err_max_relperm_converged"], err["n_relperm_converged"],
                err["k_rel_err_max_relperm_at_max_steps"], err["n_relperm_at_max_steps"]))
        if precision == lbamd.FP32S:
            out.update(leg)
        else:
            out["fp16s"] = leg
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--size", type=int, default=256)
    ap.add_argument("--layout", choices=["rows", "table"], default="rows")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-perm", action="store_true")
    ap.add_argument("--batch", type=int, default=512, help="ensemble size (0: skip)")
    ap.add_argument("--baseline-every", type=int, default=1,
                    help="time the FP64 one-at-a-time baseline on every m-th solve (1: all)")
    ap.add_argument("--serve-baseline-every", type=int, default=1,
                    help="time the serving baseline on every m-th permeability request (1: all)")
    ap.add_argument("--serve", type=int, default=32, help="requests per serving run (0: skip)")
    ap.add_argument("--no-channel", action="store_true")
    ap.add_argument("--no-trt", action="store_true")
    ap.add_argument("--no-setup", action="store_true")
    ap.add_argument("--tp-size", type=int, default=192)
    ap.add_argument("--tp-steps", type=int, default=50000, help="two-phase steps (0: skip)")
    ap.add_argument("--serve-pool-gb", type=float, default=16.0)
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU only: every box here has one GPU and the engine runs one "
              "subdomain per GPU; --gpus %d is not supported" % args.gpus, file=sys.stderr)
        sys.exit(2)

    tau, force = 1.0, np.array([1e-6, 0.0, 0.0])
    t0 = time.perf_counter()
    geo, info, topo = make_workload(args.size)
    nb, types, bulk, pos = topo.neighbors(), topo.node_types(), topo.bulk(), topo.positions()
    n_bulk = int(bulk.size)
    log("workload: %d^3, %d bulk nodes, porosity %.4f, set-up %.1f s" % (args.size, n_bulk, info["porosity"], time.perf_counter() - t0))

    t32, k32_ms, lay32 = gpu_leg(topo, nb, types, bulk, pos, args.layout, lbamd.FP32S, args.steps, args.warmup, tau, force)
    mlups32 = n_bulk * args.steps / t32 / 1e6
    log("FP32S: %.1f MLUPS, kernel %.3f ms, layout %s" % (mlups32, k32_ms, lay32))
    host_io = {}
    t64, k64_ms, lay64 = gpu_leg(topo, nb, types, bulk, pos, args.layout, lbamd.FP64, args.steps, args.warmup, tau, force,
                                 host_io=host_io)
    mlups64 = n_bulk * args.steps / t64 / 1e6
    # the same K steps with the populations handed over from host memory and back
    host_io["steps"] = args.steps
    host_io["MLUPS_host_buffers"] = round(n_bulk * args.steps / (t64 + host_io["upload_s"] + host_io["download_s"]) / 1e6, 2)
    host_io["GBs_upload"] = round(host_io["bytes_each_way"] / host_io["upload_s"] / 1e9, 2)
    host_io["GBs_download"] = round(host_io["bytes_each_way"] / host_io["download_s"] / 1e9, 2)
    log("FP64 with host buffers: %.1f MLUPS over %d steps (upload %.3f s, download %.3f s)" % (
        host_io["MLUPS_host_buffers"], args.steps, host_io["upload_s"], host_io["download_s"]))
    log("FP64 : %.1f MLUPS, kernel %.3f ms, layout %s" % (mlups64, k64_ms, lay64))
    t16, k16_ms, _ = gpu_leg(topo, nb, types, bulk, pos, args.layout, lbamd.FP16S, args.steps, args.warmup, tau, force)
    mlups16 = n_bulk * args.steps / t16 / 1e6
    log("FP16S: %.1f MLUPS, kernel %.3f ms" % (mlups16, k16_ms))

    copy_gbs, _ = lbamd.bench_copy(1 << 30, 20)
    log("plain copy: %.1f GB/s (read + write)" % copy_gbs)
    stream_gbs, _ = lbamd.bench_stream(n_bulk, lbamd.D3Q19, 20)
    log("AoSoA identity stream (same node count): %.1f GB/s (read + write)" % stream_gbs)
    bytes32 = 2 * 19 * 4
    bytes64 = 2 * 19 * 8
    ach32 = n_bulk * bytes32 / (k32_ms * 1e-3) / 1e9
    ach64 = n_bulk * bytes64 / (k64_ms * 1e-3) / 1e9

    # counters were taken on the 256^3 workload only
    traffic32 = profiled_traffic("FP32S") if args.size == 256 else None
    traffic64 = profiled_traffic("FP64") if args.size == 256 else None

    perm = None
    if not args.no_perm:
        n_box = float(geo.size)
        k64, s64, d64 = permeability(nb, types, bulk, pos, args.layout, topo.n_nodes, lbamd.FP64, tau, force, n_box)
        log("k FP64 = %.10e after %d steps (%.1f s)" % (k64, s64, d64))
        k32, s32, d32 = permeability(nb, types, bulk, pos, args.layout, topo.n_nodes, lbamd.FP32S, tau, force, n_box)
        log("k FP32S = %.10e after %d steps (%.1f s)" % (k32, s32, d32))
        k16, s16, d16 = permeability(nb, types, bulk, pos, args.layout, topo.n_nodes, lbamd.FP16S, tau, force, n_box)
        log("k FP16S = %.10e after %d steps (%.1f s)" % (k16, s16, d16))
        perm = {"k_fp64": k64, "steps_fp64": s64, "k_fp32s": k32, "steps_fp32s": s32,
                "k_rel_err_fp32s": abs(k32 - k64) / abs(k64), "k_fp16s": k16, "steps_fp16s": s16,
                "k_rel_err_fp16s": abs(k16 - k64) / abs(k64), "tol_dk_over_k": 1e-7,
                "definition": "k = nu * sum_bulk(sum_q f c_x + F_x/2) / (N_box * F_x), nu = (tau-1/2)/3"}

    setup = None if args.no_setup else setup_leg(geo, info, topo)
    trt = None if args.no_trt else trt_leg(args, geo, topo, nb, types, bulk, pos, force, n_bulk,
                                            perm["k_fp64"] if perm else None)
    batched = batched_leg(args.batch, args.baseline_every, tau, force) if args.batch > 0 else None
    channel = None if args.no_channel else channel_leg()
    if channel:
        log("channel: %.1f MLUPS, profile error vs start-up series %.2e" % (channel["MLUPS"],
                                                                          channel["profile_err_vs_startup_series"]))
    twophase = twophase_leg(args.tp_size, args.tp_steps) if args.tp_steps > 0 else None
    serving = serving_leg(args.serve, args.serve_pool_gb, args.serve_baseline_every) if args.serve > 0 else None

    cpu = None
    if not args.no_cpu:
        port = cpu_baseline(nb, types, bulk, topo.n_nodes, tau, force)
        log("CPU port: %.2f MLUPS on %d threads" % (port["value"], port["cores"]))
        try:
            cpu = cpu_reference(tau, force)
            cpu["port"] = port
            log("CPU reference: %.2f MLUPS on %d MPI ranks" % (cpu["value"], cpu["cores"]))
        except (RuntimeError, OSError, subprocess.SubprocessError) as e:
            cpu = dict(port, reference_error=str(e)[:300])
            log("CPU reference not timed (%s); the port stands in" % e)

    k_err = {}
    if perm:
        k_err = {"fp32s": perm["k_rel_err_fp32s"], "fp16s": perm["k_rel_err_fp16s"]}
    line = {
        "metric": METRIC,
        "value": round(mlups64, 2),
        "unit": "MLUPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": t64 / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded sphere pack)",
        "config": {"workload": "D3Q19 BGK+Guo collide-stream, %d^3 sphere pack r=8 porosity %.3f, "
                               "periodic x, walls y/z, tau=1, F=1e-6" % (args.size, info["porosity"]),
                   "bulk_nodes": n_bulk, "form": "FP64 (plain)", "layout": lay64,
                   "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "achieved": round(ach64, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(ach64 / HBM_PEAK_GBS, 4), "traffic": traffic64,
                     "kernel": "k_bgk<D3Q19,RowWin>", "avg_kernel_ms": k64_ms,
                     "bytes_per_lup": bytes64, "plain_copy_GBs": round(copy_gbs, 1),
                     "aosoa_stream_GBs": round(stream_gbs, 1)},
        "fp32s": {"value": round(mlups32, 2), "unit": "MLUPS", "ms_per_step": t32 / args.steps * 1e3,
                  "avg_kernel_ms": k32_ms, "kernel": "k_bgk_lp<D3Q19,float,RowWin>",
                  "achieved_GBs": round(ach32, 1), "frac": round(ach32 / HBM_PEAK_GBS, 4),
                  "bytes_per_lup": bytes32, "traffic": traffic32, "k_rel_err_vs_fp64": k_err.get("fp32s")},
        "fp16s": {"value": round(mlups16, 2), "unit": "MLUPS", "ms_per_step": t16 / args.steps * 1e3,
                  "avg_kernel_ms": k16_ms, "kernel": "k_bgk_lp<D3Q19,half,RowWin>",
                  "achieved_GBs": round(n_bulk * 76 / (k16_ms * 1e-3) / 1e9, 1),
                  "frac": round(n_bulk * 76 / (k16_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4), "bytes_per_lup": 76,
                  "traffic": profiled_traffic("FP16S") if args.size == 256 else None,
                  "k_rel_err_vs_fp64": k_err.get("fp16s")},
        "host_io": host_io,
        "permeability": perm,
        "setup": setup,
        "trt": trt,
        "batched": batched,
        "channel": channel,
        "twophase": twophase,
        "serving": serving,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
