#!/usr/bin/env python3
"""Benchmark of the MI355X cloth step (one GPU).

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload crumple513]
                    [--precond learned|amg|jacobi] [--compare jacobi,amg]

A step is one full simulation step (physics: LBVH + proximities + implicit
assembly + preconditioned CG to 1e-5 relative residual + position/geometry
update; collision: CCD LBVH + FP64 impacts + impact response + velocity
update).  The default workload is the 513 x 513 crumple (263,169 nodes,
524,288 faces, dt = 2 ms, ground plane), run over its full 300 steps from the
initial state with every input resident in HBM: with --steps K below 300, K of
the 300 steps, spread evenly over the run (free fall, landing and contact alike),
are timed, each between two completed streams, and value = N K / their summed
time; the other steps run untimed.  W warmup steps run first on the same
context, which is then reset to the initial state.

The measured run uses the learned preconditioner in its fastest form: the
aggregation-multigrid V-cycle whose finest-level smoother is the network's
factor G^T G (packaged trained weights; the network and the smoother's damping
estimate run once per solve inside the timed step).  The same steps are then
run with block-Jacobi PCG (the reference's method), with the learned factor as
the whole preconditioner (M^-1 = G^T G), and with the multigrid cycle smoothed
by block Jacobi, each on a fresh context.

Prints one JSON line: node-steps/s and CG iterations per step of the measured
run, the same for the comparison runs, the roofline of the dominant kernel (the
PCG SpMV, timed live with HIP events on the library's stream) and a CPU
baseline (the oracle restatement on the host cores, on a bounded sample: the
first step of the same workload).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "node-steps/s + CG iters to 1e-5 rel residual: learned-precond vs Jacobi PCG"
HBM_PEAK_GBS = 8000.0


def static_nnzb(faces, n_nodes):
    """Block count of the static pattern: node pairs sharing a face or a hinge stencil."""
    f = np.asarray(faces, dtype=np.int64)
    pairs = [f[:, [a, b]] for a in range(3) for b in range(3)]
    e = np.sort(np.concatenate([f[:, [0, 1]], f[:, [1, 2]], f[:, [2, 0]]]), axis=1)
    opp = np.concatenate([f[:, 2], f[:, 0], f[:, 1]])
    key = e[:, 0] * n_nodes + e[:, 1]
    order = np.argsort(key, kind="stable")
    ks, os_ = key[order], opp[order]
    same = ks[1:] == ks[:-1]
    flaps = np.stack([os_[:-1][same], os_[1:][same]], axis=1)
    pairs.append(flaps)
    pairs.append(flaps[:, ::-1])
    allp = np.concatenate(pairs)
    u = np.unique(allp[:, 0] * n_nodes + allp[:, 1])
    return int(len(u))


def read_traffic(workload, precond):
    """HBM bytes per SpMV launch measured with rocprofv3 PMC counters (committed under profiles/)."""
    p = os.path.join(ROOT, "profiles", "pcg_spmv_%s_%s_traffic.json" % (workload, PRECONDS[precond]))
    if os.path.exists(p):
        with open(p) as fh:
            return json.load(fh).get("bytes_per_launch")
    return None


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def omp_threads(n):
    """Sets the OpenMP thread count of this process (the oracle's runtime) and returns the old one."""
    import ctypes
    gomp = ctypes.CDLL("libgomp.so.1")
    old = gomp.omp_get_max_threads()
    gomp.omp_set_num_threads(int(n))
    return old


def cpu_baseline(spec, steps):
    import oracle
    threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or os.cpu_count()
    o = oracle.OracleScene(spec)
    t0 = time.perf_counter()
    for _ in range(steps):
        o.step(direct=False)
    dt = time.perf_counter() - t0
    n = spec.n_cloth_nodes
    return {"value": n * steps / dt, "unit": "node-steps/s", "cores": threads, "kind": "port",
            "cpu_model": cpu_model(),
            "sample": "first %d step(s) of %s from its initial state: oracle FP32 restatement (block-Jacobi PCG "
                      "to 1e-5, FP64 CCD), OpenMP on %d threads, %.1f s" % (steps, spec.name, threads, dt)}


def cpu_reference_direct(workload="hang33", steps=100):
    """The reference's own CPU solve (one thread, a direct sparse LL^T of every step's system,
    src/Cloth.cu:676-693) on the first `steps` steps of a workload: the oracle's FP32 step with its
    direct solver (nested-dissection multifrontal Cholesky in FP64) on one OpenMP thread."""
    import oracle
    from clothsimulator_amd import scenes
    spec = scenes.by_name(workload)
    old = omp_threads(1)
    try:
        o = oracle.OracleScene(spec)
        t0 = time.perf_counter()
        for k in range(steps):
            o.step(direct=2)
            if time.perf_counter() - t0 > 30.0 and k + 1 < steps:
                print("cpu reference column: %s step %d/%d  %.0f s" % (workload, k + 1, steps,
                      time.perf_counter() - t0), file=sys.stderr, flush=True)
        dt = time.perf_counter() - t0
    finally:
        omp_threads(old)
    n = spec.n_cloth_nodes
    span = "all %d steps" % steps if steps >= spec.steps else "the first %d of its %d steps" % (steps, spec.steps)
    return {"value": n * steps / dt, "unit": "node-steps/s", "cores": 1, "kind": "port", "cpu_model": cpu_model(),
            "workload": workload, "ms_per_step": 1e3 * dt / steps,
            "sample": "%s, %s: oracle FP32 step with a direct sparse Cholesky (nested dissection, multifrontal, "
                      "FP64) on one thread, %.1f s" % (spec.name, span, dt)}


def copy_probe(mib):
    """A copy kernel of known size: reads and writes mib MiB per launch, 5 launches."""
    import torch
    n = mib * (1 << 20) // 4
    src = torch.arange(n, dtype=torch.float32, device="cuda:0")
    dst = torch.empty_like(src)
    for _ in range(5):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"copy_probe_mib": mib, "bytes_read_per_launch": 4 * n, "bytes_written_per_launch": 4 * n}))


PRECONDS = {"learned": "learned_spai", "amg-learned": "amg_v11_learned_smoother", "jacobi": "block_jacobi",
            "amg": "amg_v11"}


def timed_steps(steps, span):
    """Which of the workload's `span` steps are timed: all of them, or, for fewer, `steps` steps
    spread evenly over the run (step floor(i span / steps) for i < steps)."""
    if steps >= span:
        return list(range(steps))
    return sorted(set(int(i * span // steps) for i in range(steps)))


def run(workload, precond, steps, warmup, timed, extra_flags=0):
    """Steps `workload` from its initial state with one preconditioner; returns the timings."""
    from clothsimulator_amd import CS_MESH_CLOTH, CS_MESH_OBSTACLE, Simulator, scenes
    from clothsimulator_amd._cabi import (CS_FLAG_TIME_KERNELS, CS_KERNEL_EVENT_PAIR, CS_KERNEL_PCG_SPMV,
                                          CS_PRECOND_AMG, CS_PRECOND_AMG_LEARNED, CS_PRECOND_BLOCK_JACOBI,
                                          CS_PRECOND_LEARNED)

    spec = scenes.by_name(workload)
    if timed:
        spec.params.flags |= CS_FLAG_TIME_KERNELS
    spec.params.flags |= extra_flags
    spec.params.precond = {"amg": CS_PRECOND_AMG, "jacobi": CS_PRECOND_BLOCK_JACOBI, "learned": CS_PRECOND_LEARNED,
                           "amg-learned": CS_PRECOND_AMG_LEARNED}[precond]
    sim = Simulator(spec)   # the learned preconditioner loads the packaged trained weights
    init = [sim.state(CS_MESH_CLOTH, i) for i in range(len(spec.cloths))]
    obs = [sim.state(CS_MESH_OBSTACLE, i) for i in range(len(spec.obstacles))]
    for _ in range(warmup):
        sim.step()
    for i, (x, x0, v) in enumerate(init):
        sim.set_state(CS_MESH_CLOTH, i, x, x0, v)
    for i, (x, x0, v) in enumerate(obs):
        sim.set_state(CS_MESH_OBSTACLE, i, x, x0, v)
    sim.set_step_count(0)  # motions restart from t = 0 with the initial state
    if timed:
        sim.kernel_time(CS_KERNEL_PCG_SPMV, reset=True)
        sim.kernel_time(CS_KERNEL_EVENT_PAIR, reset=True)
    span = max(steps, spec.steps)
    sel = set(timed_steps(steps, span))
    iters, contacts, failed, failures = [], [], 0, []
    elapsed = 0.0
    t_all = time.perf_counter()
    for k in range(span):
        # every library call returns after its stream has completed: a step's wall time is
        # bracketed by two completed streams
        t0 = time.perf_counter()
        s, st = sim.step()
        dt = time.perf_counter() - t0
        if k in sel:
            elapsed += dt
            if s != 0:  # a PCG that hit its iteration cap (status 4) counts as a failed step too
                failed += 1
                failures.append({"step": k, "status": int(s), "pcg_converged": int(st.pcg_converged),
                                 "pcg_rel_residual": float(st.pcg_rel_residual), "pcg_iters": int(st.pcg_iters)})
                print("%s step %d: status %d (pcg converged %d, relative residual %.3g)" % (
                    precond, k, s, st.pcg_converged, st.pcg_rel_residual), file=sys.stderr, flush=True)
            iters.append(st.pcg_iters)
            contacts.append(st.n_contact_blocks)
        if (k + 1) % 50 == 0:
            print("%s step %d/%d  %.1f s  cg %d" % (precond, k + 1, span, time.perf_counter() - t_all, st.pcg_iters),
                  file=sys.stderr, flush=True)
    out = {"spec": spec, "elapsed": elapsed, "iters": iters, "contacts": contacts, "failed": failed,
           "failures": failures,
           "value": spec.n_cloth_nodes * len(sel) / elapsed, "ms_per_step": 1e3 * elapsed / len(sel),
           "span": span, "timed": len(sel)}
    if timed:
        out["spmv"] = sim.kernel_time(CS_KERNEL_PCG_SPMV)
        out["pair"] = sim.kernel_time(CS_KERNEL_EVENT_PAIR)
    if len(spec.cloths) > 1:
        out["batch_tuning"] = sim.batch_tuning()
    sim.close()
    return out


def sequential_sample(spec, precond, steps, k, batched_value):
    """K instances of a batch (evenly spaced in the list), each run alone on its own context for
    the same steps: their summed node-steps over their summed wall time is the throughput of
    running the instances one after another."""
    from clothsimulator_amd import Simulator, scenes
    from clothsimulator_amd._cabi import CS_PRECOND_AMG, CS_PRECOND_BLOCK_JACOBI
    n = len(spec.cloths)
    code = CS_PRECOND_AMG if precond == "amg" else CS_PRECOND_BLOCK_JACOBI
    picks = sorted(set(int(i * n / k) for i in range(k)))
    nodes, secs = 0, 0.0
    for i in picks:
        p = scenes.default_params(dt=spec.params.dt)
        p.precond = code
        one = scenes.SceneSpec("instance%d" % i, p, [spec.cloths[i]], [spec.obstacles[i]], steps)
        sim = Simulator(one)
        sim.step()   # first-step allocations outside the timing
        t0 = time.perf_counter()
        for _ in range(steps):
            sim.step()
        secs += time.perf_counter() - t0
        nodes += one.n_cloth_nodes * steps
        sim.close()
        print("sequential sample: instance %d (%d nodes) done" % (i, one.n_cloth_nodes), file=sys.stderr, flush=True)
    seq = nodes / secs
    return {"instances_sampled": len(picks), "node_steps_per_s": seq, "batched_speedup": batched_value / seq,
            "note": "instances run one after another, each on its own context, %s PCG" % (
                "multigrid" if code == CS_PRECOND_AMG else "block-Jacobi")}


def summary(r):
    return {"value": r["value"], "ms_per_step": r["ms_per_step"], "cg_iters_mean": float(np.mean(r["iters"])),
            "cg_iters_max": int(np.max(r["iters"])), "failed_steps": r["failed"], "failures": r["failures"]}


def remesh_run(workload, frames, steps_per_frame=8):
    """A remeshing workload (cloths with a "remeshing" block): `frames` frames of `steps_per_frame`
    steps, each frame ending with the remeshing step (Simulator::simulateStep,
    src/Simulator.cu:838-897).  Reports the step and remesh wall times and the mesh sizes."""
    from clothsimulator_amd import CS_MESH_CLOTH, Simulator, scenes
    spec = scenes.by_name(workload)
    sim = Simulator(spec)
    rows, failures = [], []
    node_steps, t_steps, t_remesh = 0, 0.0, 0.0
    for f in range(frames):
        n = sim.counts(CS_MESH_CLOTH, 0)[0]
        t0 = time.perf_counter()
        for k in range(steps_per_frame):
            s, st = sim.step()
            if s != 0 or not st.pcg_converged:
                failures.append({"frame": f, "step": k, "status": int(s), "pcg_converged": int(st.pcg_converged)})
        t1 = time.perf_counter()
        sim.remesh()
        t2 = time.perf_counter()
        sz, ops, rebuild, sep, sep_solves = sim.remesh_times()
        n2 = sim.counts(CS_MESH_CLOTH, 0)[0]
        node_steps += n * steps_per_frame
        t_steps += t1 - t0
        t_remesh += t2 - t1
        rows.append({"frame": f, "nodes": n, "nodes_after": n2, "ms_per_step": 1e3 * (t1 - t0) / steps_per_frame,
                     "ms_remesh": 1e3 * (t2 - t1), "ms_sizing": sz, "ms_operators": ops, "ms_rebuild": rebuild,
                     "ms_separation": sep, "separation_solves": sep_solves})
        print("remesh run %s frame %d: %d -> %d nodes, %.2f ms/step, remesh %.1f ms (sizing %.1f, operators %.1f, "
              "rebuild %.1f, separation %.1f with %d solves)" % (workload, f, n, n2, rows[-1]["ms_per_step"],
                                                                 rows[-1]["ms_remesh"], sz, ops, rebuild, sep,
                                                                 sep_solves), file=sys.stderr, flush=True)
    return {"metric": "cloth node-steps per second, remeshing included", "workload": spec.name, "frames": frames,
            "steps_per_frame": steps_per_frame, "value": node_steps / (t_steps + t_remesh),
            "unit": "node-steps/s", "ms_per_step_mean": 1e3 * t_steps / (frames * steps_per_frame),
            "ms_per_remesh_mean": 1e3 * t_remesh / frames, "failed_steps": len(failures), "failures": failures,
            "frames_detail": rows}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="crumple513")
    ap.add_argument("--cpu-steps", type=int, default=1)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--precond", choices=sorted(PRECONDS), default="amg-learned",
                    help="PCG preconditioner of the measured run: multigrid with the learned smoother (default), "
                         "learned SPAI factor, multigrid with block-Jacobi smoothing or block-Jacobi")
    ap.add_argument("--compare", default="jacobi,learned,amg",
                    help="comma-separated preconditioners also run over the same steps ('' for none)")
    ap.add_argument("--sequential-sample", type=int, default=16, metavar="K",
                    help="batch workloads: also run K of the instances one at a time, each on its own context, "
                         "and report the sequential throughput they imply")
    ap.add_argument("--persistent", action="store_true",
                    help="batch workloads: also run block Jacobi with each poll's iterations in one cooperative "
                         "launch (CS_FLAG_PERSISTENT_BATCH)")
    ap.add_argument("--cpu-direct", default="", metavar="WORKLOAD:STEPS,...",
                    help="also time the single-thread direct-solve CPU column on these windows "
                         "(e.g. drape129:20,crumple513:3); hang33 over its 100 steps always runs")
    ap.add_argument("--remesh-run", default="", metavar="WORKLOAD:FRAMES",
                    help="only run a remeshing workload (e.g. remesh_drape129:25: 8 steps then a remesh per frame) "
                         "and print its own JSON line")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="only run a plain device copy of MIB MiB (calibrates HBM byte counters) and exit")
    a = ap.parse_args()
    if a.gpus != 1:
        print("bench.py: this build measures one GPU; --gpus %d is not supported (the cloth step has no "
              "multi-GPU path here: it runs on one GPU's share of the work)" % a.gpus, file=sys.stderr)
        sys.exit(2)

    if a.copy_probe:
        copy_probe(a.copy_probe)
        return
    if a.remesh_run:
        wl, frames = a.remesh_run.split(":")
        print(json.dumps(remesh_run(wl, int(frames))))
        return

    from clothsimulator_amd import scenes

    steps = a.steps if a.steps is not None else scenes.by_name(a.workload).steps
    main_run = run(a.workload, a.precond, steps, a.warmup, True)
    spec = main_run["spec"]
    others = {}
    for p in [q for q in a.compare.split(",") if q and q != a.precond]:
        others[p] = run(a.workload, p, steps, a.warmup, False)

    n_nodes = spec.n_cloth_nodes
    c0 = spec.cloths[0]
    nnzb = sum(static_nnzb(c.mesh.faces, c.mesh.n_nodes) for c in spec.cloths)
    nnzb_dyn = nnzb + float(np.mean(main_run["contacts"]))
    # algorithmic bytes of one SpMV launch: 36 B values + 4 B column index per block; per node
    # z and the previous direction read once (16 B each), the new direction (16 B) and q (12 B)
    # written once
    bytes_per_launch = 40.0 * nnzb_dyn + 60.0 * n_nodes
    # every sampled SpMV is bracketed by an event pair; an empty pair recorded just before it
    # measures the markers' own latency, which is subtracted
    spmv_ms, spmv_n = main_run["spmv"]
    pair_ms, pair_n = main_run["pair"]
    marker_s = (pair_ms / pair_n) * 1e-3 if pair_n else 0.0
    avg_s = (spmv_ms / spmv_n) * 1e-3 - marker_s if spmv_n else float("nan")
    achieved = bytes_per_launch / avg_s / 1e9 if spmv_n else None
    from clothsimulator_amd._cabi import CS_FLAG_INSTANCES, CS_PRECOND_AMG, CS_PRECOND_BLOCK_JACOBI
    batched = bool(spec.params.flags & CS_FLAG_INSTANCES) and len(spec.cloths) > 1 \
        and spec.params.precond in (CS_PRECOND_BLOCK_JACOBI, CS_PRECOND_AMG)  # the batched solve of instances
    roofline = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": (achieved / HBM_PEAK_GBS) if achieved else None,
                "traffic": read_traffic(a.workload, a.precond),
                "kernel": "k_bpcg_spmv" if batched else "k_pcg_spmv",
                "avg_us": avg_s * 1e6, "event_marker_us": marker_s * 1e6,
                "launches_sampled": spmv_n, "bytes_per_launch": bytes_per_launch}
    ms = summary(main_run)
    out = {
        "metric": METRIC,
        "value": ms["value"],
        "unit": "node-steps/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": a.warmup,
        "ms_per_step": ms["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": spec.name, "nodes": n_nodes, "faces": int(c0.mesh.n_faces), "dt": spec.params.dt,
                   "precond": PRECONDS[a.precond], "rel_tol": 1e-5, "obstacles": len(spec.obstacles),
                   "steps_in_run": main_run["span"],
                   "timed_steps": "all" if main_run["timed"] >= main_run["span"] else
                   "%d of the run's %d steps, spread evenly (every %g-th)" % (
                       main_run["timed"], main_run["span"], main_run["span"] / main_run["timed"])},
        "cg_iters_mean": ms["cg_iters_mean"],
        "cg_iters_max": ms["cg_iters_max"],
        "cg_iters_learned": ms["cg_iters_mean"] if a.precond == "learned" else (
            summary(others["learned"])["cg_iters_mean"] if "learned" in others else None),
        "cg_iters_jacobi": ms["cg_iters_mean"] if a.precond == "jacobi" else (
            summary(others["jacobi"])["cg_iters_mean"] if "jacobi" in others else None),
        "failed_steps": ms["failed_steps"],
        "failures": ms["failures"],
        "roofline": roofline,
    }
    for p, r in others.items():
        out[PRECONDS[p]] = summary(r)
    if a.persistent and a.workload.startswith("batch"):
        from clothsimulator_amd._cabi import CS_FLAG_PERSISTENT_BATCH
        out["block_jacobi_persistent"] = summary(run(a.workload, "jacobi", steps, a.warmup, False,
                                                     CS_FLAG_PERSISTENT_BATCH))
    if a.workload.startswith("batch") and a.sequential_sample > 0:
        out["sequential"] = sequential_sample(spec, a.precond, steps, a.sequential_sample, ms["value"])
        out["config"]["instances"] = len(spec.cloths)
        rows, chunk, trials = main_run["batch_tuning"]
        out["batch_launch"] = {"rows_per_block": rows, "iterations_per_poll": chunk, "solves_timed_by_autotuner": trials}
    if "jacobi" in others:
        out["speedup_vs_jacobi"] = ms["value"] / others["jacobi"]["value"]
    if not a.no_cpu:
        print("cpu baseline: %d step(s) of %s on the host" % (a.cpu_steps, a.workload), file=sys.stderr, flush=True)
        out["cpu_baseline"] = cpu_baseline(scenes.by_name(a.workload), a.cpu_steps)
        print("cpu reference column: hang33, 100 steps, direct solve on one thread", file=sys.stderr, flush=True)
        out["cpu_reference_direct"] = cpu_reference_direct()
        for item in [w for w in a.cpu_direct.split(",") if w]:
            wl, k = item.split(":")
            print("cpu reference column: %s, %s steps, direct solve on one thread" % (wl, k), file=sys.stderr,
                  flush=True)
            out.setdefault("cpu_reference_direct_windows", []).append(cpu_reference_direct(wl, int(k)))
    print(json.dumps(out))


if __name__ == "__main__":
    main()
