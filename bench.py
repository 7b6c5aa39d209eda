"""Benchmark: particle-substeps/s of the adaptive GIMP substep on one MI355X.

Workload (the headline configuration): 3D dynamic adaptivity, 3 levels, coarsest 64
(finest dx = 1/256), 1,000,000 particles in a jittered-lattice sphere falling onto a slab,
snow plasticity with the FP64 SVD (adaptive-gimp_amd/scenes.py ``c3_sphere_on_slab``).
A step is one full substep (refine + level mapping + P2G + grid + G2P) over all particles,
with every input resident in HBM.

Prints ONE JSON line with ``value`` = particle-substeps/s, the roofline of the dominant
transfer kernel (P2G or G2P; algorithmic bytes / average kernel time measured with HIP
events on the library's stream) and a CPU baseline (the oracle's C++ OpenMP restatement on
a bounded sample of the same workload, rank 0, N=1).

    python bench.py [--steps K] [--warmup W] [--workload c3|c5|c2|c4] [--tune] [--no-cpu]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "hbm_traffic.json")


def measured_traffic(kernel):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 --pmc summary
    (FETCH_SIZE + WRITE_SIZE, scaled by a known-size copy), or None."""
    try:
        with open(TRAFFIC_FILE) as f:
            t = json.load(f)
        return t["kernels"][kernel]["hbm_bytes_per_launch"]
    except (OSError, KeyError, ValueError):
        return None


def algorithmic_bytes(dim, n_particles, n_nodes):
    """Bytes per launch of P2G and G2P (FP32): P2G reads x, v, F-derived stress, C, m per
    particle (4(2d + 2d^2 + 1) B) and writes (m, momentum) per node (4(d+1) B); G2P reads x,
    F (4(d + d^2) B), writes x, v, F, C (4(2d + 2d^2) B) per particle and reads v per node
    (4d B)."""
    d = dim
    p2g = n_particles * 4 * (2 * d + 2 * d * d + 1) + n_nodes * 4 * (d + 1)
    g2p = n_particles * 4 * (d + d * d) + n_particles * 4 * (2 * d + 2 * d * d) + n_nodes * 4 * d
    return p2g, g2p


def _oracle_for(wl, threads, fp64=False):
    from oracle import oracle as O

    ora = O.OracleSim(wl.dim, wl.level, wl.coarsest_size, wl.n, fp64=fp64, threads=threads, **wl.params)
    st = wl.state
    ora.set_particles(st["x"], st["v"], st["F"], st["C"], st["m"], st["g"])
    if wl.grid_initializer is not None:
        import __graft_entry__ as ge

        ret, act = ge.package().scenes.static_level_maps(wl.grid_initializer, wl.dim, wl.level, wl.coarsest_size)
        ora.set_static_map(ret, act)
    return ora


def host_threads():
    """Threads for the CPU legs: OMP_NUM_THREADS when set (the GPU box sets it to the CPU share
    of one GPU), else every CPU this process may run on."""
    omp = int(os.environ.get("OMP_NUM_THREADS", "0") or 0)
    try:
        aff = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        aff = os.cpu_count() or 1
    return max(1, omp or aff), {"os_cpu_count": os.cpu_count(), "affinity_cpus": aff,
                                "OMP_NUM_THREADS": os.environ.get("OMP_NUM_THREADS")}


def _time_oracle(workloads, threads, per_scene_s, max_sub=50):
    work, elapsed, nsub = 0.0, 0.0, 0
    for wl in workloads:
        ora = _oracle_for(wl, threads)
        ora.substep(wl.dt)  # warm-up
        t_end = time.perf_counter() + per_scene_s
        k = 0
        t0 = time.perf_counter()
        while k < 1 or time.perf_counter() < t_end:
            ora.substep(wl.dt)
            k += 1
            if k >= max_sub:
                break
        elapsed += time.perf_counter() - t0
        work += wl.n * k
        nsub += k
        del ora
    return work, elapsed, nsub


def cpu_baseline(workloads, seconds_target=12.0, seconds_1t=8.0):
    """The oracle (float32 + FP64 SVD, C++ OpenMP) timed on this host's cores: per scene one
    warm-up substep, then substeps until ~seconds_target have elapsed over all scenes; then the
    same on one thread for ~seconds_1t (at least one substep per scene).  Test infrastructure,
    used here only as the reported baseline."""
    threads, host = host_threads()
    work, elapsed, nsub = _time_oracle(workloads, threads, seconds_target / len(workloads))
    w1, e1, n1 = _time_oracle(workloads[:1], 1, seconds_1t, max_sub=4)
    names = workloads[0].name if len(workloads) == 1 else f"{len(workloads)} scenes ({workloads[0].name}...)"
    return {"value": work / elapsed, "unit": "particle-substeps/s", "cores": threads, "kind": "port",
            "sample": f"{nsub} substeps of {names} after one warm-up substep each "
                      f"({work / 1e6:.1f} M particle-substeps in {elapsed:.1f} s on {threads} threads; oracle C++ "
                      "OpenMP restatement, float32 + FP64 SVD)",
            "host": host,
            "one_thread": {"value": w1 / e1, "substeps": n1, "seconds": round(e1, 1),
                           "scene": workloads[0].name}}


# kernels of each substep phase (launch names of the library's per-kernel timer)
PHASE_KERNELS = {
    "refine": ("k_clear_sgrid", "k_clear_refine", "k_activate_static", "k_activate_dynamic", "k_compact_cells",
               "k_cull", "k_node_A", "k_mark_TG", "k_copy_flags"),
    "bucket": ("k_bucket_prep", "k_levelmap", "k_scan_reduce", "k_scan_final", "k_scatter_blocks"),
    "p2g": ("k_p2g",),
    "grid": ("k_restrict", "k_gridop", "k_prolong"),
    "g2p": ("k_g2p",),
}

METRIC = "particle-substeps/s (3D 1M p, 3 levels) plain vs learned-adapt; HBM GB/s"


def _threads():
    return host_threads()[0]


def _new_sim(pkg, wl, ppc):
    return pkg.AdaptiveGIMP(wl.dim, wl.level, wl.coarsest_size, wl.n, wl.initializer(), wl.grid_initializer,
                            particles_per_chunk=ppc or None, **wl.params)


def _learned_sim(pkg, wl, ppc, model, parts, tau=0.3, window=0):
    sim = _new_sim(pkg, wl, ppc)
    sim.learned_enable(window or model[1].get("window", 20), tau)
    sim.learned_set_model(model[0])
    sim.learned_set_parts(network="network" in parts, sweep="sweep" in parts)
    return sim


def bench_learned(pkg, wl, args, np, model, plain_x):
    """The same workload with predicted octrees (agimp_learned_advance): throughput over the
    same warmup + steps from the same initial state, rebuild statistics, and the positions'
    deviation from the plain run at the end."""
    sim = _learned_sim(pkg, wl, args.ppc, model, args.learned_parts, args.tau, args.window)
    sim.advance_learned(wl.dt, args.warmup)
    sim.sync()
    sim.learned_reset_stats()
    t0 = time.perf_counter()
    sim.advance_learned(wl.dt, args.steps)
    sim.sync()
    elapsed = time.perf_counter() - t0
    stats = sim.learned_stats()
    dev = np.linalg.norm(sim.x_p.to_numpy().astype(np.float64) - plain_x, axis=1)
    counters = sim.counters()
    sim.close()
    return elapsed, stats, dev, counters


def _worst_particles(np, wl, x, v, fl, ref, k=100):
    """Where the largest velocity errors sit: for the k particles with the largest |v - v_ref|,
    the share whose transfer level differs from the oracle's, whose position lies within two
    cells (of their level) of a wall band or of the slab faces, or within 1e-3 cell of a face of
    their cell (where a rounding difference moves the base cell); and the worst one itself."""
    ev = np.linalg.norm(v - ref["v"], axis=1)
    top = np.argsort(ev)[-k:][::-1]
    fs = wl.coarsest_size * 2 ** (wl.level - 1)
    gap = wl.params.get("boundary_gap", 4)
    lvl = np.where(ref["fl"] >= 0, ref["fl"], wl.level - 1)
    dxl = 1.0 / (wl.coarsest_size * 2.0 ** lvl)
    q = x / dxl[:, None]
    face = np.abs(q - np.round(q)).min(axis=1)  # distance to the nearest cell face, in cells
    wall_d = np.minimum(x - gap / fs, (fs - gap) / fs - x).min(axis=1) / dxl
    slab = wl.params.get("slab")
    slab_d = (np.minimum(np.abs(x[:, slab[0]] - slab[1]), np.abs(x[:, slab[0]] - slab[2])) / dxl
              if slab else np.full(x.shape[0], np.inf))
    i = int(top[0])
    return {
        "k": k,
        "share_fl_differs": float(np.mean(fl[top] != ref["fl"][top])),
        "share_near_wall": float(np.mean(wall_d[top] < 2)),
        "share_near_slab": float(np.mean(slab_d[top] < 2)),
        "share_near_cell_face": float(np.mean(face[top] < 1e-3)),
        "worst": {"index": i, "x": [float(a) for a in x[i]], "v": [float(a) for a in v[i]],
                  "v_ref": [float(a) for a in ref["v"][i]], "fl": int(fl[i]), "fl_ref": int(ref["fl"][i]),
                  "cells_to_wall": float(wall_d[i]), "cells_to_slab": float(slab_d[i]),
                  "cells_to_face": float(face[i])},
    }


def _errors(np, wl, x, v, F, ref):
    ex = np.linalg.norm(x - ref["x"], axis=1)
    vs = max(float(np.abs(ref["v"]).max()), 1e-30)
    mom = np.sum(wl.state["m"][:, None] * v, axis=0)
    mom_ref = np.sum(wl.state["m"][:, None] * ref["v"], axis=0)
    ev = np.linalg.norm(v - ref["v"], axis=1) / vs
    return {"x_max_abs": float(ex.max()), "x_p99_abs": float(np.percentile(ex, 99)),
            "v_max_rel": float(ev.max()), "v_p99_rel": float(np.percentile(ev, 99)),
            "F_max_abs": float(np.abs(F - ref["F"]).max()),
            "momentum_rel": float(np.linalg.norm(mom - mom_ref) / max(np.linalg.norm(mom_ref), 1e-30))}


def accuracy(pkg, wl, args, np, model):
    """Positions and velocities of the plain and the learned GPU paths against the FP64 oracle
    (float64 particles and grid, FP64 SVD) after `n` substeps from the same float32 initial
    state, next to the noise floor: the float32 oracle against the same FP64 run (the same
    algorithm, only the precision and the summation order differ)."""
    n = args.accuracy_steps
    st = wl.state
    f32 = lambda a: np.asarray(a, np.float32)
    refs = {}
    t_ora = 0.0
    for name, fp64 in (("fp64", True), ("fp32", False)):
        ora = _oracle_for(wl, _threads(), fp64=fp64)
        ora.set_particles(f32(st["x"]), f32(st["v"]), f32(st["F"]), f32(st["C"]), f32(st["m"]), st["g"])
        t0 = time.perf_counter()
        for _ in range(n):
            ora.substep(wl.dt)
        t_ora += time.perf_counter() - t0
        refs[name] = ora.get_particles()
        del ora
    ref = refs["fp64"]
    out = {"substeps": n, "reference": "oracle FP64 (C++ restatement, double particles and grid)",
           "oracle_seconds": round(t_ora, 1),
           "noise_floor_fp32_oracle": _errors(np, wl, refs["fp32"]["x"], refs["fp32"]["v"], refs["fp32"]["F"], ref)}
    out["noise_floor_fp32_oracle"]["fl_differs"] = int(np.sum(refs["fp32"]["fl"] != ref["fl"]))
    runs = [("plain", None)] + ([("learned", model)] if model is not None else [])
    for name, mdl in runs:
        if mdl is None:
            sim = _new_sim(pkg, wl, args.ppc)
            for _ in range(n):
                sim.substep(wl.dt)
        else:
            sim = _learned_sim(pkg, wl, args.ppc, mdl, args.learned_parts, args.tau, args.window)
            sim.advance_learned(wl.dt, n)
        x = sim.x_p.to_numpy().astype(np.float64)
        v = sim.v_p.to_numpy().astype(np.float64)
        F = sim.F_p.to_numpy().astype(np.float64)
        fl = sim.fl_p.to_numpy()
        sim.close()
        out[name] = _errors(np, wl, x, v, F, ref)
        out[name]["fl_differs"] = int(np.sum(fl != ref["fl"]))
        out[name]["mass"] = "exact (particle masses are constant)"
        out[name]["largest_v_errors"] = _worst_particles(np, wl, x, v, fl, ref)
    out["noise_floor_fp32_oracle"]["largest_v_errors"] = _worst_particles(np, wl, refs["fp32"]["x"], refs["fp32"]["v"],
                                                                         refs["fp32"]["fl"], ref)
    return out


def bench_single(pkg, wl, args, np):
    sim = pkg.AdaptiveGIMP(wl.dim, wl.level, wl.coarsest_size, wl.n, wl.initializer(), wl.grid_initializer,
                           particles_per_chunk=args.ppc or None, **wl.params)
    args.ppc_used = int(sim._cfg.particles_per_chunk)
    tuned = None
    if args.tune:
        from importlib import import_module

        at = import_module(pkg.__name__ + ".autotune")
        tuned = at.tune(sim, wl.dt)
    sim.advance(wl.dt, args.warmup)
    sim.sync()
    t0 = time.perf_counter()
    sim.advance(wl.dt, args.steps)  # one captured substep replayed: the production path
    sim.sync()
    elapsed = time.perf_counter() - t0
    # per-phase breakdown: eager substeps with HIP events on the library's stream
    sim.timings()
    sim.kernel_timings()
    sim.set_timing(True)
    nt = min(args.steps, 10)
    for _ in range(nt):
        sim.substep(wl.dt)
    ktim = sim.kernel_timings()
    sim.timings()
    sim.set_timing(False)
    # per phase: the sum of its kernels' event-timed durations, per substep
    tim = {ph: [0.0, nt] for ph in PHASE_KERNELS}
    for k, (ms, calls, _, _) in ktim.items():
        for ph, names in PHASE_KERNELS.items():
            if k in names:
                tim[ph][0] += ms
    n_nodes = int(sum(int(np.count_nonzero(sim.ad_grid_m[l].to_numpy())) for l in range(wl.level)))
    counters = sim.counters()
    sim.close()
    # positions after warmup + steps substeps from the initial state (for the learned comparison)
    ref = _new_sim(pkg, wl, tuned[0] if tuned else args.ppc)
    ref.advance(wl.dt, args.warmup + args.steps)
    plain_x = ref.x_p.to_numpy().astype(np.float64)
    ref.close()
    return elapsed, tim, n_nodes, counters, tuned, plain_x, ktim


def bench_batch(pkg, wls, args):
    sims = [pkg.AdaptiveGIMP(w.dim, w.level, w.coarsest_size, w.n, w.initializer(), w.grid_initializer,
                             particles_per_chunk=256, **w.params) for w in wls]
    batch = pkg.Batch(sims)
    dts = [w.dt for w in wls]
    tuned = None
    if args.tune:
        from importlib import import_module

        at = import_module(pkg.__name__ + ".autotune")
        ch, log = at.tune_batch(batch, dts)
        tuned = {"chunks_by_class": sorted(set(int(c) for c in ch)), "trials": log}
    batch.advance(dts, args.warmup)
    t0 = time.perf_counter()
    batch.advance(dts, args.steps)
    elapsed = time.perf_counter() - t0
    counters = [s.counters() for s in sims]
    batch.close()
    for s in sims:
        s.close()
    return elapsed, counters, tuned


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--workload", default="c3", choices=["c3", "c5", "c2", "c4", "c1"],
                    help="c3: 3D 1M sphere (the headline); c5: 3D 4M three spheres; c2: 2D static band; "
                         "c4: 64 batched 2D scenes; c1: 2D two squares (~21k particles)")
    ap.add_argument("--n", type=int, default=0, help="particle count override (c3/c5)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--tune", action="store_true", help="autotune the P2G chunk size first")
    ap.add_argument("--ppc", type=int, default=0,
                    help="P2G particles per chunk (single-scene workloads; 0: the library default, 1024)")
    ap.add_argument("--mode", default="both", choices=["plain", "learned", "both"],
                    help="single-scene workloads: time the plain path, the learned path (predicted octrees, "
                         "adaptive-gimp_amd/refine_mlp.npz), or both (value = the plain path)")
    ap.add_argument("--window", type=int, default=0,
                    help="learned path: substeps per predicted octree (0: the model's training window)")
    ap.add_argument("--tau", type=float, default=0.3,
                    help="learned path: decision threshold on the cumulative class probability")
    ap.add_argument("--learned-parts", default="network,sweep",
                    help="parts of a predicted octree: network (unit levels), sweep (velocity-swept requests)")
    ap.add_argument("--accuracy-steps", type=int, default=-1,
                    help="substeps of the FP64-oracle accuracy leg (default: 20 for c3, 0 = off otherwise)")
    ap.add_argument("--copy-mb", type=int, default=0,
                    help="first run one plain device copy of this many MiB (a known-size kernel that calibrates "
                         "rocprofv3 FETCH_SIZE/WRITE_SIZE)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); the multi-GPU layout is a "
              "design note, not a built path", file=sys.stderr)
        sys.exit(2)

    import numpy as np

    import __graft_entry__ as ge

    pkg = ge.package()
    if args.copy_mb > 0:
        import torch

        a = torch.empty(args.copy_mb * (1 << 20) // 4, dtype=torch.float32, device="cuda")
        b = torch.empty_like(a)
        a.fill_(1.0)
        torch.cuda.synchronize()
        b.copy_(a)
        torch.cuda.synchronize()
        del a, b
    S = pkg.scenes
    notes = None
    if args.workload in ("c3", "c5"):
        wl = S.c3_sphere_on_slab(n=args.n or 1_000_000) if args.workload == "c3" else \
            S.c5_three_spheres(n=args.n or 4_000_000)
        wls = [wl]
    elif args.workload == "c2":
        wls = [S.c2_static_band(rho=1000.0)]
        notes = "rho=1000 instead of the stated 1 (rho=1 violates the P-wave CFL bound at dt=5e-5)"
    elif args.workload == "c1":
        wls = [S.c1_two_squares(rho=1000.0)]
        notes = "rho=1000 instead of the stated 1 (rho=1 violates the P-wave CFL bound at dt=1e-4)"
    else:
        wls = S.c4_batch(rho=1000.0)
        notes = "rho=1000 instead of the stated 1 (rho=1 violates the P-wave CFL bound at dt=5e-5)"
    total_n = sum(w.n for w in wls)
    line = {"metric": METRIC, "unit": "particle-substeps/s", "n_gpus": 1, "steps": args.steps,
            "warmup": args.warmup, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
            "dtype": "f32 (FP64 SVD/plasticity)", "data": "synthetic (seeded scene generators)"}
    if len(wls) == 1:
        wl = wls[0]
        model = None
        if args.mode in ("learned", "both") and wl.grid_initializer is None:
            from importlib import import_module

            lm = import_module(pkg.__name__ + ".learned")
            model = lm.load_model()
        elapsed, tim, n_nodes, counters, tuned, plain_x, ktim = bench_single(pkg, wl, args, np)
        b_p2g, b_g2p = algorithmic_bytes(wl.dim, wl.n, n_nodes)
        avg = {k: (v[0] / v[1] if v[1] else float("nan")) for k, v in tim.items()}
        kavg = {k: v[0] / v[1] for k, v in ktim.items() if v[1]}
        kern = {"p2g": (b_p2g, kavg["k_p2g"]), "g2p": (b_g2p, kavg["k_g2p"])}
        dom = max(kern, key=lambda k: kern[k][1])
        achieved = kern[dom][0] / (kern[dom][1] * 1e-3) / 1e9
        line.update({
            "value": total_n * args.steps / elapsed,
            "ms_per_step": elapsed * 1e3 / args.steps,
            "config": {"workload": wl.name, "dim": wl.dim, "levels": wl.level, "coarsest_size": wl.coarsest_size,
                       "n_particles": wl.n, "dt": wl.dt, "mode": "plain", "timed_path": "hipGraph replay",
                       "particles_per_chunk": tuned[0] if tuned else args.ppc_used},
            "roofline": {"bound": "hbm", "kernel": dom, "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                         "frac": achieved / HBM_PEAK_GBS,
                         "traffic": measured_traffic(dom) if wl.name.startswith("c3") else None,
                         "bytes_per_launch": kern[dom][0], "avg_kernel_ms": kern[dom][1], "active_nodes": n_nodes},
            "phase_ms_eager": {k: round(v, 4) for k, v in avg.items()},
            "kernel_ms_eager": {k: round(v, 4) for k, v in sorted(kavg.items(), key=lambda kv: -kv[1])},
            "p2g_gbs": b_p2g / (kavg["k_p2g"] * 1e-3) / 1e9,
            "g2p_gbs": b_g2p / (kavg["k_g2p"] * 1e-3) / 1e9,
            "counters": counters,
        })
        if tuned:
            line["autotune_s_per_substep"] = {str(k): v for k, v in tuned[1].items()}
        if model is not None:
            el, st, dev, lcnt = bench_learned(pkg, wl, args, np, model, plain_x)
            lv = total_n * args.steps / el
            line["learned"] = {
                "value": lv, "ms_per_step": el * 1e3 / args.steps, "speedup_vs_plain": lv / line["value"],
                "parts": args.learned_parts, "window": args.window or model[1].get("window", 20), "tau": args.tau,
                "stats": st,
                # particles whose check failed in substeps that were kept (the device counter; a
                # window rolled back restores it with the particles)
                "criterion_failures_executed": lcnt["criterion_failures"],
                "criterion_failures_detected": st["violations"],
                "x_dev_from_plain": {"max": float(dev.max()), "p99": float(np.percentile(dev, 99)),
                                     "after_substeps": args.warmup + args.steps},
                "counters": lcnt,
            }
            if args.mode == "learned":
                line["value_plain"] = line["value"]
                line["value"] = lv
                line["ms_per_step"] = el * 1e3 / args.steps
                line["config"]["mode"] = "learned"
        acc_steps = args.accuracy_steps if args.accuracy_steps >= 0 else (20 if wl.name.startswith("c3") else 0)
        if acc_steps > 0 and not args.no_cpu:
            args.accuracy_steps = acc_steps
            line["accuracy"] = accuracy(pkg, wl, args, np, model)
    else:
        elapsed, counters, tuned = bench_batch(pkg, wls, args)
        line.update({
            "value": total_n * args.steps / elapsed,
            "ms_per_step": elapsed * 1e3 / args.steps,
            "config": {"workload": "c4_batched_2d_64_scenes", "scenes": len(wls), "n_particles_total": total_n,
                       "levels": sorted(set(w.level for w in wls)), "mode": "plain",
                       "timed_path": "one hipGraph replay per batched substep"},
            "roofline": None,
            "counters_total": {k: sum(c[k] for c in counters) for k in counters[0]},
        })
        if tuned:
            line["autotune"] = tuned
    if notes:
        line["config"]["note"] = notes
    line["cpu_baseline"] = None if args.no_cpu else cpu_baseline(wls if len(wls) <= 8 else wls[:8])
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
