"""DFSPH particle-steps/s on one MI355X.

Workload: the 1,000,000-particle dam break (100^3 lattice at r = 0.025 in a
[0,16]x[0,6]x[0,6] box wall represented by a Bender2019 volume map with
0.05 m cells, Weiler2018 viscosity nu = 0.01 with CG tolerance 1e-4, XSPH
0.01, CFL factor 1 with dt <= 5 ms, max density error 0.01 %, max divergence
error 0.1 %). A step is one TimeStepDFSPH::step() of the whole scene with all
data resident in HBM. Prints one JSON line.

  python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu-baseline]
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0

# Algorithmic bytes per particle of each kernel: 12 B per vec3, 4 B per
# scalar or neighbour index, neighbour-side reads counted once per particle
# (they are the same arrays, re-read from cache); K = mean fluid neighbours.
# Volume-map boundary record (x_b, V_b) = 16 B.
BYTES = {
    "fluid_lists": lambda K: 4 * K + 16,
    "density_factor_divsetup": lambda K: 4 * K + 64,
    "density_factor": lambda K: 4 * K + 48,
    "pressure_accel": lambda K: 4 * K + 48,
    "pressure_accel_v": lambda K: 4 * K + 48,
    "jacobi": lambda K: 4 * K + 60,
    "jacobi_v": lambda K: 4 * K + 60,
    "divergence_finish": lambda K: 4 * K + 80,
    "pressure_finish": lambda K: 4 * K + 72,
    "post_forces": lambda K: 4 * K + 104,
    "pressure_setup": lambda K: 4 * K + 68,
    "weiler_diag": lambda K: 4 * K + 96,
    "cg_init": lambda K: 4 * K + 120,
    "cg_matvec": lambda K: 4 * K + 60,
    "cg_update": lambda K: 108,
    "cg_pupdate": lambda K: 36,
    "volume_map_sample": lambda K: 156,
    "finalize": lambda K: 68,
    "permute": lambda K: 124,
    "cell_hist": lambda K: 20,
    "cell_scatter": lambda K: 12,
}


# bench kernel names -> device kernel symbols in the rocprofv3 traffic record
SYMBOL = {"jacobi": "k_jacobi", "jacobi_v": "k_jacobi", "pressure_accel": "k_accel", "pressure_accel_v": "k_accel",
          "fluid_lists": "k_lists_cm"}


def measured_traffic(kernel):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 PMC
    record (FETCH_SIZE/WRITE_SIZE scaled by a copy calibration; the commands
    are in profiles/dam1m_traffic_k_jacobi.md), or None."""
    path = os.path.join(ROOT, "profiles", "dam1m_traffic.json")
    try:
        rec = json.load(open(path))
        return rec["kernels"][SYMBOL[kernel]]["hbm_bytes_per_launch"]
    except (OSError, KeyError, ValueError):
        return None


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def cpu_baseline(sim, sc, steps):
    """The FP64 oracle (OpenMP) timed on a bounded sample of the same
    workload: the same scene, started from the device state after the
    warm-up (so both see the same solver iteration regime)."""
    import numpy as np

    from oracle import oracle
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    o = oracle.OracleDFSPH(sc["params"])
    x = sim.get("position").astype(np.float64)
    v = sim.get("velocity").astype(np.float64)
    o.set_fluid(x, v)
    o.set_volume_map_box(sc["box_lo"], sc["box_hi"], sc["vm_res"])
    o.init()
    c = sim.counters()
    o.set_time(c["time"], c["time_step_size"])
    # untimed: build the volume-map cells the particles touch
    o.stage("neighborhood_search")
    o.stage("volume_maps")
    o.set("position", x)
    o.set("velocity", v)
    for nm in ("p / rho^2", "p_v / rho^2", "velocity difference"):
        o.set(nm, sim.get(nm).astype(np.float64))
    t0 = time.perf_counter()
    o.step(steps)
    dt = time.perf_counter() - t0
    oc = o.counters()
    n = len(x)
    return {"value": n * steps / dt, "unit": "particle-steps/s", "cores": threads, "kind": "port",
            "sample": f"FP64 OpenMP oracle, same {n}-particle scene from the device state after warm-up, "
                      f"{steps} step(s) in {dt:.1f} s (pressure/divergence/CG iterations "
                      f"{oc['iterations']}/{oc['iterations_v']}/{oc['visco_iterations']})"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--per-side", type=int, default=100)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-steps", type=int, default=1)
    ap.add_argument("--no-timing", action="store_true", help="skip per-kernel HIP-event timing")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              f"(a multi-GPU run would shard the particles by slabs, see DESIGN.md)", file=sys.stderr)
        sys.exit(2)

    from splishsplash_amd import capi, scenes

    sc = scenes.dam_break_1m(per_side=args.per_side)
    n = len(sc["x"])
    log(f"bench: {n} particles, volume map {list(sc['vm_res'])} cells")
    sim = capi.TimeStepDFSPH(sc["params"])
    sim.set_fluid(sc["x"], sc["v"])
    sim.set_volume_map_box(sc["box_lo"], sc["box_hi"], sc["vm_res"])
    t0 = time.perf_counter()
    sim.init()
    log(f"bench: init (volume map build) {time.perf_counter() - t0:.2f} s")
    t0 = time.perf_counter()
    sim.step(args.warmup)
    log(f"bench: {args.warmup} warm-up steps {time.perf_counter() - t0:.2f} s")
    c0 = sim.counters()
    if not args.no_timing:
        sim.enable_timing(True)
        sim.reset_kernel_stats()
    capi.synchronize()
    t0 = time.perf_counter()
    sim.step(args.steps)
    capi.synchronize()
    elapsed = time.perf_counter() - t0
    c1 = sim.counters()
    value = n * args.steps / elapsed
    log(f"bench: {args.steps} timed steps {elapsed:.3f} s -> {value:.4g} particle-steps/s")

    it_p = (c1["total_iterations"] - c0["total_iterations"]) / args.steps
    it_v = (c1["total_iterations_v"] - c0["total_iterations_v"]) / args.steps
    it_cg = (c1["total_visco_iterations"] - c0["total_visco_iterations"]) / args.steps
    K = float(sim.get("fluid neighbor count").mean())

    roofline = None
    kernels = {}
    if not args.no_timing:
        stats = sim.kernel_stats()
        for name, (launches, ms) in stats.items():
            if launches:
                kernels[name] = {"launches": launches, "avg_us": 1e3 * ms / launches, "total_ms": ms}
        best = None
        for name, s in kernels.items():
            if name in BYTES and (best is None or s["total_ms"] > kernels[best]["total_ms"]):
                best = name
        if best:
            bpp = BYTES[best](K)
            achieved = n * bpp / (kernels[best]["avg_us"] * 1e-6) / 1e9
            roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                        "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": measured_traffic(best), "kernel": best,
                        "bytes_per_particle": bpp, "avg_launch_us": round(kernels[best]["avg_us"], 2)}
            for name, s in kernels.items():
                if name in BYTES:
                    s["gbs"] = round(n * BYTES[name](K) / (s["avg_us"] * 1e-6) / 1e9, 1)

    cpu = None
    if not args.no_cpu_baseline:
        log("bench: CPU baseline (oracle) ...")
        cpu = cpu_baseline(sim, sc, args.cpu_steps)
        log(f"bench: CPU baseline {cpu['value']:.4g} particle-steps/s")

    out = {
        "metric": "DFSPH particle-steps/s at 1M particles, 0.01% density err; learned vs plain",
        "value": value,
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * elapsed / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 (FP64 reductions)",
        "data": "synthetic",
        "config": {"workload": "dam break, 1e6 particles (100^3 lattice, r=0.025) in [0,16]x[0,6]x[0,6], "
                               "Bender2019 volume map 0.05 m cells, Weiler2018 nu=0.01 tol 1e-4, XSPH 0.01, "
                               "CFL 0.4, maxError 0.01%, maxErrorV 0.1%; plain warm start",
                   "particles": n, "parallelism": "single GPU"},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "solver": {"pressure_iterations_per_step": it_p, "divergence_iterations_per_step": it_v,
                   "cg_iterations_per_step": it_cg, "mean_fluid_neighbors": K,
                   "avg_density_err": c1["avg_density_err"], "time_step_size": c1["time_step_size"],
                   "sim_time": c1["time"], "flags": c1["flags"]},
        "kernels": kernels,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
