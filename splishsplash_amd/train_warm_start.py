"""Trains the warm-start networks on the solver's own states and evaluates
them against the reference's warm start.

    python -m splishsplash_amd.train_warm_start [--out PATH] [--train 32]
        [--heldout 8] [--eval-per-side 100] [--report PATH]

Training scenes are seeded dam-break variants (scenes.dam_break_variant:
fluid block of 60-120 particles per side, box aspect 1.5-3, 0-8 static
spheres), seeds 0..train-1, stepped with the reference's warm start while the
device records (features, converged stiffness) samples. Both networks are
fitted on the device (BF16 MFMA forward/backward, Adam). Held-out seeds
1000.. are then run at `eval-per-side`^3 particles with and without the
networks, and iterations and step times are reported.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

from . import capi, learned, scenes


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def collect_variant(seed, skip, steps, samples_per_step_target, per_side=None):
    sc = scenes.dam_break_variant(seed, per_side=per_side)
    n = len(sc["x"])
    stride = max(1, n // max(samples_per_step_target, 1))
    t0 = time.perf_counter()
    d = learned.collect(sc, steps, skip=skip, stride=stride, capacity=(n // stride + 1) * steps + 1)
    v = sc["variant"]
    log(f"train seed {seed}: {n} particles, aspect {v['aspect']:.2f}, {len(v['spheres'])} spheres, "
        f"{len(d['pressure'][0])}/{len(d['divergence'][0])} samples, {time.perf_counter() - t0:.1f} s")
    return d


def evaluate_pair(sc, thetas, warmup, steps):
    plain = learned.evaluate(sc, None, steps=steps, warmup=warmup)
    lrn = learned.evaluate(sc, thetas, steps=steps, warmup=warmup)
    # the previous step's full stiffness as the start (no network): what the
    # network adds beyond that
    prev = learned.evaluate(sc, None, steps=steps, warmup=warmup, mode="previous")
    return plain, lrn, prev


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default=learned.WEIGHTS)
    ap.add_argument("--report", default=None)
    ap.add_argument("--train", type=int, default=32)
    ap.add_argument("--heldout", type=int, default=8)
    ap.add_argument("--skip", type=int, default=60)
    ap.add_argument("--steps", type=int, default=240)
    ap.add_argument("--samples-per-step", type=int, default=1500)
    ap.add_argument("--train-per-side", type=int, default=None, help="override the seeded block size")
    ap.add_argument("--epochs", type=int, default=6000, help="Adam steps per network")
    ap.add_argument("--batch", type=int, default=16384)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--eval-per-side", type=int, default=100)
    ap.add_argument("--eval-warmup", type=int, default=150)
    ap.add_argument("--eval-steps", type=int, default=150)
    args = ap.parse_args(argv)

    t0 = time.perf_counter()
    data = [collect_variant(s, args.skip, args.steps, args.samples_per_step, args.train_per_side)
            for s in range(args.train)]
    samples = learned.merge(data)
    del data
    log(f"collected {len(samples['pressure'][0])} pressure / {len(samples['divergence'][0])} divergence samples "
        f"in {time.perf_counter() - t0:.1f} s")
    t1 = time.perf_counter()
    thetas, losses = learned.train(samples, steps=args.epochs, batch=args.batch, lr=args.lr)
    log(f"trained in {time.perf_counter() - t1:.1f} s, loss {losses}")
    fit = {}
    for s in learned.SOLVERS:
        x, t = samples[s]
        y = capi.mlp_forward(thetas[s], x[:, :6]) + x[:, 2]
        fit[s] = {"mse": float(np.mean((np.maximum(y, 0) - t) ** 2)), "target_var": float(np.var(t)),
                  "reference_start_mse": float(np.mean((t - 0.5 * x[:, 2]) ** 2)),
                  "previous_stiffness_mse": float(np.mean((t - x[:, 2]) ** 2)), "samples": int(len(t))}
    learned.save(args.out, thetas, train_seeds=np.arange(args.train), epochs=args.epochs, lr=args.lr)
    log(f"saved {args.out}")

    evals = []
    for k in range(args.heldout):
        seed = 1000 + k
        sc = scenes.dam_break_variant(seed, per_side=args.eval_per_side)
        plain, lrn, prev = evaluate_pair(sc, thetas, args.eval_warmup, args.eval_steps)
        snap = learned.evaluate_snapshots(sc, {"plain": None, "learned": thetas, "previous": "previous"},
                                          warmup=args.eval_warmup, snapshots=15, every=10)
        v = sc["variant"]
        rec = {"seed": seed, "particles": len(sc["x"]), "aspect": v["aspect"], "spheres": len(v["spheres"]),
               "plain": plain, "learned": lrn, "previous_stiffness_start": prev, "from_identical_states": snap}
        evals.append(rec)
        log(f"held-out seed {seed}: pressure it {plain['pressure_iterations']:.2f} -> "
            f"{lrn['pressure_iterations']:.2f} (previous-stiffness start {prev['pressure_iterations']:.2f}), "
            f"divergence it {plain['divergence_iterations']:.2f} -> {lrn['divergence_iterations']:.2f}, "
            f"ms/step {plain['ms_per_step']:.3f} -> {lrn['ms_per_step']:.3f} ({prev['ms_per_step']:.3f})")
    summary = {}
    if evals:
        for key in ("pressure_iterations", "divergence_iterations", "ms_per_step"):
            p = np.mean([e["plain"][key] for e in evals])
            q = np.mean([e["learned"][key] for e in evals])
            z = np.mean([e["previous_stiffness_start"][key] for e in evals])
            summary[key] = {"plain": p, "learned": q, "reduction": 1.0 - q / p, "previous_stiffness_start": z}
        summary["pressure_iterations_from_identical_states"] = {
            arm: float(np.mean([e["from_identical_states"][arm]["pressure_iterations"] for e in evals]))
            for arm in ("plain", "learned", "previous")}
    report = {"fit": fit, "heldout": evals, "summary": summary,
              "settings": {k: v for k, v in vars(args).items() if k not in ("out", "report")}}
    if args.report:
        with open(args.report, "w") as f:
            json.dump(report, f, indent=1)
    print(json.dumps({"summary": summary, "fit": fit}))


if __name__ == "__main__":
    main()
