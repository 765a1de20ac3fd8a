"""Learned warm start for the DFSPH pressure and divergence solves.

The reference starts each solve from half its previous-step stiffness
(TimeStepDFSPH.cpp:292-296 pressure, :449-453 divergence). Here a per-particle
network (6 -> 64 -> 64 -> 1, SiLU, BF16 MFMA with FP32 accumulation, the
device kernels in csrc/mlp.hip) predicts the converged stiffness of the
current step from neighbour-aggregated features (as a change from the
previous step's stiffness, which the start adds back), and the first Jacobi
iteration keeps, per particle, whichever start has the smaller initial
residual. The solve still iterates to the configured tolerance, so the
prediction only changes how many iterations that takes.

Features (computed on the device, csrc/dfsph.hip k_features), per particle i:
  0  100 * s_i                 source term: 1 - rho_adv (pressure) or
                               rho_dot * h (divergence)
  1  100 * <s>_W               its kernel-weighted neighbourhood mean
  2  kappa_prev_i / kref       previous-step stiffness (p h^2 or p_v h)
  3  <kappa_prev>_W / kref     its neighbourhood mean
  4  V_b / V                   boundary volume relative to the particle's
  5  n_neighbours / 40
with kref = the step's largest previous stiffness. The target is the
converged stiffness of the step divided by the same kref.

Training data are the solver's own states: `collect` runs the plain solver
with sample recording switched on; `train` fits the network with the
device's forward, backward and Adam kernels.
"""
import os
import time

import numpy as np

from . import capi

SOLVERS = ("pressure", "divergence")
NPARAM = capi.MLP_NPARAM
WEIGHTS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "weights", "warm_start.npz")


def init_theta(seed=0):
    """He-normal weights, zero biases, in the flat layout W1 (6x64), b1, W2
    (64x64), b2, W3 (64), b3."""
    rng = np.random.default_rng(seed)
    th = np.zeros(NPARAM, np.float64)
    th[0:384] = rng.normal(0.0, np.sqrt(2.0 / 6), 384)
    th[448:4544] = rng.normal(0.0, np.sqrt(2.0 / 64), 4096)
    th[4608:4672] = rng.normal(0.0, np.sqrt(1.0 / 64), 64)
    return th.astype(np.float32)


def make_sim(scene, thetas=None, mode=None):
    """A TimeStepDFSPH set up from a scene dict (scenes.py) and initialised;
    `thetas` maps solver name -> network parameters (warm-start mode
    "network"); `mode` ("previous", "reference") sets both solvers'
    warm-start mode instead."""
    sim = capi.TimeStepDFSPH(scene["params"])
    sim.set_fluid(scene["x"], scene["v"])
    if "boundary_x" in scene:
        sim.add_boundary_particles(scene["boundary_x"])
    if "vm_res" in scene:
        sim.set_volume_map_box(scene["box_lo"], scene["box_hi"], scene["vm_res"])
    for sph in scene.get("spheres", ()):
        sim.add_volume_map_sphere(sph[:3], sph[3], scene["sphere_res"])
    sim.init()
    for name, th in (thetas or {}).items():
        sim.set_warm_start_model(name, th)
    if mode is not None:
        for name in SOLVERS:
            sim.set_warm_start_mode(name, mode)
    return sim


def collect(scene, steps, skip=0, stride=8, capacity=1 << 20, thetas=None, log=None):
    """Steps `scene` for `skip` + `steps` steps and returns, per solver, the
    (features, targets) of every `stride`-th particle of the last `steps`
    steps (sample positions rotate from step to step)."""
    sim = make_sim(scene, thetas)
    if skip:
        sim.step(skip)
    out = {s: ([], []) for s in SOLVERS}
    per_step = (len(scene["x"]) + stride - 1) // stride
    chunk = max(1, min(steps, capacity // max(per_step, 1)))
    for s in SOLVERS:
        sim.collect_training(s, stride, capacity)
    done = 0
    while done < steps:
        k = min(chunk, steps - done)
        sim.step(k)
        done += k
        for s in SOLVERS:
            x, t = sim.get_training(s)
            out[s][0].append(x)
            out[s][1].append(t)
    if log:
        c = sim.counters()
        log(f"collect: {len(scene['x'])} particles, {skip}+{steps} steps, "
            f"mean iterations {c['total_iterations'] / max(c['steps'], 1):.2f} / "
            f"{c['total_iterations_v'] / max(c['steps'], 1):.2f}")
    del sim
    return {s: (np.concatenate(out[s][0]), np.concatenate(out[s][1])) for s in SOLVERS}


def merge(samples):
    """Concatenates a list of collect() results."""
    return {s: (np.concatenate([d[s][0] for d in samples]), np.concatenate([d[s][1] for d in samples]))
            for s in SOLVERS}


def fold_normalisation(theta, mu, sd):
    """Parameters for raw inputs x of a network trained on (x - mu) / sd:
    W1[k][o] / sd[k], b1[o] - sum_k mu[k] W1[k][o] / sd[k] (exact in FP32
    because layer 1 runs in FP32 on the device)."""
    th = np.asarray(theta, np.float64).copy()
    W1 = th[0:384].reshape(6, 64) / np.asarray(sd, np.float64)[:, None]
    th[0:384] = W1.ravel()
    th[384:448] -= np.asarray(mu, np.float64) @ W1
    return th.astype(np.float32)


def train(samples, steps=4000, batch=8192, lr=2e-3, seed=0):
    """Fits one network per solver on the device (mean squared error, Adam)
    to the change of the stiffness, t - input 2 (the start adds the previous
    stiffness back), on standardised inputs; the standardisation is folded
    into the first layer. Returns ({solver: theta}, {solver: final loss})."""
    out, losses = {}, {}
    for k, s in enumerate(SOLVERS):
        x, t = samples[s]
        if len(x) == 0:
            continue
        x6 = np.asarray(x[:, :6], np.float64)
        mu = x6.mean(0)
        sd = x6.std(0)
        sd = np.where(sd > 1e-6, sd, 1.0)
        xs = ((x6 - mu) / sd).astype(np.float32)
        ts = (np.asarray(t, np.float64) - x6[:, 2]).astype(np.float32)
        th, loss = capi.mlp_train(init_theta(seed + k), xs, ts, steps=steps, batch=min(batch, len(x)), lr=lr,
                                  seed=seed + k)
        out[s], losses[s] = fold_normalisation(th, mu, sd), loss
    return out, losses


def save(path, thetas, **meta):
    os.makedirs(os.path.dirname(os.path.abspath(path)), exist_ok=True)
    arrays = {f"theta_{s}": np.asarray(th, np.float32) for s, th in thetas.items()}
    arrays.update({k: np.asarray(v) for k, v in meta.items()})
    np.savez(path, **arrays)


def load(path=WEIGHTS):
    """{solver: theta} from a save() file (plain arrays, no pickles)."""
    with np.load(path, allow_pickle=False) as f:
        return {s: f[f"theta_{s}"].astype(np.float32) for s in SOLVERS if f"theta_{s}" in f}


def evaluate(scene, thetas, steps, warmup=0, mode=None):
    """Runs `scene` for warmup + steps steps with the given networks (None =
    the reference's warm start, or `mode`) and returns iteration and timing
    figures of the last `steps` steps."""
    sim = make_sim(scene, thetas, mode)
    if warmup:
        sim.step(warmup)
    c0 = sim.counters()
    capi.synchronize()
    t0 = time.perf_counter()
    sim.step(steps)
    capi.synchronize()
    dt = time.perf_counter() - t0
    c1 = sim.counters()
    res = {"pressure_iterations": (c1["total_iterations"] - c0["total_iterations"]) / steps,
           "divergence_iterations": (c1["total_iterations_v"] - c0["total_iterations_v"]) / steps,
           "cg_iterations": (c1["total_visco_iterations"] - c0["total_visco_iterations"]) / steps,
           "ms_per_step": 1e3 * dt / steps, "sim_time": c1["time"] - c0["time"],
           "avg_density_err": c1["avg_density_err"], "flags": c1["flags"],
           "learned_particles": c1["learned_particles"], "learned_particles_v": c1["learned_particles_v"],
           "learned_solves": c1["learned_solves"] - c0["learned_solves"],
           "avg_overpressure_err": c1["avg_overpressure_err"]}
    # state at the end, to compare the motion of runs
    v = sim.get("velocity")
    speed2 = (v.astype(np.float64) ** 2).sum(1)
    res.update({"mean_height": float(sim.get("position")[:, 1].mean()), "kinetic_energy": float(0.5 * speed2.mean()),
                "max_speed": float(np.sqrt(speed2.max()))})
    del sim
    return res


FIELDS = ("position", "velocity", "p / rho^2", "p_v / rho^2", "velocity difference")


def evaluate_snapshots(scene, arms, warmup=100, snapshots=20, every=5):
    """Solve-level comparison from identical states: steps `scene` with the
    reference's start, and at `snapshots` states (every `every` steps after
    `warmup`) runs one step of each arm ({name: thetas or None}) from that
    exact state (an arm may also be a warm-start mode name such as
    "previous"). Returns {name: {"pressure_iterations": mean, ...}}. The
    arms other than the reference use their start in every solve."""
    ref = make_sim(scene)
    sims = {}
    for name, th in arms.items():
        if isinstance(th, str):
            sims[name] = make_sim(scene, None, th)
        else:
            sims[name] = make_sim(scene, th)
        if th:
            for sv in SOLVERS:
                sims[name].set_warm_start_policy(sv, 0)
    ref.step(warmup)
    acc = {name: {"pressure_iterations": [], "divergence_iterations": [], "avg_overpressure_err": []}
           for name in arms}
    for k in range(snapshots):
        state = {f: ref.get(f) for f in FIELDS}
        c = ref.counters()
        for name, sim in sims.items():
            for f, v in state.items():
                sim.set(f, v)
            sim.set_time(c["time"], c["time_step_size"])
            sim.step(1)
            cc = sim.counters()
            acc[name]["pressure_iterations"].append(cc["iterations"])
            acc[name]["divergence_iterations"].append(cc["iterations_v"])
            acc[name]["avg_overpressure_err"].append(cc["avg_overpressure_err"])
        ref.step(every)
    return {name: {key: float(np.mean(v)) for key, v in d.items()} for name, d in acc.items()}
