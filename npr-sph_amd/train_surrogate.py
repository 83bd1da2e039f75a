"""Train the learned coarse step's pair-message MLP on the GPU.

Data: plain-SPH rollouts of the dam1m scene family (1,000,000 particles, a
100^3 jittered lattice at spacing 0.5h in the box scaled to 40% fill), one
rollout per seed.  Each rollout starts from its seeded lattice, advances a
seeded number of plain steps, and then yields snapshots: the state S at time t
and the same particles 4 fine steps later.  Targets are the mean
accelerations over those 4 steps (sph_mlp_train).  All forward, backward and
optimiser work runs in the library's own kernels (libnprsph.so).

Usage:
  python npr-sph_amd/train_surrogate.py --rollouts 200 --out npr-sph_amd/weights/coarse_mlp.npz
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
sys.path.insert(0, ROOT)

NORM = dict(v_scale=0.1, rho_scale=1e-5, pi_scale=5e-4)


def glorot_init(seed: int) -> np.ndarray:
    """Scaled-uniform initialisation of W1[9][64], b1, W2[64][64], b2, W3[64][3] (x0.1), b3."""
    rng = np.random.default_rng(seed)
    parts = [rng.uniform(-1, 1, (9, 64)) * np.sqrt(6.0 / 73), np.zeros(64),
             rng.uniform(-1, 1, (64, 64)) * np.sqrt(6.0 / 128), np.zeros(64),
             rng.uniform(-1, 1, (64, 3)) * np.sqrt(6.0 / 67) * 0.1, np.zeros(3)]
    return np.concatenate([p.reshape(-1) for p in parts]).astype(np.float32)


def splitmix(seed: int, k: int) -> int:
    m = (1 << 64) - 1
    z = (seed + (k + 1) * 0x9E3779B97F4A7C15) & m
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & m
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & m
    return z ^ (z >> 31)


def target_scale(pkg, sim, later) -> float:
    """RMS of the 4-step mean accelerations of one snapshot (off-wall particles)."""
    a = sim.download()
    b = later.download()
    dt = float(np.float32(1.0) / np.float32(10000.0))
    acc = (b["vel"][:, :3].astype(np.float64) - a["vel"][:, :3]) / (4 * dt)
    acc[:, 1] -= float(np.float32(-9806.65))
    ok = np.all(np.isfinite(acc), axis=1)
    return float(np.sqrt(np.mean(acc[ok] ** 2)))


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--rollouts", type=int, default=200)
    ap.add_argument("--workload", default="dam1m")
    ap.add_argument("--max-start", type=int, default=800, help="rollout start times drawn from [0, max-start] steps")
    ap.add_argument("--snapshots", type=int, default=3, help="snapshots per rollout")
    ap.add_argument("--gap", type=int, default=60, help="plain steps between snapshots")
    ap.add_argument("--updates", type=int, default=2, help="Adam steps per snapshot (disjoint target subsets)")
    ap.add_argument("--stride", type=int, default=48, help="target subsampling stride over cell-sorted slots")
    ap.add_argument("--lr", type=float, default=3e-3)
    ap.add_argument("--lr-final", type=float, default=3e-4)
    ap.add_argument("--seed", type=int, default=2024)
    ap.add_argument("--out", default=os.path.join(HERE, "weights", "coarse_mlp.npz"))
    ap.add_argument("--time-budget", type=float, default=900.0, help="seconds; stops early when exceeded")
    args = ap.parse_args(argv)

    import __graft_entry__ as entry
    pkg = entry.load_package()
    sph = pkg.sph
    base = pkg.scenes.WORKLOADS[args.workload]()
    bnd = sph.make_boundary(base.lower, base.upper)
    sim = pkg.Simulation(base.count, boundaries=bnd)
    later = pkg.Simulation(base.count, boundaries=bnd)
    params = glorot_init(args.seed)
    norm = None
    log = []
    t_start = time.time()
    total_updates = args.rollouts * args.snapshots * args.updates
    done = 0
    for r in range(args.rollouts):
        seed = splitmix(args.seed, r)
        sc = pkg.scenes.WORKLOADS[args.workload](seed=seed)
        sim.init_lattice(sph.make_lattice(**sc.lattice_args()))
        t0 = splitmix(seed, 1) % (args.max_start + 1)
        if t0:
            sim.step(int(t0))
        for k in range(args.snapshots):
            later.copy_state_from(sim)
            later.step(4)
            if norm is None:
                acc_scale = target_scale(pkg, sim, later)
                norm = sph.MlpNorm(NORM["v_scale"], NORM["rho_scale"], NORM["pi_scale"], acc_scale)
                sim.mlp_set(params, norm)
            for u in range(args.updates):
                frac = done / max(total_updates - 1, 1)
                lr = args.lr * (args.lr_final / args.lr) ** frac
                first = (splitmix(seed, 100 + k * args.updates + u) % args.stride)
                res = sim.mlp_train(later, first=first, stride=args.stride, lr=lr)
                done += 1
                log.append({"rollout": r, "t": int(t0) + k * args.gap, "loss": res.loss, "pairs": int(res.pairs),
                            "targets": int(res.weighted), "lr": lr})
            sim.step(args.gap)
        if r % 10 == 0 or r == args.rollouts - 1:
            recent = np.mean([e["loss"] for e in log[-20:]])
            print(json.dumps({"rollout": r, "loss_recent": recent, "elapsed_s": round(time.time() - t_start, 1)}),
                  flush=True)
        if time.time() - t_start > args.time_budget:
            print(json.dumps({"stopped_early_after_rollouts": r + 1}), flush=True)
            break
    params, norm = sim.mlp_get()
    os.makedirs(os.path.dirname(args.out), exist_ok=True)
    meta = {"workload": args.workload, "rollouts": len({e["rollout"] for e in log}), "updates": len(log),
            "stride": args.stride, "lr": [args.lr, args.lr_final], "seed": args.seed,
            "norm": {"v_scale": norm.v_scale, "rho_scale": norm.rho_scale, "pi_scale": norm.pi_scale,
                     "acc_scale": norm.acc_scale},
            "loss_first20": float(np.mean([e["loss"] for e in log[:20]])),
            "loss_last20": float(np.mean([e["loss"] for e in log[-20:]])),
            "seconds": round(time.time() - t_start, 1)}
    np.savez(args.out, params=params, norm=np.array([norm.v_scale, norm.rho_scale, norm.pi_scale, norm.acc_scale],
                                                    dtype=np.float32),
             loss=np.array([e["loss"] for e in log]), meta=json.dumps(meta))
    print(json.dumps(meta), flush=True)
    sim.close()
    later.close()
    return 0


def load_weights(path=None):
    """(params float32[4995], MlpNorm, meta dict) from a training output file."""
    import importlib
    path = path or os.path.join(HERE, "weights", "coarse_mlp.npz")
    z = np.load(path, allow_pickle=False)
    sph = sys.modules.get("npr_sph_amd.sph") or importlib.import_module("npr_sph_amd.sph")
    n = z["norm"]
    return z["params"], sph.MlpNorm(*[float(v) for v in n]), json.loads(str(z["meta"]))


if __name__ == "__main__":
    sys.exit(main())
