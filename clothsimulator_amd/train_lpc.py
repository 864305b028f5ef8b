"""Trains the learned preconditioner on the device and writes its weights.

    python -m clothsimulator_amd.train_lpc --minutes 8 --out clothsimulator_amd/lpc_weights.npy

Training systems are the implicit-Euler systems the simulator itself assembles:
each round builds one context holding several cloths, far apart so they never
interact, each a hanging sheet, a drape over its own sphere or a crumple onto
its own ground patch, with a grid side drawn from {65, 129, 257, 513}, the
stretch stiffness scaled by U[0.5, 2] and dt drawn from U[2, 10] ms.  The round
steps the scene (multigrid PCG: the systems do not depend on the preconditioner)
and, after every step, takes one Adam step per
cloth on the Kaporin loss of that cloth's assembled system (cs_lpc_train_step:
the network's forward and backward kernels on the device), optionally plus the
solution term (--sol-weight) and the smoothing term (--smooth-weight: the
A-energy one sweep of the learned smoother leaves of a random probe, the figure
that sets its worth inside the multigrid cycle).  Weights carry over from round to round; the
optimiser restarts with each round.  The packaged weights come from
`--minutes 8 --seed 0` (Kaporin loss only).
"""
import argparse
import math
import sys
import time

import numpy as np

from . import CS_MESH_CLOTH, Simulator, scenes
from ._cabi import (CS_E_NOT_CONVERGED, CS_OK, CS_PRECOND_AMG, CS_PRECOND_AMG_LEARNED, CS_PRECOND_BLOCK_JACOBI,
                    CS_PRECOND_LEARNED, default_params, material_struct)
from .lpc_layout import init_weights

SIDES = (65, 129, 257, 513)


def training_scene(rng, slots):
    """One context's worth of independent cloths (each with its own obstacle, if any)."""
    cloths, obstacles = [], []
    for k in range(slots):
        n = int(rng.choice(SIDES))
        kind = rng.choice(["hang", "drape", "crumple"])
        ox = 4.0 * k
        stiff = float(rng.uniform(0.5, 2.0))
        mat = material_struct(stretch=5e3 * stiff)
        if kind == "hang":
            x, f = scenes.grid_mesh(n, 1.0, origin=(ox, 0.0, 0.0))
            cloths.append(scenes.ClothSpec(scenes.MeshArrays(x, f), mat, [(n - 1) * n, (n - 1) * n + n - 1]))
        elif kind == "drape":
            x, f = scenes.grid_mesh(n, 1.0, origin=(ox - 0.5, -0.5, 0.4))
            cloths.append(scenes.ClothSpec(scenes.MeshArrays(x, f), mat))
            sx, sf = scenes.icosphere(3, float(rng.uniform(0.2, 0.4)))
            sx = sx + np.array([ox + rng.uniform(-0.1, 0.1), rng.uniform(-0.1, 0.1), 0.0])
            obstacles.append(scenes.MeshArrays(sx, sf))
        else:
            x, f = scenes.grid_mesh(n, 1.0, origin=(ox - 0.5, -0.5, 0.2))
            x[:, 2] += rng.normal(0.0, 1e-3, len(x))
            cloths.append(scenes.ClothSpec(scenes.MeshArrays(x, f), mat))
            px = np.array([[ox - 1.5, -1.5, 0.0], [ox + 1.5, -1.5, 0.0], [ox + 1.5, 1.5, 0.0], [ox - 1.5, 1.5, 0.0]])
            obstacles.append(scenes.MeshArrays(px, np.array([[0, 1, 2], [0, 2, 3]], dtype=np.int32)))
    p = default_params(dt=float(rng.uniform(0.002, 0.010)))
    p.precond = CS_PRECOND_AMG
    return scenes.SceneSpec("train", p, cloths, obstacles, 0)


HELD_OUT = {"block Jacobi": CS_PRECOND_BLOCK_JACOBI, "learned": CS_PRECOND_LEARNED,
            "multigrid": CS_PRECOND_AMG, "multigrid + learned smoother": CS_PRECOND_AMG_LEARNED}


def held_out_iterations(w, n=129, steps=6):
    """PCG iterations of a held-out crumple system under each preconditioner."""
    out = {}
    for name, pre in HELD_OUT.items():
        spec = scenes.crumple(n, seed=123)
        spec.params.precond = pre
        sim = Simulator(spec)
        if pre in (CS_PRECOND_LEARNED, CS_PRECOND_AMG_LEARNED):
            sim.set_learned_weights(w)
        its = []
        for _ in range(steps):
            s, st = sim.step()
            its.append(st.pcg_iters)
        out[name] = its
        sim.close()
    return out


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--minutes", type=float, default=8.0)
    ap.add_argument("--out", default=None)
    ap.add_argument("--init", default=None, help="start from these weights (.npy) instead of a fresh network")
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--lr", type=float, default=5e-4)
    ap.add_argument("--slots", type=int, default=6)
    ap.add_argument("--steps-per-round", type=int, default=40)
    ap.add_argument("--sol-weight", type=float, default=0.0,
                    help="weight of the solution term (1 - cos^2_A between M^-1 b and the step's solution)")
    ap.add_argument("--smooth-weight", type=float, default=0.0,
                    help="weight of the smoothing term (|(I - G^T G A) v|_A^2 / |v|_A^2, v a random probe)")
    a = ap.parse_args(argv)
    rng = np.random.default_rng(a.seed)
    w = np.load(a.init, allow_pickle=False).astype(np.float32) if a.init else init_weights(a.seed)
    t0 = time.time()
    budget = 60.0 * a.minutes
    total = 0
    skipped = 0
    rnd = 0
    while time.time() - t0 < budget:
        spec = training_scene(rng, a.slots)
        sim = Simulator(spec)
        sim.set_learned_weights(w)
        losses = []
        good = w.copy()  # the weights after the last finite step
        for step in range(a.steps_per_round):
            s, _ = sim.step()
            if s not in (CS_OK, CS_E_NOT_CONVERGED):  # a failed step (an unresolved collision) ends the round
                break
            frac = min(1.0, (time.time() - t0) / budget)
            lr = a.lr * 0.5 * (1.0 + math.cos(math.pi * frac))
            for c in range(len(spec.cloths)):
                loss = sim.lpc_train_step(c, lr, sol_weight=a.sol_weight, smooth_weight=a.smooth_weight,
                                          seed=total)
                total += 1
                cur = sim.learned_weights()
                if not (np.isfinite(loss) and np.isfinite(cur).all()):
                    sim.set_learned_weights(good)  # drop the update of a system whose loss overflowed
                    skipped += 1
                    continue
                good = cur
                losses.append(loss)
            if time.time() - t0 >= budget:
                break
        w = good
        sim.close()
        print("round %d  systems %d (updates dropped %d)  mean loss %.4f  %.0fs" % (
            rnd, total, skipped, float(np.mean(losses)) if losses else float("nan"), time.time() - t0), flush=True)
        rnd += 1
    if a.out:
        np.save(a.out, w.astype(np.float32), allow_pickle=False)
        print("wrote", a.out, flush=True)
    for name, its in held_out_iterations(w).items():
        print("held-out crumple129 PCG iterations, %s: %s (sum %d)" % (name, its, sum(its)), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
