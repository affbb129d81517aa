"""Training of the learned demag correction on (m, exact field) pairs that
the exact solver generates: trajectories of the exact RK12 solver from seeded
vortex, domain-wall and random textures, one training step (GPU forward,
backward and Adam, mmx_surrogate_train_step) per snapshot."""
import time

import numpy as np

from . import synth


def train(sur, sim, nsamples=2000, steps_per_sample=2, samples_per_trajectory=100, lr=5e-4,
          lr_final=1e-4, alpha=0.1, log=None):
    """Trains `sur` (an mmx.Surrogate of `sim`'s grid) in place.  `sim` is used
    to run the exact trajectories (its m, t and solver state are overwritten).
    Learning rates above ~1e-3 let the ReLU layers die (the output collapses
    to zero residual).  Returns the list of (sample, loss)."""
    n = sim.n
    cell = sim.cell
    hist = []
    sim.set_surrogate(None)
    alpha0 = sim.getv("alpha")
    sim.setv("alpha", alpha)
    t0 = time.perf_counter()
    k = 0
    traj = 0
    while k < nsamples:
        sim.setarray("m", synth.training_texture(traj, n, cell))
        sim.setv("t", 0.0)
        sim.setv("dt", 1e-15)
        traj += 1
        for _ in range(samples_per_trajectory):
            if k >= nsamples:
                break
            # cosine decay of the learning rate
            f = 0.5 * (1 + np.cos(np.pi * k / max(1, nsamples - 1)))
            rate = lr_final + (lr - lr_final) * f
            loss = sur.train_step(sim.device_array("m"), rate)
            hist.append((k, loss))
            if log and k % 500 == 0:
                log("sample %d loss %.3e (%.1f s)" % (k, loss, time.perf_counter() - t0))
            k += 1
            sim.steps(steps_per_sample)
    sim.setv("alpha", alpha0)
    return hist
