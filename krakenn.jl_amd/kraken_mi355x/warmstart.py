"""Learned warm start for lid-driven cavities.

Pipeline (all heavy work on the GPU through the C ABI):
  1. steady states of N×N cavities for sampled (Re, lid amplitude A) from the batched
     solver (kmi_batch, F32S MRT, Zou–He lid of run_cavity_2d);
  2. a conv decoder (kmi_nn, BF16 MFMA) trained to map (Re, A) to (u/A, v/A, (ρ−1)/A²);
  3. held-out cavities started from the predicted fields (feq of the predicted ρ, u — the
     `f_init` injection point of src/drivers/gnf_periodic_steady.jl:114-150 — optionally
     followed by consistent-initialisation steps that hold the predicted velocity while
     density and non-equilibrium populations settle) against rest-state starts, both run by
     the same batched solver to the same residual.
Rest initialisation stays the fallback: a cavity without a prediction starts at rest.
"""
from __future__ import annotations

import math
import time

import numpy as np

from .batch import Batch
from .lattice import MI355XBackend
from .nn import WarmStartNet

RE_RANGE = (100.0, 3200.0)
A_RANGE = (0.06, 0.12)


def draw_cases(n, seed, Re_range=RE_RANGE, A_range=A_RANGE):
    """Re log-uniform, lid amplitude A uniform."""
    rng = np.random.default_rng(seed)
    Re = np.exp(rng.uniform(math.log(Re_range[0]), math.log(Re_range[1]), size=n))
    A = rng.uniform(A_range[0], A_range[1], size=n)
    return Re, A


def encode(Re, A, Re_range=RE_RANGE, A_range=A_RANGE):
    """Network input in [−1, 1]²: scaled log Re and scaled amplitude."""
    lr = (np.log(Re) - math.log(Re_range[0])) / (math.log(Re_range[1]) - math.log(Re_range[0]))
    la = (np.asarray(A) - A_range[0]) / (A_range[1] - A_range[0])
    return np.stack([2 * lr - 1, 2 * la - 1], axis=1).astype(np.float32)


def to_target(m, A):
    """Moments (arrays indexed [i, j]) → image [y, x, (u/A, v/A, (ρ−1)/A²)]."""
    return np.stack([m["ux"].T / A, m["uy"].T / A, (m["rho"].T - 1.0) / (A * A)], axis=-1).astype(np.float32)


def from_prediction(y, A):
    """Image [y, x, 3] → (ρ, ux, uy) arrays indexed [i, j]."""
    y = np.asarray(y, dtype=np.float64)
    return 1.0 + A * A * y[..., 2].T, A * y[..., 0].T, A * y[..., 1].T


def generate_steady_states(Re, A, N=128, tol=1e-6, check=100, window=100, max_steps=800_000,
                           storage="f32s", collision="mrt", backend=None, log=None):
    """Run every (Re, A) cavity from rest to r < tol; return targets of the converged ones."""
    nu = A * N / Re
    t0 = time.perf_counter()
    with Batch(np.full(len(Re), N), nu, storage=storage, collision=collision, backend=backend) as b:
        b.autotune(4)
        cap = 0
        while True:
            cap += 100_000
            b.run(tol, check, min(cap, max_steps), window=window)
            st = b.status()
            if log:
                log(f"[data] cap {cap}: converged {(st['status'] == 1).sum()} diverged {(st['status'] == 2).sum()} "
                    f"t {time.perf_counter() - t0:.1f}s")
            if cap >= max_steps or not np.any(st["status"] == 3):
                break
            b.resume()
        ok = np.where(st["status"] == 1)[0]
        y = np.stack([to_target(b.moments(int(c)), A[c]) for c in ok]) if len(ok) else np.zeros((0, N, N, 3), np.float32)
    return {"index": ok, "x": encode(Re[ok], A[ok]), "y": y, "status": st["status"], "steps": st["steps"],
            "seconds": time.perf_counter() - t0}


def train(x, y, *, epochs=300, batch=32, lr=2e-3, lr_final=1e-4, seed=0, latent_hw=None, backend=None, log=None):
    """Train the decoder on (x, y); learning rate decays geometrically over `epochs` in 10 stages."""
    hw = y.shape[1] // 8 if latent_hw is None else latent_hw
    net = WarmStartNet(latent_hw=hw, max_batch=batch, seed=seed + 1, backend=backend)
    net.set_dataset(x, y)
    stages = 10
    t0 = time.perf_counter()
    loss = float("nan")
    for k in range(stages):
        lr_k = lr * (lr_final / lr) ** (k / (stages - 1))
        loss = net.train_epochs(max(1, epochs // stages), min(batch, len(x)), lr_k, seed=seed + 17 * k)
        if log:
            log(f"[train] stage {k} lr {lr_k:.2e} loss {loss:.3e} t {time.perf_counter() - t0:.1f}s")
    return net, loss


def time_to_residual(Re, A, *, N=128, net=None, tol=1e-7, check=100, window=100, max_steps=1_000_000,
                     storage="f32s", collision="mrt", consistent=0, backend=None):
    """All cavities in one batch, from rest (net=None) or from the network's prediction.
    With `consistent` > 0 the predicted velocity is held for that many steps so that density
    and non-equilibrium populations settle to it (Batch.consistent_init) before the run;
    those steps are counted.  Returns wall seconds (including inference and initialisation),
    per-cavity steps/status and the final velocity fields."""
    backend = backend or MI355XBackend()
    nu = A * N / Re
    with Batch(np.full(len(Re), N), nu, storage=storage, collision=collision, backend=backend) as b:
        b.autotune(4)
        backend.synchronize()
        t0 = time.perf_counter()
        if net is not None:
            pred = net.predict(encode(Re, A))
            for c in range(len(Re)):
                rho, ux, uy = from_prediction(pred[c], A[c])
                b.init_equilibrium(c, rho, ux, uy)
            if consistent > 0:
                b.consistent_init(consistent)
        t_init = time.perf_counter() - t0
        ms = b.run(tol, check, max_steps, window=window)
        wall = time.perf_counter() - t0
        st = b.status()
        fields = [b.moments(c) for c in range(len(Re))]
    return {"wall_s": wall, "init_s": t_init, "device_ms": ms, "steps": st["steps"], "status": st["status"],
            "residual": st["residual"], "fields": fields}


def converged_agreement(Re, A, net, *, N=128, tol=1e-11, check=100, window=100, max_steps=3_000_000,
                        storage="f64", consistent=0, backend=None):
    """Whether the warm start changes the answer: the same cavities run from rest and from the
    prediction, in FP64, to a tight residual.  A field stopped at r < tol still sits about
    r·τ from the steady state (τ the slowest mode's decay time in steps), so two starts agree
    only to that distance; at tol = 1e-11 it is far below the fields' 1e-6 agreement bound.
    Returns per-case max |Δu|/A and relative L2 difference of u over the cases converged from
    both starts, with steps and status of both runs."""
    rest = time_to_residual(Re, A, N=N, tol=tol, check=check, window=window, max_steps=max_steps,
                            storage=storage, backend=backend)
    warm = time_to_residual(Re, A, N=N, net=net, tol=tol, check=check, window=window,
                            max_steps=max_steps, storage=storage, consistent=consistent, backend=backend)
    ok = np.where((rest["status"] == 1) & (warm["status"] == 1))[0]
    du = np.array([max(float(np.max(np.abs(warm["fields"][c][k] - rest["fields"][c][k]))) for k in ("ux", "uy"))
                   / float(A[c]) for c in ok])
    l2 = np.array([math.sqrt(sum(float(np.sum((warm["fields"][c][k] - rest["fields"][c][k]) ** 2)) for k in ("ux", "uy"))
                             / sum(float(np.sum(rest["fields"][c][k] ** 2)) for k in ("ux", "uy"))) for c in ok])
    return {"index": ok, "dU_over_A": du, "rel_L2": l2, "rest_steps": rest["steps"], "warm_steps": warm["steps"],
            "rest_status": rest["status"], "warm_status": warm["status"],
            "wall_s": rest["wall_s"] + warm["wall_s"]}


def plain_time_to_residual(Re, A, *, N=128, tol=1e-7, check=100, window=100, max_steps=1_000_000,
                           storage="f32s", backend=None):
    """Rest-initialised plain runs: one lattice at a time (the single-cavity path with its
    autotuned launch), each run to the batch's stop rule r = ‖u^k − u^(k−m)‖/(m‖u^k‖) < tol
    checked every `check` steps (m = `window`), evaluated on the device
    (kmi_step_residual_window).  Returns per-case wall seconds, steps and status (1 converged,
    2 diverged, 3 max_steps)."""
    from .drivers import cavity_desc

    backend = backend or MI355XBackend()
    walls, steps, status = [], [], []
    for Re_c, A_c in zip(Re, A):
        nu = float(A_c) * N / float(Re_c)
        d = cavity_desc(N, nu, float(A_c), storage=storage, collision="mrt")
        from .lattice import Lattice

        t0 = time.perf_counter()
        n, st = 0, 3
        with Lattice(d, backend) as lat:
            lat.autotune(64)  # the fastest single-lattice launch (several steps per launch here)
            while n < max_steps:
                r = lat.step_residual(check, window)  # on the device, FP64
                n += check
                if not math.isfinite(r):
                    st = 2
                    break
                if r < tol:
                    st = 1
                    break
        walls.append(time.perf_counter() - t0)
        steps.append(n)
        status.append(st)
    return {"wall_s": np.array(walls), "steps": np.array(steps), "status": np.array(status)}
