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

# Re log-uniform on [100, 1000] and lid amplitude A uniform on [0.06, 0.12] at N = 128 keep
# tau = 3 A N / Re + 1/2 >= 0.523.  Below tau ~ 0.522 the reference's MRT cavity (collide_mrt_2d!
# with the Zou-He lid) has no stable solution: its FP64 restatement blows up within ~1000
# steps from rest whatever the storage (tests/test_oracle_mrt.py), so no steady state exists
# there to train on or to converge to.
RE_RANGE = (100.0, 1000.0)
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
                     storage="f32s", collision="mrt", consistent=0, backend=None, log=None, chunk=500_000):
    """All cavities in one batch, from rest (net=None) or from the network's prediction.
    With `consistent` > 0 the predicted velocity is held for that many steps so that density
    and non-equilibrium populations settle to it (Batch.consistent_init) before the run;
    those steps are counted.  The batch advances in spans of `chunk` steps (a cavity stopped
    at a span's end resumes in the next; `log` gets one line per span).  Returns wall seconds
    (including inference and initialisation), per-cavity steps/status and the final velocity
    fields."""
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
        ms, cap = 0.0, 0
        while True:
            cap = min(cap + chunk, max_steps)
            ms += b.run(tol, check, cap, window=window)
            st = b.status()
            if log:
                log(f"[batch] {'warm' if net is not None else 'rest'} {storage} r<{tol:g} {cap} steps: "
                    f"converged {(st['status'] == 1).sum()}/{len(Re)} diverged {(st['status'] == 2).sum()} "
                    f"t {time.perf_counter() - t0:.1f}s")
            if cap >= max_steps or not np.any(st["status"] == 3):
                break
            b.resume()
        wall = time.perf_counter() - t0
        fields = [b.moments(c) for c in range(len(Re))]
    return {"wall_s": wall, "init_s": t_init, "device_ms": ms, "steps": st["steps"], "status": st["status"],
            "residual": st["residual"], "fields": fields}


def converged_agreement(Re, A, net, *, N=128, tol=1e-11, check=100, window=100, max_steps=3_000_000,
                        storage="f64", consistent=0, backend=None, log=None):
    """Whether the warm start changes the answer: the same cavities run from rest and from the
    prediction, in FP64, to a tight residual.  A field stopped at r < tol still sits about
    r·τ from the steady state (τ the slowest mode's decay time in steps), so two starts agree
    only to that distance; at tol = 1e-11 it is far below the fields' 1e-6 agreement bound.
    Returns per-case max |Δu|/A and relative L2 difference of u over the cases converged from
    both starts, with steps and status of both runs."""
    rest = time_to_residual(Re, A, N=N, tol=tol, check=check, window=window, max_steps=max_steps,
                            storage=storage, backend=backend, log=log)
    warm = time_to_residual(Re, A, N=N, net=net, tol=tol, check=check, window=window,
                            max_steps=max_steps, storage=storage, consistent=consistent, backend=backend, log=log)
    ok = np.where((rest["status"] == 1) & (warm["status"] == 1))[0]
    du = np.array([max(float(np.max(np.abs(warm["fields"][c][k] - rest["fields"][c][k]))) for k in ("ux", "uy"))
                   / float(A[c]) for c in ok])
    l2 = np.array([math.sqrt(sum(float(np.sum((warm["fields"][c][k] - rest["fields"][c][k]) ** 2)) for k in ("ux", "uy"))
                             / sum(float(np.sum(rest["fields"][c][k] ** 2)) for k in ("ux", "uy"))) for c in ok])
    return {"index": ok, "dU_over_A": du, "rel_L2": l2, "rest_steps": rest["steps"], "warm_steps": warm["steps"],
            "rest_status": rest["status"], "warm_status": warm["status"],
            "wall_s": rest["wall_s"] + warm["wall_s"], "rest_wall_s": rest["wall_s"], "warm_wall_s": warm["wall_s"]}


def plain_time_to_residual(Re, A, *, N=128, tol=1e-7, check=100, window=100, max_steps=1_000_000,
                           storage="f32s", backend=None):
    """Rest-initialised plain runs: one lattice at a time (the single-cavity path with its
    autotuned launch), each run to the batch's stop rule r = ‖u^k − u^(k−m)‖/(m‖u^k‖) < tol
    checked every `check` steps (m = `window`), evaluated on the device
    (kmi_step_residual_window).  The clock runs from the rest state to the stop, as for the
    batched runs (lattice creation and launch tuning excluded from both).  Returns per-case
    wall seconds, steps and status (1 converged, 2 diverged, 3 max_steps)."""
    from .drivers import cavity_desc

    backend = backend or MI355XBackend()
    from .lattice import Lattice

    walls, steps, status = [], [], []
    for Re_c, A_c in zip(Re, A):
        nu = float(A_c) * N / float(Re_c)
        d = cavity_desc(N, nu, float(A_c), storage=storage, collision="mrt")
        n, st = 0, 3
        with Lattice(d, backend) as lat:
            # set-up outside the clock, as for the batched runs: the lattice, its fastest
            # launch (several steps per launch here; autotune restores the rest state)
            lat.autotune(64)
            backend.synchronize()
            t0 = time.perf_counter()
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


def experiment(*, n_train=2048, n_test=256, N=128, epochs=1000, target_tol=1e-8, tol=1e-7,
               plain_sample=32, agree_sample=16, agree_tol=2e-13, agree_max_steps=8_000_000,
               agree_all=False, backend=None, log=None):
    """The learned warm start end to end: steady states of `n_train` cavities from the batched
    solver (F32S, windowed residual < `target_tol`) train the decoder; `n_test` held-out cases
    then run to r < `tol` as one batch from rest and one from the prediction; a sample of
    `plain_sample` of them runs as rest-initialised plain single-lattice runs (the per-case
    mean scales to all held-out cases); `agree_sample` cases (all with `agree_all`) run in
    FP64 from both starts to r < `agree_tol`, where the two converged fields are compared.
    Every clock starts from the initial state (set-up and launch tuning excluded alike).
    Returns a flat record of the measurements."""
    backend = backend or MI355XBackend()
    say = log or (lambda s: None)
    t_all = time.perf_counter()
    Re_tr, A_tr = draw_cases(n_train, seed=1)
    data = generate_steady_states(Re_tr, A_tr, N=N, tol=target_tol, backend=backend, log=say)
    t0 = time.perf_counter()
    net, loss = train(data["x"], data["y"], epochs=epochs, batch=32, lr=2e-3, backend=backend, log=say)
    backend.synchronize()
    t_train = time.perf_counter() - t0
    Re_te, A_te = draw_cases(n_test, seed=99)
    rest = time_to_residual(Re_te, A_te, N=N, tol=tol, backend=backend)
    warm = time_to_residual(Re_te, A_te, N=N, net=net, tol=tol, backend=backend)
    say(f"[warmstart] batched rest {rest['wall_s']:.2f}s warm {warm['wall_s']:.2f}s")
    conv = np.where((rest["status"] == 1) & (warm["status"] == 1))[0]
    pick = conv[np.linspace(0, len(conv) - 1, min(plain_sample, len(conv))).astype(int)] if len(conv) else conv
    plain = plain_time_to_residual(Re_te[pick], A_te[pick], N=N, tol=tol, backend=backend)
    plain_total = float(np.mean(plain["wall_s"])) * n_test if len(pick) else float("nan")
    say(f"[warmstart] plain sample {len(pick)}: mean {np.mean(plain['wall_s']) if len(pick) else float('nan'):.3f} s")
    pred = net.predict(encode(Re_te[conv], A_te[conv])) if len(conv) else []
    perr = [float(np.linalg.norm(pred[k][..., :2] - to_target(rest["fields"][c], A_te[c])[..., :2])
                  / np.linalg.norm(to_target(rest["fields"][c], A_te[c])[..., :2])) for k, c in enumerate(conv)]
    du7 = [max(float(np.max(np.abs(warm["fields"][c][k] - rest["fields"][c][k]))) for k in ("ux", "uy")) / A_te[c]
           for c in conv]
    pa = np.arange(n_test) if agree_all else (
        conv[np.linspace(0, len(conv) - 1, min(agree_sample, len(conv))).astype(int)] if len(conv) else conv)
    agree = converged_agreement(Re_te[pa], A_te[pa], net, N=N, tol=agree_tol, max_steps=agree_max_steps,
                                backend=backend, log=say)
    du = agree["dU_over_A"]
    say(f"[warmstart] FP64 r<{agree_tol:g} on {len(pa)}: converged both {len(agree['index'])}, "
        f"max dU/A {du.max() if len(du) else float('nan'):.2e}")
    return {
        "value": round(plain_total / warm["wall_s"], 3) if len(pick) else None,
        "unit": "x (rest-initialised plain runs' wall / warm-started batch wall), residual %g" % tol,
        "batched_rest_over_warm": round(rest["wall_s"] / warm["wall_s"], 3),
        "median_step_ratio_rest_over_warm": (round(float(np.median(rest["steps"][conv] / warm["steps"][conv])), 3)
                                             if len(conv) else None),
        "walls_s": {"plain_rest_estimate": round(plain_total, 2), "batched_rest": round(rest["wall_s"], 3),
                    "batched_warm": round(warm["wall_s"], 3),
                    "warm_inference_and_init": round(warm["init_s"], 3)},
        "status_rest": np.bincount(rest["status"], minlength=4).tolist(),
        "status_warm": np.bincount(warm["status"], minlength=4).tolist(),
        "plain_sample": {"cases": int(len(pick)), "mean_wall_s": round(float(np.mean(plain["wall_s"])), 4) if len(pick) else None,
                         "status": np.bincount(plain["status"], minlength=4).tolist() if len(pick) else None,
                         "median_steps": int(np.median(plain["steps"])) if len(pick) else None},
        "prediction_rel_L2_median": round(float(np.median(perr)), 5) if perr else None,
        "max_dU_over_A_at_stop": float(max(du7)) if du7 else None,
        "fp64_agreement": {
            "residual": agree_tol, "cases": int(len(pa)), "converged_both": int(len(agree["index"])),
            "max_dU_over_A": float(du.max()) if len(du) else None,
            "median_dU_over_A": float(np.median(du)) if len(du) else None,
            "below_1e-6": int(np.sum(du <= 1e-6)),
            "max_rel_L2": float(agree["rel_L2"].max()) if len(du) else None,
            "batched_rest_wall_s": round(agree["rest_wall_s"], 2),
            "batched_warm_wall_s": round(agree["warm_wall_s"], 2),
            "median_steps_rest": int(np.median(agree["rest_steps"])),
            "median_steps_warm": int(np.median(agree["warm_steps"]))},
        "train": {"generated": int(n_train), "steady_states": int(len(data["index"])),
                  "status": np.bincount(data["status"], minlength=4).tolist(), "target_residual": target_tol,
                  "median_steps": int(np.median(data["steps"])), "data_gen_s": round(data["seconds"], 1),
                  "epochs": epochs, "final_loss": loss, "train_s": round(t_train, 2)},
        "config": {"N": N, "Re": list(RE_RANGE), "A": list(A_RANGE), "train": n_train, "test": n_test,
                   "solver": "batched D2Q9 MRT, F32S storage, Zou-He lid, windowed residual (m = 100)"},
        "total_s": round(time.perf_counter() - t_all, 1),
    }
