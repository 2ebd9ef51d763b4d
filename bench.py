"""Benchmark: D3Q19 TRT (Λ = 3/16) triply periodic Taylor–Green vortex, 256³, FP32 shifted
storage (populations stored as FP32 deviations f − w_q, collision in FP32); the FP32, FP16
shifted and FP64 storages are timed alongside as variants.

One "step" = one stream–collide time step over the whole 256³ lattice (populations resident
in HBM).  The autotuner picks the launch: single-step launches, or two or three steps per
launch with the intermediate populations in LDS.  Prints ONE JSON line with MLUPS, the HBM
roofline of the dominant kernel and a CPU baseline (the FP32 oracle restatement, OpenMP on the
host cores, bounded sample).

Roofline of the dominant kernel: `achieved` = the least HBM traffic one launch can have — one
read and one write of every population, 152 B per cell in FP32 (19 × 4 B × 2), whatever the
steps the launch advances — divided by the launch's average duration, timed with HIP events on
the lattice stream over whole launches only.  `traffic` is the rocprofv3 counter figure for
the same kernel and launch shape (FETCH_SIZE and WRITE_SIZE, calibrated on a copy kernel),
`frac_counter_bytes` the counted bytes over the same time.  The per-step figure (152 B per
cell update × steps per launch) is reported apart as `per_step_GBps` and, over 8 TB/s, as
`per_step_frac`.

The same line carries the learned warm-start record (`warm_start`): 2048 steady 128² MRT
cavities from the batched solver train the decoder, 256 held-out cases run to the residual
1e-7 warm-started in one batch against rest-initialised plain single-lattice runs.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "krakenn.jl_amd"))

HBM_PEAK_GBPS = 8000.0  # MI355X HBM3E, published
BF16_PEAK_TFLOPS = 2500.0  # MI355X dense BF16 MFMA, published
BYTES_PER_POP = {"f64": 8, "f32": 4, "f32s": 4, "f16s": 2}
ARITH_DTYPE = {"f64": "f64", "f32": "f32", "f32s": "f32", "f16s": "f32"}
STORAGE_DTYPE = {"f64": "f64", "f32": "f32", "f32s": "f32 (f - w_q)", "f16s": "f16 (2^15 (f - w_q))"}
LAUNCHES_TIMED = 20
STORAGE_LABEL = {"f64": "FP64 storage", "f32": "FP32 storage",
                 "f32s": "FP32 shifted storage f-w", "f16s": "FP16 shifted storage"}


def launch_span(launch):
    """Steps one launch advances for a lattice launch record (temporal tile 0: 1, 1-5 and
    12-14: 2, 6-11: 3)."""
    t = launch.get("tile", 0)
    return 3 if 6 <= t <= 11 else (2 if t >= 1 else 1)


def traffic_record(kernel, storage, N, launch, tag="trt"):
    """Counted HBM bytes per launch for this kernel, workload tag, storage and launch shape,
    or None."""
    try:
        with open(os.path.join(ROOT, "profiles", "traffic.json")) as fh:
            recs = json.load(fh).get(f"{kernel}_{tag}_{storage}_{N}")
    except (OSError, ValueError):
        return None
    recs = recs if isinstance(recs, list) else ([recs] if recs else [])
    for rec in recs:
        shape = rec.get("launch", {})
        if all(launch.get(k) == shape.get(k) for k in ("tile", "zchunks", "nontemporal")) or \
                (not shape and not launch.get("tile")):
            return rec["bytes_per_launch"]
    return None


def cpu_baseline(N, nu, storage, budget_s=12.0, max_steps=30):
    """Oracle D3Q19 TRT pull step (dsl/bricks_3d.jl composition) on the host, timed."""
    import numpy as np

    sys.path.insert(0, ROOT)
    from oracle import pyoracle as po

    dt = np.float64 if storage == "f64" else np.float32
    import kraken_mi355x as kmi

    rho, ux, uy, uz = kmi.taylor_green_3d_fields(N, 0.05)
    f = po.equilibrium(rho, ux, uy, uz, dtype=dt)
    o = np.empty_like(f, order="F")
    sp, sm = kmi.trt_rates(nu)
    po.call("fused_trt_3d", o, f, N, N, N, 1, 1, 1, sp, sm, dtype=dt)  # warm-up (page-in)
    steps, t0 = 0, time.perf_counter()
    while steps < max_steps:
        po.call("fused_trt_3d", o, f, N, N, N, 1, 1, 1, sp, sm, dtype=dt)
        f, o = o, f
        steps += 1
        if time.perf_counter() - t0 > budget_s:
            break
    el = time.perf_counter() - t0
    return {
        "value": round(N ** 3 * steps / el / 1e6, 2),
        "unit": "MLUPS",
        "cores": po.threads(),
        "kind": "port",
        "sample": f"D3Q19 TRT {'FP64' if storage == 'f64' else 'FP32'} {N}^3 periodic Taylor-Green, "
                  f"{steps} steps ({el:.1f} s) of the C++ oracle fused pull+TRT, OpenMP",
    }


def cpu_baseline_chain(desc, N, storage, budget_s=12.0, max_steps=20):
    """The oracle's descriptor-driven chain (the reference drivers' step order) on the host
    cores from rest, timed on a bounded sample."""
    import numpy as np

    sys.path.insert(0, ROOT)
    from oracle import pyoracle as po

    dt = np.float64 if storage == "f64" else np.float32
    f = po.rest((N, N, N), dtype=dt)
    po.run(desc, f, None, 1)  # warm-up (page-in)
    steps, t0 = 0, time.perf_counter()
    while steps < max_steps and time.perf_counter() - t0 < budget_s:
        po.run(desc, f, None, 1)
        steps += 1
    el = time.perf_counter() - t0
    return {"value": round(N ** 3 * steps / el / 1e6, 2), "unit": "MLUPS", "cores": po.threads(), "kind": "port",
            "sample": f"{steps} steps ({el:.1f} s) of the C++ oracle chain, {'FP64' if dt is np.float64 else 'FP32'}, "
                      f"same configuration, OpenMP"}


def bench_cavities(args):
    """Batched ensemble: 256 D2Q9 MRT lid-driven cavities, N uniform in {64, 96, …, 512},
    Re log-uniform in [100, 3200], u_lid = 0.1, F32S storage, one launch per step over all
    running cavities, each run to r < 1e-7 (mean per-step relative change of u over the
    last 100 steps, checked every 100 steps) or 600k steps."""
    import numpy as np

    import kraken_mi355x as kmi

    n, nu, Re = kmi.ensemble_parameters(256, 2024)
    cells = int((n.astype(np.int64) ** 2).sum())
    with kmi.Batch(n, nu, storage="f32s", collision="mrt") as b:
        if args.batch_launch:
            tw, th, groups, nt, span = (int(v) for v in args.batch_launch.split(","))
            b.set_launch(tw, th, groups, bool(nt))
            b.set_temporal(span)
            b.step(2 * span)  # warm-up, then time a full-load launch sequence
            b.backend.synchronize()
            t0 = time.perf_counter()
            b.step(20 * span)
            b.backend.synchronize()
            full_ms = (time.perf_counter() - t0) * 1e3 / (20 * span)
            b.reset_status()
            for c in range(len(n)):
                b.init_equilibrium(c)
        else:
            full_ms = b.autotune(6)
        t0 = time.perf_counter()
        ms = b.run(1e-7, 100, args.cavity_max_steps, window=100)
        wall = time.perf_counter() - t0
        st = b.status()
        launch = b.launch
    work = float((st["steps"] * n.astype(np.int64) ** 2).sum())
    counts = {kmi.STATUS_NAMES[k]: int((st["status"] == k).sum()) for k in range(4)}
    conv = st["status"] == kmi._lib.CAV_CONVERGED
    out = {
        "metric": "aggregate MLUPS, 256 batched D2Q9 MRT cavities run to residual 1e-7",
        "value": round(work / wall / 1e6, 1), "unit": "MLUPS", "n_gpus": 1,
        "ms_total": round(wall * 1e3, 1), "device_ms": round(ms, 1), "higher_is_better": True,
        "dtype": "f32", "data": "synthetic (seeded N and Re draws, rest initial state)",
        "config": {"workload": "256 lid-driven cavities, N in {64..512 step 32}, Re log-uniform [100, 3200], "
                               "seed 2024, F32S MRT, Zou-He lid", "cells": cells},
        "full_load_MLUPS": round(cells / full_ms / 1e3, 1),
        "full_load_frac_of_8TBps": round(cells * 72 / (full_ms / 1e3) / 8e12, 4),
        "launch": launch, "status": counts,
        "steps_converged": {"median": float(np.median(st["steps"][conv])) if conv.any() else None,
                            "max": int(st["steps"][conv].max()) if conv.any() else None},
    }
    print(json.dumps(out), flush=True)


def bench_small(args):
    """The two small single-lattice configurations, timed as whole runs from their initial
    state: `cavity` = D2Q9 BGK lid-driven cavity 128², Re = 100 (u_lid = 0.1, ν = 0.128),
    Zou–He lid, FP64, rest init, 20000 steps (run_cavity_2d); `rb` = D2Q9 thermal DDF
    Rayleigh–Bénard 512×256, Ra = 1e5, Pr = 0.71, F32 shifted storage (g too), linear T plus
    1e-3 noise, 20000 steps.  Both are launch-bound (2–30 MB of state per step): the steps
    run as captured HIP graphs, with several steps per launch where the autotuner finds it
    faster (D2Q9 without thermal populations)."""
    import numpy as np

    import kraken_mi355x as kmi

    sys.path.insert(0, ROOT)
    if args.workload == "cavity":
        N, nu, ul = 128, 0.128, 0.1
        st = "f64" if args.storage == "f32s" else args.storage  # the configuration is FP64
        desc = kmi.cavity_desc(N, nu, ul, storage=st)
        cpu_desc = kmi.cavity_desc(N, nu, ul, storage="f64" if st == "f64" else "f32")
        cells, bpc, T0 = N * N, 9 * 2 * BYTES_PER_POP[st], None
        name = f"D2Q9 BGK lid-driven cavity {N}^2, Re=100 (u_lid=0.1, nu=0.128), Zou-He lid, rest init"
    else:
        Nx, Ny = 512, 256
        st = args.storage
        desc, _ = kmi.rayleigh_benard_desc(Nx, Ny, 1e5, 0.71, storage=st)
        cpu_desc, _ = kmi.rayleigh_benard_desc(Nx, Ny, 1e5, 0.71, storage="f64" if st == "f64" else "f32")
        T0 = kmi.rb_initial_temperature(Nx, Ny, 1.0, 0.0, seed="noise")
        # g of an F16S lattice is stored F32S
        gb = 4 if st == "f16s" else BYTES_PER_POP[st]
        cells, bpc = Nx * Ny, 9 * 2 * (BYTES_PER_POP[st] + gb)
        name = f"D2Q9 thermal DDF Rayleigh-Benard {Nx}x{Ny}, Ra=1e5, Pr=0.71, linear T + 1e-3 noise"
    steps = args.steps if args.steps != 1000 else 20000
    with kmi.Lattice(desc) as lat:
        if T0 is not None:
            lat.init_temperature(T0)
        lat.autotune(64)  # launch choice (single step or several steps per launch); state kept
        launch = {**lat.launch, **lat.temporal}
        lat.step(args.warmup)  # graph capture and warm-up, then back to the initial state
        lat.backend.synchronize()
    with kmi.Lattice(desc) as lat:
        if T0 is not None:
            lat.init_temperature(T0)
        lat.set_launch(launch["bx"], launch["by"], launch["nontemporal"])
        if launch["tile"]:
            lat.set_temporal(launch["tile"], launch["zchunks"])
        lat.backend.synchronize()
        t0 = time.perf_counter()
        ms = lat.step_timed(steps)
        lat.backend.synchronize()
        wall = time.perf_counter() - t0
        mass = lat.reduce("mass")
        extra = {"Nu": lat.reduce("nusselt", "south")} if T0 is not None else {}
    per_step = ms / 1e3 / steps
    achieved = cells * bpc / per_step / 1e9
    out = {
        "metric": f"MLUPS, {name}, {st}",
        "value": round(cells * steps / wall / 1e6, 1), "unit": "MLUPS", "n_gpus": 1, "steps": steps,
        "warmup": args.warmup, "ms_per_step": round(wall * 1e3 / steps, 5), "wall_s": round(wall, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64" if st == "f64" else "f32",
        "data": "synthetic (rest / analytic initial state)",
        "config": {"workload": name, "cells": cells, "storage": st, "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBPS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBPS, 4), "traffic": None,
                     "kernel": "k_step_tb2d" if launch["tile"] else "k_step",
                     "bytes_per_launch": cells * bpc, "avg_step_us": round(per_step * 1e6, 3),
                     "note": "launch-bound: the state is 2-30 MB, one step is a few microseconds"},
        "mass_after": mass, "launch": launch, **extra,
    }
    if not args.no_cpu_baseline:
        from oracle import pyoracle as po

        dt = np.float64 if st == "f64" else np.float32
        with kmi.Lattice(desc) as lat:
            if T0 is not None:
                lat.init_temperature(T0)
            f = np.asfortranarray(lat.download().astype(dt))
            g = np.asfortranarray(lat.download("g").astype(dt)) if T0 is not None else None
        n, t0 = 0, time.perf_counter()
        while time.perf_counter() - t0 < 10.0 and n < steps:
            po.run(cpu_desc, f, g, 200)
            n += 200
        el = time.perf_counter() - t0
        out["cpu_baseline"] = {"value": round(cells * n / el / 1e6, 2), "unit": "MLUPS", "cores": po.threads(),
                               "kind": "port",
                               "sample": f"{n} steps ({el:.1f} s) of the C++ oracle chain, "
                                         f"{'FP64' if dt is np.float64 else 'FP32'}, same configuration"}
        if args.workload == "cavity":
            # the reference's own published CPU figure for this case (Julia, one Apple M2 core)
            out["reference_published_cpu"] = {"value": "26.5-31.9", "unit": "MLUPS",
                                              "note": "D2Q9 FP64 lid-driven cavity, Kraken.jl CPU backend"}
    print(json.dumps(out), flush=True)


def decoder_forward_flop(d):
    """BF16 product FLOPs of one forward sample of the warm-start decoder: the MLP
    (in → hidden → latent) and three 3×3 convolutions after 2× upsampling."""
    lat = d.latent_hw * d.latent_hw * d.latent_c
    fwd = 2 * (d.in_dim * d.hidden + d.hidden * lat)
    for hi, ci, co in ((d.latent_hw, d.latent_c, d.c1), (2 * d.latent_hw, d.c1, d.c2), (4 * d.latent_hw, d.c2, d.out_c)):
        fwd += 2 * (2 * hi) ** 2 * 9 * ci * co
    return fwd


def bench_nn(args):
    """BF16-MFMA throughput of the warm-start decoder's own kernels: training epochs (forward,
    both gradients, Adam) over 2048 samples of 128×128×3 targets at batch 32, then inference of
    256 samples.  Targets are smooth synthetic fields of the inputs (the kernels' work does not
    depend on their values).  A step is one epoch."""
    import numpy as np

    from kraken_mi355x import warmstart as ws
    from kraken_mi355x.nn import WarmStartNet

    n, B, N = 2048, 32, 128
    Re, A = ws.draw_cases(n, seed=1)
    x = ws.encode(Re, A)
    yy, xx = np.meshgrid(np.linspace(0, 1, N), np.linspace(0, 1, N), indexing="ij")
    y = np.stack([np.stack([np.sin(np.pi * (1 + c) * xx + x[k, 0]) * np.cos(np.pi * yy * (1 + x[k, 1]))
                            for c in range(3)], axis=-1) for k in range(n)]).astype(np.float32)
    net = WarmStartNet(latent_hw=N // 8, max_batch=B, seed=1)
    net.set_dataset(x, y)
    net.train_epochs(max(args.warmup // 32, 1), B, 1e-3, seed=3)
    net.backend.synchronize()
    epochs = max(args.steps // 200, 1)
    t0 = time.perf_counter()
    loss = net.train_epochs(epochs, B, 1e-3, seed=5)
    net.backend.synchronize()
    dt = time.perf_counter() - t0
    fwd = decoder_forward_flop(net.desc)
    flop = 3.0 * fwd * n * epochs
    xt = ws.encode(*ws.draw_cases(256, seed=99))
    net.predict(xt)
    net.backend.synchronize()
    t1 = time.perf_counter()
    net.predict(xt)
    net.backend.synchronize()
    dti = time.perf_counter() - t1
    achieved = flop / dt / 1e12
    out = {
        "metric": "BF16 TFLOP/s of the warm-start decoder training (forward + gradients + Adam), batch 32",
        "value": round(achieved, 2), "unit": "TFLOP/s", "n_gpus": 1, "steps": epochs, "warmup": 1,
        "ms_per_step": round(dt * 1e3 / epochs, 2), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "bf16 products, fp32 accumulate",
        "data": "synthetic (seeded inputs, smooth synthetic 128x128x3 targets)",
        "config": {"workload": f"warm-start decoder 2 -> {net.desc.hidden} -> {N // 8}^2x{net.desc.latent_c} -> 3 conv "
                               f"(2x upsample) -> {N}x{N}x3, {n} samples, batch {B}", "parallelism": "single GPU"},
        "roofline": {"bound": "mfma", "achieved": round(achieved, 2), "peak": BF16_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": round(achieved / BF16_PEAK_TFLOPS, 4), "traffic": None,
                     "fwd_flop_per_sample": fwd, "flop_per_epoch": 3.0 * fwd * n,
                     "note": "latency-bound: 2.6e8 FLOP per sample, batch 32 (8.4 GFLOP per training step)"},
        "inference": {"samples": 256, "ms": round(dti * 1e3, 2),
                      "TFLOPs": round(fwd * 256 / dti / 1e12, 2)},
        "final_loss": loss,
    }
    print(json.dumps(out), flush=True)


def bench_warmstart(args):
    """Learned warm start alone: steady 128² cavities (Re log-uniform [100, 1000], lid
    amplitude A uniform [0.06, 0.12]) from the batched solver train the conv decoder on the
    GPU; the held-out cases then run to r < 1e-7 from rest and from the prediction (feq of the
    predicted ρ, u), batched, against rest-initialised plain single-lattice runs; FP64 runs of
    both starts to a tight residual compare the converged fields (all held-out cases with
    --ws-agree-all)."""
    from kraken_mi355x import warmstart as ws

    log = (lambda m: print(m, file=sys.stderr, flush=True))
    rec = ws.experiment(n_train=args.ws_train, n_test=args.ws_test, epochs=args.ws_epochs,
                        target_tol=args.ws_target_tol, plain_sample=args.ws_plain, agree_sample=args.ws_agree,
                        agree_tol=args.ws_agree_tol, agree_max_steps=args.ws_agree_max_steps,
                        agree_all=args.ws_agree_all, log=log)
    out = {"metric": "time to residual 1e-7, held-out 128^2 MRT cavities: warm-started batched runs vs "
                     "rest-initialised plain runs",
           "value": rec.pop("value"), "unit": rec.pop("unit"), "higher_is_better": True, "n_gpus": 1,
           "dtype": "f32 solver / bf16 network", "data": "synthetic (seeded Re, A draws; training targets "
                                                           "from the batched solver)", **rec}
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--workload", default="tgv",
                    choices=["tgv", "cavity3d", "cavities", "warmstart", "cavity", "rb", "nn"])
    ap.add_argument("--ws-train", type=int, default=2048)
    ap.add_argument("--ws-test", type=int, default=256)
    ap.add_argument("--ws-epochs", type=int, default=1000)
    ap.add_argument("--ws-plain", type=int, default=32, help="plain rest runs timed (sample of the converging set)")
    ap.add_argument("--ws-agree", type=int, default=16, help="cases run in FP64 to a tight residual from both starts")
    ap.add_argument("--ws-agree-tol", type=float, default=2e-13)
    ap.add_argument("--ws-agree-max-steps", type=int, default=8_000_000)
    ap.add_argument("--ws-target-tol", type=float, default=1e-8, help="residual of the training states")
    ap.add_argument("--ws-agree-all", action="store_true", help="FP64 agreement over every held-out case")
    ap.add_argument("--cavity-max-steps", type=int, default=600_000)
    ap.add_argument("--batch-launch", default=None,
                    help="TW,TH,GROUPS,NT,SPAN: fixed batch launch for the cavities workload (no autotune)")
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=64)
    ap.add_argument("--n", type=int, default=256)
    ap.add_argument("--storage", default="f32s", choices=sorted(BYTES_PER_POP))
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-autotune", action="store_true")
    ap.add_argument("--temporal", default=None,
                    help="TILE,ZCHUNKS[,NT]: fixed two-step launch (implies --no-autotune)")
    ap.add_argument("--no-variants", action="store_true", help="skip timing the other storages")
    ap.add_argument("--walls", default="cube", choices=["cube", "z", "y"],
                    help="cavity3d workload: run_cavity_3d's cube, or the same lattice with walls only in z "
                         "(resting bottom, moving top wall) or only in y (other axes periodic)")
    ap.add_argument("--no-warmstart", action="store_true", help="skip the learned warm-start record")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py: this build runs on one GPU only (no multi-GPU path); use --gpus 1",
              file=sys.stderr)
        sys.exit(2)
    if args.workload == "cavities":
        bench_cavities(args)
        return
    if args.workload == "warmstart":
        bench_warmstart(args)
        return
    if args.workload in ("cavity", "rb"):
        bench_small(args)
        return
    if args.workload == "nn":
        bench_nn(args)
        return

    import kraken_mi355x as kmi

    cavity = args.workload == "cavity3d"
    # the cube at Re = U0 N / nu = 256 (tau = 0.65); the periodic TGV at nu = 0.01
    N, nu, U0 = args.n, (0.05 if cavity else 0.01), 0.05
    fields = None if cavity else kmi.taylor_green_3d_fields(N, U0)

    def make(storage):
        # run_cavity_3d's cube (halfway-BB walls, Zou-He lid u = (U0, 0, 0)) or the periodic TGV
        if cavity and args.walls == "cube":
            return kmi.cavity_3d_desc(N, N, N, nu, U0, storage=storage)
        if cavity:  # one walled axis, the others periodic: how the tiles' wall cost splits
            faces = {f: {"kind": "periodic"} for f in ("west", "east", "south", "north", "bottom", "top")}
            if args.walls == "z":
                faces.update(bottom={"kind": "wall"}, top={"kind": "moving_wall", "u": (U0, 0.0, 0.0), "rho": 1.0})
            else:
                faces.update(south={"kind": "wall"}, north={"kind": "wall"})
            return kmi.make_desc("d3q19", N, N, N, nu, storage=storage, collision="bgk", faces=faces)
        return kmi.taylor_green_3d_desc(N, nu, storage)

    def run(storage, steps, warmup):
        desc = make(storage)
        with kmi.Lattice(desc) as lat:
            if fields is not None:
                lat.init_equilibrium(*fields)
            if args.temporal:
                tile, zc, *nt = (int(v) for v in args.temporal.split(","))
                lat.set_launch(256, 1, bool(nt[0]) if nt else False)
                lat.set_temporal(tile, zc)
            elif not args.no_autotune:
                lat.autotune(64)  # restores the populations afterwards
            lat.step(warmup)
            lat.backend.synchronize()
            t0 = time.perf_counter()
            ms = lat.step_timed(steps)
            lat.backend.synchronize()
            wall = time.perf_counter() - t0
            mass = lat.reduce("mass")
            launch = {**lat.launch, **lat.temporal}
            # the dominant kernel alone: whole launches only (a step count that is a multiple
            # of the steps per launch), HIP events on the lattice stream
            span = launch_span(launch)
            launch_ms = lat.step_timed(span * LAUNCHES_TIMED) / LAUNCHES_TIMED
        return ms, wall, mass, launch, launch_ms

    log = (lambda m: print(m, file=sys.stderr, flush=True))
    ms, wall, mass, launch, launch_ms = run(args.storage, args.steps, args.warmup)
    copy_gbps = kmi.MI355XBackend().copy_bandwidth(2 * 1024 ** 3, 20)
    log(f"[bench] {args.storage}: {wall * 1e3 / args.steps:.4f} ms/step, launch {launch}, "
        f"{launch_ms:.4f} ms per launch")
    # HBM bytes per launch from the rocprofv3 FETCH_SIZE / WRITE_SIZE runs (calibrated on a
    # copy kernel of known size), recorded in profiles/traffic.json per kernel, storage and
    # launch shape (tile, z-ranges, non-temporal stores): a record holds for its shape only.
    span = launch_span(launch)
    kernel = {3: "k_step3_d3q19", 2: "k_step2_d3q19", 1: "k_step"}[span]
    tag = "trt" if not cavity else ("bgk_cavity" if args.walls == "cube" else f"bgk_walls_{args.walls}")
    traffic = traffic_record(kernel, args.storage, N, launch, tag=tag)
    cells = N ** 3
    mlups = cells * args.steps / wall / 1e6
    launch_s = launch_ms / 1e3
    min_bytes = cells * 19 * 2 * BYTES_PER_POP[args.storage]  # one read + one write per launch
    achieved = min_bytes / launch_s / 1e9
    layout = {"cube": f"lid-driven cube {N}^3 (run_cavity_3d: halfway-BB walls, Zou-He lid U0={U0})",
              "z": f"{N}^3 walls in z only (resting bottom, top wall moving at U0={U0}; x, y periodic)",
              "y": f"{N}^3 walls in y only (x, z periodic)"}[args.walls]
    name = (f"D3Q19 BGK {layout}, nu={nu}, rest start" if cavity else
            f"D3Q19 TRT Lambda=3/16 triply periodic Taylor-Green {N}^3, U0={U0}, nu={nu}")
    out = {
        "metric": (("MLUPS (D3Q19 BGK 256^3 walled cavity, " if cavity else "MLUPS (D3Q19 TRT 256^3 Taylor-Green, ")
                   + STORAGE_LABEL[args.storage] + ") and % of HBM roofline"),
        "value": round(mlups, 1),
        "unit": "MLUPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(wall * 1e3 / args.steps, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": ARITH_DTYPE[args.storage],
        "storage_dtype": STORAGE_DTYPE[args.storage],
        "data": "synthetic (rest start)" if cavity else "synthetic (analytic Taylor-Green initial field, no dataset)",
        "config": {"workload": f"{name}, storage={args.storage}",
                   "cells": cells, "storage": args.storage, "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBPS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBPS, 4), "traffic": traffic,
                     "kernel": kernel, "steps_per_launch": span, "launch": launch,
                     "bytes_per_launch": min_bytes, "avg_launch_ms": round(launch_ms, 5),
                     "launches_timed": LAUNCHES_TIMED,
                     "frac_counter_bytes": (round(traffic / launch_s / 1e9 / HBM_PEAK_GBPS, 4)
                                            if traffic else None),
                     "per_step_GBps": round(span * achieved, 1),
                     # the same per cell update (152 B in FP32 D3Q19): 0.8 of the 8 TB/s
                     # roofline per update is 42 GLUPS
                     "per_step_frac": round(span * achieved / HBM_PEAK_GBPS, 4),
                     "copy_measured": round(copy_gbps, 1),
                     "frac_of_copy": round(achieved / copy_gbps, 4)},
        "launch": launch,
        "mass_after": mass,
    }
    if not args.no_variants:
        # The other storages of the same workload (FP32 is bit-identical to the reference's
        # FP32 arithmetic; FP16 shifted halves the bytes; FP64 is the reference precision).
        var = {}
        for st in ("f32", "f16s", "f64"):
            if st == args.storage:
                continue
            m2, w2, _, l2, lms2 = run(st, args.steps, args.warmup)
            var[st] = {"MLUPS": round(cells * args.steps / w2 / 1e6, 1),
                       "frac": round(cells * 38 * BYTES_PER_POP[st] / (lms2 / 1e3) / 1e9 / HBM_PEAK_GBPS, 4),
                       "avg_launch_ms": round(lms2, 5), "steps_per_launch": launch_span(l2),
                       "dtype": ARITH_DTYPE[st], "storage_dtype": STORAGE_DTYPE[st], "launch": l2}
            log(f"[bench] {st}: {var[st]['MLUPS']} MLUPS, launch {l2}")
        out["variants"] = var
    if not args.no_warmstart and not cavity:
        from kraken_mi355x import warmstart as ws

        out["warm_start"] = ws.experiment(log=log)
    if not args.no_cpu_baseline and cavity:
        out["cpu_baseline"] = cpu_baseline_chain(make("f64" if args.storage == "f64" else "f32"), N, args.storage)
    elif not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(N, nu, args.storage)
        # the FP64 restatement on the same host cores, for the FP64 variant and reference
        out["cpu_baseline_f64"] = cpu_baseline(N, nu, "f64", budget_s=6.0, max_steps=10)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
