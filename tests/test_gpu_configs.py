"""The benchmark configurations at their real sizes.

Parity at full size: FP64 populations bit-identical to the oracle (cavity 128² for the
full 20000 steps; TGV 256³ and RB 512×256 for as many steps as the CPU oracle finishes in
seconds).  Reduced-precision storage is judged against the FP64 run; every test prints
the max / L2 errors it measured and asserts the bound written in it.
"""
import json
import math
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

OUT = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "gpurun_out")


def _record(name, data):
    os.makedirs(OUT, exist_ok=True)
    path = os.path.join(OUT, "config_errors.json")
    cur = json.load(open(path)) if os.path.exists(path) else {}
    cur[name] = data
    json.dump(cur, open(path, "w"), indent=1)


def _err(a, b):
    d = (a - b).ravel()
    ref = np.abs(b).max()
    return {"max_abs": float(np.abs(d).max()), "l2_rel": float(np.linalg.norm(d) / max(np.linalg.norm(b.ravel()), 1e-300)),
            "max_rel_to_max": float(np.abs(d).max() / ref) if ref > 0 else 0.0}


def test_config1_cavity_fp64_full_run_bitwise(kmi, po, backend):
    """D2Q9 BGK cavity 128², Re = 100 (u_lid = 0.1, ν = 0.128), Zou–He lid, rest init,
    20000 steps, FP64: every population bit-identical to the oracle."""
    N, nu, ul, steps = 128, 0.128, 0.1, 20000
    print(f"[config1] cavity {N}^2 FP64 {steps} steps", flush=True)
    d = kmi.cavity_desc(N, nu, ul)
    with kmi.Lattice(d, backend) as lat:
        lat.step(steps)
        fg = lat.download()
        m = lat.moments()
    fr = po.rest((N, N))
    po.run(d, fr, None, steps)
    assert np.array_equal(fg, fr)
    mr = po.moments(fr)
    assert np.array_equal(m["ux"], mr["ux"]) and np.array_equal(m["rho"], mr["rho"])
    # the same full run with eight (16x16 tile) and four (8x8 tile) steps per launch, the
    # tile and its ring in LDS
    for tile in (4, 6):
        with kmi.Lattice(d, backend) as lat:
            lat.set_temporal(tile, 1)
            lat.step(steps)
            assert np.array_equal(lat.download(), fr), tile
    # the Ladd-lid variant (BASELINE's "moving-lid bounce-back") against its oracle chain
    d2 = kmi.cavity_desc(N, nu, ul, lid="ladd")
    with kmi.Lattice(d2, backend) as lat:
        lat.step(2000)
        fg2 = lat.download()
    fr2 = po.rest((N, N))
    po.run(d2, fr2, None, 2000)
    assert np.array_equal(fg2, fr2)


def test_config2_tgv_256_fp64_bitwise_and_reduced_precision(kmi, po, backend):
    """D3Q19 TRT Λ = 3/16, triply periodic TGV 256³, U0 = 0.05, ν = 0.01."""
    N, nu, U0 = 256, 0.01, 0.05
    fields = kmi.taylor_green_3d_fields(N, U0)
    d64 = kmi.taylor_green_3d_desc(N, nu, "f64")
    # single-step, two-step (tiles 3 and 12) and three-step (tiles 7 and 9) launches at full size
    for tile, zc, nst in ((0, 1, 2), (3, 2, 2), (12, 2, 2), (7, 2, 3), (9, 2, 3)):
        print(f"[config2] TGV 256^3 FP64 full-size parity, temporal tile {tile}, {nst} steps", flush=True)
        f0 = po.equilibrium(*fields)
        with kmi.Lattice(d64, backend) as lat:
            lat.upload(f0)
            if tile:
                lat.set_temporal(tile, zc)
            lat.step(nst)
            fg = lat.download()
        po.run(d64, f0, None, nst)
        assert np.array_equal(fg, f0), tile
        del fg, f0
    # the benchmark's launches (F32S three steps on 64x8 tiles, 832 threads; F16S two steps on
    # 64x8 tiles, 704 threads) against single-step launches of the same storage, 6 steps
    for st, tile in (("f32s", 11), ("f16s", 12)):
        print(f"[config2] TGV 256^3 {st} tile {tile} against single steps", flush=True)
        d = kmi.taylor_green_3d_desc(N, nu, st)
        out = []
        for t in (0, tile):
            with kmi.Lattice(d, backend) as lat:
                lat.init_equilibrium(*fields)
                if t:
                    lat.set_temporal(t, 2)
                lat.step(6)
                out.append(lat.download_raw())  # the stored elements verbatim
        assert np.array_equal(out[0], out[1]), (st, tile)
        del out
    steps = 1000
    res = {}
    for st in ("f64", "f32", "f32s", "f16s"):
        print(f"[config2] TGV 256^3 {st} {steps} steps", flush=True)
        with kmi.Lattice(kmi.taylor_green_3d_desc(N, nu, st), backend) as lat:
            lat.init_equilibrium(*fields)
            m0 = lat.reduce("mass")
            ke0 = lat.reduce("kinetic_energy")
            lat.step(steps)
            res[st] = lat.moments()
            res[st]["mass_drift"] = abs(lat.reduce("mass") - m0) / m0
            res[st]["ke_ratio"] = lat.reduce("kinetic_energy") / ke0
    ref = res["f64"]
    k = 2 * np.pi / N
    decay = math.exp(-2 * 3 * nu * k * k * steps)
    report = {"ke_ratio_f64": ref["ke_ratio"], "ke_linear_theory": decay, "mass_drift_f64": ref["mass_drift"]}
    assert ref["mass_drift"] < 1e-12
    assert abs(ref["ke_ratio"] - decay) < 0.01 * decay
    bounds = {"f32": 1e-4, "f32s": 1e-5, "f16s": 3e-2}
    for st in ("f32", "f32s", "f16s"):
        e = {q: _err(res[st][q], ref[q]) for q in ("ux", "uy", "uz", "rho")}
        e["mass_drift"] = res[st]["mass_drift"]
        report[st] = e
        print(f"[config2] {st}: " + ", ".join(f"{q} max {e[q]['max_abs']:.2e} L2rel {e[q]['l2_rel']:.2e}"
                                               for q in ("ux", "uy", "rho")), flush=True)
        assert e["ux"]["l2_rel"] < bounds[st] and e["uy"]["l2_rel"] < bounds[st]
        assert e["rho"]["max_abs"] < bounds[st]
    _record("config2_tgv256", report)


def test_config3_rayleigh_benard_512x256(kmi, po, backend):
    """D2Q9 thermal DDF Rayleigh–Bénard 512×256, Ra = 1e5, Pr = 0.71, hot south / cold
    north halfway anti-bounce-back, periodic x, T = linear + 1e-3 uniform noise."""
    Nx, Ny, Ra, Pr = 512, 256, 1e5, 0.71
    T0 = kmi.rb_initial_temperature(Nx, Ny, 1.0, 0.0, seed="noise")
    d64, p = kmi.rayleigh_benard_desc(Nx, Ny, Ra, Pr, storage="f64")
    print("[config3] RB 512x256 FP64 200-step full-size parity", flush=True)
    with kmi.Lattice(d64, backend) as lat:
        lat.init_temperature(T0)
        f0, g0 = lat.download(), lat.download("g")
        lat.step(200)
        fg, gg = lat.download(), lat.download("g")
    po.run(d64, f0, g0, 200)
    assert np.array_equal(fg, f0) and np.array_equal(gg, g0)
    steps = 150_000
    res = {}
    for st in ("f64", "f32", "f32s", "f16s"):
        print(f"[config3] RB {st} {steps} steps", flush=True)
        d, _ = kmi.rayleigh_benard_desc(Nx, Ny, Ra, Pr, storage=st)
        with kmi.Lattice(d, backend) as lat:
            lat.init_temperature(T0)
            lat.step(steps)
            m = lat.moments(temperature=True)
            m["Nu"] = lat.reduce("nusselt", "south")
            res[st] = m
    ref = res["f64"]
    report = {"steps": steps, "Nu_f64": ref["Nu"]}
    print(f"[config3] Nu f64 = {ref['Nu']:.6f}", flush=True)
    assert 1.0 < ref["Nu"] < 8.0
    for st in ("f32", "f32s", "f16s"):
        e = {q: _err(res[st][q], ref[q]) for q in ("ux", "uy", "rho", "Temp")}
        e["Nu"] = res[st]["Nu"]
        e["Nu_rel_err"] = abs(res[st]["Nu"] - ref["Nu"]) / ref["Nu"]
        report[st] = e
        print(f"[config3] {st}: Nu {e['Nu']:.6f} (rel err {e['Nu_rel_err']:.2e}), T max {e['Temp']['max_abs']:.2e}, "
              f"ux L2rel {e['ux']['l2_rel']:.2e}", flush=True)
        assert np.all(np.isfinite(res[st]["Temp"]))
    _record("config3_rb512x256", report)
