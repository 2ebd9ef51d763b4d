"""GPU parity: the HIP single-pass kernel against the CPU oracle on the same inputs.

FP64 and FP32 storage must be bit-identical to the oracle's restatement of the reference
chain in the same precision.  Shifted storage (F32S, F16S) is checked against the FP64
oracle with the tolerances written in each test.
"""
import math
import zlib

import numpy as np
import pytest

from test_oracle_lattice import rng_pops

pytestmark = pytest.mark.gpu


def _mask_blob(shape, centre, r):
    idx = np.indices(shape)
    d2 = sum((idx[a] - centre[a]) ** 2 for a in range(len(shape)))
    return np.asfortranarray((d2 <= r * r).astype(np.uint8))


def _case(kmi, name, storage):
    """(desc, shape, mask, thermal, steps) of a parity case."""
    per = lambda *fs: {f: {"kind": "periodic"} for f in fs}
    if name == "d2q9_cavity_zh_bgk_mask":
        return kmi.cavity_desc(24, 0.05, 0.1, storage=storage, Ny=20), (24, 20), _mask_blob((24, 20), (10, 8), 2), False, 45
    if name == "d2q9_cavity_ladd_trt":
        return kmi.cavity_desc(20, 0.03, 0.1, storage=storage, Ny=18, collision="trt", lid="ladd"), (20, 18), None, False, 40
    if name == "d2q9_cavity_mrt":
        return kmi.cavity_desc(22, 0.01, 0.1, storage=storage, Ny=16, collision="mrt"), (22, 16), None, False, 40
    if name == "d2q9_periodic_bgk":
        return kmi.make_desc("d2q9", 32, 24, 1, 0.02, storage=storage, faces=per("west", "east", "south", "north")), (32, 24), None, False, 40
    if name == "d2q9_periodic_trt":
        return kmi.make_desc("d2q9", 16, 16, 1, 0.02, storage=storage, collision="trt",
                             faces=per("west", "east", "south", "north")), (16, 16), None, False, 33
    if name == "d2q9_couette_zh":
        f = per("west", "east")
        f.update(south={"kind": "zouhe_velocity", "u": (0.05, 0, 0)}, north={"kind": "zouhe_velocity", "u": (0, 0, 0)})
        return kmi.make_desc("d2q9", 8, 16, 1, 0.1, storage=storage, faces=f, bc_order=["south", "north"]), (8, 16), None, False, 40
    if name == "d2q9_channel_inlet_outlet":
        f = {"west": {"kind": "zouhe_velocity", "u": (0.04, 0, 0)}, "east": {"kind": "zouhe_pressure", "rho": 1.0},
             "south": {"kind": "wall"}, "north": {"kind": "wall"}}
        return kmi.make_desc("d2q9", 40, 16, 1, 0.04, storage=storage, faces=f), (40, 16), _mask_blob((40, 16), (12, 8), 3), False, 50
    if name == "d2q9_rb_split":
        d, _ = kmi.rayleigh_benard_desc(32, 16, Ra=1e4, Pr=0.71, storage=storage)
        return d, (32, 16), None, True, 40
    if name == "d2q9_natconv_fused":
        return kmi.natural_convection_desc(24, 1e4, 0.71, storage=storage), (24, 24), None, True, 40
    if name == "d2q9_thermal_corners_mask":
        f = {"west": {"kind": "wall", "thermal": "dirichlet", "T": 0.9}, "south": {"kind": "wall", "thermal": "dirichlet", "T": 0.3},
             "east": {"kind": "wall"}, "north": {"kind": "moving_wall", "u": (0.03, 0, 0)}}
        d = kmi.make_desc("d2q9", 16, 12, 1, 0.05, storage=storage, faces=f, thermal=True, omega_T=1.5,
                          beta_g=1e-4, T_ref=0.5)
        return d, (16, 12), _mask_blob((16, 12), (8, 6), 1), True, 35
    if name == "d3q19_periodic_trt":
        return kmi.make_desc("d3q19", 16, 12, 10, 0.01, storage=storage, collision="trt",
                             faces=per("west", "east", "south", "north", "bottom", "top")), (16, 12, 10), None, False, 40
    if name == "d3q19_cavity_bgk_ladd_mask":
        d = kmi.make_desc("d3q19", 12, 10, 8, 0.05, storage=storage,
                          faces={"top": {"kind": "moving_wall", "u": (0.05, 0.02, 0)}})
        return d, (12, 10, 8), _mask_blob((12, 10, 8), (6, 5, 4), 2), False, 35
    if name == "d3q19_couette_trt_ladd":
        f = per("west", "east", "bottom", "top")
        f.update(south={"kind": "moving_wall", "u": (0.03, 0, 0)}, north={"kind": "moving_wall", "u": (-0.02, 0, 0)})
        return kmi.make_desc("d3q19", 8, 12, 6, 0.05, storage=storage, collision="trt", faces=f), (8, 12, 6), None, False, 40
    if name == "d3q19_pz_trt_walls":
        return kmi.make_desc("d3q19", 10, 8, 6, 0.02, storage=storage, collision="trt",
                             faces=per("bottom", "top")), (10, 8, 6), None, False, 33
    if name == "d3q19_cavity_zh_lid_mask":
        return kmi.cavity_3d_desc(12, 10, 8, 0.05, 0.08, storage=storage), (12, 10, 8), _mask_blob((12, 10, 8), (6, 5, 3), 2), False, 35
    if name == "d3q19_channel_zh_trt":
        z = {"kind": "zouhe_velocity", "u": (0.0, 0.0, 0.0)}
        f = {"west": {"kind": "zouhe_velocity", "u": (0.03, 0.01, -0.005)}, "east": {"kind": "zouhe_pressure", "rho": 1.0},
             "south": dict(z), "north": {"kind": "zouhe_velocity", "u": (0.02, 0.0, 0.01)}, "bottom": dict(z), "top": dict(z)}
        d = kmi.make_desc("d3q19", 14, 8, 6, 0.04, storage=storage, collision="trt", faces=f,
                          bc_order=["west", "east", "top", "bottom", "south", "north"])
        return d, (14, 8, 6), None, False, 40
    if name == "d3q19_zh_pressure_top_px_py":
        f = per("west", "east", "south", "north")
        f.update(bottom={"kind": "zouhe_velocity", "u": (0.01, -0.02, 0.015)}, top={"kind": "zouhe_pressure", "rho": 0.995})
        return kmi.make_desc("d3q19", 8, 6, 10, 0.03, storage=storage, faces=f), (8, 6, 10), None, False, 33
    raise KeyError(name)


CASES = ["d2q9_cavity_zh_bgk_mask", "d2q9_cavity_ladd_trt", "d2q9_cavity_mrt", "d2q9_periodic_bgk",
         "d2q9_periodic_trt", "d2q9_couette_zh", "d2q9_channel_inlet_outlet", "d2q9_rb_split",
         "d2q9_natconv_fused", "d2q9_thermal_corners_mask", "d3q19_periodic_trt",
         "d3q19_cavity_bgk_ladd_mask", "d3q19_couette_trt_ladd", "d3q19_pz_trt_walls",
         "d3q19_cavity_zh_lid_mask", "d3q19_channel_zh_trt", "d3q19_zh_pressure_top_px_py"]


def _init(shape, thermal, seed):
    Q = 9 if len(shape) == 2 else 19
    f = rng_pops(shape, Q, seed=seed, amp=0.02)
    g = None
    if thermal:
        r = np.random.default_rng(seed + 100)
        T = 0.5 + 0.3 * r.uniform(-1, 1, size=shape)
        g = np.asfortranarray(np.array([4 / 9] + [1 / 9] * 4 + [1 / 36] * 4)[None, None, :] * T[:, :, None])
        g = g * (1 + 0.01 * r.uniform(-1, 1, size=g.shape))
        g = np.asfortranarray(g)
    return f, g


def _run_gpu(kmi, backend, desc, f, g, mask, steps, dtype):
    with kmi.Lattice(desc, backend) as lat:
        if mask is not None:
            lat.set_mask(mask)
        lat.upload(f.astype(dtype, order="F"))
        if g is not None:
            lat.upload(g.astype(dtype, order="F"), field="g")
        lat.step(steps)
        fo = lat.download(dtype=dtype)
        go = lat.download("g", dtype=dtype) if g is not None else None
    return fo, go


@pytest.mark.parametrize("storage,dtype", [("f64", np.float64), ("f32", np.float32)])
@pytest.mark.parametrize("case", CASES)
def test_bitwise_parity(kmi, po, backend, case, storage, dtype):
    desc, shape, mask, thermal, steps = _case(kmi, case, storage)
    print(f"[parity] {case} {storage} shape={shape} steps={steps}", flush=True)
    f, g = _init(shape, thermal, seed=zlib.crc32(case.encode()) % 1000)
    fg, gg = _run_gpu(kmi, backend, desc, f, g, mask, steps, dtype)
    fr = f.astype(dtype, order="F")
    gr = None if g is None else g.astype(dtype, order="F")
    po.run(desc, fr, gr, steps, mask=mask)
    assert np.all(np.isfinite(fr))
    assert np.array_equal(fg, fr), f"max |Δf| = {np.max(np.abs(fg.astype(float) - fr)):.3e}"
    if thermal:
        assert np.array_equal(gg, gr), f"max |Δg| = {np.max(np.abs(gg.astype(float) - gr)):.3e}"


# F32S: FP32 arithmetic on deviations; F16S: additionally binary16 storage of 2^15·(f − w).
SHIFTED_TOL = {"f32s": 2e-6, "f16s": 3e-4}


@pytest.mark.parametrize("storage", ["f32s", "f16s"])
@pytest.mark.parametrize("case", CASES)
def test_shifted_storage_against_fp64(kmi, po, backend, case, storage):
    desc, shape, mask, thermal, steps = _case(kmi, case, storage)
    print(f"[shifted] {case} {storage}", flush=True)
    f, g = _init(shape, thermal, seed=7 + zlib.crc32(case.encode()) % 1000)
    fg, gg = _run_gpu(kmi, backend, desc, f, g, mask, steps, np.float64)
    d64, *_ = _case(kmi, case, "f64")
    fr, gr = f.copy(order="F"), (None if g is None else g.copy(order="F"))
    po.run(d64, fr, gr, steps, mask=mask)
    tol = SHIFTED_TOL[storage]
    err = np.max(np.abs(fg - fr))
    assert err < tol, f"{storage}: max |Δf| = {err:.3e} (tol {tol})"
    if thermal:
        errg = np.max(np.abs(gg - gr))
        assert errg < 4 * tol, f"{storage}: max |Δg| = {errg:.3e}"


@pytest.mark.parametrize("storage,dtype", [("f64", np.float64), ("f32", np.float32)])
def test_moments_match_macroscopic(kmi, po, backend, storage, dtype):
    for case in ("d2q9_natconv_fused", "d3q19_couette_trt_ladd"):
        desc, shape, mask, thermal, steps = _case(kmi, case, storage)
        f, g = _init(shape, thermal, seed=3)
        with kmi.Lattice(desc, backend) as lat:
            lat.upload(f.astype(dtype, order="F"))
            if thermal:
                lat.upload(g.astype(dtype, order="F"), "g")
            lat.step(5)
            m = lat.moments(temperature=thermal)
            fd = lat.download(dtype=dtype)
            gd = lat.download("g", dtype=dtype) if thermal else None
        mr = po.moments(fd)
        for k in mr:
            assert np.array_equal(m[k], mr[k].astype(np.float64)), k
        if thermal:
            assert np.array_equal(m["Temp"], po.temperature(gd).astype(np.float64))


def test_reductions(kmi, po, backend):
    desc, shape, mask, thermal, steps = _case(kmi, "d2q9_natconv_fused", "f64")
    f, g = _init(shape, thermal, seed=4)
    with kmi.Lattice(desc, backend) as lat:
        lat.upload(f)
        lat.upload(g, "g")
        lat.step(30)
        m = lat.moments(temperature=True)
        mass = lat.reduce("mass")
        mom = lat.reduce("momentum")
        ke = lat.reduce("kinetic_energy")
        th = lat.reduce("thermal")
        vmax = lat.reduce("max_speed")
        nu = lat.reduce("nusselt", "west")
        mass2 = lat.reduce("mass")
    rho, ux, uy = m["rho"], m["ux"], m["uy"]
    assert mass == mass2  # deterministic
    assert abs(mass - math.fsum(rho.ravel())) <= 1e-14 * mass
    assert abs(mom[0] - math.fsum((rho * ux).ravel())) <= 1e-15 * mass
    assert abs(mom[1] - math.fsum((rho * uy).ravel())) <= 1e-15 * mass
    assert abs(ke - math.fsum((0.5 * rho * (ux ** 2 + uy ** 2)).ravel())) <= 1e-13 * abs(ke)
    assert abs(th - math.fsum(m["Temp"].ravel())) <= 1e-14 * abs(th)
    assert vmax == np.max(np.sqrt(ux ** 2 + uy ** 2))
    nref = po.nusselt(m["Temp"], 0, 1.0, 0.0, False)
    assert abs(nu - nref) <= 1e-12 * abs(nref)


def test_residual_matches_definition(kmi, po, backend):
    """r_u = ‖u^k − u^(k−1)‖₂ / ‖u^k‖₂ over fluid cells (gnf_periodic_steady.jl:186-196)."""
    desc, shape, mask, thermal, steps = _case(kmi, "d2q9_cavity_zh_bgk_mask", "f64")
    with kmi.Lattice(desc, backend) as lat:
        lat.set_mask(mask)
        lat.step(40)
        f0 = lat.download()
        r = lat.step_residual(3)
    fr = f0.copy(order="F")
    po.run(desc, fr, None, 2, mask=mask)
    u0 = po.moments(fr)
    po.run(desc, fr, None, 1, mask=mask)
    u1 = po.moments(fr)
    fl = mask == 0
    du2 = ((u1["ux"] - u0["ux"]) ** 2 + (u1["uy"] - u0["uy"]) ** 2)[fl].sum()
    u2 = (u1["ux"] ** 2 + u1["uy"] ** 2)[fl].sum()
    assert abs(r - math.sqrt(du2 / u2)) <= 1e-10 * r


def test_graph_replay_equals_eager_steps(kmi, backend):
    """64 steps through the captured graph (2 replays) equal 64 single launches."""
    desc, shape, mask, thermal, steps = _case(kmi, "d2q9_rb_split", "f32")
    f, g = _init(shape, thermal, seed=5)
    out = []
    for chunks in ([64], [1] * 64, [32, 7, 25]):
        with kmi.Lattice(desc, backend) as lat:
            lat.upload(f)
            lat.upload(g, "g")
            for n in chunks:
                lat.step(n)
            assert lat.steps == 64
            out.append((lat.download(), lat.download("g")))
    for fo, go in out[1:]:
        assert np.array_equal(fo, out[0][0]) and np.array_equal(go, out[0][1])


def test_autotune_restores_state_and_launch_shape_is_invisible(kmi, backend):
    desc, shape, mask, thermal, steps = _case(kmi, "d2q9_rb_split", "f32")
    f, g = _init(shape, thermal, seed=6)
    outs = []
    for mode in ("default", "tuned", (64, 1, 0), (16, 4, 1), (32, 8, 0), (128, 2, 1)):
        with kmi.Lattice(desc, backend) as lat:
            lat.upload(f)
            lat.upload(g, "g")
            lat.step(10)
            before = (lat.download(), lat.download("g"))
            if mode == "tuned":
                lat.autotune(32)
                assert lat.steps == 10
                assert np.array_equal(lat.download(), before[0]) and np.array_equal(lat.download("g"), before[1])
            elif mode != "default":
                lat.set_launch(*mode)
                assert lat.launch == {"bx": mode[0], "by": mode[1], "nontemporal": bool(mode[2])}
            lat.step(50)
            outs.append((lat.download(), lat.download("g")))
    for fo, go in outs[1:]:
        assert np.array_equal(fo, outs[0][0]) and np.array_equal(go, outs[0][1])
    with kmi.Lattice(desc, backend) as lat:
        with pytest.raises(kmi.KrakenError):
            lat.set_launch(48, 1)


def test_invalid_descriptor_raises(kmi, backend):
    d = kmi.make_desc("d3q19", 8, 8, 8, collision="mrt")
    with pytest.raises(kmi.KrakenError) as e:
        kmi.Lattice(d, backend)
    assert e.value.status == 5 and "MRT" in str(e.value)


@pytest.mark.parametrize("storage", ["f64", "f32", "f32s", "f16s"])
@pytest.mark.parametrize("collision", ["trt", "bgk"])
def test_two_step_launches_bit_identical(kmi, backend, storage, collision):
    """Temporal blocking (two steps per launch, t+1 populations in LDS) reproduces single-step
    launches bit for bit, for every tile and thread count and z split, odd step counts
    included."""
    shape = (64, 16, 12)
    per = {f: {"kind": "periodic"} for f in ("west", "east", "south", "north", "bottom", "top")}
    desc = kmi.make_desc("d3q19", *shape, 0.02, storage=storage, collision=collision, faces=per)
    f = rng_pops(shape, 19, seed=5, amp=0.05)
    steps = 37
    print(f"[two-step] {storage} {collision} shape={shape} steps={steps}", flush=True)
    with kmi.Lattice(desc, backend) as lat:
        lat.upload(f)
        lat.step(steps)
        ref = lat.download()
    modes = [(1, 1), (1, 3), (2, 2), (2, 12), (3, 5), (4, 1)] + ([(5, 2)] if storage != "f64" else [])
    modes += [(12, 1), (12, 4), (13, 2), (14, 1), (14, 12)]  # one thread per extended cell
    for tile, zc in modes:
        with kmi.Lattice(desc, backend) as lat:
            lat.upload(f)
            lat.set_temporal(tile, zc)
            assert lat.temporal == {"tile": tile, "zchunks": zc}
            lat.step(steps)
            got = lat.download()
        assert np.array_equal(got, ref), (tile, zc, float(np.max(np.abs(got - ref))))


@pytest.mark.parametrize("storage", ["f64", "f32", "f32s", "f16s"])
@pytest.mark.parametrize("collision", ["trt", "bgk"])
def test_three_step_launches_bit_identical(kmi, backend, storage, collision):
    """Three steps per launch (t+1 and t+2 populations in two LDS levels) reproduce single-step
    launches bit for bit, for every tile and thread count (one thread per interior cell, tiles
    6-8, or per level-1 cell, tiles 9-11) and z split down to one plane per range, with step
    counts that leave one- and two-step remainders, and on a four-plane lattice."""
    per = {f: {"kind": "periodic"} for f in ("west", "east", "south", "north", "bottom", "top")}
    for shape in ((64, 16, 12), (32, 16, 4)):
        desc = kmi.make_desc("d3q19", *shape, 0.02, storage=storage, collision=collision, faces=per)
        f = rng_pops(shape, 19, seed=6, amp=0.05)
        for steps in (37, 65):
            print(f"[three-step] {storage} {collision} shape={shape} steps={steps}", flush=True)
            with kmi.Lattice(desc, backend) as lat:
                lat.upload(f)
                lat.step(steps)
                ref = lat.download()
            modes = [(7, 1), (7, min(5, shape[2] - 1)), (7, shape[2]), (9, 1), (9, 3)]
            if storage != "f64":
                modes += [(8, 2), (8, 3), (10, 1), (10, 4)]
                modes += [(6, 1), (6, 4), (11, 2), (11, shape[2])] if shape[0] % 64 == 0 else []
            for tile, zc in modes:
                with kmi.Lattice(desc, backend) as lat:
                    lat.upload(f)
                    lat.set_temporal(tile, zc)
                    assert lat.temporal == {"tile": tile, "zchunks": zc}
                    lat.step(steps)
                    assert lat.steps == steps
                    got = lat.download()
                assert np.array_equal(got, ref), (shape, steps, tile, zc, float(np.max(np.abs(got - ref))))
    with kmi.Lattice(kmi.make_desc("d3q19", 32, 8, 3, 0.02, storage=storage, collision=collision,
                                   faces=per), backend) as lat:
        with pytest.raises(kmi.KrakenError):
            lat.set_temporal(7, 1)  # fewer than four planes
    if storage == "f64":
        with kmi.Lattice(kmi.make_desc("d3q19", 64, 16, 8, 0.02, storage="f64", faces=per), backend) as lat:
            for tile in (6, 8, 10, 11):
                with pytest.raises(kmi.KrakenError):
                    lat.set_temporal(tile, 1)  # FP64 LDS budget


def _walled_3d_cases(kmi, storage, collision, shape):
    """D3Q19 lattices with walls in the walled multi-step kernels' scope: (name, desc, mask)."""
    nx, ny, nz = shape
    walls = {f: {"kind": "wall"} for f in ("west", "east", "south", "north", "bottom")}
    cav = dict(walls, top={"kind": "zouhe_velocity", "u": (0.08, 0.0, 0.0)})  # run_cavity_3d
    yield "cavity_zh_lid", kmi.make_desc("d3q19", *shape, 0.02, storage=storage, collision=collision, faces=cav), None
    ladd = dict(walls, top={"kind": "moving_wall", "u": (0.05, 0.02, 0.0), "rho": 1.0})
    x, y, z = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")
    ball = ((x - nx / 3) ** 2 + (y - ny / 2) ** 2 + (z - nz / 2) ** 2 < (min(ny, nz) / 3) ** 2).astype(np.uint8)
    yield "ladd_lid_ball", kmi.make_desc("d3q19", *shape, 0.03, storage=storage, collision=collision, faces=ladd), ball
    mixed = {"west": {"kind": "periodic"}, "east": {"kind": "periodic"}, "south": {"kind": "wall"},
             "north": {"kind": "moving_wall", "u": (0.04, 0.0, 0.01), "rho": 1.0},
             "bottom": {"kind": "moving_wall", "u": (-0.02, 0.03, 0.0), "rho": 1.0}, "top": {"kind": "wall"}}
    yield "periodic_x_ladd", kmi.make_desc("d3q19", *shape, 0.02, storage=storage, collision=collision,
                                           faces=mixed), None
    chan = {"west": {"kind": "zouhe_velocity", "u": (0.03, 0.0, 0.0)}, "east": {"kind": "zouhe_pressure", "rho": 1.0},
            "south": {"kind": "wall"}, "north": {"kind": "wall"}, "bottom": {"kind": "periodic"},
            "top": {"kind": "periodic"}}
    yield "zh_channel_inert_ball", kmi.make_desc("d3q19", *shape, 0.04, storage=storage, collision=collision,
                                                 faces=chan, solid_mode="inert"), ball
    yield "zh_channel", kmi.make_desc("d3q19", *shape, 0.04, storage=storage, collision=collision, faces=chan), None
    # resting x walls, periodic y, a moving z wall: the ring tiles take the lean variant
    # (no mask, x–y faces periodic or resting) with the z wall's Ladd term
    box = {"west": {"kind": "wall"}, "east": {"kind": "wall"}, "south": {"kind": "periodic"},
           "north": {"kind": "periodic"}, "bottom": {"kind": "wall"},
           "top": {"kind": "moving_wall", "u": (0.04, 0.02, 0.0), "rho": 1.0}}
    yield "x_walls_ladd_z", kmi.make_desc("d3q19", *shape, 0.03, storage=storage, collision=collision,
                                          faces=box), None
    # walls only in z (a Couette gap): every tile takes the interior variant
    couette = {"west": {"kind": "periodic"}, "east": {"kind": "periodic"}, "south": {"kind": "periodic"},
               "north": {"kind": "periodic"}, "bottom": {"kind": "wall"},
               "top": {"kind": "moving_wall", "u": (0.05, 0.01, 0.0), "rho": 1.0}}
    yield "couette_z", kmi.make_desc("d3q19", *shape, 0.03, storage=storage, collision=collision, faces=couette), None


@pytest.mark.parametrize("storage", ["f64", "f32", "f32s"])
@pytest.mark.parametrize("collision", ["trt", "bgk"])
def test_walled_multi_step_launches_bit_identical(kmi, backend, storage, collision):
    """Two- and three-step launches on walled D3Q19 lattices (halfway and moving-wall
    bounce-back, Zou–He velocity and pressure faces, bounce-back and inert solid masks, mixed
    with periodic axes) reproduce single-step launches bit for bit: the bounced links of the
    LDS levels take the cell's own opposite population of the previous level.  On the
    192 × 32 lattice the tiles clear of the x and y walls run the interior variant (z walls
    only) and the rest the full one, or its lean form when the lattice has no mask and its x
    and y faces are periodic or resting walls (run_cavity_3d's cube, the x-walled box);
    lattices with a mask run the full one throughout."""
    for shape in ((64, 16, 12), (32, 8, 5), (192, 32, 10)):
        for name, desc, mask in _walled_3d_cases(kmi, storage, collision, shape):
            f = rng_pops(shape, 19, seed=8, amp=0.05)
            for steps in (37, 65):
                print(f"[walled] {name} {storage} {collision} shape={shape} steps={steps}", flush=True)
                with kmi.Lattice(desc, backend) as lat:
                    if mask is not None:
                        lat.set_mask(mask)
                    lat.upload(f)
                    lat.step(steps)
                    ref = lat.download()
                modes = [(14, 1), (14, 3)] + ([(12, 2)] if shape[0] % 64 == 0 else [])
                if storage != "f64":
                    modes += [(9, 1), (9, 2), (9, shape[2])]
                for tile, zc in modes:
                    with kmi.Lattice(desc, backend) as lat:
                        if mask is not None:
                            lat.set_mask(mask)
                        lat.upload(f)
                        lat.set_temporal(tile, zc)
                        lat.step(steps)
                        got = lat.download()
                    assert np.array_equal(got, ref), (name, shape, steps, tile, zc,
                                                      float(np.nanmax(np.abs(got - ref))))
    # out of the walled scope: the other tiles, F16S, forcing
    walls = kmi.make_desc("d3q19", 64, 16, 8, 0.02, storage=storage, collision=collision)
    with kmi.Lattice(walls, backend) as lat:
        for tile in (1, 3, 7, 11):
            with pytest.raises(kmi.KrakenError):
                lat.set_temporal(tile, 1)
    with kmi.Lattice(kmi.make_desc("d3q19", 64, 16, 8, 0.02, storage="f16s", collision=collision), backend) as lat:
        with pytest.raises(kmi.KrakenError):
            lat.set_temporal(12, 1)
    forced = kmi.make_desc("d3q19", 64, 16, 8, 0.02, storage=storage, forcing="guo", force=(1e-5, 0.0, 0.0))
    with kmi.Lattice(forced, backend) as lat:
        with pytest.raises(kmi.KrakenError):
            lat.set_temporal(14, 1)


def test_walled_cube_256_bench_launches_bit_identical(kmi, backend):
    """The walled benchmark lattice itself: run_cavity_3d's cube at 256³ (F32S, BGK, Zou–He
    lid, rest start) on the lean ring variant with ring tiles handed out first (64×8 with 6
    z-ranges, 768 groups in three rounds; 32×8 with 3 z-ranges) and on the three-step 32×8
    tile equals single-step launches bit for bit after 30 steps."""
    N = 256
    desc = kmi.cavity_3d_desc(N, N, N, 0.05, 0.05, storage="f32s")
    with kmi.Lattice(desc, backend) as lat:
        lat.step(30)
        ref = lat.download_raw()
    for tile, zc in ((12, 6), (14, 3), (9, 2)):
        print(f"[walled] cube 256^3 f32s tile {tile} z-ranges {zc}", flush=True)
        with kmi.Lattice(desc, backend) as lat:
            lat.set_temporal(tile, zc)
            lat.step(30)
            got = lat.download_raw()
        assert np.array_equal(got, ref), (tile, zc)


def test_walled_cavity_autotune_picks_a_multi_step_launch(kmi, backend):
    """run_cavity_3d's lattice at 128³ (FP32 shifted): the autotuner now offers the walled
    multi-step launches, and its choice reproduces single-step launches bit for bit."""
    N = 128
    desc = kmi.cavity_3d_desc(N, N, N, 0.05, 0.1, storage="f32s")
    with kmi.Lattice(desc, backend) as lat:
        lat.step(30)
        ref = lat.download_raw()
    with kmi.Lattice(desc, backend) as lat:
        lat.autotune(32)
        print(f"[walled] cavity 128^3 tuned launch {lat.launch} temporal {lat.temporal}", flush=True)
        lat.step(30)
        assert np.array_equal(lat.download_raw(), ref)


def test_two_step_tuned_and_rejected_where_it_does_not_apply(kmi, backend):
    N = 64
    desc = kmi.taylor_green_3d_desc(N, 0.01, "f32")
    fields = kmi.taylor_green_3d_fields(N, 0.05)
    with kmi.Lattice(desc, backend) as lat:
        lat.init_equilibrium(*fields)
        lat.step(3)
        before = lat.download()
        lat.autotune(32)
        assert lat.steps == 3 and np.array_equal(lat.download(), before)
        lat.step(40)
        tuned = lat.download()
        print(f"[two-step] tuned launch {lat.launch} temporal {lat.temporal}", flush=True)
    with kmi.Lattice(desc, backend) as lat:
        lat.init_equilibrium(*fields)
        lat.step(43)
        assert np.array_equal(lat.download(), tuned)
    walls = kmi.make_desc("d3q19", 64, 16, 8, 0.02, collision="trt")
    with kmi.Lattice(walls, backend) as lat:
        with pytest.raises(kmi.KrakenError) as e:
            lat.set_temporal(1, 1)
        assert e.value.status == 5


def _tb2d_cases(kmi, storage):
    """D2Q9 lattices in the multi-step kernel's scope: (name, desc, shape, mask)."""
    N, M = 70, 45
    yield "cavity_zh_bgk", kmi.cavity_desc(N, 0.05, 0.1, storage=storage, Ny=M), (N, M), None
    yield "cavity_ladd_trt_mask", kmi.cavity_desc(N, 0.04, 0.1, storage=storage, collision="trt", lid="ladd",
                                                  Ny=M), (N, M), _mask_blob((N, M), (20, 18), 4)
    yield "cavity_mrt", kmi.cavity_desc(N, 0.02, 0.1, storage=storage, collision="mrt", Ny=M), (N, M), None
    per = {"west": {"kind": "periodic"}, "east": {"kind": "periodic"}}
    yield "channel_guo", kmi.make_desc("d2q9", N, M, 1, 0.05, storage=storage, faces=per, forcing="guo",
                                       force=(1e-5, 0.0, 0.0)), (N, M), _mask_blob((N, M), (25, 21), 5)
    zh = {"west": {"kind": "zouhe_velocity", "u": (0.04, 0.0, 0.0)}, "east": {"kind": "zouhe_pressure", "rho": 1.0}}
    yield "channel_zh_inert", kmi.make_desc("d2q9", N, M, 1, 0.05, storage=storage, faces=zh,
                                            solid_mode="inert"), (N, M), _mask_blob((N, M), (18, 20), 4)
    full = {f: {"kind": "periodic"} for f in ("west", "east", "south", "north")}
    yield "periodic_trt", kmi.make_desc("d2q9", N, M, 1, 0.02, storage=storage, collision="trt", faces=full), \
        (N, M), None


@pytest.mark.parametrize("storage", ["f64", "f32", "f32s", "f16s"])
def test_d2q9_multi_step_launches_bit_identical(kmi, backend, storage):
    """K steps per launch on D2Q9 lattices (tile plus a K-cell ring in LDS) reproduce single
    steps bit for bit for every variant: walls, Ladd and Zou–He lids, Zou–He inlet/outlet,
    periodic axes, masks (bounce-back and inert), Guo forcing, BGK / TRT / MRT, and lattice
    sizes that are not multiples of the tile; step counts leave remainders."""
    steps = 37
    for name, desc, shape, mask in _tb2d_cases(kmi, storage):
        f = rng_pops(shape, 9, seed=8, amp=0.02)
        print(f"[tb2d] {storage} {name} shape={shape} steps={steps}", flush=True)
        with kmi.Lattice(desc, backend) as lat:
            if mask is not None:
                lat.set_mask(mask)
            lat.upload(f)
            lat.step(steps)
            ref = lat.download()
        for tile in (1, 2, 3, 4, 5, 6, 7, 8):
            with kmi.Lattice(desc, backend) as lat:
                if mask is not None:
                    lat.set_mask(mask)
                lat.upload(f)
                lat.set_temporal(tile, 1)
                lat.step(steps)
                assert lat.steps == steps
                got = lat.download()
            assert np.array_equal(got, ref), (name, tile, float(np.max(np.abs(got - ref))))


def test_d2q9_multi_step_rejected_out_of_scope(kmi, backend):
    # thermal populations, per-node faces and lattices smaller than tile + ring stay single-step
    d, _ = kmi.rayleigh_benard_desc(64, 32, 1e4, 0.71)
    with kmi.Lattice(d, backend) as lat:
        with pytest.raises(kmi.KrakenError):
            lat.set_temporal(1, 1)
    with kmi.Lattice(kmi.cavity_desc(40, 0.05, 0.1), backend) as lat:
        lat.set_temporal(1, 1)  # 40 >= 28 + 4
        with pytest.raises(kmi.KrakenError):
            lat.set_temporal(3, 1)  # 40 < 56 + 8
    prof = {"west": {"kind": "zouhe_velocity_profile"}}
    with kmi.Lattice(kmi.make_desc("d2q9", 40, 40, 1, 0.05, faces=prof), backend) as lat:
        with pytest.raises(kmi.KrakenError):
            lat.set_temporal(1, 1)


def test_d2q9_multi_step_autotuned_cavity_matches(kmi, backend):
    """The autotuner may pick a multi-step variant for a 128² cavity; the run is unchanged."""
    d = kmi.cavity_desc(128, 0.128, 0.1)
    with kmi.Lattice(d, backend) as lat:
        lat.step(10)
        lat.autotune(64)
        print(f"[tb2d] cavity 128^2 tuned temporal {lat.temporal}", flush=True)
        lat.step(290)
        a = lat.download()
    with kmi.Lattice(d, backend) as lat:
        lat.step(300)
        assert np.array_equal(lat.download(), a)


@pytest.mark.parametrize("storage", ["f32", "f32s", "f16s"])
def test_d2q9_thermal_multi_step_launches_bit_identical(kmi, backend, storage):
    """The thermal double distribution in the multi-step kernel: Rayleigh–Bénard (split
    coupling, periodic x, Dirichlet south/north) and natural convection in a box (fused
    coupling, Dirichlet west/east with the corner rule, adiabatic south/north), f and g
    bit-identical to single steps for every variant."""
    N, M, steps = 70, 45, 29
    rb, _ = kmi.rayleigh_benard_desc(N, M, 1e4, 0.71, storage=storage)
    nc = kmi.natural_convection_desc(64, 1e5, 0.71, storage=storage)
    x = (np.arange(64)[:, None] + 0.5) / 64 + 0.0 * np.arange(64)[None, :]
    cases = [("rb", rb, kmi.rb_initial_temperature(N, M, 1.0, 0.0, seed="noise")),
             ("natconv", nc, np.asfortranarray(1.0 - x))]
    for name, desc, T0 in cases:
        print(f"[tb2d thermal] {storage} {name}", flush=True)
        outs = []
        for tile in (0, 1, 2, 3, 4, 5, 6, 7, 8):
            with kmi.Lattice(desc, backend) as lat:
                if T0 is not None:
                    lat.init_temperature(T0)
                if tile:
                    lat.set_temporal(tile, 1)
                lat.step(steps)
                outs.append((lat.download(), lat.download("g")))
        for tile, (f, g) in zip((1, 2, 3, 4, 5, 6, 7, 8), outs[1:]):
            assert np.array_equal(f, outs[0][0]) and np.array_equal(g, outs[0][1]), (name, tile)
    with kmi.Lattice(kmi.rayleigh_benard_desc(N, M, 1e4, 0.71, storage="f64")[0], backend) as lat:
        with pytest.raises(kmi.KrakenError):
            lat.set_temporal(1, 1)  # FP64 thermal lattices keep single steps
