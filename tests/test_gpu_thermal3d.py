"""GPU parity of the D3Q19 thermal double distribution (run_natural_convection_3d chain,
thermal_3d.jl) and of the Guo body force in 3D and per node (collide_guo_3d.jl,
collide_guo_2d.jl:93-160), against the oracle on the same inputs.

FP64 and FP32 storage are bit-identical to the oracle; F32S / F16S are checked against the
FP64 oracle with the tolerances of test_gpu_parity.py.  The Nusselt reduction is compared
with the oracle's estimator on the downloaded temperature (FP64, relative 1e-12).
"""
import math
import zlib

import numpy as np
import pytest

from test_gpu_parity import SHIFTED_TOL, _mask_blob, _run_gpu
from test_oracle_lattice import rng_pops

pytestmark = pytest.mark.gpu

W19 = np.array([1 / 3] + [1 / 18] * 6 + [1 / 36] * 12)


def _thermal_case(kmi, name, storage):
    """(desc, shape, mask, steps) of a D3Q19 thermal case."""
    if name == "natconv_box":
        d, _ = kmi.natural_convection_3d_desc(10, 2e3, 0.71, storage=storage)  # β_g = 7e-3
        return d, (10, 10, 10), None, 30
    if name == "all_dirichlet_edges_mask":
        T = dict(west=0.9, east=0.1, south=0.7, north=0.3, bottom=0.6, top=0.2)
        f = {k: {"kind": "wall", "thermal": "dirichlet", "T": v} for k, v in T.items()}
        f["top"]["kind"] = "moving_wall"
        f["top"]["u"] = (0.03, 0.01, 0.0)
        d = kmi.make_desc("d3q19", 10, 8, 9, 0.05, storage=storage, faces=f, thermal=True,
                          omega_T=1.4, beta_g=2e-3, T_ref=0.5)
        return d, (10, 8, 9), _mask_blob((10, 8, 9), (5, 4, 4), 1.5), 30
    if name == "rayleigh_benard_pxz":
        f = {"west": {"kind": "periodic"}, "east": {"kind": "periodic"},
             "bottom": {"kind": "periodic"}, "top": {"kind": "periodic"},
             "south": {"kind": "wall", "thermal": "dirichlet", "T": 1.0},
             "north": {"kind": "wall", "thermal": "dirichlet", "T": 0.0}}
        d = kmi.make_desc("d3q19", 12, 8, 10, 0.05, storage=storage, faces=f, thermal=True,
                          omega_T=1.0 / (3 * 0.05 / 0.71 + 0.5), beta_g=5e-3, T_ref=0.5)
        return d, (12, 8, 10), None, 30
    raise KeyError(name)


THERMAL = ["natconv_box", "all_dirichlet_edges_mask", "rayleigh_benard_pxz"]


def _thermal_init(shape, seed):
    f = rng_pops(shape, 19, seed=seed, amp=0.02)
    r = np.random.default_rng(seed + 100)
    T = 0.5 + 0.3 * r.uniform(-1, 1, size=shape)
    g = W19[None, None, None, :] * T[..., None] * (1 + 0.01 * r.uniform(-1, 1, size=shape + (19,)))
    return f, np.asfortranarray(g)


@pytest.mark.parametrize("storage,dtype", [("f64", np.float64), ("f32", np.float32)])
@pytest.mark.parametrize("case", THERMAL)
def test_thermal3d_bitwise(kmi, po, backend, case, storage, dtype):
    desc, shape, mask, steps = _thermal_case(kmi, case, storage)
    print(f"[thermal3d] {case} {storage} shape={shape} steps={steps}", flush=True)
    f, g = _thermal_init(shape, zlib.crc32(case.encode()) % 1000)
    fg, gg = _run_gpu(kmi, backend, desc, f, g, mask, steps, dtype)
    fr, gr = f.astype(dtype, order="F"), g.astype(dtype, order="F")
    po.run(desc, fr, gr, steps, mask=mask)
    assert np.all(np.isfinite(fr)) and not np.array_equal(fr, f.astype(dtype))
    assert np.array_equal(fg, fr), f"max |Δf| = {np.max(np.abs(fg.astype(float) - fr)):.3e}"
    assert np.array_equal(gg, gr), f"max |Δg| = {np.max(np.abs(gg.astype(float) - gr)):.3e}"


@pytest.mark.parametrize("storage", ["f32s", "f16s"])
@pytest.mark.parametrize("case", THERMAL)
def test_thermal3d_shifted_against_fp64(kmi, po, backend, case, storage):
    desc, shape, mask, steps = _thermal_case(kmi, case, storage)
    print(f"[thermal3d shifted] {case} {storage}", flush=True)
    f, g = _thermal_init(shape, 7 + zlib.crc32(case.encode()) % 1000)
    fg, gg = _run_gpu(kmi, backend, desc, f, g, mask, steps, np.float64)
    d64, *_ = _thermal_case(kmi, case, "f64")
    fr, gr = f.copy(order="F"), g.copy(order="F")
    po.run(d64, fr, gr, steps, mask=mask)
    tol = SHIFTED_TOL[storage]
    err, errg = np.max(np.abs(fg - fr)), np.max(np.abs(gg - gr))
    assert err < tol, f"{storage}: max |Δf| = {err:.3e} (tol {tol})"
    # g is stored F32S on F16S lattices, so its tolerance is the F32S one in both cases
    assert errg < 4 * SHIFTED_TOL["f32s"], f"{storage}: max |Δg| = {errg:.3e}"


@pytest.mark.parametrize("storage,dtype", [("f64", np.float64), ("f32", np.float32)])
def test_thermal3d_moments_and_nusselt(kmi, po, backend, storage, dtype):
    desc, shape, mask, steps = _thermal_case(kmi, "natconv_box", storage)
    f, g = _thermal_init(shape, 11)
    with kmi.Lattice(desc, backend) as lat:
        lat.upload(f.astype(dtype, order="F"))
        lat.upload(g.astype(dtype, order="F"), "g")
        lat.step(7)
        m = lat.moments(temperature=True)
        Nu = lat.reduce("nusselt", "west")
        Nu_e = lat.reduce("nusselt", "east")
        fd, gd = lat.download(dtype=dtype), lat.download("g", dtype=dtype)
    mr = po.moments(fd)
    for k in mr:
        assert np.array_equal(m[k], mr[k].astype(np.float64)), k
    T = po.temperature(gd).astype(np.float64)
    assert np.array_equal(m["Temp"], T)
    ref = po.nusselt_3d(T, kmi._lib.WEST, 1.0, 0.0)
    assert abs(Nu - ref) <= 1e-12 * max(1.0, abs(ref)), (Nu, ref)
    ref_e = po.nusselt_3d(T, kmi._lib.EAST, 0.0, 1.0)
    assert abs(Nu_e - ref_e) <= 1e-12 * max(1.0, abs(ref_e)), (Nu_e, ref_e)


def test_box_pure_conduction_driver(kmi, po, backend):
    """run_natural_convection_3d(N = 12, Ra = 0, 6000 steps): test_thermal.jl:332-336 bounds, and
    the same Nu as the oracle's run."""
    N = 12
    r = kmi.run_natural_convection_3d(N=N, Ra=0.0, Pr=0.71, max_steps=6000, backend=backend)
    d3 = max(abs(r["Temp"][i, 5, 5] - (1 - (i + 0.5) / N)) for i in range(N))
    assert abs(r["Nu"] - 1) < 3e-3, r["Nu"]
    assert d3 < 4e-4, d3
    d, _ = kmi.natural_convection_3d_desc(N, 0.0, 0.71)
    f = po.rest((N, N, N))
    g = np.asfortranarray(W19[None, None, None, :] * kmi.natural_convection_3d_initial_temperature(N)[..., None])
    po.run(d, f, g, 6000)
    T = po.temperature(g)
    assert np.array_equal(r["Temp"], T)
    assert abs(r["Nu"] - po.nusselt_3d(T, kmi._lib.WEST, 1.0, 0.0)) < 1e-12


# ---- Guo body force: constant in 3D, per-node field in 2D and 3D ------------------------------

def _force_case(kmi, name, storage):
    """(desc, shape, mask, field or None, steps)."""
    r = np.random.default_rng(zlib.crc32(name.encode()) % 1000)
    if name == "d3q19_poiseuille_guo_mask":
        f = {"west": {"kind": "periodic"}, "east": {"kind": "periodic"},
             "bottom": {"kind": "periodic"}, "top": {"kind": "periodic"}}
        d = kmi.make_desc("d3q19", 10, 12, 8, 0.05, storage=storage, faces=f, forcing="guo",
                          force=(1e-4, 2e-5, -3e-5))
        return d, (10, 12, 8), _mask_blob((10, 12, 8), (5, 6, 4), 1.5), None, 30
    if name == "d3q19_guo_field_walls":
        shape = (9, 8, 7)
        d = kmi.make_desc("d3q19", *shape, 0.04, storage=storage, forcing="guo_field",
                          faces={"north": {"kind": "moving_wall", "u": (0.02, 0, 0)}})
        fld = tuple(np.asfortranarray(r.uniform(-2e-4, 2e-4, shape)) for _ in range(3))
        return d, shape, _mask_blob(shape, (4, 4, 3), 1), fld, 30
    if name == "d2q9_guo_field_cavity":
        shape = (20, 14)
        d = kmi.make_desc("d2q9", 20, 14, 1, 0.05, storage=storage, forcing="guo_field",
                          faces={"north": {"kind": "zouhe_velocity", "u": (0.05, 0, 0)}})
        fld = tuple(np.asfortranarray(r.uniform(-2e-4, 2e-4, shape)) for _ in range(2))
        return d, shape, _mask_blob(shape, (10, 6), 2), fld, 30
    raise KeyError(name)


FORCE = ["d3q19_poiseuille_guo_mask", "d3q19_guo_field_walls", "d2q9_guo_field_cavity"]


def _run_force_gpu(kmi, backend, desc, f, mask, fld, steps, dtype, moments=False):
    with kmi.Lattice(desc, backend) as lat:
        if mask is not None:
            lat.set_mask(mask)
        if fld is not None:
            lat.set_force_field(*fld)
        lat.upload(f.astype(dtype, order="F"))
        lat.step(steps)
        out = lat.download(dtype=dtype)
        m = lat.moments() if moments else None
    return out, m


def _run_force_oracle(po, desc, f, mask, fld, steps):
    if fld is None:
        return po.run(desc, f, None, steps, mask=mask)[0]
    z = fld[2] if len(fld) == 3 else None
    return po.run_field(desc, f, fld[0], fld[1], z, nsteps=steps, mask=mask)


@pytest.mark.parametrize("storage,dtype", [("f64", np.float64), ("f32", np.float32)])
@pytest.mark.parametrize("case", FORCE)
def test_guo_force_bitwise(kmi, po, backend, case, storage, dtype):
    desc, shape, mask, fld, steps = _force_case(kmi, case, storage)
    print(f"[guo] {case} {storage} shape={shape}", flush=True)
    Q = 19 if len(shape) == 3 else 9
    f = rng_pops(shape, Q, seed=5, amp=0.02)
    fg, m = _run_force_gpu(kmi, backend, desc, f, mask, fld, steps, dtype, moments=True)
    fr = _run_force_oracle(po, desc, f.astype(dtype, order="F"), mask, fld, steps)
    assert np.array_equal(fg, fr), f"max |Δf| = {np.max(np.abs(fg.astype(float) - fr)):.3e}"
    if Q == 19:
        fa = None if fld is None else tuple(a.astype(dtype, order="F") for a in fld)
        F = desc.force
        mr = po.moments_forced_3d(fr, F[0], F[1], F[2], field=fa)
        for k in ("rho", "ux", "uy", "uz"):
            assert np.array_equal(m[k], mr[k].astype(np.float64)), k


@pytest.mark.parametrize("case", FORCE)
def test_guo_force_f32s_against_fp64(kmi, po, backend, case):
    desc, shape, mask, fld, steps = _force_case(kmi, case, "f32s")
    Q = 19 if len(shape) == 3 else 9
    f = rng_pops(shape, Q, seed=6, amp=0.02)
    fg, _ = _run_force_gpu(kmi, backend, desc, f, mask, fld, steps, np.float64)
    d64, *_ = _force_case(kmi, case, "f64")
    fr = _run_force_oracle(po, d64, f.copy(order="F"), mask, fld, steps)
    err = np.max(np.abs(fg - fr))
    assert err < SHIFTED_TOL["f32s"], err


def test_force_field_errors(kmi, backend):
    d = kmi.make_desc("d3q19", 6, 5, 4, 0.1, forcing="guo")
    with kmi.Lattice(d, backend) as lat:
        with pytest.raises(kmi.KrakenError, match="not KMI_FORCE_GUO_FIELD"):
            lat.set_force_field(*(np.zeros((6, 5, 4)) for _ in range(3)))
    d = kmi.make_desc("d3q19", 6, 5, 4, 0.1, forcing="guo_field")
    with kmi.Lattice(d, backend) as lat:
        with pytest.raises(kmi.ArgumentError):
            lat.set_force_field(np.zeros((6, 5, 4)), np.zeros((6, 5, 4)))
        with pytest.raises(kmi.ArgumentError):
            lat.set_force_field(*(np.zeros((5, 6, 4)) for _ in range(3)))
        # a field of zeros is no force: the rest state stays put
        lat.set_force_field(*(np.zeros((6, 5, 4)) for _ in range(3)))
        lat.step(3)
        assert abs(lat.reduce("mass") - 120.0) < 1e-12
        assert lat.reduce("max_speed") < 1e-15
