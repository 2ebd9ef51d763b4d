"""GPU parity of the interpolated bounce-back step (LI-BB V2 pre-phase, inert solids) and of
the cut-link drag reductions against the oracle."""
import numpy as np
import pytest

from test_oracle_lattice import rng_pops
from test_oracle_libb import couette_errors, couette_libb

pytestmark = pytest.mark.gpu


def _cylinder(kmi, storage, collision="trt"):
    Nx, Ny, cx, cy, R = 36, 28, 13.3, 13.8, 5.2
    qw, solid = kmi.precompute_q_wall_cylinder(Nx, Ny, cx, cy, R)
    uwx, uwy = kmi.wall_velocity_rotating_cylinder(qw, cx, cy, 0.003)
    d = kmi.make_desc("d2q9", Nx, Ny, 1, 0.05, storage=storage, collision=collision,
                      solid_mode="inert")
    return d, (Nx, Ny), solid, qw, (uwx, uwy, None)


def _sphere(kmi, storage, collision="trt"):
    n, c, R = (18, 14, 12), (7.4, 6.6, 5.8), 3.7
    qw, solid = kmi.precompute_q_wall_sphere_3d(*n, *c, R)
    r = np.random.default_rng(3)
    uw = tuple(np.where(qw > 0, 0.01 * r.uniform(-1, 1, qw.shape), 0.0) for _ in range(3))
    faces = {"west": {"kind": "periodic"}, "east": {"kind": "periodic"}}
    d = kmi.make_desc("d3q19", *n, 0.04, storage=storage, collision=collision, faces=faces,
                      solid_mode="inert")
    return d, n, solid, qw, uw


def _gpu(kmi, backend, d, f, solid, qw, uw, steps, dtype):
    with kmi.Lattice(d, backend) as lat:
        lat.set_mask(solid)
        lat.set_cut_links(*kmi.cut_links_from_dense(qw, *uw))
        lat.upload(f.astype(dtype, order="F"))
        lat.step(steps)
        out = lat.download(dtype=dtype)
        drag = (lat.reduce("drag_mea"), lat.reduce("drag_halfway"))
    return out, drag


def _oracle(po, d, f, solid, qw, uw, steps, dtype):
    fr = f.astype(dtype, order="F")
    uwr = tuple(None if u is None else u.astype(dtype) for u in uw)
    po.run_libb(d, fr, solid, qw.astype(dtype), uwr, nsteps=steps)
    return fr


@pytest.mark.parametrize("storage,dtype", [("f64", np.float64), ("f32", np.float32)])
@pytest.mark.parametrize("geom,collision", [("cylinder", "trt"), ("cylinder", "bgk"), ("sphere", "trt"),
                                            ("sphere", "bgk")])
def test_libb_bitwise(kmi, po, backend, geom, collision, storage, dtype):
    d, shape, solid, qw, uw = (_cylinder if geom == "cylinder" else _sphere)(kmi, storage, collision)
    f = rng_pops(shape, 9 if len(shape) == 2 else 19, seed=41, amp=0.03)
    print(f"[libb] {geom} {collision} {storage}", flush=True)
    fg, drag = _gpu(kmi, backend, d, f, solid, qw, uw, 37, dtype)
    fr = _oracle(po, d, f, solid, qw, uw, 37, dtype)
    assert np.array_equal(fg, fr), float(np.max(np.abs(fg.astype(float) - fr)))
    if geom == "cylinder":
        ref = po.drag_mei_2d(fr, qw.astype(dtype), uw[0].astype(dtype), uw[1].astype(dtype))
        np.testing.assert_allclose(drag[0][:2], ref, rtol=1e-10 if dtype == np.float64 else 1e-5)
    ref = po.drag_halfway(d, fr, qw.astype(dtype))
    np.testing.assert_allclose(drag[1], ref, rtol=1e-10 if dtype == np.float64 else 1e-5, atol=1e-12)


@pytest.mark.parametrize("storage", ["f32s", "f16s"])
def test_libb_shifted_storage(kmi, po, backend, storage):
    tol = {"f32s": 2e-6, "f16s": 3e-4}[storage]
    for geom in (_cylinder, _sphere):
        d, shape, solid, qw, uw = geom(kmi, storage)
        f = rng_pops(shape, 9 if len(shape) == 2 else 19, seed=42, amp=0.03)
        fg, _ = _gpu(kmi, backend, d, f, solid, qw, uw, 30, np.float64)
        fr = _oracle(po, geom(kmi, "f64")[0], f, solid, qw, uw, 30, np.float64)
        assert np.max(np.abs(fg - fr)) < tol


@pytest.mark.parametrize("stencil,q_w", [("d2q9", 0.7), ("d3q19", 0.3)])
def test_couette_ginzburg_exact_on_gpu(kmi, backend, stencil, q_w):
    Ny, u_top = (33 if stencil == "d2q9" else 17), 0.01
    d, shape, solid, qw, uwx = couette_libb(kmi, stencil, Ny, u_top, q_w)
    with kmi.Lattice(d, backend) as lat:
        lat.set_mask(solid)
        lat.set_cut_links(*kmi.cut_links_from_dense(qw, uwx))
        lat.step(10000)
        m = lat.moments()
    L2, Linf = couette_errors(m["ux"], Ny, u_top, q_w)
    assert L2 < 1e-4 and Linf < 1e-4, (L2, Linf)


def test_cut_link_validation(kmi, backend):
    d, shape, solid, qw, uw = _cylinder(kmi, "f64")
    cells, q, w, u = kmi.cut_links_from_dense(qw, *uw)
    with kmi.Lattice(d, backend) as lat:
        lat.set_mask(solid)
        bad = cells.copy()
        bad[0] = int(np.ravel_multi_index(np.argwhere(solid)[0], shape, order="F"))
        with pytest.raises(kmi.KrakenError):
            lat.set_cut_links(bad, q, w, u)
        with pytest.raises(kmi.KrakenError):
            lat.set_cut_links(cells, q, np.full(len(cells), 1.5), u)
        lat.set_cut_links(cells, q, w, u)
        with pytest.raises(kmi.KrakenError):
            lat.set_mask(np.ones(shape, dtype=np.uint8))
        lat.set_cut_links([], [], [])
        assert lat.reduce("drag_mea") == (0.0, 0.0, 0.0)


# ---- rebuilt inlet / outlet (run_cylinder_libb_2d / run_sphere_libb_3d) ------------------------

def _rebuild_case(kmi, geom, storage):
    faces = {"west": {"kind": "zouhe_velocity_rebuild"},
             "east": {"kind": "zouhe_pressure_rebuild", "rho": 1.001}}
    if geom == "cylinder":
        Nx, Ny = 48, 20
        qw, solid = kmi.precompute_q_wall_cylinder(Nx, Ny, 12, 10, 3.5)
        j = np.arange(1, Ny + 1)
        prof = 4 * 0.04 * (j - 1) * (Ny - j) / (Ny - 1) ** 2
        d = kmi.make_desc("d2q9", Nx, Ny, 1, 0.04, storage=storage, collision="trt", faces=faces,
                          solid_mode="inert")
        return d, (Nx, Ny), solid, qw, prof
    n = (20, 12, 10)
    qw, solid = kmi.precompute_q_wall_sphere_3d(*n, 5, 6, 5, 2.5)
    yy, zz = np.meshgrid(np.arange(n[1]), np.arange(n[2]), indexing="ij")
    prof = np.asfortranarray(0.03 + 0.005 * np.sin(yy) * np.cos(zz))
    d = kmi.make_desc("d3q19", *n, 0.04, storage=storage, collision="trt", faces=faces,
                      solid_mode="inert")
    return d, n, solid, qw, prof


def _rebuild_gpu(kmi, backend, d, f, solid, qw, prof, steps, dtype):
    with kmi.Lattice(d, backend) as lat:
        lat.set_mask(solid)
        lat.set_cut_links(*kmi.cut_links_from_dense(qw))
        lat.set_face_profile("west", prof)
        lat.upload(f.astype(dtype, order="F"))
        lat.step(steps)
        return lat.download(dtype=dtype)


@pytest.mark.parametrize("storage,dtype", [("f64", np.float64), ("f32", np.float32)])
@pytest.mark.parametrize("geom", ["cylinder", "sphere"])
def test_rebuild_bitwise(kmi, po, backend, geom, storage, dtype):
    d, shape, solid, qw, prof = _rebuild_case(kmi, geom, storage)
    print(f"[rebuild] {geom} {storage}", flush=True)
    f = rng_pops(shape, 9 if len(shape) == 2 else 19, seed=44, amp=0.03)
    fg = _rebuild_gpu(kmi, backend, d, f, solid, qw, prof, 33, dtype)
    fr = f.astype(dtype, order="F")
    po.run_libb(d, fr, solid, qw.astype(dtype), nsteps=33, west_profile=prof)
    assert np.all(np.isfinite(fr))
    assert np.array_equal(fg, fr), float(np.max(np.abs(fg.astype(float) - fr)))


@pytest.mark.parametrize("geom", ["cylinder", "sphere"])
def test_rebuild_f32s_against_fp64(kmi, po, backend, geom):
    d, shape, solid, qw, prof = _rebuild_case(kmi, geom, "f32s")
    f = rng_pops(shape, 9 if len(shape) == 2 else 19, seed=45, amp=0.03)
    fg = _rebuild_gpu(kmi, backend, d, f, solid, qw, prof, 33, np.float64)
    d64, *_ = _rebuild_case(kmi, geom, "f64")
    fr = f.copy(order="F")
    po.run_libb(d64, fr, solid, qw, nsteps=33, west_profile=prof)
    assert np.max(np.abs(fg - fr)) < 2e-6


def test_cylinder_driver_schaefer_turek_scaffold(kmi, backend):
    """test/test_cylinder_libb.jl: 300×80, R = 10, uniform u_in = 0.04, Re = 20, 20000 steps."""
    Nx, Ny, R, u_in = 300, 80, 10, 0.04
    nu = u_in * 2 * R / 20.0
    r = kmi.run_cylinder_libb_2d(Nx=Nx, Ny=Ny, radius=R, cx=Nx // 4, cy=Ny // 2, u_in=u_in, nu=nu,
                                 inlet="uniform", max_steps=20_000, avg_window=5_000, backend=backend)
    print(f"[cylinder] Cd={r['Cd']:.4f} Fy/Fx={r['Fy'] / r['Fx']:.2e}", flush=True)
    assert np.all(np.isfinite(r["ρ"])) and np.all(np.isfinite(r["ux"]))
    assert abs(r["Fy"] / r["Fx"]) < 0.05
    assert r["ρ"].max() < 1.3 and r["ρ"].min() > 0.8
    cx = Nx // 4
    gap = [r["ux"][cx, j] for j in range(Ny) if not r["is_solid"][cx, j]]
    expected = u_in * Ny / (Ny - 2 * R)
    assert 0.7 * expected < np.mean(gap) < 1.3 * expected
    assert 3.0 < r["Cd"] < 8.0


def test_sphere_driver_scaffold(kmi, backend):
    """test/test_sphere_libb.jl (GPU branch): 120×60×60, R = 8, Re = 20, 10000 steps."""
    Nx, Ny, Nz, R, u_in = 120, 60, 60, 8, 0.04
    nu = u_in * 2 * R / 20.0
    r = kmi.run_sphere_libb_3d(Nx=Nx, Ny=Ny, Nz=Nz, radius=R, u_in=u_in, nu=nu, max_steps=10_000,
                               avg_window=2_000, backend=backend)
    print(f"[sphere] Cd={r['Cd']:.4f}", flush=True)
    for k in ("ρ", "ux", "uy", "uz"):
        assert np.all(np.isfinite(r[k]))
    assert abs(r["Fy"] / r["Fx"]) < 0.1 and abs(r["Fz"] / r["Fx"]) < 0.1
    assert r["ρ"].max() < 1.6 and r["ρ"].min() > 0.6
    sol = r["is_solid"]
    cx = Nx // 4
    gap = r["ux"][cx][~sol[cx]]
    expected = u_in * Ny * Nz / (Ny * Nz - np.pi * R * R)
    assert 0.7 * expected < gap.mean() < 1.3 * expected
    assert 1.0 < r["Cd"] < 8.0
