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
