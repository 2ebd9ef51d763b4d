"""GPU parity of the Guo body force and the per-node (profile) Zou–He faces against the
oracle: bit-identical in FP64 / FP32 storage, within the shifted-storage tolerances
otherwise; time-dependent profiles and scalar faces updated between steps."""
import numpy as np
import pytest

from test_gpu_parity import SHIFTED_TOL, _mask_blob
from test_oracle_lattice import rng_pops

pytestmark = pytest.mark.gpu


def _forced_desc(kmi, storage, forcing="guo"):
    per = {"west": {"kind": "periodic"}, "east": {"kind": "periodic"}}
    return kmi.make_desc("d2q9", 24, 16, 1, 0.04, storage=storage, faces=per, forcing=forcing,
                         force=(1e-4, -3e-5))


def _profile_case(kmi, storage):
    nx, ny = 30, 14
    r = np.random.default_rng(21)
    prof = {"west": (0.03 + 0.01 * r.uniform(-1, 1, ny), 0.005 * r.uniform(-1, 1, ny)),
            "east": (1.0 + 0.01 * r.uniform(-1, 1, ny), None),
            "south": (0.02 * r.uniform(-1, 1, nx), 0.004 * r.uniform(-1, 1, nx))}
    faces = {"west": {"kind": "zouhe_velocity_profile"}, "east": {"kind": "zouhe_pressure_profile", "rho": 1.0},
             "south": {"kind": "zouhe_velocity_profile"}, "north": {"kind": "zouhe_velocity", "u": (0.02, 0, 0)}}
    d = kmi.make_desc("d2q9", nx, ny, 1, 0.05, storage=storage, faces=faces,
                      bc_order=["west", "east", "south", "north"])
    return d, (nx, ny), prof


def _run(kmi, backend, desc, f, steps, dtype, mask=None, prof=None):
    with kmi.Lattice(desc, backend) as lat:
        if mask is not None:
            lat.set_mask(mask)
        for face, (a, b) in (prof or {}).items():
            lat.set_face_profile(face, a, b)
        lat.upload(f.astype(dtype, order="F"))
        lat.step(steps)
        return lat.download(dtype=dtype), lat.moments()


@pytest.mark.parametrize("storage,dtype", [("f64", np.float64), ("f32", np.float32)])
@pytest.mark.parametrize("forcing", ["guo", "guo_plain"])
def test_guo_force_bitwise(kmi, po, backend, storage, dtype, forcing):
    d = _forced_desc(kmi, storage, forcing)
    mask = _mask_blob((24, 16), (11, 7), 2)
    f = rng_pops((24, 16), 9, seed=31, amp=0.03)
    print(f"[force] {forcing} {storage}", flush=True)
    fg, mg = _run(kmi, backend, d, f, 41, dtype, mask=mask)
    fr = f.astype(dtype, order="F")
    po.run(d, fr, None, 41, mask=mask)
    assert np.array_equal(fg, fr), float(np.max(np.abs(fg.astype(float) - fr)))
    mr = po.moments_forced(fr, 1e-4, -3e-5) if forcing == "guo" else po.moments(fr)
    for k in ("rho", "ux", "uy"):
        assert np.array_equal(mg[k], mr[k].astype(np.float64)), k


@pytest.mark.parametrize("storage,dtype", [("f64", np.float64), ("f32", np.float32)])
def test_profile_faces_bitwise(kmi, po, backend, storage, dtype):
    d, shape, prof = _profile_case(kmi, storage)
    f = rng_pops(shape, 9, seed=32, amp=0.03)
    print(f"[profile] {storage}", flush=True)
    fg, _ = _run(kmi, backend, d, f, 37, dtype, prof=prof)
    fr = f.astype(dtype, order="F")
    po.run_profiles(d, fr, prof, 37)
    assert np.array_equal(fg, fr), float(np.max(np.abs(fg.astype(float) - fr)))


@pytest.mark.parametrize("storage", ["f32s", "f16s"])
def test_forcing_and_profiles_shifted_storage(kmi, po, backend, storage):
    tol = SHIFTED_TOL[storage]
    d = _forced_desc(kmi, storage)
    f = rng_pops((24, 16), 9, seed=33, amp=0.03)
    fg, _ = _run(kmi, backend, d, f, 40, np.float64)
    fr = f.copy(order="F")
    po.run(_forced_desc(kmi, "f64"), fr, None, 40)
    assert np.max(np.abs(fg - fr)) < tol
    d, shape, prof = _profile_case(kmi, storage)
    f = rng_pops(shape, 9, seed=34, amp=0.03)
    fg, _ = _run(kmi, backend, d, f, 40, np.float64, prof=prof)
    fr = f.copy(order="F")
    po.run_profiles(_profile_case(kmi, "f64")[0], fr, prof, 40)
    assert np.max(np.abs(fg - fr)) < tol


def test_time_dependent_profiles_and_faces(kmi, po, backend):
    """Profiles and a scalar face refreshed before every step (handlers.jl:203-218)."""
    d, shape, prof = _profile_case(kmi, "f64")
    nx, ny = shape
    f = rng_pops(shape, 9, seed=35, amp=0.03)
    fr = f.copy(order="F")
    y = np.arange(ny) + 0.5
    with kmi.Lattice(d, backend) as lat:
        lat.upload(f)
        lat.set_face_profile("east", prof["east"][0])
        lat.set_face_profile("south", *prof["south"])
        for step in range(1, 13):
            uin = 0.03 * (1 + 0.1 * np.sin(0.3 * step)) * y * (ny - y) * 4 / ny ** 2
            vin = 0.002 * np.cos(0.2 * step) * np.ones(ny)
            lid = 0.02 * np.cos(0.1 * step)
            lat.set_face_profile("west", uin, vin)
            lat.set_face("north", {"kind": "zouhe_velocity", "u": (lid, 0, 0)})
            lat.step(1)
            d.faces[3].u[0] = lid
            po.run_profiles(d, fr, {"west": (uin, vin), "east": prof["east"], "south": prof["south"]}, 1)
        fg = lat.download()
    assert np.array_equal(fg, fr), float(np.max(np.abs(fg - fr)))


def test_face_updates_are_validated(kmi, backend):
    d, shape, prof = _profile_case(kmi, "f64")
    with kmi.Lattice(d, backend) as lat:
        with pytest.raises(kmi.KrakenError):
            lat.set_face("north", {"kind": "wall"})  # kind cannot change
        with pytest.raises(kmi.KrakenError):
            lat.set_face_profile("north", np.zeros(shape[0]), np.zeros(shape[0]))  # not a profile face
        with pytest.raises(kmi.KrakenError):
            lat.set_face_profile("east", np.zeros(shape[1]))  # density must be positive
