"""The C-ABI library loads, exports every symbol include/kraken_mi355x.h declares, and
rejects what the reference rejects — no device needed."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_symbols():
    txt = open(os.path.join(ROOT, "include", "kraken_mi355x.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(kmi_[a-z0-9_]+)\s*\(", txt)))


def test_every_declared_symbol_is_exported(kmi):
    lib = kmi.load()
    syms = header_symbols()
    assert len(syms) >= 20
    missing = [s for s in syms if not hasattr(lib, s)]
    assert not missing, missing
    # and the binding knows each of them
    assert set(syms) <= set(kmi._lib.SIGNATURES)


def test_abi_version_and_layout(kmi):
    lib = kmi.load()
    assert lib.kmi_abi_version() == 2
    assert lib.kmi_desc_sizeof() == ctypes.sizeof(kmi._lib.KmiDesc)


def test_desc_defaults(kmi):
    d = kmi._lib.desc_init(kmi._lib.D2Q9, 128, 128, 7, 0.128)
    assert d.nz == 1 and d.storage == kmi._lib.STORE_F64 and d.collision == kmi._lib.BGK
    assert abs(d.omega - 1 / (3 * 0.128 + 0.5)) == 0.0
    assert all(d.faces[f].kind == kmi._lib.FACE_WALL for f in range(6))
    assert kmi.load().kmi_desc_validate(ctypes.byref(d)) == 0


def _validate(kmi, d):
    lib = kmi.load()
    rc = lib.kmi_desc_validate(ctypes.byref(d))
    return rc, lib.kmi_last_error(None).decode()


@pytest.mark.parametrize("case,status,needle", [
    ("unpaired_periodic", 1, "periodic"),
    ("mrt_3d", 5, "MRT"),
    ("zh_velocity_east", 5, "east"),
    ("pressure_north", 5, "Pressure boundary condition on face north"),
    ("thermal_trt", 5, "BGK"),
    ("thermal_3d_fused", 5, "unforced velocity"),
    ("guo_trt", 5, "BGK"),
    ("guo_thermal", 5, "thermal"),
    ("zh_3d_normal", 1, "1 + sign*u_n"),
    ("dirichlet_periodic", 1, "periodic"),
    ("bad_size", 1, "positive"),
    ("too_many_populations", 5, "2^31"),
])
def test_desc_validation_errors(kmi, case, status, needle):
    mk = kmi.make_desc
    if case == "unpaired_periodic":
        d = mk("d2q9", 8, 8, faces={"west": {"kind": "periodic"}})
    elif case == "mrt_3d":
        d = mk("d3q19", 8, 8, 8, collision="mrt")
    elif case == "zh_velocity_east":
        d = mk("d2q9", 8, 8, faces={"east": {"kind": "zouhe_velocity", "u": (0.1, 0, 0)}})
    elif case == "pressure_north":
        d = mk("d2q9", 8, 8, faces={"north": {"kind": "zouhe_pressure", "rho": 1.0}})
    elif case == "thermal_trt":
        d = mk("d2q9", 8, 8, collision="trt", thermal=True)
    elif case == "thermal_3d_fused":
        d = mk("d3q19", 8, 8, 8, thermal=True, coupling="fused")
    elif case == "guo_trt":
        d = mk("d3q19", 8, 8, 8, collision="trt", forcing="guo_field")
    elif case == "guo_thermal":
        d = mk("d3q19", 8, 8, 8, thermal=True, forcing="guo")
    elif case == "zh_3d_normal":
        d = mk("d3q19", 8, 8, 8, faces={"top": {"kind": "zouhe_velocity", "u": (0.1, 0, -1.0)}})
    elif case == "dirichlet_periodic":
        d = mk("d2q9", 8, 8, thermal=True, faces={"west": {"kind": "periodic", "thermal": "dirichlet"},
                                                  "east": {"kind": "periodic"}})
    elif case == "too_many_populations":
        d = mk("d3q19", 1024, 1024, 128)  # 2.55e9 populations: 32-bit plane offsets would overflow
    else:
        d = mk("d2q9", 0, 8)
    rc, msg = _validate(kmi, d)
    assert rc == status and needle in msg, (rc, msg)


def _has_device(kmi):
    n = ctypes.c_int()
    kmi.load().kmi_device_count(ctypes.byref(n))
    return n.value > 0


def test_no_device_is_an_error_not_a_fallback(kmi):
    """Backend selection raises when the device is missing (test_simulation_runner.jl:327-400)."""
    if _has_device(kmi):
        pytest.skip("a device is visible here")
    lib = kmi.load()
    h = ctypes.c_void_p()
    rc = lib.kmi_ctx_create(0, ctypes.byref(h))
    assert rc == 2 and "unavailable" in lib.kmi_last_error(None).decode()
    with pytest.raises(kmi.ArgumentError, match="mi355x"):
        kmi.MI355XBackend(0)
    with pytest.raises(kmi.ArgumentError):
        kmi.run_cavity_2d(kmi.lbm_config(kmi.D2Q9(), Nx=8, Ny=8, nu=0.1, max_steps=1))


def test_d3q19_zou_he_faces_validate(kmi):
    """Zou–He velocity and pressure are accepted on every D3Q19 face (boundary_3d.jl:152-290)."""
    faces = {"west": {"kind": "zouhe_velocity", "u": (0.02, 0.01, 0)}, "east": {"kind": "zouhe_pressure", "rho": 1.0},
             "south": {"kind": "zouhe_velocity", "u": (0.01, 0, 0)}, "north": {"kind": "zouhe_pressure", "rho": 0.99},
             "bottom": {"kind": "zouhe_velocity", "u": (0, 0, 0.01)}, "top": {"kind": "zouhe_velocity", "u": (0.1, 0, 0)}}
    rc, msg = _validate(kmi, kmi.make_desc("d3q19", 8, 8, 8, faces=faces))
    assert rc == 0, msg


def test_3d_thermal_and_force_fields_validate(kmi):
    """The D3Q19 thermal chain (split coupling, Dirichlet on any face) and the Guo force in 3D
    and per node are accepted."""
    faces = {"bottom": {"kind": "wall", "thermal": "dirichlet", "T": 1.0},
             "top": {"kind": "wall", "thermal": "dirichlet", "T": 0.0}}
    for d in (kmi.make_desc("d3q19", 8, 6, 5, thermal=True, faces=faces),
              kmi.make_desc("d3q19", 8, 6, 5, forcing="guo", force=(1e-5, 0.0, 2e-5)),
              kmi.make_desc("d3q19", 8, 6, 5, forcing="guo_field"),
              kmi.make_desc("d2q9", 8, 6, 1, forcing="guo_field")):
        rc, msg = _validate(kmi, d)
        assert rc == 0, msg
