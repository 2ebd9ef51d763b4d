"""Checkpoint / restart and VTK output of device lattices.

A lattice saved after n steps, written to an HDF5 checkpoint, read back and restored into a
fresh lattice continues bit for bit: after m more steps its populations equal those of the
uninterrupted run, in every storage kind (the state is exported as the storage elements
themselves), with thermal populations, masks, cut links, force fields and face profiles.
VTK snapshots hold the lattice's moments.
"""
import numpy as np
import pytest

from test_gpu_parity import _mask_blob
from test_oracle_lattice import rng_pops

pytestmark = pytest.mark.gpu

ck = pytest.importorskip("kraken_mi355x.checkpoint")


def _cases(kmi, storage):
    d3 = kmi.natural_convection_3d_desc(8, 1e3, 0.71, storage=storage)[0]
    yield "thermal3d", d3, (8, 8, 8), None
    d = kmi.cavity_desc(20, 0.05, 0.1, storage=storage, Ny=16)
    yield "cavity_mask", d, (20, 16), _mask_blob((20, 16), (9, 7), 2)
    f = {"west": {"kind": "periodic"}, "east": {"kind": "periodic"}}
    d = kmi.make_desc("d2q9", 16, 12, 1, 0.05, storage=storage, faces=f, forcing="guo_field",
                      solid_mode="inert")
    yield "libb_field", d, (16, 12), None


@pytest.mark.parametrize("storage", ["f64", "f32", "f32s", "f16s"])
def test_restart_is_bit_exact(kmi, backend, tmp_path, storage):
    try:
        ck._libhdf5()
    except ck.CheckpointError as e:
        pytest.skip(f"no HDF5 C library: {e}")
    for name, desc, shape, mask in _cases(kmi, storage):
        print(f"[restart] {name} {storage}", flush=True)
        Q = 9 if len(shape) == 2 else 19
        with kmi.Lattice(desc, backend) as a:
            if mask is not None:
                a.set_mask(mask)
            if name == "libb_field":
                qw, solid = kmi.precompute_q_wall_cylinder(16, 12, 7.5, 5.5, 2.5)
                a.set_mask(solid)
                a.set_cut_links(*kmi.cut_links_from_dense(qw))
                r = np.random.default_rng(3)
                a.set_force_field(r.uniform(-1e-4, 1e-4, shape), r.uniform(-1e-4, 1e-4, shape))
            a.upload(rng_pops(shape, Q, seed=9, amp=0.02))
            a.step(13)
            p = str(tmp_path / f"{name}_{storage}.h5")
            ck.save_checkpoint(p, a)
            a.step(11)
            fa = a.download_raw("f")
            ga = a.download_raw("g") if desc.thermal else None
            assert np.all(np.isfinite(a.download()))
            steps_a = a.steps
        snap = ck.read_checkpoint(p)
        assert snap.cycle == 13
        with ck.restore_lattice(snap, backend) as b:
            assert b.steps == 13
            b.step(11)
            assert b.steps == steps_a == 24
            assert np.array_equal(b.download_raw("f"), fa)
            if ga is not None:
                assert np.array_equal(b.download_raw("g"), ga)
        # restoring into a configuration with another identity is refused before allocating
        other = kmi.make_desc("d2q9", 5, 5, 1, 0.1)
        with pytest.raises(ck.CheckpointError, match="identity/"):
            ck.restore_lattice(snap, backend, desc=other)


def test_vtk_snapshot_holds_the_moments(kmi, backend, tmp_path):
    from kraken_mi355x import vtk

    d, _ = kmi.natural_convection_3d_desc(6, 1e4, 0.71)
    with kmi.Lattice(d, backend) as lat:
        lat.init_temperature(kmi.natural_convection_3d_initial_temperature(6))
        lat.step(20)
        pvd = vtk.create_pvd(str(tmp_path / "run"))
        path = vtk.write_lattice_snapshot(lat, str(tmp_path), pvd=pvd, velocity_vector=True)
        m = lat.moments(temperature=True)
    assert path.endswith("snapshot_0000020.vtr")
    back = vtk.read_vtr(path)
    assert back["cells"] == (6, 6, 6)
    assert np.array_equal(back["rho"], m["rho"]) and np.array_equal(back["Temp"], m["Temp"])
    for k, comp in zip(("ux", "uy", "uz"), back["velocity"]):
        assert np.array_equal(comp, m[k])
    assert vtk.read_pvd(pvd.path) == [(20.0, path)]
