"""The oracle's generic descriptor-driven step equals the restated reference kernel chains
bit for bit, for every face / collision / thermal combination the library implements."""
import numpy as np
import pytest

from test_oracle_lattice import rng_pops


def faces(**kw):
    return {k: (v if isinstance(v, dict) else {"kind": v}) for k, v in kw.items()}


PER2 = dict(west="periodic", east="periodic", south="periodic", north="periodic")


def test_pull_generic_equals_reference_streams_2d(po, kmi):
    nx, ny = 9, 7
    f = rng_pops((nx, ny), 9, seed=11)
    cases = [
        ({}, "stream_2d"),  # all walls: collide_stream_2d.jl:75-101
        (PER2, "stream_fully_periodic_2d"),  # stream_periodic_2d.jl:79-98
        (dict(west="periodic", east="periodic"), "stream_periodic_x_wall_y_2d"),  # :5-55
    ]
    for fc, ref_name in cases:
        d = kmi.make_desc("d2q9", nx, ny, 1, 0.1, faces=faces(**fc))
        ref = np.empty_like(f, order="F")
        po.call(ref_name, ref, f, nx, ny)
        assert np.array_equal(po.pull(d, f), ref), ref_name


def test_pull_generic_equals_reference_streams_3d(po, kmi):
    nx, ny, nz = 6, 5, 4
    f = rng_pops((nx, ny, nz), 19, seed=12)
    for px, py, pz in [(0, 0, 0), (1, 0, 1), (0, 0, 1), (1, 1, 1)]:
        fc = {}
        for ax, p in zip((("west", "east"), ("south", "north"), ("bottom", "top")), (px, py, pz)):
            if p:
                fc[ax[0]] = fc[ax[1]] = "periodic"
        d = kmi.make_desc("d3q19", nx, ny, nz, 0.1, faces=faces(**fc))
        ref = np.empty_like(f, order="F")
        po.call("stream_3d", ref, f, nx, ny, nz, px, py, pz)
        assert np.array_equal(po.pull(d, f), ref), (px, py, pz)
    # moving y walls (stream_periodic_3d.jl:106-166)
    Ub, Ut = 0.04, -0.01
    d = kmi.make_desc("d3q19", nx, ny, nz, 0.1, faces=faces(
        west="periodic", east="periodic", bottom="periodic", top="periodic",
        south={"kind": "moving_wall", "u": (Ub, 0, 0)}, north={"kind": "moving_wall", "u": (Ut, 0, 0)}))
    ref = np.empty_like(f, order="F")
    po.call("stream_periodic_xz_movingwall_y_3d", ref, f, Ub, Ut, 1.0, nx, ny, nz)
    assert np.array_equal(po.pull(d, f), ref)


def test_fused_kernels_equal_stream_then_collide(po):
    nx, ny = 10, 8
    f = rng_pops((nx, ny), 9, seed=13)
    solid = np.zeros((nx, ny), dtype=np.uint8, order="F")
    solid[4:6, 3:5] = 1
    st = np.empty_like(f, order="F")
    po.call("stream_2d", st, f, nx, ny)
    a = st.copy(order="F")
    po.call("collide_2d", a, solid, nx, ny, 1.2)
    b = np.empty_like(f, order="F")
    po.call("fused_bgk_step_2d", b, f, None, None, None, solid, nx, ny, 1.2)
    assert np.array_equal(a, b)
    a = st.copy(order="F")
    po.call("collide_trt_2d", a, solid, nx, ny, 1.3, 0.7)
    po.call("fused_trt_step_2d", b, f, None, None, None, solid, nx, ny, 1.3, 0.7)
    assert np.array_equal(a, b)
    g = rng_pops((5, 4, 6), 19, seed=14)
    st3 = np.empty_like(g, order="F")
    po.call("stream_3d", st3, g, 5, 4, 6, 1, 1, 1)
    po.call("collide_trt_3d", st3, None, 5, 4, 6, 1.88, 0.15)
    fu = np.empty_like(g, order="F")
    po.call("fused_trt_3d", fu, g, 5, 4, 6, 1, 1, 1, 1.88, 0.15)
    assert np.array_equal(st3, fu)


def test_generic_cavity_equals_run_cavity_2d_chain(po, kmi):
    """basic.jl:121-136: stream_2d! → apply_zou_he_north_2d! → collide_2d!, per step."""
    nx, ny, nu, ul = 12, 10, 0.05, 0.1
    d = kmi.cavity_desc(nx, nu, ul, Ny=ny)
    f = rng_pops((nx, ny), 9, seed=15)
    solid = np.zeros((nx, ny), dtype=np.uint8, order="F")
    solid[5, 4] = 1
    gen = f.copy(order="F")
    po.run(d, gen, None, 25, mask=solid)
    a = f.copy(order="F")
    b = np.empty_like(a, order="F")
    om = 1.0 / (3 * nu + 0.5)
    for _ in range(25):
        po.call("stream_2d", b, a, nx, ny)
        po.call("zou_he_north_2d", b, ul, nx, ny)
        po.call("collide_2d", b, solid, nx, ny, om)
        a, b = b, a
    assert np.array_equal(gen, a)


def test_generic_mrt_equals_test_mrt_loop(po, kmi):
    """test/test_mrt.jl:19-25: stream_2d! → apply_zou_he_north_2d! → collide_mrt_2d!."""
    nx, ny, nu, ul = 11, 9, 0.02, 0.1
    d = kmi.cavity_desc(nx, nu, ul, Ny=ny, collision="mrt")
    f = rng_pops((nx, ny), 9, seed=16)
    gen = f.copy(order="F")
    po.run(d, gen, None, 20)
    a, b = f.copy(order="F"), np.empty_like(f, order="F")
    for _ in range(20):
        po.call("stream_2d", b, a, nx, ny)
        po.call("zou_he_north_2d", b, ul, nx, ny)
        po.call("collide_mrt_2d", b, None, nx, ny, 1.4, 1.4, 1.2, 1.0 / (3 * nu + 0.5))
        a, b = b, a
    assert np.array_equal(gen, a)


def test_generic_channel_closures_in_order(po, kmi):
    """Zou–He velocity inlet west + pressure outlet east, applied in .krk order."""
    nx, ny = 14, 8
    d = kmi.make_desc("d2q9", nx, ny, 1, 0.1, faces=faces(
        west={"kind": "zouhe_velocity", "u": (0.03, 0, 0)},
        east={"kind": "zouhe_pressure", "rho": 1.0}, south="wall", north="wall"),
        bc_order=["east", "west"])
    f = rng_pops((nx, ny), 9, seed=17)
    gen = f.copy(order="F")
    po.run(d, gen, None, 10)
    a, b = f.copy(order="F"), np.empty_like(f, order="F")
    for _ in range(10):
        po.call("stream_2d", b, a, nx, ny)
        po.call("zou_he_pressure_east_2d", b, 1.0, nx, ny)
        po.call("zou_he_west_2d", b, 0.03, nx, ny)
        po.call("collide_2d", b, None, nx, ny, 1.0 / 0.8)
        a, b = b, a
    assert np.array_equal(gen, a)


def test_generic_thermal_split_equals_rayleigh_benard_chain(po, kmi):
    """thermal.jl:100-133 (vertical orientation)."""
    Nx, Ny = 16, 8
    d, p = kmi.rayleigh_benard_desc(Nx, Ny, Ra=5e3, Pr=0.71)
    f = rng_pops((Nx, Ny), 9, seed=18, amp=0.01)
    T0 = kmi.rb_initial_temperature(Nx, Ny, 1.0, 0.0)
    g = np.asfortranarray(po.W9[None, None, :] * T0[:, :, None])
    gf, gg = f.copy(order="F"), g.copy(order="F")
    po.run(d, gf, gg, 15)
    fa, fb = f.copy(order="F"), np.empty_like(f, order="F")
    ga, gb = g.copy(order="F"), np.empty_like(g, order="F")
    Temp = np.empty((Nx, Ny), order="F")
    rho, ux, uy = (np.empty((Nx, Ny), order="F") for _ in range(3))
    om_f = 1.0 / (3 * p["nu"] + 0.5)
    om_T = 1.0 / (3 * p["alpha"] + 0.5)
    import ctypes
    dirf = np.array([0, 0, 1, 1], dtype=np.int32)
    Tf = np.array([0.0, 0.0, 1.0, 0.0])
    for _ in range(15):
        po.call("stream_periodic_x_wall_y_2d", fb, fa, Nx, Ny)
        po.call("stream_periodic_x_wall_y_2d", gb, ga, Nx, Ny)
        po.lib().ko_apply_fixed_temp_2d_d(gb.ctypes.data_as(ctypes.c_void_p), dirf.ctypes.data_as(ctypes.c_void_p),
                                          Tf.ctypes.data_as(ctypes.c_void_p), ctypes.c_int64(Nx), ctypes.c_int64(Ny))
        po.call("compute_temperature_2d", Temp, gb, Nx, Ny)
        po.call("macroscopic_2d", rho, ux, uy, fb, Nx, Ny)
        po.call("collide_thermal_2d", gb, ux, uy, Nx, Ny, om_T)
        po.call("collide_boussinesq_2d", fb, Temp, None, Nx, Ny, om_f, p["beta_g"], p["T_ref"])
        fa, fb = fb, fa
        ga, gb = gb, ga
    assert np.array_equal(gf, fa) and np.array_equal(gg, ga)


def test_generic_thermal_fused_equals_natconv_kernel(po, kmi):
    """fused_thermal_2d.jl:178-207 (hot west, cold east, adiabatic south/north)."""
    N = 12
    d = kmi.natural_convection_desc(N, 1e4, 0.71)
    f = rng_pops((N, N), 9, seed=19, amp=0.01)
    T0 = kmi.natural_convection_initial_temperature(N)
    g = np.asfortranarray(po.W9[None, None, :] * T0[:, :, None])
    gf, gg = f.copy(order="F"), g.copy(order="F")
    po.run(d, gf, gg, 12)
    fa, fb = f.copy(order="F"), np.empty_like(f, order="F")
    ga, gb = g.copy(order="F"), np.empty_like(g, order="F")
    for _ in range(12):
        po.call("fused_natconv_step_2d", fb, fa, gb, ga, None, N, N, d.omega, d.omega_T, d.beta_g,
                d.T_ref, 1.0, 0.0)
        fa, fb = fb, fa
        ga, gb = gb, ga
    assert np.array_equal(gf, fa) and np.array_equal(gg, ga)


def test_oracle_float32_path_runs_in_float32(po, kmi):
    d = kmi.cavity_desc(8, 0.1, 0.1)
    f = rng_pops((8, 8), 9, seed=20).astype(np.float32, order="F")
    f64 = f.astype(np.float64, order="F")
    po.run(d, f, None, 30)
    po.run(d, f64, None, 30)
    err = np.max(np.abs(f - f64))
    assert 0 < err < 1e-5
