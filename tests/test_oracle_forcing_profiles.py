"""Guo body force (collide_guo_2d.jl:5-72, macroscopic.jl:84-98) and per-node Zou–He faces
(boundary_spatial_2d.jl) in the oracle, pinned by what they are built to satisfy:

  Guo BGK:          mass is conserved and the cell momentum grows by exactly F per
                    collision (Σ_q S_q = 0, Σ_q c_q S_q = (1 − ω/2) F and the F/2 velocity
                    shift); F = 0 reduces to plain BGK;
  velocity profile: at every non-corner face node the closure imposes ρ_w and momentum
                    ρ_w (u_x, u_y), derived from the lattice vectors; corners are untouched;
                    u_y = 0 reduces to the scalar closures of boundary_2d.jl;
  pressure profile: a constant profile equals the scalar pressure closure off the corners;
  physics:          the reference's Poiseuille test (test/test_poiseuille.jl) through the
                    descriptor chain, and a parabolic inflow profile carried down a channel.
"""
import numpy as np
import pytest

from test_oracle_lattice import C9, rng_pops

C = np.array(C9, dtype=float)


def test_guo_collision_adds_the_force_to_the_momentum(po):
    f = rng_pops((7, 5), 9, seed=3, amp=0.05)
    g = f.copy(order="F")
    fx, fy, om = 3e-4, -2e-4, 1.37
    po.call("collide_guo_2d", g, None, 7, 5, om, fx, fy)
    np.testing.assert_allclose(g.sum(-1), f.sum(-1), rtol=0, atol=1e-15)
    np.testing.assert_allclose(g @ C[:, 0], f @ C[:, 0] + fx, rtol=0, atol=1e-15)
    np.testing.assert_allclose(g @ C[:, 1], f @ C[:, 1] + fy, rtol=0, atol=1e-15)


def test_guo_without_force_is_bgk(po):
    f = rng_pops((6, 6), 9, seed=4, amp=0.05)
    a, b = f.copy(order="F"), f.copy(order="F")
    po.call("collide_guo_2d", a, None, 6, 6, 1.2, 0.0, 0.0)
    po.call("collide_2d", b, None, 6, 6, 1.2)
    np.testing.assert_allclose(a, b, rtol=0, atol=2e-16)


def test_forced_moments_shift_velocity_by_half_force(po):
    f = rng_pops((5, 4), 9, seed=5, amp=0.05)
    m = po.moments(f)
    mf = po.moments_forced(f, 2e-3, -1e-3)
    np.testing.assert_array_equal(mf["rho"], m["rho"])
    np.testing.assert_allclose(mf["ux"] * mf["rho"], m["ux"] * m["rho"] + 1e-3, rtol=0, atol=1e-16)
    np.testing.assert_allclose(mf["uy"] * mf["rho"], m["uy"] * m["rho"] - 5e-4, rtol=0, atol=1e-16)


FACES = {"west": (0, -1), "south": (1, -1), "north": (1, 1)}  # (normal axis, outward sign)


@pytest.mark.parametrize("face", ["west", "south", "north"])
def test_velocity_profile_imposes_node_moments(po, face):
    nx, ny = 7, 6
    f = rng_pops((nx, ny), 9, seed=10, amp=0.05)
    f0 = f.copy(order="F")
    ax, s = FACES[face]
    n = ny if face == "west" else nx
    r = np.random.default_rng(1)
    ux, uy = r.uniform(-0.05, 0.05, n), r.uniform(-0.05, 0.05, n)
    po.call("zou_he_velocity_spatial_2d", f, {"west": 0, "south": 2, "north": 3}[face], ux, uy, nx, ny)
    cn = C[:, ax] * s
    par, out = np.where(cn == 0)[0], np.where(cn > 0)[0]
    for p in range(1, n - 1):
        idx = (0, p) if face == "west" else (p, 0 if face == "south" else ny - 1)
        pre, post = f0[idx], f[idx]
        un = (ux[p], uy[p])[ax] * s  # outward normal velocity
        rho_w = (pre[par].sum() + 2 * pre[out].sum()) / (1 + un)
        assert abs(post.sum() - rho_w) < 2e-15
        assert abs(post @ C[:, 0] - rho_w * ux[p]) < 2e-15
        assert abs(post @ C[:, 1] - rho_w * uy[p]) < 2e-15
    for p in (0, n - 1):  # corners keep the pulled populations
        idx = (0, p) if face == "west" else (p, 0 if face == "south" else ny - 1)
        assert np.array_equal(f[idx], f0[idx])


@pytest.mark.parametrize("face", ["west", "south", "north"])
def test_velocity_profile_with_zero_normal_part_is_the_scalar_closure(po, face):
    nx, ny = 6, 5
    f = rng_pops((nx, ny), 9, seed=11, amp=0.05)
    a, b = f.copy(order="F"), f.copy(order="F")
    n = ny if face == "west" else nx
    u = 0.04
    ux = np.full(n, u if face != "west" else u)
    uy = np.zeros(n)
    po.call("zou_he_velocity_spatial_2d", a, {"west": 0, "south": 2, "north": 3}[face], ux, uy, nx, ny)
    po.call(f"zou_he_{face}_2d", b, u, nx, ny)
    sl = (0, slice(1, n - 1)) if face == "west" else (slice(1, n - 1), 0 if face == "south" else ny - 1)
    np.testing.assert_allclose(a[sl], b[sl], rtol=0, atol=2e-16)


def test_pressure_profile_constant_equals_scalar_closure(po):
    nx, ny = 6, 7
    f = rng_pops((nx, ny), 9, seed=12, amp=0.05)
    a, b = f.copy(order="F"), f.copy(order="F")
    po.call("zou_he_pressure_east_spatial_2d", a, np.full(ny, 1.02), nx, ny)
    po.call("zou_he_pressure_east_2d", b, 1.02, nx, ny)
    assert np.array_equal(a[-1, 1:-1], b[-1, 1:-1])
    assert np.array_equal(a[-1, [0, -1]], f[-1, [0, -1]])


def test_poiseuille_body_force(kmi, po):
    """test/test_poiseuille.jl: 4×32, ν = 0.1, F_x = 1e-5, 10000 steps, periodic x, walls y:
    L∞ error against F/(2ν)(j − ½)(N_y + ½ − j) below 2 %, mass conserved to 1e-3."""
    Nx, Ny, nu, Fx = 4, 32, 0.1, 1e-5
    d = kmi.make_desc("d2q9", Nx, Ny, 1, nu, faces={"west": {"kind": "periodic"}, "east": {"kind": "periodic"}},
                      forcing="guo_plain", force=(Fx, 0.0))
    f = po.rest((Nx, Ny))
    po.run(d, f, None, 10000)
    m = po.moments(f)
    j = np.arange(1, Ny + 1)
    ua = Fx / (2 * nu) * (j - 0.5) * (Ny + 0.5 - j)
    err = np.max(np.abs(m["ux"][1, 1:-1] - ua[1:-1])) / ua.max()
    assert err < 0.02, err
    assert abs(m["rho"].sum() - Nx * Ny) / (Nx * Ny) < 1e-3


def test_parabolic_inflow_profile_reaches_the_channel_centre(kmi, po):
    """West velocity profile u_x = 4U y(H − y)/H² at cell centres (handlers.jl:152-187
    convention), east pressure ρ = 1, no-slip walls: the developed profile midway down the
    channel matches the inflow to a few per cent."""
    Nx, Ny, U, nu = 48, 16, 0.04, 0.05
    y = (np.arange(Ny) + 0.5)
    uin = 4 * U * y * (Ny - y) / Ny ** 2
    faces = {"west": {"kind": "zouhe_velocity_profile"}, "east": {"kind": "zouhe_pressure", "rho": 1.0},
             "south": {"kind": "wall"}, "north": {"kind": "wall"}}
    d = kmi.make_desc("d2q9", Nx, Ny, 1, nu, faces=faces)
    f = po.rest((Nx, Ny))
    po.run_profiles(d, f, {"west": (uin, np.zeros(Ny))}, 6000)
    m = po.moments(f)
    mid = m["ux"][Nx // 2, 1:-1]
    assert np.all(np.isfinite(mid))
    assert np.max(np.abs(mid - uin[1:-1])) / U < 0.06
