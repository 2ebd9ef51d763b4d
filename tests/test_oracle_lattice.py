"""Pins of the CPU oracle's per-cell algebra against what the mathematics fixes.

Each test derives its expectation independently of the oracle's code: lattice vectors
from the reference's documented velocity ordering (src/lattice/d2q9.jl:6-15,
equilibrium_helpers_3d.jl:4-23), moment identities of the second-order equilibrium,
the relaxation laws of BGK / TRT / MRT modes, and the moment constraints the Zou–He
closures are built to satisfy.
"""
import numpy as np
import pytest

# Documented velocity sets (reference docstrings), in q order.
C9 = [(0, 0), (1, 0), (0, 1), (-1, 0), (0, -1), (1, 1), (-1, 1), (-1, -1), (1, -1)]
C19 = [(0, 0, 0), (1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0), (0, 0, 1), (0, 0, -1),
       (1, 1, 0), (-1, 1, 0), (1, -1, 0), (-1, -1, 0), (1, 0, 1), (-1, 0, 1), (1, 0, -1),
       (-1, 0, -1), (0, 1, 1), (0, -1, 1), (0, 1, -1), (0, -1, -1)]


def rng_pops(shape, Q, seed=0, amp=0.02):
    r = np.random.default_rng(seed)
    W = {9: [4 / 9] + [1 / 9] * 4 + [1 / 36] * 4, 19: [1 / 3] + [1 / 18] * 6 + [1 / 36] * 12}[Q]
    f = np.empty(tuple(shape) + (Q,), order="F")
    for q in range(Q):
        f[..., q] = W[q] * (1.0 + amp * r.uniform(-1, 1, size=shape))
    return f


def bgk_equilibrium_2d(po, f):
    """ω = 1 turns a BGK collision into f ← feq(ρ(f), u(f))."""
    g = f.copy(order="F")
    nx, ny, _ = f.shape
    po.call("collide_2d", g, None, nx, ny, 1.0)
    return g


def test_equilibrium_moments_d2q9(po):
    f = rng_pops((6, 5), 9, seed=1)
    feq = bgk_equilibrium_2d(po, f)
    c = np.array(C9, dtype=float)
    rho = f.sum(-1)
    j = f @ c
    u = j / rho[..., None]
    assert np.allclose(feq.sum(-1), rho, rtol=0, atol=1e-15)
    assert np.allclose(feq @ c, j, rtol=0, atol=1e-15)
    # second moment Σ feq c_a c_b = ρ (δ_ab/3 + u_a u_b)
    for a in range(2):
        for b in range(2):
            m2 = (feq * (c[:, a] * c[:, b])).sum(-1)
            assert np.allclose(m2, rho * ((a == b) / 3.0 + u[..., a] * u[..., b]), atol=2e-15)


def test_equilibrium_moments_d3q19(po):
    f = rng_pops((4, 3, 5), 19, seed=2)
    g = f.copy(order="F")
    po.call("collide_3d", g, None, 4, 3, 5, 1.0)
    c = np.array(C19, dtype=float)
    rho = f.sum(-1)
    u = (f @ c) / rho[..., None]
    assert np.allclose(g.sum(-1), rho, atol=1e-15)
    assert np.allclose(g @ c, rho[..., None] * u, atol=1e-15)
    for a in range(3):
        for b in range(3):
            m2 = (g * (c[:, a] * c[:, b])).sum(-1)
            assert np.allclose(m2, rho * ((a == b) / 3.0 + u[..., a] * u[..., b]), atol=2e-15)


def test_rest_weights_sum_to_one(po):
    for shape, Q in (((3, 3), 9), ((3, 3, 3), 19)):
        f = po.rest(shape)
        g = f.copy(order="F")
        if Q == 9:
            po.call("collide_2d", g, None, 3, 3, 1.7)
        else:
            po.call("collide_3d", g, None, 3, 3, 3, 1.7)
        assert np.allclose(f.sum(-1), 1.0, atol=1e-15)
        # rest equilibrium is a fixed point of the collision, to round-off
        # (test/test_fused_trt_2d.jl:260-278 uses 1e-12)
        assert np.max(np.abs(g - f)) < 1e-15


@pytest.mark.parametrize("Q", [9, 19])
def test_delta_streaming_and_bounce_back(po, Q):
    """A lone population moves by c_q in the interior and comes back reversed at a wall."""
    C = C9 if Q == 9 else C19
    shape = (7, 6) if Q == 9 else (7, 6, 5)
    center = tuple(s // 2 for s in shape)
    for q in range(1, Q):
        f = np.zeros(shape + (Q,), order="F")
        f[center + (q,)] = 1.0
        out = np.zeros_like(f, order="F")
        if Q == 9:
            po.call("stream_2d", out, f, *shape)
        else:
            po.call("stream_3d", out, f, *shape, 0, 0, 0)
        dest = tuple(center[a] + C[q][a] for a in range(len(shape)))
        assert out[dest + (q,)] == 1.0 and out.sum() == 1.0
        # at the wall the link leaving the domain returns in the opposite slot
        edge = tuple(0 if C[q][a] < 0 else (shape[a] - 1 if C[q][a] > 0 else center[a])
                     for a in range(len(shape)))
        f = np.zeros(shape + (Q,), order="F")
        f[edge + (q,)] = 1.0
        out = np.zeros_like(f, order="F")
        if Q == 9:
            po.call("stream_2d", out, f, *shape)
        else:
            po.call("stream_3d", out, f, *shape, 0, 0, 0)
        qbar = C.index(tuple(-x for x in C[q]))
        assert out[edge + (qbar,)] == 1.0 and out.sum() == 1.0


def test_bgk_relaxes_nonequilibrium(po):
    f = rng_pops((5, 4), 9, seed=3)
    feq = bgk_equilibrium_2d(po, f)
    om = 1.3
    g = f.copy(order="F")
    po.call("collide_2d", g, None, 5, 4, om)
    assert np.allclose(g - feq, (1 - om) * (f - feq), atol=1e-16)


@pytest.mark.parametrize("Q", [9, 19])
def test_trt_relaxes_even_and_odd_modes(po, Q):
    """s⁺ relaxes (f_q + f_q̄)/2, s⁻ relaxes (f_q − f_q̄)/2 (fused_trt_2d.jl:79-98)."""
    C = C9 if Q == 9 else C19
    opp = [C.index(tuple(-x for x in c)) for c in C]
    shape = (5, 4) if Q == 9 else (4, 3, 3)
    f = rng_pops(shape, Q, seed=4)
    feq = f.copy(order="F")
    if Q == 9:
        po.call("collide_2d", feq, None, *shape, 1.0)
    else:
        po.call("collide_3d", feq, None, *shape, 1.0)
    sp, sm = 1.6, 0.4
    g = f.copy(order="F")
    if Q == 9:
        po.call("collide_trt_2d", g, None, *shape, sp, sm)
    else:
        po.call("collide_trt_3d", g, None, *shape, sp, sm)
    even = lambda a: 0.5 * (a + a[..., opp])
    odd = lambda a: 0.5 * (a - a[..., opp])
    assert np.allclose(even(g) - even(feq), (1 - sp) * (even(f) - even(feq)), atol=2e-16)
    assert np.allclose(odd(g) - odd(feq), (1 - sm) * (odd(f) - odd(feq)), atol=2e-16)


def test_trt_equals_bgk_at_equal_rates(po):
    """Λ = (1/s − ½)² makes s⁻ = s⁺ (test/test_fused_trt_2d.jl:226-258)."""
    nu = 0.04
    s = 1 / (3 * nu + 0.5)
    Lam = (1 / s - 0.5) ** 2
    sp, sm = 1.0 / (3 * nu + 0.5), 1.0 / (Lam / (3 * nu) + 0.5)
    assert abs(sp - sm) < 1e-14
    f = rng_pops((8, 8), 9, seed=5)
    a, b = f.copy(order="F"), f.copy(order="F")
    tmp = np.empty_like(f, order="F")
    for _ in range(50):
        po.call("stream_2d", tmp, a, 8, 8)
        po.call("collide_2d", tmp, None, 8, 8, s)
        a, tmp = tmp, a
    tmp2 = np.empty_like(f, order="F")
    for _ in range(50):
        po.call("stream_2d", tmp2, b, 8, 8)
        po.call("collide_trt_2d", tmp2, None, 8, 8, sp, sm)
        b, tmp2 = tmp2, b
    assert np.max(np.abs(a - b)) < 1e-12


def mrt_matrix():
    """Lallemand–Luo (2000) moment basis from its polynomial definitions in c."""
    M = np.zeros((9, 9))
    for q, (cx, cy) in enumerate(C9):
        c2 = cx * cx + cy * cy
        M[:, q] = [1, -4 + 3 * c2, 4 - 10.5 * c2 + 4.5 * c2 * c2, cx, (-5 + 3 * c2) * cx, cy,
                   (-5 + 3 * c2) * cy, cx * cx - cy * cy, cx * cy]
    return M


def test_mrt_moment_space_relaxation(po):
    M = mrt_matrix()
    f = rng_pops((5, 4), 9, seed=6)
    se, seps, sq, snu = 1.4, 1.4, 1.2, 1.1
    g = f.copy(order="F")
    po.call("collide_mrt_2d", g, None, 5, 4, se, seps, sq, snu)
    m = f @ M.T
    ms = g @ M.T
    rho, jx, jy = m[..., 0], m[..., 3], m[..., 5]
    ux, uy = jx / rho, jy / rho
    usq = ux * ux + uy * uy
    meq = np.stack([rho, -2 * rho + 3 * rho * usq, rho - 3 * rho * usq, jx, -rho * ux, jy,
                    -rho * uy, rho * (ux * ux - uy * uy), rho * ux * uy], -1)
    S = np.array([0, se, seps, 0, sq, 0, sq, snu, snu])
    assert np.allclose(ms, m - S * (m - meq), atol=3e-15)


def test_trt_rates_known_answer(kmi):
    """trt_rates(1/6) = (1, 8/7) (test/test_fused_trt_2d.jl:206-215)."""
    sp, sm = kmi.trt_rates(1 / 6)
    assert abs(sp - 1.0) < 1e-15 and abs(sm - 8 / 7) < 1e-15


def _node_moments(f):
    c = np.array(C9, dtype=float)
    rho = f.sum(-1)
    return rho, (f @ c) / rho[..., None]


def test_zou_he_velocity_closures_impose_velocity(po):
    f = rng_pops((6, 5), 9, seed=7)
    u = 0.07
    g = f.copy(order="F")
    po.call("zou_he_north_2d", g, u, 6, 5)
    rho, vel = _node_moments(g[:, -1, :])
    assert np.allclose(vel[:, 0], u, atol=1e-15) and np.allclose(vel[:, 1], 0.0, atol=1e-15)
    g = f.copy(order="F")
    po.call("zou_he_south_2d", g, u, 6, 5)
    rho, vel = _node_moments(g[:, 0, :])
    assert np.allclose(vel[:, 0], u, atol=1e-15) and np.allclose(vel[:, 1], 0.0, atol=1e-15)
    g = f.copy(order="F")
    po.call("zou_he_west_2d", g, u, 6, 5)
    rho, vel = _node_moments(g[0, :, :])
    assert np.allclose(vel[:, 0], u, atol=1e-15) and np.allclose(vel[:, 1], 0.0, atol=1e-15)


@pytest.mark.parametrize("face", ["east", "west"])
def test_zou_he_pressure_closures_impose_density(po, face):
    f = rng_pops((6, 5), 9, seed=8)
    rho0 = 1.013
    g = f.copy(order="F")
    po.call(f"zou_he_pressure_{face}_2d", g, rho0, 6, 5)
    node = g[-1] if face == "east" else g[0]
    rho, vel = _node_moments(node)
    assert np.allclose(rho, rho0, atol=1e-15) and np.allclose(vel[:, 1], 0.0, atol=1e-15)


def test_ladd_moving_wall_term(po, kmi):
    """Ladd: the reflected edge population gains 6 w ρ_w (c·u_w) (stream_periodic_3d.jl:97-104)."""
    nx, ny, nz = 4, 5, 3
    f = rng_pops((nx, ny, nz), 19, seed=9)
    Ub, Ut = 0.03, -0.02
    out = np.empty_like(f, order="F")
    po.call("stream_periodic_xz_movingwall_y_3d", out, f, Ub, Ut, 1.0, nx, ny, nz)
    ref = np.empty_like(f, order="F")
    po.call("stream_3d", ref, f, nx, ny, nz, 1, 0, 1)
    d = out - ref
    # only the four xy edge slots at the two walls change, by ±(1/6) U
    assert np.allclose(d[:, 0, :, 7], Ub / 6, atol=1e-17)
    assert np.allclose(d[:, 0, :, 8], -Ub / 6, atol=1e-17)
    assert np.allclose(d[:, -1, :, 9], Ut / 6, atol=1e-17)
    assert np.allclose(d[:, -1, :, 10], -Ut / 6, atol=1e-17)
    d[:, 0, :, 7:9] = 0
    d[:, -1, :, 9:11] = 0
    assert not d.any()


def test_abb_slot_identity(po):
    """g_q = −g*_q̄ + 2 w_q T_w on the three wall slots (test/test_thermal.jl:145-240)."""
    nx, ny = 5, 4
    g = rng_pops((nx, ny), 9, seed=10)
    out = g.copy(order="F")
    dirf = np.array([0, 0, 1, 0], dtype=np.int32)
    Tf = np.array([0.0, 0.0, 0.8, 0.0])
    po.lib().ko_apply_fixed_temp_2d_d(out.ctypes.data_as(__import__("ctypes").c_void_p),
                                      dirf.ctypes.data_as(__import__("ctypes").c_void_p),
                                      Tf.ctypes.data_as(__import__("ctypes").c_void_p),
                                      __import__("ctypes").c_int64(nx), __import__("ctypes").c_int64(ny))
    for q, w in ((2, 1 / 9), (5, 1 / 36), (6, 1 / 36)):
        assert np.allclose(out[:, 0, q], -g[:, 0, q] + 2 * w * 0.8, atol=1e-16)
    mask = np.ones_like(g, dtype=bool)
    mask[:, 0, [2, 5, 6]] = False
    assert np.array_equal(out[mask], g[mask])
