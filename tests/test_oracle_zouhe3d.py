"""D3Q19 Zou–He closures of the oracle (boundary_3d.jl:152-290), pinned by what the
closures are built to satisfy, derived here from the lattice vectors alone:

  velocity face: the node's density is ρ_w = (Σ_{c_n=0} f + 2 Σ_{c_n=out} f)/(1 + s u_n) and its
                 momentum is ρ_w u_wall in all three components;
  pressure face: density ρ_out, normal momentum ρ_out u_n with u_n = −s(1 − (Σ_par + 2Σ_out)/ρ_out),
                 tangential momentum that of the wall-parallel diagonals alone (the closure
                 carries no diagonal correction, boundary_3d.jl:270-286);
  both:          only populations entering through the face change, nodes off the face do not.

A wrong descriptor entry (population index, tangent pair, sign) breaks one of these.  The
chain tests restate test/test_cavity_3d.jl's three cases through the descriptor-driven step.
"""
import numpy as np
import pytest

from test_oracle_lattice import C19, rng_pops

FACES = ["west", "east", "south", "north", "bottom", "top"]
# (normal axis, sign) per face, and the tangent axes in the order of _apply_zou_he_*_3d!.
AXIS = {0: (0, -1), 1: (0, 1), 2: (1, -1), 3: (1, 1), 4: (2, -1), 5: (2, 1)}
TANG = {0: (1, 2), 1: (0, 2), 2: (0, 1)}
C = np.array(C19, dtype=float)


def _face_slice(face, shape):
    ax, s = AXIS[face]
    sl = [slice(None)] * 3
    sl[ax] = 0 if s < 0 else shape[ax] - 1
    return tuple(sl)


def _classes(face):
    ax, s = AXIS[face]
    cn = C[:, ax] * s  # component along the outward normal
    return np.where(cn == 0)[0], np.where(cn > 0)[0], np.where(cn < 0)[0]


@pytest.mark.parametrize("face", range(6))
def test_velocity_closure_imposes_density_and_momentum(po, face):
    shape = (5, 4, 6)
    f = rng_pops(shape, 19, seed=40 + face, amp=0.05)
    f0 = f.copy(order="F")
    ax, s = AXIS[face]
    t1, t2 = TANG[ax]
    r = np.random.default_rng(face)
    u = np.zeros(3)
    u[ax], u[t1], u[t2] = r.uniform(-0.05, 0.05, size=3)
    po.call("zou_he_velocity_3d", f, face, u[ax], u[t1], u[t2], *shape)
    par, out, unk = _classes(face)
    sl = _face_slice(face, shape)
    pre = f0[sl]
    rho_w = (pre[..., par].sum(-1) + 2 * pre[..., out].sum(-1)) / (1 + s * u[ax])
    post = f[sl]
    np.testing.assert_allclose(post.sum(-1), rho_w, rtol=0, atol=2e-15)
    for a in range(3):
        np.testing.assert_allclose(post @ C[:, a], rho_w * u[a], rtol=0, atol=2e-15)
    known = np.setdiff1d(np.arange(19), unk)
    assert np.array_equal(post[..., known], pre[..., known])
    off = np.ones(shape, bool)
    off[sl] = False
    assert np.array_equal(f[off], f0[off])


@pytest.mark.parametrize("face", range(6))
def test_pressure_closure_imposes_density(po, face):
    shape = (4, 6, 5)
    f = rng_pops(shape, 19, seed=60 + face, amp=0.05)
    f0 = f.copy(order="F")
    ax, s = AXIS[face]
    t1, t2 = TANG[ax]
    rho_out = 1.013
    po.call("zou_he_pressure_3d", f, face, rho_out, *shape)
    par, out, unk = _classes(face)
    sl = _face_slice(face, shape)
    pre, post = f0[sl], f[sl]
    un = -s * (1 - (pre[..., par].sum(-1) + 2 * pre[..., out].sum(-1)) / rho_out)
    np.testing.assert_allclose(post.sum(-1), rho_out, rtol=0, atol=2e-15)
    np.testing.assert_allclose(post @ C[:, ax], rho_out * un, rtol=0, atol=2e-15)
    diag_par = [q for q in par if np.count_nonzero(C[q]) == 2]
    for a in (t1, t2):
        np.testing.assert_allclose(post @ C[:, a], pre[..., diag_par] @ C[diag_par, a], rtol=0, atol=2e-15)
    known = np.setdiff1d(np.arange(19), unk)
    assert np.array_equal(post[..., known], pre[..., known])


def _step_chain(po, desc, f, steps):
    po.run(desc, f, None, steps)
    m = po.moments(f)
    return m["rho"], m["ux"], m["uy"], m["uz"]


def test_all_faces_zero_velocity_keeps_rest(kmi, po):
    """test_cavity_3d.jl:47-92: Zou–He u = 0 on all six faces, 500 steps from rest (order
    top, bottom, west, east, south, north): the fluid stays at rest."""
    N = 8
    z = {"kind": "zouhe_velocity", "u": (0.0, 0.0, 0.0)}
    d = kmi.make_desc("d3q19", N, N, N, 0.1, faces={f: dict(z) for f in FACES},
                      bc_order=["top", "bottom", "west", "east", "south", "north"])
    f = po.rest((N, N, N))
    rho, ux, uy, uz = _step_chain(po, d, f, 500)
    assert np.all(np.isfinite(rho))
    assert abs(rho.mean() - 1.0) < 1e-3
    assert max(np.abs(ux).max(), np.abs(uy).max(), np.abs(uz).max()) < 1e-10


def test_pressure_outlet_channel_stays_finite(kmi, po):
    """test_cavity_3d.jl:94-135: west velocity inlet, east pressure outlet, the other faces
    Zou–He u = 0, 200 steps: finite densities."""
    N = 8
    z = {"kind": "zouhe_velocity", "u": (0.0, 0.0, 0.0)}
    faces = {f: dict(z) for f in ("south", "north", "bottom", "top")}
    faces["west"] = {"kind": "zouhe_velocity", "u": (0.01, 0.0, 0.0)}
    faces["east"] = {"kind": "zouhe_pressure", "rho": 1.0}
    d = kmi.make_desc("d3q19", N, N, N, 0.1, faces=faces,
                      bc_order=["west", "east", "top", "bottom", "south", "north"])
    f = po.rest((N, N, N))
    rho, ux, *_ = _step_chain(po, d, f, 200)
    assert np.all(np.isfinite(rho))
    assert ux[N // 2, N // 2, N // 2] > 0


def test_cavity_3d_lid_no_overshoot(kmi, po):
    """test_cavity_3d.jl:7-45 (run_cavity_3d, N = 32, Re = 100, 2000 steps): mean lid u_x
    within 10 % of u_lid, max |u_x| ≤ 1.05 u_lid, mean density within 0.05 of 1."""
    N, u_lid, nu = 32, 0.1, 0.032
    d = kmi.cavity_3d_desc(N, N, N, nu, u_lid)
    f = po.rest((N, N, N))
    rho, ux, uy, uz = _step_chain(po, d, f, 2000)
    assert np.all(np.isfinite(rho)) and np.all(np.isfinite(ux))
    assert abs(rho.mean() - 1.0) < 0.05
    top = ux[1:N - 1, 1:N - 1, N - 1]
    assert abs(top.mean() - u_lid) / u_lid < 0.1
    assert np.abs(ux).max() <= 1.05 * u_lid
