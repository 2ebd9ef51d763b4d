"""Interpolated bounce-back (LI-BB V2: Bouzidi pre-phase + TRT, li_bb_2d_v2.jl,
li_bb_3d_v2.jl) in the oracle, pinned by the reference's own acceptance tests:

  geometry:   cut fractions in (0, 1], a link is cut iff it points into the solid, and the
              two links across one wall segment add up to one (test/test_li_bb_2d.jl:6-42);
  rest state: zero flow past a cylinder stays zero to 1e-10 (test_li_bb_2d.jl:44-69);
  Couette:    halfway-BB + TRT Λ = 3/16 is Ginzburg-exact — the linear profile to L2 < 1e-4 at
              q_w = 0.5, 0.3, 0.7 (both branches of the Bouzidi rule, moving-wall term;
              test/test_couette_libb_canary.jl:151-260), in D2Q9 and D3Q19;
  rotation:   a rotating cylinder drives a positive swirl (test_li_bb_2d.jl:71-124);
  drag:       the Mei momentum exchange equals 2 c_q f_q at q_w = 1/2 on a still wall.
"""
import numpy as np
import pytest


def couette_libb(kmi, stencil, Ny, u_top, q_w, nu=0.1, storage="f64", N=8):
    """Plates on rows j = 0 and Ny−1 (solid), fluid rows 1..Ny−2, links into the plates cut at
    q_w, the top plate moving at u_top; periodic in x (and z)."""
    d3 = stencil == "d3q19"
    Nx, Nz = N, (4 if d3 else 1)
    shape = (Nx, Ny, Nz) if d3 else (Nx, Ny)
    Q = 19 if d3 else 9
    C = kmi.libb.C19 if d3 else kmi.libb.C9
    solid = np.zeros(shape, dtype=np.uint8, order="F")
    solid[:, 0] = 1
    solid[:, Ny - 1] = 1
    qw = np.zeros(shape + (Q,), order="F")
    uwx = np.zeros_like(qw)
    for q in range(1, Q):
        if C[q][1] < 0:
            qw[:, 1, ..., q] = q_w
        if C[q][1] > 0:
            qw[:, Ny - 2, ..., q] = q_w
            uwx[:, Ny - 2, ..., q] = u_top
    faces = {"west": {"kind": "periodic"}, "east": {"kind": "periodic"}}
    if d3:
        faces.update(bottom={"kind": "periodic"}, top={"kind": "periodic"})
    desc = kmi.make_desc(stencil, Nx, Ny, Nz, nu, storage=storage, collision="trt", faces=faces,
                         solid_mode="inert")
    return desc, shape, solid, qw, uwx


def couette_errors(ux, Ny, u_top, q_w):
    j = np.arange(1, Ny - 1)
    H = (Ny - 3) + 2 * q_w
    ua = u_top * (j - 1 + q_w) / H
    un = ux[ux.shape[0] // 2, 1:Ny - 1]
    if un.ndim > 1:
        un = un[:, 0]
    e = un - ua
    return np.sqrt((e ** 2).sum() / (ua ** 2).sum()), np.abs(e).max() / u_top


def test_cylinder_geometry(kmi):
    Nx = Ny = 32
    qw, solid = kmi.precompute_q_wall_cylinder(Nx, Ny, 15.5, 15.5, 4.0)
    assert qw.shape == (Nx, Ny, 9) and solid.shape == (Nx, Ny)
    assert not qw[0, 0, 1:].any() and not solid[0, 0] and solid[15, 15]
    cut = qw[qw > 0]
    assert cut.size and np.all((cut > 0) & (cut <= 1))
    C = kmi.libb.C9
    for (i, j, q) in zip(*np.nonzero(qw > 0)):
        assert not solid[i, j] and solid[i + C[q][0], j + C[q][1]]
    for j in range(Ny):
        for i in range(Nx - 1):
            a, b = qw[i, j, 1], qw[i + 1, j, 3]
            if a > 0 and b > 0 and not solid[i, j] and not solid[i + 1, j]:
                assert abs(a + b - 1) < 1e-10


def test_sphere_geometry(kmi):
    qw, solid = kmi.precompute_q_wall_sphere_3d(12, 12, 12, 5.3, 5.7, 6.1, 3.2)
    C = kmi.libb.C19
    idx = np.nonzero(qw > 0)
    assert len(idx[0]) > 50
    for (i, j, k, q) in zip(*idx):
        assert not solid[i, j, k] and solid[i + C[q][0], j + C[q][1], k + C[q][2]]
        assert 0 < qw[i, j, k, q] <= 1
    # every fluid node next to the sphere has its link to it cut
    fluid_next = 0
    for (i, j, k) in zip(*np.nonzero(~solid)):
        for q in range(1, 19):
            n = (i + C[q][0], j + C[q][1], k + C[q][2])
            if all(0 <= n[a] < 12 for a in range(3)) and solid[n]:
                fluid_next += 1
                assert qw[i, j, k, q] > 0
    assert fluid_next == len(idx[0])


def test_zero_flow_stays_zero(kmi, po):
    N = 32
    qw, solid = kmi.precompute_q_wall_cylinder(N, N, 15.5, 15.5, 4.0)
    d = kmi.make_desc("d2q9", N, N, 1, 0.1, collision="trt", solid_mode="inert")
    f = po.rest((N, N))
    po.run_libb(d, f, solid, qw, nsteps=200)
    m = po.moments(f)
    fl = ~solid
    assert np.abs(m["ux"][fl]).max() < 1e-10 and np.abs(m["uy"][fl]).max() < 1e-10
    assert np.abs(m["rho"][fl] - 1).max() < 1e-10


@pytest.mark.parametrize("stencil,q_w,steps", [("d2q9", 0.5, 20000), ("d2q9", 0.3, 10000),
                                                ("d2q9", 0.7, 10000), ("d3q19", 0.5, 6000),
                                                ("d3q19", 0.3, 6000)])
def test_couette_ginzburg_exact(kmi, po, stencil, q_w, steps):
    Ny, u_top = (33 if stencil == "d2q9" else 17), 0.01
    d, shape, solid, qw, uwx = couette_libb(kmi, stencil, Ny, u_top, q_w)
    f = po.rest(shape)
    po.run_libb(d, f, solid, qw, (uwx, None, None), nsteps=steps)
    m = po.moments(f)
    L2, Linf = couette_errors(m["ux"], Ny, u_top, q_w)
    assert L2 < 1e-4 and Linf < 1e-4, (L2, Linf)
    assert np.abs(m["uy"][:, 1:Ny - 1]).max() / u_top < 1e-6


def test_rotating_cylinder_swirl(kmi, po):
    N, cx, R, Om = 48, 23.5, 4.0, 0.002
    qw, solid = kmi.precompute_q_wall_cylinder(N, N, cx, cx, R)
    uwx, uwy = kmi.wall_velocity_rotating_cylinder(qw, cx, cx, Om)
    d = kmi.make_desc("d2q9", N, N, 1, 0.1, collision="trt", solid_mode="inert")
    f = po.rest((N, N))
    po.run_libb(d, f, solid, qw, (uwx, uwy, None), nsteps=2000)
    m = po.moments(f)
    ut = []
    for th in np.arange(0, 2 * np.pi + 1e-9, np.pi / 8):
        i, j = int(round(cx + (R + 2) * np.cos(th))), int(round(cx + (R + 2) * np.sin(th)))
        if 0 <= i < N and 0 <= j < N and not solid[i, j]:
            X, Y = i - cx, j - cx
            ut.append((-Y * m["ux"][i, j] + X * m["uy"][i, j]) / np.hypot(X, Y))
    assert np.all(np.isfinite(m["ux"])) and np.mean(ut) > 0
    assert np.sqrt(m["ux"] ** 2 + m["uy"] ** 2).max() < 0.02


def test_mei_drag_is_halfway_at_half_cut(kmi, po):
    from test_oracle_lattice import rng_pops

    N = 16
    qw, solid = kmi.precompute_q_wall_cylinder(N, N, 7.5, 7.5, 3.0)
    half = np.where(qw > 0, 0.5, 0.0)
    f = rng_pops((N, N), 9, seed=3, amp=0.05)
    d = kmi.make_desc("d2q9", N, N, 1, 0.1)
    mea = po.drag_mei_2d(f, half)
    hw = po.drag_halfway(d, f, half)
    np.testing.assert_allclose(mea, hw[:2], rtol=1e-14, atol=1e-16)


# ---- rebuilt inlet / outlet faces (apply_bc_rebuild_2d!/_3d!, src/bc/rebuild_*.jl) ----------
# Pins: the regularised collision's output depends on ρ, u and the stress Π only (adding the
# D2Q9 ghost modes ε and q_x changes nothing) and keeps equilibria fixed; both local
# collisions conserve momentum (the 3D one mass too, and agrees with the bulk TRT); the
# rebuilt node carries the Zou–He values — the normal momentum ρ_w u_n (and in 3D the mass
# ρ_w), ρ_w from the textbook Zou–He relation on the populations streamed in from inside, and
# zero tangential momentum in 2D (in 3D, the in-face diagonals' share, which the reference's
# closure leaves uncorrected).

def _local(po, f, sp, sm):
    o = np.empty_like(f)
    po.call("trt_collide_local", f, o, len(f), sp, sm)
    return o


def test_regularised_collision_sees_only_the_stress(po):
    r = np.random.default_rng(41)
    W = po.W9
    f = np.ascontiguousarray(W * (1 + 0.05 * r.uniform(-1, 1, 9)))
    sp, sm = 1.7, 0.9
    o = _local(po, f, sp, sm)
    eps = np.array([4, -2, -2, -2, -2, 1, 1, 1, 1], dtype=float)
    qx = np.array([0, -2, 0, 2, 0, 1, -1, -1, 1], dtype=float)
    qy = np.array([0, 0, -2, 0, 2, 1, 1, -1, -1], dtype=float)
    for g in (eps, qx, qy):
        assert abs(g.sum()) + abs((g * po.CX9).sum()) + abs((g * po.CY9).sum()) == 0
        assert (g * po.CX9 * po.CX9).sum() == 0 and (g * po.CY9 * po.CY9).sum() == 0
        assert (g * po.CX9 * po.CY9).sum() == 0
        o2 = _local(po, np.ascontiguousarray(f + 1e-3 * g), sp, sm)
        assert np.max(np.abs(o2 - o)) < 1e-15
    for a in (po.CX9, po.CY9):
        assert abs((o * a).sum() - (f * a).sum()) < 1e-15
    feq = po.equilibrium(np.array([1.02]), np.array([0.03]), np.array([-0.01]))[0]
    assert np.max(np.abs(_local(po, np.ascontiguousarray(feq), sp, sm) - feq)) < 1e-15


def test_local_trt_3d_is_the_bulk_trt(po, kmi):
    r = np.random.default_rng(42)
    f = np.asfortranarray(po.W19 * (1 + 0.05 * r.uniform(-1, 1, 19)))
    sp, sm = 1.6, 0.4
    o = _local(po, np.ascontiguousarray(f), sp, sm)
    g = f.reshape(1, 1, 1, 19).copy(order="F")
    po.call("collide_trt_3d", g, None, 1, 1, 1, sp, sm)
    assert np.max(np.abs(o - g.ravel())) < 1e-15
    assert abs(o.sum() - f.sum()) < 1e-15
    for a in (po.CX19, po.CY19, po.CZ19):
        assert abs((o * a).sum() - (f * a).sum()) < 1e-15


def _rebuild_desc(kmi, stencil, shape, rho_out=1.002):
    faces = {"west": {"kind": "zouhe_velocity_rebuild"},
             "east": {"kind": "zouhe_pressure_rebuild", "rho": rho_out}}
    nz = shape[2] if len(shape) == 3 else 1
    return kmi.make_desc(stencil, shape[0], shape[1], nz, 0.05, collision="trt", faces=faces)


@pytest.mark.parametrize("stencil,shape", [("d2q9", (6, 7)), ("d3q19", (5, 6, 7))])
def test_rebuilt_nodes_carry_the_zou_he_values(po, kmi, stencil, shape):
    d3 = stencil == "d3q19"
    Q = 19 if d3 else 9
    C = np.array([po.CX19, po.CY19, po.CZ19] if d3 else [po.CX9, po.CY9])
    W = po.W19 if d3 else po.W9
    r = np.random.default_rng(43)
    f0 = np.asfortranarray(W * (1 + 0.03 * r.uniform(-1, 1, shape + (Q,))))
    tshape = shape[1:]
    prof = np.asfortranarray(0.03 + 0.01 * r.uniform(-1, 1, tshape))
    d = _rebuild_desc(kmi, stencil, shape)
    f = f0.copy(order="F")
    po.run_profiles(d, f, {"west": (prof, None)}, nsteps=1)
    inner = [range(1, n - 1) for n in shape[1:]]
    for side in ("west", "east"):
        i = 0 if side == "west" else shape[0] - 1
        for t in np.ndindex(*[len(x) for x in inner]):
            node = (i,) + tuple(rg[k] for rg, k in zip(inner, t))
            o = f[node]
            # populations streamed in from inside the domain: f0(x − c_q) for inside sources
            par, out = 0.0, 0.0
            diag = np.zeros(len(shape))  # in-face diagonals (c_x = 0, two tangents nonzero)
            for q in range(Q):
                src = tuple(node[a] - C[a][q] for a in range(len(shape)))
                if 0 <= src[0] < shape[0]:
                    v = f0[src + (q,)]
                    if C[0][q] == 0:
                        par += v
                        if d3 and C[1][q] != 0 and C[2][q] != 0:
                            diag += v * C[:, q]
                    elif (C[0][q] < 0) == (side == "west"):
                        out += v
            # Tangential momentum: zero in D2Q9; in D3Q19 the reference's closure corrects the
            # axis populations only (rebuild_3d.jl:22-29), leaving the in-face diagonals' part.
            for a in range(1, len(shape)):
                assert abs((o * C[a]).sum() - diag[a]) < 1e-15, (side, node, a)
            if side == "west":
                un = prof[node[1:]]
                rho_w = (par + 2 * out) / (1 - un)
                assert abs((o * C[0]).sum() - rho_w * un) < 1e-15, node
                if d3:
                    assert abs(o.sum() - rho_w) < 1e-14
            else:
                rho = d.faces[1].rho
                u = -1 + (par + 2 * out) / rho
                assert abs((o * C[0]).sum() - rho * u) < 1e-15, node
                if d3:
                    assert abs(o.sum() - rho) < 1e-14


def test_rebuilt_channel_keeps_the_inlet_profile(po, kmi):
    """A channel with the rebuilt parabolic inlet and pressure outlet settles on a profile
    close to the inlet's (halfway walls, so within a few per cent off the wall rows)."""
    Nx, Ny, u = 40, 17, 0.04
    j = np.arange(1, Ny + 1)
    prof = 4 * u * (j - 1) * (Ny - j) / (Ny - 1) ** 2
    faces = {"west": {"kind": "zouhe_velocity_rebuild"}, "east": {"kind": "zouhe_pressure_rebuild", "rho": 1.0}}
    d = kmi.make_desc("d2q9", Nx, Ny, 1, 0.1, collision="trt", faces=faces)
    from kraken_mi355x.libb import _equilibrium_x
    f = _equilibrium_x(9, np.broadcast_to(prof[None, :], (Nx, Ny)))
    po.run_profiles(d, f, {"west": (prof, None)}, nsteps=4000)
    m = po.moments(f)
    mid = m["ux"][Nx // 2, 2:Ny - 2]
    assert np.max(np.abs(mid - prof[2:Ny - 2])) / u < 0.06
    assert np.all(np.isfinite(f))
