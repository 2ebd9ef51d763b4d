"""D3Q19 thermal double distribution and 3D Guo forcing in the oracle, pinned by:

  anti-bounce-back: every face alone equals its single-face kernel bit for bit, and every face
                    set equals the geometric rule of test/test_thermal.jl:150-173 (a slot
                    whose link leaves through Dirichlet faces gets −g + 2w·mean(T_crossed)),
                    as the reference tests at test_thermal.jl:209-238;
  conduction:       the 6³ cube with hot W/S/B and cold E/N/T walls is steady, symmetric
                    under the axis swaps and within the wall temperatures
                    (test_thermal.jl:268-302); the differentially heated box at Ra = 0
                    gives |Nu − 1| < 3e-3 and a mid-line deviation < 4e-4 (test_thermal.jl:332-336);
  thermal BGK:      Σg conserved and Σcg relaxed to (1 − ω)Σcg + ω T u (closed form);
  Guo source:       per cell Σf conserved and Σcf grown by exactly F, constant and per-node
                    force (the moments of S_q); a uniform field equals the constant force bit
                    for bit; 3D Poiseuille (test/test_poiseuille_3d.jl, L∞ < 2 %) and the
                    Kolmogorov flow u_x = F0 sin(ky)/(ν k²) of a per-node force;
  chains:           the generic step equals the restated run_natural_convection_3d chain and
                    the Poiseuille 3D chain bit for bit.
"""
import ctypes
import math

import numpy as np
import pytest

from test_oracle_lattice import rng_pops

FACES = ["west", "east", "south", "north", "bottom", "top"]


def expected_fixed_temp(po, g0, temps):
    """Geometric restatement of the Dirichlet rule (test_thermal.jl:150-173)."""
    nx, ny, nz, _ = g0.shape
    dims = (nx, ny, nz)
    C = (po.CX19, po.CY19, po.CZ19)
    g = g0.copy(order="F")
    for q in range(19):
        for idx in np.ndindex(nx, ny, nz):
            crossed = []
            for n, face in enumerate(FACES):
                if face not in temps:
                    continue
                d, low = n // 2, n % 2 == 0
                src = idx[d] - C[d][q]
                if (src < 0) if low else (src > dims[d] - 1):
                    crossed.append(temps[face])
            if crossed:
                g[idx + (q,)] = -g0[idx + (q,)] + 2 * po.W19[q] * sum(crossed) / len(crossed)
    return g


def test_abb_single_faces(po):
    g0 = rng_pops((4, 5, 3), 19, seed=31, amp=0.5)
    temps = dict(west=0.9, east=0.1, south=0.7, north=0.3, bottom=0.6, top=0.2)
    for n, face in enumerate(FACES):
        one = po.apply_fixed_temp_face_3d(g0, n, temps[face])
        multi = po.apply_fixed_temp_3d(g0, {face: temps[face]})
        assert np.array_equal(one, multi)
        assert np.max(np.abs(multi - expected_fixed_temp(po, g0, {face: temps[face]}))) < 1e-15
        assert not np.array_equal(one, g0)


@pytest.mark.parametrize("names", [FACES, ["west", "south", "top"], ["east", "north", "bottom"],
                                   ["south", "bottom", "top"]])
def test_abb_edge_rule(po, names):
    g0 = rng_pops((4, 5, 3), 19, seed=32, amp=0.5)
    allt = dict(west=0.9, east=0.1, south=0.7, north=0.3, bottom=0.6, top=0.2)
    temps = {k: allt[k] for k in names}
    got = po.apply_fixed_temp_3d(g0, temps)
    assert np.max(np.abs(got - expected_fixed_temp(po, g0, temps))) < 1e-15


def thermal_box_desc(kmi, N, temps, omega_T, beta_g=0.0):
    faces = {f: {"kind": "wall"} for f in FACES}
    for f, t in temps.items():
        faces[f] = {"kind": "wall", "thermal": "dirichlet", "T": t}
    return kmi.make_desc("d3q19", N, N, N, 0.1, faces=faces, thermal=True, omega_T=omega_T,
                         beta_g=beta_g, T_ref=0.5)


def test_conduction_cube_symmetric_steady(po, kmi):
    N, Tw, Tc = 6, 0.8, 0.2
    d = thermal_box_desc(kmi, N, dict(west=Tw, south=Tw, bottom=Tw, east=Tc, north=Tc, top=Tc),
                         1.0 / (3 * 0.1 + 0.5))
    f = po.rest((N, N, N))
    g = np.asfortranarray(po.W19[None, None, None, :] * 0.5 * np.ones((N, N, N, 1)))
    po.run(d, f, g, 3000)
    T0 = po.temperature(g)
    po.run(d, f, g, 1)
    T1 = po.temperature(g)
    assert np.max(np.abs(T1 - T0)) < 1e-13
    for perm in ((1, 0, 2), (2, 1, 0), (0, 2, 1)):
        assert np.max(np.abs(T1 - np.transpose(T1, perm))) < 1e-13
    assert T1.min() >= Tc - 1e-13 and T1.max() <= Tw + 1e-13
    assert np.max(np.abs(f - po.rest((N, N, N)))) < 1e-14  # no buoyancy: the flow stays at rest


def test_box_pure_conduction(po, kmi):
    """test_thermal.jl:332-336: N = 12, Ra = 0, 6000 steps (reference measured 1.40e-3 and 1.78e-4)."""
    N = 12
    d, _ = kmi.natural_convection_3d_desc(N, 0.0, 0.71)
    f = po.rest((N, N, N))
    T0 = kmi.natural_convection_3d_initial_temperature(N)
    g = np.asfortranarray(po.W19[None, None, None, :] * T0[..., None])
    po.run(d, f, g, 6000)
    T = po.temperature(g)
    Nu = po.nusselt_3d(T, kmi._lib.WEST, 1.0, 0.0)
    d3 = max(abs(T[i, 5, 5] - (1 - (i + 0.5) / N)) for i in range(N))
    assert abs(Nu - 1) < 3e-3, Nu
    assert d3 < 4e-4, d3


def test_thermal_collision_moments(po):
    N = 3
    g = rng_pops((N, N, N), 19, seed=33, amp=0.3)
    r = np.random.default_rng(34)
    u = [np.asfortranarray(r.uniform(-0.05, 0.05, (N, N, N))) for _ in range(3)]
    om = 1.3
    g2 = g.copy(order="F")
    po.call("collide_thermal_3d", g2, u[0], u[1], u[2], N, N, N, om)
    C = (po.CX19, po.CY19, po.CZ19)
    T = g.sum(-1)
    assert np.max(np.abs(g2.sum(-1) - T)) < 1e-14
    for a in range(3):
        j0 = (g * C[a]).sum(-1)
        j1 = (g2 * C[a]).sum(-1)
        assert np.max(np.abs(j1 - ((1 - om) * j0 + om * T * u[a]))) < 1e-14


def test_guo_3d_source_moments(po):
    N = 3
    f = rng_pops((N, N, N), 19, seed=35, amp=0.1)
    F = (1e-3, -2e-3, 5e-4)
    C = (po.CX19, po.CY19, po.CZ19)
    f2 = f.copy(order="F")
    po.call("collide_guo_3d", f2, None, N, N, N, 1.2, *F)
    assert np.max(np.abs(f2.sum(-1) - f.sum(-1))) < 1e-14
    for a in range(3):
        assert np.max(np.abs((f2 * C[a]).sum(-1) - (f * C[a]).sum(-1) - F[a])) < 1e-15
    # per-node field: the same identity cell by cell; a uniform field is the constant force
    r = np.random.default_rng(36)
    fa = [np.asfortranarray(r.uniform(-1e-3, 1e-3, (N, N, N))) for _ in range(3)]
    f3 = f.copy(order="F")
    po.call("collide_guo_field_3d", f3, None, fa[0], fa[1], fa[2], N, N, N, 1.2)
    for a in range(3):
        assert np.max(np.abs((f3 * C[a]).sum(-1) - (f * C[a]).sum(-1) - fa[a])) < 1e-15
    uni = [np.full((N, N, N), v, order="F") for v in F]
    f4 = f.copy(order="F")
    po.call("collide_guo_field_3d", f4, None, uni[0], uni[1], uni[2], N, N, N, 1.2)
    assert np.array_equal(f4, f2)


def test_guo_2d_field_is_constant_force_when_uniform(po):
    f = rng_pops((5, 4), 9, seed=37, amp=0.1)
    a, b = f.copy(order="F"), f.copy(order="F")
    po.call("collide_guo_2d", a, None, 5, 4, 1.1, 2e-3, -1e-3)
    fx, fy = np.full((5, 4), 2e-3, order="F"), np.full((5, 4), -1e-3, order="F")
    po.call("collide_guo_field_2d", b, None, fx, fy, 5, 4, 1.1)
    assert np.array_equal(a, b)


def poiseuille_3d_desc(kmi, Nx, Ny, Nz, nu, Fx, forcing="guo"):
    faces = {"west": {"kind": "periodic"}, "east": {"kind": "periodic"},
             "south": {"kind": "wall"}, "north": {"kind": "wall"},
             "bottom": {"kind": "periodic"}, "top": {"kind": "periodic"}}
    return kmi.make_desc("d3q19", Nx, Ny, Nz, nu, faces=faces, forcing=forcing, force=(Fx, 0.0, 0.0))


def test_poiseuille_3d(po, kmi):
    """test/test_poiseuille_3d.jl: 4×32×4, ν = 0.1, F_x = 1e-5, 10000 steps, L∞ < 2 %."""
    Nx, Ny, Nz, nu, Fx = 4, 32, 4, 0.1, 1e-5
    d = poiseuille_3d_desc(kmi, Nx, Ny, Nz, nu, Fx)
    f = po.rest((Nx, Ny, Nz))
    po.run(d, f, None, 10000)
    m = po.moments_forced_3d(f, Fx, 0.0, 0.0)
    j = np.arange(1, Ny + 1)
    ua = Fx / (2 * nu) * (j - 0.5) * (Ny + 0.5 - j)
    un = m["ux"][1, :, 1]
    err = np.max(np.abs(un[1:-1] - ua[1:-1])) / ua.max()
    assert err < 0.02, err
    assert abs(m["rho"].sum() - Nx * Ny * Nz) / (Nx * Ny * Nz) < 1e-3


def test_generic_3d_guo_equals_poiseuille_chain(po, kmi):
    Nx, Ny, Nz, nu, Fx = 4, 8, 3, 0.1, 1e-4
    d = poiseuille_3d_desc(kmi, Nx, Ny, Nz, nu, Fx)
    f = rng_pops((Nx, Ny, Nz), 19, seed=38, amp=0.02)
    gen = f.copy(order="F")
    po.run(d, gen, None, 20)
    a, b = f.copy(order="F"), np.empty_like(f, order="F")
    for _ in range(20):
        po.call("stream_3d", b, a, Nx, Ny, Nz, 1, 0, 1)
        po.call("collide_guo_3d", b, None, Nx, Ny, Nz, d.omega, Fx, 0.0, 0.0)
        a, b = b, a
    assert np.array_equal(gen, a)


@pytest.mark.parametrize("stencil", ["d2q9", "d3q19"])
def test_kolmogorov_force_field(po, kmi, stencil):
    """Per-node force F_x = F0 sin(k y) on a periodic box: steady u_x = F0 sin(k y)/(ν k²)."""
    d3 = stencil == "d3q19"
    Nx, Ny, Nz, nu, F0 = 4, 32, 4, 0.1, 1e-5
    shape = (Nx, Ny, Nz) if d3 else (Nx, Ny)
    k = 2 * math.pi / Ny
    y = np.arange(Ny)
    prof = F0 * np.sin(k * y)
    fx = np.asfortranarray(np.broadcast_to(prof[None, :, None] if d3 else prof[None, :], shape))
    zero = np.zeros(shape, order="F")
    faces = {n: {"kind": "periodic"} for n in (FACES if d3 else FACES[:4])}
    d = kmi.make_desc(stencil, Nx, Ny, Nz if d3 else 1, nu, faces=faces, forcing="guo_field")
    f = po.rest(shape)
    po.run_field(d, f, fx, zero, zero, nsteps=6000)
    if d3:
        ux = po.moments_forced_3d(f, field=(fx, zero, zero))["ux"][1, :, 1]
    else:
        ux = po.moments_forced(f, 0.0, 0.0)["ux"][1, :] + fx[1, :] / 2 / f.sum(-1)[1, :]
    ua = prof / (nu * k * k)
    err = np.max(np.abs(ux - ua)) / np.max(np.abs(ua))
    assert err < 0.01, err
    # a uniform field is the constant force, bit for bit, through the generic step
    dc = kmi.make_desc(stencil, Nx, Ny, Nz if d3 else 1, nu, faces=faces, forcing="guo",
                       force=(F0, -F0, F0 / 2))
    fc = rng_pops(shape, 19 if d3 else 9, seed=39, amp=0.02)
    ff = fc.copy(order="F")
    po.run(dc, fc, None, 10)
    u = [np.full(shape, v, order="F") for v in (F0, -F0, F0 / 2)]
    po.run_field(d, ff, u[0], u[1], u[2], nsteps=10)
    assert np.array_equal(fc, ff)


def test_generic_3d_thermal_equals_natural_convection_chain(po, kmi):
    """run_natural_convection_3d (thermal.jl:500-515): stream f and g, ABB west then east,
    T, unforced moments, thermal BGK, Boussinesq collision."""
    N = 6
    d, p = kmi.natural_convection_3d_desc(N, 1e5, 0.71)
    f = rng_pops((N, N, N), 19, seed=40, amp=0.01)
    T0 = kmi.natural_convection_3d_initial_temperature(N)
    g = np.asfortranarray(po.W19[None, None, None, :] * T0[..., None])
    gf, gg = f.copy(order="F"), g.copy(order="F")
    po.run(d, gf, gg, 12)
    fa, fb = f.copy(order="F"), np.empty_like(f, order="F")
    ga, gb = g.copy(order="F"), np.empty_like(g, order="F")
    Temp = np.empty((N, N, N), order="F")
    rho, ux, uy, uz = (np.empty((N, N, N), order="F") for _ in range(4))
    for _ in range(12):
        po.call("stream_3d", fb, fa, N, N, N, 0, 0, 0)
        po.call("stream_3d", gb, ga, N, N, N, 0, 0, 0)
        gb[...] = po.apply_fixed_temp_face_3d(gb, 0, 1.0)
        gb[...] = po.apply_fixed_temp_face_3d(gb, 1, 0.0)
        po.call("compute_temperature_3d", Temp, gb, N, N, N)
        po.call("macroscopic_3d", rho, ux, uy, uz, fb, N, N, N)
        po.call("collide_thermal_3d", gb, ux, uy, uz, N, N, N, d.omega_T)
        po.call("collide_boussinesq_3d", fb, Temp, None, N, N, N, d.omega, p["beta_g"], p["T_ref"])
        fa, fb = fb, fa
        ga, gb = gb, ga
    assert np.array_equal(gf, fa) and np.array_equal(gg, ga)
    assert not np.array_equal(fa, f)
