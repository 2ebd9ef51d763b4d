"""The oracle reproduces the reference's known answers (its own test tolerances)."""
import json
import math
import os

import numpy as np

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def test_cavity_re100_matches_ghia(po, kmi):
    """test/test_cavity.jl:6-40: 128², Re = 100, 30000 steps, max centreline error < 0.1."""
    g = json.load(open(os.path.join(GOLDEN, "ghia_re100.json")))
    N, Re, ul = 128, 100.0, 0.1
    nu = ul * N / Re
    d = kmi.cavity_desc(N, nu, ul)
    f = po.rest((N, N))
    po.run(d, f, None, 30000)
    m = po.moments(f)
    u = m["ux"][N // 2 - 1, :] / ul  # i_center = N ÷ 2 (1-based)
    err = max(abs(u[min(max(int(round(y * (N - 1))), 0), N - 1)] - ug) for y, ug in zip(g["y"], g["u"]))
    assert err < g["tolerance"], err


def test_couette_zou_he_exact(po, kmi):
    """test/test_couette.jl:263-285: on-node Zou–He walls give the linear profile exactly."""
    Nx, Ny, uw = 4, 16, 0.05
    d = kmi.make_desc("d2q9", Nx, Ny, 1, 0.1, faces={
        "west": {"kind": "periodic"}, "east": {"kind": "periodic"},
        "south": {"kind": "zouhe_velocity", "u": (uw, 0, 0)}, "north": {"kind": "zouhe_velocity", "u": (0, 0, 0)}},
        bc_order=["south", "north"])
    f = po.rest((Nx, Ny))
    po.run(d, f, None, 20000)
    ux = po.moments(f)["ux"][0]
    exact = uw * (Ny - 1 - np.arange(Ny)) / (Ny - 1)
    assert np.max(np.abs(ux - exact)) < 1e-10


def test_couette_ladd_halfway_exact(po, kmi):
    """Halfway moving-wall bounce-back (stream_periodic_3d.jl:97-104 on D2Q9): walls at
    y = 1/2 and Ny + 1/2, so u_j = U (j − 1/2)/Ny for a moving north wall."""
    Nx, Ny, U = 4, 12, 0.04
    d = kmi.make_desc("d2q9", Nx, Ny, 1, 1 / 6, faces={
        "west": {"kind": "periodic"}, "east": {"kind": "periodic"},
        "south": {"kind": "wall"}, "north": {"kind": "moving_wall", "u": (U, 0, 0)}})
    f = po.rest((Nx, Ny))
    po.run(d, f, None, 8000)
    ux = po.moments(f)["ux"][0]
    exact = U * (np.arange(1, Ny + 1) - 0.5) / Ny
    assert np.max(np.abs(ux - exact)) < 1e-10


def test_taylor_green_2d_decay(po, kmi):
    """test/test_taylor_green.jl:4-38: L2 error of u_x < 5 % after 1000 steps."""
    N, nu, u0, steps = 64, 0.01, 0.01, 1000
    k = 2 * np.pi / N
    x = np.arange(N, dtype=float)
    X, Y = np.meshgrid(x, x, indexing="ij")
    ux = -u0 * np.cos(k * X) * np.sin(k * Y)
    uy = u0 * np.sin(k * X) * np.cos(k * Y)
    rho = 1 - 3 * u0 ** 2 / 4 * (np.cos(2 * k * X) + np.cos(2 * k * Y))
    f = po.equilibrium(rho, ux, uy)
    d = kmi.make_desc("d2q9", N, N, 1, nu, faces={s: {"kind": "periodic"} for s in ("west", "east", "south", "north")})
    po.run(d, f, None, steps)
    res = po.moments(f)["ux"]
    ana = ux * math.exp(-2 * nu * k * k * steps)
    l2 = math.sqrt(((res - ana) ** 2).sum() / (ana ** 2).sum())
    assert l2 < 0.05


def test_taylor_green_3d_low_amplitude_decay(po, kmi):
    """Linear regime: each velocity component of the 3D TGV decays as exp(−3νk²t)."""
    N, nu, U0, steps = 24, 0.02, 1e-4, 300
    k = 2 * np.pi / N
    d = kmi.taylor_green_3d_desc(N, nu, storage="f64")
    rho, ux, uy, uz = kmi.taylor_green_3d_fields(N, U0)
    f = po.equilibrium(rho, ux, uy, uz)
    po.run(d, f, None, steps)
    m = po.moments(f)
    decay = math.exp(-3 * nu * k * k * steps)
    for got, init in ((m["ux"], ux), (m["uy"], uy)):
        l2 = math.sqrt(((got - init * decay) ** 2).sum() / ((init * decay) ** 2).sum())
        assert l2 < 0.01, l2
    assert np.max(np.abs(m["uz"])) < 1e-3 * U0


def test_conduction_nusselt_one(po, kmi):
    """test/test_thermal.jl:48-140: halfway Dirichlet walls, no buoyancy → linear T,
    Nu = 1 to round-off."""
    Nx, Ny = 3, 8
    Ta, Tb = 0.8, 0.2
    d, p = kmi.rayleigh_benard_desc(Nx, Ny, Ra=0.0, Pr=1.0, T_hot=Ta, T_cold=Tb, nu=0.1, alpha=0.1)
    f = po.rest((Nx, Ny))
    g = np.asfortranarray(po.W9[None, None, :] * 0.5 * np.ones((Nx, Ny, 1)))
    po.run(d, f, g, 3000)
    T = po.temperature(g)
    lin = Ta + (Tb - Ta) * (np.arange(1, Ny + 1) - 0.5) / Ny
    assert np.max(np.abs(T - lin[None, :])) < 1e-12
    assert abs(po.nusselt(T, kmi._lib.SOUTH, Ta, Tb, True) - 1) < 1e-10


def test_natural_convection_ra1e3_nusselt(po, kmi):
    """test/test_thermal.jl:388-402: N = 64, Ra = 1e3, Pr = 0.71, 30000 steps,
    |Nu − 1.118|/1.118 < 0.005 (reference measured 1.11851)."""
    N = 64
    d = kmi.natural_convection_desc(N, 1e3, 0.71)
    f = po.rest((N, N))
    T0 = kmi.natural_convection_initial_temperature(N)
    g = np.asfortranarray(po.W9[None, None, :] * T0[:, :, None])
    po.run(d, f, g, 30000)
    Nu = po.nusselt(po.temperature(g), kmi._lib.WEST, 1.0, 0.0, False)
    assert abs(Nu - 1.118) / 1.118 < 0.005, Nu
    assert abs(Nu - 1.11851) < 5e-4, Nu
