"""run_simulation and the basic drivers on the GPU against the oracle replay of the same
chain (bit-identical in FP64), plus the reference's own acceptance checks."""
import numpy as np
import pytest

from test_runner import _replay, _setup

pytestmark = pytest.mark.gpu


def test_run_simulation_time_dependent_matches_oracle(kmi, po, backend):
    Nx, Ny = 40, 16
    s = _setup(kmi, [("west", "velocity", {"ux": lambda y, t, Ly: 0.04 * min(t / 50, 1) * 4 * y * (Ly - y) / Ly ** 2,
                                           "uy": lambda t: 1e-3 * np.sin(t / 7)}),
                     ("east", "pressure", {"rho": lambda t: 1.0 + 1e-4 * np.sin(t / 10)}),
                     ("south", "wall", {}), ("north", "velocity", {"ux": 0.01})], Nx=Nx, Ny=Ny, steps=120,
               regions=[kmi.GeometryRegion("obstacle", lambda x, y: (x - 12) ** 2 + (y - 8) ** 2 < 9)],
               initial=kmi.InitialSetup({"ux": lambda x, y: 0.01, "rho": lambda x: 1.0 + 1e-3 * np.cos(x / 5)}))
    print("[runner] time-dependent channel", flush=True)
    seen = []
    res = kmi.run_simulation(s, backend=backend, callback=lambda k, st: seen.append(k), callback_every=40)
    assert seen == [40, 80, 120]
    p = kmi.plan_simulation(s)
    f = po.equilibrium(*p.initial)
    _replay(po, p, s, f)
    m = po.moments(f)
    for k, kk in (("ρ", "rho"), ("ux", "ux"), ("uy", "uy")):
        assert np.array_equal(res[k], m[kk]), k


def test_run_simulation_body_force_matches_oracle(kmi, po, backend):
    Nx, Ny, Fx = 8, 24, 2e-5
    s = _setup(kmi, [("west", "periodic", {}), ("east", "periodic", {})], Nx=Nx, Ny=Ny, nu=0.1, steps=500,
               force={"Fx": Fx, "Fy": -1e-6})
    seen = []
    res = kmi.run_simulation(s, backend=backend, callback=lambda k, st: seen.append((k, st["ux"].max())),
                             callback_every=200)
    assert [k for k, _ in seen] == [200, 400]
    p = kmi.plan_simulation(s)
    f = po.rest((Nx, Ny))
    po.run(p.desc, f, None, s.max_steps)
    m = po.moments_forced(f, Fx, -1e-6)
    assert np.array_equal(res["ux"], m["ux"]) and np.array_equal(res["uy"], m["uy"])


def test_run_poiseuille_2d(kmi, po, backend):
    """test/test_poiseuille.jl on the backend, and bit-identical to the oracle chain."""
    Nx, Ny, nu, Fx = 4, 32, 0.1, 1e-5
    r = kmi.run_poiseuille_2d(Nx=Nx, Ny=Ny, ν=nu, Fx=Fx, max_steps=10000, backend=backend)
    j = np.arange(1, Ny + 1)
    ua = Fx / (2 * nu) * (j - 0.5) * (Ny + 0.5 - j)
    assert np.max(np.abs(r["ux"][1, 1:-1] - ua[1:-1])) / ua.max() < 0.02
    assert abs(r["ρ"].sum() - Nx * Ny) / (Nx * Ny) < 1e-3
    d = kmi.make_desc("d2q9", Nx, Ny, 1, nu, faces={"west": {"kind": "periodic"}, "east": {"kind": "periodic"}},
                      forcing="guo_plain", force=(Fx, 0.0))
    f = po.rest((Nx, Ny))
    po.run(d, f, None, 10000)
    assert np.array_equal(r["ux"], po.moments(f)["ux"])


def test_run_couette_2d(kmi, po, backend):
    Nx, Ny, nu, uw = 4, 32, 0.1, 0.05
    r = kmi.run_couette_2d(Nx=Nx, Ny=Ny, ν=nu, u_wall=uw, max_steps=4000, backend=backend)
    faces = {"west": {"kind": "periodic"}, "east": {"kind": "periodic"},
             "south": {"kind": "zouhe_velocity", "u": (uw, 0, 0)}, "north": {"kind": "zouhe_velocity", "u": (0, 0, 0)}}
    d = kmi.make_desc("d2q9", Nx, Ny, 1, nu, faces=faces, bc_order=["south", "north"])
    f = po.rest((Nx, Ny))
    po.run(d, f, None, 4000)
    assert np.array_equal(r["ux"], po.moments(f)["ux"])


def test_run_taylor_green_2d(kmi, po, backend):
    """test/test_taylor_green.jl: amplitude decays as exp(−2νk²t) within 5 %."""
    N, nu, u0, steps = 64, 0.01, 0.01, 1000
    r = kmi.run_taylor_green_2d(N=N, ν=nu, u0=u0, max_steps=steps, backend=backend)
    k = 2 * np.pi / N
    amp = np.max(np.abs(r["ux"]))
    assert abs(amp / (u0 * np.exp(-2 * nu * k * k * steps)) - 1) < 0.05
    per = {f: {"kind": "periodic"} for f in ("west", "east", "south", "north")}
    f = po.equilibrium(*kmi.taylor_green_2d_fields(N, u0))
    po.run(kmi.make_desc("d2q9", N, N, 1, nu, faces=per), f, None, steps)
    np.testing.assert_allclose(r["ux"], po.moments(f)["ux"], rtol=0, atol=1e-15)
