"""run_simulation's setup translation (host logic) and its chain on the oracle.

The plan (descriptor, mask, initial fields, boundary handlers) is checked against the rules
of the reference's generic 2D loop (src/simulation_runner.jl:279-443, src/bc/handlers.jl);
the planned chain, replayed on the oracle, reproduces the reference's Poiseuille test and a
developed channel profile from a spatial inflow."""
import numpy as np
import pytest


def _setup(kmi, boundaries, *, Nx=48, Ny=16, nu=0.05, steps=100, force=None, regions=(), initial=None,
           collision="bgk", modules=()):
    return kmi.SimulationSetup(
        name="t", lattice="D2Q9", domain=kmi.DomainSetup(Lx=float(Nx), Ly=float(Ny), Nx=Nx, Ny=Ny),
        physics=kmi.PhysicsSetup(params={"nu": nu}, body_force=force or {}),
        boundaries=[kmi.BoundarySetup(*b) for b in boundaries], max_steps=steps,
        initial=initial, regions=list(regions), collision=collision, modules=list(modules))


def _replay(po, plan, setup, f):
    """The planned chain on the oracle, with the same per-step handler refresh."""
    prof = {}
    for face, kind, v in plan.face_values(setup, 0.0, only_time_dependent=False):
        if kind == "profile":
            prof[face] = v
    for s in range(1, setup.max_steps + 1):
        if plan.time_dependent:
            for face, kind, v in plan.face_values(setup, float(s), only_time_dependent=True):
                if kind == "profile":
                    prof[face] = v
                else:
                    plan.desc.faces[{"west": 0, "east": 1}[face]].rho = v["rho"]
        po.run_profiles(plan.desc, f, {k: (a, b) for k, (a, b) in prof.items()}, 1, mask=plan.mask)
    return f


def test_plan_follows_the_reference_rules(kmi):
    L = kmi._lib
    par = lambda x, y, Ly: 4 * 0.05 * y * (Ly - y) / Ly ** 2
    s = _setup(kmi, [("west", "velocity", {"ux": par}), ("east", "pressure", {"rho": 1.0}),
                     ("south", "wall", {}), ("north", "velocity", {"ux": 0.02})])
    p = kmi.plan_simulation(s)
    kinds = [p.desc.faces[i].kind for i in range(4)]
    assert kinds == [L.FACE_ZOUHE_VELOCITY_PROFILE, L.FACE_ZOUHE_PRESSURE, L.FACE_WALL, L.FACE_ZOUHE_VELOCITY]
    assert list(p.desc.bc_order[:p.desc.n_bc_order]) == [L.WEST, L.EAST, L.NORTH]
    assert not p.time_dependent
    (face, kind, (ux, uy)), = p.face_values(s, 0.0, only_time_dependent=False)
    y = np.arange(16) + 0.5
    assert face == "west" and np.allclose(ux, 4 * 0.05 * y * (16 - y) / 256) and not uy.any()
    # a periodic y pair without periodic x streams with walls on every edge
    s = _setup(kmi, [("south", "periodic", {}), ("north", "periodic", {})])
    assert [kmi.plan_simulation(s).desc.faces[i].kind for i in range(4)] == [L.FACE_WALL] * 4
    # time-dependent scalar velocity goes through the per-node route
    s = _setup(kmi, [("west", "velocity", {"ux": lambda t: 0.01 * t})])
    p = kmi.plan_simulation(s)
    assert p.time_dependent and p.desc.faces[0].kind == L.FACE_ZOUHE_VELOCITY_PROFILE
    assert np.allclose(p.face_values(s, 3.0, True)[0][2][0], 0.03)


@pytest.mark.parametrize("bcs,kw,needle", [
    ([("north", "pressure", {"rho": 1.0})], {}, "north"),
    ([("east", "velocity", {"ux": 0.1})], {}, "east"),
    ([("west", "velocity", {"ux": 0.1})], {"collision": "trt"}, "static wall"),
    ([], {"collision": "trt", "force": {"Fx": 1e-5}}, "body force"),
    ([], {"modules": ["thermal"]}, "thermal"),
])
def test_plan_rejects_what_the_path_does_not_run(kmi, bcs, kw, needle):
    with pytest.raises(kmi.ArgumentError, match=needle):
        kmi.plan_simulation(_setup(kmi, bcs, **kw))


def test_poiseuille_setup_on_the_oracle(kmi, po):
    """test/test_poiseuille.jl through a body-force setup (generic runner: forced moments)."""
    Nx, Ny, nu, Fx = 4, 32, 0.1, 1e-5
    s = _setup(kmi, [("west", "periodic", {}), ("east", "periodic", {})], Nx=Nx, Ny=Ny, nu=nu,
               steps=10000, force={"Fx": Fx})
    p = kmi.plan_simulation(s)
    f = po.rest((Nx, Ny))
    po.run(p.desc, f, None, s.max_steps)
    ux = po.moments_forced(f, Fx, 0.0)["ux"][1]
    j = np.arange(1, Ny + 1)
    ua = Fx / (2 * nu) * (j - 0.5) * (Ny + 0.5 - j)
    assert np.max(np.abs(ux[1:-1] - ua[1:-1])) / ua.max() < 0.02


def test_time_dependent_inflow_on_the_oracle(kmi, po):
    """A ramped parabolic inflow (t-dependent profile) with an obstacle and a t-dependent
    outlet density stays finite and carries flow past the obstacle."""
    Nx, Ny = 40, 16
    s = _setup(kmi, [("west", "velocity", {"ux": lambda y, t, Ly: 0.04 * min(t / 50, 1) * 4 * y * (Ly - y) / Ly ** 2}),
                     ("east", "pressure", {"rho": lambda t: 1.0 + 1e-4 * np.sin(t / 10)}),
                     ("south", "wall", {}), ("north", "wall", {})], Nx=Nx, Ny=Ny, steps=300,
               regions=[kmi.GeometryRegion("obstacle", lambda x, y: (x - 12) ** 2 + (y - 8) ** 2 < 9)])
    p = kmi.plan_simulation(s)
    assert p.mask is not None and p.mask.sum() > 10 and p.time_dependent
    f = _replay(po, p, s, po.rest((Nx, Ny)))
    m = po.moments(f)
    assert np.all(np.isfinite(m["ux"])) and m["ux"][30, 4:12].mean() > 0.01
