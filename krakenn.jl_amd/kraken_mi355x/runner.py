"""`run_simulation` for the MI355X backend: the reference's default single-phase 2D loop
(src/simulation_runner.jl:235-443) on the single-pass kernel.

The setup mirrors the parsed `.krk` `SimulationSetup` (src/io/krk/parser.jl:8-199) on the
fields that loop reads.  Expressions are numbers or Python callables taking keyword
arguments; like the reference's `KrakenExpr`, one is *spatial* when it reads `x` or `y` and
*time-dependent* when it reads `t` (src/bc/handlers.jl:40-51).

Per step the reference streams (periodic or halfway bounce-back, simulation_runner.jl's
`_select_streaming_kernel`), applies the boundary handlers in setup order
(handlers.jl:190-219), collides (BGK, Guo BGK with a body force, or fused TRT) and takes the
moments; here all of it is one launch.  Handlers that change in time are refreshed on the
device before every step (`kmi_set_face_profile`, `kmi_set_face`), with the reference's
t = step.  Modules other than the default path (thermal, VOF, refinement, rheology, meshes,
LI-BB walls) raise ArgumentError: they are outside this backend.
"""
from __future__ import annotations

import inspect
import warnings
from dataclasses import dataclass, field
from typing import Any, Callable

import numpy as np

from .drivers import _backend, _storage
from .lattice import ArgumentError, Lattice, make_desc


@dataclass
class DomainSetup:
    Lx: float
    Ly: float
    Nx: int
    Ny: int
    Lz: float = 1.0
    Nz: int = 1


@dataclass
class PhysicsSetup:
    params: dict                                   # :nu, ...
    body_force: dict = field(default_factory=dict)  # "Fx", "Fy" -> expression


@dataclass
class BoundarySetup:
    face: str     # west, east, south, north
    type: str     # wall, velocity, pressure, periodic
    values: dict = field(default_factory=dict)  # "ux", "uy", "rho" -> expression


@dataclass
class InitialSetup:
    fields: dict  # "rho", "ux", "uy" -> expression


@dataclass
class GeometryRegion:
    kind: str                 # obstacle or fluid
    condition: Callable       # (x, y, ...) -> bool


@dataclass
class SimulationSetup:
    name: str
    lattice: str
    domain: DomainSetup
    physics: PhysicsSetup
    boundaries: list
    max_steps: int
    initial: InitialSetup | None = None
    regions: list = field(default_factory=list)
    modules: list = field(default_factory=list)
    collision: str = "bgk"
    wall_bc: str | None = None
    backend: Any = None


def _params(e) -> set:
    if not callable(e):
        return set()
    try:
        return set(inspect.signature(e).parameters)
    except (TypeError, ValueError):
        return set()


def is_spatial(e) -> bool:
    return bool(_params(e) & {"x", "y", "z"})


def is_time_dependent(e) -> bool:
    return "t" in _params(e)


def evaluate(e, **kw):
    """Value of an expression; a callable receives the keywords it declares."""
    if not callable(e):
        return float(e)
    ps = inspect.signature(e).parameters
    if any(p.kind == p.VAR_KEYWORD for p in ps.values()):
        return e(**kw)
    return e(**{k: v for k, v in kw.items() if k in ps})


def _geometry_mask(setup, dx, dy):
    """_apply_geometry! (src/geometry/mask_apply.jl:2-35), expression regions only."""
    if not setup.regions:
        return None
    dom = setup.domain
    has_fluid = any(r.kind == "fluid" for r in setup.regions)
    solid = np.full((dom.Nx, dom.Ny), has_fluid, dtype=bool)
    for r in setup.regions:
        for j in range(dom.Ny):
            for i in range(dom.Nx):
                x, y = (i + 0.5) * dx, (j + 0.5) * dy
                if bool(evaluate(r.condition, x=x, y=y, z=0.0, Lx=dom.Lx, Ly=dom.Ly, dx=dx, dy=dy)):
                    solid[i, j] = r.kind == "obstacle"
    return np.asfortranarray(solid.astype(np.uint8))


def _face_arrays(h, setup, dx, dy, t):
    """_fill_bc_arrays! (handlers.jl:152-187): values at the face's cell centres."""
    dom = setup.domain
    Nx, Ny = dom.Nx, dom.Ny
    if h.face in ("north", "south"):
        y = dy / 2 if h.face == "south" else dom.Ly - dy / 2
        pts = [((i + 0.5) * dx, y) for i in range(Nx)]
    else:
        x = dx / 2 if h.face == "west" else dom.Lx - dx / 2
        pts = [(x, (j + 0.5) * dy) for j in range(Ny)]
    kw = dict(Lx=dom.Lx, Ly=dom.Ly, Nx=float(Nx), Ny=float(Ny), dx=dx, dy=dy, t=t)
    ux = np.array([evaluate(h.values["ux"], x=x, y=y, **kw) if "ux" in h.values else 0.0 for x, y in pts])
    uy = np.array([evaluate(h.values["uy"], x=x, y=y, **kw) if "uy" in h.values else 0.0 for x, y in pts])
    return ux, uy


def _resolve_backend(setup, backend):
    """backend= always wins over the setup's Backend directive (simulation_runner.jl:151-160)."""
    if backend is not None:
        return _backend(backend)
    return _backend(setup.backend)


@dataclass
class SimulationPlan:
    """What one run needs on the device: the descriptor, the solid mask, the initial fields
    and the boundary handlers (setup entry, 'profile' | 'scalar' | 'pressure', time-dependent)."""
    desc: Any
    mask: Any
    initial: Any
    handlers: list
    dx: float
    dy: float

    def face_values(self, setup, t, only_time_dependent):
        """(face, kind, values) updates due at time t: profile arrays or a scalar face."""
        out = []
        for b, kind, td in self.handlers:
            if only_time_dependent and not td:
                continue
            if kind == "profile":
                out.append((b.face, "profile", _face_arrays(b, setup, self.dx, self.dy, t)))
            elif kind == "pressure" and td:
                out.append((b.face, "face", {"kind": "zouhe_pressure", "rho": evaluate(b.values["rho"], t=t)}))
        return out

    @property
    def time_dependent(self):
        return any(td for _, _, td in self.handlers)


def plan_simulation(setup: SimulationSetup, T=np.float64) -> SimulationPlan:
    """Translate the setup into a kmi descriptor (host logic only, no device work)."""
    if setup.modules:
        raise ArgumentError(f"modules {setup.modules} are outside the MI355X backend's scope")
    if str(setup.lattice).lstrip(":").upper() != "D2Q9":
        raise ArgumentError("the MI355X run_simulation path is the default D2Q9 loop")
    if setup.wall_bc not in (None, "bounce_back", "halfway_bb"):
        raise ArgumentError(f"wall_bc={setup.wall_bc} (LI-BB) is outside the MI355X backend's scope")
    dom = setup.domain
    Nx, Ny = dom.Nx, dom.Ny
    dx, dy = dom.Lx / Nx, dom.Ly / Ny
    nu = float(setup.physics.params["nu"])
    collision = str(setup.collision).lstrip(":").lower()

    # streaming selection (_select_streaming_kernel): periodic x needs a periodic west or
    # east face, periodic y a periodic south or north face AND periodic x; otherwise the
    # halfway bounce-back stream on all four edges.
    types = {b.face: b.type for b in setup.boundaries}
    per_x = types.get("west") == "periodic" or types.get("east") == "periodic"
    per_y = per_x and (types.get("south") == "periodic" or types.get("north") == "periodic")
    faces = {f: {"kind": "wall"} for f in ("west", "east", "south", "north")}
    if per_x:
        faces["west"] = faces["east"] = {"kind": "periodic"}
    if per_y:
        faces["south"] = faces["north"] = {"kind": "periodic"}

    has_force = bool(setup.physics.body_force)
    if collision == "trt":
        # _ensure_generic_trt_supported! (simulation_runner.jl:203-216)
        if has_force:
            raise ArgumentError("Physics collision=trt in the generic 2D path currently supports only "
                                "Newtonian flow without body force.")
        bad = [b for b in setup.boundaries if b.type != "wall"]
        if bad:
            raise ArgumentError("Physics collision=trt in the generic 2D path currently supports only "
                                f"static wall boundaries; got {bad[0].face} {bad[0].type}.")

    handlers, order = [], []
    for b in setup.boundaries:
        if b.type in ("periodic", "wall"):
            continue  # walls: the bounce-back of the pull (handlers.jl:222-227)
        if faces[b.face]["kind"] == "periodic":
            continue
        td = any(is_time_dependent(v) for v in b.values.values())
        if b.type == "velocity":
            if b.face == "east":
                raise ArgumentError("a velocity condition on the east face is not applied by the "
                                    "reference's generic 2D route (handlers.jl:228-247)")
            sp = any(is_spatial(b.values.get(k)) or is_time_dependent(b.values.get(k)) for k in ("ux", "uy"))
            if sp:
                faces[b.face] = {"kind": "zouhe_velocity_profile"}
            else:
                ux = evaluate(b.values.get("ux", 0.0), t=0.0)
                if "uy" in b.values and evaluate(b.values["uy"], t=0.0) != 0.0:
                    warnings.warn(f"{b.face}: the scalar Zou–He route imposes u_x only; u_y is "
                                  "ignored as in the reference (handlers.jl:228-237)")
                faces[b.face] = {"kind": "zouhe_velocity", "u": (ux, 0.0, 0.0)}
            handlers.append((b, "profile" if sp else "scalar", td))
        elif b.type == "pressure":
            if b.face not in ("west", "east"):
                raise ArgumentError(f"Pressure boundary condition on face ':{b.face}' is not supported "
                                    "by the generic 2D runner. Supported pressure faces: :west, :east.")
            rho = evaluate(b.values["rho"], t=0.0) if "rho" in b.values else 1.0
            faces[b.face] = {"kind": "zouhe_pressure", "rho": rho}
            handlers.append((b, "pressure", td))
        else:
            raise ArgumentError(f"unknown boundary type {b.type!r}")
        order.append(b.face)

    force = (0.0, 0.0)
    if has_force:
        force = (evaluate(setup.physics.body_force.get("Fx", 0.0)),
                 evaluate(setup.physics.body_force.get("Fy", 0.0)))
    desc = make_desc("d2q9", Nx, Ny, 1, nu, storage=_storage(T), collision=collision, faces=faces,
                     bc_order=order or None, forcing="guo" if has_force else None, force=force)
    initial = None
    if setup.initial is not None:
        # _apply_initial_conditions! (simulation_runner.jl:785-810): cell centres
        x = (np.arange(Nx) + 0.5) * dx
        y = (np.arange(Ny) + 0.5) * dy
        kw = dict(Lx=dom.Lx, Ly=dom.Ly, Nx=float(Nx), Ny=float(Ny), dx=dx, dy=dy)
        initial = []
        for name, dflt in (("rho", 1.0), ("ux", 0.0), ("uy", 0.0)):
            e = setup.initial.fields.get(name)
            initial.append(np.array([[dflt if e is None else evaluate(e, x=xi, y=yj, **kw) for yj in y]
                                     for xi in x]))
    return SimulationPlan(desc, _geometry_mask(setup, dx, dy), initial, handlers, dx, dy)


def _apply(lat, updates):
    for face, kind, v in updates:
        if kind == "profile":
            lat.set_face_profile(face, *v)
        else:
            lat.set_face(face, v)


def _fields(lat, plan, setup):
    """The moment planes the reference's loop hands back.  The default path computes them
    from the post-collision populations (compute_macroscopic_2d!, simulation_runner.jl:404-408),
    which is kmi_moments.  With collision = trt the fused kernel writes its pre-collision
    moments instead, and ρ = 1, u = 0 at solid cells (fused_trt_2d.jl:53-66); a collision
    conserves ρ and ρu, so at fluid cells the two differ by rounding only, and the solid
    cells are set here."""
    m = lat.moments()
    if str(setup.collision).lstrip(":").lower() == "trt" and plan.mask is not None:
        solid = np.asarray(plan.mask, dtype=bool)
        m["rho"][solid], m["ux"][solid], m["uy"][solid] = 1.0, 0.0, 0.0
    return m


def run_simulation(setup: SimulationSetup, *, backend=None, T=np.float64,
                   callback: Callable | None = None, callback_every: int = 100):
    """Default single-phase D2Q9 path of run_simulation (simulation_runner.jl:279-443).
    Returns {"ρ", "ux", "uy", "setup"}; `callback(step, {"rho", "ux", "uy"})` every
    `callback_every` steps."""
    be = _resolve_backend(setup, backend)
    plan = plan_simulation(setup, T)
    with Lattice(plan.desc, be) as lat:
        if plan.mask is not None:
            lat.set_mask(plan.mask)
        if plan.initial is not None:
            lat.init_equilibrium(*plan.initial)
        _apply(lat, plan.face_values(setup, 0.0, only_time_dependent=False))
        step = 0
        while step < setup.max_steps:
            if plan.time_dependent:
                # _apply_boundary_conditions! at step s uses t = s (handlers.jl:197-208)
                _apply(lat, plan.face_values(setup, float(step + 1), only_time_dependent=True))
                n = 1
            else:
                n = setup.max_steps - step
                if callback is not None:
                    n = min(n, callback_every - step % callback_every)
            lat.step(n)
            step += n
            if callback is not None and step % callback_every == 0:
                m = _fields(lat, plan, setup)
                callback(step, {"rho": m["rho"], "ux": m["ux"], "uy": m["uy"]})
        m = _fields(lat, plan, setup)
    return {"ρ": m["rho"], "ux": m["ux"], "uy": m["uy"], "setup": setup}
