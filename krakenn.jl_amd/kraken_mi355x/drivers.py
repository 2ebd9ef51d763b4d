"""Host drivers mirroring Kraken.jl's hot-path API on the MI355X backend.

Names, arguments and return values follow the reference (src/drivers/basic.jl,
src/drivers/thermal.jl, src/kernels/fused_trt_2d.jl); every time step runs in the
single-pass HIP kernel of libkraken_mi355x.so.
"""
from __future__ import annotations

import math
import time
from dataclasses import dataclass

import numpy as np

from . import _lib as L
from .lattice import ArgumentError, Lattice, MI355XBackend, make_desc

W9 = np.array([4 / 9, 1 / 9, 1 / 9, 1 / 9, 1 / 9, 1 / 36, 1 / 36, 1 / 36, 1 / 36])


class D2Q9:
    """Lattice marker (src/lattice/d2q9.jl:17)."""


class D3Q19:
    """Lattice marker (src/lattice/d3q19.jl:11)."""


@dataclass
class LBMConfig:
    """LBMConfig of src/drivers/basic.jl:6-23."""

    lattice: object
    Nx: int
    Ny: int
    Nz: int = 1
    nu: float = 0.1
    u_lid: float = 0.1
    max_steps: int = 10000
    output_interval: int = 1000


def lbm_config(lattice, *, Nx, Ny, Nz=1, nu=None, ν=None, u_lid=0.1, max_steps=10000,
               output_interval=1000) -> LBMConfig:
    nu = ν if nu is None else nu
    if nu is None:
        raise ArgumentError("LBMConfig needs a viscosity nu")
    return LBMConfig(lattice, Nx, Ny, 1 if isinstance(lattice, D2Q9) else Nz, float(nu), float(u_lid),
                     int(max_steps), int(output_interval))


def omega(config: LBMConfig) -> float:
    """ω = 1/(3ν + 1/2) (basic.jl:30)."""
    return 1.0 / (3.0 * config.nu + 0.5)


def reynolds(config: LBMConfig) -> float:
    """Re = u_lid·N/ν with N = Ny in 2D, Nz in 3D (basic.jl:37-40)."""
    n = config.Ny if isinstance(config.lattice, D2Q9) else config.Nz
    return config.u_lid * n / config.nu


def trt_rates(nu: float, Lambda: float = 3.0 / 16.0):
    """(s_plus, s_minus) (fused_trt_2d.jl:129-133)."""
    return L.trt_rates(nu, Lambda)


def _storage(T) -> str:
    if isinstance(T, str):
        return T
    if T in (np.float64, float):
        return "f64"
    if T == np.float32:
        return "f32"
    raise ArgumentError(f"unsupported element type {T!r}")


def _backend(backend) -> MI355XBackend:
    if backend is None or backend in ("mi355x", ":mi355x"):
        return MI355XBackend()
    if isinstance(backend, MI355XBackend):
        return backend
    raise ArgumentError(f"unknown backend {backend!r}; this package provides MI355XBackend() only")


def cavity_desc(N, nu, u_lid, *, storage="f64", collision="bgk", lid="zouhe", Ny=None):
    """Lid-driven cavity: halfway-BB walls, lid on the north face (basic.jl:121-136).
    lid='zouhe' is the reference's Zou–He lid (boundary_2d.jl:13-44); lid='ladd' is the
    moving-wall bounce-back of stream_periodic_3d.jl:97-166 restated on D2Q9."""
    Ny = N if Ny is None else Ny
    top = {"kind": "zouhe_velocity" if lid == "zouhe" else "moving_wall", "u": (u_lid, 0.0, 0.0)}
    return make_desc("d2q9", N, Ny, 1, nu, storage=storage, collision=collision,
                     faces={"west": {"kind": "wall"}, "east": {"kind": "wall"},
                            "south": {"kind": "wall"}, "north": top})


def _tune(lat, steps):
    """Pick a D2Q9 lattice's launch (single-step shape, or several steps per launch) for a
    long run: small 2D lattices are bound by the fixed cost of a launch, and the ~30 trials
    of the tuner (a few thousand steps) pay off over tens of thousands.  The populations are
    restored and every choice gives bit-identical results.  D3Q19 runs keep the default
    launch here (their tuner times ~130 variants at full size)."""
    if lat.Q == 9 and steps >= 2000:
        lat.autotune(64)


def run_cavity_2d(config: LBMConfig, *, backend=None, T=np.float64, collision="bgk", lid="zouhe"):
    """run_cavity_2d (basic.jl:112-144): rest start, `max_steps` steps, final moments."""
    if not isinstance(config.lattice, D2Q9):
        raise ArgumentError("run_cavity_2d needs a D2Q9 config")
    d = cavity_desc(config.Nx, config.nu, config.u_lid, storage=_storage(T), collision=collision,
                    lid=lid, Ny=config.Ny)
    with Lattice(d, _backend(backend)) as lat:
        _tune(lat, config.max_steps)
        lat.step(config.max_steps)
        m = lat.moments()
    return {"ρ": m["rho"], "ux": m["ux"], "uy": m["uy"], "config": config}


def cavity_3d_desc(Nx, Ny, Nz, nu, u_lid, *, storage="f64", collision="bgk"):
    """Lid-driven cube of run_cavity_3d (basic.jl:153-175): halfway-BB walls from the pull,
    Zou–He velocity lid (u_lid, 0, 0) on the top face (boundary_3d.jl:317-320)."""
    walls = {f: {"kind": "wall"} for f in ("west", "east", "south", "north", "bottom")}
    walls["top"] = {"kind": "zouhe_velocity", "u": (u_lid, 0.0, 0.0)}
    return make_desc("d3q19", Nx, Ny, Nz, nu, storage=storage, collision=collision, faces=walls)


def run_cavity_3d(config: LBMConfig, *, backend=None, T=np.float64, collision="bgk"):
    """run_cavity_3d (basic.jl:153-175): rest start, `max_steps` steps, final moments."""
    if not isinstance(config.lattice, D3Q19):
        raise ArgumentError("run_cavity_3d needs a D3Q19 config")
    d = cavity_3d_desc(config.Nx, config.Ny, config.Nz, config.nu, config.u_lid,
                       storage=_storage(T), collision=collision)
    with Lattice(d, _backend(backend)) as lat:
        _tune(lat, config.max_steps)
        lat.step(config.max_steps)
        m = lat.moments()
    return {"ρ": m["rho"], "ux": m["ux"], "uy": m["uy"], "uz": m["uz"], "config": config}


def run_poiseuille_2d(*, Nx=4, Ny=32, nu=None, ν=0.1, Fx=1e-5, max_steps=10000, backend=None,
                      T=np.float64):
    """run_poiseuille_2d (basic.jl:185-204): periodic x, halfway-BB walls in y, BGK with the
    Guo force (F_x, 0); moments as compute_macroscopic_2d! (no F/2 shift)."""
    nu = ν if nu is None else nu
    per = {"west": {"kind": "periodic"}, "east": {"kind": "periodic"}}
    d = make_desc("d2q9", Nx, Ny, 1, nu, storage=_storage(T), faces=per, forcing="guo_plain",
                  force=(Fx, 0.0))
    with Lattice(d, _backend(backend)) as lat:
        _tune(lat, max_steps)
        lat.step(max_steps)
        m = lat.moments()
    return {"ρ": m["rho"], "ux": m["ux"], "uy": m["uy"],
            "config": lbm_config(D2Q9(), Nx=Nx, Ny=Ny, nu=nu, u_lid=0.0, max_steps=max_steps)}


def run_couette_2d(*, Nx=4, Ny=32, nu=None, ν=0.1, u_wall=0.05, max_steps=10000, backend=None,
                   T=np.float64):
    """run_couette_2d (basic.jl:213-233): periodic x, Zou–He south wall moving at u_wall
    then Zou–He north wall at rest, BGK."""
    nu = ν if nu is None else nu
    faces = {"west": {"kind": "periodic"}, "east": {"kind": "periodic"},
             "south": {"kind": "zouhe_velocity", "u": (u_wall, 0.0, 0.0)},
             "north": {"kind": "zouhe_velocity", "u": (0.0, 0.0, 0.0)}}
    d = make_desc("d2q9", Nx, Ny, 1, nu, storage=_storage(T), faces=faces, bc_order=["south", "north"])
    with Lattice(d, _backend(backend)) as lat:
        _tune(lat, max_steps)
        lat.step(max_steps)
        m = lat.moments()
    return {"ρ": m["rho"], "ux": m["ux"], "uy": m["uy"],
            "config": lbm_config(D2Q9(), Nx=Nx, Ny=Ny, nu=nu, u_lid=0.0, max_steps=max_steps)}


def taylor_green_2d_fields(N, u0, T=np.float64):
    """Initial (ρ, u_x, u_y) of initialize_taylor_green_2d (basic.jl:242-274), evaluated in
    the element type T at nodes x = i − 1: u_x = −u0 cos kx sin ky, u_y = u0 sin kx cos ky,
    ρ = 1 − 3u0²/4 (cos 2kx + cos 2ky), k = 2π/N."""
    T = np.dtype(T).type
    k = T(2 * np.pi / N)
    x = np.arange(N, dtype=T)[:, None] * np.ones((1, N), dtype=T)
    y = np.ones((N, 1), dtype=T) * np.arange(N, dtype=T)[None, :]
    u0 = T(u0)
    ux = -u0 * np.cos(k * x) * np.sin(k * y)
    uy = u0 * np.sin(k * x) * np.cos(k * y)
    rho = T(1) - T(3) * u0 ** 2 / T(4) * (np.cos(T(2) * k * x) + np.cos(T(2) * k * y))
    return rho, ux, uy


def run_taylor_green_2d(*, N=64, nu=None, ν=0.01, u0=0.01, max_steps=1000, backend=None,
                        T=np.float64):
    """run_taylor_green_2d (basic.jl:281-300): fully periodic BGK decay from the analytic
    vortex; returns the final moments with `k` and `u0`."""
    nu = ν if nu is None else nu
    per = {f: {"kind": "periodic"} for f in ("west", "east", "south", "north")}
    d = make_desc("d2q9", N, N, 1, nu, storage=_storage(T), faces=per)
    with Lattice(d, _backend(backend)) as lat:
        lat.init_equilibrium(*taylor_green_2d_fields(N, u0, T))
        _tune(lat, max_steps)
        lat.step(max_steps)
        m = lat.moments()
    return {"ρ": m["rho"], "ux": m["ux"], "uy": m["uy"], "u0": u0, "k": 2 * np.pi / N,
            "max_steps": max_steps,
            "config": lbm_config(D2Q9(), Nx=N, Ny=N, nu=nu, u_lid=0.0, max_steps=0)}


def taylor_green_3d_fields(N, U0):
    """u = U0 sin x cos y cos z, v = −U0 cos x sin y cos z, w = 0, ρ = 1, x = 2π(i−1)/N
    (node convention of basic.jl:259-262)."""
    x = 2 * np.pi * np.arange(N) / N
    X, Y, Z = np.meshgrid(x, x, x, indexing="ij")
    ux = U0 * np.sin(X) * np.cos(Y) * np.cos(Z)
    uy = -U0 * np.cos(X) * np.sin(Y) * np.cos(Z)
    uz = np.zeros_like(ux)
    return np.ones_like(ux), ux, uy, uz


def taylor_green_3d_desc(N, nu, storage="f32", Lambda=3.0 / 16.0):
    """Triply periodic D3Q19 TRT (PullHalfwayBB_3D with periodic wrap + CollideTRTDirect_3D,
    dsl/bricks_3d.jl:64-229)."""
    per = {f: {"kind": "periodic"} for f in ("west", "east", "south", "north", "bottom", "top")}
    return make_desc("d3q19", N, N, N, nu, storage=storage, collision="trt", faces=per, Lambda=Lambda)


def run_taylor_green_3d(*, N=256, nu=0.01, U0=0.05, max_steps=1000, storage="f32", backend=None):
    d = taylor_green_3d_desc(N, nu, storage)
    with Lattice(d, _backend(backend)) as lat:
        lat.init_equilibrium(*taylor_green_3d_fields(N, U0))
        _tune(lat, max_steps)
        lat.step(max_steps)
        m = lat.moments()
    m["max_steps"] = max_steps
    return m


def rb_initial_temperature(Nx, Ny, T_hot, T_cold, seed="sine", noise=1e-3, rng_seed=12345):
    """Linear conduction profile between halfway walls plus a seed (thermal.jl:74-97).
    seed='sine': the reference's 0.01ΔT sin(2πi/Nx) sin(πj/Ny); seed='noise': uniform
    noise of amplitude `noise`·ΔT from numpy's PCG64 with `rng_seed`."""
    dT = T_hot - T_cold
    i = np.arange(1, Nx + 1)[:, None]
    j = np.arange(1, Ny + 1)[None, :]
    T = T_hot - dT * (j - 0.5) / Ny + 0.0 * i
    if seed == "sine":
        T = T + (0.01 * dT) * np.sin(2 * np.pi * i / Nx) * np.sin(np.pi * j / Ny)
    elif seed == "noise":
        rng = np.random.default_rng(rng_seed)
        T = T + noise * dT * rng.uniform(-1.0, 1.0, size=(Nx, Ny))
    return np.asfortranarray(T)


def rayleigh_benard_desc(Nx, Ny, Ra, Pr, T_hot=1.0, T_cold=0.0, nu=None, alpha=None, storage="f64"):
    """Rayleigh–Bénard (thermal.jl:31-141): periodic x, hot south / cold north halfway
    anti-bounce-back, Boussinesq F_y = β_g (T − T_ref), unfused-chain velocity coupling."""
    dT = T_hot - T_cold
    if dT == 0:
        raise ArgumentError("run_rayleigh_benard_2d: T_hot and T_cold are equal")
    H = Ny
    nu = 0.05 if nu is None else float(nu)
    a = nu / Pr if alpha is None else float(alpha)
    beta_g = Ra * nu * a / (dT * H ** 3)
    T_ref = (T_hot + T_cold) / 2
    faces = {"west": {"kind": "periodic"}, "east": {"kind": "periodic"},
             "south": {"kind": "wall", "thermal": "dirichlet", "T": T_hot},
             "north": {"kind": "wall", "thermal": "dirichlet", "T": T_cold}}
    d = make_desc("d2q9", Nx, Ny, 1, nu, storage=storage, collision="bgk", faces=faces, thermal=True,
                  omega_T=1.0 / (3.0 * a + 0.5), beta_g=beta_g, T_ref=T_ref, coupling="split")
    return d, {"nu": nu, "alpha": a, "beta_g": beta_g, "T_ref": T_ref}


def run_rayleigh_benard_2d(*, Nx=128, Ny=32, Ra=2000.0, Pr=1.0, T_hot=1.0, T_cold=0.0,
                           max_steps=20000, nu=None, alpha=None, storage="f64", seed="sine",
                           backend=None):
    d, p = rayleigh_benard_desc(Nx, Ny, Ra, Pr, T_hot, T_cold, nu, alpha, storage)
    with Lattice(d, _backend(backend)) as lat:
        lat.init_temperature(rb_initial_temperature(Nx, Ny, T_hot, T_cold, seed))
        _tune(lat, max_steps)
        lat.step(max_steps)
        m = lat.moments(temperature=True)
        Nu = lat.reduce("nusselt", "south")
    return {"ρ": m["rho"], "ux": m["ux"], "uy": m["uy"], "Temp": m["Temp"], "Nu": Nu, "Ra": Ra,
            "Pr": p["nu"] / p["alpha"], "ν": p["nu"], "α": p["alpha"], "T_hot": T_hot, "T_cold": T_cold}


def natural_convection_desc(N, Ra, Pr, T_hot=1.0, T_cold=0.0, storage="f64"):
    """Differentially heated cavity (thermal.jl:161-238, fused_thermal_2d.jl:178-207):
    hot west / cold east anti-bounce-back, adiabatic south/north, fused-kernel coupling."""
    dT = T_hot - T_cold
    nu = 0.05
    a = nu / Pr
    beta_g = Ra * nu * a / (dT * float(N) ** 3)
    faces = {"west": {"kind": "wall", "thermal": "dirichlet", "T": T_hot},
             "east": {"kind": "wall", "thermal": "dirichlet", "T": T_cold},
             "south": {"kind": "wall"}, "north": {"kind": "wall"}}
    return make_desc("d2q9", N, N, 1, nu, storage=storage, collision="bgk", faces=faces, thermal=True,
                     omega_T=1.0 / (3.0 * a + 0.5), beta_g=beta_g, T_ref=(T_hot + T_cold) / 2,
                     coupling="fused")


def natural_convection_initial_temperature(N, T_hot=1.0, T_cold=0.0):
    """Horizontal linear profile + sine seed (thermal.jl:196-205)."""
    dT = T_hot - T_cold
    i = np.arange(1, N + 1)[:, None]
    j = np.arange(1, N + 1)[None, :]
    T = T_hot - dT * (i - 0.5) / N + (0.01 * dT) * np.sin(2 * np.pi * i / N) * np.sin(np.pi * j / N)
    return np.asfortranarray(T)


def run_natural_convection_2d(*, N=128, Ra=1e4, Pr=0.71, T_hot=1.0, T_cold=0.0, max_steps=50000,
                              storage="f64", backend=None):
    d = natural_convection_desc(N, Ra, Pr, T_hot, T_cold, storage)
    with Lattice(d, _backend(backend)) as lat:
        lat.init_temperature(natural_convection_initial_temperature(N, T_hot, T_cold))
        _tune(lat, max_steps)
        lat.step(max_steps)
        m = lat.moments(temperature=True)
        Nu = lat.reduce("nusselt", "west")
    return {"ρ": m["rho"], "ux": m["ux"], "uy": m["uy"], "Temp": m["Temp"], "Nu": Nu, "Ra": Ra, "Pr": Pr}


def natural_convection_3d_desc(N, Ra, Pr, T_hot=1.0, T_cold=0.0, storage="f64"):
    """Differentially heated cube (thermal.jl:455-538): D3Q19, halfway bounce-back on all six
    faces, hot west / cold east anti-bounce-back, adiabatic elsewhere, Boussinesq
    F_y = β_g (T − T_ref), the unfused chain's velocity coupling."""
    dT = T_hot - T_cold
    nu = 0.05
    a = nu / Pr
    beta_g = Ra * nu * a / (dT * float(N) ** 3)
    faces = {"west": {"kind": "wall", "thermal": "dirichlet", "T": T_hot},
             "east": {"kind": "wall", "thermal": "dirichlet", "T": T_cold},
             "south": {"kind": "wall"}, "north": {"kind": "wall"},
             "bottom": {"kind": "wall"}, "top": {"kind": "wall"}}
    d = make_desc("d3q19", N, N, N, nu, storage=storage, collision="bgk", faces=faces, thermal=True,
                  omega_T=1.0 / (3.0 * a + 0.5), beta_g=beta_g, T_ref=(T_hot + T_cold) / 2,
                  coupling="split")
    return d, {"nu": nu, "alpha": a, "beta_g": beta_g, "T_ref": (T_hot + T_cold) / 2}


def natural_convection_3d_initial_temperature(N, T_hot=1.0, T_cold=0.0):
    """Linear profile between the halfway walls, T = T_hot − ΔT (i − ½)/N (thermal.jl:489-496)."""
    dT = T_hot - T_cold
    i = np.arange(1, N + 1, dtype=np.float64)
    T = T_hot - dT * (i - 0.5) / N
    return np.asfortranarray(np.broadcast_to(T[:, None, None], (N, N, N)))


def run_natural_convection_3d(*, N=32, Ra=1e3, Pr=0.71, T_hot=1.0, T_cold=0.0, max_steps=10000,
                              storage="f64", backend=None):
    """run_natural_convection_3d (thermal.jl:455-538) on the device: every step one launch;
    Nu from the hot-wall estimator averaged over the face without its edge rows."""
    d, p = natural_convection_3d_desc(N, Ra, Pr, T_hot, T_cold, storage)
    with Lattice(d, _backend(backend)) as lat:
        lat.init_temperature(natural_convection_3d_initial_temperature(N, T_hot, T_cold))
        _tune(lat, max_steps)
        lat.step(max_steps)
        m = lat.moments(temperature=True)
        Nu = lat.reduce("nusselt", "west")
    return {"ρ": m["rho"], "ux": m["ux"], "uy": m["uy"], "uz": m["uz"], "Temp": m["Temp"],
            "Nu": Nu, "Ra": Ra, "Pr": Pr, "ν": p["nu"], "α": p["alpha"]}


def run_to_residual(lat: Lattice, tol=1e-7, check_every=100, max_steps=10_000_000, max_wall=math.inf):
    """Step until r_u = ‖u^k − u^(k−1)‖₂/‖u^k‖₂ < tol, checked every `check_every` steps
    (the stop rule of gnf_periodic_steady.jl:114-227).  Returns (steps, r_u, converged)."""
    t0 = time.perf_counter()
    steps, r = 0, math.inf
    while steps < max_steps:
        n = min(check_every, max_steps - steps)
        r = lat.step_residual(n)
        steps += n
        if not math.isfinite(r):
            return steps, r, False
        if r < tol:
            return steps, r, True
        if time.perf_counter() - t0 > max_wall:
            break
    return steps, r, False
