"""Cut-link geometry for interpolated bounce-back (host side).

Mirrors the reference's precomputations (src/kernels/li_bb_2d.jl:250-312, 441-470,
src/kernels/li_bb_3d_v2.jl:104-160): dense wall-fraction arrays q_wall[i, j, (k,) q] in
(0, 1] (0 = link not cut) and the solid mask, plus the conversion to the sparse link list
`Lattice.set_cut_links` takes.  Node (i, j, k) sits at (i, j, k) (0-based), as the
reference's (i−1, j−1, k−1).
"""
from __future__ import annotations

import numpy as np

C9 = np.array([(0, 0), (1, 0), (0, 1), (-1, 0), (0, -1), (1, 1), (-1, 1), (-1, -1), (1, -1)])
C19 = np.array([(0, 0, 0), (1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0), (0, 0, 1), (0, 0, -1),
                (1, 1, 0), (-1, 1, 0), (1, -1, 0), (-1, -1, 0), (1, 0, 1), (-1, 0, 1), (1, 0, -1),
                (-1, 0, -1), (0, 1, 1), (0, -1, 1), (0, 1, -1), (0, -1, -1)])


def _q_wall_ball(shape, centre, R, C, FT):
    """Smallest t in (0, 1] with |x_f + t c_q − c|² = R² for links leaving fluid nodes
    towards nodes inside the ball, evaluated in the element type FT as the reference does."""
    FT = np.dtype(FT).type
    D = len(shape)
    grids = np.meshgrid(*[np.arange(n, dtype=FT) for n in shape], indexing="ij")
    d = [grids[a] - FT(centre[a]) for a in range(D)]
    R2 = FT(R) * FT(R)
    r2 = d[0] * d[0]
    for a in range(1, D):
        r2 = r2 + d[a] * d[a]
    solid = r2 <= R2
    q_wall = np.zeros(tuple(shape) + (len(C),), dtype=FT)
    with np.errstate(invalid="ignore", divide="ignore"):
        for q in range(1, len(C)):
            c = [FT(C[q][a]) for a in range(D)]
            dn = [d[a] + c[a] for a in range(D)]
            rn = dn[0] * dn[0]
            for a in range(1, D):
                rn = rn + dn[a] * dn[a]
            a2 = c[0] * c[0]
            b = d[0] * c[0]
            for k in range(1, D):
                a2 = a2 + c[k] * c[k]
                b = b + d[k] * c[k]
            b = FT(2) * b
            cc = r2 - R2
            disc = b * b - FT(4) * a2 * cc
            sd = np.sqrt(np.maximum(disc, FT(0)))
            t1 = (-b - sd) / (FT(2) * a2)
            t2 = (-b + sd) / (FT(2) * a2)
            t = np.where(t1 > FT(0), t1, t2)
            ok = (~solid) & (rn <= R2) & (disc >= FT(0)) & (t > FT(0)) & (t <= FT(1))
            q_wall[..., q] = np.where(ok, t, FT(0))
    return np.asfortranarray(q_wall), np.asfortranarray(solid)


def precompute_q_wall_cylinder(Nx, Ny, cx, cy, R, FT=np.float64):
    """precompute_q_wall_cylinder (li_bb_2d.jl:271-312) → (q_wall (Nx, Ny, 9), is_solid)."""
    return _q_wall_ball((Nx, Ny), (cx, cy), R, C9, FT)


def precompute_q_wall_sphere_3d(Nx, Ny, Nz, cx, cy, cz, R, FT=np.float64):
    """precompute_q_wall_sphere_3d (li_bb_3d_v2.jl:116-160) → (q_wall (Nx, Ny, Nz, 19), is_solid)."""
    return _q_wall_ball((Nx, Ny, Nz), (cx, cy, cz), R, C19, FT)


def wall_velocity_rotating_cylinder(q_wall, cx, cy, omega):
    """wall_velocity_rotating_cylinder (li_bb_2d.jl:453-470): u_w = Ω × (x_w − c) at the
    intersection x_w = x_f + q_w c_q of every cut link; (uw_x, uw_y) dense arrays."""
    FT = q_wall.dtype.type
    Nx, Ny, _ = q_wall.shape
    i = np.arange(Nx, dtype=FT)[:, None, None]
    j = np.arange(Ny, dtype=FT)[None, :, None]
    xw = i + q_wall * C9[:, 0].astype(FT)[None, None, :]
    yw = j + q_wall * C9[:, 1].astype(FT)[None, None, :]
    cut = q_wall != 0
    uwx = np.where(cut, -FT(omega) * (yw - FT(cy)), FT(0))
    uwy = np.where(cut, FT(omega) * (xw - FT(cx)), FT(0))
    return np.asfortranarray(uwx), np.asfortranarray(uwy)


def cut_links_from_dense(q_wall, uw_x=None, uw_y=None, uw_z=None):
    """Sparse link list (cell, q, q_w, u_w) of the nonzero entries of dense arrays in the
    reference layout; cell = i + nx (j + ny k), q 0-based."""
    shape = q_wall.shape[:-1]
    idx = np.nonzero(q_wall > 0)
    cells = np.ravel_multi_index(idx[:-1], shape, order="F").astype(np.int32)
    q = idx[-1].astype(np.int32)
    qw = q_wall[idx].astype(np.float64)
    uw = np.zeros((len(q), 3))
    for a, arr in enumerate((uw_x, uw_y, uw_z)):
        if arr is not None:
            uw[:, a] = arr[idx]
    order = np.lexsort((q, cells))
    return cells[order], q[order], qw[order], uw[order]


# ---- drivers (src/drivers/cylinder_libb.jl:179-256, src/kernels/li_bb_3d_v2.jl:263-365) -------
def _equilibrium_x(Q, u, dtype=np.float64):
    """f_q = feq(ρ = 1, u_x = u, 0, 0) per node, as the drivers initialise (Kraken.equilibrium)."""
    from .lattice import Lattice  # noqa: F401  (keeps the package import order)
    C = C9 if Q == 9 else C19
    W = ([4 / 9] + [1 / 9] * 4 + [1 / 36] * 4) if Q == 9 else ([1 / 3] + [1 / 18] * 6 + [1 / 36] * 12)
    u = np.asarray(u, dtype=dtype)
    out = np.empty(u.shape + (Q,), dtype=dtype, order="F")
    usq = u * u
    for q in range(Q):
        cu = dtype(C[q][0]) * u
        out[..., q] = dtype(W[q]) * (dtype(1) + dtype(3) * cu + dtype(4.5) * cu * cu - dtype(1.5) * usq)
    return out


def run_cylinder_libb_2d(*, Nx=300, Ny=80, cx=None, cy=None, radius=10, u_in=0.04, nu=None,
                         ν=0.04, max_steps=50_000, avg_window=5_000, inlet="parabolic",
                         rho_out=1.0, storage="f64", backend=None):
    """run_cylinder_libb_2d on the device: LI-BB V2 cylinder (Bouzidi pre-phase, TRT Λ = 3/16,
    inert solid), rebuilt Zou–He velocity inlet (west) and pressure outlet (east), halfway
    bounce-back channel walls; Mei drag averaged over the last `avg_window` steps.
    Cd = 2 F_x / (u_ref² D), u_ref = 2/3 u_in for the parabolic inlet, u_in for the uniform one."""
    from .drivers import _backend
    from .lattice import Lattice, make_desc

    nu = ν if nu is None else nu
    cx = Nx // 4 if cx is None else float(cx)
    cy = Ny // 2 if cy is None else float(cy)
    qw, solid = precompute_q_wall_cylinder(Nx, Ny, cx, cy, radius)
    j = np.arange(1, Ny + 1, dtype=np.float64)
    if inlet == "parabolic":
        prof = 4.0 * u_in * (j - 1) * (Ny - j) / float(Ny - 1) ** 2
        u_ref = 2.0 / 3.0 * u_in
    elif inlet == "uniform":
        prof = np.full(Ny, float(u_in))
        u_ref = float(u_in)
    else:
        raise ValueError(f"unknown inlet {inlet!r}; expected 'parabolic' or 'uniform'")
    faces = {"west": {"kind": "zouhe_velocity_rebuild"},
             "east": {"kind": "zouhe_pressure_rebuild", "rho": rho_out},
             "south": {"kind": "wall"}, "north": {"kind": "wall"}}
    d = make_desc("d2q9", Nx, Ny, 1, nu, storage=storage, collision="trt", faces=faces,
                  solid_mode="inert")
    Fx = Fy = 0.0
    n = 0
    with Lattice(d, _backend(backend)) as lat:
        lat.set_mask(solid)
        lat.set_cut_links(*cut_links_from_dense(qw))
        lat.set_face_profile("west", prof)
        lat.upload(_equilibrium_x(9, np.broadcast_to(prof[None, :], (Nx, Ny))))
        lat.step(max_steps - avg_window)
        for _ in range(avg_window):
            lat.step(1)
            F = lat.reduce("drag_mea")
            Fx += F[0]
            Fy += F[1]
            n += 1
        m = lat.moments()
    Fx, Fy = Fx / max(n, 1), Fy / max(n, 1)
    D = 2.0 * radius
    return {"ρ": m["rho"], "ux": m["ux"], "uy": m["uy"], "Cd": 2.0 * Fx / (u_ref ** 2 * D),
            "Cl": 2.0 * Fy / (u_ref ** 2 * D), "Fx": Fx, "Fy": Fy, "q_wall": qw, "is_solid": solid,
            "u_ref": u_ref, "D": D, "inlet": inlet}


def run_sphere_libb_3d(*, Nx=120, Ny=60, Nz=60, cx=None, cy=None, cz=None, radius=8, u_in=0.04,
                       nu=None, ν=0.04, max_steps=20_000, avg_window=5_000, inlet="uniform",
                       rho_out=1.0, storage="f64", backend=None):
    """run_sphere_libb_3d on the device: LI-BB V2 sphere in a duct, rebuilt inlet / outlet,
    halfway bounce-back side walls, halfway-BB drag Σ 2 c_q f_q averaged over the window;
    Cd = 2 F_x / (u_ref² π R²)."""
    from .drivers import _backend
    from .lattice import Lattice, make_desc

    nu = ν if nu is None else nu
    cx = Nx // 4 if cx is None else float(cx)
    cy = Ny // 2 if cy is None else float(cy)
    cz = Nz // 2 if cz is None else float(cz)
    qw, solid = precompute_q_wall_sphere_3d(Nx, Ny, Nz, cx, cy, cz, radius)
    yy = np.arange(Ny, dtype=np.float64)[:, None]
    zz = np.arange(Nz, dtype=np.float64)[None, :]
    Hy, Hz = float(Ny - 1), float(Nz - 1)
    if inlet == "uniform":
        prof, u_ref = np.full((Ny, Nz), float(u_in)), float(u_in)
    elif inlet == "parabolic":
        prof = 16.0 * u_in * yy * (Hy - yy) * zz * (Hz - zz) / (Hy ** 2 * Hz ** 2)
        u_ref = 4.0 / 9.0 * u_in
    elif inlet == "parabolic_y":
        prof = np.broadcast_to(4.0 * u_in * yy * (Hy - yy) / Hy ** 2, (Ny, Nz)).copy()
        u_ref = 2.0 / 3.0 * u_in
    else:
        raise ValueError(f"unknown inlet {inlet!r}; expected 'uniform', 'parabolic' or 'parabolic_y'")
    faces = {"west": {"kind": "zouhe_velocity_rebuild"},
             "east": {"kind": "zouhe_pressure_rebuild", "rho": rho_out}}
    d = make_desc("d3q19", Nx, Ny, Nz, nu, storage=storage, collision="trt", faces=faces,
                  solid_mode="inert")
    F = np.zeros(3)
    n = 0
    with Lattice(d, _backend(backend)) as lat:
        lat.set_mask(solid)
        lat.set_cut_links(*cut_links_from_dense(qw))
        lat.set_face_profile("west", prof)
        lat.upload(_equilibrium_x(19, np.broadcast_to(prof[None, :, :], (Nx, Ny, Nz))))
        lat.step(max_steps - avg_window)
        for _ in range(avg_window):
            lat.step(1)
            F += np.asarray(lat.reduce("drag_halfway"))
            n += 1
        m = lat.moments()
    F /= max(n, 1)
    A = np.pi * float(radius) ** 2
    return {"ρ": m["rho"], "ux": m["ux"], "uy": m["uy"], "uz": m["uz"], "Cd": 2.0 * F[0] / (u_ref ** 2 * A),
            "Fx": F[0], "Fy": F[1], "Fz": F[2], "q_wall": qw, "is_solid": solid, "u_ref": u_ref,
            "D": 2.0 * radius, "A": A}
