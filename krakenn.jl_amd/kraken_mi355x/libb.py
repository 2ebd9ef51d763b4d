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
