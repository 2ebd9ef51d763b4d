"""ctypes binding of the CPU oracle (oracle/kraken_oracle.cpp).

TEST INFRASTRUCTURE: imported only by tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg, as the checker / CPU baseline.  Arrays are numpy arrays in the
reference layout f[i, j, (k,) q] (Fortran order).
"""
from __future__ import annotations

import ctypes
import os
import subprocess
import sys
from ctypes import POINTER, c_double, c_int, c_int64, c_uint8, c_void_p

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "libkraken_oracle.so")
_lib = None


def build(force=False):
    if force or not os.path.exists(LIB):
        subprocess.run(["make", "-C", HERE, "-B" if force else "all"], check=True,
                       stdout=subprocess.DEVNULL)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        _lib = ctypes.CDLL(LIB)
        _lib.ko_nusselt_d.restype = c_double
        _lib.ko_nusselt_3d_d.restype = c_double
        _lib.ko_ksum_d.restype = c_double
        _lib.ko_omp_threads.restype = c_int
    return _lib


def _p(a):
    return None if a is None else a.ctypes.data_as(c_void_p)


def _sfx(a):
    if a.dtype == np.float64:
        return "d"
    if a.dtype == np.float32:
        return "f"
    raise TypeError(a.dtype)


def _mask(m):
    return None if m is None else np.asfortranarray(m, dtype=np.uint8).ctypes.data_as(c_void_p)


def call(name, *args, dtype=np.float64):
    """Call ko_<name>_<d|f>; numpy arrays pass as pointers, ints as int64, floats as double."""
    sfx = "d" if dtype == np.float64 else "f"
    fn = getattr(lib(), f"ko_{name}_{sfx}")
    cargs = []
    for a in args:
        if a is None:
            cargs.append(None)
        elif isinstance(a, np.ndarray):
            if not a.flags.f_contiguous and a.ndim > 1:
                raise ValueError("oracle arrays must be Fortran-ordered")
            cargs.append(a.ctypes.data_as(c_void_p))
        elif isinstance(a, (bool, np.bool_)):
            cargs.append(c_int(int(a)))
        elif isinstance(a, (int, np.integer)):
            cargs.append(c_int64(int(a)))
        elif isinstance(a, ctypes.Structure):
            cargs.append(ctypes.byref(a))
        else:
            cargs.append(c_double(float(a)))
    fn.restype = None
    return fn(*cargs)


def run(desc, f, g=None, nsteps=1, mask=None):
    """Advance f (and g) `nsteps` steps of the generic descriptor-driven chain in place."""
    sfx = _sfx(f)
    fn = getattr(lib(), f"ko_run_{sfx}")
    fn.restype = None
    m = None if mask is None else np.asfortranarray(mask, dtype=np.uint8)
    fn(ctypes.byref(desc), None if m is None else m.ctypes.data_as(c_void_p), _p(f), _p(g),
       c_int64(int(nsteps)))
    return f, g


def run_profiles(desc, f, profiles, nsteps=1, mask=None):
    """As run(), for D2Q9 lattices with profile faces: `profiles` maps face name -> (a, b)
    arrays (u_x, u_y of a velocity profile, or rho and None of a pressure profile)."""
    names = {"west": 0, "east": 1, "south": 2, "north": 3}
    profiles = {k: (v if isinstance(v, tuple) else (v, None)) for k, v in profiles.items()}
    keep, ptrs = [], (ctypes.c_void_p * 8)()
    for name, (a, b) in profiles.items():
        for k, arr in enumerate((a, b)):
            if arr is None:
                continue
            arr = np.asfortranarray(arr, dtype=np.float64)  # (Ny, Nz) profiles: j + Ny k
            keep.append(arr)
            ptrs[2 * names[name] + k] = arr.ctypes.data
    fn = getattr(lib(), f"ko_run_profiles_{_sfx(f)}")
    fn.restype = None
    m = None if mask is None else np.asfortranarray(mask, dtype=np.uint8)
    fn(ctypes.byref(desc), None if m is None else m.ctypes.data_as(c_void_p), _p(f), ptrs,
       c_int64(int(nsteps)))
    return f


def run_libb(desc, f, solid, q_wall, uw=(None, None, None), nsteps=1, west_profile=None):
    """fused_trt_libb_v2_step! chain (LI-BB pre-phase, inert solids) on the descriptor's faces;
    q_wall and the wall velocities are dense arrays of f's element type and shape.  Rebuilt
    faces (run_cylinder_libb_2d / run_sphere_libb_3d) take the west inlet profile
    `west_profile` (Ny, or (Ny, Nz))."""
    wp = None if west_profile is None else np.asfortranarray(west_profile, dtype=np.float64)
    z = np.zeros_like(q_wall, order="F")
    arrs = [np.asfortranarray(a if a is not None else z, dtype=f.dtype) for a in uw]
    qw = np.asfortranarray(q_wall, dtype=f.dtype)
    m = None if solid is None else np.asfortranarray(solid, dtype=np.uint8)
    fn = getattr(lib(), f"ko_run_libb_{_sfx(f)}")
    fn.restype = None
    fn(ctypes.byref(desc), None if m is None else m.ctypes.data_as(c_void_p), _p(f), _p(qw),
       _p(arrs[0]), _p(arrs[1]), _p(arrs[2]), c_int64(int(nsteps)), _p(wp))
    return f


def drag_mei_2d(f, q_wall, uwx=None, uwy=None):
    z = np.zeros_like(q_wall, order="F")
    out = np.zeros(2)
    nx, ny, _ = f.shape
    fn = getattr(lib(), f"ko_drag_mei_2d_{_sfx(f)}")
    fn.restype = None
    a = [np.asfortranarray(v if v is not None else z, dtype=f.dtype) for v in (q_wall, uwx, uwy)]
    fn(_p(f), _p(a[0]), _p(a[1]), _p(a[2]), c_int64(nx), c_int64(ny), _p(out))
    return out


def drag_halfway(desc, f, q_wall):
    out = np.zeros(3)
    fn = getattr(lib(), f"ko_drag_halfway_{_sfx(f)}")
    fn.restype = None
    qw = np.asfortranarray(q_wall, dtype=f.dtype)
    fn(ctypes.byref(desc), _p(f), _p(qw), _p(out))
    return out


def moments_forced(f, fx, fy):
    """compute_macroscopic_forced_2d of a D2Q9 population array."""
    nx, ny, _ = f.shape
    r, u, v = (np.empty((nx, ny), dtype=f.dtype, order="F") for _ in range(3))
    call("macroscopic_forced_2d", r, u, v, f, nx, ny, float(fx), float(fy), dtype=f.dtype)
    return {"rho": r, "ux": u, "uy": v}


def run_field(desc, f, fx, fy, fz=None, nsteps=1, mask=None):
    """As run(), for a KMI_FORCE_GUO_FIELD descriptor: per-node force arrays (f's element
    type and the lattice shape)."""
    T = f.dtype
    arr = [np.asfortranarray(a, dtype=T) for a in (fx, fy, fz if fz is not None else fx)]
    fn = getattr(lib(), f"ko_run_field_{_sfx(f)}")
    fn.restype = None
    m = None if mask is None else np.asfortranarray(mask, dtype=np.uint8)
    fn(ctypes.byref(desc), None if m is None else m.ctypes.data_as(c_void_p), _p(f), _p(arr[0]),
       _p(arr[1]), _p(arr[2]), c_int64(int(nsteps)))
    return f


def moments_forced_3d(f, fx=0.0, fy=0.0, fz=0.0, field=None):
    """compute_macroscopic_forced_3d! (constant force) or _forced_field_3d! (field = (Fx, Fy, Fz))."""
    nx, ny, nz, _ = f.shape
    r, u, v, w = (np.empty((nx, ny, nz), dtype=f.dtype, order="F") for _ in range(4))
    fa = [None] * 3 if field is None else [np.asfortranarray(a, dtype=f.dtype) for a in field]
    call("macroscopic_forced_3d", r, u, v, w, f, nx, ny, nz, float(fx), float(fy), float(fz),
         fa[0], fa[1], fa[2], dtype=f.dtype)
    return {"rho": r, "ux": u, "uy": v, "uz": w}


def apply_fixed_temp_3d(g, temps):
    """apply_fixed_temp_3d! on a copy of g: temps maps face name -> wall temperature."""
    names = ["west", "east", "south", "north", "bottom", "top"]
    dirs = np.array([int(n in temps) for n in names], dtype=np.int32)
    Tf = np.array([float(temps.get(n, 0.0)) for n in names])
    out = np.array(g, order="F", copy=True)
    nx, ny, nz, _ = g.shape
    fn = getattr(lib(), f"ko_apply_fixed_temp_3d_{_sfx(g)}")
    fn.restype = None
    fn(_p(out), dirs.ctypes.data_as(c_void_p), Tf.ctypes.data_as(c_void_p), c_int64(nx),
       c_int64(ny), c_int64(nz))
    return out


def apply_fixed_temp_face_3d(g, face, Tw):
    """The single-face kernel (apply_fixed_temp_<face>_3d!) on a copy of g."""
    out = np.array(g, order="F", copy=True)
    nx, ny, nz, _ = g.shape
    call("apply_fixed_temp_face_3d", out, int(face), float(Tw), nx, ny, nz, dtype=g.dtype)
    return out


def nusselt_3d(Temp, face, T_hot, T_cold, per_t1=False, per_t2=False):
    nx, ny, nz = Temp.shape
    t = np.asfortranarray(Temp, dtype=np.float64)
    return lib().ko_nusselt_3d_d(t.ctypes.data_as(POINTER(c_double)), c_int64(nx), c_int64(ny),
                                 c_int64(nz), c_int(face), c_double(T_hot), c_double(T_cold),
                                 c_int(int(per_t1)), c_int(int(per_t2)))


def pull(desc, f):
    out = np.empty_like(f, order="F")
    fn = getattr(lib(), f"ko_pull_generic_{_sfx(f)}")
    fn.restype = None
    fn(ctypes.byref(desc), _p(out), _p(f))
    return out


def moments(f):
    """compute_macroscopic_2d/3d of a population array (storage precision)."""
    if f.ndim == 3:
        nx, ny, _ = f.shape
        r, u, v = (np.empty((nx, ny), dtype=f.dtype, order="F") for _ in range(3))
        call("macroscopic_2d", r, u, v, f, nx, ny, dtype=f.dtype)
        return {"rho": r, "ux": u, "uy": v}
    nx, ny, nz, _ = f.shape
    r, u, v, w = (np.empty((nx, ny, nz), dtype=f.dtype, order="F") for _ in range(4))
    call("macroscopic_3d", r, u, v, w, f, nx, ny, nz, dtype=f.dtype)
    return {"rho": r, "ux": u, "uy": v, "uz": w}


def temperature(g):
    if g.ndim == 4:
        nx, ny, nz, _ = g.shape
        t = np.empty((nx, ny, nz), dtype=g.dtype, order="F")
        call("compute_temperature_3d", t, g, nx, ny, nz, dtype=g.dtype)
        return t
    nx, ny, _ = g.shape
    t = np.empty((nx, ny), dtype=g.dtype, order="F")
    call("compute_temperature_2d", t, g, nx, ny, dtype=g.dtype)
    return t


def nusselt(Temp, face, T_hot, T_cold, periodic_parallel):
    nx, ny = Temp.shape
    t = np.asfortranarray(Temp, dtype=np.float64)
    return lib().ko_nusselt_d(t.ctypes.data_as(POINTER(c_double)), c_int64(nx), c_int64(ny),
                              c_int(face), c_double(T_hot), c_double(T_cold), c_int(int(periodic_parallel)))


def threads():
    return lib().ko_omp_threads()


# --- reference-convention initial states (host side, same formulas as the library's) ----
W9 = np.array([4 / 9, 1 / 9, 1 / 9, 1 / 9, 1 / 9, 1 / 36, 1 / 36, 1 / 36, 1 / 36])
CX9 = np.array([0, 1, 0, -1, 0, 1, -1, -1, 1])
CY9 = np.array([0, 0, 1, 0, -1, 1, 1, -1, -1])
OPP9 = np.array([0, 3, 4, 1, 2, 7, 8, 5, 6])
W19 = np.array([1 / 3] + [1 / 18] * 6 + [1 / 36] * 12)
CX19 = np.array([0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1, 0, 0, 0, 0])
CY19 = np.array([0, 0, 0, 1, -1, 0, 0, 1, 1, -1, -1, 0, 0, 0, 0, 1, -1, 1, -1])
CZ19 = np.array([0, 0, 0, 0, 0, 1, -1, 0, 0, 0, 0, 1, 1, -1, -1, 1, 1, -1, -1])
OPP19 = np.array([0, 2, 1, 4, 3, 6, 5, 10, 9, 8, 7, 14, 13, 12, 11, 18, 17, 16, 15])


def equilibrium(rho, ux, uy, uz=None, dtype=np.float64):
    """f_q = w ρ (1 + 3cu + 4.5cu² − 1.5u²) with cu = Σ c_a u_a (lattice/d2q9.jl:60-67)."""
    d3 = uz is not None
    W, C = (W19, (CX19, CY19, CZ19)) if d3 else (W9, (CX9, CY9))
    T = dtype
    rho, ux, uy = (np.asarray(a, dtype=T) for a in (rho, ux, uy))
    uz = np.asarray(uz, dtype=T) if d3 else None
    f = np.empty(rho.shape + (len(W),), dtype=T, order="F")
    usq = ux * ux + uy * uy
    if d3:
        usq = usq + uz * uz
    for q in range(len(W)):
        cu = T(C[0][q]) * ux + T(C[1][q]) * uy
        if d3:
            cu = cu + T(C[2][q]) * uz
        f[..., q] = T(W[q]) * rho * (T(1) + T(3) * cu + T(4.5) * cu * cu - T(1.5) * usq)
    return f


def rest(shape, dtype=np.float64):
    W = W9 if len(shape) == 2 else W19
    f = np.empty(tuple(shape) + (len(W),), dtype=dtype, order="F")
    for q in range(len(W)):
        f[..., q] = dtype(W[q])
    return f


def consistent_init(desc, f, iterations):
    """Consistent initialisation of a prescribed velocity field (Mei, Luo, Lallemand and
    d'Humières, Comput. Fluids 35 (2006) 855-862), FP64, on a D2Q9 state in place: the
    velocity of f is held while the chain of `desc` runs `iterations` steps, each followed by
    f_q += feq_q(ρ, u_held) − feq_q(ρ, u) with (ρ, u) the moments of f.  No reference
    counterpart (the reference starts from feq, src/drivers/basic.jl:51-70)."""
    m = moments(f)
    uh, vh = m["ux"].copy(order="F"), m["uy"].copy(order="F")
    for _ in range(iterations):
        run(desc, f, nsteps=1)
        m = moments(f)
        f += equilibrium(m["rho"], uh, vh) - equilibrium(m["rho"], m["ux"], m["uy"])
    return f
