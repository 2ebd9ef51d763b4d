"""Backend value and device-resident lattice, over the C ABI.

`MI355XBackend` plays the role of the KernelAbstractions backend object that Kraken's
drivers take as `backend=` (src/drivers/basic.jl:51-52, src/simulation_runner.jl:110-160):
asking for it where no device is visible raises `ArgumentError`, never a silent CPU run
(the reference's contract, test/test_simulation_runner.jl:327-400).
"""
from __future__ import annotations

import ctypes
from ctypes import POINTER, byref, c_double, c_int, c_int64, c_uint8, c_void_p

import numpy as np

from . import _lib as L


class ArgumentError(ValueError):
    """Invalid argument or unavailable backend (Julia's ArgumentError)."""


def _dptr(a):
    return None if a is None else a.ctypes.data_as(POINTER(c_double))


class MI355XBackend:
    """One HIP device + stream (a kmi_ctx).  Shared by every lattice built on it."""

    _cache: dict = {}

    def __new__(cls, device: int = 0):
        if device in cls._cache:
            return cls._cache[device]
        self = super().__new__(cls)
        lib = L.load()
        ctx = c_void_p()
        rc = lib.kmi_ctx_create(device, byref(ctx))
        if rc != L.KMI_OK:
            msg = lib.kmi_last_error(None).decode()
            raise ArgumentError(f"Backend mi355x (device {device}) is unavailable: {msg}")
        self.device = device
        self.ctx = ctx
        cls._cache[device] = self
        return self

    def __repr__(self):
        return f"MI355XBackend(device={self.device})"

    def synchronize(self):
        L.check(L.load().kmi_ctx_synchronize(self.ctx), self.ctx)

    def copy_bandwidth(self, nbytes=2 * 1024 ** 3, reps=20):
        """Measured device copy bandwidth in GB/s (read + write counted)."""
        g = c_double()
        L.check(L.load().kmi_measure_copy_bandwidth(self.ctx, int(nbytes), int(reps), byref(g)), self.ctx)
        return g.value

    @staticmethod
    def device_count() -> int:
        n = c_int()
        L.check(L.load().kmi_device_count(byref(n)))
        return n.value


def make_desc(stencil, nx, ny, nz=1, nu=0.1, storage="f64", collision="bgk", faces=None,
              thermal=False, omega=None, s_plus=None, s_minus=None, Lambda=3.0 / 16.0,
              s_e=1.4, s_eps=1.4, s_q=1.2, omega_T=None, beta_g=0.0, T_ref=0.0,
              coupling="split", bc_order=None, forcing=None, force=(0.0, 0.0, 0.0),
              solid_mode="bounce_back") -> L.KmiDesc:
    """Build a kmi_desc.  `faces` maps face name -> dict(kind=..., u=..., rho=..., T=...,
    thermal='dirichlet'|'adiabatic').  forcing: None, 'guo' (moments with the F/2 shift) or
    'guo_plain' (moments without it), with the body force `force`, or 'guo_field' (per-node
    force, Lattice.set_force_field).  solid_mode: masked cells
    swap their populations ('bounce_back') or hold the rest equilibrium ('inert', LI-BB)."""
    st = {"d2q9": L.D2Q9, "d3q19": L.D3Q19}.get(str(stencil).lower(), stencil)
    d = L.desc_init(st, nx, ny, nz, nu)
    d.storage = L.STORAGE_NAMES[storage] if isinstance(storage, str) else storage
    d.collision = L.COLLISION_NAMES[collision] if isinstance(collision, str) else collision
    d.thermal = 1 if thermal else 0
    if omega is not None:
        d.omega = omega
    sp, sm = L.trt_rates(nu, Lambda)
    d.s_plus = sp if s_plus is None else s_plus
    d.s_minus = sm if s_minus is None else s_minus
    d.s_e, d.s_eps, d.s_q = s_e, s_eps, s_q
    d.s_nu = 1.0 / (3.0 * nu + 0.5)
    d.omega_T = d.omega if omega_T is None else omega_T
    d.beta_g = beta_g
    d.T_ref = T_ref
    d.coupling = L.COUPLING_FUSED if coupling == "fused" else L.COUPLING_SPLIT
    for name, spec in (faces or {}).items():
        _fill_face(d.faces[L.FACE_NAMES[name]], spec)
    if bc_order:
        d.n_bc_order = len(bc_order)
        for n, name in enumerate(bc_order):
            d.bc_order[n] = L.FACE_NAMES[name] if isinstance(name, str) else name
    d.forcing = {None: L.FORCE_NONE, "none": L.FORCE_NONE, "guo": L.FORCE_GUO,
                 "guo_plain": L.FORCE_GUO_PLAIN_MOMENTS, "guo_field": L.FORCE_GUO_FIELD}[forcing]
    fv = list(force) + [0.0, 0.0, 0.0]
    for c in range(3):
        d.force[c] = float(fv[c])
    d.solid_mode = {"bounce_back": L.SOLID_BOUNCE_BACK, "inert": L.SOLID_INERT}[solid_mode]
    return d


FACE_KINDS = {"periodic": L.FACE_PERIODIC, "wall": L.FACE_WALL, "moving_wall": L.FACE_MOVING_WALL,
              "zouhe_velocity": L.FACE_ZOUHE_VELOCITY, "zouhe_pressure": L.FACE_ZOUHE_PRESSURE,
              "zouhe_velocity_profile": L.FACE_ZOUHE_VELOCITY_PROFILE,
              "zouhe_pressure_profile": L.FACE_ZOUHE_PRESSURE_PROFILE,
              "zouhe_velocity_rebuild": L.FACE_ZOUHE_VELOCITY_REBUILD,
              "zouhe_pressure_rebuild": L.FACE_ZOUHE_PRESSURE_REBUILD}


def _fill_face(f, spec):
    k = spec.get("kind", "wall")
    f.kind = FACE_KINDS[k] if isinstance(k, str) else k
    u = list(spec.get("u", (0.0, 0.0, 0.0))) + [0.0, 0.0, 0.0]
    for c in range(3):
        f.u[c] = float(u[c])
    f.rho = float(spec.get("rho", 1.0))
    f.T = float(spec.get("T", 0.0))
    f.thermal = L.THERMAL_DIRICHLET if spec.get("thermal") == "dirichlet" else L.THERMAL_ADIABATIC


class Lattice:
    """A device-resident lattice (kmi_lattice).  Host arrays use the reference layout
    f[i, j, (k,) q]: numpy arrays of shape (nx, ny, Q) / (nx, ny, nz, Q) in Fortran order."""

    def __init__(self, desc: L.KmiDesc, backend: MI355XBackend | None = None):
        self.backend = backend or MI355XBackend()
        self.desc = desc
        self.Q = 9 if desc.stencil == L.D2Q9 else 19
        self.shape = (desc.nx, desc.ny) if self.Q == 9 else (desc.nx, desc.ny, desc.nz)
        self.ncells = int(np.prod(self.shape))
        lib = L.load()
        h = c_void_p()
        L.check(lib.kmi_lattice_create(self.backend.ctx, byref(desc), byref(h)), self.backend.ctx)
        self._h = h
        # Host copies of the state set through the setters below (mask, cut links, force
        # field, face profiles and face values), for checkpoints of the whole lattice.
        self.aux = {}

    def close(self):
        if getattr(self, "_h", None):
            L.load().kmi_lattice_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def _chk(self, rc):
        L.check(rc, self.backend.ctx)

    # --- populations ---
    def upload(self, f, field="f"):
        fld = L.FIELD_G if field == "g" else L.FIELD_F
        a = np.asarray(f)
        dt = L.HOST_F32 if a.dtype == np.float32 else L.HOST_F64
        a = np.asfortranarray(a, dtype=np.float32 if dt == L.HOST_F32 else np.float64)
        if a.size != self.ncells * self.Q:
            raise ArgumentError(f"population array has {a.size} values, expected {self.ncells * self.Q}")
        self._chk(L.load().kmi_upload(self._h, fld, a.ctypes.data_as(c_void_p), dt))

    def download(self, field="f", dtype=np.float64):
        fld = L.FIELD_G if field == "g" else L.FIELD_F
        out = np.empty(self.shape + (self.Q,), dtype=dtype, order="F")
        dt = L.HOST_F32 if dtype == np.float32 else L.HOST_F64
        self._chk(L.load().kmi_download(self._h, fld, out.ctypes.data_as(c_void_p), dt))
        return out

    def storage_dtype(self, field="f"):
        """numpy type of the stored elements of a field (KMI_HOST_STORAGE transfers)."""
        st = self.desc.storage
        if field == "g" and st == L.STORE_F16S:
            st = L.STORE_F32S  # thermal populations of an F16S lattice are F32S
        return {L.STORE_F64: np.float64, L.STORE_F32: np.float32, L.STORE_F32S: np.float32,
                L.STORE_F16S: np.uint16}[st]

    def download_raw(self, field="f"):
        """The stored elements verbatim (exact state: F32S/F16S deviations, F16S as binary16 bits)."""
        fld = L.FIELD_G if field == "g" else L.FIELD_F
        out = np.empty(self.shape + (self.Q,), dtype=self.storage_dtype(field), order="F")
        self._chk(L.load().kmi_download(self._h, fld, out.ctypes.data_as(c_void_p), L.HOST_STORAGE))
        return out

    def upload_raw(self, a, field="f"):
        fld = L.FIELD_G if field == "g" else L.FIELD_F
        want = self.storage_dtype(field)
        a = np.asarray(a)
        if a.dtype != want or a.size != self.ncells * self.Q:
            raise ArgumentError(f"raw {field} needs {self.ncells * self.Q} values of {np.dtype(want)}")
        a = np.asfortranarray(a.reshape(self.shape + (self.Q,), order="F"))
        self._chk(L.load().kmi_upload(self._h, fld, a.ctypes.data_as(c_void_p), L.HOST_STORAGE))

    def set_mask(self, is_solid):
        if is_solid is None:
            self._chk(L.load().kmi_set_mask(self._h, None))
            self.aux.pop("is_solid", None)
            return
        m = np.asfortranarray(np.asarray(is_solid, dtype=np.uint8))
        if m.size != self.ncells:
            raise ArgumentError("mask size does not match the lattice")
        self._chk(L.load().kmi_set_mask(self._h, m.ctypes.data_as(POINTER(c_uint8))))
        self.aux["is_solid"] = m.reshape(self.shape, order="F") != 0

    def init_equilibrium(self, rho=None, ux=None, uy=None, uz=None):
        arrs = [None if a is None else np.asfortranarray(np.broadcast_to(a, self.shape), dtype=np.float64)
                for a in (rho, ux, uy, uz)]
        self._chk(L.load().kmi_init_equilibrium(self._h, *[_dptr(a) for a in arrs]))

    def set_face_profile(self, face, a, b=None):
        """Per-node values of a profile face: (u_x, u_y) arrays for a velocity profile, the
        density array `a` for a pressure profile, or the normal velocity `a` of a rebuilt
        inlet (length nx on south/north, ny on west/east; on a D3Q19 face an array of the two
        tangent extents, e.g. (ny, nz) on west)."""
        fid = L.FACE_NAMES[face] if isinstance(face, str) else int(face)
        if self.Q == 9:
            tshape = (self.shape[0],) if fid in (L.SOUTH, L.NORTH) else (self.shape[1],)
        else:
            ax = fid // 2
            tshape = tuple(self.shape[t] for t in range(3) if t != ax)
        flat = lambda v: np.asfortranarray(np.broadcast_to(v, tshape), dtype=np.float64).ravel(order="F")
        pa = flat(a)
        pb = None if b is None else flat(b)
        self._chk(L.load().kmi_set_face_profile(self._h, fid, _dptr(pa), _dptr(pb)))
        self.aux.setdefault("profiles", {})[fid] = (pa.copy(), None if pb is None else pb.copy())

    def set_face(self, face, spec):
        """Replace a face's values (u, rho, T) between steps; its kind must not change."""
        fid = L.FACE_NAMES[face] if isinstance(face, str) else int(face)
        f = L.KmiFace()
        _fill_face(f, spec)
        self._chk(L.load().kmi_set_face(self._h, fid, byref(f)))
        self.aux.setdefault("faces", {})[fid] = f

    def init_temperature(self, T=None):
        a = None if T is None else np.asfortranarray(np.broadcast_to(T, self.shape), dtype=np.float64)
        self._chk(L.load().kmi_init_temperature(self._h, _dptr(a)))

    # --- time stepping ---
    def step(self, n=1):
        self._chk(L.load().kmi_step(self._h, int(n)))

    def step_timed(self, n):
        ms = c_double()
        self._chk(L.load().kmi_step_timed(self._h, int(n), byref(ms)))
        return ms.value

    def step_residual(self, n, window=1):
        """Advance n steps; the relative velocity change over the last `window` steps,
        ‖u^k − u^(k−m)‖/(m‖u^k‖) (m = 1: the reference's one-step r_u)."""
        r = c_double()
        if window == 1:
            self._chk(L.load().kmi_step_residual(self._h, int(n), byref(r)))
        else:
            self._chk(L.load().kmi_step_residual_window(self._h, int(n), int(window), byref(r)))
        return r.value

    @property
    def steps(self):
        s = c_int64()
        self._chk(L.load().kmi_lattice_steps(self._h, byref(s)))
        return s.value

    @steps.setter
    def steps(self, n):
        self._chk(L.load().kmi_lattice_set_steps(self._h, int(n)))

    # --- launch tuning (results never depend on it) ---
    def set_launch(self, bx, by=1, nontemporal=False):
        self._chk(L.load().kmi_lattice_set_launch(self._h, int(bx), int(by), int(bool(nontemporal))))

    @property
    def launch(self):
        bx, by, nt = ctypes.c_int32(), ctypes.c_int32(), ctypes.c_int32()
        self._chk(L.load().kmi_lattice_get_launch(self._h, byref(bx), byref(by), byref(nt)))
        return {"bx": bx.value, "by": by.value, "nontemporal": bool(nt.value)}

    def set_temporal(self, tile, zchunks=1):
        """Temporal blocking.  Fully periodic D3Q19 BGK/TRT lattice: two steps per launch
        (tile 1: 32×8, 2: 64×4, 3: 64×8, 4: 32×16, 5: 64×16 cells with a thread per cell;
        12: 64×8, 13: 32×16, 14: 32×8 with a thread per cell of the tile and its ring) or three (6: 64×8, 7: 32×8,
        8: 32×16 with a thread per cell; 9: 32×8, 10: 32×16, 11: 64×8 with a thread per cell
        of the tile and its two-cell ring; FP64 only 7 and 9).  D2Q9 lattice without thermal populations, cut links or
        per-node faces: 1: 28×28 two steps, 2: 24×24 four, 3: 56×8 four, 4: 16×16 eight,
        5: 60×12 two, 6: 8×8 four, 7: 16×16 four, 8: 8×8 eight.  0: off.  Bit-identical to single steps."""
        self._chk(L.load().kmi_lattice_set_temporal(self._h, int(tile), int(zchunks)))

    @property
    def temporal(self):
        t, z = ctypes.c_int32(), ctypes.c_int32()
        self._chk(L.load().kmi_lattice_get_temporal(self._h, byref(t), byref(z)))
        return {"tile": t.value, "zchunks": z.value}

    def autotune(self, steps_per_trial=64):
        """Pick the fastest launch shape / store policy / two-step variant on this lattice;
        state is restored."""
        ms = c_double()
        self._chk(L.load().kmi_lattice_autotune(self._h, int(steps_per_trial), byref(ms)))
        return ms.value

    @property
    def device_bytes(self):
        b = c_int64()
        self._chk(L.load().kmi_lattice_device_bytes(self._h, byref(b)))
        return b.value

    # --- observables ---
    def moments(self, temperature=False):
        shp = self.shape
        rho = np.empty(shp, order="F")
        ux = np.empty(shp, order="F")
        uy = np.empty(shp, order="F")
        uz = np.empty(shp, order="F") if self.Q == 19 else None
        T = np.empty(shp, order="F") if temperature else None
        self._chk(L.load().kmi_moments(self._h, *[_dptr(a) for a in (rho, ux, uy, uz, T)]))
        out = {"rho": rho, "ux": ux, "uy": uy}
        if uz is not None:
            out["uz"] = uz
        if T is not None:
            out["Temp"] = T
        return out

    def reduce(self, kind, arg=0):
        out = (c_double * 3)()
        k = {"mass": L.RED_MASS, "momentum": L.RED_MOMENTUM, "kinetic_energy": L.RED_KINETIC_ENERGY,
             "thermal": L.RED_THERMAL, "nusselt": L.RED_NUSSELT, "max_speed": L.RED_MAX_SPEED,
             "drag_mea": L.RED_DRAG_MEA, "drag_halfway": L.RED_DRAG_HALFWAY}.get(kind, kind)
        a = L.FACE_NAMES[arg] if isinstance(arg, str) else arg
        self._chk(L.load().kmi_reduce(self._h, k, a, out))
        if k in (L.RED_MOMENTUM, L.RED_DRAG_MEA, L.RED_DRAG_HALFWAY):
            return tuple(out)
        return out[0]

    def set_force_field(self, fx, fy, fz=None):
        """Per-node body force of a 'guo_field' lattice (collide_guo_field_2d!/_3d!), arrays
        of the lattice shape in the reference layout; rounded once to the collision
        arithmetic."""
        shp = self.shape
        arrs = []
        for a in (fx, fy, fz):
            if a is None:
                arrs.append(None)
                continue
            a = np.asarray(a, dtype=np.float64)
            if a.shape != shp:
                raise ArgumentError(f"force field shape {a.shape} != lattice shape {shp}")
            arrs.append(np.asfortranarray(a))
        if arrs[0] is None or arrs[1] is None or (len(shp) == 3 and arrs[2] is None):
            raise ArgumentError("set_force_field needs every component of the lattice's dimension")
        self._chk(L.load().kmi_set_force_field(self._h, _dptr(arrs[0]), _dptr(arrs[1]),
                                               _dptr(arrs[2]) if arrs[2] is not None else None))
        self.aux["force_field"] = tuple(a.copy(order="F") for a in arrs[:len(shp)])

    def set_cut_links(self, cell, q, qw, uw=None):
        """Interpolated bounce-back links: cell index (x fastest), direction q (0-based),
        wall fraction in (0, 1], wall velocity (n × 3, default zero)."""
        cell = np.ascontiguousarray(cell, dtype=np.int32)
        q = np.ascontiguousarray(q, dtype=np.int32)
        qw = np.ascontiguousarray(qw, dtype=np.float64)
        n = len(cell)
        uw = np.zeros((n, 3)) if uw is None else np.asarray(uw, dtype=np.float64).reshape(n, -1)
        if uw.shape[1] < 3:
            uw = np.hstack([uw, np.zeros((n, 3 - uw.shape[1]))])
        uw = np.ascontiguousarray(uw)
        ip = lambda a: a.ctypes.data_as(POINTER(ctypes.c_int32))
        self._chk(L.load().kmi_set_cut_links(self._h, n, ip(cell), ip(q), _dptr(qw), _dptr(uw)))
        if n:
            self.aux["cut_links"] = (cell.copy(), q.copy(), qw.copy(), uw.copy())
        else:
            self.aux.pop("cut_links", None)
