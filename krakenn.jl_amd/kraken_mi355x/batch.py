"""Batched lid-driven cavities (kmi_batch): hundreds of uneven D2Q9 cavities per launch.

The reference runs one cavity per driver call (run_cavity_2d, src/drivers/basic.jl:112-144;
the MRT loop of test/test_mrt.jl:19-25).  Here every cavity of the ensemble advances in the
same single launch per step, each checked for the one-step residual r_u
(src/drivers/gnf_periodic_steady.jl:186-196) and retired once converged or diverged.
"""
from __future__ import annotations

import ctypes
from ctypes import POINTER, byref, c_double, c_int32, c_int64, c_void_p

import numpy as np

from . import _lib as L
from .lattice import ArgumentError, MI355XBackend

STATUS_NAMES = {L.CAV_RUNNING: "running", L.CAV_CONVERGED: "converged", L.CAV_DIVERGED: "diverged",
                L.CAV_MAX_STEPS: "max_steps"}


class Batch:
    def __init__(self, n, nu, *, storage="f32s", collision="mrt", lid="zouhe", u_lid=0.1,
                 s_e=1.4, s_eps=1.4, s_q=1.2, Lambda=3.0 / 16.0, backend=None):
        self.backend = backend or MI355XBackend()
        self.n = np.asarray(n, dtype=np.int32)
        self.nu = np.asarray(nu, dtype=np.float64)
        if self.n.shape != self.nu.shape or self.n.ndim != 1:
            raise ArgumentError("n and nu must be 1-D arrays of the same length")
        d = L.KmiBatchDesc()
        d.storage = L.STORAGE_NAMES[storage] if isinstance(storage, str) else storage
        d.collision = L.COLLISION_NAMES[collision] if isinstance(collision, str) else collision
        d.lid = L.FACE_ZOUHE_VELOCITY if lid == "zouhe" else L.FACE_MOVING_WALL
        d.u_lid, d.s_e, d.s_eps, d.s_q, d.Lambda = u_lid, s_e, s_eps, s_q, Lambda
        self.desc = d
        h = c_void_p()
        L.check(L.load().kmi_batch_create(self.backend.ctx, byref(d), len(self.n),
                                          self.n.ctypes.data_as(POINTER(c_int32)),
                                          self.nu.ctypes.data_as(POINTER(c_double)), byref(h)),
                self.backend.ctx)
        self._h = h

    def close(self):
        if getattr(self, "_h", None):
            L.load().kmi_batch_destroy(self._h)
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

    def __len__(self):
        return len(self.n)

    def _chk(self, rc):
        L.check(rc, self.backend.ctx)

    def upload(self, c, f):
        a = np.asarray(f)
        dt = L.HOST_F32 if a.dtype == np.float32 else L.HOST_F64
        a = np.asfortranarray(a, dtype=np.float32 if dt == L.HOST_F32 else np.float64)
        if a.shape != (self.n[c], self.n[c], 9):
            raise ArgumentError(f"cavity {c} populations must have shape {(self.n[c], self.n[c], 9)}")
        self._chk(L.load().kmi_batch_upload(self._h, int(c), a.ctypes.data_as(c_void_p), dt))

    def download(self, c, dtype=np.float64):
        out = np.empty((self.n[c], self.n[c], 9), dtype=dtype, order="F")
        dt = L.HOST_F32 if dtype == np.float32 else L.HOST_F64
        self._chk(L.load().kmi_batch_download(self._h, int(c), out.ctypes.data_as(c_void_p), dt))
        return out

    def init_equilibrium(self, c, rho=None, ux=None, uy=None):
        shp = (self.n[c], self.n[c])
        arrs = [None if a is None else np.asfortranarray(np.broadcast_to(a, shp), dtype=np.float64)
                for a in (rho, ux, uy)]
        ptrs = [None if a is None else a.ctypes.data_as(POINTER(c_double)) for a in arrs]
        self._chk(L.load().kmi_batch_init_equilibrium(self._h, int(c), *ptrs))

    def moments(self, c):
        shp = (self.n[c], self.n[c])
        rho, ux, uy = (np.empty(shp, order="F") for _ in range(3))
        self._chk(L.load().kmi_batch_moments(self._h, int(c), *[a.ctypes.data_as(POINTER(c_double)) for a in (rho, ux, uy)]))
        return {"rho": rho, "ux": ux, "uy": uy}

    def step(self, n=1):
        self._chk(L.load().kmi_batch_step(self._h, int(n)))

    def consistent_init(self, iterations, hold=None):
        """Hold the current velocity of the cavities flagged in `hold` (all when None) for
        `iterations` steps, re-imposing it after each step, so that density and the
        non-equilibrium populations become consistent with it (Mei et al. 2006)."""
        h = None
        if hold is not None:
            h = np.ascontiguousarray(hold, dtype=np.int32)
            if h.shape != self.n.shape:
                raise ArgumentError("hold must have one entry per cavity")
        self._chk(L.load().kmi_batch_consistent_init(
            self._h, None if h is None else h.ctypes.data_as(c_void_p), int(iterations)))

    def run(self, tol=1e-7, check_every=100, max_steps=10_000_000, window=1):
        """Run every cavity to r < tol (or divergence / max_steps); device ms elapsed.
        r = ‖u^k − u^(k−m)‖/(m‖u^k‖) over the last m = `window` steps (m = 1: the
        reference's one-step r_u)."""
        ms = c_double()
        self._chk(L.load().kmi_batch_run(self._h, float(tol), int(check_every), int(window),
                                         int(max_steps), byref(ms)))
        return ms.value

    def reset_status(self):
        self._chk(L.load().kmi_batch_reset_status(self._h))

    def resume(self):
        """Cavities stopped at max_steps run again (step counts kept)."""
        self._chk(L.load().kmi_batch_resume(self._h))

    def status(self):
        k = len(self.n)
        st = np.empty(k, dtype=np.int32)
        steps = np.empty(k, dtype=np.int64)
        r = np.empty(k, dtype=np.float64)
        self._chk(L.load().kmi_batch_status(self._h, st.ctypes.data_as(POINTER(c_int32)),
                                            steps.ctypes.data_as(POINTER(c_int64)),
                                            r.ctypes.data_as(POINTER(c_double))))
        return {"status": st, "steps": steps, "residual": r}

    def set_launch(self, tw, th, groups=0, nontemporal=False):
        self._chk(L.load().kmi_batch_set_launch(self._h, int(tw), int(th), int(groups), int(bool(nontemporal))))

    @property
    def launch(self):
        v = [c_int32() for _ in range(4)]
        self._chk(L.load().kmi_batch_get_launch(self._h, *[byref(x) for x in v]))
        sp = c_int32()
        self._chk(L.load().kmi_batch_get_temporal(self._h, byref(sp)))
        return {"tw": v[0].value, "th": v[1].value, "groups": v[2].value, "nontemporal": bool(v[3].value),
                "steps_per_launch": sp.value}

    def set_temporal(self, steps_per_launch):
        """1: one step per launch; 2: two steps per launch (step t+1 of each tile and its ring
        kept in LDS); 3 or 4: the tile and a 3- or 4-cell ring advanced in LDS (tile width 16,
        32 or 64).  Bit-identical to single steps."""
        self._chk(L.load().kmi_batch_set_temporal(self._h, int(steps_per_launch)))

    def autotune(self, steps_per_trial=8):
        ms = c_double()
        self._chk(L.load().kmi_batch_autotune(self._h, int(steps_per_trial), byref(ms)))
        return ms.value

    @property
    def active_cells(self):
        c = c_int64()
        self._chk(L.load().kmi_batch_active_cells(self._h, byref(c)))
        return c.value


def ensemble_parameters(ncav=256, seed=2024, u_lid=0.1, sizes=tuple(range(64, 513, 32)),
                        Re_range=(100.0, 3200.0)):
    """Side N uniform in {64, 96, …, 512}, Re log-uniform in [100, 3200], ν = u_lid·N/Re."""
    rng = np.random.default_rng(seed)
    n = rng.choice(np.asarray(sizes), size=ncav)
    Re = np.exp(rng.uniform(np.log(Re_range[0]), np.log(Re_range[1]), size=ncav))
    nu = u_lid * n / Re
    return n.astype(np.int32), nu, Re
