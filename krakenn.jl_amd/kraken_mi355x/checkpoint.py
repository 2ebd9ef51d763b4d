"""Resumable lattice state: the state-snapshot contract of Kraken.jl (src/platform/state.jl)
and its HDF5 container (src/io/checkpoint_hdf5.jl), fed from the device.

Three layers, as in the reference:

- ``StateSnapshot``: a plain host image of a solver state (arrays, histories, scalars and
  four configuration dictionaries) with the reference's content rules
  (``validate_content``, ``check_finite``, ``check_compatible``: the contract of
  state.jl:58-215, 367-445, written here as a table of rules);
- ``write_checkpoint`` / ``read_checkpoint`` / ``checkpoint_info``: one HDF5 file per
  checkpoint in the reference's container layout, version 1 (checkpoint_hdf5.jl:23-257):
  root attributes ``container_version, schema_version, solver, cycle, kraken_version,
  julia_version, created_unix``; groups ``/fields`` and ``/series`` holding one chunked,
  Fletcher32-checksummed dataset per entry (a ``/`` in a name is a nested group; chunk
  shape of _checkpoint_chunk); groups ``/scalars, /identity, /run_control, /parameters,
  /derived`` holding one attribute per entry plus a newline-joined ``keys`` attribute;
  fixed-length strings only; superblock v3 with checksummed object headers (library
  version bounds 1.10, 1.10); write to ``path.tmp``, keep the old file as ``path.prev``,
  rename into place.  Arrays are stored in the reference's column-major order (HDF5 sees
  reversed dimensions, as for a Julia array);
- ``snapshot(lattice)`` / ``restore_lattice(snap)``: the lattice client.  The populations
  are exported with ``Lattice.download_raw`` (the storage elements themselves, binary16
  bits widened to Int32 for F16S), so a restored lattice continues bit for bit.

HDF5 is reached through the C library of the image (``libhdf5``, 1.10 API) with ctypes;
without it, writing and reading raise ``CheckpointError``.
"""
from __future__ import annotations

import ctypes
import math
import os
import sys
import time
from ctypes import POINTER, byref, c_char_p, c_double, c_int, c_int64, c_uint, c_uint64, c_void_p
from dataclasses import dataclass, field as dfield

import numpy as np

from . import _lib as L

CHECKPOINT_CONTAINER_VERSION = 1
SNAPSHOT_ELTYPES = (np.float32, np.float64, np.int32, np.int64, np.bool_)
SNAPSHOT_CONFIG_CLASSES = ("identity", "run_control", "parameters")
_ARRAY_GROUPS = ("fields", "series")
_SCALAR_GROUPS = ("scalars", "identity", "run_control", "parameters", "derived")
_KEYS_ATTR = "keys"
_VECTOR_CHUNK = 65_536
_MATRIX_CHUNK_BYTES = 1 << 20
_MAX_CHUNK_BYTES = 1 << 30


class CheckpointError(Exception):
    """A snapshot or checkpoint file was refused; the message names the key or object."""


# --------------------------------------------------------------------------------------
# StateSnapshot and its content rules (state.jl:58-215)
# --------------------------------------------------------------------------------------
@dataclass
class StateSnapshot:
    solver: str
    schema_version: int
    cycle: int
    fields: dict = dfield(default_factory=dict)
    series: dict = dfield(default_factory=dict)
    scalars: dict = dfield(default_factory=dict)
    identity: dict = dfield(default_factory=dict)
    run_control: dict = dfield(default_factory=dict)
    parameters: dict = dfield(default_factory=dict)
    derived: dict = dfield(default_factory=dict)

    def __post_init__(self):
        self.solver = str(self.solver)
        self.schema_version = int(self.schema_version)
        self.cycle = int(self.cycle)
        validate_content(self)


# A snapshot entry is checked against a table of rules.  Each rule names the groups it
# covers, a test that is true when the entry breaks it, and the reason given in the error
# ("<group>/<name>: <reason>").  The contract the rules encode is the reference's (allowed
# element and scalar types, the reserved attribute name, path-like array names that may not
# double as group prefixes, vectors only in `series`, None only in the configuration
# classes); the table form and the wording are this module's.
_CONTROL_CHARS = {chr(c) for c in range(32)} | {chr(127)}


@dataclass(frozen=True)
class _Rule:
    groups: tuple
    broken: object   # (name, value) -> bool
    reason: object   # (name, value) -> str


def _bad_path(name) -> bool:
    if not isinstance(name, str) or not name or _CONTROL_CHARS & set(name):
        return True
    return any(seg in ("", ".", "..") for seg in name.split("/"))


def _storable_scalar(v, group) -> bool:
    if v is None:
        return group in SNAPSHOT_CONFIG_CLASSES
    return isinstance(v, (str, bool, int, float, np.floating, np.integer, np.bool_))


_ARRAY_RULES = (
    _Rule(_ARRAY_GROUPS, lambda n, a: _bad_path(n),
          lambda n, a: "array names are `/`-joined paths of non-empty segments other than "
                       "`.` and `..`, without control characters"),
    _Rule(_ARRAY_GROUPS, lambda n, a: not isinstance(a, np.ndarray) or a.dtype.type not in SNAPSHOT_ELTYPES,
          lambda n, a: f"dtype {getattr(a, 'dtype', type(a).__name__)} cannot be stored "
                       "(Float32, Float64, Int32, Int64 or Bool arrays only)"),
    _Rule(_ARRAY_GROUPS, lambda n, a: a.ndim == 0,
          lambda n, a: "a 0-d array is a scalar: put it in `scalars`"),
    _Rule(("series",), lambda n, a: a.ndim != 1,
          lambda n, a: f"series entries are 1-D, this one has shape {a.shape}"),
)

_SCALAR_RULES = (
    _Rule(_SCALAR_GROUPS, lambda k, v: not isinstance(k, str) or not k or "/" in k or bool(_CONTROL_CHARS & set(k)),
          lambda k, v: "scalar keys are single non-empty names without `/` or control characters"),
    _Rule(_SCALAR_GROUPS, lambda k, v: k == _KEYS_ATTR,
          lambda k, v: f"`{_KEYS_ATTR}` is the group's key-list attribute and cannot name an entry"),
)


def _scalar_type_rule(group):
    extra = ", or None" if group in SNAPSHOT_CONFIG_CLASSES else ""
    return _Rule((group,), lambda k, v: not _storable_scalar(v, group),
                 lambda k, v: f"{type(v).__name__} is not a storable scalar (str, bool, int or float{extra})")


def _enforce(rules, group, entries):
    for name, value in entries.items():
        for r in rules:
            if group in r.groups and r.broken(name, value):
                raise CheckpointError(f"{group}/{name}: {r.reason(name, value)}")


def _enforce_no_prefix_clash(group, names):
    # an array name cannot also be the group that holds another array ("a" and "a/b")
    ordered = sorted(names)
    for i, name in enumerate(ordered):
        for later in ordered[i + 1:]:
            if later.startswith(name + "/"):
                raise CheckpointError(f"{group}/{name}: the name is also the parent group of {group}/{later}")
            if not later.startswith(name):
                break


def validate_content(snap: StateSnapshot) -> StateSnapshot:
    """Refuse a snapshot whose names, keys, element or value types the container cannot hold
    (the content contract of state.jl:139-215)."""
    if not snap.solver:
        raise CheckpointError("snapshot: the solver name is empty")
    if snap.cycle < 0:
        raise CheckpointError(f"snapshot: cycle {snap.cycle} is negative")
    for group in _ARRAY_GROUPS:
        entries = getattr(snap, group)
        _enforce(_ARRAY_RULES, group, entries)
        _enforce_no_prefix_clash(group, entries)
    for group in _SCALAR_GROUPS:
        _enforce(_SCALAR_RULES + (_scalar_type_rule(group),), group, getattr(snap, group))
    return snap


def check_finite(snap: StateSnapshot) -> StateSnapshot:
    """Refuse NaN or Inf in any floating-point array; the error gives the first bad element
    with 1-based (column-major, Julia) indices."""
    for group in _ARRAY_GROUPS:
        for name, a in getattr(snap, group).items():
            if a.dtype.kind != "f":
                continue
            bad = np.argwhere(~np.isfinite(a))
            if len(bad):
                at = tuple(int(v) for v in bad[0])
                raise CheckpointError(f"{group}/{name}: {a[at]} at 1-based index {tuple(v + 1 for v in at)}; "
                                      "a snapshot must hold finite values only")
    return snap


def _same_scalar(a, b):
    if type(a) is not type(b):
        return False
    if isinstance(a, float) and math.isnan(a) and math.isnan(b):
        return True
    return a == b


def _typed(v):
    return f"{v!r}::{type(v).__name__}"


def check_compatible(snap: StateSnapshot, *, solver, schema_version, identity, identity_defaults=None):
    """A snapshot may restore a run only if it comes from the same solver and schema version
    and its identity entries equal the run's, type included; keys the run declares with a
    default may be absent.  run_control and parameters are free to differ (the compatibility
    contract of state.jl:409-441)."""
    defaults = identity_defaults or {}
    if snap.solver != solver:
        raise CheckpointError(f"solver: the snapshot was written by {snap.solver!r}, this run is {solver!r}")
    if snap.schema_version != schema_version:
        raise CheckpointError(f"schema_version: the snapshot has {snap.schema_version}, "
                              f"this run reads {schema_version}")
    for key, want in identity.items():
        have = snap.identity.get(key, defaults.get(key, _MISSING))
        if have is _MISSING:
            raise CheckpointError(f"identity/{key}: the run needs {_typed(want)}; the snapshot lacks the key "
                                  "and the run declares no default")
        if not _same_scalar(have, want):
            raise CheckpointError(f"identity/{key}: the snapshot has {_typed(have)}, the run needs {_typed(want)}")
    extra = [k for k in snap.identity if k not in identity]
    if extra:
        raise CheckpointError(f"identity/{extra[0]}: the snapshot has {_typed(snap.identity[extra[0]])}, "
                              "a key this run does not declare")
    return snap


_MISSING = object()


# --------------------------------------------------------------------------------------
# HDF5 through the C library (1.10 API)
# --------------------------------------------------------------------------------------
_H5 = None
_HID = c_int64
_HSIZE = c_uint64
_H5F_ACC_RDONLY, _H5F_ACC_TRUNC = 0, 2
_H5F_LIBVER_V110 = 2
_H5S_SCALAR = 0
_H5T_INTEGER, _H5T_FLOAT, _H5T_STRING, _H5T_ENUM = 0, 1, 3, 8
_H5T_STR_NULLTERM, _H5T_CSET_UTF8 = 0, 1
_H5I_GROUP, _H5I_DATASET = 2, 5
_H5_INDEX_NAME, _H5_ITER_INC = 0, 0
_ITER_CB = ctypes.CFUNCTYPE(c_int, _HID, c_char_p, c_void_p, c_void_p)


def _libhdf5():
    global _H5
    if _H5 is not None:
        return _H5
    cands = [os.environ.get("KRAKEN_MI355X_LIBHDF5", "")]
    cands += [os.path.join(p, "lib", n) for p in (sys.prefix, "/opt/conda", "/usr")
              for n in ("libhdf5.so", "libhdf5.so.103", "libhdf5_serial.so")]
    cands += ["/usr/lib/x86_64-linux-gnu/libhdf5_serial.so", "libhdf5.so"]
    err = None
    for c in cands:
        if not c:
            continue
        try:
            lib = ctypes.CDLL(c)
            break
        except OSError as e:
            err = e
    else:
        raise CheckpointError(f"the HDF5 C library is not available ({err})")
    if lib.H5open() < 0:
        raise CheckpointError("H5open failed")
    maj, mi, rel = c_uint(), c_uint(), c_uint()
    lib.H5get_libversion(byref(maj), byref(mi), byref(rel))
    if (maj.value, mi.value) < (1, 10):
        raise CheckpointError(f"HDF5 {maj.value}.{mi.value} is too old (1.10 needed)")
    lib.H5Eset_auto2(_HID(0), None, None)  # errors surface as CheckpointError, not stderr dumps
    sig = {
        "H5Pcreate": (_HID, [_HID]), "H5Pclose": (c_int, [_HID]),
        "H5Pset_libver_bounds": (c_int, [_HID, c_int, c_int]),
        "H5Pset_chunk": (c_int, [_HID, c_int, POINTER(_HSIZE)]), "H5Pset_fletcher32": (c_int, [_HID]),
        "H5Fcreate": (_HID, [c_char_p, c_uint, _HID, _HID]), "H5Fopen": (_HID, [c_char_p, c_uint, _HID]),
        "H5Fclose": (c_int, [_HID]),
        "H5Gcreate2": (_HID, [_HID, c_char_p, _HID, _HID, _HID]), "H5Gopen2": (_HID, [_HID, c_char_p, _HID]),
        "H5Gclose": (c_int, [_HID]),
        "H5Oopen": (_HID, [_HID, c_char_p, _HID]), "H5Oclose": (c_int, [_HID]), "H5Iget_type": (c_int, [_HID]),
        "H5Literate": (c_int, [_HID, c_int, c_int, POINTER(_HSIZE), _ITER_CB, c_void_p]),
        "H5Screate_simple": (_HID, [c_int, POINTER(_HSIZE), POINTER(_HSIZE)]), "H5Screate": (_HID, [c_int]),
        "H5Sclose": (c_int, [_HID]), "H5Sget_simple_extent_ndims": (c_int, [_HID]),
        "H5Sget_simple_extent_dims": (c_int, [_HID, POINTER(_HSIZE), POINTER(_HSIZE)]),
        "H5Dcreate2": (_HID, [_HID, c_char_p, _HID, _HID, _HID, _HID, _HID]),
        "H5Dopen2": (_HID, [_HID, c_char_p, _HID]), "H5Dclose": (c_int, [_HID]),
        "H5Dwrite": (c_int, [_HID, _HID, _HID, _HID, _HID, c_void_p]),
        "H5Dread": (c_int, [_HID, _HID, _HID, _HID, _HID, c_void_p]),
        "H5Dget_space": (_HID, [_HID]), "H5Dget_type": (_HID, [_HID]),
        "H5Acreate2": (_HID, [_HID, c_char_p, _HID, _HID, _HID, _HID]),
        "H5Aopen": (_HID, [_HID, c_char_p, _HID]), "H5Aclose": (c_int, [_HID]),
        "H5Awrite": (c_int, [_HID, _HID, c_void_p]), "H5Aread": (c_int, [_HID, _HID, c_void_p]),
        "H5Aget_type": (_HID, [_HID]), "H5Aget_space": (_HID, [_HID]), "H5Aexists": (c_int, [_HID, c_char_p]),
        "H5Aiterate2": (c_int, [_HID, c_int, c_int, POINTER(_HSIZE), _ITER_CB, c_void_p]),
        "H5Tcopy": (_HID, [_HID]), "H5Tclose": (c_int, [_HID]), "H5Tset_size": (c_int, [_HID, ctypes.c_size_t]),
        "H5Tset_strpad": (c_int, [_HID, c_int]), "H5Tset_cset": (c_int, [_HID, c_int]),
        "H5Tget_class": (c_int, [_HID]), "H5Tget_size": (ctypes.c_size_t, [_HID]),
        "H5Tis_variable_str": (c_int, [_HID]), "H5Tget_sign": (c_int, [_HID]),
        "H5Tenum_create": (_HID, [_HID]), "H5Tenum_insert": (c_int, [_HID, c_char_p, c_void_p]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype, fn.argtypes = res, args
    g = lambda n: _HID.in_dll(lib, n).value
    lib.T = {"f64": g("H5T_NATIVE_DOUBLE_g"), "f32": g("H5T_NATIVE_FLOAT_g"),
             "i64": g("H5T_NATIVE_LLONG_g"), "i32": g("H5T_NATIVE_INT_g"),
             "i8": g("H5T_NATIVE_SCHAR_g"), "c_s1": g("H5T_C_S1_g")}
    lib.P_FILE_ACCESS = g("H5P_CLS_FILE_ACCESS_ID_g")
    lib.P_DATASET_CREATE = g("H5P_CLS_DATASET_CREATE_ID_g")
    _H5 = lib
    return lib


def _ok(rc, what):
    if rc < 0:
        raise CheckpointError(what)
    return rc


class _Types:
    """Memory / file types of one session (the enum and string types are created per use)."""

    def __init__(self, h5):
        self.h5 = h5

    def bool_enum(self):
        t = _ok(self.h5.H5Tenum_create(self.h5.T["i8"]), "H5Tenum_create")
        for name, v in ((b"FALSE", 0), (b"TRUE", 1)):
            val = ctypes.c_int8(v)
            _ok(self.h5.H5Tenum_insert(t, name, byref(val)), "H5Tenum_insert")
        return t

    def string(self, nbytes):
        t = _ok(self.h5.H5Tcopy(self.h5.T["c_s1"]), "H5Tcopy")
        self.h5.H5Tset_size(t, max(1, nbytes))
        self.h5.H5Tset_strpad(t, _H5T_STR_NULLTERM)
        self.h5.H5Tset_cset(t, _H5T_CSET_UTF8)
        return t


_NP_TO_H5 = {np.dtype(np.float64): "f64", np.dtype(np.float32): "f32",
             np.dtype(np.int64): "i64", np.dtype(np.int32): "i32"}


def checkpoint_chunk(shape, elsize):
    """_checkpoint_chunk (checkpoint_hdf5.jl:43-57), in the array's own (column-major) order."""
    c = [max(int(n), 1) for n in shape]
    N = len(c)
    if N == 1:
        c[0] = min(c[0], _VECTOR_CHUNK)
    elif N == 2:
        c[1] = min(max(_MATRIX_CHUNK_BYTES // (c[0] * elsize), 1), c[1])
    else:
        c[-1] = 1
    while math.prod(c) * elsize > _MAX_CHUNK_BYTES:
        i = max(k for k in range(N) if c[k] > 1)
        c[i] = -(-c[i] // 2)
    return tuple(c)


def _write_array(h5, T, root, name, a):
    segs = name.split("/")
    opened = []
    g = root
    for seg in segs[:-1]:
        b = seg.encode()
        sub = h5.H5Gopen2(g, b, 0)  # an existing group, else a new one
        g = sub if sub >= 0 else _ok(h5.H5Gcreate2(g, b, 0, 0, 0), f"create group {seg}")
        opened.append(g)
    a = np.asarray(a)
    is_bool = a.dtype == np.bool_
    buf = np.asfortranarray(a.astype(np.int8) if is_bool else a)
    dims = (_HSIZE * a.ndim)(*reversed([max(0, n) for n in a.shape]))
    chunk = (_HSIZE * a.ndim)(*reversed(checkpoint_chunk(a.shape, buf.dtype.itemsize)))
    space = _ok(h5.H5Screate_simple(a.ndim, dims, None), "H5Screate_simple")
    dcpl = _ok(h5.H5Pcreate(h5.P_DATASET_CREATE), "H5Pcreate")
    _ok(h5.H5Pset_chunk(dcpl, a.ndim, chunk), "H5Pset_chunk")
    _ok(h5.H5Pset_fletcher32(dcpl), "H5Pset_fletcher32")
    t = T.bool_enum() if is_bool else h5.T[_NP_TO_H5[buf.dtype]]
    d = _ok(h5.H5Dcreate2(g, segs[-1].encode(), t, space, 0, dcpl, 0), f"create dataset {name}")
    if buf.size:
        _ok(h5.H5Dwrite(d, t, 0, 0, 0, buf.ctypes.data_as(c_void_p)), f"write dataset {name}")
    h5.H5Dclose(d)
    if is_bool:
        h5.H5Tclose(t)
    h5.H5Pclose(dcpl)
    h5.H5Sclose(space)
    for o in reversed(opened):
        h5.H5Gclose(o)


def _write_attr(h5, T, obj, key, v):
    space = _ok(h5.H5Screate(_H5S_SCALAR), "H5Screate")
    close_t = False
    if isinstance(v, (bool, np.bool_)):
        t, buf, close_t = T.bool_enum(), ctypes.c_int8(int(v)), True
    elif isinstance(v, (int, np.integer)):
        t, buf = h5.T["i64"], c_int64(int(v))
    elif isinstance(v, np.float32):
        t, buf = h5.T["f32"], ctypes.c_float(float(v))
    elif isinstance(v, (float, np.floating)):
        t, buf = h5.T["f64"], c_double(float(v))
    else:
        raw = str(v).encode()
        t, buf, close_t = T.string(len(raw) + 1), ctypes.create_string_buffer(raw, len(raw) + 1), True
    a = _ok(h5.H5Acreate2(obj, key.encode(), t, space, 0, 0), f"create attribute {key}")
    _ok(h5.H5Awrite(a, t, byref(buf) if not isinstance(buf, ctypes.Array) else buf), f"write attribute {key}")
    h5.H5Aclose(a)
    if close_t:
        h5.H5Tclose(t)
    h5.H5Sclose(space)


def _write_scalars(h5, T, g, d):
    names = sorted(d)
    _write_attr(h5, T, g, _KEYS_ATTR, "\n".join(names))
    for name in names:
        if d[name] is not None:  # `nothing`: listed in `keys`, attribute absent
            _write_attr(h5, T, g, name, d[name])


def _package_version():
    try:
        from . import __version__
        return f"kraken_mi355x {__version__}"
    except ImportError:
        return "kraken_mi355x"


def write_checkpoint(path, snap: StateSnapshot, keep_previous=True):
    """Write `snap` as one HDF5 file (checkpoint_hdf5.jl:111-155)."""
    validate_content(snap)
    check_finite(snap)
    h5 = _libhdf5()
    T = _Types(h5)
    path = os.fspath(path)
    tmp = path + ".tmp"
    f = -1
    try:
        fapl = _ok(h5.H5Pcreate(h5.P_FILE_ACCESS), "H5Pcreate")
        _ok(h5.H5Pset_libver_bounds(fapl, _H5F_LIBVER_V110, _H5F_LIBVER_V110), "H5Pset_libver_bounds")
        f = _ok(h5.H5Fcreate(tmp.encode(), _H5F_ACC_TRUNC, 0, fapl), f"create {tmp}")
        h5.H5Pclose(fapl)
        for key, v in (("container_version", CHECKPOINT_CONTAINER_VERSION),
                       ("schema_version", snap.schema_version), ("solver", snap.solver),
                       ("cycle", snap.cycle), ("kraken_version", _package_version()),
                       ("julia_version", "none (MI355X host, Python " + sys.version.split()[0] + ")"),
                       ("created_unix", time.time())):
            _write_attr(h5, T, f, key, v)
        for cls in _ARRAY_GROUPS:
            g = _ok(h5.H5Gcreate2(f, cls.encode(), 0, 0, 0), f"create group {cls}")
            d = getattr(snap, cls)
            for name in sorted(d):
                _write_array(h5, T, g, name, d[name])
            h5.H5Gclose(g)
        for cls in _SCALAR_GROUPS:
            g = _ok(h5.H5Gcreate2(f, cls.encode(), 0, 0, 0), f"create group {cls}")
            _write_scalars(h5, T, g, getattr(snap, cls))
            h5.H5Gclose(g)
        _ok(h5.H5Fclose(f), f"close {tmp}")
        f = -1
    except BaseException as err:
        if f >= 0:
            h5.H5Fclose(f)
        if os.path.exists(tmp):
            os.remove(tmp)
        if isinstance(err, KeyboardInterrupt):
            raise
        raise CheckpointError(f"{tmp}: write failed, {path!r} left untouched: {err}") from None
    if keep_previous and os.path.exists(path):
        os.rename(path, path + ".prev")
    os.rename(tmp, path)
    return path


def _iter_names(h5, fn, obj):
    names = []

    def cb(_loc, name, _info, _data):
        names.append(name.decode())
        return 0

    cbf = _ITER_CB(cb)
    idx = _HSIZE(0)
    _ok(fn(obj, _H5_INDEX_NAME, _H5_ITER_INC, byref(idx), cbf, None), "iterate")
    return names


def _read_attr(h5, path, obj, objname, key):
    a = h5.H5Aopen(obj, key.encode(), 0)
    if a < 0:
        raise CheckpointError(f"{path}: cannot read attribute {objname}@{key}")
    try:
        t = h5.H5Aget_type(a)
        cls = h5.H5Tget_class(t)
        size = h5.H5Tget_size(t)
        if cls == _H5T_STRING:
            if h5.H5Tis_variable_str(t) > 0:
                raise CheckpointError(f"{path}: attribute {objname}@{key} is a variable-length string")
            buf = ctypes.create_string_buffer(size + 1)
            _ok(h5.H5Aread(a, t, buf), f"{path}: cannot read attribute {objname}@{key}")
            v = buf.value.decode()
        elif cls == _H5T_ENUM and size == 1:
            b = ctypes.c_int8()
            _ok(h5.H5Aread(a, t, byref(b)), f"{path}: cannot read attribute {objname}@{key}")
            v = bool(b.value)
        elif cls == _H5T_INTEGER:
            b = c_int64()
            _ok(h5.H5Aread(a, h5.T["i64"], byref(b)), f"{path}: cannot read attribute {objname}@{key}")
            v = int(b.value)
        elif cls == _H5T_FLOAT:
            if size == 4:
                b = ctypes.c_float()
                _ok(h5.H5Aread(a, h5.T["f32"], byref(b)), f"{path}: cannot read attribute {objname}@{key}")
                v = np.float32(b.value)
            else:
                b = c_double()
                _ok(h5.H5Aread(a, h5.T["f64"], byref(b)), f"{path}: cannot read attribute {objname}@{key}")
                v = float(b.value)
        else:
            raise CheckpointError(f"{path}: attribute {objname}@{key} has an unsupported type")
        h5.H5Tclose(t)
        return v
    finally:
        h5.H5Aclose(a)


def _read_dataset(h5, path, d, objname, cls):
    space = h5.H5Dget_space(d)
    t = h5.H5Dget_type(d)
    try:
        nd = h5.H5Sget_simple_extent_ndims(space)
        if nd < 1:
            raise CheckpointError(f"{path}: dataset {objname} has unsupported rank {nd}")
        dims = (_HSIZE * nd)()
        h5.H5Sget_simple_extent_dims(space, dims, None)
        shape_c = tuple(int(v) for v in dims)
        tc, ts = h5.H5Tget_class(t), h5.H5Tget_size(t)
        if tc == _H5T_ENUM and ts == 1:
            mem, np_t, is_bool = t, np.int8, True
        elif tc == _H5T_FLOAT:
            mem, np_t, is_bool = (h5.T["f32"], np.float32, False) if ts == 4 else (h5.T["f64"], np.float64, False)
        elif tc == _H5T_INTEGER and ts in (4, 8):
            mem, np_t, is_bool = (h5.T["i32"], np.int32, False) if ts == 4 else (h5.T["i64"], np.int64, False)
        else:
            raise CheckpointError(f"{path}: dataset {objname} has unsupported type")
        buf = np.empty(shape_c, dtype=np_t)  # C order of the reversed dimensions
        if buf.size and h5.H5Dread(d, mem, 0, 0, 0, buf.ctypes.data_as(c_void_p)) < 0:
            raise CheckpointError(f"{path}: cannot read dataset {objname} (damaged, truncated or "
                                  "foreign file: checksum or read failure)")
        a = buf.T  # column-major view in the writer's index order
        if is_bool:
            a = a != 0
        if cls == "series" and a.ndim != 1:
            raise CheckpointError(f"{path}: dataset {objname} must be a vector, found size {a.shape}")
        return np.asfortranarray(a)
    finally:
        h5.H5Tclose(t)
        h5.H5Sclose(space)


def _read_arrays(h5, path, g, prefix, cls, out):
    for name in _iter_names(h5, h5.H5Literate, g):
        full = name if not prefix else prefix + "/" + name
        objname = f"/{cls}/{full}"
        o = h5.H5Oopen(g, name.encode(), 0)
        if o < 0:
            raise CheckpointError(f"{path}: cannot read {objname}")
        try:
            kind = h5.H5Iget_type(o)
            if kind == _H5I_GROUP:
                _read_arrays(h5, path, o, full, cls, out)
            elif kind == _H5I_DATASET:
                out[full] = _read_dataset(h5, path, o, objname, cls)
            else:
                raise CheckpointError(f"{path}: object {objname} is neither a group nor a dataset")
        finally:
            h5.H5Oclose(o)
    return out


def _read_scalars(h5, path, f, cls):
    g = h5.H5Gopen2(f, cls.encode(), 0)
    if g < 0:
        raise CheckpointError(f"{path}: cannot read group /{cls}")
    try:
        joined = _read_attr(h5, path, g, f"/{cls}", _KEYS_ATTR)
        if not isinstance(joined, str):
            raise CheckpointError(f"{path}: attribute /{cls}@keys must be a string")
        listed = [k for k in joined.split("\n") if k]
        out = {}
        for key in listed:
            out[key] = _read_attr(h5, path, g, f"/{cls}", key) if h5.H5Aexists(g, key.encode()) > 0 else None
        for key in _iter_names(h5, h5.H5Aiterate2, g):
            if key != _KEYS_ATTR and key not in out:
                raise CheckpointError(f"{path}: attribute /{cls}@{key} is not listed in /{cls}@keys")
        return out
    finally:
        h5.H5Gclose(g)


def _open(path):
    h5 = _libhdf5()
    if not os.path.isfile(path):
        raise CheckpointError(f"{path}: no such checkpoint file")
    f = h5.H5Fopen(os.fspath(path).encode(), _H5F_ACC_RDONLY, 0)
    if f < 0:
        raise CheckpointError(f"{path}: cannot read the file (damaged, truncated or foreign file)")
    return h5, f


def _read_root(h5, path, f):
    v = _read_attr(h5, path, f, "/", "container_version")
    if v != CHECKPOINT_CONTAINER_VERSION:
        raise CheckpointError(f"{path}: container_version: expected {CHECKPOINT_CONTAINER_VERSION}, found {v!r}")
    root = {k: _read_attr(h5, path, f, "/", k) for k in
            ("solver", "schema_version", "cycle", "kraken_version", "julia_version", "created_unix")}
    if not (isinstance(root["solver"], str) and isinstance(root["schema_version"], int)
            and isinstance(root["cycle"], int)):
        raise CheckpointError(f"{path}: root attributes solver / schema_version / cycle have unexpected types")
    return root


def read_checkpoint(path) -> StateSnapshot:
    """Read a file written by write_checkpoint (checkpoint_hdf5.jl:226-241)."""
    h5, f = _open(path)
    try:
        root = _read_root(h5, path, f)
        arrays = []
        for cls in _ARRAY_GROUPS:
            g = h5.H5Gopen2(f, cls.encode(), 0)
            if g < 0:
                raise CheckpointError(f"{path}: cannot read group /{cls}")
            try:
                arrays.append(_read_arrays(h5, path, g, "", cls, {}))
            finally:
                h5.H5Gclose(g)
        scal = [_read_scalars(h5, path, f, cls) for cls in _SCALAR_GROUPS]
    finally:
        h5.H5Fclose(f)
    try:
        return StateSnapshot(root["solver"], root["schema_version"], root["cycle"], arrays[0], arrays[1], *scal)
    except CheckpointError as err:
        raise CheckpointError(f"{path}: invalid content: {err}") from None


def checkpoint_info(path) -> dict:
    """Root attributes only (checkpoint_hdf5.jl:253-261)."""
    h5, f = _open(path)
    try:
        root = _read_root(h5, path, f)
    finally:
        h5.H5Fclose(f)
    return {"container_version": CHECKPOINT_CONTAINER_VERSION, **root}


# --------------------------------------------------------------------------------------
# The lattice client
# --------------------------------------------------------------------------------------
SOLVER = "kraken_mi355x_lattice"
SCHEMA_VERSION = 1
_FACE_NAMES = ("west", "east", "south", "north", "bottom", "top")


def _identity(desc):
    d = desc
    idn = {"stencil": "d2q9" if d.stencil == L.D2Q9 else "d3q19", "nx": int(d.nx), "ny": int(d.ny),
           "nz": int(d.nz), "storage": {v: k for k, v in L.STORAGE_NAMES.items()}[d.storage],
           "collision": int(d.collision), "thermal": bool(d.thermal), "coupling": int(d.coupling),
           "forcing": int(d.forcing), "solid_mode": int(d.solid_mode), "n_bc_order": int(d.n_bc_order)}
    for n in range(6):
        idn[f"bc_order_{n}"] = int(d.bc_order[n])
    for f, name in enumerate(_FACE_NAMES):
        idn[f"{name}_kind"] = int(d.faces[f].kind)
        idn[f"{name}_thermal"] = int(d.faces[f].thermal)
    return idn


_PARAMS = ("omega", "s_plus", "s_minus", "s_e", "s_eps", "s_q", "s_nu", "omega_T", "beta_g", "T_ref")


def _parameters(desc, faces_override=None):
    p = {k: float(getattr(desc, k)) for k in _PARAMS}
    for c in range(3):
        p[f"force_{c}"] = float(desc.force[c])
    for f, name in enumerate(_FACE_NAMES):
        F = (faces_override or {}).get(f, desc.faces[f])
        for c in range(3):
            p[f"{name}_u{c}"] = float(F.u[c])
        p[f"{name}_rho"] = float(F.rho)
        p[f"{name}_T"] = float(F.T)
    return p


def snapshot(lat) -> StateSnapshot:
    """Everything needed to resume `lat`: populations (storage elements verbatim), the step
    counter as the cycle, the configuration, and the state set through the setters."""
    fields = {}
    raw = lat.download_raw("f")
    fields["f"] = raw.astype(np.int32) if raw.dtype == np.uint16 else raw
    if lat.desc.thermal:
        fields["g"] = lat.download_raw("g")
    aux = lat.aux
    if "is_solid" in aux:
        fields["is_solid"] = np.asfortranarray(aux["is_solid"])
    if "cut_links" in aux:
        cell, q, qw, uw = aux["cut_links"]
        fields.update({"cut_links/cell": cell, "cut_links/q": q, "cut_links/qw": qw,
                       "cut_links/uw": np.asfortranarray(uw.T)})
    if "force_field" in aux:
        for c, a in enumerate(aux["force_field"]):
            fields[f"force_field/{'xyz'[c]}"] = a
    for fid, (pa, pb) in aux.get("profiles", {}).items():
        fields[f"profiles/{_FACE_NAMES[fid]}/a"] = pa
        if pb is not None:
            fields[f"profiles/{_FACE_NAMES[fid]}/b"] = pb
    snap = StateSnapshot(SOLVER, SCHEMA_VERSION, lat.steps, fields=fields,
                         identity=_identity(lat.desc),
                         parameters=_parameters(lat.desc, aux.get("faces")),
                         run_control={}, derived={"Q": int(lat.Q)})
    return snap


def export_state(lat) -> StateSnapshot:
    """snapshot + the platform's refusals (state.jl:367-377)."""
    snap = snapshot(lat)
    validate_content(snap)
    check_finite(snap)
    return snap


def save_checkpoint(path, lat, keep_previous=True):
    return write_checkpoint(path, export_state(lat), keep_previous=keep_previous)


def _desc_from(snap):
    idn, p = snap.identity, snap.parameters
    d = L.desc_init(L.D2Q9 if idn["stencil"] == "d2q9" else L.D3Q19, idn["nx"], idn["ny"], idn["nz"], 0.1)
    d.storage = L.STORAGE_NAMES[idn["storage"]]
    for k in ("collision", "coupling", "forcing", "solid_mode", "n_bc_order"):
        setattr(d, k, idn[k])
    d.thermal = 1 if idn["thermal"] else 0
    for n in range(6):
        d.bc_order[n] = idn[f"bc_order_{n}"]
    for k in _PARAMS:
        setattr(d, k, p[k])
    for c in range(3):
        d.force[c] = p[f"force_{c}"]
    for f, name in enumerate(_FACE_NAMES):
        F = d.faces[f]
        F.kind, F.thermal = idn[f"{name}_kind"], idn[f"{name}_thermal"]
        for c in range(3):
            F.u[c] = p[f"{name}_u{c}"]
        F.rho, F.T = p[f"{name}_rho"], p[f"{name}_T"]
    return d


def restore_lattice(snap: StateSnapshot, backend=None, desc=None):
    """Rebuild a lattice from `snap` (restore_state, state.jl:286-296).  With `desc`, the
    snapshot must match its identity (check_compatible) before anything is allocated."""
    from .lattice import Lattice

    if desc is not None:
        check_compatible(snap, solver=SOLVER, schema_version=SCHEMA_VERSION, identity=_identity(desc))
    else:
        check_compatible(snap, solver=SOLVER, schema_version=SCHEMA_VERSION, identity=snap.identity)
        desc = _desc_from(snap)
    Q = 9 if snap.identity["stencil"] == "d2q9" else 19
    shape = (snap.identity["nx"], snap.identity["ny"]) + ((snap.identity["nz"],) if Q == 19 else ())
    fs = snap.fields
    if "f" not in fs or fs["f"].shape != shape + (Q,):
        raise CheckpointError(f"fields/f: expected size {shape + (Q,)}, found "
                              f"{None if 'f' not in fs else fs['f'].shape}")
    if snap.identity["thermal"] and ("g" not in fs or fs["g"].shape != shape + (Q,)):
        raise CheckpointError(f"fields/g: expected size {shape + (Q,)}")
    lat = Lattice(desc, backend)
    try:
        if "is_solid" in fs:
            lat.set_mask(fs["is_solid"])
        if "cut_links/cell" in fs:
            lat.set_cut_links(fs["cut_links/cell"], fs["cut_links/q"], fs["cut_links/qw"],
                              np.ascontiguousarray(fs["cut_links/uw"].T))
        if "force_field/x" in fs:
            lat.set_force_field(*(fs[f"force_field/{c}"] for c in "xyz"[:len(shape)]))
        for fid, name in enumerate(_FACE_NAMES):
            if f"profiles/{name}/a" in fs:
                lat.set_face_profile(name, fs[f"profiles/{name}/a"], fs.get(f"profiles/{name}/b"))
        f = fs["f"]
        want = lat.storage_dtype("f")
        lat.upload_raw(f.astype(want) if want == np.uint16 else f, "f")
        if snap.identity["thermal"]:
            lat.upload_raw(fs["g"], "g")
        lat.steps = snap.cycle
    except Exception:
        lat.close()
        raise
    return lat


def load_checkpoint(path, backend=None, desc=None):
    return restore_lattice(read_checkpoint(path), backend, desc)
