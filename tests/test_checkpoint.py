"""State snapshots and the HDF5 checkpoint container (no GPU): the content rules of
src/platform/state.jl and the container of src/io/checkpoint_hdf5.jl, mirrored by
kraken_mi355x.checkpoint.

Pinned by the reference's own rules (test/platform/state_contract_*.jl restate them):
round trip of every allowed element and scalar type, refusals naming the key, non-finite
refusal leaving the previous file untouched, the `.prev` generation, Fletcher32 detection
of a flipped payload byte, the chunk shapes of _checkpoint_chunk, and check_compatible's
key-by-key identity comparison.
"""
import ctypes
import os
import re

import numpy as np
import pytest

ck = pytest.importorskip("kraken_mi355x.checkpoint")


def _snap(**kw):
    r = np.random.default_rng(0)
    fields = {
        "f": np.asfortranarray(r.standard_normal((5, 4, 3, 19))),
        "f32": np.asfortranarray(r.standard_normal((7, 6)).astype(np.float32)),
        "block003/g": np.arange(12, dtype=np.int32).reshape(3, 4, order="F"),
        "block003/h": np.arange(5, dtype=np.int64),
        "mask": np.asfortranarray(r.uniform(size=(4, 5)) > 0.5),
    }
    base = dict(solver="test_client", schema_version=3, cycle=1234, fields=fields,
                series={"drag": np.linspace(0, 1, 17), "steps": np.arange(4, dtype=np.int64)},
                scalars={"t": 0.25, "n": 7, "ok": True, "name": "cavité", "x32": np.float32(1.5)},
                identity={"N": 64, "Re": 100.0, "mode": "bgk", "absent": None},
                run_control={"max_steps": 10_000, "horizon": float("inf")},
                parameters={"nu": 0.01}, derived={"omega": 1.8})
    base.update(kw)
    return ck.StateSnapshot(**base)


def _libhdf5_or_skip():
    try:
        return ck._libhdf5()
    except ck.CheckpointError as e:
        pytest.skip(f"no HDF5 C library: {e}")


def test_round_trip(tmp_path):
    _libhdf5_or_skip()
    snap = _snap()
    p = str(tmp_path / "a.h5")
    ck.write_checkpoint(p, snap)
    back = ck.read_checkpoint(p)
    assert (back.solver, back.schema_version, back.cycle) == ("test_client", 3, 1234)
    for cls in ("fields", "series"):
        a, b = getattr(snap, cls), getattr(back, cls)
        assert sorted(a) == sorted(b)
        for k in a:
            assert b[k].dtype == a[k].dtype and b[k].shape == a[k].shape, k
            assert np.array_equal(b[k], a[k]), k
    for cls in ("scalars", "identity", "run_control", "parameters", "derived"):
        a, b = getattr(snap, cls), getattr(back, cls)
        assert sorted(a) == sorted(b), cls
        for k in a:
            assert type(b[k]) is type(a[k]) or (isinstance(a[k], np.float32) and isinstance(b[k], np.float32)), (k, b[k])
            assert b[k] == a[k] or (a[k] is None and b[k] is None), k
    info = ck.checkpoint_info(p)
    assert info["container_version"] == 1 and info["cycle"] == 1234 and info["solver"] == "test_client"


def test_layout_seen_through_the_c_library(tmp_path):
    """Root attributes, nested groups, reversed dimensions and the `keys` attribute."""
    h5 = _libhdf5_or_skip()
    p = str(tmp_path / "b.h5")
    snap = _snap()
    ck.write_checkpoint(p, snap)
    f = h5.H5Fopen(p.encode(), 0, 0)
    assert f >= 0
    try:
        for k in ("container_version", "schema_version", "solver", "cycle", "kraken_version",
                  "julia_version", "created_unix"):
            assert h5.H5Aexists(f, k.encode()) > 0, k
        d = h5.H5Dopen2(f, b"/fields/f", 0)
        sp = h5.H5Dget_space(d)
        dims = (ck._HSIZE * 4)()
        h5.H5Sget_simple_extent_dims(sp, dims, None)
        assert tuple(dims) == (19, 3, 4, 5)  # column-major (5, 4, 3, 19) seen reversed
        h5.H5Sclose(sp)
        h5.H5Dclose(d)
        g = h5.H5Gopen2(f, b"/fields/block003", 0)
        assert g >= 0
        h5.H5Gclose(g)
        ident = h5.H5Gopen2(f, b"/identity", 0)
        keys = ck._read_attr(h5, p, ident, "/identity", "keys")
        assert keys.split("\n") == sorted(snap.identity)
        assert h5.H5Aexists(ident, b"absent") == 0  # `nothing`: listed, no attribute
        h5.H5Gclose(ident)
    finally:
        h5.H5Fclose(f)


def test_keep_previous_and_refusal_leaves_file(tmp_path):
    _libhdf5_or_skip()
    p = str(tmp_path / "c.h5")
    ck.write_checkpoint(p, _snap(cycle=1))
    ck.write_checkpoint(p, _snap(cycle=2))
    assert ck.checkpoint_info(p)["cycle"] == 2 and ck.checkpoint_info(p + ".prev")["cycle"] == 1
    bad = _snap(cycle=3)
    bad.fields["f"][1, 2, 0, 4] = np.nan
    with pytest.raises(ck.CheckpointError, match=r"fields/f: nan at 1-based index \(2, 3, 1, 5\)"):
        ck.write_checkpoint(p, bad)
    assert ck.checkpoint_info(p)["cycle"] == 2 and not os.path.exists(p + ".tmp")


def test_flipped_payload_byte_is_detected(tmp_path):
    _libhdf5_or_skip()
    p = str(tmp_path / "d.h5")
    marker = np.full(4096, 1234.5678)
    ck.write_checkpoint(p, ck.StateSnapshot("x", 1, 0, fields={"m": marker}))
    raw = bytearray(open(p, "rb").read())
    at = raw.find(marker[:4].tobytes())
    assert at > 0
    raw[at + 1000] ^= 0x10
    open(p, "wb").write(bytes(raw))
    with pytest.raises(ck.CheckpointError, match="fields/m"):
        ck.read_checkpoint(p)
    open(p, "wb").write(bytes(raw[: len(raw) // 2]))
    with pytest.raises(ck.CheckpointError):
        ck.read_checkpoint(p)
    with pytest.raises(ck.CheckpointError, match="no such checkpoint"):
        ck.read_checkpoint(str(tmp_path / "missing.h5"))


@pytest.mark.parametrize("kw,needle", [
    (dict(solver=""), "the solver name is empty"),
    (dict(cycle=-1), "cycle -1 is negative"),
    (dict(fields={"a": np.zeros(3, dtype=np.uint16)}), "fields/a: dtype uint16 cannot be stored"),
    (dict(fields={"a": np.zeros(())}), "fields/a: a 0-d array is a scalar"),
    (dict(fields={"a": np.zeros(3), "a/b": np.zeros(2)}), "fields/a: the name is also the parent group of fields/a/b"),
    (dict(fields={"a//b": np.zeros(3)}), "fields/a//b: array names are"),
    (dict(series={"s": np.zeros((2, 2))}), "series/s: series entries are 1-D, this one has shape (2, 2)"),
    (dict(scalars={"keys": 1}), "scalars/keys: `keys` is the group's key-list attribute"),
    (dict(scalars={"a/b": 1}), "scalars/a/b: scalar keys are single"),
    (dict(scalars={"v": None}), "scalars/v: NoneType is not a storable scalar"),
    (dict(derived={"v": [1, 2]}), "derived/v: list is not a storable scalar"),
])
def test_content_refusals(kw, needle):
    with pytest.raises(ck.CheckpointError, match=re.escape(needle)):
        _snap(**kw)


def test_chunk_shapes():
    assert ck.checkpoint_chunk((256, 256, 256, 19), 4) == (256, 256, 256, 1)
    assert ck.checkpoint_chunk((128, 128, 9), 8) == (128, 128, 1)
    assert ck.checkpoint_chunk((100,), 8) == (100,)
    assert ck.checkpoint_chunk((200_000,), 8) == (65_536,)
    assert ck.checkpoint_chunk((1000, 1000), 8) == (1000, 131)  # ≈ 1 MiB of whole columns
    assert ck.checkpoint_chunk((0, 5), 8) == (1, 5)
    c = ck.checkpoint_chunk((2048, 2048, 512, 19), 4)  # one slab is 8 GiB: halved to ≤ 1 GiB
    assert np.prod(c) * 4 <= 1 << 30 and c[-1] == 1


def test_check_compatible():
    snap = _snap()
    idn = dict(snap.identity)
    assert ck.check_compatible(snap, solver="test_client", schema_version=3, identity=idn) is snap
    with pytest.raises(ck.CheckpointError, match="solver: the snapshot was written by"):
        ck.check_compatible(snap, solver="other", schema_version=3, identity=idn)
    with pytest.raises(ck.CheckpointError, match="schema_version: the snapshot has 3, this run reads 4"):
        ck.check_compatible(snap, solver="test_client", schema_version=4, identity=idn)
    with pytest.raises(ck.CheckpointError, match=r"identity/N: the snapshot has 64::int, the run needs 64.0::float"):
        ck.check_compatible(snap, solver="test_client", schema_version=3, identity={**idn, "N": 64.0})
    with pytest.raises(ck.CheckpointError, match="identity/new: the run needs 1::int; the snapshot lacks the key"):
        ck.check_compatible(snap, solver="test_client", schema_version=3, identity={**idn, "new": 1})
    ck.check_compatible(snap, solver="test_client", schema_version=3, identity={**idn, "new": 1},
                        identity_defaults={"new": 1})
    less = {k: v for k, v in idn.items() if k != "Re"}
    with pytest.raises(ck.CheckpointError, match="identity/Re: the snapshot has .*a key this run does not declare"):
        ck.check_compatible(snap, solver="test_client", schema_version=3, identity=less)


def test_container_bytes_follow_the_format_specification(tmp_path):
    """The reference creates its checkpoints with libver_bounds = (1.10, 1.10), which selects
    superblock v3 (checkpoint_hdf5.jl:8-13, 116): read from the file's bytes as the HDF5 file
    format specification lays them out (signature at offset 0, superblock version at byte 8,
    then the size of offsets and lengths), independent of the library."""
    _libhdf5_or_skip()
    p = str(tmp_path / "c.h5")
    ck.write_checkpoint(p, _snap())
    head = open(p, "rb").read(16)
    assert head[:8] == b"\x89HDF\r\n\x1a\n"
    assert head[8] == 3  # superblock version
    assert head[9] == 8 and head[10] == 8  # 8-byte offsets and lengths
