"""The benchmark script's own bookkeeping (no GPU): steps per launch of a launch record, the
decoder's FLOP count, the traffic records it reads, and its refusal of several GPUs."""
import json
import re
import os
import subprocess
import sys
from types import SimpleNamespace

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
import bench  # noqa: E402


def test_launch_span():
    assert bench.launch_span({"tile": 0}) == 1
    assert bench.launch_span({}) == 1
    assert all(bench.launch_span({"tile": t}) == 2 for t in range(1, 6))
    assert all(bench.launch_span({"tile": t}) == 3 for t in range(6, 12))
    assert all(bench.launch_span({"tile": t}) == 2 for t in (1, 5, 12, 14))


def test_decoder_flop_count():
    # 2·(in·hidden + hidden·latent) for the MLP, 2·(2h)²·9·ci·co per convolution after a ×2 upsample
    d = SimpleNamespace(in_dim=2, hidden=256, latent_hw=16, latent_c=64, c1=64, c2=32, out_c=3)
    mlp = 2 * (2 * 256 + 256 * 16 * 16 * 64)
    conv = 2 * (32 ** 2 * 9 * 64 * 64 + 64 ** 2 * 9 * 64 * 32 + 128 ** 2 * 9 * 32 * 3)
    assert bench.decoder_forward_flop(d) == mlp + conv == 263193600


def test_traffic_records_are_consistent():
    with open(os.path.join(ROOT, "profiles", "traffic.json")) as fh:
        recs = json.load(fh)
    for key, entry in recs.items():
        # a kernel's record, or one record per launch it was measured with
        for rec in entry if isinstance(entry, list) else [entry]:
            assert os.path.isfile(os.path.join(ROOT, rec["source"])), key
            calib = rec["fetch_kib"] * 1024 * rec["fetch_scale"] + rec["write_kib"] * 1024 * rec["write_scale"]
            assert abs(calib - rec["bytes_per_launch"]) <= 1024, key
            # every multi-step record names the launch it was measured with
            if key.startswith(("k_step2", "k_step3")):
                span = 3 if key.startswith("k_step3") else 2
                assert bench.launch_span(rec["launch"]) == span, key


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2 and "one GPU" in r.stderr


def test_traffic_record_holds_for_its_launch_shape_only():
    """A counter record applies to the tile, z-ranges and store policy it was measured with."""
    with open(os.path.join(ROOT, "profiles", "traffic.json")) as fh:
        recs = json.load(fh)
    for key, entry in recs.items():
        if not key.startswith(("k_step2", "k_step3")):
            continue
        m = re.match(r"^(k_step[23]_d3q19)_(.+)_(f64|f32s|f32|f16s)_(\d+)$", key)
        assert m, key
        kernel, tag, storage, n = m.group(1), m.group(2), m.group(3), int(m.group(4))
        for rec in entry if isinstance(entry, list) else [entry]:
            shape = dict(rec["launch"])
            assert bench.traffic_record(kernel, storage, n, shape, tag=tag) is not None, key
            flipped = {**shape, "nontemporal": not shape["nontemporal"]}
            got = bench.traffic_record(kernel, storage, n, flipped, tag=tag)
            others = [r for r in (entry if isinstance(entry, list) else [entry]) if r["launch"] == flipped]
            assert got == (others[0]["bytes_per_launch"] if others else None), key
