"""Batched cavities: each cavity of the ensemble is bit-identical to the oracle's cavity chain
(src/drivers/basic.jl:121-136 / test/test_mrt.jl:19-25) run alone."""
import numpy as np
import pytest

from test_oracle_lattice import rng_pops

pytestmark = pytest.mark.gpu

SIZES = [24, 40, 17, 64, 33, 50, 8]
NUS = [0.05, 0.02, 0.1, 0.01, 0.03, 0.08, 0.2]


@pytest.mark.parametrize("collision", ["mrt", "bgk", "trt"])
@pytest.mark.parametrize("storage,dtype", [("f64", np.float64), ("f32", np.float32)])
@pytest.mark.parametrize("lid", ["zouhe", "ladd"])
def test_batch_cavities_bitwise(kmi, po, backend, collision, storage, dtype, lid):
    steps = 37
    print(f"[batch] {collision} {storage} {lid}", flush=True)
    f0 = [rng_pops((n, n), 9, seed=10 + c, amp=0.02).astype(dtype, order="F") for c, n in enumerate(SIZES)]
    with kmi.Batch(SIZES, NUS, storage=storage, collision=collision, lid=lid, backend=backend) as b:
        for c in range(len(SIZES)):
            b.upload(c, f0[c])
        b.step(steps)
        got = [b.download(c, dtype=dtype) for c in range(len(SIZES))]
        st = b.status()
    assert np.all(st["steps"] == steps)
    for c, (n, nu) in enumerate(zip(SIZES, NUS)):
        d = kmi.cavity_desc(n, nu, 0.1, storage=storage, collision=collision, lid=lid)
        ref = f0[c].copy(order="F")
        po.run(d, ref, None, steps)
        assert np.array_equal(got[c], ref), f"cavity {c} (n={n}) max |Δ| {np.max(np.abs(got[c] - ref)):.3e}"


def test_batch_tile_shapes_and_groups_are_invisible(kmi, backend):
    f0 = [rng_pops((n, n), 9, seed=30 + c) for c, n in enumerate(SIZES)]
    outs = []
    for launch in [None, (256, 1, 0, 0), (32, 8, 16, 1), (128, 2, 8, 0), (64, 4, 1024, 1)]:
        with kmi.Batch(SIZES, NUS, storage="f32s", backend=backend) as b:
            if launch:
                b.set_launch(*launch)
            for c in range(len(SIZES)):
                b.upload(c, f0[c])
            b.step(25)
            outs.append([b.download(c) for c in range(len(SIZES))])
    for o in outs[1:]:
        for c in range(len(SIZES)):
            assert np.array_equal(o[c], outs[0][c])


@pytest.mark.parametrize("check", [50, 51])
def test_batch_run_to_residual_matches_single_lattices(kmi, backend, check):
    """Odd check intervals included: a cavity retired while the others keep stepping still
    reads back its final populations."""
    sizes = [16, 24, 20, 32]
    nus = [0.1, 0.08, 0.05, 0.12]
    tol = 1e-6
    with kmi.Batch(sizes, nus, storage="f64", collision="mrt", backend=backend) as b:
        b.run(tol, check, 200_000)
        st = b.status()
        got = [b.download(c) for c in range(len(sizes))]
    assert np.all(st["status"] == kmi._lib.CAV_CONVERGED), st
    assert np.all(st["steps"] % check == 0) and np.all(st["residual"] < tol)
    for c, (n, nu) in enumerate(zip(sizes, nus)):
        with kmi.Lattice(kmi.cavity_desc(n, nu, 0.1, collision="mrt"), backend) as lat:
            steps, r, ok = kmi.run_to_residual(lat, tol, check, 200_000)
            assert ok and steps == st["steps"][c]
            assert abs(r - st["residual"][c]) <= 1e-9 * r
            assert np.array_equal(lat.download(), got[c])


def test_batch_windowed_residual_definition(kmi, po, backend):
    """r = ‖u^k − u^(k−m)‖₂ / (m ‖u^k‖₂) with m = window = check_every."""
    sizes, nus, m = [16, 24], [0.1, 0.05], 10
    f0 = [rng_pops((n, n), 9, seed=70 + c) for c, n in enumerate(sizes)]
    with kmi.Batch(sizes, nus, storage="f64", backend=backend) as b:
        for c in range(2):
            b.upload(c, f0[c])
        b.run(1.0, m, 10 * m, window=m)  # tol 1: retires at the first check
        st = b.status()
    assert np.all(st["steps"] == m)
    for c, (n, nu) in enumerate(zip(sizes, nus)):
        u0 = po.moments(f0[c])
        f = f0[c].copy(order="F")
        po.run(kmi.cavity_desc(n, nu, 0.1, collision="mrt"), f, None, m)
        u1 = po.moments(f)
        d2 = ((u1["ux"] - u0["ux"]) ** 2 + (u1["uy"] - u0["uy"]) ** 2).sum()
        u2 = (u1["ux"] ** 2 + u1["uy"] ** 2).sum()
        assert abs(st["residual"][c] - np.sqrt(d2 / u2) / m) <= 1e-10 * st["residual"][c]


def test_batch_diverged_cavity_is_retired_alone(kmi, backend):
    sizes = [16, 24, 20]
    nus = [0.1, 0.08, 0.05]
    clean = []
    with kmi.Batch(sizes, nus, storage="f32s", backend=backend) as b:
        b.run(1e-5, 20, 50_000)
        clean = [b.download(c) for c in (0, 2)]
        st0 = b.status()
    with kmi.Batch(sizes, nus, storage="f32s", backend=backend) as b:
        f = b.download(1)
        f[3, 4, 2] = np.nan  # poison one cavity from the host
        b.upload(1, f)
        b.run(1e-5, 20, 50_000)
        st = b.status()
        assert st["status"][1] == kmi._lib.CAV_DIVERGED
        assert st["status"][0] == st["status"][2] == kmi._lib.CAV_CONVERGED
        assert np.array_equal(st["steps"][[0, 2]], st0["steps"][[0, 2]])
        assert np.array_equal(b.download(0), clean[0]) and np.array_equal(b.download(2), clean[1])


def test_batch_autotune_restores_state(kmi, backend):
    f0 = [rng_pops((n, n), 9, seed=50 + c) for c, n in enumerate(SIZES)]
    with kmi.Batch(SIZES, NUS, storage="f32", backend=backend) as b:
        for c in range(len(SIZES)):
            b.upload(c, f0[c].astype(np.float32, order="F"))
        b.step(5)
        before = [b.download(c) for c in range(len(SIZES))]
        ms = b.autotune(4)
        assert ms > 0
        assert all(np.array_equal(b.download(c), before[c]) for c in range(len(SIZES)))
        assert np.all(b.status()["steps"] == 5)


def test_batch_rejects_bad_input(kmi, backend):
    with pytest.raises(kmi.KrakenError):
        kmi.Batch([16, 1], [0.1, 0.1], backend=backend)
    with kmi.Batch([16], [0.1], backend=backend) as b:
        with pytest.raises(kmi.KrakenError):
            b.set_launch(48, 4)
        with pytest.raises(kmi.ArgumentError):
            b.upload(0, np.zeros((8, 8, 9)))


@pytest.mark.parametrize("storage,tol", [("f64", 1e-13), ("f32s", 2e-6)])
def test_batch_consistent_init_matches_oracle(kmi, po, backend, storage, tol):
    """Held cavities follow the oracle's consistent initialisation (Mei et al. 2006), the
    others plain steps; FP64 within round-off (the hold is evaluated in another order),
    F32S within its storage tolerance."""
    iters = 9
    hold = [1, 0, 1, 1, 0, 1, 1]
    f0 = []
    for c, n in enumerate(SIZES):
        r = np.random.default_rng(50 + c)
        rho = np.asfortranarray(1 + 1e-3 * r.standard_normal((n, n)))
        u, v = (np.asfortranarray(0.02 * r.standard_normal((n, n))) for _ in range(2))
        f0.append(po.equilibrium(rho, u, v))
    print(f"[batch consistent] {storage}", flush=True)
    with kmi.Batch(SIZES, NUS, storage=storage, collision="mrt", backend=backend) as b:
        for c in range(len(SIZES)):
            b.upload(c, f0[c])
        b.consistent_init(iters, hold=hold)
        got = [b.download(c) for c in range(len(SIZES))]
        st = b.status()
    assert np.all(st["steps"] == iters)
    for c, (n, nu) in enumerate(zip(SIZES, NUS)):
        d = kmi.cavity_desc(n, nu, 0.1, storage="f64", collision="mrt")
        ref = f0[c].copy(order="F")
        if hold[c]:
            po.consistent_init(d, ref, iters)
        else:
            po.run(d, ref, None, iters)
        err = np.max(np.abs(got[c] - ref))
        assert err < tol, f"cavity {c} (n={n}, hold={hold[c]}) max |Δ| {err:.3e}"
    if storage == "f64":
        # the held velocity is the uploaded one
        for c in (0, 2):
            m = po.moments(np.asfortranarray(got[c]))
            m0 = po.moments(f0[c])
            assert np.max(np.abs(m["ux"] - m0["ux"])) < 1e-15


def test_batch_consistent_init_arguments(kmi, backend):
    with kmi.Batch([8, 9], [0.1, 0.1], backend=backend) as b:
        with pytest.raises(kmi.ArgumentError):
            b.consistent_init(3, hold=[1, 0, 1])
        with pytest.raises(kmi.KrakenError):
            b.consistent_init(-1)
        b.consistent_init(0)
        assert np.all(b.status()["steps"] == 0)


@pytest.mark.parametrize("storage", ["f64", "f32", "f32s", "f16s"])
@pytest.mark.parametrize("collision,lid", [("mrt", "zouhe"), ("bgk", "ladd"), ("trt", "zouhe")])
def test_batch_two_step_launches_bit_identical(kmi, backend, storage, collision, lid):
    """Two steps per launch (step t+1 of a tile and its ring in LDS) reproduce single steps bit
    for bit on every tile shape, with uneven cavity sizes, odd step counts and persistent
    groups."""
    f0 = [rng_pops((n, n), 9, seed=70 + c, amp=0.02) for c, n in enumerate(SIZES)]
    print(f"[batch two-step] {storage} {collision} {lid}", flush=True)
    with kmi.Batch(SIZES, NUS, storage=storage, collision=collision, lid=lid, backend=backend) as b:
        for c in range(len(SIZES)):
            b.upload(c, f0[c])
        b.step(37)
        ref = [b.download(c) for c in range(len(SIZES))]
    for tw, th, groups, span in ((64, 4, 0, 2), (32, 8, 16, 2), (16, 16, 0, 2), (256, 1, 0, 2), (128, 2, 8, 2),
                                 (32, 8, 0, 3), (64, 4, 16, 3), (16, 16, 0, 4), (32, 8, 8, 4)):
        with kmi.Batch(SIZES, NUS, storage=storage, collision=collision, lid=lid, backend=backend) as b:
            b.set_launch(tw, th, groups)
            b.set_temporal(span)
            assert b.launch["steps_per_launch"] == span
            for c in range(len(SIZES)):
                b.upload(c, f0[c])
            b.step(37)
            assert np.all(b.status()["steps"] == 37)
            for c in range(len(SIZES)):
                assert np.array_equal(b.download(c), ref[c]), (tw, th, groups, span, c)


def test_batch_two_step_run_to_residual(kmi, backend):
    """Graph-replayed check blocks with two-step launches (odd and even block lengths) give
    the single-step run's steps, residuals and fields."""
    sizes, nus = [16, 24, 20, 32], [0.1, 0.08, 0.05, 0.12]
    out = []
    for span in (1, 2, 3):
        for check, window in ((50, 50), (51, 7)):
            with kmi.Batch(sizes, nus, storage="f32s", collision="mrt", backend=backend) as b:
                b.set_temporal(span)
                b.run(1e-6, check, 60_000, window=window)
                st = b.status()
                out.append((span, check, st, [b.download(c) for c in range(len(sizes))]))
    for span, check, st, f in out[2:]:  # spans 2 and 3 against span 1
        ref = next(o for o in out[:2] if o[1] == check)
        assert np.array_equal(st["steps"], ref[2]["steps"]) and np.array_equal(st["residual"], ref[2]["residual"])
        assert np.array_equal(st["status"], ref[2]["status"])
        for a, r in zip(f, ref[3]):
            assert np.array_equal(a, r)


def test_batch_autotune_considers_two_step(kmi, backend):
    sizes = [64, 96, 128]
    with kmi.Batch(sizes, [0.02, 0.03, 0.04], backend=backend) as b:
        b.step(5)
        before = [b.download(c) for c in range(len(sizes))]
        b.autotune(4)
        print(f"[batch two-step] tuned {b.launch}", flush=True)
        assert b.launch["steps_per_launch"] in (1, 2, 3, 4)
        assert all(np.array_equal(b.download(c), before[c]) for c in range(len(sizes)))
        with pytest.raises(kmi.KrakenError):
            b.set_temporal(5)
        b.set_temporal(1)
        b.set_launch(256, 1)
        with pytest.raises(kmi.KrakenError):
            b.set_temporal(3)  # a 262 x 7 region exceeds one work-group


def test_batch_retired_cavity_keeps_written_state(kmi, po, backend):
    """A cavity that is no longer running reads back what was last written to it, however
    many steps the running cavities take afterwards (odd counts included)."""
    with kmi.Batch([8, 16], [0.2, 0.01], storage="f64", collision="mrt", backend=backend) as b:
        b.run(1e-3, 9, 300, window=3)  # the small viscous cavity converges first
        st = b.status()
        assert st["status"][0] == kmi._lib.CAV_CONVERGED
        assert st["status"][1] == kmi._lib.CAV_MAX_STEPS, st
        f0 = rng_pops((8, 8), 9, seed=91, amp=0.02)
        b.upload(0, f0)
        b.resume()  # cavity 1 (max_steps) runs on, cavity 0 stays retired
        b.step(7)
        assert np.array_equal(b.download(0), f0)
        b.init_equilibrium(0)
        b.step(3)
        assert np.array_equal(b.download(0), po.rest((8, 8)))


def test_single_lattice_windowed_residual_matches_batch(kmi, backend):
    """kmi_step_residual_window follows the batch's stop rule: the same cavity reaches
    r < tol after the same number of steps alone and in a batch (F32S, m = 100)."""
    n, nu, tol, check, window = 48, 0.05, 1e-6, 200, 100
    with kmi.Batch([n], [nu], storage="f32s", collision="mrt", backend=backend) as b:
        b.run(tol, check, 200_000, window=window)
        st = b.status()
    assert st["status"][0] == kmi._lib.CAV_CONVERGED
    with kmi.Lattice(kmi.cavity_desc(n, nu, 0.1, storage="f32s", collision="mrt"), backend) as lat:
        steps, r = 0, 1.0
        while r >= tol:
            r = lat.step_residual(check, window)
            steps += check
        assert steps == st["steps"][0]
        assert abs(r - st["residual"][0]) <= 1e-6 * st["residual"][0]
        with pytest.raises(kmi.KrakenError):
            lat.step_residual(10, 20)
