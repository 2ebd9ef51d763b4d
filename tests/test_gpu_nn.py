"""The warm-start network's BF16-MFMA kernels against the NumPy oracle."""
import numpy as np
import pytest

from oracle import nn_oracle as no

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("M,N,K,splits,trans", [(64, 64, 16, 1, False), (37, 70, 45, 1, True),
                                                (1, 300, 2, 1, False), (130, 3, 1000, 7, True),
                                                (5, 17, 4096, 13, False), (200, 96, 777, 3, True)])
def test_gemm_bf16_mfma(kmi, backend, M, N, K, splits, trans):
    from kraken_mi355x.nn import gemm_bf16

    r = np.random.default_rng(M * 1000 + N + K)
    A = r.normal(size=(K, M)).T if trans else r.normal(size=(M, K))
    B = r.normal(size=(N, K)).T if trans else r.normal(size=(K, N))
    A = A.astype(np.float32)
    B = B.astype(np.float32)
    C = gemm_bf16(A, B, splits=splits, backend=backend)
    ref = no.bf16(A) @ no.bf16(B)
    bound = (np.abs(no.bf16(A)) @ np.abs(no.bf16(B))) * (K * 2.0 ** -23)
    assert np.all(np.abs(C - ref) <= bound + 1e-30), np.max(np.abs(C - ref) / (bound + 1e-30))


def _net(kmi, backend, **kw):
    from kraken_mi355x.nn import WarmStartNet

    return WarmStartNet(backend=backend, **kw)


def _rel(a, b):
    return float(np.linalg.norm((a - b).ravel()) / max(np.linalg.norm(b.ravel()), 1e-30))


def _grads_match(net, arch, x, t, tol=1e-2, tag=""):
    assert net.nparam == arch.nparam
    p = net.get_params().astype(np.float64)
    y = net.forward(x)
    y_ref, _ = no.forward(arch, p, x.astype(np.float64), bf=True)
    assert _rel(y, y_ref) < 2e-3, _rel(y, y_ref)
    loss, g = net.loss_grad(x, t)
    loss_ref, g_ref, _ = no.loss_grad(arch, p, x.astype(np.float64), t.astype(np.float64), bf=True)
    assert abs(loss - loss_ref) <= 1e-4 * loss_ref
    off = 0
    for name, shape in arch.shapes:
        n = int(np.prod(shape))
        e = _rel(g[off:off + n], g_ref[off:off + n])
        print(f"[nn-grad {tag}] {name} rel L2 {e:.2e}", flush=True)
        assert e < tol, (name, e)
        off += n


def test_forward_and_gradients_match_oracle(kmi, backend):
    net = _net(kmi, backend, max_batch=4, seed=7)
    r = np.random.default_rng(3)
    x = r.uniform(-1, 1, size=(3, 2)).astype(np.float32)
    t = r.normal(scale=0.3, size=(3, 128, 128, 3)).astype(np.float32)
    _grads_match(net, no.Arch(), x, t, tag="default")


@pytest.mark.parametrize("shape", [
    # every stage on the general products (channel counts not multiples of 32)
    dict(latent_hw=8, latent_c=24, c1=40, c2=16, out_c=3),
    # other direct-kernel instantiations (32 → 32, 32 → 64, 64 → 5 channels) and an 8-pixel
    # latent, whose weight gradient takes the general product
    dict(latent_hw=8, latent_c=32, c1=32, c2=64, out_c=5),
])
def test_other_channel_shapes(kmi, backend, shape):
    net = _net(kmi, backend, max_batch=3, seed=21, **shape)
    arch = no.Arch(**shape)
    r = np.random.default_rng(5)
    hw = 8 * shape["latent_hw"]
    x = r.uniform(-1, 1, size=(3, 2)).astype(np.float32)
    t = r.normal(scale=0.3, size=(3, hw, hw, shape["out_c"])).astype(np.float32)
    _grads_match(net, arch, x, t, tag=str(shape["c1"]))


def test_wide_hidden_layer_small_batch(kmi, backend):
    """hidden = 2048 with the default latent (16·16·64): the dense-2 weight gradient
    ((hidden + 1) × L = 33.6M floats) exceeds the split workspace's 16M-float floor, and at
    batch 2 the product's K is the batch, so the workspace must be sized from the gradient."""
    net = _net(kmi, backend, hidden=2048, max_batch=2, seed=9)
    r = np.random.default_rng(12)
    x = r.uniform(-1, 1, size=(2, 2)).astype(np.float32)
    t = r.normal(scale=0.3, size=(2, 128, 128, 3)).astype(np.float32)
    _grads_match(net, no.Arch(hidden=2048), x, t, tag="hidden=2048")


def test_training_reduces_loss(kmi, backend):
    net = _net(kmi, backend, max_batch=16, seed=11)
    r = np.random.default_rng(4)
    n = 64
    x = r.uniform(-1, 1, size=(n, 2)).astype(np.float32)
    yy, xx = np.meshgrid(np.linspace(0, 1, 128), np.linspace(0, 1, 128), indexing="ij")
    t = np.stack([np.sin(np.pi * (xx[None] * (1 + x[:, :1, None]))) * np.sin(np.pi * yy[None]),
                  x[:, 1:2, None] * xx[None] * yy[None],
                  0.1 * np.cos(np.pi * xx[None]) * x[:, :1, None]], axis=-1).astype(np.float32)
    net.set_dataset(x, t)
    first = net.loss_grad(x[:16], t[:16], want_grad=False)[0]
    last = net.train_epochs(60, 16, 2e-3, seed=5)
    print(f"[nn-train] loss {first:.4e} -> {last:.4e}", flush=True)
    assert last < first / 5


def test_create_rejects_shapes_beyond_32_bit_indices(kmi, backend):
    from kraken_mi355x.nn import WarmStartNet

    with pytest.raises(kmi.KrakenError):
        WarmStartNet(latent_hw=512, max_batch=64, backend=backend)
    with pytest.raises(kmi.KrakenError):
        WarmStartNet(latent_c=12, backend=backend)  # not a multiple of 8
