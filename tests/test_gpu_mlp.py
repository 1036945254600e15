"""HIP learned-tail kernels against the FP64 MLP oracle (oracle/mlp.c) and
against what the tail reconstruction must satisfy.  Bars (max|gpu - oracle| /
max|oracle|): FP32 forward 1e-5; FP32 gradient 1e-4 (sums over the batch in
FP32); bf16x3 inference 1e-4; plain bf16 inference 3e-2 (8-bit mantissa)."""
import math
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
DIMS = [21, 256, 256, 21]  # 3 layers, width 256


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def relerr(a, b):
    return float(np.abs(np.asarray(a) - np.asarray(b)).max() / np.abs(np.asarray(b)).max())


def rand_params(rng, dims):
    import oracle as orc

    p = np.empty(orc.mlp_nparam(dims))
    o = 0
    for l in range(len(dims) - 1):
        din, dout = dims[l], dims[l + 1]
        p[o:o + dout * din] = rng.standard_normal(dout * din) * math.sqrt(2.0 / din)
        o += dout * din
        p[o:o + dout] = 0.1 * rng.standard_normal(dout)
        o += dout
    return p


@pytest.fixture(scope="module")
def h():
    import ecalj_amd

    hd = ecalj_amd.Handle(0, "f64")
    yield hd
    hd.close()


def rowmajor(h, a, dtype=np.float32):
    """DevArray holding a (B x d) row-major array (Fortran shape d x B)."""
    return h.to_dev(np.asarray(a, dtype).T)


@pytest.mark.parametrize("dims,B", [(DIMS, 300), ([5, 37, 3], 1), ([21, 64, 21], 64)])
def test_mlp_forward_grad_vs_oracle(h, orc, dims, B):
    import ecalj_amd

    say("mlp forward/grad", dims, B)
    rng = np.random.default_rng(B)
    p = rand_params(rng, dims).astype(np.float32).astype(np.float64)
    x = rng.standard_normal((B, dims[0])).astype(np.float32).astype(np.float64)
    t = rng.standard_normal((B, dims[-1])).astype(np.float32).astype(np.float64)
    n = ecalj_amd.mlp_nparam(dims)
    assert n == orc.mlp_nparam(dims) == p.size
    dp, dx, dt = h.to_dev(p.astype(np.float32)), rowmajor(h, x), rowmajor(h, t)
    dy = h.array((dims[-1], B), np.float32)
    h.mlp_forward_dev(dims, dp, dx, B, dy)
    y = dy.download().T
    assert relerr(y, orc.mlp_forward(dims, p, x)) < 1e-5
    dg = h.array(n, np.float32)
    dl = h.array(1, np.float32)
    h.mlp_grad_dev(dims, dp, dx, dt, B, dg, dl)
    loss, g = orc.mlp_loss_grad(dims, p, x, t)
    assert abs(dl.download()[0] - loss) / loss < 1e-5
    gg = dg.download()
    assert relerr(gg, g) < 1e-4
    # per layer too, so a wrong small block cannot hide behind a large one
    o = 0
    for l in range(len(dims) - 1):
        nw = dims[l + 1] * dims[l]
        assert relerr(gg[o:o + nw], g[o:o + nw]) < 1e-4
        assert relerr(gg[o + nw:o + nw + dims[l + 1]], g[o + nw:o + nw + dims[l + 1]]) < 1e-4
        o += nw + dims[l + 1]
    # deterministic: a second evaluation is bit-identical
    h.mlp_grad_dev(dims, dp, dx, dt, B, dg, dl)
    assert np.array_equal(dg.download(), gg)


def test_mlp_adam_step(h):
    rng = np.random.default_rng(3)
    n = 1000
    p, g = rng.standard_normal(n).astype(np.float32), rng.standard_normal(n).astype(np.float32)
    m, v = rng.standard_normal(n).astype(np.float32) * 0.1, rng.random(n).astype(np.float32)
    dp, dg, dm, dv = (h.to_dev(a) for a in (p, g, m, v))
    step, lr, b1, b2, eps = 7, 1e-3, 0.9, 0.999, 1e-8
    h.mlp_adam_dev(n, dp, dg, dm, dv, step, lr, b1, b2, eps)
    p64, g64, m64, v64 = (a.astype(np.float64) for a in (p, g, m, v))
    m1 = b1 * m64 + (1 - b1) * g64          # Kingma & Ba, Algorithm 1
    v1 = b2 * v64 + (1 - b2) * g64 ** 2
    p1 = p64 - lr * (m1 / (1 - b1 ** step)) / (np.sqrt(v1 / (1 - b2 ** step)) + eps)
    assert relerr(dm.download(), m1) < 1e-6 and relerr(dv.download(), v1) < 1e-6
    assert np.abs(dp.download() - p1).max() < 1e-6


def test_mlp_bf16_inference(h, orc):
    import ecalj_amd

    say("mlp bf16 inference")
    rng = np.random.default_rng(11)
    B = 517
    p = rand_params(rng, DIMS).astype(np.float32).astype(np.float64)
    x = rng.standard_normal((B, DIMS[0])).astype(np.float32).astype(np.float64)
    n = ecalj_amd.mlp_nparam(DIMS)
    dp, dx = h.to_dev(p.astype(np.float32)), rowmajor(h, x)
    hi, lo = h.array(n, np.uint16), h.array(n, np.uint16)
    h.mlp_to_bf16_dev(n, dp, hi, lo)
    # the split is exact to ~2^-16: hi + lo reproduces the FP32 weights
    up = lambda a: (a.astype(np.uint32) << 16).view(np.float32).astype(np.float64)
    assert np.abs(up(hi.download()) + up(lo.download()) - p).max() <= 2.0 ** -16 * np.abs(p).max()
    ref = orc.mlp_forward(DIMS, p, x)
    dy = h.array((DIMS[-1], B), np.float32)
    h.mlp_infer_bf16_dev(DIMS, dp, hi, lo, dx, B, dy)
    assert relerr(dy.download().T, ref) < 1e-4
    h.mlp_infer_bf16_dev(DIMS, dp, hi, None, dx, B, dy)
    e1 = relerr(dy.download().T, ref)
    assert 1e-4 < e1 < 3e-2  # plain bf16: visibly coarser, still close


def test_mlp_training_reduces_loss(h):
    """Adam on the GPU kernels fits a smooth map R^4 -> R^3 (loss falls 20x)."""
    import ecalj_amd

    dims = [4, 64, 64, 3]
    rng = np.random.default_rng(1)
    N = 2048
    x = rng.uniform(-1, 1, (N, 4))
    t = np.stack([np.sin(2 * x[:, 0]) * x[:, 1], x[:, 2] ** 2 - x[:, 3], np.cos(x[:, 0] + x[:, 3])], axis=1)
    n = ecalj_amd.mlp_nparam(dims)
    p = rand_params(rng, dims).astype(np.float32)
    dp, dx, dt = h.to_dev(p), rowmajor(h, x), rowmajor(h, t)
    dg, dm, dv, dl = h.array(n, np.float32), h.array(n, np.float32), h.array(n, np.float32), h.array(1, np.float32)
    dm.zero()
    dv.zero()
    h.mlp_grad_dev(dims, dp, dx, dt, N, dg, dl)
    l0 = float(dl.download()[0])
    B = 256
    step = 0
    for ep in range(60):
        for b0 in range(0, N, B):
            step += 1
            h.mlp_grad_dev(dims, dp, dx.ptr + b0 * dims[0] * 4, dt.ptr + b0 * dims[-1] * 4, B, dg)
            h.mlp_adam_dev(n, dp, dg, dm, dv, step, 3e-3)
    h.mlp_grad_dev(dims, dp, dx, dt, N, dg, dl)
    l1 = float(dl.download()[0])
    say("training loss", l0, "->", l1)
    assert l1 < l0 / 20


def _chi_nsd(rng, ni, nw, npair=400):
    """A negative semidefinite chi(I,J,j) = -sum_p a_p(j) z_p z_p^H (j = 0 largest)."""
    z = rng.standard_normal((ni, npair)) + 1j * rng.standard_normal((ni, npair))
    d = rng.uniform(0.2, 4.0, npair)
    nu = np.concatenate([[0.0], np.linspace(0.3, 6.0, nw - 1)])
    chi = np.zeros((ni, ni, nw), complex, order="F")
    for j in range(nw):
        a = 2 * d / (d ** 2 + nu[j] ** 2)
        chi[:, :, j] = -(z * a) @ z.conj().T / npair
    return chi


def test_tail_features_and_apply(h):
    rng = np.random.default_rng(4)
    ni, nw, iref = 37, 6, 5
    chi = _chi_nsd(rng, ni, nw)
    chi[3, 3, :] = 0.0  # a row without spectral weight: its features and tail are zero
    d = -chi[np.arange(ni), np.arange(ni), :].real   # ni x nw
    mu = rng.standard_normal(nw).astype(np.float32)
    sd = rng.uniform(0.5, 2.0, nw).astype(np.float32)
    dchi = h.to_dev(chi)
    feat, aux = h.array((nw, ni), np.float32), h.array((2, ni), np.float64)
    h.tail_features_dev(dchi, ni, nw, iref, h.to_dev(mu), h.to_dev(sd), feat, aux)
    f = feat.download().T
    ok = np.ones(ni, bool)
    ok[3] = False
    raw = np.concatenate([d[:, 1:] / d[:, :1], np.log(np.maximum(d[:, :1], 1e-300))], axis=1)
    assert np.allclose(f[ok], ((raw - mu) / sd)[ok], rtol=2e-6, atol=2e-6)
    assert np.all(f[3] == 0)
    a = aux.download().T
    assert np.allclose(a[ok, 0], d[ok, 0], rtol=1e-15) and np.allclose(a[ok, 1], d[ok, iref], rtol=1e-15)
    assert a[3, 0] == 0.0
    # reconstruction
    pred = rng.standard_normal((ni, nw)).astype(np.float32)
    mu_y = np.abs(rng.standard_normal(nw)).astype(np.float32) + 0.5
    sd_y = rng.uniform(0.05, 0.2, nw).astype(np.float32)
    out = h.array((ni, ni, nw), np.complex128)
    dpred, dmu, dsd = rowmajor(h, pred), h.to_dev(mu_y), h.to_dev(sd_y)
    h.tail_apply_dev(dchi, ni, nw, iref, dpred, dmu, dsd, aux, out)
    co = out.download()
    t = np.maximum(0.0, a[:, :1] * (mu_y.astype(np.float64) + sd_y.astype(np.float64) * pred.astype(np.float64)))
    tail = co - chi
    for j in range(nw):
        tj = tail[:, :, j]
        assert np.abs(tj - tj.conj().T).max() < 1e-13                     # Hermitian
        assert np.linalg.eigvalsh(tj).max() < 1e-12 * np.abs(tj).max()   # negative semidefinite
        assert np.allclose(-np.diag(tj).real, t[:, j], rtol=1e-12, atol=1e-300)  # diagonal = predicted tail
    # in place gives the same bits
    h.tail_apply_dev(dchi, ni, nw, iref, dpred, dmu, dsd, aux, dchi)
    assert np.array_equal(dchi.download(), co)
