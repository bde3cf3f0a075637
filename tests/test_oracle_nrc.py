"""Pins of the CPU restatement of the learned radiance cache (oracle/nrc.cpp).

Nothing in the reference corresponds to this part, so every check here comes from the
mathematics: hash constants and XOR structure, orthonormality of the spherical
harmonics, trilinear partition of unity and linear precision, finite differences of the
loss for every gradient, and Adam's closed-form first step."""
import math

import numpy as np
import pytest

import oracle

LEVELS, FEATS = 12, 2


def _res(levels=LEVELS, nmin=16, nmax=2048):
    b = math.exp((math.log(nmax) - math.log(nmin)) / (levels - 1))
    return [int(math.floor(nmin * b ** l)) for l in range(levels)] + [0] * (16 - levels)


def _samples(n, rng, lo=(0, 0, 0), hi=(64, 64, 64)):
    pos = rng.uniform(lo, hi, (n, 3)).astype(np.float32)
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    nrm = rng.normal(size=(n, 3))
    nrm /= np.linalg.norm(nrm, axis=1, keepdims=True)
    alb = rng.uniform(0.1, 0.9, (n, 3)).astype(np.float32)
    return pos, d.astype(np.float32), nrm.astype(np.float32), alb


def _net(rng, log2_T=10, scale=0.3, table_scale=0.1, loss=1, lr=1e-2, res=None):
    w = np.zeros(oracle.N_WEIGHTS, np.float32)
    w[: 4 * 64 * 64] = rng.uniform(-scale, scale, 4 * 64 * 64)
    w5 = rng.uniform(-scale, scale, (64, 16)).astype(np.float32)
    w5[:, 3:] = 0
    w[4 * 64 * 64:] = w5.ravel()
    t = rng.uniform(-table_scale, table_scale, LEVELS * (1 << log2_T) * FEATS).astype(np.float32)
    return oracle.Nrc(LEVELS, FEATS, log2_T, res or _res(), (0, 0, 0), (64, 64, 64), w, t, lr=lr, loss=loss)


def test_hash_known_answers():
    assert oracle.nrc_hash(1, 0, 0, 18) == 1
    assert oracle.nrc_hash(0, 1, 0, 18) == 2654435761 & 0x3FFFF
    assert oracle.nrc_hash(0, 0, 1, 18) == 805459861 & 0x3FFFF
    assert oracle.nrc_hash(0, 1, 0, 32) == 2654435761
    rng = np.random.default_rng(0)
    for x, y, z in rng.integers(0, 4096, (50, 3)):
        h = oracle.nrc_hash(int(x), int(y), int(z), 20)
        assert h == oracle.nrc_hash(int(x), 0, 0, 20) ^ oracle.nrc_hash(0, int(y), 0, 20) ^ oracle.nrc_hash(0, 0, int(z), 20)
        assert h == (int(x) ^ ((int(y) * 2654435761) & 0xFFFFFFFF) ^ ((int(z) * 805459861) & 0xFFFFFFFF)) & 0xFFFFF


def test_sh_orthonormal():
    """Gram matrix of the 16 functions over the sphere (Gauss-Legendre in cos(theta),
    uniform in phi) is the identity: pins every constant up to sign; the l = 0 and l = 1
    terms are checked by value."""
    xg, wg = np.polynomial.legendre.leggauss(24)
    nphi = 48
    phis = (np.arange(nphi) + 0.5) * 2 * np.pi / nphi
    gram = np.zeros((16, 16))
    for ct, wt in zip(xg, wg):
        st = math.sqrt(1 - ct * ct)
        for ph in phis:
            y = oracle.nrc_sh16([st * math.cos(ph), st * math.sin(ph), ct])
            gram += np.outer(y, y) * wt * (2 * np.pi / nphi)
    assert np.abs(gram - np.eye(16)).max() < 1e-12
    y = oracle.nrc_sh16([0.0, 0.0, 1.0])
    assert y[0] == pytest.approx(0.5 / math.sqrt(math.pi), rel=1e-15)
    assert y[2] == pytest.approx(math.sqrt(3 / (4 * math.pi)), rel=1e-15)
    assert y[1] == 0 and y[3] == 0


def test_grid_partition_of_unity_and_linear_precision():
    rng = np.random.default_rng(1)
    log2_T = 14
    res = _res()
    T = 1 << log2_T
    pos, d, n, a = _samples(200, rng)
    # all-ones table: every grid feature is the sum of the trilinear weights = 1
    w = np.zeros(oracle.N_WEIGHTS, np.float32)
    net = oracle.Nrc(LEVELS, FEATS, log2_T, res, (0, 0, 0), (64, 64, 64), w, np.ones(LEVELS * T * FEATS, np.float32))
    enc = net.encode(pos, d, n, a)
    assert np.abs(enc[:, :24] - 1).max() < 1e-12
    # dense levels: table = corner x coordinate (feature 0), z (feature 1) -> the
    # interpolated features are u_x * N and u_z * N exactly (linear precision)
    tab = np.zeros((LEVELS, T, FEATS), np.float32)
    dense = [l for l in range(LEVELS) if (res[l] + 1) ** 3 <= T]
    assert dense, "no dense level at this table size"
    for l in dense:
        N = res[l]
        idx = np.arange((N + 1) ** 3)
        tab[l, idx, 0] = idx % (N + 1)
        tab[l, idx, 1] = idx // ((N + 1) ** 2)
    net2 = oracle.Nrc(LEVELS, FEATS, log2_T, res, (0, 0, 0), (64, 64, 64), w, tab.ravel())
    enc = net2.encode(pos, d, n, a)
    for l in dense:
        u = pos.astype(np.float64) / 64.0
        np.testing.assert_allclose(enc[:, 2 * l], u[:, 0] * res[l], rtol=0, atol=1e-9)
        np.testing.assert_allclose(enc[:, 2 * l + 1], u[:, 2] * res[l], rtol=0, atol=1e-9)
    # the rest of the encoding: SH of direction and normal, albedo, 1, zeros
    assert np.allclose(enc[:, 24:40], np.stack([oracle.nrc_sh16(v) for v in d.astype(np.float64)]), atol=1e-7)
    assert np.allclose(enc[:, 40:56], np.stack([oracle.nrc_sh16(v) for v in n.astype(np.float64)]), atol=1e-7)
    assert np.array_equal(enc[:, 56:59], a.astype(np.float64))
    assert np.all(enc[:, 59] == 1) and np.all(enc[:, 60:] == 0)


def _fd_check(net, inputs, target, loss_fn, idx_w, idx_t, h=1e-5):
    _, gw, gt = net.grads(*inputs, target)
    w0, t0 = net.get()
    for k in idx_w:
        wp, wm = w0.copy(), w0.copy()
        wp[k] += h
        wm[k] -= h
        net.set(weights=wp)
        lp = loss_fn()
        net.set(weights=wm)
        lm = loss_fn()
        net.set(weights=w0)
        fd = (lp - lm) / (2 * h)
        assert fd == pytest.approx(gw[k], rel=1e-5, abs=1e-9), ("weight", k)
    for k in idx_t:
        tp, tm = t0.copy(), t0.copy()
        tp[k] += h
        tm[k] -= h
        net.set(table=tp)
        lp = loss_fn()
        net.set(table=tm)
        lm = loss_fn()
        net.set(table=t0)
        fd = (lp - lm) / (2 * h)
        assert fd == pytest.approx(gt[k], rel=1e-5, abs=1e-9), ("table", k)


@pytest.mark.parametrize("loss", [1, 0])
def test_gradients_finite_differences(loss):
    """Every weight-matrix and grid gradient equals the central difference of the loss.
    For the relative L2 loss the denominator is a stop-gradient, so the difference is
    taken of the loss with the denominator frozen at the unperturbed predictions."""
    rng = np.random.default_rng(2)
    net = _net(rng, loss=loss)
    inputs = _samples(6, rng)
    target = rng.uniform(0, 2, (6, 3)).astype(np.float32)
    y0 = net.forward(*inputs)
    den = y0 ** 2 + 0.01 if loss == 0 else np.ones_like(y0)

    def loss_fn():
        y = net.forward(*inputs)
        return float(np.mean((y - target) ** 2 / den))

    _, gw, gt = net.grads(*inputs, target)
    # weights with a nonzero gradient in every layer, padding columns of layer 5 excluded
    idx_w = []
    for layer in range(5):
        off = layer * 64 * 64
        cols = 64 if layer < 4 else 16
        seg = np.arange(off, off + 64 * cols)
        if layer == 4:
            seg = seg[(seg - off) % 16 < 3]
        nz = seg[np.abs(gw[seg]) > 1e-6]
        idx_w += list(rng.choice(nz, min(8, len(nz)), replace=False))
    nz_t = np.flatnonzero(np.abs(gt) > 1e-7)
    idx_t = list(rng.choice(nz_t, 24, replace=False))
    _fd_check(net, inputs, target, loss_fn, idx_w, idx_t)
    # padding columns carry no gradient
    g5 = gw[4 * 64 * 64:].reshape(64, 16)
    assert np.all(g5[:, 3:] == 0)


def test_adam_first_step_closed_form():
    """At t = 1 the bias-corrected moments are g and g^2, so each parameter moves by
    lr * g / (|g| + eps): lr against the sign of its gradient."""
    rng = np.random.default_rng(3)
    net = _net(rng, loss=0, lr=1e-2)
    inputs = _samples(16, rng)
    target = rng.uniform(0, 2, (16, 3)).astype(np.float32)
    _, gw, gt = net.grads(*inputs, target)
    w0, t0 = net.get()
    net.train_step(*inputs, target)
    w1, t1 = net.get()
    lr, eps = float(np.float32(1e-2)), float(np.float32(1e-10))  # the config holds float32
    exp_w = w0 - lr * gw / (np.abs(gw) + eps)
    exp_t = t0 - lr * gt / (np.abs(gt) + eps)
    np.testing.assert_allclose(w1, exp_w, rtol=0, atol=1e-12)
    np.testing.assert_allclose(t1, exp_t, rtol=0, atol=1e-12)


def test_training_reduces_loss():
    rng = np.random.default_rng(4)
    net = _net(rng, scale=0.15, table_scale=1e-4, loss=0, lr=1e-2)
    pos, d, n, a = _samples(256, rng)
    target = (1.0 + np.sin(pos[:, :1] / 8.0) * a).astype(np.float32)
    first = net.train_step(pos, d, n, a, target)
    for _ in range(60):
        last = net.train_step(pos, d, n, a, target)
    assert last < 0.1 * first


def test_bf16_emulation_close_to_fp64():
    rng = np.random.default_rng(5)
    net = _net(rng, scale=0.2)
    w, t = net.get()
    res = _res()
    em = oracle.Nrc(LEVELS, FEATS, 10, res, (0, 0, 0), (64, 64, 64), w.astype(np.float32), t.astype(np.float32),
                    emulate_bf16=True)
    inputs = _samples(64, rng)
    y64, y16 = net.forward(*inputs), em.forward(*inputs)
    assert np.abs(y16 - y64).max() <= 0.05 * np.abs(y64).max()


def test_fp16_rounding_matches_numpy():
    """The oracle's binary16 emulation (double -> float -> half, round to nearest even,
    overflow to infinity, subnormals) against numpy's float16 conversion, an independent
    implementation, on special values and 20000 random magnitudes over the whole range;
    bf16 against its bit-level definition (round the float's upper 16 bits, ties to even)."""
    import oracle
    special = [0.0, -0.0, 1.0, -1.0, 1 + 2 ** -11, 1 + 3 * 2 ** -11, 65504.0, 65519.99, 65520.0, -70000.0,
               2 ** -24, 2 ** -25, 1.5 * 2 ** -25, 2 ** -14, 2 ** -14 - 2 ** -25, 6e-8, 1e-3, 3.14159, 1e5]
    rng = np.random.default_rng(0)
    vals = special + list(rng.choice([-1, 1], 20000) * np.exp(rng.uniform(np.log(1e-9), np.log(1e6), 20000)))
    for v in vals:
        with np.errstate(over="ignore"):
            want = float(np.float16(np.float32(v)))
        got = oracle.nrc_round16(v, 2)
        assert (np.isinf(want) and got == want) or got == want, (v, got, want)
    for v in vals[:2000]:
        u = int(np.array([v], np.float32).view(np.uint32)[0])
        u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
        want = float(np.array([u], np.uint32).view(np.float32)[0])
        assert oracle.nrc_round16(v, 1) == want, v
