"""Pins the oracle's learned target function (hash-grid encoding + MLP + relative L2
loss gradient): level resolutions in closed form, the forward pass against an
independent vectorised numpy implementation, and the analytic gradient against
central finite differences of the loss."""
import numpy as np
import pytest

import orc

LO, HI = (-1.0, -2.0, 0.5), (3.0, 2.0, 4.5)
LOG2T, BASE, MAXR = 12, 4, 96   # small table: coarse levels dense, fine levels hashed


@pytest.fixture(scope="module")
def net():
    m = orc.OracleMlp(LOG2T, BASE, MAXR, LO, HI)
    g = np.random.default_rng(11)
    T = 1 << LOG2T
    grid = g.uniform(-0.5, 0.5, 16 * T * 2).astype(np.float32)
    w = np.concatenate([g.uniform(-1, 1, 32 * 64) * np.sqrt(6 / 96), g.uniform(-1, 1, 64 * 64) * np.sqrt(6 / 128),
                        g.uniform(-1, 1, 64 * 3) * np.sqrt(6 / 67)]).astype(np.float32)
    pos = g.uniform(LO, HI, (400, 3)).astype(np.float32)
    pos[:4] = HI          # upper faces: the cell clamp puts them on the last corner
    pos[4:8] = LO
    pos[8:12] = np.array(HI) + 1.0  # outside: clamped to the box
    target = np.exp(g.normal(0, 1, (400, 3))).astype(np.float32)
    target[::7] = -1.0  # masked samples
    return m, grid, w, pos, target


def test_level_resolutions_closed_form():
    m = orc.OracleMlp(19, 16, 2048, (0, 0, 0), (1, 1, 1))
    res, dense = m.levels()
    b = np.exp((np.log(2048) - np.log(16)) / 15)
    want = np.floor(16 * b ** np.arange(16) + 1e-9).astype(np.uint32)
    assert np.array_equal(res, want)
    assert res[0] == 16 and res[15] == 2048
    dense_want = sum(1 << l for l in range(16) if (int(res[l]) + 1) ** 3 <= 1 << 19)
    assert dense == dense_want and dense & 1 and not dense >> 15


def numpy_forward(grid, w, pos, res, dense, T):
    """Hash-grid encoding + MLP, vectorised over samples (float64 sums of the float32 weights)."""
    lo, hi = np.array(LO, np.float32), np.array(HI, np.float32)
    inv = (np.float32(1.0) / (hi - lo)).astype(np.float32)
    u = np.clip((pos - lo) * inv, np.float32(0), np.float32(1)).astype(np.float32)
    enc = np.zeros((len(pos), 32))
    for l in range(16):
        N = np.float32(res[l])
        x = (u * N).astype(np.float32)
        f = np.minimum(np.floor(x), N - np.float32(1))
        fr = (x - f).astype(np.float32)
        ci = f.astype(np.uint64)
        for c in range(8):
            bits = np.array([c & 1, (c >> 1) & 1, (c >> 2) & 1], np.uint64)
            cc = ci + bits
            if (dense >> l) & 1:
                r = np.uint64(res[l] + 1)
                idx = cc[:, 0] + r * (cc[:, 1] + r * cc[:, 2])
            else:
                h = (cc[:, 0] ^ ((cc[:, 1] * np.uint64(2654435761)) & np.uint64(0xFFFFFFFF))
                     ^ ((cc[:, 2] * np.uint64(805459861)) & np.uint64(0xFFFFFFFF)))
                idx = h & np.uint64(T - 1)
            wt = np.where(bits.astype(bool), fr, np.float32(1) - fr).astype(np.float32)
            wt = ((wt[:, 0] * wt[:, 1]).astype(np.float32) * wt[:, 2]).astype(np.float32)
            slot = np.uint64(l * T) + idx
            enc[:, 2 * l] += wt.astype(np.float64) * grid[2 * slot]
            enc[:, 2 * l + 1] += wt.astype(np.float64) * grid[2 * slot + 1]
    W1 = w[:2048].reshape(32, 64).astype(np.float64)
    W2 = w[2048:6144].reshape(64, 64).astype(np.float64)
    W3 = w[6144:].reshape(64, 3).astype(np.float64)
    h1 = np.maximum(enc @ W1, 0)
    h2 = np.maximum(h1 @ W2, 0)
    return np.exp(np.clip(h2 @ W3, -20, 20))


def test_forward_matches_independent_numpy(net):
    m, grid, w, pos, _ = net
    res, dense = m.levels()
    assert dense not in (0, 0xFFFF), "the fixture should exercise dense and hashed levels"
    want = numpy_forward(grid, w, pos, res, dense, 1 << LOG2T)
    got = m.forward(grid, w, pos, mode=0)
    np.testing.assert_allclose(got, want, rtol=1e-6)


def test_fp16_mode_close_to_exact(net):
    m, grid, w, pos, _ = net
    exact = m.forward(grid, w, pos, mode=0)
    half = m.forward(grid, w, pos, mode=1)
    rel = np.abs(half - exact) / exact
    assert np.median(rel) < 5e-3 and rel.max() < 5e-2


def test_gradient_matches_finite_differences(net):
    m, grid, w, pos, target = net
    loss, gg, gw = m.gradients(grid, w, pos, target, mode=0)
    den = np.zeros((len(pos), 3))
    assert abs(m.loss(grid, w, pos, target, None, den) - loss) < 1e-12 * max(1.0, loss)
    rng = np.random.default_rng(3)

    def fd(params, which, i, eps):
        # central difference over the float32-representable step actually taken (small enough
        # that no ReLU changes side for this batch)
        a = params.copy(); a[i] += np.float32(eps)
        b = params.copy(); b[i] -= np.float32(eps)
        h = float(a[i]) - float(b[i])
        if which == "w":
            return (m.loss(grid, a, pos, target, den) - m.loss(grid, b, pos, target, den)) / h
        return (m.loss(a, w, pos, target, den) - m.loss(b, w, pos, target, den)) / h

    # weights of every layer
    for lo, hi in ((0, 2048), (2048, 6144), (6144, 6336)):
        for i in rng.integers(lo, hi, 6):
            num = fd(w, "w", int(i), 2e-5)
            assert abs(num - gw[i]) <= 1e-3 * abs(gw[i]) + 1e-8, (int(i), num, gw[i])
    # grid entries that the batch touches
    touched = np.nonzero(gg)[0]
    assert len(touched) > 100
    for i in rng.choice(touched, 12, replace=False):
        num = fd(grid, "g", int(i), 2e-5)
        assert abs(num - gg[i]) <= 1e-3 * abs(gg[i]) + 1e-8, (int(i), num, gg[i])


def test_masked_samples_do_not_contribute(net):
    m, grid, w, pos, target = net
    keep = target[:, 0] >= 0
    l1, gg1, gw1 = m.gradients(grid, w, pos, target, mode=0)
    l2, gg2, gw2 = m.gradients(grid, w, pos[keep], target[keep], mode=0)
    assert abs(l1 - l2) < 1e-12 and np.allclose(gw1, gw2, rtol=1e-12, atol=0) and np.allclose(gg1, gg2, rtol=1e-12,
                                                                                                atol=0)
