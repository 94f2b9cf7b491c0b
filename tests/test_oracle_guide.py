"""Pins the oracle's learned directional proposal (no counterpart in the reference):

* its exp and log against numpy's float64 functions (a few ulp over the range the mixture uses,
  special values exact);
* the mixture's density: integrates to 1 over the sphere (quadrature), and equals an
  independent numpy evaluation from the raw network outputs (softmax weights, concentrations
  through the clamp, mean directions through the sigmoids);
* the mixture's sampler: the histogram of its directions matches the density (chi-square), for
  broad and concentrated lobes;
* the lobe conversion's invariants (weights sum to 1, cdf monotone ending at 1, unit means);
* the loss gradient against central finite differences of the loss, through all three layers
  and the hash grid, and masked samples leave it unchanged.
"""
import numpy as np
import pytest

import orc
from test_oracle_mlp import BASE, HI, LO, LOG2T, MAXR


def test_exp_log_against_numpy():
    g = np.random.default_rng(1)
    x = np.concatenate([g.uniform(-86, 88, 200000), g.uniform(-1, 1, 100000), g.uniform(-1e-3, 1e-3, 1000)])
    x = x.astype(np.float32)
    got = orc.explog(x).astype(np.float64)
    want = np.exp(x.astype(np.float64))
    rel = np.abs(got - want) / want
    assert rel.max() < 4e-7, rel.max()
    y = np.concatenate([np.exp(g.uniform(-80, 80, 200000)), g.uniform(0.5, 2.0, 100000), [1.0]]).astype(np.float32)
    got = orc.explog(y, log=True).astype(np.float64)
    want = np.log(y.astype(np.float64))
    err = np.abs(got - want)
    assert (err <= 4e-7 * np.maximum(np.abs(want), 1e-2)).all(), err.max()
    special = orc.explog(np.array([0.0, -100.0, 100.0, np.nan], np.float32))
    assert special[0] == 1.0 and special[1] == 0.0 and np.isinf(special[2]) and np.isnan(special[3])
    special = orc.explog(np.array([1.0, 0.0, -1.0, np.inf], np.float32), log=True)
    assert special[0] == 0.0 and special[1] == -np.inf and np.isnan(special[2]) and special[3] == np.inf


def octants():
    k = np.arange(8)
    return np.stack([np.where(k & 1, -1.0, 1.0), np.where(k & 2, -1.0, 1.0), np.where(k & 4, -1.0, 1.0)], 1) / np.sqrt(3)


def numpy_mixture(z):
    """(pi, kappa, mu) of raw outputs z [32], restated independently: the outputs offset so that zero
    outputs give lobes of concentration 4 at the octant directions."""
    z = np.asarray(z, np.float64).reshape(8, 4)
    o = octants()
    th0 = np.arccos(o[:, 2])
    ph0 = np.mod(np.arctan2(o[:, 1], o[:, 0]), 2 * np.pi)
    a = np.exp(z[:, 0] - z[:, 0].max())
    pi = a / a.sum()
    kap = np.exp(0.5 + 4.5 * np.tanh((z[:, 1] + np.log(4) - 0.5) / 4.5))
    th = np.pi / (1 + np.exp(-(z[:, 2] + np.log(th0 / (np.pi - th0)))))
    ph = 2 * np.pi / (1 + np.exp(-(z[:, 3] + np.log(ph0 / (2 * np.pi - ph0)))))
    mu = np.stack([np.sin(th) * np.cos(ph), np.sin(th) * np.sin(ph), np.cos(th)], 1)
    return pi, kap, mu


def numpy_pdf(pi, kap, mu, w):
    c = kap / (2 * np.pi * (1 - np.exp(-2 * kap)))
    return (pi * c * np.exp(kap * (w @ mu.T - 1))).sum(-1)


def sphere_grid(nt=720, nphi=1440):
    """Midpoint quadrature of the sphere in (cos theta, phi): directions and equal weights."""
    ct = -1 + (np.arange(nt) + 0.5) * (2.0 / nt)
    phi = (np.arange(nphi) + 0.5) * (2 * np.pi / nphi)
    st = np.sqrt(1 - ct ** 2)
    w = np.stack(np.broadcast_arrays(st[:, None] * np.cos(phi)[None], st[:, None] * np.sin(phi)[None],
                                     ct[:, None] + 0 * phi[None]), -1).reshape(-1, 3)
    return w, 4 * np.pi / (nt * nphi)


def raw_outputs(seed, log_kappa):
    """Raw outputs whose lobes have the given log concentrations (inverting the soft clamp)."""
    g = np.random.default_rng(seed)
    z = g.normal(0, 1.2, (8, 4))
    z[:, 1] = 0.5 + 4.5 * np.arctanh((np.asarray(log_kappa) - 0.5) / 4.5) - np.log(4)
    return z.reshape(-1).astype(np.float32)


def test_zero_outputs_give_octant_lobes():
    lobes = orc.guide_lobes(np.zeros((1, 32), np.float32))[0]
    assert np.allclose(lobes[0::2, :3], octants(), atol=1e-6)
    assert np.allclose(lobes[0::2, 3], 4.0, rtol=2e-2) and np.allclose(lobes[1::2, 3], 1 / 8)


@pytest.mark.parametrize("log_kappa", [np.linspace(-3.5, 3, 8), np.linspace(0.5, 3.5, 8)])
def test_mixture_density_normalised_and_independent(log_kappa):
    z = raw_outputs(5, log_kappa)
    lobes = orc.guide_lobes(z[None])[0]
    dirs, dw = sphere_grid()
    pdf, _ = orc.mixture(lobes, dirs.astype(np.float32), np.zeros((0, 3)))
    assert abs(pdf.astype(np.float64).sum() * dw - 1.0) < 2e-3
    pi, kap, mu = numpy_mixture(z)
    want = numpy_pdf(pi, kap, mu, dirs)
    np.testing.assert_allclose(pdf, want, rtol=2e-5, atol=1e-7)


def test_lobe_conversion_invariants():
    g = np.random.default_rng(9)
    raw = g.normal(0, 3, (500, 32)).astype(np.float32)
    L = orc.guide_lobes(raw).astype(np.float64)
    mu, kap = L[:, 0::2, :3], L[:, 0::2, 3]
    c, e2, cdf, pi = L[:, 1::2, 0], L[:, 1::2, 1], L[:, 1::2, 2], L[:, 1::2, 3]
    assert np.allclose(np.linalg.norm(mu, axis=-1), 1, atol=1e-6)
    assert (kap >= np.exp(-4) * (1 - 1e-6)).all() and (kap <= np.exp(5) * (1 + 1e-6)).all() and kap.std() > 1
    assert np.allclose(pi.sum(1), 1, atol=1e-6) and (np.diff(cdf, axis=1) >= 0).all() and (cdf[:, -1] == 1).all()
    assert np.allclose(cdf, np.cumsum(pi, 1), atol=1e-6)
    # e^{-2 kappa} of the unrounded kappa: relative error up to 2 kappa x float rounding of kappa
    assert (np.abs(e2 - np.exp(-2 * kap)) <= 2e-7 * (1 + 2 * kap) * np.exp(-2 * kap) + 1e-37).all()
    assert np.allclose(c, pi * kap / (2 * np.pi * (1 - np.exp(-2 * kap))), rtol=1e-5)


@pytest.mark.parametrize("log_kappa", [np.linspace(-3, 2, 8), np.linspace(3, 4.9, 8)], ids=["broad", "sharp"])
def test_mixture_sampler_matches_density(log_kappa):
    z = raw_outputs(7, log_kappa)
    lobes = orc.guide_lobes(z[None])[0]
    g = np.random.default_rng(3)
    n = 400000
    _, w = orc.mixture(lobes, np.zeros((0, 3)), g.random((n, 3)).astype(np.float32))
    assert np.allclose(np.linalg.norm(w, axis=1), 1, atol=1e-5)
    # bins: 24 x 48 in (cos theta, phi); expected counts from the independent density
    nt, nphi = 24, 48
    ti = np.clip(((w[:, 2] + 1) / 2 * nt).astype(int), 0, nt - 1)
    pj = np.clip((np.mod(np.arctan2(w[:, 1], w[:, 0]), 2 * np.pi) / (2 * np.pi) * nphi).astype(int), 0, nphi - 1)
    counts = np.bincount(ti * nphi + pj, minlength=nt * nphi).astype(np.float64)
    dirs, dw = sphere_grid(nt * 20, nphi * 20)
    pi, kap, mu = numpy_mixture(z)
    dens = numpy_pdf(pi, kap, mu, dirs).reshape(nt, 20, nphi, 20).sum((1, 3)).reshape(-1) * dw
    expected = dens / dens.sum() * n
    big = expected > 20
    chi2 = (((counts - expected) ** 2) / expected)[big].sum()
    dof = big.sum() - 1
    assert chi2 < dof + 6 * np.sqrt(2 * dof), (chi2, dof)
    assert counts[~big].sum() <= expected[~big].sum() + 6 * np.sqrt(expected[~big].sum() + 1) + 10


@pytest.fixture(scope="module")
def guide_net():
    m = orc.OracleGuide(LOG2T, BASE, MAXR, LO, HI)
    g = np.random.default_rng(21)
    T = 1 << LOG2T
    grid = g.uniform(-0.5, 0.5, 16 * T * 2).astype(np.float32)
    w = np.concatenate([g.uniform(-1, 1, 32 * 64) * np.sqrt(6 / 96), g.uniform(-1, 1, 64 * 64) * np.sqrt(6 / 128),
                        g.uniform(-1, 1, 64 * 32) * np.sqrt(6 / 96)]).astype(np.float32)
    pos = g.uniform(LO, HI, (300, 3)).astype(np.float32)
    d = g.normal(0, 1, (300, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    tgt = np.concatenate([d, np.exp(g.normal(0, 1, (300, 1)))], 1).astype(np.float32)
    tgt[::9, 3] = -1.0   # masked
    tgt[1::11, 3] = 0.0  # zero weight: masked
    return m, grid, w, pos, tgt


def test_guide_forward_fp16_close_to_exact(guide_net):
    m, grid, w, pos, _ = guide_net
    exact = m.forward(grid, w, pos, mode=0)
    half = m.forward(grid, w, pos, mode=1)
    assert np.abs(half - exact).max() < 2e-2 * max(1.0, np.abs(exact).max())


def test_guide_gradient_matches_finite_differences(guide_net):
    m, grid, w, pos, tgt = guide_net
    loss, gg, gw = m.gradients(grid, w, pos, tgt, mode=0)
    valid = tgt[:, 3] > 0
    wmean = float(tgt[valid, 3].astype(np.float64).mean())
    assert abs(m.loss(grid, w, pos, tgt, wmean) - loss) < 1e-12 * max(1.0, abs(loss))
    rng = np.random.default_rng(4)

    def fd(params, which, i, eps):
        a = params.copy(); a[i] += np.float32(eps)
        b = params.copy(); b[i] -= np.float32(eps)
        h = float(a[i]) - float(b[i])
        if which == "w":
            return (m.loss(grid, a, pos, tgt, wmean) - m.loss(grid, b, pos, tgt, wmean)) / h
        return (m.loss(a, w, pos, tgt, wmean) - m.loss(b, w, pos, tgt, wmean)) / h

    for lo, hi in ((0, 2048), (2048, 6144), (6144, 8192)):
        for i in rng.integers(lo, hi, 6):
            num = fd(w, "w", int(i), 2e-5)
            assert abs(num - gw[i]) <= 1e-3 * abs(gw[i]) + 1e-8, (int(i), num, gw[i])
    touched = np.nonzero(gg)[0]
    assert len(touched) > 100
    for i in rng.choice(touched, 10, replace=False):
        num = fd(grid, "g", int(i), 2e-5)
        assert abs(num - gg[i]) <= 1e-3 * abs(gg[i]) + 1e-8, (int(i), num, gg[i])


def test_guide_masked_samples_do_not_contribute(guide_net):
    m, grid, w, pos, tgt = guide_net
    keep = tgt[:, 3] > 0
    l1, gg1, gw1 = m.gradients(grid, w, pos, tgt, mode=0)
    l2, gg2, gw2 = m.gradients(grid, w, pos[keep], tgt[keep], mode=0)
    assert abs(l1 - l2) < 1e-12 * max(1, abs(l1)) and np.allclose(gw1, gw2, rtol=1e-12, atol=0)
    assert np.allclose(gg1, gg2, rtol=1e-12, atol=0)
