"""Pins the oracle's BSDFs (material.glsl) against closed forms and sampling statistics."""
import numpy as np

import orc
import rtpkg

rt = rtpkg.rt
L = orc.lib()


def material(kind, rough=0.5, metallic=0.0, color=(0.6, 0.4, 0.2)):
    m = rt.Material()
    m.baseColor[:] = color
    m.type = kind
    m.textureIdx = 0xFFFFFFFF
    m.roughness, m.metallic, m.ior = rough, metallic, 1.5
    return m


def unit(v):
    return v / np.linalg.norm(v, axis=-1, keepdims=True)


def random_dirs(n, g, upper=True):
    d = unit(g.normal(size=(n, 3)))
    if upper:
        d[:, 2] = np.abs(d[:, 2])
    return d.astype(np.float32)


def eval_bsdf(m, albedo, n, wo, wi):
    q = np.ascontiguousarray(np.concatenate([n, wo, wi], axis=1).astype(np.float32))
    f = np.zeros((len(q), 3), np.float32)
    pdf = np.zeros(len(q), np.float32)
    L.orc_eval_bsdf(m, np.asarray(albedo, np.float32).ctypes.data, q.ctypes.data, len(q), f.ctypes.data,
                    pdf.ctypes.data)
    return f, pdf


def sample_bsdf(m, albedo, n, wo, r):
    q = np.ascontiguousarray(np.concatenate([n, wo, r], axis=1).astype(np.float32))
    k = len(q)
    wi = np.zeros((k, 3), np.float32)
    f = np.zeros((k, 3), np.float32)
    pdf = np.zeros(k, np.float32)
    typ = np.zeros(k, np.uint32)
    ok = np.zeros(k, np.uint32)
    L.orc_sample_bsdf(m, np.asarray(albedo, np.float32).ctypes.data, q.ctypes.data, k, wi.ctypes.data, f.ctypes.data,
                      pdf.ctypes.data, typ.ctypes.data, ok.ctypes.data)
    return wi, f, pdf, typ, ok


def test_lambert_closed_form_and_sampling():
    g = np.random.default_rng(0)
    m = material(1)
    albedo = np.array([0.6, 0.4, 0.2], np.float32)
    N = 200000
    n = np.tile(np.array([[0, 0, 1]], np.float32), (N, 1))
    wo = random_dirs(N, g)
    wi = random_dirs(N, g)
    f, pdf = eval_bsdf(m, albedo, n, wo, wi)
    assert np.allclose(f, albedo / np.pi, rtol=1e-6)
    assert np.allclose(pdf, np.abs(wi[:, 2]) / np.pi, rtol=1e-6, atol=1e-7)
    # cosine-weighted sampling: cos(theta) has density 2c on [0, 1]  =>  c^2 uniform
    r = g.random((N, 3)).astype(np.float32)
    s_wi, s_f, s_pdf, typ, ok = sample_bsdf(m, albedo, n, wo, r)
    assert ok.all()
    c2 = s_wi[:, 2].astype(np.float64) ** 2
    h, _ = np.histogram(c2, bins=20, range=(0, 1))
    e = N / 20
    assert ((h - e) ** 2 / e).sum() < 60
    # estimator f cos / pdf equals the albedo for every sample (white furnace of one bounce)
    est = s_f * np.abs(s_wi[:, 2:3]) / s_pdf[:, None]
    ok = np.isfinite(est).all(axis=1)
    assert ok.mean() > 0.9999
    assert np.allclose(est[ok], albedo, rtol=1e-5)


def test_metallic_workflow_reciprocity():
    g = np.random.default_rng(1)
    N = 20000
    n = np.tile(np.array([[0, 0, 1]], np.float32), (N, 1))
    wo = random_dirs(N, g)
    wi = random_dirs(N, g)
    for rough, metal in ((0.3, 0.0), (0.7, 1.0), (0.06, 0.0)):
        m = material(2, rough, metal)
        a = np.array([0.8, 0.5, 0.3], np.float32)
        f1, _ = eval_bsdf(m, a, n, wo, wi)
        f2, _ = eval_bsdf(m, a, n, wi, wo)
        assert np.allclose(f1, f2, rtol=2e-5, atol=1e-7)


def smith_g1_ggx(cos_v, alpha):
    t2 = (1 - cos_v ** 2) / cos_v ** 2
    return 2.0 / (1.0 + np.sqrt(1.0 + alpha ** 2 * t2))


def test_gtr2_sampler_is_visible_normal_sampling():
    """GTR2Sample (material.glsl:84-101) draws half vectors with the GGX visible
    normal density D_wo(h) = G1(wo) max(0, wo.h) D(h) / (wo.n) (Heitz 2018)."""
    g = np.random.default_rng(2)
    N = 400000
    alpha = 0.35
    theta_o = 0.9
    wo = np.array([np.sin(theta_o), 0.0, np.cos(theta_o)], np.float32)
    q = np.zeros((N, 8), np.float32)
    q[:, 2] = 1.0
    q[:, 3:6] = wo
    q[:, 6:8] = g.random((N, 2))
    wh = np.zeros((N, 3), np.float32)
    L.orc_gtr2_sample(q.ctypes.data, N, alpha, wh.ctypes.data)
    ok = np.isfinite(wh).all(axis=1)
    assert ok.mean() > 0.9999
    ch = wh[ok, 2].astype(np.float64)
    # density of cos(theta_h) after integrating D_wo over phi numerically
    edges = np.linspace(0.0, 1.0, 41)
    cos_mid = np.linspace(0.0, 1.0, 4001)[1:]
    phi = np.linspace(0, 2 * np.pi, 721)[:-1]
    ct = cos_mid[:, None]
    st = np.sqrt(1 - ct ** 2)
    shape = (ct.shape[0], phi.shape[0])
    h = np.stack([st * np.cos(phi), st * np.sin(phi), np.broadcast_to(ct, shape)], -1)
    D = alpha ** 2 / (np.pi * (ct ** 2 * (alpha ** 2 - 1) + 1) ** 2)
    dot = np.maximum(0.0, h @ wo.astype(np.float64))
    dens = smith_g1_ggx(np.cos(theta_o), alpha) * dot * D / np.cos(theta_o)  # per solid angle
    marg = dens.mean(axis=1) * 2 * np.pi  # d(omega) = d(cos) d(phi)
    cdf = np.concatenate([[0], np.cumsum(marg) * (cos_mid[1] - cos_mid[0])])
    assert abs(cdf[-1] - 1.0) < 5e-3  # the visible normal density integrates to one
    grid = np.concatenate([[0.0], cos_mid])
    p_bins = np.diff(np.interp(edges, grid, cdf))
    p_bins /= p_bins.sum()
    hist, _ = np.histogram(ch, bins=edges)
    e = p_bins * ok.sum()
    m = e > 20
    chi2 = ((hist[m] - e[m]) ** 2 / e[m]).sum()
    assert chi2 < 3.0 * m.sum(), chi2


def test_gtr2_pdf_uses_schlick_g1():
    """evalPdf of the GGX lobe uses the Schlick-Smith G1 (material.glsl:59-82), not the
    exact GGX G1 the sampler draws from: the recorded reference property."""
    m = material(3, 0.5)
    wo = np.array([[np.sin(1.2), 0.0, np.cos(1.2)]], np.float32)
    n = np.array([[0, 0, 1]], np.float32)
    wi = np.array([[-np.sin(1.2), 0.0, np.cos(1.2)]], np.float32)  # mirror: h = n
    _, pdf = eval_bsdf(m, np.ones(3, np.float32), n, wo, wi)
    alpha = 0.25  # roughness^2
    D = 1.0 / (np.pi * alpha ** 2)  # GGX D at the normal: alpha^2 / (pi alpha^4)
    cos_o = np.cos(1.2)
    schlick = cos_o / (cos_o * (1 - alpha / 2) + alpha / 2)
    want = D * schlick * cos_o / cos_o / (4 * cos_o)
    assert np.isclose(pdf[0], want, rtol=1e-5)
    exact = D * smith_g1_ggx(cos_o, alpha) / (4 * cos_o)
    assert not np.isclose(pdf[0], exact, rtol=1e-3)
