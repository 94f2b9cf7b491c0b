"""Pins the oracle's ReSTIR GI temporal pass (gi_resample_temporal.glsl) against what the mathematics
fixes:
  - without reuse (the clear flag on every frame) the pass is an unbiased estimator of the indirect
    radiance with NEE + MIS, the same integral the GRIS estimator without reuse estimates on a
    Lambertian scene (reconnection at the first secondary vertex): the image means over seeds agree;
  - on the floor under a quad light there is no indirect light (the floor cannot see itself and the
    light does not reflect): the output is exactly zero;
  - a still view merges the previous reservoir each frame: M grows by one per frame up to the cap of 40."""
import ctypes

import numpy as np

import orc
import rtpkg
from test_oracle_di import seeded, still_pair

rt = rtpkg.rt


def gi_frame(osc, cams, dn, am, mv, dn_p, am_p, this, prev, W, H):
    out = np.zeros((H, W, 4), np.float32)
    osc.gi_temporal(cams, dn, am, mv, dn_p, am_p, this, prev, out)
    return out


def test_gi_without_reuse_matches_gris_without_reuse():
    hs = rt.HostScene(0, 2)
    osc = orc.OracleScene(hs.arrays)
    W, H = 32, 24
    n = W * H
    base = hs.camera(0, W, H)[0]
    dn, am, mv, _ = osc.gbuffer(base)
    s = rt.settings(shift=2, temporal=False, spatial=False, max_depth=15)
    gi, gris = [], []
    for f in range(96):
        cams = seeded(base, f)
        out = gi_frame(osc, cams, dn, am, mv, dn, am, orc.reset_gi_reservoirs(n), orc.reset_gi_reservoirs(n), W, H)
        gi.append(out[..., :3].astype(np.float64).mean())
        r = orc.reset_reservoirs(n)
        t = orc.reset_reservoirs(n)
        o2 = np.zeros((H, W, 4), np.float32)
        osc.gris_path_trace(cams, dn, am, s, r)
        osc.gris_temporal(cams, dn, am, mv, dn, am, s, r, orc.reset_reservoirs(n), t)
        osc.gris_spatial(cams, dn, am, s, t, r, o2)
        gris.append(o2[..., :3].astype(np.float64).mean())
    gi, gris = np.array(gi), np.array(gris)
    se = np.sqrt(gi.var(ddof=1) / len(gi) + gris.var(ddof=1) / len(gris))
    assert gi.mean() > 0
    assert abs(gi.mean() - gris.mean()) < 4 * se, (gi.mean(), gris.mean(), se)


def test_gi_floor_has_no_indirect_light():
    hs = rt.HostScene(2, 0)
    osc = orc.OracleScene(hs.arrays)
    W, H = 24, 24
    n = W * H
    base = hs.camera(0, W, H)[0]
    dn, am, mv, _ = osc.gbuffer(base)
    prev = orc.reset_gi_reservoirs(n)
    for f in range(4):
        cams = seeded(base, f)
        cams[0].frameIndex = 0x80000000 if f == 0 else 0
        this = orc.reset_gi_reservoirs(n)
        out = gi_frame(osc, cams, dn, am, mv, dn, am, this, prev, W, H)
        assert (out[..., :3] == 0).all()
        prev = this


def test_gi_temporal_count_grows_to_the_cap():
    hs = rt.HostScene(0, 2)
    osc = orc.OracleScene(hs.arrays)
    W, H = 16, 12
    n = W * H
    base = hs.camera(0, W, H)[0]
    dn, am, _, _ = osc.gbuffer(base)
    mv = np.zeros((H, W, 2), np.uint16)
    bufs = [orc.reset_gi_reservoirs(n), orc.reset_gi_reservoirs(n)]
    fg = dn.reshape(n, 4)[:, 0] > 0
    for f in range(45):
        cams = still_pair(base)
        cams[0].seed = (f * 2654435761 + 7) & 0xFFFFFFFF
        cams[0].frameIndex = 0x80000000 if f == 0 else f
        cur = f & 1
        this = orc.reset_gi_reservoirs(n)
        gi_frame(osc, cams, dn, am, mv, dn, am, this, bufs[cur ^ 1], W, H)
        bufs[cur] = this
        m = this["sampleCount"][fg]
        assert (m <= min(f + 1, 40)).all() and (m <= 40).all()
    assert (m == 40).mean() > 0.9
    assert (this["sampleCount"][~fg] == 0).all()
