"""What the trained network gains: the learned vMF-mixture proposal against the same estimator
without it (BSDF sampling at the primary vertex), on a reduced view of the terrain workload.

Both forms run GRIS with temporal + spatial reuse and the reference's resampling weights over the
same frame count, from reset reservoirs, on paired seed sets; each is judged by its relMSE against
a converged image of the GRIS estimator without reuse (1024 frames, FP64 mean), relMSE = mean over
pixels of sum_rgb (I - R)^2 / (R^2 + 1e-2). The proposal network is trained on the view beforehand
(128 frames with other seeds) and keeps training. Stated bound: the geometric mean over 8 seed sets
of relMSE(plain) / relMSE(proposal) exceeds 1.5 (the 1080p equal-time study measures 2.08, 95 %
interval [1.37, 3.15], over 16 sets), and the proposal is better in at least 5 of the 8 sets.
The learned target function has no such gain (its study is in bench.py --learned-study), so it has
no test of one."""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

import gpuhelp  # noqa: E402
from test_gpu_bias import still  # noqa: E402
from test_gpu_guide import GuidedRun  # noqa: E402
from test_gpu_learned import scene_bounds  # noqa: E402

rt = gpuhelp.rt
W, H = 320, 180


def test_trained_proposal_lowers_the_error():
    hs = rt.HostScene(1, 1)
    ds = rt.DeviceScene(hs.arrays)
    base = hs.camera(0, W, H)[0]
    run = GuidedRun(ds, W, H)
    L = rt.core()
    lo, hi = scene_bounds(ds)
    guide = rt.Mlp(rt.mlp_config(lo, hi, log2_table=19, seed=4), guide=True)
    kw = dict(shift=2, max_depth=6, cap=20, neighbors=3, radius=30.0 * W / 1920.0)
    s_ref = rt.settings(temporal=False, spatial=False, **kw)
    s_plain = rt.settings(temporal=True, spatial=True, **kw)
    s_guided = rt.settings(temporal=True, spatial=True, guided=True, **kw)

    def render(s, frames, seed, train=False):
        for b in run.resv + [run.temp]:
            rt.check(L.rt_gris_reservoir_reset(b.data_ptr(), W, H, gpuhelp.stream()), "reset")
        run.cur = 0
        acc = torch.zeros((H, W, 3), dtype=torch.float64, device="cuda")
        for i in range(frames):
            cams = still(base, i, seed)
            run.gbuffer(cams[0])
            if s.guidedProposal:
                run.guide_refresh(guide, cams)
            run.path_trace(cams, s)
            run.temporal(cams, s)
            run.spatial(cams, s)
            if s.guidedProposal and train:
                run.guide_train(guide, cams)
            acc += run.indirect.view(torch.float32).view(H, W, 4)[..., :3].double()
            run.flip()
        return acc / frames

    print("[gpu] gain: converged reference, 1024 frames", flush=True)
    ref = render(s_ref, 1024, 7)

    def relmse(img):
        return float((((img - ref) ** 2) / (ref ** 2 + 1e-2)).sum(-1).mean())

    print("[gpu] gain: training the proposal on 128 frames", flush=True)
    render(s_guided, 128, 29, train=True)
    ratios = []
    for k in range(8):
        seed = 500 + 11 * k
        e_plain = relmse(render(s_plain, 32, seed))
        e_guided = relmse(render(s_guided, 32, seed, train=True))
        ratios.append(e_plain / e_guided)
        print(f"[gpu] gain set {k}: relMSE plain {e_plain:.4f} proposal {e_guided:.4f}", flush=True)
    geo = math.exp(np.mean(np.log(ratios)))
    better = sum(r > 1 for r in ratios)
    print(f"[gpu] gain: geometric mean plain/proposal {geo:.3f}, better in {better} of 8", flush=True)
    assert geo > 1.5 and better >= 5
