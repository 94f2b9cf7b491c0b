"""GPU learned target function through the C ABI: MFMA inference and gradients against
the oracle (FP16 rounding points mirrored; tolerances cover FP32 summation order and
occasional one-ulp FP16 flips of an activation), determinism of the weight gradients,
and online training on a synthetic radiance field."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

import gpuhelp  # noqa: E402
import orc  # noqa: E402

rt = gpuhelp.rt
LO, HI = (-1.0, -2.0, 0.5), (3.0, 2.0, 4.5)
LOG2T, BASE, MAXR = 14, 4, 256


def make(seed=5):
    print(f"[gpu] mlp seed {seed}", flush=True)
    m = rt.Mlp(rt.mlp_config(LO, HI, log2_table=LOG2T, base_res=BASE, max_res=MAXR, seed=seed))
    g = np.random.default_rng(seed)
    grid = g.uniform(-0.5, 0.5, m.grid_params).astype(np.float32)
    w = np.concatenate([g.uniform(-1, 1, 2048) * np.sqrt(6 / 96), g.uniform(-1, 1, 4096) * np.sqrt(6 / 128),
                        g.uniform(-1, 1, 192) * np.sqrt(6 / 67)]).astype(np.float32)
    m.set_params(grid, w)
    return m, grid, w


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a, np.float32)).cuda()


def infer(m, pos):
    d_pos = dev(pos)
    out = torch.zeros((len(pos), 3), dtype=torch.float32, device="cuda")
    rt.check(rt.core().rt_mlp_infer(m.handle, d_pos.data_ptr(), len(pos), out.data_ptr(), gpuhelp.stream()), "infer")
    torch.cuda.synchronize()
    return out.cpu().numpy()


def gradients(m, pos, target):
    d_pos, d_t = dev(pos), dev(target)
    loss = torch.zeros(1, dtype=torch.float32, device="cuda")
    rt.check(rt.core().rt_mlp_gradients(m.handle, d_pos.data_ptr(), d_t.data_ptr(), len(pos), loss.data_ptr(),
                                        gpuhelp.stream()), "gradients")
    torch.cuda.synchronize()
    gg, gw = m.read_gradients()
    return float(loss.item()), gg, gw


def test_params_roundtrip_and_init():
    m = rt.Mlp(rt.mlp_config(LO, HI, log2_table=LOG2T, base_res=BASE, max_res=MAXR, seed=1))
    grid, w = m.params()
    assert grid.shape == (16 * (1 << LOG2T) * 2,) and w.shape == (6336,)
    assert np.abs(grid).max() <= 1e-4 and np.abs(grid).max() > 0
    assert np.abs(w[:2048]).max() <= np.sqrt(6 / 96) and np.abs(w[6144:]).max() <= np.sqrt(6 / 67)
    m2 = rt.Mlp(rt.mlp_config(LO, HI, log2_table=LOG2T, base_res=BASE, max_res=MAXR, seed=1))
    assert np.array_equal(m2.params()[1], w), "initialisation is a function of the seed"


def test_inference_matches_oracle():
    m, grid, w = make(5)
    g = np.random.default_rng(1)
    pos = g.uniform(np.array(LO) - 0.3, np.array(HI) + 0.3, (20000, 3)).astype(np.float32)  # includes clamped ones
    got = infer(m, pos)
    want = orc.OracleMlp(LOG2T, BASE, MAXR, LO, HI).forward(grid, w, pos, mode=1)
    rel = np.abs(got - want) / want
    print(f"[gpu] inference rel err median {np.median(rel):.2e} max {rel.max():.2e}", flush=True)
    assert np.median(rel) < 1e-5 and np.quantile(rel, 0.999) < 1e-2 and rel.max() < 5e-2


def test_inference_ragged_sizes():
    m, grid, w = make(6)
    osc = orc.OracleMlp(LOG2T, BASE, MAXR, LO, HI)
    g = np.random.default_rng(2)
    for n in (1, 31, 33, 127, 129):
        pos = g.uniform(LO, HI, (n, 3)).astype(np.float32)
        got = infer(m, pos)
        want = osc.forward(grid, w, pos, mode=1)
        assert np.allclose(got, want, rtol=5e-2), n


def test_gradients_match_oracle():
    m, grid, w = make(7)
    g = np.random.default_rng(3)
    n = 4096 + 17
    pos = g.uniform(LO, HI, (n, 3)).astype(np.float32)
    target = np.exp(g.normal(0, 1, (n, 3))).astype(np.float32)
    target[::9] = -1.0
    loss, gg, gw = gradients(m, pos, target)
    oloss, ogg, ogw = orc.OracleMlp(LOG2T, BASE, MAXR, LO, HI).gradients(grid, w, pos, target, mode=1)
    print(f"[gpu] loss gpu {loss:.6f} oracle {oloss:.6f}", flush=True)
    assert abs(loss - oloss) <= 1e-3 * oloss
    for name, a, b in (("W1", gw[:2048], ogw[:2048]), ("W2", gw[2048:6144], ogw[2048:6144]),
                       ("W3", gw[6144:], ogw[6144:]), ("grid", gg, ogg)):
        err = np.linalg.norm(a - b) / np.linalg.norm(b)
        print(f"[gpu] {name} gradient relative error {err:.2e}", flush=True)
        assert err < 2e-2, name
    # the grid gradient touches exactly the entries the oracle touches
    assert np.array_equal(gg != 0, ogg != 0)


def test_weight_gradients_are_deterministic():
    m, grid, w = make(8)
    g = np.random.default_rng(4)
    pos = g.uniform(LO, HI, (3000, 3)).astype(np.float32)
    target = np.exp(g.normal(0, 1, (3000, 3))).astype(np.float32)
    _, _, a = gradients(m, pos, target)
    _, _, b = gradients(m, pos, target)
    assert np.array_equal(a.view(np.uint32), b.view(np.uint32))


def test_training_fits_a_radiance_field():
    """Online Adam steps on a smooth positive field over the box: the loss drops by far and the
    fitted field is within a few percent of the truth."""
    m = rt.Mlp(rt.mlp_config(LO, HI, log2_table=16, base_res=4, max_res=256, seed=9))
    g = np.random.default_rng(5)

    def field(p):
        q = (p - np.array(LO)) / (np.array(HI) - np.array(LO))
        r = 0.5 + 2.0 * np.exp(-8 * ((q - 0.4) ** 2).sum(1))[:, None] * np.array([1.0, 0.6, 0.3])
        return (r + 0.3 * np.sin(6 * q[:, :1]) ** 2).astype(np.float32)

    d_loss = torch.zeros(1, dtype=torch.float32, device="cuda")
    losses = []
    for it in range(300):
        pos = g.uniform(LO, HI, (1 << 14, 3)).astype(np.float32)
        d_pos, d_t = dev(pos), dev(field(pos))
        rt.check(rt.core().rt_mlp_train_step(m.handle, d_pos.data_ptr(), d_t.data_ptr(), len(pos), d_loss.data_ptr(),
                                             gpuhelp.stream()), "train")
        if it % 50 == 0 or it == 299:
            torch.cuda.synchronize()
            losses.append(float(d_loss.item()))
            print(f"[gpu] step {it} loss {losses[-1]:.5f}", flush=True)
    assert losses[-1] < 0.05 * losses[0]
    test = g.uniform(LO, HI, (5000, 3)).astype(np.float32)
    pred = infer(m, test)
    rel = np.abs(pred - field(test)) / field(test)
    print(f"[gpu] fitted field mean rel err {rel.mean():.4f}", flush=True)
    assert rel.mean() < 0.05
