"""The cost model's MLP kernels against a plain PyTorch fp32 reference on the
CPU: forward, gradients of the group-centred loss, Adam steps."""
import numpy as np
import pytest
import torch

from graphindor_amd import native as N

pytestmark = pytest.mark.gpu


def unpack(p, F, H):
    o = 0
    out = []
    for shape in [(F, H), (H,), (H, H), (H,), (H,), (1,)]:
        n = int(np.prod(shape))
        out.append(torch.tensor(p[o:o + n].reshape(shape), dtype=torch.float32))
        o += n
    return out


def ref_forward(params, x):
    W1, b1, W2, b2, W3, b3 = params
    h1 = torch.relu(x @ W1 + b1)
    h2 = torch.relu(h1 @ W2 + b2)
    return h2 @ W3 + b3


def ref_loss(params, x, t, goff):
    y = ref_forward(params, x)
    r = y - t
    total = 0.0
    for g in range(len(goff) - 1):
        seg = r[goff[g]:goff[g + 1]]
        if len(seg):
            total = total + ((seg - seg.mean()) ** 2).sum()
    return total / goff[-1]


def data(rows_per_group, F, seed):
    rng = np.random.default_rng(seed)
    goff = np.concatenate([[0], np.cumsum(rows_per_group)]).astype(np.int32)
    x = rng.normal(size=(goff[-1], F)).astype(np.float32)
    t = (np.sin(x[:, 0]) + 0.5 * x[:, 1] * x[:, 2] + 0.1 * rng.normal(size=goff[-1])).astype(np.float32)
    return x, t, goff


@pytest.mark.parametrize("F,H,rows", [(40, 256, 300), (64, 256, 1), (17, 96, 130)])
def test_forward_matches_torch(F, H, rows):
    m = N.Mlp(F, H, seed=3)
    p = m.get_params()
    x = np.random.default_rng(1).normal(size=(rows, F)).astype(np.float32)
    got = m.predict(x)
    want = ref_forward(unpack(p, F, H), torch.tensor(x)).numpy()
    np.testing.assert_allclose(got, want, rtol=1e-4, atol=1e-4)


def test_gradients_match_torch():
    F, H = 48, 256
    x, t, goff = data([1, 7, 50, 200, 3], F, 2)
    m = N.Mlp(F, H, seed=5)
    p0 = m.get_params()
    m.train(x, t, goff, steps=1, lr=0.0)
    g = m.get_grads()
    params = [q.requires_grad_() for q in unpack(p0, F, H)]
    loss = ref_loss(params, torch.tensor(x), torch.tensor(t), goff)
    loss.backward()
    want = np.concatenate([q.grad.numpy().ravel() for q in params])
    scale = np.abs(want).max()
    np.testing.assert_allclose(g, want, rtol=1e-3, atol=1e-4 * scale)
    np.testing.assert_array_equal(m.get_params(), p0)  # lr 0: no change


def test_adam_steps_match_torch():
    F, H, steps, lr, wd = 24, 64, 5, 1e-3, 1e-4
    x, t, goff = data([30, 30, 40], F, 4)
    m = N.Mlp(F, H, seed=8)
    p0 = m.get_params()
    m.train(x, t, goff, steps=steps, lr=lr, weight_decay=wd)
    got = m.get_params()
    params = unpack(p0, F, H)
    m1 = [torch.zeros_like(q) for q in params]
    m2 = [torch.zeros_like(q) for q in params]
    for step in range(1, steps + 1):
        ps = [q.clone().requires_grad_() for q in params]
        ref_loss(ps, torch.tensor(x), torch.tensor(t), goff).backward()
        new = []
        for i, q in enumerate(ps):
            g = q.grad
            m1[i] = 0.9 * m1[i] + 0.1 * g
            m2[i] = 0.999 * m2[i] + 0.001 * g * g
            mh = m1[i] / (1 - 0.9 ** step)
            vh = m2[i] / (1 - 0.999 ** step)
            new.append((q - lr * (mh / (vh.sqrt() + 1e-8) + wd * q)).detach())
        params = new
    want = np.concatenate([q.numpy().ravel() for q in params])
    # Adam normalises each coordinate, so a gradient near zero may round
    # either way; the bulk must agree closely.
    diff = np.abs(got - want)
    assert np.mean(diff < 1e-5) > 0.999
    assert diff.max() < 2 * lr * steps


def test_training_reduces_loss_and_ranks_within_groups():
    F, H = 16, 256
    x, t, goff = data([40] * 25, F, 7)
    m = N.Mlp(F, H, seed=1)
    first = m.train(x, t, goff, steps=1, lr=1e-3)
    last = m.train(x, t, goff, steps=400, lr=1e-3)
    assert last < 0.2 * first
    y = m.predict(x)
    hits = sum(int(np.argmin(y[goff[g]:goff[g + 1]]) == np.argmin(t[goff[g]:goff[g + 1]]) or
                   t[goff[g] + np.argmin(y[goff[g]:goff[g + 1]])] <= np.sort(t[goff[g]:goff[g + 1]])[2])
               for g in range(len(goff) - 1))
    assert hits >= 20


def test_cost_model_save_load_and_selection(tmp_path):
    """A cost model trained on recorded timings round-trips through its file,
    and whatever it picks counts the same as the analytic pick."""
    # R-MAT 2^10, edge factor 4: every candidate of these patterns counts in
    # milliseconds (on R-MAT 2^12 ef16 one 5-cycle count is ~1 TB of set work).
    g = N.Graph.rmat(10, 4, 5)
    v, e = g.csr()
    v, e = v.copy(), e.copy()
    inf = g.info()
    dg = N.DeviceGraph(v, e)
    pats = [(5, "0101110101010101010011000"), (5, "0100110100010100010110010"), (5, "0111110110110101110010000")]
    xs, ts, goff = [], [], [0]
    for n, adj in pats:
        cands = N.schedule_candidates(n, adj, inf["v_cnt"], len(e), inf["tri_cnt"])
        x = N.encode_candidates(n, adj, v, inf["tri_cnt"])
        ms = []
        print(f"timing 24 candidates of {adj}", flush=True)
        for c in cands[:24]:
            s = N.Schedule.from_candidate(n, adj, c)
            _, st = N.count(dg, N.DevicePlan(s.plan_desc()))
            ms.append(st.kernel_ms)
        xs.append(x[:24])
        ts.append(np.log(np.maximum(ms, 1e-4)).astype(np.float32))
        goff.append(goff[-1] + len(ms))
    m = N.CostModel.new(256, 6)
    m.train(np.concatenate(xs), np.concatenate(ts), goff, steps=200)
    path = tmp_path / "cm.gscm"
    m.save(path)
    m2 = N.CostModel.load(path)
    x = np.concatenate(xs)
    np.testing.assert_array_equal(m.predict(x), m2.predict(x))
    for n, adj in pats:
        sa, ia = N.select_schedule(n, adj, v, inf["tri_cnt"], None)
        sl, il = N.select_schedule(n, adj, v, inf["tri_cnt"], m2)
        assert not ia["learned"] and il["learned"] and ia["analytic"] == il["analytic"]
        ra, _ = N.count(dg, N.DevicePlan(sa.plan_desc()))
        rl, _ = N.count(dg, N.DevicePlan(sl.plan_desc()))
        assert ra // sa.redundancy == rl // sl.redundancy


def test_measured_selection_keeps_counts_and_never_loses_to_analytic_on_its_sample():
    """The model proposes, a sampled run decides: the chosen schedule counts
    what the analytic one counts."""
    g = N.Graph.rmat(10, 4, 5)
    v, e = g.csr()
    v, e = v.copy(), e.copy()
    inf = g.info()
    dg = N.DeviceGraph(v, e)
    model = N.CostModel.new(64, 3)
    for n, adj in [(5, "0101110101010101010011000"), (5, "0100110100010100010110010"),
                   (6, "011100101010110001100011010101001110")]:
        sa, ia = N.select_schedule(n, adj, v, inf["tri_cnt"], None)
        sm, im = N.select_schedule_measured(n, adj, g, model, top_k=3, sample_roots=512, time_limit_ms=2000)
        assert im["analytic"] == ia["analytic"] and 0 <= im["chosen"] < im["candidates"]
        ra, _ = N.count(dg, N.DevicePlan(sa.plan_desc()))
        rm, _ = N.count(dg, N.DevicePlan(sm.plan_desc()))
        assert ra // sa.redundancy == rm // sm.redundancy
    # without a model: the analytic pick, no measurement
    s0, i0 = N.select_schedule_measured(5, "0100110100010100010110010", g, None)
    assert i0["chosen"] == i0["analytic"] and not i0["learned"]
