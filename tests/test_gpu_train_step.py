"""The optimizer side of the training step on the GPU (dpmi_lr_tick,
dpmi_adam_*) against the CPU oracle (oracle/optim_ref.py), the Trainer's
device Adam against torch.optim.Adam on the same gradients, and the whole
training step captured as a HIP graph against eager steps."""

import numpy as np
import pytest
import torch

from dpmi import _lib, ops, systems

pytestmark = pytest.mark.gpu
DEV = torch.device("cuda:0")


def _say(msg):
    print(f"[gpu case] {msg}", flush=True)


@pytest.fixture(scope="module", autouse=True)
def _need_gpu():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    _lib.load()


SCHEDULES = [
    dict(start_lr=1e-3, num_steps=60, stop_lr=1e-6, decay_steps=7),
    dict(start_lr=2e-3, num_steps=60, stop_lr_ratio=0.01, decay_steps=5, warmup_steps=6, warmup_start_factor=0.1),
    dict(start_lr=1e-3, num_steps=60, stop_lr=1e-5, decay_steps=4, smooth=True),
    dict(start_lr=1e-3, num_steps=60, stop_lr=4e-4, decay_steps=3, decay_rate=0.5),  # clipped at stop_lr
]


@pytest.mark.parametrize("kw", SCHEDULES, ids=["stepped", "warmup", "smooth", "clipped"])
def test_lr_tick_matches_oracle(kw):
    from dpmi import train
    from oracle import optim_ref

    _say(f"lr tick {kw}")
    loss = train.EnergyStdLoss(0.02, 1000.0, 1.0, 2.0, starter_learning_rate=kw["start_lr"])
    sched = train.LearningRateExp(**kw).sched(loss)
    ref = optim_ref.LrExp(**kw)
    hyper = torch.zeros(4, dtype=torch.float64, device=DEV)
    step = torch.zeros(1, dtype=torch.int64, device=DEV)
    for k in range(70):
        ops.lr_tick(hyper, step, sched)
        h = hyper.cpu().numpy()
        lr = ref.optimizer_lr(k)
        pe, pf = optim_ref.prefactors(lr, kw["start_lr"], 0.02, 1.0, 1000.0, 2.0)
        assert abs(h[0] - lr) <= 4e-16 * lr, (k, h[0], lr)
        assert abs(h[1] - pe) <= 1e-15 and abs(h[2] - pf) <= 1e-12 * pf, (k, h, pe, pf)
        assert h[3] == k + 1 and int(step.item()) == k + 1


@pytest.mark.parametrize("n", [0, 1, 100003], ids=["empty", "one", "ragged"])
def test_adam_kernel_matches_oracle(n):
    """Four float64 steps with new gradients over five decades and a moving
    rate (weight decay on the one-element case)."""
    from oracle import optim_ref

    _say(f"adam kernel f64 n={n}")
    rng = np.random.default_rng(n + 3)
    p = rng.normal(size=n)
    m, v = np.zeros(n), np.zeros(n)
    tp = torch.tensor(p, dtype=torch.float64, device=DEV)
    tm, tv = torch.zeros_like(tp), torch.zeros_like(tp)
    hyper = torch.zeros(4, dtype=torch.float64, device=DEV)
    wd = 0.01 if n == 1 else 0.0
    for t in range(1, 5):
        lr = 1e-3 * 0.8 ** t
        g = rng.normal(size=n) * 10.0 ** rng.uniform(-4, 1, size=n)
        hyper.copy_(torch.tensor([lr, 0.0, 0.0, float(t)], dtype=torch.float64))
        ops.adam_step(tp, torch.tensor(g, dtype=torch.float64, device=DEV), tm, tv, hyper, weight_decay=wd)
        optim_ref.adam_step(p, g, m, v, t, lr, weight_decay=wd)
    np.testing.assert_allclose(tp.cpu().numpy(), p, rtol=2e-15, atol=1e-18)
    np.testing.assert_allclose(tm.cpu().numpy(), m, rtol=2e-15, atol=1e-300)
    np.testing.assert_allclose(tv.cpu().numpy(), v, rtol=2e-15, atol=1e-300)


@pytest.mark.parametrize("n", [1, 4097])
def test_adam_float32_update_within_rounding(n):
    """One float32 step from the same state: the update (p_new - p) agrees
    with the float64 oracle within float32 rounding of its terms."""
    from oracle import optim_ref

    _say(f"adam kernel f32 single step n={n}")
    rng = np.random.default_rng(9)
    p = rng.normal(size=n).astype(np.float32).astype(np.float64)
    g = (rng.normal(size=n) * 10.0 ** rng.uniform(-3, 1, size=n)).astype(np.float32).astype(np.float64)
    tp = torch.tensor(p, dtype=torch.float32, device=DEV)
    tm = torch.zeros_like(tp)
    tv = torch.zeros_like(tp)
    hyper = torch.tensor([1e-3, 0.0, 0.0, 1.0], dtype=torch.float64, device=DEV)
    ops.adam_step(tp, torch.tensor(g, dtype=torch.float32, device=DEV), tm, tv, hyper)
    ref = optim_ref.adam_step(p.copy(), g, np.zeros(n), np.zeros(n), 1, 1e-3)
    du, dr = tp.double().cpu().numpy() - p, ref - p
    assert np.abs(du - dr).max() <= 1e-6 * np.abs(p).max() + 1e-9


def test_adam_rejects_mismatched_buffers():
    a = torch.zeros(10, dtype=torch.float64, device=DEV)
    h = torch.zeros(4, dtype=torch.float64, device=DEV)
    with pytest.raises(ValueError):
        ops.adam_step(a, torch.zeros(9, dtype=torch.float64, device=DEV), a.clone(), a.clone(), h)
    with pytest.raises(ValueError):
        ops.adam_step(a, a.float(), a.clone(), a.clone(), h)


def _train_case(precision="fp64"):
    from dpmi.model import SeAEnergyModel

    spec = systems.WATER_SPEC
    params = systems.make_params(spec, seed=7)
    lab = SeAEnergyModel.from_params(spec, systems.make_params(spec, seed=8), device=DEV)
    frames = [systems.water192(k) for k in range(4)]
    coords = [torch.tensor(f.coord, device=DEV) for f in frames]
    atypes = [torch.tensor(f.atype, device=DEV) for f in frames]
    boxes = [f.box for f in frames]
    outs = [lab.compute(c, t, b) for c, t, b in zip(coords, atypes, boxes)]
    e_l = torch.stack([o["energy"] for o in outs])
    f_l = torch.stack([o["force"] for o in outs])
    return spec, params, coords, atypes, boxes, e_l, f_l


def test_trainer_adam_matches_torch_adam():
    """Three steps of the Trainer (dpmi Adam, decaying rate) and of
    torch.optim.Adam on the same gradients: parameters within 1e-12."""
    from dpmi import train
    from dpmi.model import SeAEnergyModel

    _say("trainer adam vs torch adam, 4 x 192 water")
    spec, params, coords, atypes, boxes, e_l, f_l = _train_case()
    sched = dict(start_lr=1e-3, num_steps=10, stop_lr=1e-5, decay_steps=1)
    ma = SeAEnergyModel.from_params(spec, params, device=DEV)
    mb = SeAEnergyModel.from_params(spec, params, device=DEV)
    ta = train.Trainer(ma, schedule=train.LearningRateExp(**sched))
    tb = train.Trainer(mb, schedule=train.LearningRateExp(**sched))
    try:
        topt = torch.optim.Adam(tb.params, lr=1e-3, fused=True)
    except (RuntimeError, TypeError):
        topt = torch.optim.Adam(tb.params, lr=1e-3)
    from oracle import optim_ref

    ref = optim_ref.LrExp(**sched)
    for k in range(3):
        batch = train.prepare(ma, coords, atypes, boxes)
        la = float(ta.step(batch, e_l, f_l)[0])
        lb = float(tb.loss_and_grad(batch, e_l, f_l)[0])
        tb.opt.step_count.add_(1)  # the schedule index of the next step
        for grp in topt.param_groups:
            grp["lr"] = ref.optimizer_lr(k)
        topt.step()
        assert abs(la - lb) <= 1e-9 * abs(lb), (k, la, lb)
    for pa, pb in zip(ta.params, tb.params):
        np.testing.assert_allclose(pa.detach().cpu().numpy(), pb.detach().cpu().numpy(), rtol=0,
                                   atol=1e-12 * max(1.0, float(pb.abs().max())))


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_captured_train_step_matches_eager(precision):
    """The captured step (lists, env-mat, forward, double backward, Adam)
    replayed with eager work and synchronisations between replays, and after
    new coordinates are copied in: every loss, and the parameters at the
    end, equal eager Trainer steps from the same start."""
    from dpmi import train
    from dpmi.model import SeAEnergyModel

    _say(f"captured train step {precision}, 4 x 192 water")
    spec, params, coords, atypes, boxes, e_l, f_l = _train_case()
    sched = dict(start_lr=1e-3, num_steps=20, stop_lr=1e-5, decay_steps=2)
    me = SeAEnergyModel.from_params(spec, params, device=DEV, precision=precision)
    mg = SeAEnergyModel.from_params(spec, params, device=DEV, precision=precision)
    te = train.Trainer(me, schedule=train.LearningRateExp(**sched))
    tg = train.Trainer(mg, schedule=train.LearningRateExp(**sched))
    cap = train.CapturedTrainStep(tg, coords, atypes, boxes, e_l, f_l, warmup=2)
    tol = 1e-10 if precision == "fp64" else 1e-4
    for _ in range(2):  # the capture's warm-up steps
        te.step(train.prepare(me, coords, atypes, boxes), e_l, f_l)
    rng = np.random.default_rng(4)
    for k in range(4):
        if k == 3:  # new coordinates in the captured inputs
            coords = [c + torch.tensor(rng.uniform(-0.02, 0.02, c.shape), device=DEV) for c in coords]
            for dst, src in zip(cap.coords, coords):
                dst.copy_(src)
        lg = float(cap.replay())
        torch.cuda.synchronize()
        cap.check()
        le = float(te.step(train.prepare(me, coords, atypes, boxes), e_l, f_l)[0])
        torch.cuda.synchronize()
        assert abs(lg - le) <= tol * abs(le), (k, lg, le)
    assert int(tg.opt.step_count.item()) == int(te.opt.step_count.item()) == 6
    num = sum(float(((a - b) ** 2).sum()) for a, b in zip(te.params, tg.params))
    den = sum(float((a ** 2).sum()) for a in te.params)
    assert (num / den) ** 0.5 <= tol


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32], ids=["f64", "f32"])
@pytest.mark.parametrize("R,F", [(0, 5), (1, 1), (1000, 25), (105984, 100), (3001, 300)])
def test_row_sum_matches_exact_sums(dtype, R, F):
    """Column sums against math.fsum of the same values; bitwise equal on
    two calls (fixed order)."""
    import math

    _say(f"row sum {dtype} R={R} F={F}")
    rng = np.random.default_rng(R + F)
    x = rng.normal(size=(R, F)) * 10.0 ** rng.uniform(-2, 2, size=(R, F))
    tx = torch.tensor(x, dtype=dtype, device=DEV)
    xr = tx.double().cpu().numpy()
    out = ops.row_sum(tx)
    ref = np.array([math.fsum(xr[:, c]) for c in range(F)])
    scale = np.abs(xr).sum(axis=0) if R else np.zeros(F)
    eps = 1e-15 if dtype == torch.float64 else 1e-6
    assert np.all(np.abs(out.double().cpu().numpy() - ref) <= 4 * eps * scale + 1e-300)
    assert torch.equal(out, ops.row_sum(tx))


def test_row_broadcast_pair_gradients():
    """RowBroadcast / RowSum: first and second derivatives against torch's
    own broadcast on float64 (gradcheck-sized)."""
    _say("row broadcast autograd")
    v = torch.randn(7, dtype=torch.float64, device=DEV, requires_grad=True)
    x = torch.randn(40000, 7, dtype=torch.float64, device=DEV)
    y = (torch.tanh(x * ops.RowBroadcast.apply(v, 40000)) ** 2).sum()
    (g,) = torch.autograd.grad(y, v, create_graph=True)
    y2 = (torch.tanh(x * v) ** 2).sum()
    (g2,) = torch.autograd.grad(y2, v, create_graph=True)
    torch.testing.assert_close(g, g2, rtol=1e-12, atol=1e-12)
    (h,) = torch.autograd.grad(g.sum(), v)
    (h2,) = torch.autograd.grad(g2.sum(), v)
    torch.testing.assert_close(h, h2, rtol=1e-12, atol=1e-12)
