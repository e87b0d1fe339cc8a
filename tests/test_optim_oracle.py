"""CPU pins of the optimizer oracle (oracle/optim_ref.py) and of the host-side
schedule of dpmi.train: the exponential-decay schedule against its closed
forms, Adam against torch.optim.Adam on the CPU (the optimizer class the
reference trainer builds, training.py:1028-1030)."""

import numpy as np
import pytest
import torch

from oracle import optim_ref


def test_lr_exp_closed_forms():
    s = optim_ref.LrExp(1e-3, num_steps=100000, stop_lr=1e-8, decay_steps=5000)
    r = (1e-8 / 1e-3) ** (5000 / 100000)
    assert abs(s.decay_rate - r) <= 1e-15 * r
    assert s.value(0) == 1e-3
    assert s.value(4999) == 1e-3
    assert abs(s.value(5000) - 1e-3 * r) <= 1e-18
    assert abs(s.value(12345) - 1e-3 * r ** 2) <= 1e-18
    # the stop rate is reached at num_steps and clips beyond
    assert abs(s.value(100000) - 1e-8) <= 1e-12 * 1e-8
    assert s.value(10 ** 7) == 1e-8
    # smooth decay: r^(k / s)
    sm = optim_ref.LrExp(1e-3, num_steps=100000, stop_lr=1e-8, decay_steps=5000, smooth=True)
    assert abs(sm.value(2500) - 1e-3 * r ** 0.5) <= 1e-15 * 1e-3
    # decay_steps beyond the decay phase is replaced (learning_rate.py:345-349)
    assert optim_ref.LrExp(1e-3, num_steps=1000, stop_lr=1e-6, decay_steps=5000).decay_steps == 11
    assert optim_ref.LrExp(1e-3, num_steps=200000, stop_lr=1e-6, decay_steps=500000).decay_steps == 100
    # an explicit rate keeps decay_steps; stop_lr is then only the floor
    e = optim_ref.LrExp(1e-3, num_steps=1000, stop_lr=5e-4, decay_steps=10, decay_rate=0.5)
    assert e.decay_steps == 10 and e.value(10) == 5e-4 and e.value(9) == 1e-3
    # warmup: a line from warmup_start_factor * start_lr, then the decay restarted
    w = optim_ref.LrExp(1e-3, num_steps=1000, stop_lr=1e-5, decay_steps=100, warmup_steps=10,
                        warmup_start_factor=0.1)
    assert abs(w.value(0) - 1e-4) <= 1e-19
    assert abs(w.value(5) - (1e-4 + 9e-4 * 0.5)) <= 1e-18
    assert w.value(10) == 1e-3 and w.value(109) == 1e-3
    assert abs(w.value(110) - 1e-3 * w.decay_rate) <= 1e-18
    assert abs(w.decay_rate - (1e-5 / 1e-3) ** (100 / 990)) <= 1e-15


def test_prefactors_follow_the_rate():
    pe, pf = optim_ref.prefactors(1e-3, 1e-3, 0.02, 1.0, 1000.0, 1.0)
    assert abs(pe - 0.02) <= 1e-16 and pf == 1000.0  # 1 + (0.02 - 1) * 1 rounds
    pe, pf = optim_ref.prefactors(0.0, 1e-3, 0.02, 1.0, 1000.0, 1.0)
    assert (pe, pf) == (1.0, 1.0)
    pe, pf = optim_ref.prefactors(5e-4, 1e-3, 0.02, 1.0, 1000.0, 1.0)
    assert abs(pe - 0.51) <= 1e-15 and abs(pf - 500.5) <= 1e-12


@pytest.mark.parametrize("weight_decay", [0.0, 0.01])
def test_adam_matches_torch(weight_decay):
    """Five steps with new gradients and a moving rate; float64."""
    rng = np.random.default_rng(5)
    shapes = [(7, 3), (11,), (4, 5, 2)]
    ps = [rng.normal(size=s) for s in shapes]
    tp = [torch.tensor(p, requires_grad=True) for p in ps]
    opt = torch.optim.Adam(tp, lr=1e-3, betas=(0.9, 0.999), eps=1e-8, weight_decay=weight_decay)
    flat = np.concatenate([p.ravel() for p in ps])
    m, v = np.zeros_like(flat), np.zeros_like(flat)
    for t in range(1, 6):
        lr = 1e-3 * 0.7 ** (t - 1)
        gs = [rng.normal(size=s) * 10.0 ** rng.integers(-3, 2) for s in shapes]
        for x, g in zip(tp, gs):
            x.grad = torch.tensor(g)
        for grp in opt.param_groups:
            grp["lr"] = lr
        opt.step()
        optim_ref.adam_step(flat, np.concatenate([g.ravel() for g in gs]), m, v, t, lr, weight_decay=weight_decay)
        ref = np.concatenate([x.detach().numpy().ravel() for x in tp])
        np.testing.assert_allclose(flat, ref, rtol=1e-14, atol=1e-17)


def test_adam_oracle_detects_a_dropped_bias_correction():
    """A plausible slip (no second-moment bias correction) moves the first
    step by a factor sqrt(1 - b2) ~ 0.03: far outside the tolerance above."""
    rng = np.random.default_rng(1)
    p = rng.normal(size=20)
    g = rng.normal(size=20)
    good = optim_ref.adam_step(p.copy(), g, np.zeros(20), np.zeros(20), 1, 1e-3)
    m = 0.1 * g
    v = 0.001 * g * g
    bad = p - (1e-3 / 0.1) * m / (np.sqrt(v) + 1e-8)
    assert np.abs((good - p) / (bad - p) - np.sqrt(0.001)).max() < 1e-4  # eps shifts it slightly


def test_host_schedule_matches_oracle():
    """dpmi.train.LearningRateExp derives the same schedule parameters (the
    values the GPU tick reads) as the oracle."""
    from dpmi import train

    for kw in (dict(start_lr=1e-3, num_steps=100000, stop_lr=1e-8, decay_steps=5000),
               dict(start_lr=1e-3, num_steps=1000, stop_lr=1e-6, decay_steps=5000),
               dict(start_lr=2e-3, num_steps=1000, stop_lr_ratio=0.01, decay_steps=100, warmup_ratio=0.05,
                    warmup_start_factor=0.2, smooth=True),
               dict(start_lr=1e-3, num_steps=1000, stop_lr=5e-4, decay_steps=10, decay_rate=0.5)):
        a, b = train.LearningRateExp(**kw), optim_ref.LrExp(**kw)
        for k in ("start_lr", "stop_lr", "decay_rate", "decay_steps", "warmup_steps", "warmup_start_lr", "smooth"):
            assert getattr(a, k) == getattr(b, k), (kw, k)
    loss = train.EnergyStdLoss(0.02, 1000.0, 1.0, 1.0, starter_learning_rate=1e-3)
    s = train.LearningRateExp(1e-3, 100, stop_lr=1e-5, decay_steps=10).sched(loss)
    assert s.shape == (11,) and s.dtype == np.float64
    assert list(s[7:]) == [0.02, 1.0, 1000.0, 1.0]
    with pytest.raises(ValueError):
        train.LearningRateExp(1e-3, 100)
    with pytest.raises(ValueError):
        train.LearningRateExp(1e-3, 100, stop_lr=1e-5, warmup_steps=100)
