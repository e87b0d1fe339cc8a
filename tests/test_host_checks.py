"""Host-side logic that runs without a GPU: atom-type validation of the
model path and the learning-rate schedule of the loss prefactors."""

import numpy as np
import pytest
import torch

from dpmi import systems, train


def _cpu_model():
    from dpmi.model import SeAEnergyModel

    spec = systems.WATER_SPEC
    return SeAEnergyModel.from_params(spec, systems.make_params(spec, seed=7), device="cpu")


@pytest.mark.parametrize("bad", [-1, 2, 7])
def test_model_rejects_types_outside_range(bad):
    """Types outside [0, ntypes) raise a ValueError naming the range before
    any kernel runs (the reference masks type < 0 as virtual atoms and
    cannot index its nets with type >= ntypes)."""
    m = _cpu_model()
    s = systems.water192(0)
    atype = torch.as_tensor(s.atype.copy())
    atype[5] = bad
    with pytest.raises(ValueError, match=r"\[0, 2\)"):
        m.compute(torch.as_tensor(s.coord), atype, s.box)
    ext_t = torch.as_tensor(np.concatenate([s.atype, [bad]]).astype(np.int32))
    with pytest.raises(ValueError, match=r"\[0, 2\)"):
        m.compute_lower(torch.zeros((s.nloc + 1, 3), dtype=torch.float64), ext_t,
                        torch.full((s.nloc, 4), -1, dtype=torch.int32), s.nloc)


def test_model_accepts_valid_types_once():
    m = _cpu_model()
    t = torch.as_tensor(systems.water192(0).atype)
    m.check_types(t)
    assert m._types_checked[0] is t
    m.check_types(t)  # cached: same tensor, same version


def test_loss_prefactor_schedule_matches_reference_formula():
    """pref = limit + (start - limit) * lr / starter_lr
    (deepmd/pt/loss/ener.py:229-231): start at the starter rate, limit at
    lr = 0, linear in between; the oracle's schedule agrees."""
    from oracle.model_ref import loss_prefactors

    loss = train.EnergyStdLoss(0.02, 1000.0, limit_pref_e=1.0, limit_pref_f=1.0, starter_learning_rate=1e-3)
    assert loss.prefactors(1e-3) == pytest.approx((0.02, 1000.0), rel=1e-15)
    assert loss.prefactors(0.0) == pytest.approx((1.0, 1.0), rel=1e-15)
    assert loss.prefactors(2.5e-4) == pytest.approx((1.0 - 0.98 * 0.25, 1.0 + 999.0 * 0.25), rel=1e-14)
    for lr in (1e-3, 3e-4, 1e-5, 0.0):
        assert loss.prefactors(lr) == pytest.approx(loss_prefactors(0.02, 1000.0, 1.0, 1.0, lr, 1e-3), rel=1e-15)
    # without limits the prefactors stay at their start values
    assert train.EnergyStdLoss(0.02, 1000.0, starter_learning_rate=1e-3).prefactors(1e-5) == \
        pytest.approx((0.02, 1000.0), rel=1e-15)


def test_loss_value_at_scheduled_prefactors():
    rng = np.random.default_rng(3)
    nf, nloc = 3, 5
    e, el = rng.normal(size=nf), rng.normal(size=nf)
    f, fl = rng.normal(size=(nf, nloc, 3)), rng.normal(size=(nf, nloc, 3))
    loss = train.EnergyStdLoss(0.02, 1000.0, limit_pref_e=2.0, limit_pref_f=3.0, starter_learning_rate=1e-3)
    v, _ = loss(*(torch.as_tensor(x) for x in (e, f, el, fl)), nloc, learning_rate=4e-4)
    pe, pf = 2.0 + (0.02 - 2.0) * 0.4, 3.0 + (1000.0 - 3.0) * 0.4
    want = pe / nloc * np.mean((e - el) ** 2) + pf * np.mean((f - fl) ** 2)
    assert float(v) == pytest.approx(want, rel=1e-13)
