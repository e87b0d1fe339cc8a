"""CPU oracle of the training step's optimizer side (test infrastructure only:
tests and the benchmark's CPU baseline use it as the checker; nothing in the
library imports it).

* learning rate: the reference trainer's LambdaLR over LearningRateExp
  (deepmd/pt/train/training.py:1138-1150; deepmd/dpmodel/utils/
  learning_rate.py:30-166 BaseLR.__init__, 167-217 BaseLR.value, 219-397
  LearningRateExp), restated in float64 numpy;
* loss prefactors: deepmd/pt/loss/ener.py:229-231;
* Adam: torch.optim.Adam as the reference builds it (training.py:1028-1030:
  betas from adam_beta1/2, weight_decay, fused on the GPU; no amsgrad), in the
  published form of the update (Kingma and Ba, Algorithm 1, with the bias
  corrections folded into the step size and the denominator as torch does).

Pinned by tests/test_optim_oracle.py: the schedule against its closed forms
(k = 0, every decay boundary, the stop rate reached at num_steps, the warmup
line, the smooth form) and Adam against torch.optim.Adam on the CPU over
several steps, with and without weight decay.
"""

from __future__ import annotations

import numpy as np


class LrExp:
    """Parameters of LearningRateExp as its constructor derives them
    (learning_rate.py:30-166, 258-367)."""

    def __init__(self, start_lr, num_steps, stop_lr=None, stop_lr_ratio=None, decay_steps=5000, decay_rate=None,
                 warmup_steps=0, warmup_ratio=None, warmup_start_factor=0.0, smooth=False):
        self.start_lr = float(start_lr)
        self.stop_lr = start_lr * stop_lr_ratio if stop_lr_ratio is not None else float(stop_lr)
        self.warmup_steps = int(warmup_ratio * num_steps) if warmup_ratio is not None else int(warmup_steps)
        self.warmup_start_lr = warmup_start_factor * start_lr
        decay_total = num_steps - self.warmup_steps
        self.decay_steps = int(decay_steps)
        if decay_rate is None and self.decay_steps >= decay_total:
            self.decay_steps = 100 if decay_total // 10 > 100 else decay_total // 100 + 1
        if decay_rate is not None:
            self.decay_rate = float(decay_rate)
        elif decay_total == 0:
            self.decay_rate = 1.0
        else:
            self.decay_rate = np.exp(np.log(max(self.stop_lr, 1e-10) / self.start_lr)
                                     / (decay_total / self.decay_steps)).item()
        self.smooth = bool(smooth)

    def _decay(self, step: float) -> float:
        # learning_rate.py:367-397
        e = step / self.decay_steps if self.smooth else float(np.floor(step / self.decay_steps))
        lr = self.start_lr * np.power(self.decay_rate, e)
        return float(max(lr, self.stop_lr))

    def value(self, k: int) -> float:
        """BaseLR.value (learning_rate.py:167-217)."""
        if self.warmup_steps == 0:
            return self._decay(float(k))
        warm = self.warmup_start_lr + (self.start_lr - self.warmup_start_lr) * (float(k) / self.warmup_steps)
        decay = self._decay(max(float(k) - self.warmup_steps, 0.0))
        return warm if k < self.warmup_steps else decay

    def optimizer_lr(self, k: int) -> float:
        """The rate LambdaLR hands the optimizer at step k:
        base_lr * (value(k) / base_lr) (training.py:1138-1150)."""
        return self.start_lr * (self.value(k) / self.start_lr)


def prefactors(lr, start_lr, start_pref_e, limit_pref_e, start_pref_f, limit_pref_f):
    """ener.py:229-231."""
    coef = lr / start_lr
    return (limit_pref_e + (start_pref_e - limit_pref_e) * coef, limit_pref_f + (start_pref_f - limit_pref_f) * coef)


def adam_step(p, g, m, v, t, lr, betas=(0.9, 0.999), eps=1e-8, weight_decay=0.0):
    """One Adam update (t = steps taken including this one) on float64
    arrays, in place; returns p."""
    b1, b2 = betas
    if weight_decay != 0.0:
        g = g + p * weight_decay
    m[...] = b1 * m + (1.0 - b1) * g
    v[...] = b2 * v + (1.0 - b2) * g * g
    step_size = lr / (1.0 - b1 ** t)
    denom = np.sqrt(v) / np.sqrt(1.0 - b2 ** t) + eps
    p[...] = p - step_size * m / denom
    return p
