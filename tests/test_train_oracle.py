"""The training-step oracle (E + F loss and its parameter gradient by a
double backward) pinned by central finite differences of the loss, h = 1e-5
(FP64), with the embedding as a network and as a table."""

import numpy as np
import pytest

from dpmi import systems, table


def _frames(n, natoms):
    out = []
    for k in range(n):
        s = systems.water192(k)
        out.append((s.coord[:natoms], s.atype[:natoms], s.box))
    return out


def _setup(tabulated):
    from oracle.model_ref import Model

    spec = systems.WATER_SPEC
    params = systems.make_params(spec, seed=7)
    tabs = None
    if tabulated:
        tabs = table.build_tables(params["embedding"], params["davg"], params["dstd"], spec.rcut, spec.rcut_smth,
                                  0.9, spec.sel)
    frames = _frames(2, 48)
    lab = Model.from_params(spec, systems.make_params(spec, seed=8))
    el, fl = [], []
    rng = np.random.default_rng(5)
    for c, t, b in frames:
        o = lab.evaluate(c, t, b)
        # label noise keeps the residuals, and so the gradient, well above
        # the finite-difference noise floor
        el.append(o["energy"] + rng.normal())
        fl.append(o["force"] + 0.5 * rng.normal(size=o["force"].shape))
    return spec, params, tabs, frames, np.array(el), np.array(fl)


def _perturbed(params, path, h):
    import copy

    p = copy.deepcopy(params)
    kind, net, layer, name, idx = path
    arr = np.array(p[kind][net][layer][name], dtype=np.float64)
    arr[idx] += h
    p[kind][net][layer][name] = arr
    return p


# (params key, net, layer, array, index); the index order follows Model.parameters()
PATHS = [
    ("fitting", 0, 0, "w", (5, 3)),
    ("fitting", 1, 2, "b", (7,)),
    ("fitting", 1, 1, "idt", (11,)),
    ("fitting", 0, 3, "w", (17, 0)),
    ("embedding", 0, 0, "w", (0, 4)),
    ("embedding", 1, 2, "w", (10, 20)),
    ("embedding", 1, 1, "idt", (3,)),
]


def _grad_of(names, grads, path):
    kind, net, layer, name, idx = path
    return grads[names.index((kind, net, layer, name))][idx]


@pytest.mark.parametrize("tabulated", [False, True], ids=["network", "table"])
def test_train_grad_matches_finite_difference(tabulated):
    from oracle.model_ref import Model

    spec, params, tabs, frames, el, fl = _setup(tabulated)
    m = Model.from_params(spec, params, tables=tabs)
    named = m.named_parameters()
    loss, grads = m.train_loss(frames, el, fl, params=[t for _, t in named])
    names = [k for k, _ in named]
    assert loss > 0
    h = 1e-5
    for path in PATHS:
        if tabulated and path[0] == "embedding":
            continue
        lp, _ = Model.from_params(spec, _perturbed(params, path, h), tables=tabs).train_loss(frames, el, fl)
        lm, _ = Model.from_params(spec, _perturbed(params, path, -h), tables=tabs).train_loss(frames, el, fl)
        fd = (lp - lm) / (2 * h)
        g = _grad_of(names, grads, path)
        # FD noise here is below 1e-8 absolute; dropping the tabulate
        # grad's em_x term from grad_grad moves these gradients by >= 4e-4
        # relative, 1e-6 absolute
        assert abs(fd - g) <= 1e-5 * abs(g) + 1e-8, (path, fd, g)
