"""Training step on the GPU (8 frames x 192 water atoms, E + F loss with
pref_e 0.02 and pref_f 1000): loss and parameter gradients against the
training oracle, central finite differences (h = 1e-5, FP64), the FP32 form
against FP64, and the force/virial gradient kernels against the oracle and
the reference's golden vectors."""

import numpy as np
import pytest
import torch

from dpmi import systems, table

pytestmark = pytest.mark.gpu
DEV = torch.device("cuda:0")


def _say(msg):
    print(f"[gpu case] {msg}", flush=True)


@pytest.fixture(scope="module", autouse=True)
def _need_gpu():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")


# --------------------------------------------------------------------------- kernels


def test_prod_force_grad_golden(oracle, golden):
    from dpmi import ops

    _say("prod_force_grad / prod_virial_grad golden 6 atoms")
    for name, cls in (("prod_force_grad_a", "TestProdForceGradA"), ("prod_virial_grad_a", "TestProdVirialGradA")):
        g = golden(f"{name}.json")[cls]
        posi = np.array(g["posi"]).reshape(-1, 3)
        atype = np.array(g["atype"], np.int32)
        box = np.diag([13.0, 13.0, 13.0]).ravel()
        ec, et, _, nlist, _ = oracle.build_nlist_pbc(posi, atype, box, g["rc"], [5, 5])
        _, ed, rij = oracle.env_mat_a(nlist, ec, et, np.zeros((2, 10, 4)), np.ones((2, 10, 4)), g["rc"],
                                      g["rc_smth"], [5, 5])
        t = lambda a, dt=torch.float64: torch.as_tensor(np.ascontiguousarray(a), dtype=dt, device=DEV)
        if name == "prod_force_grad_a":
            grad = 10 - np.arange(18) * 0.1
            gn = ops.prod_force_grad(t(np.concatenate([grad, grad]).reshape(2, -1)),
                                     t(np.concatenate([ed.ravel(), ed.ravel()]).reshape(2, -1)),
                                     t(np.concatenate([nlist, nlist]), torch.int32), 6, 6, 2)
            np.testing.assert_allclose(gn.cpu().numpy().ravel(), g["expected_grad_net"] * 2, atol=1e-5, rtol=0)
        else:
            grad = 10 - np.arange(9) * 1.0
            gn = ops.prod_virial_grad(t(grad.reshape(1, 9)), t(ed.reshape(1, -1)), t(rij.reshape(1, -1)),
                                      t(nlist, torch.int32), 6, 1)
            np.testing.assert_allclose(gn.cpu().numpy().ravel(), g["expected_grad_net"], atol=1e-5, rtol=0)


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32], ids=["f64", "f32"])
def test_prod_grad_kernels_match_oracle(oracle, dtype):
    """Reference rule (grad over local atoms, ghost index j -> j % nloc) on
    an extended list, the exact transpose with nall rows, and the virial
    gradient, on water192 against the oracle."""
    from dpmi import ops

    _say(f"prod grad kernels water192 {dtype}")
    s = systems.water192(4)
    sel = [46, 92]
    ec, et, mp, nlist, _ = oracle.build_nlist_pbc(s.coord, s.atype, s.box, 6.0, sel)
    _, ed, rij = oracle.env_mat_a(nlist, ec, et, np.zeros((2, 138, 4)), np.ones((2, 138, 4)), 6.0, 0.5, sel)
    rng = np.random.default_rng(6)
    nloc, nall = s.nloc, ec.shape[0]
    gf = rng.normal(size=(nloc, 3))
    gv = rng.normal(size=9)
    t = lambda a, dt=dtype: torch.as_tensor(np.ascontiguousarray(a), dtype=dt, device=DEV)
    tol = dict(rtol=1e-12, atol=1e-12) if dtype == torch.float64 else dict(rtol=1e-5, atol=1e-5)
    got = ops.prod_force_grad(t(gf.reshape(1, -1)), t(ed.reshape(1, -1)), t(nlist, torch.int32), nloc, nloc, 1)
    np.testing.assert_allclose(got.cpu().double().numpy().reshape(nloc, -1), oracle.prod_force_grad_a(gf, ed, nlist),
                               **tol)
    # exact transpose of prod_force over the extended atoms
    gfe = rng.normal(size=(nall, 3))
    got = ops.prod_force_grad(t(gfe.reshape(1, -1)), t(ed.reshape(1, -1)), t(nlist, torch.int32), nloc, nall, 1)
    nd = rng.normal(size=(nloc, 4 * 138))
    f = oracle.prod_force_a(nd, ed, nlist, nall)[0]
    lhs = float(np.sum(gfe * f))
    rhs = float(np.sum(got.cpu().double().numpy().reshape(nloc, -1) * nd))
    assert abs(lhs - rhs) <= (1e-10 if dtype == torch.float64 else 1e-4) * abs(lhs)
    got = ops.prod_virial_grad(t(gv.reshape(1, 9)), t(ed.reshape(1, -1)), t(rij.reshape(1, -1)),
                               t(nlist, torch.int32), nloc, 1)
    np.testing.assert_allclose(got.cpu().double().numpy().reshape(nloc, -1),
                               oracle.prod_virial_grad_a(gv, ed, rij, nlist), **tol)


# --------------------------------------------------------------------------- training step

NF, NLOC = 8, 192


@pytest.fixture(scope="module")
def train_case():
    """8 water192 frames (seeds 0-7), model weights seed 7, E/F labels from
    the same architecture with weight seed 8 (oracle, FP64)."""
    from oracle.model_ref import Model

    spec = systems.WATER_SPEC
    params = systems.make_params(spec, seed=7)
    tabs = table.build_tables(params["embedding"], params["davg"], params["dstd"], spec.rcut, spec.rcut_smth, 0.9,
                              spec.sel)
    frames = []
    for k in range(NF):
        s = systems.water192(k)
        frames.append((s.coord, s.atype, s.box))
    lab = Model.from_params(spec, systems.make_params(spec, seed=8))
    el, fl = [], []
    for c, t, b in frames:
        o = lab.evaluate(c, t, b)
        el.append(o["energy"])
        fl.append(o["force"])
    return spec, params, tabs, frames, np.array(el), np.array(fl)


def _gpu_loss_grad(spec, params, tabs, frames, el, fl, precision="fp64"):
    from dpmi import train
    from dpmi.model import SeAEnergyModel

    m = SeAEnergyModel.from_params(spec, params, device=DEV, precision=precision, tables=tabs)
    named = m.named_parameters()
    tr = train.Trainer(m)
    batch = train.prepare(m, *zip(*frames))
    loss, _ = tr.loss_and_grad(batch, torch.as_tensor(el, device=DEV), torch.as_tensor(fl, device=DEV))
    return float(loss), {k: t.grad.detach().double().cpu().numpy() for k, t in named}


@pytest.mark.parametrize("tabulated", [False, True], ids=["network", "table"])
def test_train_grads_match_oracle(train_case, tabulated):
    """FP64 loss and every parameter gradient within 1e-10 relative (of the
    largest entry of that tensor's gradient) of the CPU oracle."""
    from oracle.model_ref import Model

    spec, params, tabs, frames, el, fl = train_case
    tabs = tabs if tabulated else None
    _say(f"train step 8x192 fp64 {'table' if tabulated else 'network'} vs oracle")
    loss, grads = _gpu_loss_grad(spec, params, tabs, frames, el, fl)
    ref = Model.from_params(spec, params, tables=tabs)
    named = ref.named_parameters()
    rloss, rgrads = ref.train_loss(frames, el, fl, params=[t for _, t in named])
    assert abs(loss - rloss) <= 1e-10 * abs(rloss)
    assert set(grads) == {k for k, _ in named}
    for (k, _), rg in zip(named, rgrads):
        scale = np.abs(rg).max()
        np.testing.assert_allclose(grads[k], rg, rtol=0, atol=1e-10 * scale, err_msg=str(k))


def _perturbed(params, key, idx, h):
    import copy

    p = copy.deepcopy(params)
    kind, net, layer, name = key
    arr = np.array(p[kind][net][layer][name], dtype=np.float64)
    arr[idx] += h
    p[kind][net][layer][name] = arr
    return p


def _gpu_loss(spec, params, tabs, frames, el, fl):
    from dpmi import train
    from dpmi.model import SeAEnergyModel

    m = SeAEnergyModel.from_params(spec, params, device=DEV, tables=tabs)
    batch = train.prepare(m, *zip(*frames))
    e, f = train.energy_force(m, batch, create_graph=False)
    loss, _ = train.EnergyStdLoss()(e, f, torch.as_tensor(el, device=DEV), torch.as_tensor(fl, device=DEV), NLOC)
    return float(loss.detach())


@pytest.mark.parametrize("tabulated", [False, True], ids=["network", "table"])
def test_train_grads_finite_difference(train_case, tabulated):
    """Central differences of the GPU loss, h = 1e-5 (FP64)."""
    spec, params, tabs, frames, el, fl = train_case
    tabs = tabs if tabulated else None
    _say(f"train step FD 8x192 fp64 {'table' if tabulated else 'network'}")
    _, grads = _gpu_loss_grad(spec, params, tabs, frames, el, fl)
    h = 1e-5
    keys = [(("fitting", 0, 0, "w"), (5, 3)), (("fitting", 1, 2, "b"), (7,)), (("fitting", 1, 1, "idt"), (11,)),
            (("fitting", 0, 3, "w"), (17, 0))]
    if not tabulated:
        keys += [(("embedding", 0, 0, "w"), (0, 4)), (("embedding", 1, 2, "w"), (10, 20)),
                 (("embedding", 1, 1, "idt"), (3,))]
    for key, idx in keys:
        lp = _gpu_loss(spec, _perturbed(params, key, idx, h), tabs, frames, el, fl)
        lm = _gpu_loss(spec, _perturbed(params, key, idx, -h), tabs, frames, el, fl)
        fd = (lp - lm) / (2 * h)
        g = grads[key][idx]
        assert abs(fd - g) <= 1e-5 * abs(g) + 1e-8, (key, idx, fd, g)


@pytest.mark.parametrize("tabulated", [False, True], ids=["network", "table"])
def test_train_fp32_tracks_fp64(train_case, tabulated):
    """The FP32 form: loss within 1e-4 and the gradient within 1e-3 (relative
    L2 norm over all parameters) of the FP64 step; reported."""
    spec, params, tabs, frames, el, fl = train_case
    tabs = tabs if tabulated else None
    _say(f"train step 8x192 fp32 vs fp64 {'table' if tabulated else 'network'}")
    l64, g64 = _gpu_loss_grad(spec, params, tabs, frames, el, fl)
    tabs32 = None if tabs is None else [(t.astype(np.float32), i) for t, i in tabs]
    l32, g32 = _gpu_loss_grad(spec, params, tabs32, frames, el, fl, precision="fp32")
    num = np.sqrt(sum(np.sum((g32[k] - g64[k]) ** 2) for k in g64))
    den = np.sqrt(sum(np.sum(g64[k] ** 2) for k in g64))
    print(f"fp32 vs fp64: loss {l32:.8e} / {l64:.8e}, gradient rel L2 {num / den:.3e}", flush=True)
    assert abs(l32 - l64) <= 1e-4 * abs(l64)
    assert num / den <= 1e-3


def test_train_steps_reduce_loss(train_case):
    """Ten Adam steps (lr 1e-3) on one batch lower the loss."""
    from dpmi import train
    from dpmi.model import SeAEnergyModel

    spec, params, tabs, frames, el, fl = train_case
    _say("train 10 Adam steps 8x192 fp64")
    m = SeAEnergyModel.from_params(spec, params, device=DEV)
    tr = train.Trainer(m)
    batch = train.prepare(m, *zip(*frames))
    e_l, f_l = torch.as_tensor(el, device=DEV), torch.as_tensor(fl, device=DEV)
    losses = [float(tr.step(batch, e_l, f_l)[0]) for _ in range(10)]
    print("losses", losses, flush=True)
    assert all(np.isfinite(losses))
    assert losses[-1] < losses[0]
