"""The drop-ins used inside the reference's own PyTorch composition: the
GpuNeighborList strategy (NeighborList.build) and dpmi.ops.prod_env_mat in
place of deepmd.pt's prod_env_mat under the descriptor.  The network part
here is the oracle's CPU restatement of DescrptBlockSeA / the fitting
(oracle.model_ref.Model), fed the GPU env-mat through a differentiable copy,
so forces, virials and a training step's parameter gradients flow back
through the HIP gradient kernels and are compared with the oracle's own
all-CPU evaluation."""

import numpy as np
import pytest
import torch

from dpmi import ops, systems
from dpmi.neighbor_list import GpuNeighborList

pytestmark = pytest.mark.gpu

DEV = torch.device("cuda:0")


def _say(msg):
    print(f"[gpu case] {msg}", flush=True)


@pytest.fixture(scope="module", autouse=True)
def _lib():
    from dpmi import _lib as L

    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    L.load()


def _batch(systems_list):
    coord = torch.as_tensor(np.stack([s.coord for s in systems_list]), device=DEV)
    atype = torch.as_tensor(np.stack([s.atype for s in systems_list]).astype(np.int64), device=DEV)
    box = torch.as_tensor(np.stack([s.box for s in systems_list]), device=DEV)
    return coord, atype, box


@pytest.mark.parametrize("kind", ["water", "alloy_triclinic"])
def test_neighbor_list_strategy_matches_oracle(oracle, kind):
    _say(f"GpuNeighborList.build {kind}")
    if kind == "water":
        frames, rc, sel = [systems.water192(0), systems.water192(1)], 6.0, [46, 92]
    else:
        frames, rc, sel = [systems.alloy262k(4, ncell=8), systems.alloy262k(5, ncell=8)], 8.0, [60, 60, 60, 60]
    coord, atype, box = _batch(frames)
    ec, et, nl, mp = GpuNeighborList().build(coord.reshape(len(frames), -1), atype, box, rc, sel)
    assert nl.dtype == torch.int64 and mp.dtype == torch.int64 and nl.shape == (len(frames), frames[0].nloc, sum(sel))
    for f, s in enumerate(frames):
        oc, ot, om, onl, _ = oracle.build_nlist_pbc(s.coord, s.atype, s.box, rc, sel)
        n = oc.shape[0]
        np.testing.assert_array_equal(ec[f, :n].cpu().numpy(), oc)
        np.testing.assert_array_equal(et[f, :n].cpu().numpy(), ot)
        np.testing.assert_array_equal(mp[f, :n].cpu().numpy(), om)
        np.testing.assert_array_equal(nl[f].cpu().numpy(), onl)
        assert (et[f, n:] == -1).all()


def _composition(model, ext_coord, ext_atype, nlist, mapping, nloc):
    """E, F (folded to local atoms), V of one frame: GPU env-mat drop-in,
    then the oracle's descriptor and fitting on the CPU."""
    em, _, _ = ops.prod_env_mat(ext_coord.reshape(1, -1), nlist.reshape(1, nloc, -1), ext_atype[:nloc].reshape(1, -1),
                                torch.as_tensor(model.davg.numpy(), device=DEV),
                                torch.as_tensor(model.dstd.numpy(), device=DEV), model.rcut, model.rcut_smth)
    em_cpu = em[0].cpu()
    at = ext_atype[:nloc].cpu().numpy()
    return model.atom_energy(model.descriptor(em_cpu), at).sum()


def test_env_mat_dropin_energy_force_virial(oracle):
    _say("prod_env_mat drop-in: E, F, V of the 192-atom water model")
    from oracle.model_ref import Model

    spec = systems.WATER_SPEC
    params = systems.make_params(spec, seed=7)
    model = Model.from_params(spec, params)
    s = systems.water192(2)
    want = model.evaluate(s.coord, s.atype, s.box)
    coord, atype, box = _batch([s])
    ec, et, nl, mp = GpuNeighborList().build(coord, atype, box, spec.rcut, list(spec.sel))
    n = int((et[0] >= 0).sum())
    x = ec[0, :n].clone().requires_grad_(True)
    e = _composition(model, x, et[0, :n], nl[0], mp[0, :n], s.nloc)
    (g,) = torch.autograd.grad(e, x)
    fext = -g
    force = torch.zeros((s.nloc, 3), dtype=torch.float64, device=DEV).index_add_(0, mp[0, :n], fext)
    virial = (fext.unsqueeze(-1) * x.detach().unsqueeze(1)).sum(0).reshape(9)
    assert abs(e.item() - want["energy"]) <= 1e-10 * abs(want["energy"])
    np.testing.assert_allclose(force.cpu().numpy(), want["force"], rtol=1e-10,
                               atol=1e-10 * np.abs(want["force"]).max())
    np.testing.assert_allclose(virial.cpu().numpy(), want["virial"], rtol=1e-10,
                               atol=1e-10 * np.abs(want["virial"]).max())


def test_env_mat_dropin_training_gradient(oracle):
    """Parameter gradient of the E + F loss through the drop-in's double
    backward (the JVP kernel) against the oracle's all-CPU training loss."""
    _say("prod_env_mat drop-in: training-step parameter gradient")
    from oracle.model_ref import Model

    spec = systems.WATER_SPEC
    params = systems.make_params(spec, seed=7)
    frames = [systems.water192(k) for k in range(2)]
    lab = Model.from_params(spec, systems.make_params(spec, seed=8))
    rng = np.random.default_rng(5)
    el, fl = [], []
    for s in frames:
        o = lab.evaluate(s.coord, s.atype, s.box)
        el.append(o["energy"] + rng.normal())
        fl.append(o["force"] + 0.5 * rng.normal(size=o["force"].shape))
    el, fl = np.array(el), np.array(fl)
    ref = Model.from_params(spec, params)
    rp = ref.parameters()
    loss_ref, g_ref = ref.train_loss([(s.coord, s.atype, s.box) for s in frames], el, fl, params=rp)

    model = Model.from_params(spec, params)
    mparams = model.parameters()
    coord, atype, box = _batch(frames)
    ec, et, nl, mp = GpuNeighborList().build(coord, atype, box, spec.rcut, list(spec.sel))
    es, fs = [], []
    nloc = frames[0].nloc
    for f in range(len(frames)):
        n = int((et[f] >= 0).sum())
        x = ec[f, :n].clone().requires_grad_(True)
        e = _composition(model, x, et[f, :n], nl[f], mp[f, :n], nloc)
        (g,) = torch.autograd.grad(e, x, create_graph=True)
        force = torch.zeros((nloc, 3), dtype=torch.float64, device=DEV).index_add(0, mp[f, :n], -g)
        es.append(e)
        fs.append(force.cpu())
    energy, force = torch.stack(es), torch.stack(fs)
    l2_e = torch.mean(torch.square(energy - torch.as_tensor(el)))
    l2_f = torch.mean(torch.square((torch.as_tensor(fl) - force).reshape(-1)))
    loss = (1.0 / nloc) * (0.02 * l2_e) + 1000.0 * l2_f
    grads = torch.autograd.grad(loss, mparams)
    assert abs(loss.item() - loss_ref) <= 1e-10 * abs(loss_ref)
    for a, b in zip(grads, g_ref):
        np.testing.assert_allclose(a.numpy(), b, rtol=1e-9, atol=1e-9 * max(np.abs(b).max(), 1e-30))
