"""The FP64 fused descriptor + fitting kernel (dpmi_se_a_fitting_gr_f64)
against float64 torch autograd of the same composition (se_a.py:851-858,
fitting.py:734-900): atom energies and dE/d(sum of G^T R partials), and the
compensated energy sum against exact rational summation."""

from fractions import Fraction

import numpy as np
import pytest
import torch

from dpmi import fitting, systems

pytestmark = pytest.mark.gpu
DEV = torch.device("cuda:0")


def _say(msg):
    print(f"[gpu case] {msg}", flush=True)


@pytest.fixture(scope="module", autouse=True)
def _need_gpu():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from dpmi import _lib

    _lib.load()


def _fit_params(ntypes, M, H, seed):
    spec = systems.ModelSpec(ntypes=ntypes, sel=tuple([8] * ntypes), neuron=(M // 4, M // 2, M), axis_neuron=16,
                             fitting=(H, H, H))
    return systems.make_params(spec, seed=seed)


def _reference(params, parts, atype, nnei):
    """float64 torch autograd: GR = sum(parts) / nnei, D = GR^T GR[:, :16],
    own-type fitting net; (atom energies, dE/d(sum of parts))."""
    from dpmi.model import Net

    fit = [Net(layers, DEV, torch.float64, last_linear=True) for layers in params["fitting"]]
    p = parts.clone().requires_grad_(True)
    gr = p.sum(dim=0) / nnei
    d = torch.bmm(gr.transpose(1, 2), gr[:, :, :16].contiguous()).reshape(gr.shape[0], -1)
    ae = torch.zeros(gr.shape[0], dtype=torch.float64, device=DEV)
    for t, net in enumerate(fit):
        y = net(d)[:, 0] + float(params["bias_atom_e"][t])
        ae = torch.where(atype == t, y, ae)
    (g,) = torch.autograd.grad(ae.sum(), p)
    return ae.detach(), g[0]


@pytest.mark.parametrize("ntypes,M,H,nloc,nparts", [
    (2, 100, 240, 1000, 2),    # the water model's shapes; partial 32-atom tiles
    (4, 128, 240, 777, 4),     # the alloy's
    (2, 100, 100, 300, 1),     # half the waves without hidden features
    (3, 64, 50, 97, 2),        # hidden width not a multiple of 16 (nor of 4 K steps' 16)
    (1, 16, 240, 33, 1),       # narrowest G^T R (M = axis)
])
def test_fit64_matches_torch(ntypes, M, H, nloc, nparts):
    _say(f"fit64 ntypes={ntypes} M={M} H={H} nloc={nloc} nparts={nparts}")
    params = _fit_params(ntypes, M, H, seed=M + H)
    g = torch.Generator(device="cpu").manual_seed(nloc)
    nnei = 138
    # partials of the magnitude the embedding produces (GR entries O(1))
    parts = (torch.randn((nparts, nloc, 4, M), generator=g, dtype=torch.float64) * 0.5 * nnei / nparts).to(DEV)
    atype = torch.randint(0, ntypes, (nloc,), generator=g).to(torch.int32).to(DEV)
    ff = fitting.FusedFitting64(params["fitting"], params["bias_atom_e"], None, M, DEV)
    ae, dpart = ff.from_parts(parts, atype, nnei)
    ae_ref, d_ref = _reference(params, parts, atype.long(), nnei)
    np.testing.assert_allclose(ae.cpu().numpy(), ae_ref.cpu().numpy(), rtol=1e-12,
                               atol=1e-12 * ae_ref.abs().max().item())
    np.testing.assert_allclose(dpart.cpu().numpy(), d_ref.cpu().numpy(), rtol=1e-11,
                               atol=1e-12 * d_ref.abs().max().item())
    # deterministic: a second launch gives the same bits
    ae2, d2 = ff.from_parts(parts, atype, nnei)
    assert torch.equal(ae2, ae) and torch.equal(d2, dpart)


def test_energy_sum_compensated():
    """Neumaier sum against the exact rational sum of the same doubles, on
    values whose plain sum loses most digits (large terms that cancel)."""
    _say("energy sum")
    rng = np.random.default_rng(0)
    n = 110592
    x = rng.normal(size=n) * 1e-3
    x[::2] += 1e8
    x[1::2] -= 1e8
    x = x[rng.permutation(n)]
    exact = float(sum((Fraction(v) for v in x), Fraction(0)))
    got = fitting.energy_sum(torch.as_tensor(x, device=DEV)).item()
    plain = float(np.sum(x))
    print(f"energy sum: compensated error {abs(got - exact):.3e}, plain {abs(plain - exact):.3e}", flush=True)
    # Neumaier: error <= eps |S| + O(n eps^2) sum |x|, about 1e-14 here
    assert abs(got - exact) <= 1e-12
    assert fitting.energy_sum(torch.zeros(0, dtype=torch.float64, device=DEV)).item() == 0.0
    assert fitting.energy_sum(torch.as_tensor([3.5], dtype=torch.float64, device=DEV)).item() == 3.5
