"""E+F+V of the GPU model against the model oracle and the reference's
end-to-end golden values."""

import numpy as np
import pytest
import torch

from dpmi import systems

pytestmark = pytest.mark.gpu
DEV = torch.device("cuda:0")


def _say(msg):
    print(f"[gpu case] {msg}", flush=True)


@pytest.fixture(scope="module", autouse=True)
def _need_gpu():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")


@pytest.mark.parametrize("block", [0, 1, 2], ids=["pbc13", "nopbc", "pbc19x13x13"])
def test_deeppot_sea_golden_gpu(golden, block):
    from dpmi.model import SeAEnergyModel

    _say(f"deeppot_sea golden block {block}")
    m = SeAEnergyModel.from_serialized(golden("deeppot_sea_model.json"), device=DEV)
    r = golden("deeppot_sea_cases.json")["results"][block]
    out = m.compute(torch.tensor(r["coord"], dtype=torch.float64, device=DEV).reshape(-1, 3),
                    torch.tensor(r["atype"], device=DEV),
                    r["box"], need_atom_energy=True)
    np.testing.assert_allclose(out["atom_energy"].cpu().numpy(), r["atomic_energy"], rtol=0, atol=1e-10)
    np.testing.assert_allclose(out["force"].cpu().numpy().ravel(), r["force"], rtol=0, atol=1e-11)
    virial = np.array(r["atomic_virial"]).reshape(-1, 9).sum(axis=0)
    np.testing.assert_allclose(out["virial"].cpu().numpy(), virial, rtol=0, atol=1e-11)


@pytest.fixture(scope="module")
def water_model():
    from oracle.model_ref import Model

    from dpmi.stat import gpu_model_params

    spec = systems.WATER_SPEC
    params = gpu_model_params(spec, systems.water192(0))
    return spec, params, Model.from_params(spec, params)


def test_fp64_model_matches_oracle(water_model):
    """FP64 GPU path vs the CPU model oracle: <= 1e-10 relative."""
    from dpmi.model import SeAEnergyModel

    spec, params, ref = water_model
    s = systems.water192(0)
    _say("fp64 model water192 vs oracle")
    m = SeAEnergyModel.from_params(spec, params, device=DEV)
    out = m.compute(torch.tensor(s.coord, device=DEV), torch.tensor(s.atype, device=DEV), s.box,
                    need_atom_energy=True)
    o = ref.evaluate(s.coord, s.atype, s.box)
    assert abs(out["energy"].item() - o["energy"]) <= 1e-10 * abs(o["energy"])
    f = out["force"].cpu().numpy()
    np.testing.assert_allclose(f, o["force"], rtol=1e-10, atol=1e-10 * np.abs(o["force"]).max())
    np.testing.assert_allclose(out["virial"].cpu().numpy(), o["virial"], rtol=1e-10,
                               atol=1e-10 * np.abs(o["virial"]).max())


@pytest.mark.parametrize("precision", ["bf16", "fp16"])
def test_low_precision_forms_error_budget(water_model, precision):
    from dpmi.model import SeAEnergyModel

    spec, params, ref = water_model
    s = systems.water192(0)
    _say(f"{precision} model water192 error")
    m64 = SeAEnergyModel.from_params(spec, params, device=DEV)
    mlo = SeAEnergyModel.from_params(spec, params, device=DEV, precision=precision)
    c = torch.tensor(s.coord, device=DEV)
    t = torch.tensor(s.atype, device=DEV)
    a = m64.compute(c, t, s.box)
    b = mlo.compute(c, t, s.box)
    rmse = torch.sqrt(((a["force"] - b["force"]) ** 2).mean()).item()
    de = abs(a["energy"].item() - b["energy"].item()) / s.nloc
    frms = torch.sqrt((a["force"] ** 2).mean()).item()
    print(f"{precision}: force RMSE {rmse:.3e} eV/A ({rmse / frms:.2e} of RMS force), |dE|/atom {de:.3e} eV",
          flush=True)
    # measured on MI355X: bf16 4.2e-5 eV/A and 2.3e-5 eV/atom, fp16 5.3e-6 and
    # 1.4e-5; the budgets are a few times that (the benchmark's own bound,
    # 1e-3 eV/A, would let a tenfold regression through)
    assert rmse <= {"bf16": 1e-4, "fp16": 2e-5}[precision]
    assert de <= {"bf16": 1e-4, "fp16": 5e-5}[precision]


def test_tabulated_fp64_model_matches_oracle():
    """Compressed model (tabulated embedding, grid [0, 2) / 0.01 + [2, 10) /
    0.1, 280 intervals) at 12,288 water atoms, FP64: E, F, V within 1e-10
    relative of the oracle running the same tables on the CPU."""
    from oracle.model_ref import Model

    from dpmi import table
    from dpmi.model import SeAEnergyModel

    spec = systems.WATER_SPEC
    s = systems.water12k()
    params = systems.make_params(spec, seed=7)  # davg 0, dstd 1: table range [0, 2)
    tabs = table.build_tables(params["embedding"], params["davg"], params["dstd"], spec.rcut, spec.rcut_smth,
                              0.9, spec.sel)
    assert tabs[0][0].shape == (280, 600)
    _say("tabulated fp64 model water12k vs oracle")
    m = SeAEnergyModel.from_params(spec, params, device=DEV, tables=tabs)
    out = m.compute(torch.tensor(s.coord, device=DEV), torch.tensor(s.atype, device=DEV), s.box,
                    need_atom_energy=True)
    o = Model.from_params(spec, params, tables=tabs).evaluate(s.coord, s.atype, s.box)
    assert abs(out["energy"].item() - o["energy"]) <= 1e-10 * abs(o["energy"])
    np.testing.assert_allclose(out["atom_energy"].cpu().numpy(), o["atom_energy"], rtol=1e-10,
                               atol=1e-10 * np.abs(o["atom_energy"]).max())
    f = out["force"].cpu().numpy()
    np.testing.assert_allclose(f, o["force"], rtol=1e-10, atol=1e-10 * np.abs(o["force"]).max())
    np.testing.assert_allclose(out["virial"].cpu().numpy(), o["virial"], rtol=1e-10,
                               atol=1e-10 * np.abs(o["virial"]).max())
    # the table stands in for the network: report how far the compressed
    # model is from the uncompressed one
    full = SeAEnergyModel.from_params(spec, params, device=DEV).compute(
        torch.tensor(s.coord, device=DEV), torch.tensor(s.atype, device=DEV), s.box)
    df = torch.sqrt(((full["force"] - out["force"]) ** 2).mean()).item()
    de = abs(full["energy"].item() - out["energy"].item()) / s.nloc
    print(f"table vs network: force RMSE {df:.3e} eV/A, |dE|/atom {de:.3e} eV", flush=True)


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_tabulated_direct_path_matches_op_chain(precision):
    """compute() of the compressed model launches the table kernels directly
    and runs the fitting on type-sorted G^T R; the same E+F+V through the
    autograd chain of the registered ops (descriptor -> atom_energy ->
    autograd, the reference composition) on the same list: FP64 within
    1e-12, FP32 within 1e-5 (relative)."""
    from dpmi import table
    from dpmi.model import SeAEnergyModel

    spec = systems.WATER_SPEC
    s = systems.water192(3)
    params = systems.make_params(spec, seed=7)
    tabs = table.build_tables(params["embedding"], params["davg"], params["dstd"], spec.rcut, spec.rcut_smth,
                              0.9, spec.sel)
    if precision == "fp32":
        tabs = [(t.astype(np.float32), i.astype(np.float32)) for t, i in tabs]
    _say(f"tabulated direct path vs op chain {precision} water192")
    m = SeAEnergyModel.from_params(spec, params, device=DEV, precision=precision, tables=tabs)
    c = torch.tensor(s.coord, device=DEV)
    t = torch.tensor(s.atype, device=DEV)
    out = m.compute(c, t, s.box, need_atom_energy=True)
    at = t.to(torch.int32)
    ext_c, ext_t, mapping, nlist, _, _, cells = m.neighbor_list(c, at, s.box)
    em = m.env_mat(nlist, ext_c, ext_t).detach().requires_grad_(True)
    with torch.enable_grad():
        ae = m.atom_energy(m.descriptor(em), at)
        (nd,) = torch.autograd.grad(ae.sum(), em)
    force, virial = m.forces(nd, nlist, cells, mapping, ext_c, ext_t, s.box)
    tol = 1e-12 if precision == "fp64" else 1e-5
    assert abs(out["energy"].item() - ae.sum().item()) <= tol * abs(ae.sum().item())
    assert ((out["force"] - force).abs().max() <= tol * force.abs().max()).item()
    assert ((out["virial"] - virial).abs().max() <= tol * virial.abs().max()).item()
    assert ((out["atom_energy"] - ae.detach()).abs().max() <= tol * ae.detach().abs().max()).item()


@pytest.fixture(scope="module")
def alloy_model():
    from dpmi.stat import gpu_model_params

    spec = systems.ALLOY_SPEC
    return spec, gpu_model_params(spec, systems.alloy262k(4, ncell=8))


def test_alloy_fp64_matches_oracle(alloy_model):
    """Four-type FCC alloy (triclinic box, rcut 8, sel [60]*4, embedding
    [32, 64, 128]) at 1,000 atoms, FP64: E, F, V within 1e-10 relative."""
    from oracle.model_ref import Model

    from dpmi.model import SeAEnergyModel

    spec, params = alloy_model
    s = systems.alloy262k(5, ncell=10)
    _say("fp64 model alloy1000 vs oracle")
    out = SeAEnergyModel.from_params(spec, params, device=DEV).compute(
        torch.tensor(s.coord, device=DEV), torch.tensor(s.atype, device=DEV), s.box)
    o = Model.from_params(spec, params).evaluate(s.coord, s.atype, s.box)
    assert abs(out["energy"].item() - o["energy"]) <= 1e-10 * abs(o["energy"])
    np.testing.assert_allclose(out["force"].cpu().numpy(), o["force"], rtol=1e-10,
                               atol=1e-10 * np.abs(o["force"]).max())
    np.testing.assert_allclose(out["virial"].cpu().numpy(), o["virial"], rtol=1e-10,
                               atol=1e-10 * np.abs(o["virial"]).max())


def test_alloy_bf16_error_budget(alloy_model):
    """BF16-MFMA form on the alloy (4,096 atoms): force RMSE <= 1e-3 eV/A
    against FP64; energy/atom and virial errors reported."""
    from dpmi.model import SeAEnergyModel

    spec, params = alloy_model
    s = systems.alloy262k(4, ncell=16)
    _say("bf16 model alloy4096 error")
    c = torch.tensor(s.coord, device=DEV)
    t = torch.tensor(s.atype, device=DEV)
    a = SeAEnergyModel.from_params(spec, params, device=DEV).compute(c, t, s.box)
    m = SeAEnergyModel.from_params(spec, params, device=DEV, precision="bf16")
    assert m.fused is not None and m.fused_fit is not None
    b = m.compute(c, t, s.box)
    rmse = torch.sqrt(((a["force"] - b["force"]) ** 2).mean()).item()
    de = abs(a["energy"].item() - b["energy"].item()) / s.nloc
    dv = (torch.abs(a["virial"] - b["virial"]).max() / torch.abs(a["virial"]).max()).item()
    print(f"alloy bf16: force RMSE {rmse:.3e} eV/A, |dE|/atom {de:.3e} eV, virial max rel {dv:.3e}", flush=True)
    # measured 2.4e-5 eV/A, 1.9e-5 eV/atom, 1.5e-4
    assert rmse <= 1e-4
    assert de <= 1e-4
    assert dv <= 1e-3


@pytest.mark.parametrize("precision", ["fp64", "bf16"])
def test_compute_lower_external_list(oracle, water_model, precision):
    """forward_lower semantics: extended coordinates + an external (shuffled,
    unsorted) neighbor list give the same E, F, V as compute(); without a
    mapping the forces stay on the extended atoms and fold to the same."""
    from dpmi.model import SeAEnergyModel

    spec, params, _ = water_model
    s = systems.water192(1)
    _say(f"compute_lower external list water192 {precision}")
    m = SeAEnergyModel.from_params(spec, params, device=DEV, precision=precision)
    c = torch.tensor(s.coord, device=DEV)
    t = torch.tensor(s.atype, device=DEV)
    ref = m.compute(c, t, s.box, need_atom_energy=True)
    oc, ot, om, _, _ = oracle.build_nlist_pbc(s.coord, s.atype, s.box, spec.rcut, spec.sel)
    nn, cand, mem = oracle.candidates(oc, s.nloc, spec.rcut, box=s.box)
    rng = np.random.default_rng(3)
    for i in range(s.nloc):
        rng.shuffle(cand[i, :nn[i]])
    ext_c = torch.tensor(oc, device=DEV)
    ext_t = torch.tensor(ot, device=DEV)
    out = m.compute_lower(ext_c, ext_t, torch.tensor(cand, device=DEV), s.nloc, numneigh=torch.tensor(nn, device=DEV),
                          mapping=torch.tensor(om, device=DEV))
    tol = 1e-12 if precision == "fp64" else 1e-5
    assert abs(out["energy"].item() - ref["energy"].item()) <= tol * abs(ref["energy"].item())
    np.testing.assert_allclose(out["force"].cpu().numpy(), ref["force"].cpu().numpy(), rtol=tol,
                               atol=tol * ref["force"].abs().max().item())
    np.testing.assert_allclose(out["virial"].cpu().numpy(), ref["virial"].cpu().numpy(), rtol=tol,
                               atol=tol * ref["virial"].abs().max().item())
    ext = m.compute_lower(ext_c, ext_t, torch.tensor(cand, device=DEV), s.nloc, numneigh=torch.tensor(nn, device=DEV))
    fext = ext["force"].cpu().numpy()
    assert fext.shape == (oc.shape[0], 3)
    folded = np.zeros((s.nloc, 3))
    np.add.at(folded, om, fext)
    np.testing.assert_allclose(folded, ref["force"].cpu().numpy(), rtol=tol, atol=tol * ref["force"].abs().max().item())


def test_table_builder_on_gpu():
    """The table builder's net chain on the GPU gives the CPU tables: every
    coefficient's contribution over its interval, a_k t^k, to 1e-12 of the
    largest (a_3..a_5 themselves divide last-bit differences of the node
    values by t^3..t^5, so they are compared at that scale)."""
    from dpmi import table

    spec = systems.WATER_SPEC
    params = systems.make_params(spec, seed=7)
    _say("table builder on the GPU")
    for fold in (False, True):
        a = table.build_tables(params["embedding"], params["davg"], params["dstd"], spec.rcut, spec.rcut_smth, 0.9,
                               spec.sel, fold_idt=fold)
        b = table.build_tables(params["embedding"], params["davg"], params["dstd"], spec.rcut, spec.rcut_smth, 0.9,
                               spec.sel, fold_idt=fold, device=DEV)
        for (ta, ia), (tb, ib) in zip(a, b):
            np.testing.assert_array_equal(ia, ib)
            n = ta.shape[0]
            tt = np.where(np.arange(n) < round((ia[1] - ia[0]) / ia[3]), ia[3], ia[4])
            scale = tt[:, None, None] ** np.arange(6)[None, None, :]
            xa = ta.reshape(n, -1, 6) * scale
            xb = tb.reshape(n, -1, 6) * scale
            np.testing.assert_allclose(xb, xa, rtol=0, atol=1e-12 * np.abs(xa).max())


@pytest.mark.parametrize("neuron,resnet", [((32, 64, 128), True), ((6, 13, 26, 26), True), ((5, 7, 9), False)],
                         ids=["alloy", "four_layers", "no_skip"])
def test_table_builder_kernel_node_values(neuron, resnet):
    """dpmi_tabulate_build_f64's value / slope / curvature at the grid nodes
    against the host chain (itself pinned to torch autograd), for nets with
    and without skip connections and timesteps, and its coefficients."""
    from dpmi import table

    spec = systems.ModelSpec(ntypes=1, sel=(10,), neuron=neuron, resnet_dt=resnet)
    layers = systems.make_params(spec, seed=3)["embedding"][0]
    _say(f"table builder kernel {neuron}")
    xx = np.append(np.arange(-1.0, 2.0, 0.01), np.arange(2.0, 10.0, 0.1))
    tt = np.diff(xx)
    for fold in (False, True):
        tab, vdd = table.build_on_gpu(xx, tt, layers, fold_idt=fold, device=DEV)
        ref = table.mlp_value_derivs(xx, layers, fold)
        for k in range(3):
            np.testing.assert_allclose(vdd[k].cpu().numpy(), ref[k], rtol=1e-13, atol=1e-13 * np.abs(ref[k]).max())
        want = table.hermite_table(*ref, tt[:, None])
        n = tt.shape[0]
        scale = tt[:, None, None] ** np.arange(6)[None, None, :]
        np.testing.assert_allclose(tab.cpu().numpy().reshape(n, -1, 6) * scale, want.reshape(n, -1, 6) * scale,
                                   rtol=0, atol=1e-12 * np.abs(want.reshape(n, -1, 6) * scale).max())
