"""Pin the model oracle to the reference's end-to-end golden values
(source/tests/infer/deeppot-testcase.yaml for deeppot_sea.yaml) and to
finite differences."""

import numpy as np
import pytest


@pytest.fixture(scope="module")
def model(golden):
    from oracle.model_ref import Model

    return Model.from_serialized(golden("deeppot_sea_model.json"))


@pytest.mark.parametrize("block", [0, 1, 2], ids=["pbc13", "nopbc", "pbc19x13x13"])
def test_deeppot_sea_golden(model, golden, block):
    r = golden("deeppot_sea_cases.json")["results"][block]
    out = model.evaluate(np.array(r["coord"]), np.array(r["atype"]), None if r["box"] is None else np.array(r["box"]))
    # reference tolerance: 12 decimal places (source/tests/infer/test_models.py:30-32)
    np.testing.assert_allclose(out["atom_energy"], r["atomic_energy"], rtol=0, atol=1e-10)
    np.testing.assert_allclose(out["force"].ravel(), r["force"], rtol=0, atol=1e-12)
    virial = np.array(r["atomic_virial"]).reshape(-1, 9).sum(axis=0)
    np.testing.assert_allclose(out["virial"], virial, rtol=0, atol=1e-12)
    if "descriptor" in r:
        np.testing.assert_allclose(out["descriptor"].ravel(), r["descriptor"], rtol=0, atol=1e-12)


def test_force_is_minus_energy_gradient(model):
    from dpmi import systems

    s = systems.water192(5)
    s.coord = s.coord[:60]
    s.atype = s.atype[:60]
    base = model.evaluate(s.coord, s.atype, s.box)
    h = 1e-5
    for i, d in [(0, 0), (17, 1), (59, 2)]:
        cp = s.coord.copy()
        cm = s.coord.copy()
        cp[i, d] += h
        cm[i, d] -= h
        fd = -(model.evaluate(cp, s.atype, s.box)["energy"] - model.evaluate(cm, s.atype, s.box)["energy"]) / (2 * h)
        assert abs(fd - base["force"][i, d]) < 1e-6 * max(1, abs(fd))


def test_virial_is_cell_derivative(model):
    """W = -dE/d(eps) for a homogeneous strain x -> (1+eps) x of atoms and box
    (the reference FD check, source/tests/pt/model/test_autodiff.py:165-215)."""
    from dpmi import systems

    s = systems.water192(6)
    s.coord = s.coord[:60]
    s.atype = s.atype[:60]
    base = model.evaluate(s.coord, s.atype, s.box)
    h = 1e-5
    box = s.box.reshape(3, 3)
    for a, b in [(0, 0), (1, 2), (2, 1)]:
        def e(eps):
            strain = np.eye(3)
            strain[a, b] += eps
            return model.evaluate(s.coord @ strain, s.atype, (box @ strain).ravel())["energy"]

        fd = -(e(h) - e(-h)) / (2 * h)
        assert abs(fd - base["virial"].reshape(3, 3)[b, a]) < 1e-5 * max(1, abs(fd))


def test_tabulated_oracle_tracks_network():
    """The compressed oracle (C tabulate_fusion_se_a + its grad in place of
    embedding + G^T R) with a table of the network as evaluated
    (timestep folded in) agrees with the uncompressed oracle to the table's
    interpolation error, and its forces are minus its energy gradient."""
    from oracle.model_ref import Model

    from dpmi import systems, table

    spec = systems.WATER_SPEC
    params = systems.make_params(spec, seed=3)
    tabs = table.build_tables(params["embedding"], params["davg"], params["dstd"], spec.rcut, spec.rcut_smth,
                              0.9, spec.sel, fold_idt=True)
    s = systems.water192(2)
    s.coord = s.coord[:48]
    s.atype = s.atype[:48]
    net = Model.from_params(spec, params).evaluate(s.coord, s.atype, s.box)
    tm = Model.from_params(spec, params, tables=tabs)
    tab = tm.evaluate(s.coord, s.atype, s.box)
    assert abs(tab["energy"] - net["energy"]) < 1e-7 * s.nloc
    np.testing.assert_allclose(tab["force"], net["force"], rtol=0, atol=1e-6)
    np.testing.assert_allclose(tab["virial"], net["virial"], rtol=0, atol=1e-5)
    h = 1e-5
    for i, d in [(3, 0), (30, 2)]:
        cp = s.coord.copy()
        cm = s.coord.copy()
        cp[i, d] += h
        cm[i, d] -= h
        fd = -(tm.evaluate(cp, s.atype, s.box)["energy"] - tm.evaluate(cm, s.atype, s.box)["energy"]) / (2 * h)
        assert abs(fd - tab["force"][i, d]) < 1e-6 * max(1, abs(fd))
