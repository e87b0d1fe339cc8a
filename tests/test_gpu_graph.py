"""The neighbor-list pipeline without host waits (dpmi_*_nowait_f64) against
the waiting one, its error path, and whole E+F+V steps captured in a HIP graph
and replayed (with eager work between replays and new coordinates) against
eager steps."""

import ctypes

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


def _clustered():
    rng = np.random.default_rng(11)
    spread = rng.uniform(0, 40, (3000, 3))
    v = rng.normal(size=(600, 3))
    ball = 20 + v / np.linalg.norm(v, axis=1, keepdims=True) * 2.5 * rng.uniform(0, 1, (600, 1)) ** (1 / 3)
    coord = np.concatenate([spread, ball])
    atype = rng.integers(0, 2, coord.shape[0]).astype(np.int32)
    return systems.System("clustered", coord, atype, np.diag([40.0, 40.0, 40.0]).ravel())


CASES = [
    ("water192", lambda: systems.water192(0), 6.0, [46, 92]),
    ("water12k", lambda: systems.water12k(1), 6.0, [46, 92]),
    ("alloy_triclinic_4k", lambda: systems.alloy262k(4, ncell=16), 8.0, [60, 60, 60, 60]),
    ("clustered", _clustered, 6.0, [200, 200]),          # atoms through the per-atom pass
    ("water12k_rcut9", lambda: systems.water12k(1), 9.0, [110, 220]),  # cells over the first table
    ("box_just_over_8rc", lambda: systems._water_like("w48", 16, 48.0001, 1, True), 6.0, [46, 92]),
]


@pytest.mark.parametrize("name,make,rcut,sel", CASES, ids=[c[0] for c in CASES])
def test_nowait_pipeline_matches(name, make, rcut, sel):
    s = make()
    _say(f"nlist without waits {name}")
    c, t = torch.tensor(s.coord, device=DEV), torch.tensor(s.atype, device=DEV)
    a = ops.build_nlist(c, t, s.box, rcut, sel, mapped=True, cells=True)
    b = ops.build_nlist(c, t, s.box, rcut, sel, mapped=True, cells=True, wait=False)
    nall = a[0].shape[0]
    assert b[0].shape[0] >= nall
    for k in range(3):
        np.testing.assert_array_equal(b[k][:nall].cpu().numpy(), a[k].cpu().numpy())
    for k in (3, 4, 5):
        np.testing.assert_array_equal(b[k].cpu().numpy(), a[k].cpu().numpy())
    np.testing.assert_array_equal(b[6][0].cpu().numpy(), a[6][0].cpu().numpy())
    assert int(ops.nlist_status(b[6][1]).item()) == 0


def test_nowait_foreign_workspace_poisons():
    """A workspace prepared for another system: status bit 2, every list
    entry -1 (safe indices for later kernels) and NaN forces."""
    big, small = systems.water12k(1), systems.water192(0)
    _say("nlist without waits: foreign workspace")
    sel = np.array([46, 92], dtype=np.int32)
    *_, (cslot_b, ws_b) = ops.build_nlist(torch.tensor(big.coord, device=DEV), torch.tensor(big.atype, device=DEV),
                                          big.box, 6.0, sel, cells=True, wait=False)
    ext_c, ext_t, mp, *_ = ops.build_nlist(torch.tensor(small.coord, device=DEV),
                                           torch.tensor(small.atype, device=DEV), small.box, 6.0, sel, wait=False)
    nl = torch.full((small.nloc, 138), 5, dtype=torch.int32, device=DEV)
    cs = torch.full((small.nloc, 138), 5, dtype=torch.int16, device=DEV)
    _lib.call("dpmi_build_nlist_nowait_f64", _lib.ptr(nl), None, None, _lib.ptr(cs), _lib.ptr(ext_c), _lib.ptr(ext_t),
              _lib.ptr(mp), small.nloc, 6.0, _lib.ptr(sel), 2, _lib.ptr(ws_b), ws_b.numel(), ops._stream())
    assert int(ops.nlist_status(ws_b).item()) & 2
    assert (nl == -1).all() and (cs == -1).all()
    std = torch.ones((2, 138, 4), dtype=torch.float32, device=DEV)
    nd = torch.ones((small.nloc, 138, 4), dtype=torch.float32, device=DEV)
    f, v = ops.force_virial_cells(nd, nl, (cs, ws_b), mp, ext_c, ext_t, std, 6.0, 0.5, sel, small.box)
    assert torch.isnan(f).all()


def _models(kind):
    from dpmi import table
    from dpmi.model import SeAEnergyModel
    from dpmi.stat import gpu_model_params

    spec = systems.WATER_SPEC
    if kind == "tabulated":
        params = systems.make_params(spec, seed=7)
        tabs = table.build_tables(params["embedding"], params["davg"], params["dstd"], spec.rcut, spec.rcut_smth,
                                  0.9, spec.sel)
        return SeAEnergyModel.from_params(spec, params, device=DEV, tables=tabs)
    params = gpu_model_params(spec, systems.water192(0))
    return SeAEnergyModel.from_params(spec, params, device=DEV, precision=kind)


@pytest.mark.parametrize("kind", ["bf16", "fp64", "tabulated"])
def test_graph_replays_match_eager_steps(kind):
    """The whole step (list rebuilt, E, F, V) captured once; replays with
    eager steps and synchronisations between them, and after new coordinates
    are copied into the captured input, equal eager steps bit for bit."""
    s = systems.water12k(1)
    _say(f"graph capture {kind} water12k")
    m = _models(kind)
    coord = torch.tensor(s.coord, device=DEV)
    atype = torch.tensor(s.atype, device=DEV)
    ref = m.compute(coord, atype, s.box)
    m.check_status()
    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(side):
        for _ in range(2):
            m.compute(coord, atype, s.box)
    torch.cuda.current_stream().wait_stream(side)
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        out = m.compute(coord, atype, s.box)

    def same(a, b):
        for k in ("energy", "force", "virial"):
            np.testing.assert_array_equal(a[k].cpu().numpy(), b[k].cpu().numpy())

    for _ in range(3):
        g.replay()
        torch.cuda.synchronize()
        same(out, ref)
        eager = m.compute(coord, atype, s.box)  # eager work between replays
        torch.cuda.synchronize()
        same(eager, ref)
    # new coordinates in the captured input: the replay rebuilds the list
    rng = np.random.default_rng(3)
    c1 = s.coord + rng.uniform(-0.05, 0.05, s.coord.shape)
    coord.copy_(torch.tensor(c1, device=DEV))
    g.replay()
    torch.cuda.synchronize()
    ref1 = m.compute(torch.tensor(c1, device=DEV), atype, s.box)
    m.check_status()
    same(out, ref1)
    assert abs(out["energy"].item() - ref["energy"].item()) > 0
