"""Seeded synthetic systems and model parameters for the se_e2_a workloads.

Everything here is deterministic from an integer seed through SplitMix64, so
coordinates and weights are bit-identical on every machine.  SplitMix64 is the
published generator of Steele, Lea and Flood ("Fast splittable pseudorandom
number generators", OOPSLA 2014); uniforms take the top 53 bits, normals use
the Box-Muller transform.

Workloads (sizes fixed by the benchmark definition):

* ``water192``     192 atoms (64 O-like type 0, 128 H-like type 1), cubic box
  12.42 A: 4x4x4 cells of 3 sites each, sites at fractional (0,0,0),
  (1/3,1/3,1/3), (2/3,2/3,2/3) of the cell; the first site of each cell is
  type 0.
* ``water12k``     12,288 atoms, 16x16x16 cells of 3 sites, box 49.68 A.
* ``water110k``    110,592 atoms on a 48^3 simple-cubic lattice, box 99.36 A;
  site s has type 0 when s % 3 == 0 (36,864 / 73,728).
* ``alloy262k``    262,144 atoms, 64^3 FCC primitive cells (a = 3.6 A,
  triclinic box with rows 64*(a/2)*(0,1,1), (1,0,1), (1,1,0)), four types
  drawn uniformly.

Water-like systems are jittered by U(-0.3, 0.3) A per component; any atom
closer than 0.9 A to another (minimum image) is re-jittered from the next
draws until no such pair is left.  The alloy uses U(-0.15, 0.15) A.
"""

from __future__ import annotations

import dataclasses

import numpy as np

_GAMMA = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


class SplitMix64:
    """Vectorised SplitMix64 stream."""

    def __init__(self, seed: int) -> None:
        self.state = np.uint64(seed & 0xFFFFFFFFFFFFFFFF)

    def next_u64(self, n: int) -> np.ndarray:
        with np.errstate(over="ignore"):
            steps = np.arange(1, n + 1, dtype=np.uint64)
            z = self.state + steps * _GAMMA
            self.state = self.state + np.uint64(n) * _GAMMA
            z = (z ^ (z >> np.uint64(30))) * _M1
            z = (z ^ (z >> np.uint64(27))) * _M2
            z = z ^ (z >> np.uint64(31))
        return z

    def uniform(self, n: int) -> np.ndarray:
        """Doubles in [0, 1) from the top 53 bits."""
        return (self.next_u64(n) >> np.uint64(11)).astype(np.float64) * (2.0**-53)

    def normal(self, n: int) -> np.ndarray:
        m = (n + 1) // 2
        u1 = self.uniform(m)
        u2 = self.uniform(m)
        r = np.sqrt(-2.0 * np.log1p(-u1))
        z = np.concatenate([r * np.cos(2 * np.pi * u2), r * np.sin(2 * np.pi * u2)])
        return z[:n]


@dataclasses.dataclass
class System:
    name: str
    coord: np.ndarray  # [nloc, 3] float64
    atype: np.ndarray  # [nloc] int32
    box: np.ndarray  # [9] float64, rows are cell vectors

    @property
    def nloc(self) -> int:
        return int(self.atype.shape[0])


def _min_image_close_pairs(coord: np.ndarray, L: float, dmin: float) -> np.ndarray:
    """Atoms (higher index of each pair) closer than dmin, cubic box."""
    from scipy.spatial import cKDTree

    wrapped = np.mod(coord, L)
    wrapped[wrapped >= L] = 0.0
    tree = cKDTree(wrapped, boxsize=L)
    pairs = tree.query_pairs(dmin, output_type="ndarray")
    if pairs.size == 0:
        return np.zeros(0, dtype=np.int64)
    return np.unique(pairs.max(axis=1))


def _water_like(name: str, ncell: int, L: float, seed: int, three_site: bool) -> System:
    rng = SplitMix64(seed)
    if three_site:
        a = L / ncell
        g = np.arange(ncell, dtype=np.float64)
        cx, cy, cz = np.meshgrid(g, g, g, indexing="ij")
        cells = np.stack([cx.ravel(), cy.ravel(), cz.ravel()], axis=1) * a
        offs = np.array([[0.0, 0.0, 0.0], [1 / 3, 1 / 3, 1 / 3], [2 / 3, 2 / 3, 2 / 3]]) * a
        sites = (cells[:, None, :] + offs[None, :, :]).reshape(-1, 3)
        atype = np.tile(np.array([0, 1, 1], dtype=np.int32), ncell**3)
    else:
        a = L / ncell
        g = np.arange(ncell, dtype=np.float64)
        cx, cy, cz = np.meshgrid(g, g, g, indexing="ij")
        sites = np.stack([cx.ravel(), cy.ravel(), cz.ravel()], axis=1) * a
        atype = (np.arange(sites.shape[0]) % 3 != 0).astype(np.int32)
    n = sites.shape[0]
    jit = (rng.uniform(3 * n).reshape(n, 3) - 0.5) * 0.6
    coord = sites + jit
    for _ in range(1000):
        bad = _min_image_close_pairs(coord, L, 0.9)
        if bad.size == 0:
            break
        jit = (rng.uniform(3 * bad.size).reshape(bad.size, 3) - 0.5) * 0.6
        coord[bad] = sites[bad] + jit
    else:  # pragma: no cover - would need an absurd density
        raise RuntimeError("could not place atoms with min distance 0.9 A")
    box = np.array([L, 0, 0, 0, L, 0, 0, 0, L], dtype=np.float64)
    return System(name, np.ascontiguousarray(coord), atype, box)


def water192(seed: int = 0) -> System:
    return _water_like("water192", 4, 12.42, seed, True)


def water12k(seed: int = 1) -> System:
    return _water_like("water12k", 16, 49.68, seed, True)


def water110k(seed: int = 3) -> System:
    return _water_like("water110k", 48, 99.36, seed, False)


def alloy262k(seed: int = 4, ncell: int = 64) -> System:
    rng = SplitMix64(seed)
    a = 3.6
    v = 0.5 * a * np.array([[0.0, 1.0, 1.0], [1.0, 0.0, 1.0], [1.0, 1.0, 0.0]])
    g = np.arange(ncell, dtype=np.float64)
    i0, i1, i2 = np.meshgrid(g, g, g, indexing="ij")
    frac = np.stack([i0.ravel(), i1.ravel(), i2.ravel()], axis=1)
    sites = frac @ v
    n = sites.shape[0]
    jit = (rng.uniform(3 * n).reshape(n, 3) - 0.5) * 0.3
    atype = np.minimum((rng.uniform(n) * 4).astype(np.int32), 3)
    box = (ncell * v).reshape(9)
    return System(f"alloy{n}", np.ascontiguousarray(sites + jit), atype, box)


SYSTEMS = {
    "water192": water192,
    "water12k": water12k,
    "water110k": water110k,
    "alloy262k": alloy262k,
}


@dataclasses.dataclass
class ModelSpec:
    """Hyper-parameters of an se_e2_a energy model (type_one_side)."""

    ntypes: int = 2
    rcut: float = 6.0
    rcut_smth: float = 0.5
    sel: tuple = (46, 92)
    neuron: tuple = (25, 50, 100)
    axis_neuron: int = 16
    fitting: tuple = (240, 240, 240)
    resnet_dt: bool = True
    fitting_resnet_dt: bool = True


WATER_SPEC = ModelSpec()
ALLOY_SPEC = ModelSpec(
    ntypes=4,
    rcut=8.0,
    rcut_smth=2.0,
    sel=(60, 60, 60, 60),
    neuron=(32, 64, 128),
    axis_neuron=16,
    fitting=(240, 240, 240),
)


def make_params(spec: ModelSpec, seed: int = 7, bias_atom_e=None) -> dict:
    """Random parameters: W ~ N(0, 1/fan_in), b ~ N(0, 1), idt ~ N(0.1, 1e-3).

    Layout follows the reference serialization: every layer is a dict with
    ``w`` [in, out], ``b`` [out], ``idt`` [out] or None.  A layer carries a
    timestep only where it has a skip connection (out == in or out == 2*in),
    as in deepmd/dpmodel/utils/network.py NativeLayer.__init__.
    davg = 0, dstd = 1.
    """
    rng = SplitMix64(seed)

    def layer(nin: int, nout: int, use_idt: bool) -> dict:
        w = rng.normal(nin * nout).reshape(nin, nout) / np.sqrt(nin)
        b = rng.normal(nout)
        idt = 0.1 + 0.001 * rng.normal(nout) if use_idt else None
        return {"w": w, "b": b, "idt": idt}

    emb = []
    for _t in range(spec.ntypes):
        layers = []
        nin = 1
        for nout in spec.neuron:
            # the timestep exists only where a skip connection does
            use_idt = spec.resnet_dt and (nout == nin or nout == 2 * nin)
            layers.append(layer(nin, nout, use_idt))
            nin = nout
        emb.append(layers)
    dim_descrpt = spec.neuron[-1] * spec.axis_neuron
    fit = []
    for _t in range(spec.ntypes):
        layers = []
        nin = dim_descrpt
        for nout in spec.fitting:
            use_idt = spec.fitting_resnet_dt and (nout == nin or nout == 2 * nin)
            layers.append(layer(nin, nout, use_idt))
            nin = nout
        layers.append(layer(nin, 1, False))
        fit.append(layers)
    nnei = int(sum(spec.sel))
    if bias_atom_e is None:
        bias_atom_e = np.array([-93.0, -187.0, -150.0, -120.0][: spec.ntypes])
    return {
        "embedding": emb,
        "fitting": fit,
        "davg": np.zeros((spec.ntypes, nnei, 4)),
        "dstd": np.ones((spec.ntypes, nnei, 4)),
        "bias_atom_e": np.asarray(bias_atom_e, dtype=np.float64).reshape(spec.ntypes),
    }
