"""Per-type environment-matrix statistics (davg, dstd) for se_e2_a.

Restates EnvMatStatSe (deepmd/dpmodel/utils/env_mat_stat.py:137-340) and
StatItem (deepmd/utils/env_mat_stat.py:23-100): for centre type t, every
neighbor slot (padding included, as zeros) contributes to the radial
statistic r_t (component 0) and, three values per slot, to the angular
statistic a_t (components 1..3).  davg = [mean(r_t), 0, 0, 0] and
dstd = [std(r_t), std(a_t), std(a_t), std(a_t)], tiled over the nnei slots;
a std below 1e-2 is replaced by 1e-2, an empty type gets mean 0 / std 0.1.
"""

from __future__ import annotations

import numpy as np
import torch


def env_mat_stat(em_raw, atype, ntypes: int, nnei: int):
    """em_raw: un-normalised em [nloc, nnei, 4] (davg = 0, dstd = 1), any device."""
    em = torch.as_tensor(em_raw).to(torch.float64).reshape(-1, nnei, 4)
    at = torch.as_tensor(atype).reshape(-1).to(em.device)
    davg = np.zeros((ntypes, nnei, 4))
    dstd = np.ones((ntypes, nnei, 4))
    for t in range(ntypes):
        dd = em[at == t].reshape(-1, 4)
        stats = []
        for part in (dd[:, :1], dd[:, 1:]):
            n = part.numel()
            s = part.sum().item()
            s2 = (part * part).sum().item()
            stats.append((n, s, s2))
        (nr, sr, s2r), (na, sa, s2a) = stats

        def std(n, s, s2):
            if n == 0:
                return 1e-1
            v = np.sqrt(max(s2 / n - (s / n) * (s / n), 0.0))
            return 1e-2 if abs(v) < 1e-2 else v

        avg_r = sr / nr if nr else 0.0
        davg[t, :, 0] = avg_r
        dstd[t, :, 0] = std(nr, sr, s2r)
        dstd[t, :, 1:] = std(na, sa, s2a)
    return davg, dstd


def rounded(a: np.ndarray, digits: int = 6) -> np.ndarray:
    """Round to `digits` significant digits, so statistics computed on the
    GPU or on the CPU give bit-identical model parameters."""
    a = np.asarray(a, dtype=np.float64)
    return np.vectorize(lambda v: float(f"{v:.{digits}g}"))(a)


def model_params_with_stats(spec, em_raw, atype, seed: int = 7) -> dict:
    """dpmi.systems.make_params(spec, seed) with davg/dstd from the
    environment statistics of one frame (em_raw computed with davg = 0,
    dstd = 1), rounded to six significant digits."""
    from . import systems

    params = systems.make_params(spec, seed=seed)
    davg, dstd = env_mat_stat(em_raw, atype, spec.ntypes, int(sum(spec.sel)))
    params["davg"], params["dstd"] = rounded(davg), rounded(dstd)
    return params


def gpu_model_params(spec, system, seed: int = 7, device="cuda") -> dict:
    """Statistics from `system` evaluated with the GPU kernels."""
    from . import ops

    nnei = int(sum(spec.sel))
    dev = torch.device(device)
    c = torch.as_tensor(system.coord, device=dev)
    t = torch.as_tensor(system.atype, device=dev)
    ext_c, ext_t, mapping, nlist, _, _ = ops.build_nlist(c, t, system.box, spec.rcut, spec.sel)
    z = torch.zeros((spec.ntypes, nnei, 4), dtype=torch.float64, device=dev)
    em, _, _ = ops.env_mat_a(nlist, ext_c, ext_t, z, z + 1, spec.rcut, spec.rcut_smth, spec.sel,
                             with_deriv=False, with_rij=False)
    return model_params_with_stats(spec, em.cpu(), system.atype, seed=seed)
