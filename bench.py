#!/usr/bin/env python
"""E+F+V atom-steps/s of the se_e2_a path on one MI355X.

A step = neighbor-list rebuild + energy + forces + virial of one frame, all on
the GPU, inputs resident in HBM.  Both forms run on the same GPU and the same
coordinates: the plain FP64 path and the BF16-MFMA form (embedding/fitting
GEMM operands bf16 with fp32 accumulation, env-mat fp32, energy and virial
reductions fp64).  `value` is the BF16 form's throughput; the FP64 figure,
the speedup and the BF16 error against FP64 ride along, and so does the
FP16-MFMA form (same kernels with fp16 operands: throughput, speedup, error).

    python bench.py [--steps K] [--warmup W] [--workload water110k]
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
for _p in (str(ROOT), str(ROOT / "dpmd-public_amd")):
    if _p not in sys.path:
        sys.path.insert(0, _p)

METRIC = "E+F+V atom-steps/s/GPU: FP64 plain vs BF16-MFMA form at force RMSE≤1e-3 eV/Å"

WORKLOADS = {
    "water110k": dict(desc="110,592-atom water-like box (36,864 type-0 / 73,728 type-1), L=99.36 A, jittered "
                           "48^3 lattice, seed 3; se_e2_a rcut 6.0, rcut_smth 0.5, sel [46,92], embedding "
                           "[25,50,100] tanh resnet_dt, axis 16, fitting [240,240,240]",
                      spec="water", stat="water192"),
    "water12k": dict(desc="12,288-atom water-like box, L=49.68 A, seed 1; same model", spec="water",
                     stat="water192"),
    "water192": dict(desc="192-atom water-like box, L=12.42 A, seed 0; same model", spec="water", stat="water192"),
    "alloy262k": dict(desc="262,144-atom four-type FCC alloy (64^3 primitive cells, a=3.6 A), seed 4; rcut 8.0, "
                           "rcut_smth 2.0, sel [60]*4, embedding [32,64,128], axis 16, fitting [240]*3",
                      spec="alloy", stat="alloy_small"),
}

# algorithmic HBM bytes per local atom of each libdpmi kernel (recompute design):
# nlist row (4*nnei) + em or dE/dem row (4*nnei*elem) + centre and neighbor coordinates
# counted once per atom (2 * 3 * 8).
def _bytes_per_atom(name: str, nnei: int) -> float:
    elem = 8 if name.endswith("f64") else 4
    if name.startswith("dpmi_force_virial") or name.startswith("dpmi_prod_env_mat_a"):
        return 4 * nnei + 4 * nnei * elem + 48
    if name.startswith("dpmi_build_nlist"):
        return 4 * nnei + 4 * nnei + 24 + 4  # nlist + mapped/overflow, own coordinate
    return 0.0


HBM_PEAK_GBS = 8000.0
# entry point -> kernel it launches (for the measured-traffic table)
KERNEL_OF = {"dpmi_force_virial_cells_f32": "k_force_cells<float, 1536, false, 3>"}
for _h in ("bf16", "f16"):
    KERNEL_OF.update({f"dpmi_se_a_embed_bwd_{_h}": f"k_embed_bwd<{_h}, skip>",
                      f"dpmi_se_a_embed_fwd_{_h}": f"k_embed_fwd<{_h}, skip>",
                      f"dpmi_se_a_fitting_{_h}": f"k_fit<{_h}, 0>",
                      f"dpmi_se_a_fitting_gr_{_h}": f"k_fit<{_h}, 1>"})


def _measured_traffic(workload: str, entry: str):
    """HBM bytes per launch (FETCH_SIZE + WRITE_SIZE, copy-calibrated) of the
    kernel behind `entry`, from the committed rocprofv3 counter summary of
    this workload, or None."""
    f = ROOT / "profiles" / f"hbm_traffic_{workload}_bf16.txt"
    k = KERNEL_OF.get(entry)
    if not f.exists() or k is None:
        return None, None
    for line in f.read_text().splitlines():
        # table rows: kernel name, read MB, write MB, algorithmic read/write MB, ratio
        if line.startswith(k + " "):
            cols = line[len(k):].split()
            return int((float(cols[0]) + float(cols[1])) * 1e6), str(f.relative_to(ROOT))
    return None, None
MFMA_PEAK_TFLOPS = {"bf16": 2500.0, "fp16": 2500.0, "fp64": 78.6}


def _embed_flops_per_row(neuron) -> float:
    """Algorithmic FLOPs of one neighbor row: layer GEMMs 2*(n1*n2 + n2*n3)
    (the 1 -> n1 layer and the activations are not GEMM work)."""
    n1, n2, n3 = neuron
    return 2.0 * (n1 * n2 + n2 * n3)


def _mfma_flops(name: str, spec, nloc: int, nrows: int) -> float:
    """Per-launch algorithmic FLOPs of the fused MFMA kernels.
    Embedding forward: layer GEMMs + em^T G (2*4*n3 per neighbor row);
    embedding backward: the recomputed forward, its tangent in s (the same
    two GEMMs again), and the contractions G dGR^T and (dG/ds) dGR^T
    (2 * 2*4*n3 per row).
    ``nrows`` counts the rows the algorithm needs: the listed neighbors when
    the padding rows (all equal within a type section) are folded into one
    closed-form term, else every slot.
    Fitting: forward 2*(k1*H + 2*H*H + H) and adjoint 2*(2*H*H + H*k1) per
    atom (k1 = n3 * axis, H = hidden width)."""
    n3 = spec.neuron[2]
    f = _embed_flops_per_row(spec.neuron)
    if name.startswith("dpmi_se_a_embed_fwd"):
        return nrows * (f + 8.0 * n3)
    if name.startswith("dpmi_se_a_embed_bwd"):
        return nrows * (2 * f + 16.0 * n3)
    if name.startswith("dpmi_se_a_fitting"):
        k1, h = n3 * spec.axis_neuron, spec.fitting[0]
        return nloc * (2.0 * (k1 * h + 2 * h * h + h) + 2.0 * (2 * h * h + h * k1))
    return 0.0


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="water110k", choices=sorted(WORKLOADS))
    ap.add_argument("--precision", default="bf16", choices=["bf16", "fp16"])
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-fp64", action="store_true",
                    help="skip the FP64 leg (profiling runs of the low-precision step only)")
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--copy-bytes", type=int, default=0,
                    help="calibration for HBM counters: only run 5 device-to-device copies of this many bytes")
    ap.add_argument("--form", default="precision", choices=["precision", "tabulated", "train"],
                    help="precision: FP64 plain vs the bf16/fp16 form (default); tabulated: the compressed "
                         "FP64 model (use with --workload water12k); train: the 8 x 192 E+F training step")
    args = ap.parse_args()
    if args.copy_bytes:
        return copy_calibration(args.copy_bytes)
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print(f"bench.py measures one MI355X; --gpus {max(args.gpus, world)} is not supported (the multi-GPU "
              "form, spatial shards with a halo exchange, is not built)",
              file=sys.stderr)
        return 2

    if args.form == "tabulated":
        return bench_tabulated(args)
    if args.form == "train":
        return bench_train(args)

    import numpy as np
    import torch

    from dpmi import ops, systems
    from dpmi.model import CapturedStep, SeAEnergyModel
    from dpmi.stat import gpu_model_params

    dev = torch.device("cuda:0")
    wl = WORKLOADS[args.workload]
    spec = systems.WATER_SPEC if wl["spec"] == "water" else systems.ALLOY_SPEC
    stat_sys = systems.water192(0) if wl["stat"] == "water192" else systems.alloy262k(4, ncell=8)
    params = gpu_model_params(spec, stat_sys)
    sysm = systems.SYSTEMS[args.workload]()
    coord = torch.tensor(sysm.coord, device=dev)
    atype = torch.tensor(sysm.atype, device=dev)
    nloc = sysm.nloc
    nnei = int(sum(spec.sel))

    peak_mem = {}
    eager_ms = {}

    def run(model, names):
        torch.cuda.synchronize()
        torch.cuda.reset_peak_memory_stats()
        for _ in range(args.warmup):
            model.compute(coord, atype, sysm.box)
        torch.cuda.synchronize()
        peak_mem[model.precision] = torch.cuda.max_memory_allocated()
        with ops.KernelTimer(*names) as kt:
            t0 = time.perf_counter()
            for _ in range(args.steps):
                out = model.compute(coord, atype, sysm.box)
            torch.cuda.synchronize()
            t1 = time.perf_counter()
        eager = (t1 - t0) / args.steps
        # the same step captured in a HIP graph (list rebuilt in every replay)
        cap = CapturedStep(model, coord, atype, sysm.box)
        for _ in range(args.warmup):
            cap.replay()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            cap.replay()
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        graph = (t1 - t0) / args.steps
        if not torch.isfinite(cap.out["force"]).all():
            raise RuntimeError("captured step produced non-finite forces")
        eager_ms[model.precision] = round(eager * 1e3, 3)
        return graph, cap.out, kt.summary()

    names64 = ["dpmi_copy_coord_nowait_f64", "dpmi_build_nlist_nowait_f64", "dpmi_prod_env_mat_a_f64",
               "dpmi_force_virial_cells_f64", "dpmi_se_a_embed_fwd_f64", "dpmi_se_a_embed_bwd_f64",
               "dpmi_se_a_fitting_gr_f64", "dpmi_energy_sum_f64"]
    h = {"bf16": "bf16", "fp16": "f16"}[args.precision]
    names32 = ["dpmi_copy_coord_nowait_f64", "dpmi_build_nlist_nowait_f64", "dpmi_prod_env_mat_a_f32c64",
               "dpmi_force_virial_cells_f32", f"dpmi_se_a_embed_fwd_{h}", "dpmi_se_a_descriptor_fwd",
               "dpmi_se_a_descriptor_bwd", f"dpmi_se_a_embed_bwd_{h}", f"dpmi_se_a_fitting_{h}",
               f"dpmi_se_a_fitting_gr_{h}", "dpmi_energy_sum_f64"]
    t64, out64, k64 = None, None, {}
    if not args.no_fp64:
        m64 = SeAEnergyModel.from_params(spec, params, device=dev, precision="fp64")
        t64, out64, k64 = run(m64, names64)
        del m64
    mlo = SeAEnergyModel.from_params(spec, params, device=dev, precision=args.precision)
    tlo, outlo, klo = run(mlo, names32)
    # the other 16-bit form beside the FP64 leg (bf16 line: fp16 too)
    other = None
    if out64 is not None and args.precision == "bf16":
        m16 = SeAEnergyModel.from_params(spec, params, device=dev, precision="fp16")
        other = run(m16, [n.replace("_bf16", "_f16") for n in names32])
        del m16

    def accuracy_of(out):
        frms = torch.sqrt((out64["force"] ** 2).mean()).item()
        rmse = torch.sqrt(((out64["force"] - out["force"]) ** 2).mean()).item()
        de_atom = abs(out64["energy"].item() - out["energy"].item()) / nloc
        vscale = torch.abs(out64["virial"]).max().item()
        dv = torch.abs(out64["virial"] - out["virial"]).max().item() / max(vscale, 1e-300)
        return {"force_rmse_eV_per_A": rmse, "force_rms_eV_per_A": frms,
                "force_rmse_relative": rmse / frms if frms else None,
                "energy_per_atom_abs_eV": de_atom, "virial_max_rel": dv, "force_rmse_ok": rmse <= 1e-3}

    accuracy = accuracy_of(outlo) if out64 is not None else None

    # rows and 16-row tiles the embedding kernels evaluate: listed neighbors
    # per (atom, type section) when the padding shortcut is on, else all slots
    nlist = mlo.neighbor_list(coord, atype.to(torch.int32), sysm.box)[3]
    if mlo.pad_em is not None:
        offs = [0]
        for s_t in spec.sel:
            offs.append(offs[-1] + s_t)
        listed = [(nlist[:, offs[t]:offs[t + 1]] >= 0).sum(1) for t in range(len(spec.sel))]
        nrows = int(sum(c.sum().item() for c in listed))
        tiles = int(sum(((c + 15) // 16).sum().item() for c in listed))
    else:
        nrows = nloc * nnei
        tiles = nloc * sum((s_t + 15) // 16 for s_t in spec.sel)
    del nlist

    # dominant libdpmi kernel of the low-precision step and its roofline
    dom = max(klo, key=lambda n: klo[n]["total_ms"])
    avg_s = klo[dom]["avg_ms"] / 1e3
    flops = _mfma_flops(dom, spec, nloc, nrows)
    traffic, traffic_src = _measured_traffic(args.workload, dom)
    if flops > 0:
        achieved = flops / avg_s / 1e12 if avg_s > 0 else 0.0
        peak = MFMA_PEAK_TFLOPS[args.precision]
        roofline = {"bound": "mfma", "kernel": dom, "achieved": round(achieved, 2), "peak": peak,
                    "unit": "TFLOP/s", "frac": round(achieved / peak, 4), "traffic": traffic,
                    "traffic_source": traffic_src,
                    "algorithmic_flops_per_launch": flops, "embedding_rows_per_launch": nrows,
                    "embedding_tiles_per_launch": tiles,
                    "avg_launch_ms": round(klo[dom]["avg_ms"], 4)}
    else:
        algo = _bytes_per_atom(dom, nnei) * nloc
        achieved = algo / avg_s / 1e9 if avg_s > 0 else 0.0
        roofline = {"bound": "hbm", "kernel": dom, "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                    "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                    "traffic_source": traffic_src,
                    "algorithmic_bytes_per_launch": int(algo), "avg_launch_ms": round(klo[dom]["avg_ms"], 4)}
    # whole-step roofline: the step's algorithmic FLOPs at the MFMA peak plus
    # its HBM-bound kernels' algorithmic bytes at the HBM peak, over the
    # measured step.  FLOPs per atom: E+F = 2 x forward, forward = embedding
    # 2 nnei (n1 + n1 n2 + n2 n3) + G^T R 2 nnei 4 M + D 2 4 M axis + own-type
    # fitting 2 (k1 H + 2 H^2 + H); bytes per atom: list 4 nnei, env-mat and
    # recompute force 4 nnei (1 + 4) + 48 each (fp32 form)
    n1, n2, n3 = spec.neuron
    hid = spec.fitting[0]
    fwd_flops = (2 * nnei * (n1 + n1 * n2 + n2 * n3) + 2 * nnei * 4 * n3 + 2 * 4 * n3 * spec.axis_neuron
                 + 2 * (n3 * spec.axis_neuron * hid + 2 * hid * hid + hid))
    step_flops = 2.0 * fwd_flops * nloc
    step_bytes = (4 * nnei + 2 * (4 * nnei * 5 + 48)) * nloc
    bound_s = step_flops / (MFMA_PEAK_TFLOPS[args.precision] * 1e12) + step_bytes / (HBM_PEAK_GBS * 1e9)
    step_roofline = {"flops_per_step": step_flops, "hbm_bytes_per_step": step_bytes,
                     "bound_ms": round(bound_s * 1e3, 4), "frac": round(bound_s / tlo, 4)}
    # the HBM-bound kernels of the same step, for reference
    line_hbm = {}
    for n, v in klo.items():
        b = _bytes_per_atom(n, nnei) * nloc
        if b > 0 and v["avg_ms"] > 0:
            line_hbm[n] = {"GB_per_s": round(b / (v["avg_ms"] / 1e3) / 1e9, 1), "avg_ms": round(v["avg_ms"], 4)}

    line = {
        "metric": METRIC,
        "value": round(nloc / tlo, 1),
        "unit": "atom-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(tlo * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": args.precision,
        "data": "synthetic (seeded SplitMix64 lattice, random-init weights, davg/dstd from the 192-atom frame)",
        "config": {"workload": args.workload, "description": wl["desc"], "atoms": nloc,
                   "form": f"{args.precision} GEMM operands / fp32 accumulate, fp32 env-mat, fp64 E and virial"},
        "fp64": None if t64 is None else {"value": round(nloc / t64, 1), "ms_per_step": round(t64 * 1e3, 3)},
        "speedup_vs_fp64": None if t64 is None else round(t64 / tlo, 3),
        "accuracy_vs_fp64": accuracy,
        "fp16": None if other is None else {
            "value": round(nloc / other[0], 1), "ms_per_step": round(other[0] * 1e3, 3),
            "speedup_vs_fp64": round(t64 / other[0], 3), "accuracy_vs_fp64": accuracy_of(other[1])},
        "step_timing": "ms_per_step: the step (list rebuilt, E, F, V) captured once as a HIP graph and replayed; "
                       "eager_ms_per_step: the same step launched from Python each time",
        "eager_ms_per_step": eager_ms,
        "kernels_ms_per_step": {"fp64": {k: round(v["total_ms"] / args.steps, 4) for k, v in k64.items()},
                                args.precision: {k: round(v["total_ms"] / args.steps, 4) for k, v in klo.items()}}
        | ({} if other is None else {"fp16": {k: round(v["total_ms"] / args.steps, 4) for k, v in other[2].items()}}),
        "roofline": roofline,
        "step_roofline": step_roofline,
        "hbm_kernels": line_hbm,
        # the env-mat derivative is recomputed inside the force kernel, never
        # stored; what storing it would take (nloc x nnei x 4 x 3 values)
        "memory_GB": {"peak_" + k: round(v / 1e9, 3) for k, v in peak_mem.items()} | {
            "env_deriv_stored": 0.0,
            "env_deriv_if_stored_fp64": round(nloc * nnei * 12 * 8 / 1e9, 3),
            "env_deriv_if_stored_fp32": round(nloc * nnei * 12 * 4 / 1e9, 3)},
    }

    if not args.no_cpu_baseline:
        line["cpu_baseline"] = cpu_baseline(spec, params, args.cpu_seconds, args.workload, sysm)
    print(json.dumps(line), flush=True)
    return 0


def _line(args, value, ms, dtype, config, **extra) -> dict:
    return {"metric": METRIC, "value": round(value, 1), "unit": "atom-steps/s", "n_gpus": 1, "steps": args.steps,
            "warmup": args.warmup, "ms_per_step": round(ms, 3), "higher_is_better": True, "scaling": "weak",
            "vs_baseline": None, "dtype": dtype, "data": "synthetic (seeded SplitMix64 lattice, random-init weights)",
            "config": config, **extra}


def bench_tabulated(args) -> int:
    """The compressed FP64 model: E+F+V steps with the tabulated embedding
    (tables from the model, grid [0, 2) / 0.01 + [2, 10) / 0.1), beside the
    same model with the network embedding."""
    import torch

    from dpmi import ops, systems, table
    from dpmi.model import CapturedStep, SeAEnergyModel

    dev = torch.device("cuda:0")
    spec = systems.WATER_SPEC
    params = systems.make_params(spec, seed=7)  # davg 0, dstd 1
    tabs = table.build_tables(params["embedding"], params["davg"], params["dstd"], spec.rcut, spec.rcut_smth,
                              0.9, spec.sel)
    sysm = systems.SYSTEMS[args.workload]()
    coord = torch.tensor(sysm.coord, device=dev)
    atype = torch.tensor(sysm.atype, device=dev)
    names = ["dpmi_build_nlist_nowait_f64", "dpmi_prod_env_mat_a_f64", "dpmi_tabulate_fusion_se_a_f64",
             "dpmi_tabulate_fusion_se_a_grad_f64", "dpmi_se_a_fitting_gr_f64", "dpmi_force_virial_cells_f64"]

    eager_ms = {}

    def run(model, label):
        for _ in range(args.warmup):
            model.compute(coord, atype, sysm.box)
        torch.cuda.synchronize()
        with ops.KernelTimer(*names) as kt:
            t0 = time.perf_counter()
            for _ in range(args.steps):
                out = model.compute(coord, atype, sysm.box)
            torch.cuda.synchronize()
            eager_ms[label] = round((time.perf_counter() - t0) / args.steps * 1e3, 3)
        cap = CapturedStep(model, coord, atype, sysm.box)
        for _ in range(args.warmup):
            cap.replay()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            cap.replay()
        torch.cuda.synchronize()
        dt = (time.perf_counter() - t0) / args.steps
        if not torch.isfinite(cap.out["force"]).all():
            raise RuntimeError("captured step produced non-finite forces")
        return dt, cap.out, kt.summary()

    tt, ot, kt = run(SeAEnergyModel.from_params(spec, params, device=dev, tables=tabs), "tabulated")
    tn, on, _ = run(SeAEnergyModel.from_params(spec, params, device=dev), "network_fp64")
    frms = torch.sqrt((on["force"] ** 2).mean()).item()
    line = _line(args, sysm.nloc / tt, tt * 1e3, "f64",
                 {"workload": args.workload, "atoms": sysm.nloc, "form": "FP64, tabulated embedding "
                  "(info [0, 2, 10, 0.01, 0.1, -1], 280 intervals)"},
                 step_timing="ms_per_step: the step captured once as a HIP graph and replayed; "
                             "eager_ms_per_step: launched from Python each time",
                 eager_ms_per_step=eager_ms,
                 network_fp64={"value": round(sysm.nloc / tn, 1), "ms_per_step": round(tn * 1e3, 3)},
                 table_vs_network={"force_rmse_eV_per_A": torch.sqrt(((ot["force"] - on["force"]) ** 2).mean()).item(),
                                   "force_rms_eV_per_A": frms,
                                   "energy_per_atom_abs_eV": abs(ot["energy"].item() - on["energy"].item()) / sysm.nloc,
                                   "note": "tables built as the reference's DPTabulate builds them: the timestep "
                                           "(idt) of the resnet_dt layers is not folded in, so this is the "
                                           "reference's own table-vs-network difference (build_tables(fold_idt=True) "
                                           "reproduces the network)"},
                 kernels_ms_per_step={k: round(v["total_ms"] / args.steps, 4) for k, v in kt.items()})
    print(json.dumps(line), flush=True)
    return 0


def bench_train(args) -> int:
    """Training steps on 8 frames x 192 water atoms (E+F loss, pref_e 0.02,
    pref_f 1000, labels from the same architecture with weight seed 8): the
    neighbor lists and env-mat of the batch, forward with the force graph,
    loss, double backward and an Adam step, in each of the four forms."""
    import numpy as np
    import torch

    from dpmi import systems, table, train
    from dpmi.model import SeAEnergyModel

    dev = torch.device("cuda:0")
    spec = systems.WATER_SPEC
    params = systems.make_params(spec, seed=7)
    lab = SeAEnergyModel.from_params(spec, systems.make_params(spec, seed=8), device=dev)
    frames = [systems.water192(k) for k in range(8)]
    coords = [torch.tensor(f.coord, device=dev) for f in frames]
    atypes = [torch.tensor(f.atype, device=dev) for f in frames]
    boxes = [f.box for f in frames]
    outs = [lab.compute(c, t, b) for c, t, b in zip(coords, atypes, boxes)]
    e_l = torch.stack([o["energy"] for o in outs])
    f_l = torch.stack([o["force"] for o in outs])
    tabs = table.build_tables(params["embedding"], params["davg"], params["dstd"], spec.rcut, spec.rcut_smth,
                              0.9, spec.sel)
    res = {}
    for name, prec, tb in (("fp64_network", "fp64", None), ("fp64_table", "fp64", tabs),
                           ("fp32_network", "fp32", None),
                           ("fp32_table", "fp32", [(t.astype(np.float32), i) for t, i in tabs])):
        m = SeAEnergyModel.from_params(spec, params, device=dev, precision=prec, tables=tb)
        tr = train.Trainer(m)

        def step():
            batch = train.prepare(m, coords, atypes, boxes)
            return tr.step(batch, e_l, f_l)[0]

        for _ in range(args.warmup):
            step()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            loss = step()
        torch.cuda.synchronize()
        eager = (time.perf_counter() - t0) / args.steps
        # the same step captured once as a HIP graph (a fresh model and optimizer)
        m = SeAEnergyModel.from_params(spec, params, device=dev, precision=prec, tables=tb)
        cap = train.CapturedTrainStep(train.Trainer(m), coords, atypes, boxes, e_l, f_l, warmup=args.warmup)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            gloss = cap.replay()
        torch.cuda.synchronize()
        dt = (time.perf_counter() - t0) / args.steps
        cap.check()
        res[name] = {"atom_steps_per_s": round(8 * 192 / dt, 1), "ms_per_step": round(dt * 1e3, 3),
                     "eager_ms_per_step": round(eager * 1e3, 3), "last_loss": float(gloss),
                     "eager_last_loss": float(loss)}
    main_ = res["fp64_network"]
    line = _line(args, main_["atom_steps_per_s"], main_["ms_per_step"], "f64",
                 {"workload": "train8x192", "atoms": 8 * 192, "form": "training step (E+F loss, double backward, "
                  "Adam) FP64 network embedding; the other forms below"},
                 step_timing="ms_per_step: the whole step (lists, env-mat, forward, double backward, lr tick, Adam) "
                             "captured once as a HIP graph and replayed; eager_ms_per_step: launched from Python",
                 forms=res)
    print(json.dumps(line), flush=True)
    return 0


def copy_calibration(nbytes: int) -> int:
    """5 plain device-to-device copies of nbytes (reads nbytes, writes
    nbytes each): the reference for scaling FETCH_SIZE / WRITE_SIZE counters."""
    import torch

    src = torch.ones(nbytes // 4, dtype=torch.float32, device="cuda:0")
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(5):
        dst.copy_(src)
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / 5
    print(json.dumps({"copy_bytes": nbytes, "ms_per_copy": round(dt * 1e3, 4),
                      "GB_per_s_read_plus_write": round(2 * nbytes / dt / 1e9, 1)}), flush=True)
    return 0


def _cpu_model_name() -> str:
    try:
        for line in Path("/proc/cpuinfo").read_text().splitlines():
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def _whole_step_sample(m, s, budget_s: float):
    """Whole E+F+V steps of system s with the CPU oracle (C neighbor list +
    float64 torch-CPU model, autograd forces) until budget_s has elapsed (at
    least one, at most 8): (atom-steps/s, steps, seconds)."""
    m.evaluate(s.coord[:600], s.atype[:600], None)  # warm up allocators
    steps, t0 = 0, time.perf_counter()
    while True:
        m.evaluate(s.coord, s.atype, s.box)
        steps += 1
        el = time.perf_counter() - t0
        if el >= budget_s or steps >= 8:
            break
    return s.nloc * steps / el, steps, el


def cpu_baseline(spec, params, budget_s: float, workload: str, sysm) -> dict:
    """The CPU oracle on the benchmark's own system: the C restatement of the
    reference CPU kernels (cell-list neighbor list, env-mat with its
    derivative, prod_force_a, prod_virial_a: OpenMP over atoms) on all atoms,
    and the float64 torch-CPU network with autograd dE/d em (the reference PT
    composition) on blocks of atoms until budget_s has elapsed, its time per
    atom extrapolated to the whole system.  Beside it: the reference's own
    O(nloc * nall) candidate search (build_nlist_cpu, neighbor_list.cc:873-916,
    one thread) on a sample of centres, and whole steps of the smaller box of
    the same model."""
    import numpy as np
    import torch

    from dpmi import systems
    from oracle import oracle
    from oracle.model_ref import Model

    threads = torch.get_num_threads()
    os.environ.setdefault("OMP_NUM_THREADS", str(threads))
    oracle.lib()
    m = Model.from_params(spec, params)
    s = sysm
    sel = list(spec.sel)
    t0 = time.perf_counter()
    ec, et, mp, nlist, _ = oracle.build_nlist_pbc(s.coord, s.atype, s.box, spec.rcut, sel)
    t_nlist = time.perf_counter() - t0
    avg, std = m.davg.numpy(), m.dstd.numpy()
    t0 = time.perf_counter()
    em, ed, rij = oracle.env_mat_a(nlist, ec, et, avg, std, spec.rcut, spec.rcut_smth, sel)
    t_env = time.perf_counter() - t0
    nnei = nlist.shape[1]
    nd = np.zeros((s.nloc, nnei * 4))
    blk, done, t0 = 2048, 0, time.perf_counter()
    while done < s.nloc:
        b = min(s.nloc, done + blk)
        x = torch.tensor(em[done:b].reshape(b - done, nnei, 4), requires_grad=True)
        ae = m.atom_energy(m.descriptor(x), s.atype[done:b])
        (g,) = torch.autograd.grad(ae.sum(), x)
        nd[done:b] = g.numpy().reshape(b - done, -1)
        done = b
        if time.perf_counter() - t0 >= budget_s:
            break
    t_net_done = time.perf_counter() - t0
    t_net = t_net_done * s.nloc / done
    t0 = time.perf_counter()
    oracle.prod_force_a(nd, ed.reshape(s.nloc, -1), nlist, ec.shape[0])
    t_force = time.perf_counter() - t0
    t0 = time.perf_counter()
    oracle.prod_virial_a(nd, ed.reshape(s.nloc, -1), rij.reshape(s.nloc, -1), nlist, ec.shape[0])
    t_virial = time.perf_counter() - t0
    step = t_nlist + t_env + t_net + t_force + t_virial
    # the reference's all-pairs candidate search (its algorithm, run here on
    # the oracle's threads) for the first ncent centres against every
    # extended atom; its cost is linear in the number of centres
    ncent = min(s.nloc, 8192)
    t0 = time.perf_counter()
    oracle.candidates(ec, ncent, spec.rcut, brute=True)
    t_brute = (time.perf_counter() - t0) * s.nloc / max(ncent, 1)
    small = systems.water12k(1) if spec.ntypes == 2 else systems.alloy262k(4, ncell=24)
    ws_rate, ws_steps, ws_el = _whole_step_sample(m, small, min(budget_s, 8.0))
    return {"value": round(s.nloc / step, 1), "unit": "atom-steps/s", "cores": threads, "kind": "port",
            "nproc": os.cpu_count(), "cpu_model": _cpu_model_name(),
            "sample": f"one FP64 E+F+V step of the {workload} system ({s.nloc} atoms) on the CPU oracle: C "
                      f"cell-list neighbor list {t_nlist:.2f} s, C env-mat with derivative {t_env:.2f} s, "
                      f"C prod_force {t_force:.2f} s and prod_virial {t_virial:.2f} s on all atoms; the "
                      f"float64 torch-CPU network with autograd dE/d em on {done} of the atoms in "
                      f"{t_net_done:.1f} s, extrapolated to {t_net:.1f} s for all",
            "parts_s": {"nlist": round(t_nlist, 3), "env_mat": round(t_env, 3), "network": round(t_net, 3),
                        "prod_force": round(t_force, 3), "prod_virial": round(t_virial, 3)},
            "reference_all_pairs_candidates": {
                "seconds_for_all_atoms": round(t_brute, 2), "threads": threads,
                "sample": f"build_nlist_cpu's O(nloc * nall) search for {ncent} centres against all "
                          f"{ec.shape[0]} extended atoms, scaled to {s.nloc} centres"},
            "whole_step_smaller_box": {"atoms": small.nloc, "atom_steps_per_s": round(ws_rate, 1), "steps": ws_steps,
                                       "seconds": round(ws_el, 1)}}


if __name__ == "__main__":
    sys.exit(main())
