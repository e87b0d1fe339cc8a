"""Write the reference's own test vectors for the se_e2_a path as JSON fixtures.

Reads the reference checkout (default /root/reference) as text and extracts
only data: the input arrays and expected outputs its tests assert against.

    source/lib/tests/test_fmt_nlist.cc      -> fmt_nlist.json
    source/lib/tests/test_env_mat_a.cc      -> env_mat_a.json
    source/lib/tests/test_tabulate_se_a.cc  -> tabulate_se_a.json
    source/lib/tests/test_prod_force_a.cc   -> prod_force_a.json
    source/lib/tests/test_prod_virial_a.cc  -> prod_virial_a.json
    source/lib/tests/test_prod_force_grad_a.cc  -> prod_force_grad_a.json
    source/lib/tests/test_prod_virial_grad_a.cc -> prod_virial_grad_a.json
    source/lib/tests/test_coord.cc          -> coord.json
    source/tests/pt/test_tabulate_fusion_se_a.py -> tabulate_fusion_se_a_pt.json
    source/tests/infer/deeppot_sea.yaml     -> deeppot_sea_model.json
    source/tests/infer/deeppot-testcase.yaml -> deeppot_sea_cases.json

Run once in a container that has the reference; the JSON files are committed
and the tests read only them.
"""

from __future__ import annotations

import json
import re
import sys
from pathlib import Path

import yaml

OUT = Path(__file__).resolve().parent


def _numbers(body: str) -> list:
    body = re.sub(r"//[^\n]*", "", body)
    toks = [t for t in re.split(r"[\s,]+", body) if t]
    vals = []
    for t in toks:
        t = t.rstrip("fUL")
        if re.fullmatch(r"[-+]?\d+", t):
            vals.append(int(t))
        else:
            vals.append(float(t))
    return vals


def cpp_classes(path: Path) -> dict:
    """{class_name: {array_name: [values]}} for std::vector<...> x = {...};"""
    text = path.read_text()
    out = {}
    parts = re.split(r"\nclass (\w+) : public ::testing::Test", text)
    for i in range(1, len(parts), 2):
        name, body = parts[i], parts[i + 1]
        body = body.split("\nTEST_F")[0]
        arrs = {}
        for m in re.finditer(r"std::vector<(double|int|uint_64)> (\w+) = \{([^;{}]*?)\};", body):
            if "<<" in m.group(3):  # key-encoding cases of the reference GPU sort: not data we use
                continue
            arrs[m.group(2)] = _numbers(m.group(3))
        for m in re.finditer(r"\n\s*(?:const )?(double|int) (\w+) = ([-+0-9.e]+);", body):
            arrs[m.group(2)] = _numbers(m.group(3))[0]
        out[name] = arrs
    return out


def py_tensors(path: Path) -> dict:
    text = path.read_text()
    out = {}
    for m in re.finditer(r"self\.(\w+) = torch\.tensor\(\s*\[(.*?)\]\s*,", text, re.S):
        out[m.group(1)] = _numbers(m.group(2))
    for m in re.finditer(r"self\.(\w+) = (\d+)\n", text):
        out[m.group(1)] = int(m.group(2))
    return out


def _yaml_arrays(x):
    """Strip the serialization wrappers: ndarray dicts become nested lists."""
    if isinstance(x, dict):
        if x.get("@class") == "np.ndarray":
            return x["value"]
        return {k: _yaml_arrays(v) for k, v in x.items()}
    if isinstance(x, list):
        return [_yaml_arrays(v) for v in x]
    return x


def main(ref: Path) -> None:
    t = ref / "source/lib/tests"
    jobs = {
        "fmt_nlist.json": cpp_classes(t / "test_fmt_nlist.cc"),
        "env_mat_a.json": cpp_classes(t / "test_env_mat_a.cc"),
        "tabulate_se_a.json": cpp_classes(t / "test_tabulate_se_a.cc"),
        "prod_force_a.json": cpp_classes(t / "test_prod_force_a.cc"),
        "prod_virial_a.json": cpp_classes(t / "test_prod_virial_a.cc"),
        "prod_force_grad_a.json": cpp_classes(t / "test_prod_force_grad_a.cc"),
        "prod_virial_grad_a.json": cpp_classes(t / "test_prod_virial_grad_a.cc"),
        "coord.json": cpp_classes(t / "test_coord.cc"),
        "tabulate_fusion_se_a_pt.json": py_tensors(ref / "source/tests/pt/test_tabulate_fusion_se_a.py"),
    }
    model = yaml.safe_load((ref / "source/tests/infer/deeppot_sea.yaml").read_text())
    jobs["deeppot_sea_model.json"] = _yaml_arrays(model["model"])
    cases = yaml.safe_load((ref / "source/tests/infer/deeppot-testcase.yaml").read_text())
    jobs["deeppot_sea_cases.json"] = {
        "rcut": cases["rcut"],
        "results": [
            {k: v for k, v in r.items() if k in ("coord", "atype", "box", "atomic_energy", "force", "atomic_virial", "descriptor")}
            for r in cases["results"]
        ],
    }
    for name, data in jobs.items():
        (OUT / name).write_text(json.dumps(data, separators=(",", ":")))
        print("wrote", name)


if __name__ == "__main__":
    main(Path(sys.argv[1]) if len(sys.argv) > 1 else Path("/root/reference"))
