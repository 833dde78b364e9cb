"""Oracle reference of the 15-mode heat-bath workload (c1).

Runs the CPU oracle on the synthetic 15-mode field (about 20 minutes on
5-8 cores for the variational stage; with PT2 on, the gather over the 5.4 M
external states adds about 40 minutes) and stores, in c1_oracle.json, the
basis size after every heat-bath iteration, the 100 variational energies
(%.17g), the checksum of the final basis (the same checksum vhci_run_ctx
reports), and, from the run with PT2 on, the external-space size and the
100 PT2-corrected energies.  With existing oracle run directories, pass the
paths of the JSON summary and basis file of the PT2-off run and then the
JSON summary of the PT2-on run to skip the runs; `--pt2-only JSON` adds the
PT2 results of an existing PT2-on run to the stored file.
"""
import json
import os
import subprocess
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
OUT = os.path.join(HERE, "c1_oracle.json")


def checksum(q):
    v = np.ascontiguousarray(q, dtype=np.int32).ravel().astype(np.int64).astype(np.uint64)
    with np.errstate(over="ignore"):
        pw = np.cumprod(np.full(v.size, 1099511628211, dtype=np.uint64), dtype=np.uint64)
        return int(np.sum(v * pw, dtype=np.uint64))


def oracle_run(pt2):
    from vhci_amd import synth
    work = os.path.join(HERE, "_c1_run")
    os.makedirs(work, exist_ok=True)
    tag = "c1_pt2" if pt2 else "c1_var"
    inp = os.path.join(work, tag + ".inp")
    synth.config_input(1, inp, pt2=pt2)
    js, basis = os.path.join(work, tag + ".json"), os.path.join(work, tag + ".basis")
    subprocess.run([os.path.join(ROOT, "oracle", "bin", "vhci_cpu"), "-n", "8", "-i", inp, "-o",
                    os.path.join(work, tag + ".freqs"), "--json", js, "--basis", basis, "--quiet"], check=True)
    return js, basis


def pt2_fields(js):
    d = json.load(open(js))
    return {"pt2_eps2": 0.01, "pt2_space": d["pt2_space"], "E_pt2": d["E"], "oracle_pt2_seconds": d["seconds"]}


def main():
    if len(sys.argv) == 3 and sys.argv[1] == "--pt2-only":
        out = json.load(open(OUT))
        out.update(pt2_fields(sys.argv[2]))
        out["workload"] = "synthetic c1 (seed 0x5648434901); sizes, E_var and checksum from the PT2-off run"
        json.dump(out, open(OUT, "w"), indent=1)
        return
    if len(sys.argv) == 4:
        js, basis, js2 = sys.argv[1], sys.argv[2], sys.argv[3]
    else:
        js, basis = oracle_run(False)
        js2, _ = oracle_run(True)
    d = json.load(open(js))
    q = np.loadtxt(basis, dtype=np.int32)
    out = {"workload": "synthetic c1 (seed 0x5648434901); sizes, E_var and checksum from the PT2-off run",
           "sizes": [d["N0"]] + [it["N"] for it in d["iters"]],
           "E_var": d["E_var"], "checksum": str(checksum(q)), "oracle_seconds": d["seconds"]}
    out.update(pt2_fields(js2))
    with open(OUT, "w") as f:
        json.dump(out, f, indent=1)


if __name__ == "__main__":
    main()
