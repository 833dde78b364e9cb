"""Oracle reference of the 15-mode heat-bath workload (variational stage).

Runs the CPU oracle (about 20 minutes on 5-8 cores) on the synthetic
15-mode field with PT2 switched off, and stores the basis size after every
heat-bath iteration, the 100 variational energies (%.17g) and the checksum
of the final basis (the same checksum vhci_run_ctx reports), in
c1_oracle.json.  With an existing oracle run directory, pass the paths of
its JSON summary and basis file to skip the run.
"""
import json
import os
import subprocess
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)


def checksum(q):
    v = np.ascontiguousarray(q, dtype=np.int32).ravel().astype(np.int64).astype(np.uint64)
    with np.errstate(over="ignore"):
        pw = np.cumprod(np.full(v.size, 1099511628211, dtype=np.uint64), dtype=np.uint64)
        return int(np.sum(v * pw, dtype=np.uint64))


def main():
    from vhci_amd import synth
    if len(sys.argv) == 3:
        js, basis = sys.argv[1], sys.argv[2]
    else:
        work = os.path.join(HERE, "_c1_run")
        os.makedirs(work, exist_ok=True)
        inp = os.path.join(work, "c1_var.inp")
        synth.config_input(1, inp, pt2=False)
        js, basis = os.path.join(work, "c1_var.json"), os.path.join(work, "c1_var.basis")
        subprocess.run([os.path.join(ROOT, "oracle", "bin", "vhci_cpu"), "-n", "8", "-i", inp, "-o",
                        os.path.join(work, "c1_var.freqs"), "--json", js, "--basis", basis, "--quiet"], check=True)
    d = json.load(open(js))
    q = np.loadtxt(basis, dtype=np.int32)
    out = {"workload": "synthetic c1 (seed 0x5648434901), PT2 off",
           "sizes": [d["N0"]] + [it["N"] for it in d["iters"]],
           "E_var": d["E_var"], "checksum": str(checksum(q)), "oracle_seconds": d["seconds"]}
    with open(os.path.join(HERE, "c1_oracle.json"), "w") as f:
        json.dump(out, f, indent=1)


if __name__ == "__main__":
    main()
