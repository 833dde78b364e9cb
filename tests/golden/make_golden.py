"""Copy the reference's own regression data into tests/golden/.

The reference (berkelbach-group/VHCI) has no unit tests; its regression data
are the example runs under examples/ (input, expected freqs.dat, terminal log)
and the force fields under input/.  This script copies the inputs and the
expected freqs.dat files verbatim and extracts the integer pins of the logs
(initial basis, basis size per heat-bath iteration, PT2 space size) into
manifest.json.  It also copies the published runs under results/ (the
freqs.dat of each force field at the eps1, Nstate, PT2 and eps2 that its
file name states, total quanta 2).  It reads the reference checkout, so it runs only where that
checkout exists; the tests use the committed copies.
"""
import json
import os
import re
import shutil
import sys

REF = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
OUT = os.path.dirname(os.path.abspath(__file__))

CASES = {
    "clo2_vci": "examples/CLO2/VCI",
    "clo2_vhci": "examples/CLO2/VHCI",
    "clo2_pt2": "examples/CLO2/VHCI+PT2",
    "acn_vci": "examples/Acetonitrile/VCI",
    "acn_vhci": "examples/Acetonitrile/VHCI",
    "acn_pt2": "examples/Acetonitrile/VHCI+PT2",
}
FIELDS = ["CLO2.inp", "CLO2_quartic.inp", "C2H4.inp", "C2H4_quartic.inp", "acetonitrile.inp",
          "ethylene_oxide.inp", "naphthalene.inp"]


def log_pins(text):
    pins = {}
    m = re.search(r"Basis functions: (\d+)", text)
    pins["basis_functions"] = int(m.group(1))
    pins["iterations"] = [int(x) for x in re.findall(r"Basis size after iteration \d+ is: (\d+)", text)]
    m = re.search(r"converged after (\d+) iterations and contains (\d+)", text)
    if m:
        pins["converged_iterations"] = int(m.group(1))
        pins["converged_size"] = int(m.group(2))
    m = re.search(r"Perturbative space contains (\d+) states", text)
    if m:
        pins["pt2_space"] = int(m.group(1))
    pins["sparse_pct"] = [float(x) for x in re.findall(r"The Hamiltonian is ([\d.]+)% sparse", text)]
    m = re.search(r"Normal modes: (\d+)", text)
    pins["modes"] = int(m.group(1))
    return pins


def main():
    manifest = {}
    for name, d in CASES.items():
        src = os.path.join(REF, d)
        inp = [f for f in os.listdir(src) if f.endswith(".inp")][0]
        shutil.copyfile(os.path.join(src, inp), os.path.join(OUT, name + ".inp"))
        shutil.copyfile(os.path.join(src, "freqs.dat"), os.path.join(OUT, name + ".freqs"))
        with open(os.path.join(src, "output.log")) as f:
            manifest[name] = log_pins(f.read())
        manifest[name]["source"] = d
    os.makedirs(os.path.join(OUT, "fields"), exist_ok=True)
    for f in FIELDS:
        shutil.copyfile(os.path.join(REF, "input", f), os.path.join(OUT, "fields", f))
    # the published runs on the force fields above (freqs.dat of each)
    os.makedirs(os.path.join(OUT, "results"), exist_ok=True)
    for f in sorted(os.listdir(os.path.join(REF, "results"))):
        if f.endswith(".dat"):
            shutil.copyfile(os.path.join(REF, "results", f), os.path.join(OUT, "results", f))
    with open(os.path.join(OUT, "manifest.json"), "w") as f:
        json.dump(manifest, f, indent=1, sort_keys=True)


if __name__ == "__main__":
    main()
