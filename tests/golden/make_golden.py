"""Writes tests/golden/reference_examples.json from the reference tree.

Run where the reference source tree is present (tests read only the fixture):
    python tests/golden/make_golden.py /root/reference/src/examples

The fixture holds, for each example run that Moldy ships with its output
(src/examples/*-example.out, written by Moldy 2.10), the input data files
(control and sys-spec text) and the values the output log prints that the
force path determines: Ewald self energy, K-vector count, distant potential
and pressure corrections, subcell division and neighbour-cell count, the
framework/system charge corrections and, for the static clay run (step=0,
lattice start), the real- and reciprocal-space potential energies.
"""
import json
import os
import re
import sys

RUNS = {
    # name: (control file, output log)
    "water": ("control.water", "water-example.out"),
    "tips2": ("control.tips2", "tips2-example.out"),
    "tip4p": ("control.tip4p", "tip4p-example.out"),
    "mgclh2o": ("control.mgclh2o", "mgclh2o-example.out"),
    "quartz": ("control.quartz", "quartz-example.out"),
    "argon": ("control.argon", "argon-example.out"),
    "clay": ("control.clay", "clay-example.out"),
}

NUM = r"([-+]?[0-9]*\.?[0-9]+(?:[eE][-+]?[0-9]+)?)"


def grab(pattern, text, conv=float):
    m = re.search(pattern, text)
    return [conv(g) for g in m.groups()] if m else None


def main(exdir, out):
    fix = {}
    for name, (ctlf, logf) in RUNS.items():
        ctl = open(os.path.join(exdir, ctlf)).read()
        log = open(os.path.join(exdir, logf)).read()
        m = re.search(r"sys-spec-file[ \t]*=[ \t]*([^\s#]+)", ctl)
        if m:
            sysspec = open(os.path.join(exdir, m.group(1))).read()
        else:
            # control.clay carries its sys-spec after the control 'end' line
            sysspec = ctl.split("\nend\n", 1)[1]
            ctl = ctl.split("\nend\n", 1)[0] + "\nend\n"
        exp = {}
        exp["self_energy_kjmol"] = grab(r"Ewald self-energy = " + NUM, log)
        exp["n_kvectors"] = grab(r"\*I\* (\d+) K-vectors", log, int)
        exp["distant"] = grab(r"Distant potential correction = " + NUM +
                              r", Pressure correction = " + NUM, log)
        exp["subcells"] = grab(r"divided into (\d+) subcells \((\d+)x(\d+)x(\d+)\)", log, int)
        exp["neighbour_cells"] = grab(r"Neighbour list contains (\d+) cells", log, int)
        exp["alpha"] = grab(r"Alpha parameter for Ewald sum\s+=\s+" + NUM, log)
        exp["k_cutoff"] = grab(r"Reciprocal space cut-off\s+=\s+" + NUM, log)
        exp["cutoff"] = grab(r"Interaction cut-off\s+=\s+" + NUM, log)
        exp["framework_charge"] = grab(r"Framework has net electric charge of " + NUM +
                                       r" - correction of " + NUM, log)
        exp["system_charge"] = grab(r"System has net electric charge of " + NUM +
                                    r" - correction of " + NUM, log)
        cell = re.findall(r"^\s+([abc])\s+=\s+" + NUM + r"\s+" + NUM + r"\s+" + NUM + " A",
                          log, re.M)
        exp["cell_vectors"] = {c[0]: [float(x) for x in c[1:]] for c in cell}
        if name == "clay":
            rows = re.search(r"Current values =+\n(.*)\n(.*)\n", log)
            r1 = rows.group(1).split()
            r2 = rows.group(2).split()
            exp["pe_real_kjmol"] = float(r1[2])
            exp["pe_recip_kjmol"] = float(r2[2])
        fix[name] = dict(control=ctl, sysspec=sysspec, expected=exp)
    with open(out, "w") as f:
        json.dump(fix, f, indent=1, sort_keys=True)


if __name__ == "__main__":
    exdir = sys.argv[1] if len(sys.argv) > 1 else "/root/reference/src/examples"
    main(exdir, os.path.join(os.path.dirname(os.path.abspath(__file__)),
                             "reference_examples.json"))
