#!/usr/bin/env python3
"""Copy four of the reference's test inputs into tests/golden/inputs.

Run in a checkout that has the reference tree (default /root/reference):

    python tests/golden/copy_inputs.py [--ref /root/reference]

Each case keeps its YAML input as it is and, of each alpha(omega) table it
names, the rows at the case's frequencies (the reader matches rows by
frequency, so the rest is not read).  Data only; nothing is run.
"""
import argparse
import os
import shutil
import sys

import yaml

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
from plasmonx_amd.inputs import AU_TO_EV, frequencies  # noqa: E402

CASES = ["wfq_sodium_inversion", "wfqfmu_ag_inversion", "wfqfmu_hetero_agau_inversion", "wfqfmu_ag_diff_units"]


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--ref", default="/root/reference")
    a = ap.parse_args()
    for case in CASES:
        src = os.path.join(a.ref, "tests", case, case + ".yaml")
        dst = os.path.join(HERE, "inputs", case)
        os.makedirs(dst, exist_ok=True)
        shutil.copy(src, os.path.join(dst, case + ".yaml"))
        with open(src) as f:
            y = yaml.safe_load(f)
        ev = [w * AU_TO_EV for w in frequencies(y.get("field", {}))]
        for p in y["parameters"].values():
            if isinstance(p, dict) and p.get("wfqfmu file", "none") != "none":
                with open(os.path.join(a.ref, "tests", case, p["wfqfmu file"])) as f:
                    rows = [ln for ln in f if ln.strip() and not ln.startswith("#") and
                            any(abs(float(ln.replace(",", " ").split()[0]) - e) <= 1e-8 for e in ev)]
                with open(os.path.join(dst, p["wfqfmu file"]), "w") as f:
                    f.writelines(rows)


if __name__ == "__main__":
    main()
