"""Copies the reference's own regression-test inputs and golden values into
tests/golden/ref_cases/ (data only: input files and expected numbers).

Run in the build container, where the reference is mounted:
    python tests/golden/make_fixtures.py /root/reference
Runs on the GPU never read the reference; they use the committed copies.
"""
import json
import os
import shutil
import sys

import yaml

CASES = {  # case dir -> files to copy
    "Solvers": ["inpsd.dat.base", "jfile", "momfile", "posfile"],
    "bccFe": ["inpsd.dat.base", "jfile", "momfile", "posfile"],
    "HeisStripe": ["inpsd.dat", "jfile", "kfile", "momfile", "posfile"],
    "kagome": ["inpsd.dat", "jfile", "dmfile", "momfile", "posfile"],
    "SCsurf": ["inpsd.dat", "jij", "dmdata", "kdata", "momfile", "posfile"],
    "HeisChain": ["inpsd.dat", "jfile", "momfile", "posfile", "startmom"],
    "HeisChainAF": ["inpsd.dat", "jfile", "momfile", "posfile"],
    "FeCo": ["inpsd.dat", "jASD", "momfile", "posfile"],
    "Regression": ["inpsd.dat", "jfile", "momfile", "posfile"],
}

# Cases whose golden values come from another descriptor file than
# regulartests.yaml, and the outputs of them that the accelerated path covers.
EXTRA = {
    "Regression": ("regressionResaro.yaml", ("averages", "cumulants", "totenergy")),
}


def main(ref):
    here = os.path.dirname(os.path.abspath(__file__))
    out = os.path.join(here, "ref_cases")
    for case, files in CASES.items():
        d = os.path.join(out, case)
        os.makedirs(d, exist_ok=True)
        for f in files:
            src = os.path.join(ref, "tests", case, f)
            if os.path.exists(src):
                shutil.copyfile(src, os.path.join(d, f))
    with open(os.path.join(ref, "tests", "regulartests.yaml")) as fh:
        y = yaml.safe_load(fh)
    golden = {}
    for c in y["caselist"]:
        case = os.path.basename(os.path.normpath(c["workingdir"]))
        if case not in CASES:
            continue
        entries = golden.setdefault(case, [])
        for comp in c["comparisons"]:
            for t in comp["testlist"]:
                entries.append({"name": c["name"], "outfile": os.path.basename(comp["outfile"]),
                                "headers": comp.get("headers", []),
                                "select": t["select"], "extract": t["extract"],
                                "expected": t["expected"],
                                "comparison": t.get("comparison_func", "similar"),
                                "type": t.get("type", "")})
    for case, (yml, kinds) in EXTRA.items():
        with open(os.path.join(ref, "tests", yml)) as fh:
            y = yaml.safe_load(fh)
        entries = golden.setdefault(case, [])
        for c in y["caselist"]:
            if os.path.basename(os.path.normpath(c["workingdir"])) != case:
                continue
            for comp in c["comparisons"]:
                if os.path.basename(comp["outfile"]).split(".")[0] not in kinds:
                    continue
                for t in comp["testlist"]:
                    entries.append({"name": c["name"], "outfile": os.path.basename(comp["outfile"]),
                                    "headers": comp.get("headers", []),
                                    "select": t["select"], "extract": t["extract"],
                                    "expected": t["expected"],
                                    "comparison": t.get("comparison_func", "similar"),
                                    "type": t.get("type", ""), "source": yml})
    with open(os.path.join(out, "golden.json"), "w") as fh:
        json.dump(golden, fh, indent=1, sort_keys=True)


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
