"""Extract the 200-bit Sage values of tests/test_rholut.cpp into
tests/golden/rholut_golden.json (data only: inputs and expected outputs).

Run from the repository root in a container that holds the reference:

    python tests/golden/make_rholut_golden.py /root/reference

The reference file is read as text.  Two things are recorded:
  * "inputs": the point data of rholut_test_data() (test_rholut.cpp:49-131),
    name -> value string, with vector/matrix entries as "name(i)" and
    "name(i,j)";
  * "cases": for every BOOST_AUTO_TEST_CASE, the high-precision literals in
    source order with their line number and the text that precedes the
    literal on its line (so each value can be audited against the source).
"""
import json
import os
import re
import sys

NUM = re.compile(r"(-\s*)?\b(\d+\.\d{12,}(?:[eE][-+]?\d+)?)L?")
ASSIGN = re.compile(r"^\s*([A-Za-z_]+(?:\(\d(?:,\d)?\))?)\s*=\s*(.+?);")


def main(ref):
    path = os.path.join(ref, "tests", "test_rholut.cpp")
    with open(path) as fh:
        lines = fh.read().splitlines()
    inputs = {}
    in_data = False
    for ln in lines:
        if "void rholut_test_data(" in ln:
            in_data = True
            continue
        if in_data:
            if ln.startswith("}"):
                break
            m = ASSIGN.match(ln)
            if m:
                val = m.group(2).replace(" ", "").rstrip("L").rstrip(".")
                inputs[m.group(1)] = val
    cases = {}
    current = None
    for no, ln in enumerate(lines, 1):
        m = re.match(r"BOOST_AUTO_TEST_CASE\(\s*(\w+)\s*\)", ln)
        if m:
            current = m.group(1)
            cases[current] = []
            continue
        if current is None:
            continue
        for mm in NUM.finditer(ln):
            sign = "-" if mm.group(1) else ""
            cases[current].append({
                "line": no,
                "context": ln[:mm.start()].strip(),
                "value": sign + mm.group(2),
            })
    out = {"source": "tests/test_rholut.cpp (200-bit Sage values)",
           "inputs": inputs, "cases": cases}
    dst = os.path.join(os.path.dirname(os.path.abspath(__file__)),
                       "rholut_golden.json")
    with open(dst, "w") as fh:
        json.dump(out, fh, indent=1)
    print(dst, len(inputs), sum(len(v) for v in cases.values()))


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
