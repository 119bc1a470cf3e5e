"""Extract the reference's own known-answer values for the Laplacian into a fixture.

    python tests/golden/make_golden.py /root/reference

Reads testsuite/components/03-derivatives_3d.test (expected error norms and their
precision) and 03-derivatives_3d.01.inp (box, spacing, block size) as text and writes
tests/golden/derivatives_3d.json.  Nothing of the reference is executed.
"""
import json
import re
import sys
from pathlib import Path


def main(ref: Path) -> None:
    comp = ref / "testsuite" / "components"
    test = (comp / "03-derivatives_3d.test").read_text()
    inp = (comp / "03-derivatives_3d.01.inp").read_text()
    first = test.split("Input: 03-derivatives_3d.02")[0]
    prec = float(re.search(r"Precision\s*:\s*([0-9.eE+-]+)", first).group(1))
    vals = {}
    for kind in ("real", "complex"):
        m = re.search(rf"Laplacian {kind} bsize =\s+1', 9\) ; ([0-9.eE+-]+)", first)
        vals[kind] = float(m.group(1))
    num = lambda k: float(re.search(rf"^{k}\s*=\s*([0-9.eE+-]+)", inp, re.M).group(1))
    out = {
        "source": "testsuite/components/03-derivatives_3d.test + 03-derivatives_3d.01.inp",
        "routine": "src/grid/derivatives_inc.F90:378-501",
        "radius": num("Radius"),
        "spacing": num("Spacing"),
        "states_block_size": int(num("StatesBlockSize")),
        "precision": prec,
        "laplacian_error": vals,
    }
    dst = Path(__file__).resolve().parent / "derivatives_3d.json"
    dst.write_text(json.dumps(out, indent=2) + "\n")
    print(dst.read_text())


if __name__ == "__main__":
    main(Path(sys.argv[1] if len(sys.argv) > 1 else "/root/reference"))
