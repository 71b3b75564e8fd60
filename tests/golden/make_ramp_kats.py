"""Writes ramp_kats.json: the Ram-Lak known-answer vectors of aprict's test
``src/aprict/tests/ramp_kernels/test_ramp_kernels_precision.py:33-153``.

The vectors (inputs and float128 expected outputs) are parsed from that file
as text; nothing from the reference is imported or executed.  Run once in
the development container, where the reference is mounted:

    python tests/golden/make_ramp_kats.py /root/reference
"""

from __future__ import annotations

import json
import pathlib
import re
import sys


def main(ref_root: str) -> None:
    src = pathlib.Path(ref_root) / "src/aprict/tests/ramp_kernels/test_ramp_kernels_precision.py"
    text = src.read_text()
    body = text[text.index("argvalues=[") : text.index("ids=[")]
    cases = []
    # each case: (factory, input array, expected array)
    for m in re.finditer(r"\(\s*ramp_kernels\.(\w+),(.*?)dtype=np\.float128,\s*\),\s*\)", body, re.S):
        factory, chunk = m.group(1), m.group(2)
        arrays = re.findall(r"np\.array\((.*?)dtype=np\.float(?:32|128)", chunk + "dtype=np.float128", re.S)
        inp_src = arrays[0]
        exp_src = arrays[1]
        tup = re.match(r"\s*\(0,\)\s*\*\s*(\d+)\s*\+\s*\(1,\)\s*\*\s*(\d+)", inp_src)
        if tup:
            inp = [0.0] * int(tup.group(1)) + [1.0] * int(tup.group(2))
        else:
            inp = [float(x) for x in re.findall(r"-?\d+(?:\.\d*)?(?:e[-+]?\d+)?", inp_src)]
        # keep the expected values as decimal strings (they carry ~20 digits)
        exp = re.findall(r"-?\d+\.\d+(?:e-?\d+)?", exp_src)
        cases.append({"kernel": factory, "input": inp, "expected": exp})
    ids = re.findall(r'"(ramlak-[\w-]+)"', text[text.index("ids=[") :])
    for c, i in zip(cases, ids):
        c["id"] = i
    out = pathlib.Path(__file__).with_name("ramp_kats.json")
    out.write_text(json.dumps({"source": str(src.relative_to(ref_root)), "physpixsize": 1.0,
                               "cases": cases}, indent=1))
    print(f"wrote {out} with {len(cases)} cases: {[len(c['input']) for c in cases]}")


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
