"""Write the example-run fixtures used by tests/test_oracle_example.py.

  example_out_vmax.txt     the max-|V| trace of the reference's own output
                           file example/example.out:17-36 (data, as printed)
  example_vmax_oracle.txt  the same trace from the oracle, mixed (16) and
                           FP32 (32) variants: float32 bit patterns of the
                           kernel__vmax maxima (normalised units) and the
                           printed values (x UC x M0, ES9.2)
  example_medium.txt       the set-up summary lines of example.out:9-12
                           (stability c, min/max velocity) as printed

Run in the build container (the reference is mounted there); the oracle run
takes about 25 min (mixed) plus 15 min (FP32) on 8 cores:

    python tests/golden/make_example_trace.py [--prec 16,32]
"""
import argparse
import os
import re
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(HERE))

from oracle import oracle as O  # noqa: E402
from example_case import NT, es92, example_config, report_steps  # noqa: E402

REF_OUT = "/root/reference/example/example.out"


def write_reference_trace():
    text = open(REF_OUT).read()
    rows = re.findall(r"it=(\d+),.*\(\s*(\S+)\s+(\S+)\s+(\S+)\s*\)", text)
    with open(os.path.join(HERE, "example_out_vmax.txt"), "w") as f:
        f.write("# max|V| trace printed by the reference, example/example.out:17-36\n")
        f.write("# it vx vy vz (m/s, ES9.2 as printed)\n")
        for it, a, b, c in rows:
            f.write(f"{int(it)} {a} {b} {c}\n")
    med = {}
    for key, pat in (("c", r"Stability\s+Condition c\s*:\s*(\S+)"), ("vmin", r"Minimum velocity\s*:\s*(\S+)"),
                     ("vmax", r"Maximum velocity\s*:\s*(\S+)"), ("fmax", r"Maximum frequency\s*:\s*(\S+)"),
                     ("r", r"Wavelength Condition r\s*:\s*(\S+)")):
        med[key] = re.search(pat, text).group(1)
    with open(os.path.join(HERE, "example_medium.txt"), "w") as f:
        f.write("# set-up summary printed by the reference, example/example.out:9-13 (F15.3)\n")
        for k, v in med.items():
            f.write(f"{k} {v}\n")
    print(f"reference trace: {len(rows)} rows", flush=True)


def oracle_trace(prec):
    cfg = example_config()
    t0 = time.time()
    m = O.build_model(cfg)
    run = O.OracleRun(m, prec)
    steps = set(report_steps(NT))
    rows = []
    for it in range(1, NT + 1):
        if it in steps:
            v = run.vmax()
            va = v.astype(np.float32) * np.float32(m.UC) * np.float32(m.M0)
            rows.append((it, v.copy(), va))
            print(f"prec {prec} it={it:07d} ({es92(va[0])} {es92(va[1])} {es92(va[2])}) "
                  f"{time.time() - t0:.0f}s", flush=True)
        if it < NT:
            run.step()
    return rows


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--prec", default="16,32")
    args = ap.parse_args()
    write_reference_trace()
    out = os.path.join(HERE, "example_vmax_oracle.txt")
    lines = ["# kernel__vmax trace of the oracle on the reference example (tests/example_case.py)",
             "# prec it hex(vx) hex(vy) hex(vz) printed(vx vy vz)"]
    for prec in [int(p) for p in args.prec.split(",")]:
        for it, v, va in oracle_trace(prec):
            hx = " ".join(f"{int(x):08x}" for x in v.astype(np.float32).view(np.uint32))
            lines.append(f"{prec} {it} {hx} {es92(va[0])} {es92(va[1])} {es92(va[2])}")
    with open(out, "w") as f:
        f.write("\n".join(lines) + "\n")
    print(f"wrote {out}")


if __name__ == "__main__":
    main()
