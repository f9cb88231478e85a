"""Records the reference's MSM configuration choices as golden rows.

The reference picks an MSM configuration (get_best_config, core/src/msm.rs:35-98)
by running its stdlib-only cost model, core/src/msm/tuning/cost_model.py, with
--k (log2 of the point count, rounded up), --n (batches), --l / --p (scalar and
point bits: 254 / 254 for bn254, msm.rs:23-33) and --mem (memory limit), and
reading five lines: alpha, s, c, w, h.  Nothing else in the reference's tests
fixes those choices (the argmin's tie-breaking follows the model's search
order), so this script runs the model itself, in this container only, over a
grid of (k, n, mem), and writes msm_cost_model_rows.json beside it.  The rows
are data: inputs and the five outputs per row.

Usage: python tests/golden/make_msm_config_rows.py [/path/to/reference]
"""
from __future__ import annotations

import json
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
KS = (10, 12, 14, 16, 18, 20, 22, 24, 26)
NS = (1, 2, 3, 4, 8)
MEMS = (1 << 30, 4 << 30, 16 << 30, 40 << 30, 288 << 30)
BITS = 254  # bn254 scalar and point bits (core/src/msm.rs:23-33)


def main(ref: str) -> None:
    script = os.path.join(ref, "core", "src", "msm", "tuning", "cost_model.py")
    rows = []
    for k in KS:
        for n in NS:
            for mem in MEMS:
                out = subprocess.run(
                    [sys.executable, script, f"--k={k}", f"--n={n}", f"--l={BITS}", f"--p={BITS}", f"--mem={mem}"],
                    stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True, cwd=ref)
                if out.returncode != 0:
                    rows.append({"k": k, "n": n, "mem": mem, "error": out.stderr.strip().splitlines()[-1]})
                    continue
                vals = [int(x) for x in out.stdout.split()]
                alpha, s, c, w, h = vals[:5]
                rows.append({"k": k, "n": n, "mem": mem, "alpha": alpha, "s": s, "c": c, "w": w, "h": h})
    path = os.path.join(HERE, "msm_cost_model_rows.json")
    with open(path, "w") as f:
        json.dump({"source": "core/src/msm/tuning/cost_model.py (args --k --n --l=254 --p=254 --mem)",
                   "rows": rows}, f, indent=0)
        f.write("\n")
    print(f"{len(rows)} rows -> {path}")


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
