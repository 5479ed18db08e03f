"""Writes tests/golden/golden.json: FP64 oracle results for the small seeded
books of cases.py, raw Sobol integers and inverse-normal values.  The oracle
is the CPU restatement of the reference engine (oracle/); the fixtures pin it
against later changes and give the device tests a committed reference.

    python tests/golden/make_golden.py
"""
import json
import subprocess
import sys
from pathlib import Path

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))
sys.path.insert(0, str(Path(__file__).resolve().parent))

from thoth_amd.book import BuiltRequest  # noqa: E402

import cases  # noqa: E402

FIELDS = ("premium", "premium_trust", "delta", "gamma", "vega", "rho")
CONTRACT_FIELDS = ("contract_premium", "contract_trust", "contract_delta", "contract_vega")


def result_dict(built, res):
    n = max(1, len(built.spec.contracts))
    d = {k: getattr(res, k) for k in FIELDS}
    for k in CONTRACT_FIELDS:
        d[k] = built.contract_values(res, k, n)
    d["n_dates"] = res.n_dates
    return d


def main():
    subprocess.run(["make", "-C", str(ROOT / "oracle")], check=True, stdout=subprocess.DEVNULL)
    from oracle import pyoracle
    out = {"books": {}, "sobol": [], "normal": []}
    for name, spec in cases.cases().items():
        b = BuiltRequest(spec)
        out["books"][name] = result_dict(b, pyoracle.price(b))
    for dim, skip, n in cases.SOBOL:
        pts = pyoracle.sobol_int(dim, n, skip)
        out["sobol"].append({"dim": dim, "skip": skip, "n": n, "points": pts.astype(int).tolist()})
    z = pyoracle.normal_cdf_inv(cases.NORMAL_P)
    out["normal"] = [{"p": p, "z": float(v)} for p, v in zip(cases.NORMAL_P, z)]
    path = Path(__file__).resolve().parent / "golden.json"
    path.write_text(json.dumps(out, sort_keys=True, separators=(",", ":")) + "\n")
    print("wrote", path)


if __name__ == "__main__":
    main()
