"""Collect the reference's own 512x512 Riemann run results as a fixture.

The reference ships the configuration and the end-of-run diagnostics of its
CuPy runs (results/riemann_config*/<run>/config.json and diagnostics.json).
This script copies those numbers, unchanged, into reference_run_anchors.json
next to itself, so GPU tests can compare full-size device runs with them
without the reference tree.

Run:  python tests/golden/make_anchors.py
"""

from __future__ import annotations

import json
from pathlib import Path

REF = Path("/root/reference")
OUT = Path(__file__).resolve().parent / "reference_run_anchors.json"


def main():
    runs = []
    for diag in sorted(REF.glob("results/riemann_config*/*/diagnostics.json")):
        cfg = json.loads((diag.parent / "config.json").read_text())
        runs.append({"source": str(diag.relative_to(REF)), "config": cfg, "diagnostics": json.loads(diag.read_text())})
    OUT.write_text(json.dumps({"runs": runs}, indent=1, sort_keys=True) + "\n")
    print(f"wrote {OUT} ({len(runs)} runs)")


if __name__ == "__main__":
    main()
