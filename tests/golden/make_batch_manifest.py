"""Writes tests/golden/batch_manifest.json: the 64-chunk uneven batch (seed 0)."""
import json
import os
import sys

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__)))))
from unnatural_planets_amd.spec import batch_manifest  # noqa: E402

if __name__ == "__main__":
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "batch_manifest.json")
    with open(out, "w") as f:
        json.dump({"seed": 0, "chunks": batch_manifest(0, 64)}, f, indent=1)
    print(out)
