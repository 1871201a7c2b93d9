"""Writes c4_sizes.json: the 256 instance sizes of the batched workload
(log-uniform on [4096, 524288], seed 42, rescaled to 2^24 particles), as
tests/scenes.py:c4_sizes draws them.  Run from the repository root."""
import json
import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, os.path.join(ROOT, "tests"))
import scenes  # noqa: E402

sizes = scenes.c4_sizes()
with open(os.path.join(ROOT, "tests", "golden", "c4_sizes.json"), "w") as f:
    json.dump({"n": len(sizes), "total": sum(sizes), "seed": 42, "range": [4096, 524288], "sizes": sizes}, f)
print(len(sizes), sum(sizes), min(sizes), max(sizes))
