"""The launch shape and distance grid that the committed tuning cache picks for
the 81,920-triangle SH-icosphere workload (binary32 CROSS, eps 1e-4): the
configuration bench.py measures."""

import json
from pathlib import Path

CACHE = Path(__file__).resolve().parents[1] / "walking-into-holes_amd" / "twos" / "tuning_cache.json"
KEY = "dev0|sh_icosphere_81920|p32|c2|e0.0001|g1|m"
SHAPE_KEYS = ("block", "waves_per_simd", "lds_stack", "step_batch", "refill_batch", "grid_kappa")


def tuned_choice() -> dict:
    return json.loads(CACHE.read_text())[KEY]["choice"]


def tuned_shape() -> dict:
    c = tuned_choice()
    return {k: c[k] for k in SHAPE_KEYS if k in c}
