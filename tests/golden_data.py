"""Loaders for the committed fixtures under tests/golden/."""
import json
from pathlib import Path

import numpy as np

GOLDEN = Path(__file__).resolve().parent / "golden"

# The reference's only fixture: 20 f32 points "close to the surface"
# (src/shape/mod.rs:110-133), inputs without expected outputs.
BENCH_POINTS = np.array(json.loads((GOLDEN / "bench_points.json").read_text())["points"], np.float32)


def load_case(name: str) -> dict:
    """A golden case: `<name>.json` manifest plus `<name>.npz` arrays."""
    meta = json.loads((GOLDEN / f"{name}.json").read_text())
    with np.load(GOLDEN / f"{name}.npz", allow_pickle=False) as z:
        arrays = {k: z[k] for k in z.files}
    return {"meta": meta, **arrays}
