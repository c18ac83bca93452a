"""Loaders for the committed fixtures under tests/golden/."""
import json
from pathlib import Path

import numpy as np

GOLDEN = Path(__file__).resolve().parent / "golden"

# The reference's only fixture: 20 f32 points "close to the surface"
# (src/shape/mod.rs:110-133), inputs without expected outputs.
BENCH_POINTS = np.array(json.loads((GOLDEN / "bench_points.json").read_text())["points"], np.float32)

