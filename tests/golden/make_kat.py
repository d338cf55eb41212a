"""Write tests/golden/kat_vectors.json from the numpy restatement (tests/np_ref.py).

violet ships no test vectors; these known-answer vectors are produced by an
implementation written independently of the C oracle and of the GPU library,
and committed so that a later change to either is caught even if np_ref.py
changes with it.  Run: python tests/golden/make_kat.py
"""
import json
import sys
from pathlib import Path

import numpy as np

sys.path.insert(0, str(Path(__file__).resolve().parents[1]))
import np_ref  # noqa: E402

rng = np.random.default_rng(2024)
xs = [0, 1, 97, 255, 256, 0xFFFFFFFF] + [int(v) for v in rng.integers(0, 2**32, 26, dtype=np.uint64)]
data = {
    "jenkins": [[x, int(np_ref.jenkins(np.uint32(x)))] for x in xs],
    "lcg_init": [],
    "xxh": [],
}
for _ in range(16):
    px, py, seed = int(rng.integers(0, 4096)), int(rng.integers(0, 4096)), int(rng.integers(0, 2**32))
    data["lcg_init"].append([[px, py, seed], int(np_ref.lcg_init(px, py, seed))])
for n in (2, 3, 4):
    for _ in range(8):
        w = [int(v) for v in rng.integers(0, 2**32, n, dtype=np.uint64)]
        data["xxh"].append([w, int(np_ref.xxh([np.uint32(v) for v in w]))])
out = Path(__file__).with_name("kat_vectors.json")
out.write_text(json.dumps(data, indent=1))
print("wrote", out)
