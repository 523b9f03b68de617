"""Reading the trajectory fixtures written by tests/golden/make_golden.py."""
import json
import os

import numpy as np

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def names():
    return sorted(f[:-4] for f in os.listdir(GOLDEN) if f.endswith(".npz"))


def load(name):
    with np.load(os.path.join(GOLDEN, name + ".npz"), allow_pickle=False) as z:
        d = {k: z[k] for k in z.files}
    d["meta"] = json.loads(str(d["meta"]))
    return d


def seeded_theta(model, H, rng_seed, sigma=1e-2, nm=8):
    """Same generator as tests/golden/make_golden.py (kept in sync by a test)."""
    rng = np.random.default_rng(rng_seed)
    X, Z = np.meshgrid(model.x, model.z, indexing="ij")
    g = 0.5 * H * Z * (1.0 - Z)
    for m in range(nm):
        for n in range(1, nm + 1):
            a = rng.normal(0.0, sigma) / float(m * m + n * n)
            g = g + a * np.cos(model.kx[m] * X) * np.sin(n * np.pi * Z)
    return model.coeffs(g)


def rel(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))
