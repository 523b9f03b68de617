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


def guard_theta(model, H, amp):
    """Heated conductive profile + ``amp cos(pi x) sin(pi z)`` (the state of
    the positivity-guard fixture; same generator as make_golden.py)."""
    X, Z = np.meshgrid(model.x, model.z, indexing="ij")
    return model.coeffs(0.5 * H * Z * (1.0 - Z)
                        + amp * np.cos(np.pi * X) * np.sin(np.pi * Z))


def initial_theta(model, g):
    """The initial state of fixture ``g`` rebuilt on ``model`` from its
    recipe (a stored state is taken from the fixture itself)."""
    from specmc_amd.core import bench_ic
    meta = g["meta"]
    ic = meta["ic"]
    H = float(meta["fields"].get("internal_heating", 0.0))
    kind = ic["kind"]
    if kind == "conductive":
        return model.seed(amp=ic["amp"], m=ic.get("m", 1), n=ic.get("n", 1))
    if kind == "analytic_bl":
        return bench_ic.analytic_bl_theta(model)
    if kind == "guard":
        return guard_theta(model, H, ic["amp"])
    if kind == "state":
        return np.array(g["theta0"])
    return seeded_theta(model, H, ic["seed"])


def branch_names():
    """Fixtures that record the coupled model's counters at every step."""
    out = []
    for n in names():
        if n.startswith("oracle_"):
            continue
        with np.load(os.path.join(GOLDEN, n + ".npz"), allow_pickle=False) as z:
            if "counters" in z.files:
                out.append(n)
    return out
