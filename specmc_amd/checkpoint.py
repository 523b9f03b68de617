"""Run checkpoints: the reference's state files and device snapshots that
resume a march bit for bit.

State files keep the reference's contract (io/checkpoint.py:1-103): a state
is ``<dir>/<prefix>_<t:%.6f>.npz`` with the complex coefficient array
``that`` and the scalars ``t``, ``n`` and ``fac``; it is written atomically
through a sibling ``<name>.npz.tmp.npz`` that is then renamed; reading is
tolerant (a missing ``t`` / ``n`` / ``fac`` reads as nan / None / None, the
keys are reported in ``meta['keys']``).  Reading never unpickles: a field
stored as a Python object (the reference writes ``n=None`` that way) reads
as None.

A snapshot (``save_run`` / ``load_run``, ``Engine.snapshot`` /
``Engine.restore``) is a state file with more fields: what one run's march
needs to continue exactly as if it had not stopped -- the BDF2 history
(previous theta, previous stream function, previous dt, the history flag),
the step and solver counters, and on the lagged-LU route (or the mixed
rescue) the frozen factor, its pivots, the frozen eta and the refresh
counters.  Its ``that``/``t``/``n``/``fac`` fields are the reference's, so
the reference's reader and ``load`` read it as a plain state.  A snapshot
restores into any slot of an engine of the same model kind, shape and
precision, which is how a served run is preempted and resumed
(``serve.Server.preempt``).  In FP64 the resumed march equals the uninterrupted
one bit for bit in any such engine.  In mixed precision it does too when the
runs around it are the same; otherwise its reused FP32 rescue factor may be
refreshed at other steps (runs past half of their factor's life join another
run's refresh), which changes its answers within the rescue's 1e-12 residual,
not beyond.
"""
from __future__ import annotations

import os

import numpy as np

__all__ = ["DEFAULT_PREFIX", "KEEP_PREFIX", "filename", "path_for", "save", "list_",
           "latest", "load", "save_run", "load_run", "SNAP_PREFIX"]

#: state prefix of the production march (``t_<time>.npz``)
DEFAULT_PREFIX = "t"
#: state prefix of the periodically kept (resume) states
KEEP_PREFIX = "keep"
#: keys of a device snapshot inside a state file
SNAP_PREFIX = "snap__"


def filename(prefix, t):
    """-> the checkpoint file name of ``prefix`` at time ``t``
    (io/checkpoint.py:41-43)."""
    return "%s_%.6f.npz" % (prefix, float(t))


def path_for(directory, prefix, t):
    return os.path.join(directory, filename(prefix, t))


def _write(directory, prefix, t, fields):
    os.makedirs(directory, exist_ok=True)
    fn = path_for(directory, prefix, t)
    tmp = fn + ".tmp.npz"          # savez appends .npz to a name without it
    np.savez_compressed(tmp, **fields)
    os.replace(tmp, fn)
    return fn


def _state_fields(that, t, n, fac):
    f = dict(that=np.asarray(that, dtype=complex), t=float(t))
    # None is left out rather than stored as an object: the reader treats a
    # missing n / fac as None
    if n is not None:
        f["n"] = int(n)
    if fac is not None:
        f["fac"] = float(fac)
    return f


def save(directory, that, t, n=None, fac=None, prefix=DEFAULT_PREFIX):
    """Atomically write one state to ``directory``; -> the final path
    (io/checkpoint.py:51-58)."""
    return _write(directory, prefix, t, _state_fields(that, t, n, fac))


def list_(directory, prefix=DEFAULT_PREFIX):
    """-> ``[(t, path)]`` of the ``prefix`` states in ``directory``, by time
    (io/checkpoint.py:61-73)."""
    out = []
    if not os.path.isdir(directory):
        return out
    head = prefix + "_"
    for f in os.listdir(directory):
        if f.startswith(head) and f.endswith(".npz") and not f.endswith(".tmp.npz"):
            try:
                out.append((float(f[len(head):-4]), os.path.join(directory, f)))
            except ValueError:
                pass
    return sorted(out)


def latest(directory, prefix=DEFAULT_PREFIX):
    """-> ``(t, path)`` of the newest state, or None (io/checkpoint.py:76-79)."""
    states = list_(directory, prefix)
    return states[-1] if states else None


def _scalar(z, key, kind):
    if key not in z.files:
        return None
    try:
        v = z[key]
    except ValueError:             # an object field (None): not unpickled
        return None
    if v.dtype == object or v.size != 1:
        return None
    return kind(v.reshape(()))


def _open(path):
    if not os.path.exists(path):
        raise FileNotFoundError(path)
    return np.load(path, allow_pickle=False)


def load(path):
    """-> ``(theta_coefficients, meta)`` from one state file
    (io/checkpoint.py:82-103): FileNotFoundError for a missing file,
    ValueError for a file without ``that``."""
    with _open(path) as z:
        files = sorted(k for k in z.files if not k.startswith(SNAP_PREFIX))
        if "that" not in z.files:
            raise ValueError("%s has no 'that' field (keys: %s)" % (path, files))
        that = np.asarray(z["that"], dtype=complex)
        t = _scalar(z, "t", float)
        meta = dict(path=os.path.abspath(path), keys=files,
                    t=float("nan") if t is None else t,
                    n=_scalar(z, "n", int), fac=_scalar(z, "fac", float),
                    snapshot=any(k.startswith(SNAP_PREFIX) for k in z.files))
    return that, meta


def save_run(directory, model, that, t, n=None, fac=None, prefix=KEEP_PREFIX):
    """Write the state ``that`` of ``model`` (a device model that has just
    returned it) together with the model's device snapshot; -> the path."""
    snap = model.eng.snapshot(0)
    fields = _state_fields(that, t, n, fac)
    for k, v in snap.items():
        fields[SNAP_PREFIX + k] = v
    return _write(directory, prefix, t, fields)


def read_snapshot(path):
    """-> ``(that, meta, snapshot dict)`` of a snapshot file."""
    that, meta = load(path)
    with _open(path) as z:
        snap = {k[len(SNAP_PREFIX):]: np.array(z[k]) for k in z.files
                if k.startswith(SNAP_PREFIX)}
    if not snap:
        raise ValueError("%s is a plain state file, not a device snapshot" % (path,))
    return that, meta, snap


def load_run(path, model):
    """Restore ``model``'s device state from a snapshot file; -> ``(that,
    meta)``.  Marching on from ``that`` continues the saved march exactly."""
    that, meta, snap = read_snapshot(path)
    model.eng.restore(0, snap)
    return model._resume_from_device(), meta
