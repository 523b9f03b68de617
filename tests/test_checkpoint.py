"""State files (the reference's io/checkpoint.py contract) on the CPU: names,
atomic write, ordering, tolerant and pickle-free reading."""
import os

import numpy as np
import pytest

from specmc_amd import api, checkpoint as ck


def test_filename_and_roundtrip(tmp_path):
    th = (np.arange(12.0).reshape(3, 4) + 0.5j).astype(complex)
    assert ck.filename("t", 0.0020051) == "t_0.002005.npz"
    p = ck.save(str(tmp_path), th, 0.0020051, n=25, fac=40.0)
    assert os.path.basename(p) == "t_0.002005.npz"
    assert sorted(os.listdir(tmp_path)) == ["t_0.002005.npz"]      # no .tmp left
    got, meta = ck.load(p)
    assert np.array_equal(got, th)
    assert meta["t"] == 0.0020051 and meta["n"] == 25 and meta["fac"] == 40.0
    assert meta["keys"] == ["fac", "n", "t", "that"] and not meta["snapshot"]
    assert meta["path"] == os.path.abspath(p)
    # api.load_checkpoint delegates here
    got2, meta2 = api.load_checkpoint(p)
    assert np.array_equal(got2, th) and meta2["n"] == 25


def test_list_latest_and_prefixes(tmp_path):
    d = str(tmp_path)
    for t in (3e-3, 1e-3, 2e-3):
        ck.save(d, np.zeros((2, 2)), t)
    ck.save(d, np.zeros((2, 2)), 5e-3, prefix=ck.KEEP_PREFIX)
    open(os.path.join(d, "t_junk.npz"), "w").close()
    ts = [t for t, _ in ck.list_(d)]
    assert ts == [1e-3, 2e-3, 3e-3]
    assert ck.latest(d)[0] == 3e-3
    assert ck.latest(d, ck.KEEP_PREFIX)[0] == 5e-3
    assert ck.list_(os.path.join(d, "missing")) == [] and ck.latest(str(tmp_path / "x")) is None


def test_tolerant_and_pickle_free_read(tmp_path):
    th = np.ones((2, 3), dtype=complex)
    p = tmp_path / "a.npz"
    np.savez_compressed(p, that=th)
    got, meta = ck.load(str(p))
    assert np.isnan(meta["t"]) and meta["n"] is None and meta["fac"] is None
    # the reference writes n=None / fac=None as object fields: read as None,
    # never unpickled
    q = tmp_path / "b.npz"
    np.savez_compressed(q, that=th, t=1e-3, n=None, fac=None)
    got, meta = ck.load(str(q))
    assert meta["t"] == 1e-3 and meta["n"] is None and meta["fac"] is None
    # None is omitted when written here
    r = ck.save(str(tmp_path), th, 2e-3)
    assert ck.load(r)[1]["keys"] == ["t", "that"]


def test_errors(tmp_path):
    with pytest.raises(FileNotFoundError):
        ck.load(str(tmp_path / "none.npz"))
    p = tmp_path / "c.npz"
    np.savez_compressed(p, theta=np.zeros(3))
    with pytest.raises(ValueError, match="no 'that' field"):
        ck.load(str(p))
    with pytest.raises(ValueError, match="plain state file"):
        ck.read_snapshot(ck.save(str(tmp_path), np.zeros((2, 2)), 1.0))
