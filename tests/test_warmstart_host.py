"""Host-side pieces of the warm-start module (no GPU): the weights file
round trip, the weights-path override and the committed weights' metadata."""
import json
import os

import numpy as np

from pyseistr_amd import warmstart as W


def test_weights_round_trip_and_env_override(tmp_path, monkeypatch):
    p = np.random.default_rng(0).standard_normal(W.NPARAM).astype(np.float32)
    path = str(tmp_path / "w.npz")
    W.WarmStart(p, {"target_niter": 7, "dilation": [1] * 6}).save(path)
    a = W.WarmStart.load(path)
    assert np.array_equal(a.params_host, p) and a.meta["target_niter"] == 7
    monkeypatch.setenv("SH_WARMSTART_WEIGHTS", path)
    b = W.WarmStart.load()
    assert np.array_equal(b.params_host, p)


def test_committed_weights_metadata():
    if not os.path.exists(W.WEIGHTS):
        return
    with np.load(W.WEIGHTS, allow_pickle=False) as z:
        meta = json.loads(str(z["meta"]))
        assert z["params"].shape == (W.NPARAM,) and z["params"].dtype == np.float32
    assert np.all(np.isfinite(np.load(W.WEIGHTS)["params"]))
    # trained on converged-length targets of seeded volumes disjoint from the
    # held-out benchmark and test volumes (seed 5)
    assert meta["target_niter"] >= 400 and meta["seeds"][0] >= 10000
