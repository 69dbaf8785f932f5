"""The committed scheduler network (openmxp_amd/data/sched_mlp.f32) against
the data it was trained on (openmxp_amd/data/sched_train_data.npz, written
by the same training run), evaluated with the oracle's MLP forward pass."""
import os

import numpy as np
import pytest

DATA = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "openmxp_amd", "data")


def _load():
    p = os.path.join(DATA, "sched_mlp.f32")
    d = os.path.join(DATA, "sched_train_data.npz")
    if not (os.path.exists(p) and os.path.exists(d)):
        pytest.skip("no trained scheduler committed")
    params = np.fromfile(p, dtype="<f4")
    z = np.load(d)  # allow_pickle defaults to False
    return params, z["F"], z["y"]


def test_committed_network_fits_its_labels(orc):
    params, F, y = _load()
    assert params.size == orc.MLP_NPARAMS
    _, _, Z = orc.mlp_forward(params, F)
    pred = (Z[:, 1] > Z[:, 0]).astype(np.int32)
    acc = np.mean(pred == y)
    # both classes are learned, not just the majority (FP8)
    rec16 = np.mean(pred[y == 0] == 0)
    rec8 = np.mean(pred[y == 1] == 1)
    print(f"accuracy {acc:.3f}, FP16 recall {rec16:.3f}, FP8 recall {rec8:.3f} over {y.size} steps")
    assert acc >= 0.95 and rec16 >= 0.75 and rec8 >= 0.95


def test_features_are_in_range():
    _, F, _ = _load()
    assert F.shape[1] == 16 and np.all(np.isfinite(F)) and np.all(np.abs(F) <= 8.0)
