"""Host-side checks of the learned preconditioner's shipped artefacts (no GPU)."""
import json
import os

import numpy as np

from oracle import net_oracle as no

WEIGHTS = os.path.join(os.path.dirname(__file__), "..", "caffe_old_amd", "weights", "pcorr_net_cavity128.npy")


def test_trained_weights_load_safely():
    th = np.load(WEIGHTS, allow_pickle=False)
    assert th.dtype == np.float32 and th.shape == (no.param_count(),)
    assert np.all(np.isfinite(th))


def test_training_record():
    meta = json.load(open(os.path.splitext(WEIGHTS)[0] + ".json"))
    assert meta["grid"] == 128
    assert meta["loss_train_after"] < 0.2 * meta["loss_before"]
    assert meta["loss_heldout_mean"] < 0.1


def test_trained_network_is_linear_and_finite_on_oracle():
    """The shipped parameters define a finite linear map (checked on a small coarse grid
    with the FP64 oracle)."""
    th = np.load(WEIGHTS, allow_pickle=False).astype(np.float64)
    r = np.random.default_rng(0)
    x1, x2 = r.standard_normal((8, 8, 8)), r.standard_normal((8, 8, 8))
    y1, y2 = no.forward(th, x1), no.forward(th, x2)
    assert np.all(np.isfinite(y1))
    assert np.allclose(no.forward(th, x1 + 2 * x2), y1 + 2 * y2, rtol=1e-9, atol=1e-9 * np.abs(y1).max())
