"""A-posteriori LES study of the learned closure at a small size (fine 16^3 p=2, coarse 8^3 p=2): the whole pipeline
runs on the GPU (plain fine run, differential filter, coarse projections, model-space samples, FP32 training, coarse
BDF2 runs with the FP32 Jacobian with and without the closure) and gives finite, decaying, filtered-fine-consistent
spectra; the closure run ends closer to the filtered fine spectrum than the run without a model.  The full-size study (fine 96^3, coarse 24^3) is bench.py --les."""
import json
import sys

import numpy as np
import pytest

from pyhonee import les as LES

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def test_les_study_small():
    _say("les study: fine 16^3 -> coarse 8^3")
    out = LES.study(n_fine=16, n_coarse=8, t_end=0.5, n_out=2, coarse_dt=0.05, train_offsets=8, epochs=20, batch=1024, log=_say)
    _say(json.dumps({k: out[k] for k in ("fine_gpu_s", "training_samples", "training_loss", "filtered_ke")}))
    for name in ("no_model", "closure"):
        _say(name + " " + json.dumps({k: out[name][k] for k in ("gpu_s", "spectrum_error", "ke_error", "krylov_its")}))
    assert out["training_samples"] == 8 * 16 ** 3 * 2
    assert out["training_loss"][1] < 0.7 * out["training_loss"][0]
    ke = out["filtered_ke"]
    assert ke[0] > ke[1] > ke[2] > 0
    for name in ("no_model", "closure"):
        r = out[name]
        assert np.all(np.isfinite(r["spectrum_error"])) and r["spectrum_error"][0] < 1e-12  # same initial state
        assert max(r["spectrum_error"]) < 0.5 and max(r["ke_error"]) < 0.5
        assert r["ke"][0] > r["ke"][1] > r["ke"][2]
    # the closure's gain a posteriori: at the final time its spectrum is closer to the filtered fine run's
    # (measured on MI355X: 0.12 against 0.30 without a model)
    assert out["closure"]["spectrum_error"][-1] < out["no_model"]["spectrum_error"][-1]


def test_les_study_is_reproducible():
    """Two runs of the whole pipeline give bitwise identical results: every reduction on the path (operator scatters,
    diagonal and point-block assembly, projections, filter solves, dot products, training gradients) sums in a fixed
    order."""
    _say("les study twice: fine 12^3 -> coarse 6^3")
    kw = dict(n_fine=12, n_coarse=6, t_end=0.25, n_out=1, coarse_dt=0.05, train_offsets=4, epochs=5, batch=512, log=lambda m: None)
    a, b = LES.study(**kw), LES.study(**kw)
    assert a["training_loss"] == b["training_loss"]
    assert a["filtered_ke"] == b["filtered_ke"]
    for name in ("no_model", "closure"):
        assert a[name]["newton_its"] == b[name]["newton_its"] and a[name]["krylov_its"] == b[name]["krylov_its"]
        assert a[name]["spectrum_error"] == b[name]["spectrum_error"]
