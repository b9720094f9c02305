"""Learned refinement predictor end to end, small: solver runs on the GPU
produce the samples, the GPU MLP is trained on two runs and judged on a
third against the estimator-driven mesh h steps later.  The gain over the
method without the network (the estimator's present flags) is the
missed-refinement rate: the estimator flags a cell only once its error has
crossed the threshold, the network is asked to flag it h steps earlier."""
import numpy as np
import pytest

from dacti_amd import predictor, workloads
from dacti_amd.mlp import MLP

pytestmark = pytest.mark.gpu


def _data(mach, steps, horizon):
    case = workloads.shock_bubble(base=4, extra=2, mach=mach)
    print(f"[predictor] run mach={mach}", flush=True)
    s = predictor.run_samples(case, steps, horizon)
    return (np.concatenate([a for a, _, _ in s]), np.concatenate([b for _, b, _ in s]),
            np.concatenate([c for _, _, c in s]))


def test_predictor_forecasts_refinement():
    h = 4
    Xa, ya, _ = _data(1.6, 14, h)
    Xb, yb, _ = _data(2.0, 14, h)
    Xt, yt, gt = _data(1.8, 14, h)
    X, y = np.concatenate([Xa, Xb]), np.concatenate([ya, yb])
    std = predictor.Standardizer(X)
    m = MLP(X.shape[1], 128, 3, 3, seed=0)
    loss = m.train(std(X), y, epochs=40, batch=256, lr=2e-3, seed=1)
    pred = m.predict(std(Xt))
    acc = float((pred == yt).mean())
    majority = float(np.bincount(yt, minlength=3).max() / len(yt))
    miss_net = predictor.missed_refinement(pred, yt)
    miss_ind = predictor.missed_refinement(gt, yt)
    print(f"[predictor] loss {loss:.3f} acc {acc:.3f} (majority {majority:.3f}) missed refinement: "
          f"net {miss_net:.3f} indicator {miss_ind:.3f}", flush=True)
    assert acc > majority
    assert miss_net < miss_ind
    m.close()


def test_remesh_pause_and_mark_split():
    from dacti_amd import scene
    from dacti_amd.integrator import Integrator
    # thresholds that never split or coarsen by themselves: only flagged
    # leaves (and the 2:1 balance they need) split
    case = workloads.shock_bubble(base=4, extra=2, mach=1.8, error_threshold=(-1.0, 1e300))
    g = Integrator(case, scene.build_tree(case))
    g.integrate()
    a = g.get_nodes()
    g.set_remesh(False)
    g.integrate()
    b = g.get_nodes()
    # no split or unsplit while paused: same pool and leaf pattern
    assert len(a["flags"]) == len(b["flags"])
    np.testing.assert_array_equal(a["flags"] & 1, b["flags"] & 1)
    lv = predictor._active_leaves(b)
    cand = lv[(b["tree_level"][lv] < case.max_refinement_level) & ((b["flags"][lv] & 8) != 0)][:20]
    assert len(cand) == 20
    g.mark_split(cand)
    g.set_remesh(True)
    g.integrate()
    c = g.get_nodes()
    split = (c["flags"][cand] & 1) == 0
    print(f"[predictor] {split.sum()} of {len(cand)} flagged leaves split", flush=True)
    assert split.all()
    g.close()
