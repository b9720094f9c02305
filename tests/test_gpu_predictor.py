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


@pytest.mark.parametrize("which", ["multi2d", "multi3d", "euler3d"])
def test_gpu_features_equal_host_restatement(which):
    """dacti_features (one thread per leaf, a tree descent per neighbour
    position) against predictor.features_raw on the exported pool: the same
    leaves in the same order, every feature bitwise except log10 of the
    error (libm vs the device's log10, to an FP32 ulp); the standardised form
    equals (x - mu) / sd in FP32; leaf centres and levels as exported."""
    from dacti_amd import scene
    from dacti_amd.devmem import DeviceArray
    from dacti_amd.integrator import Integrator
    if which == "multi2d":
        case = workloads.shock_bubble(base=4, extra=2, mach=1.8)
    elif which == "multi3d":
        case = workloads.shock_drop(base=3, extra=2)
    else:
        from tests.helpers import tiny_box
        case = tiny_box(level=3, dim=3)
    g = Integrator(case, scene.build_tree(case))
    for _ in range(2):
        g.integrate()
    print(f"[predictor] features {which}", flush=True)
    ids, X, ctr, lev = predictor.gpu_features(g, case, with_centers=True)
    nodes = g.get_nodes()
    lv, Xh = predictor.features_raw(nodes, case)
    np.testing.assert_array_equal(ids.numpy(), lv)
    Xg = X.numpy()
    assert Xg.shape == Xh.shape
    np.testing.assert_array_equal(Xg[:, :-1], Xh[:, :-1])
    np.testing.assert_allclose(Xg[:, -1], Xh[:, -1], rtol=2e-7, atol=1e-7)
    np.testing.assert_array_equal(ctr.numpy(), nodes["center"][lv])
    np.testing.assert_array_equal(lev.numpy(), nodes["tree_level"][lv])
    std = predictor.Standardizer(Xh)
    mu = DeviceArray.from_numpy(std.mu.astype(np.float32))
    sd = DeviceArray.from_numpy(std.sd.astype(np.float32))
    _, Xs = predictor.gpu_features(g, case, mu, sd)
    np.testing.assert_array_equal(Xs.numpy()[:, :-1], std(Xg)[:, :-1])
    # point location on the device
    levs = predictor.gpu_locate_levels(g, ctr)
    np.testing.assert_array_equal(levs, nodes["tree_level"][predictor.locate(nodes, nodes["center"][lv])])
    g.close()


def test_gpu_mark_predicted_with_rings():
    """dacti_mark_predicted flags the leaves whose logits pick refine, plus
    one ring of active face neighbours (predictor.dilate), as should_split."""
    from dacti_amd import scene
    from dacti_amd.devmem import DeviceArray
    from dacti_amd.integrator import Integrator
    case = workloads.shock_bubble(base=4, extra=2, mach=1.8, error_threshold=(-1.0, 1e300))
    g = Integrator(case, scene.build_tree(case))
    g.integrate()
    ids, X = predictor.gpu_features(g, case)
    n = len(ids)
    rng = np.random.default_rng(3)
    cls = rng.choice(3, n, p=[0.45, 0.45, 0.1])
    logits = np.zeros((n, 3), np.float32)
    logits[np.arange(n), cls] = 1.0
    nodes = g.get_nodes()
    marked = predictor.gpu_mark(g, DeviceArray.from_numpy(logits), ids, rings=1)
    after = g.get_nodes()
    lv = ids.numpy()
    before = (nodes["flags"] & 32) != 0
    want = predictor.dilate(nodes, lv[cls == 2], rings=1)
    want = want[((nodes["flags"][want] & 8) != 0)]  # can_split
    got = np.nonzero(((after["flags"] & 32) != 0) & ~before)[0]
    np.testing.assert_array_equal(np.sort(got), np.sort(want[~before[want]]))
    assert marked == len(want) > 0
    g.close()
