"""Partial-convergence (dynamic) tile classes: the device classifier against the
oracle's threshold restatement, and solves with dynamic classes reaching 1e-10."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch


def _deltas():
    rng = np.random.default_rng(21)
    edges = np.array([2e-5, 1e-5, np.nextafter(1e-5, 1), 5e-6, 1e-6, np.nextafter(1e-6, 0), 5e-7, 1e-7,
                      np.nextafter(1e-7, 1), 5e-8, 1e-8, np.nextafter(1e-8, 1), 5e-9, 0.0, 1.0])
    segs = []
    for e in edges:               # one segment per edge value, the rest tiny
        d = np.full(16, 1e-12)
        d[rng.integers(16)] = e
        segs.append(d)
    for _ in range(200):          # random mixtures
        segs.append(10.0 ** rng.uniform(-10, -4, 16) * (rng.random(16) < 0.5))
    d = np.concatenate(segs)
    p_old = rng.standard_normal(d.size)
    sign = np.where(rng.random(d.size) < 0.5, -1.0, 1.0)
    return p_old, p_old + sign * d


def test_dynamic_classes_reference_thresholds(mf, orc, torch):
    p_old, p_new = _deltas()
    # compare on the computed differences (p_new - p_old as rounded in binary64)
    got = mf.dyn_classify(p_old, p_new, mode=2)
    want = orc.dyn_classes_ref(p_old, p_new)
    assert np.array_equal(got, want)
    # ragged tail
    got = mf.dyn_classify(p_old[:37], p_new[:37], mode=2)
    assert np.array_equal(got, orc.dyn_classes_ref(p_old[:37], p_new[:37]))


def test_dynamic_classes_magnitude_rule(mf, orc, torch):
    rng = np.random.default_rng(22)
    p = rng.standard_normal(16 * 64) * 10.0 ** rng.integers(-12, 1, 16 * 64)
    rms = float(np.sqrt(np.mean(p ** 2)))
    got = mf.dyn_classify(np.zeros_like(p), p, mode=1, rms=rms)
    assert np.array_equal(got, orc.dyn_classes_rel(p, rms))


@pytest.mark.parametrize("mode", [1, 2])
def test_cg_with_dynamic_classes_reaches_tolerance(mf, orc, torch, mode):
    rng = np.random.default_rng(3)
    rp, ci, va = mf.workloads.poisson2d(64)
    b, _ = mf.workloads.rhs_from_solution(rp, ci, va, 0)
    A = mf.TiledMatrix(rp, ci, va, class_policy="rowscaled")
    x, info = A.cg(b, tol=1e-10, precision="mixed", dynamic_precision=mode)
    print(f"dynamic mode {mode}: status {info['status']} iterations {info['iterations']} "
          f"replacements {info['replacements']} relres_true {info['relres_true']:.3e}", flush=True)
    assert info["status"] == 0, info
    assert orc.true_relres(rp, ci, va, x.cpu().numpy(), b) <= 1.2e-10
    A.close()
