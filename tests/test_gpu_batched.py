"""Batched persistent CG and the multi-kernel fallback against the oracle."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch


def _systems(mf, count, seed, nmin=200, nmax=6000):
    return mf.workloads.batched_poisson(count=count, seed=seed, nmin=nmin, nmax=nmax)


@pytest.mark.parametrize("rows_per_wg", [1024, 2048, 8192])
@pytest.mark.parametrize("graph", [False, True])
def test_batched_cg_matches_oracle(mf, orc, torch, rows_per_wg, graph):
    systems = _systems(mf, 24, 5)
    B = mf.Batch([(r, c, v) for r, c, v, _ in systems])
    B.set_rows_per_workgroup(rows_per_wg)
    b = B.pack([s[3] for s in systems])
    for rep in range(2 if graph else 1):  # replays of the captured graph give the same answer
        x, infos, ms = B.cg(b, use_graph=graph, tol=1e-10)
        xs = B.unpack(x)
        for (rp, ci, va, bs), xi, inf in zip(systems, xs, infos):
            assert inf["status"] == 0, inf
            ref = orc.cg(rp, ci, va, bs, tol=1e-10)
            assert abs(inf["iterations"] - ref["iterations"]) <= max(1, int(0.02 * ref["iterations"]))
            assert orc.true_relres(rp, ci, va, xi, bs) <= 1.2e-10
    B.close()


def test_batched_mixed_reaches_tolerance(mf, orc, torch):
    systems = _systems(mf, 16, 6)
    B = mf.Batch([(r, c, v) for r, c, v, _ in systems], class_policy="rowscaled")
    b = B.pack([s[3] for s in systems])
    x, infos, ms = B.cg(b, precision="mixed", tol=1e-10)
    for (rp, ci, va, bs), xi, inf in zip(systems, B.unpack(x), infos):
        assert inf["status"] == 0, inf
        assert orc.true_relres(rp, ci, va, xi, bs) <= 1.2e-10
    B.close()


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_multikernel_cg_and_bicgstab(mf, orc, torch, prec):
    rp, ci, va = mf.workloads.poisson2d(40, 37)
    b, _ = mf.workloads.rhs_from_solution(rp, ci, va, 3)
    A = mf.TiledMatrix(rp, ci, va, class_policy="rowscaled")
    x, info = A.cg_multikernel(b, tol=1e-10, precision=prec)
    ref = orc.cg(rp, ci, va, b, tol=1e-10)
    assert info["status"] == 0
    if prec == "fp64":
        assert abs(info["iterations"] - ref["iterations"]) <= max(1, int(0.02 * ref["iterations"]))
    assert orc.true_relres(rp, ci, va, x.cpu().numpy(), b) <= 1.2e-10
    rp2, ci2, va2 = mf.workloads.convdiff3d(10)
    b2, _ = mf.workloads.rhs_from_solution(rp2, ci2, va2, 3)
    A2 = mf.TiledMatrix(rp2, ci2, va2, class_policy="rowscaled")
    x2, info2 = A2.bicgstab_multikernel(b2, tol=1e-10, precision=prec)
    ref2 = orc.bicgstab(rp2, ci2, va2, b2, tol=1e-10)
    assert info2["status"] == 0
    if prec == "fp64":
        assert abs(info2["iterations"] - ref2["iterations"]) <= max(1, int(0.02 * ref2["iterations"]))
    assert orc.true_relres(rp2, ci2, va2, x2.cpu().numpy(), b2) <= 1.2e-10
    A.close()
    A2.close()


def test_value_update_and_pattern_outer(mf, orc, torch):
    rp, ci, va = mf.workloads.poisson2d(20)
    b, _ = mf.workloads.rhs_from_solution(rp, ci, va, 1)
    A = mf.TiledMatrix(rp, ci, va)
    va2 = va * 1.5
    va2[::7] += 0.25
    A.update_values(va2)
    x, info = A.cg(b, tol=1e-10)
    ref = orc.cg(rp, ci, va2, b, tol=1e-10)
    assert abs(info["iterations"] - ref["iterations"]) <= max(1, int(0.02 * ref["iterations"]))
    n = rp.size - 1
    u = torch.as_tensor(np.random.default_rng(1).standard_normal(n), device="cuda")
    v = torch.as_tensor(np.random.default_rng(2).standard_normal(n), device="cuda")
    out = A.pattern_outer(u, v, scale=-1.0).cpu().numpy()
    rows = np.repeat(np.arange(n), np.diff(rp))
    assert np.array_equal(out, -u.cpu().numpy()[rows] * v.cpu().numpy()[ci])
    A.close()
