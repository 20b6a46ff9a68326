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


def test_dynamic_classes_only_in_single_kernel_cg(mf, torch):
    """Partial-convergence classes exist in the single-kernel CG; BiCGSTAB and the
    multi-kernel solvers refuse the option instead of ignoring it."""
    rp, ci, va = mf.workloads.poisson2d(24)
    b, _ = mf.workloads.rhs_from_solution(rp, ci, va, 1)
    A = mf.TiledMatrix(rp, ci, va)
    for fn in (A.bicgstab, A.cg_multikernel, A.bicgstab_multikernel):
        with pytest.raises(mf.MfError) as e:
            fn(b, tol=1e-10, precision="mixed", dynamic_precision=1)
        assert "UNSUPPORTED" in str(e.value)
    x, info = A.cg(b, tol=1e-10, precision="mixed", dynamic_precision=1)
    assert info["status"] == 0
    A.close()


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


def test_graph_replay_follows_changed_options(mf, orc, torch):
    """A cached graph must not be replayed after an option it bakes in changes:
    fixed_iterations, max_iter, max_replacements and the tolerance each re-capture."""
    systems = _systems(mf, 6, 9)
    B = mf.Batch([(r, c, v) for r, c, v, _ in systems], class_policy="lossless")
    b = B.pack([s[3] for s in systems])
    _, inf_a, _ = B.cg(b, use_graph=True, tol=1e-10, precision="mixed")
    assert all(i["status"] == 0 for i in inf_a)
    # fixed budget of 7 iterations: every system stops at exactly 7
    _, inf_b, _ = B.cg(b, use_graph=True, tol=1e-10, precision="mixed", fixed_iterations=True, max_iter=7)
    assert [i["iterations"] for i in inf_b] == [7] * len(systems)
    # back to tolerance mode: the first graph's answer again
    _, inf_c, _ = B.cg(b, use_graph=True, tol=1e-10, precision="mixed")
    assert [i["iterations"] for i in inf_c] == [i["iterations"] for i in inf_a]
    # no residual replacements allowed: the count reported is 0
    _, inf_d, _ = B.cg(b, use_graph=True, tol=1e-10, precision="mixed", max_replacements=1, replace_delta=1e-30)
    assert all(i["replacements"] <= 1 for i in inf_d)
    # a looser tolerance needs fewer iterations
    _, inf_e, _ = B.cg(b, use_graph=True, tol=1e-4, precision="mixed")
    assert all(e["iterations"] < a["iterations"] for e, a in zip(inf_e, inf_a))
    B.close()


@pytest.mark.parametrize("opt", [{"use_x0": True}, {"transpose": True}, {"dynamic_precision": 1}])
def test_batched_refuses_options_it_cannot_honour(mf, torch, opt):
    """Every batched launch starts from x = 0 and has no transposed or dynamic-class path:
    asking for one is an error, not a silently different solve."""
    systems = _systems(mf, 3, 11)
    B = mf.Batch([(r, c, v) for r, c, v, _ in systems], class_policy="lossless")
    b = B.pack([s[3] for s in systems])
    for graph in (False, True):
        with pytest.raises(mf.MfError) as e:
            B.cg(b, use_graph=graph, tol=1e-10, precision="mixed", **opt)
        assert "UNSUPPORTED" in str(e.value)
    x, infos, _ = B.cg(b, use_graph=False, tol=1e-10, precision="mixed")  # still usable
    assert all(i["status"] == 0 for i in infos)
    B.close()


@pytest.mark.parametrize("fn", ["mf_cg_solve_multikernel", "mf_cg_solve"])
def test_cg_breakdown_keeps_last_finite_iterate(mf, torch, fn):
    """p'Ap = 0 at the first step (blocks [[0, 1], [1, 0]], b = e_1): MF_BREAKDOWN, and
    x keeps its last finite iterate (x0 = 0) instead of Inf/NaN."""
    import ctypes
    n = 64
    rp = np.arange(0, n + 1, dtype=np.int32)
    ci = np.array([i + 1 if i % 2 == 0 else i - 1 for i in range(n)], np.int32)
    va = np.ones(n)
    b = np.zeros(n)
    b[0::2] = 1.0
    A = mf.TiledMatrix(rp, ci, va)
    lb = mf._lib.lib()
    o = mf.api.solve_options(tol=1e-10, precision="fp64")
    bt = torch.as_tensor(b, device="cuda:0")
    x = torch.zeros_like(bt)
    info = mf._lib.SolveInfo()
    st = getattr(lb, fn)(A.h, ctypes.c_void_p(bt.data_ptr()), ctypes.c_void_p(x.data_ptr()), ctypes.byref(o),
                         ctypes.byref(info), None)
    torch.cuda.synchronize()
    assert mf._lib.STATUS_NAMES[st] == "MF_BREAKDOWN", (st, info.as_dict())
    xh = x.cpu().numpy()
    assert np.all(np.isfinite(xh)) and not xh.any()
    A.close()


def test_batch_timeline_and_layout(mf, orc, torch):
    """The per-system solve intervals of a launch (the kernel's clock): every system starts
    before it ends and within the launch, the group sizes follow the rows-per-workgroup
    bound, and the default SpMV reads the sliced-ELL streams (layout 2/3); layout 1 forces
    the tile chunks with the same iteration counts."""
    systems = _systems(mf, 12, 8)
    B = mf.Batch([(r, c, v) for r, c, v, _ in systems])
    B.set_rows_per_workgroup(1024)
    b = B.pack([s[3] for s in systems])
    x, infos, ms = B.cg(b, use_graph=False, tol=1e-10)
    st, en, gs = B.timeline()
    n = np.array([s[0].size - 1 for s in systems])
    assert np.all(st >= 0) and np.all(en > st) and en.max() <= 1.05 * ms + 0.05
    # the smallest power of two W with ceil(tile rows / W) * 16 <= 1024
    want = []
    for ni in n:
        tr, W = (ni + 15) // 16, 1
        while -(-tr // W) * 16 > 1024:
            W *= 2
        want.append(W)
    assert list(gs) == want, (list(gs), want)
    assert all(inf["layout"] in (2, 3) for inf in infos)
    x1, infos1, _ = B.cg(b, use_graph=False, tol=1e-10, layout="chunks")
    assert all(inf["layout"] == 1 and inf["status"] == 0 for inf in infos1)
    for inf, inf1 in zip(infos, infos1):
        assert abs(inf["iterations"] - inf1["iterations"]) <= 1
    B.close()
