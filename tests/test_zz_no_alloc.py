"""mf_cg_solve_noalloc against mf_cg_solve, through the C-ABI (ctypes).

The no-alloc entry point runs the same kernel on the library's static scratch array, on
two calls in a row (the second on a smaller system), each repeated after the scratch holds
the previous call's contents.  The sliced-ELL kernels (tile-mixed solves) sum every row in
a fixed order, so there the results equal mf_cg_solve's bit for bit.  The all-FP64 solves
of stencils run the tile-chunk SpMV, whose products meet in LDS through atomic adds: their
last bits vary from call to call of either entry point, so there both solutions are held
to the tolerance (same iteration count within one, both true residuals at 1e-10, and the
solutions within kappa * 2e-10 of each other)."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _solve(mf, torch, fn, A, b, **kw):
    lb = mf._lib.lib()
    o = mf.api.solve_options(**kw)
    bt = torch.as_tensor(b, device="cuda:0")
    x = torch.zeros_like(bt)
    info = mf._lib.SolveInfo()
    st = getattr(lb, fn)(A.h, ctypes.c_void_p(bt.data_ptr()), ctypes.c_void_p(x.data_ptr()), ctypes.byref(o),
                         ctypes.byref(info), None)
    torch.cuda.synchronize()
    return st, x.cpu().numpy(), info.as_dict()


CASES = [
    # (generator, size, precision): the second system is smaller than the first
    ("poisson3d", 40, "mixed"),
    ("poisson2d", 48, "fp64"),
]


def test_noalloc_matches_cg_solve_twice(mf, orc, torch):
    print("noalloc: two systems, the second smaller", flush=True)
    for gen, size, prec in CASES:
        rp, ci, va = getattr(mf.workloads, gen)(size)
        b, _ = mf.workloads.rhs_from_solution(rp, ci, va, seed=1)
        A = mf.TiledMatrix(rp, ci, va)
        try:
            # the existing entry point (builds the launch plan and the matrix workspace)
            st0, x0, i0 = _solve(mf, torch, "mf_cg_solve", A, b, precision=prec)
            assert st0 == 0, i0
            for rep in range(2):
                st1, x1, i1 = _solve(mf, torch, "mf_cg_solve_noalloc", A, b, precision=prec)
                print(f"  {gen}({size}) {prec} rep {rep}: iters {i1['iterations']} vs {i0['iterations']}",
                      flush=True)
                assert st1 == 0, (st1, i1)
                assert i1["grid"] == i0["grid"] and i1["variant"] == i0["variant"]
                assert i1["layout"] == i0["layout"] and i1["precision"] == i0["precision"]
                assert i1["time_ms"] == 0.0
                if i0["layout"] != 1:  # sliced-ELL: fixed summation order
                    assert i1["iterations"] == i0["iterations"]
                    assert i1["replacements"] == i0["replacements"]
                    assert i1["relres_true"] == i0["relres_true"]
                    assert np.array_equal(x1, x0)
                else:  # tile chunks: LDS atomic adds
                    assert abs(i1["iterations"] - i0["iterations"]) <= 1
                    assert max(i0["relres_true"], i1["relres_true"]) <= 1.01e-10
                    # 2D Poisson 48^2: kappa < 1000
                    assert np.linalg.norm(x1 - x0) <= 2e-7 * np.linalg.norm(x0)
                assert orc.true_relres(rp, ci, va, x1, b) <= 1.5e-10
            ref = orc.cg(rp, ci, va, b, tol=1e-10)
            assert abs(i0["iterations"] - ref["iterations"]) <= max(1, 0.02 * ref["iterations"])
            assert orc.true_relres(rp, ci, va, x0, b) <= 1.5e-10
        finally:
            A.close()


def test_noalloc_refuses_what_it_cannot_hold(mf, torch):
    print("noalloc: refusals", flush=True)
    rp, ci, va = mf.workloads.poisson2d(40)
    b, _ = mf.workloads.rhs_from_solution(rp, ci, va, seed=2)
    A = mf.TiledMatrix(rp, ci, va)
    try:
        # no launch plan yet for these options on this handle
        st, _, _ = _solve(mf, torch, "mf_cg_solve_noalloc", A, b, precision="fp64", grid=24)
        assert mf._lib.STATUS_NAMES[st] == "MF_ERR_UNSUPPORTED"
        # pipelined variant needs a second publish buffer pair
        st, _, _ = _solve(mf, torch, "mf_cg_solve", A, b, precision="fp64", grid=24, variant=3)
        assert st == 0
        st, _, _ = _solve(mf, torch, "mf_cg_solve_noalloc", A, b, precision="fp64", grid=24, variant=3)
        assert mf._lib.STATUS_NAMES[st] == "MF_ERR_UNSUPPORTED"
        # beyond the scratch's record capacity
        st, _, _ = _solve(mf, torch, "mf_cg_solve_noalloc", A, b, precision="fp64", grid=513)
        assert mf._lib.STATUS_NAMES[st] in ("MF_ERR_TOO_LARGE", "MF_ERR_INVALID")
    finally:
        A.close()
    # more rows than the scratch holds
    n = (1 << 21) + 16
    rpd = np.arange(n + 1, dtype=np.int32)
    cid = np.arange(n, dtype=np.int32)
    D = mf.TiledMatrix(rpd, cid, np.full(n, 2.0))
    try:
        st, _, _ = _solve(mf, torch, "mf_cg_solve_noalloc", D, np.ones(n), precision="fp64")
        assert mf._lib.STATUS_NAMES[st] == "MF_ERR_TOO_LARGE"
    finally:
        D.close()


def test_noalloc_transposed_after_transposed_solve(mf, torch):
    """With transpose set the no-alloc entry point runs on the transposed matrix, which an
    earlier transposed mf_cg_solve must have built (it allocates)."""
    print("noalloc: transposed", flush=True)
    rp, ci, va = mf.workloads.poisson3d(20)
    b, _ = mf.workloads.rhs_from_solution(rp, ci, va, seed=3)
    A = mf.TiledMatrix(rp, ci, va)
    try:
        st, _, _ = _solve(mf, torch, "mf_cg_solve_noalloc", A, b, precision="mixed", transpose=True)
        assert mf._lib.STATUS_NAMES[st] == "MF_ERR_UNSUPPORTED"
        st0, x0, i0 = _solve(mf, torch, "mf_cg_solve", A, b, precision="mixed", transpose=True)
        assert st0 == 0, i0
        st1, x1, i1 = _solve(mf, torch, "mf_cg_solve_noalloc", A, b, precision="mixed", transpose=True)
        assert st1 == 0, i1
        assert i1["layout"] == i0["layout"] and i1["layout"] != 1
        assert i1["iterations"] == i0["iterations"] and np.array_equal(x1, x0)
    finally:
        A.close()
