"""The launch autotuners: workgroups and recurrence variant of the single-system CG
(mf_cg_autotune) and rows per workgroup / chunk groups per wave of the batched launch
(mf_batch_autotune).  A tuned solve must keep the oracle's iteration count (2 %) and must
not be slower than the default launch (10 % allowance for timing noise); the tuners must
not write the caller's vectors."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _median_time(fn, reps=5):
    ts, last = [], None
    for _ in range(reps):
        last = fn()
        ts.append(last[1]["time_ms"] if isinstance(last[1], dict) else last[2])
    return float(np.median(ts)), last


@pytest.mark.parametrize("precision", ["mixed", "fp64"])
def test_cg_autotune_keeps_iterations_and_speed(mf, orc, torch, precision):
    rp, ci, va = mf.workloads.poisson3d(48)
    b, _ = mf.workloads.rhs_from_solution(rp, ci, va, seed=3)
    A = mf.TiledMatrix(rp, ci, va)
    bt = torch.as_tensor(b, device="cuda:0")
    t_def, (x0, i0) = _median_time(lambda: A.cg(bt, tol=1e-10, precision=precision))
    # the C-ABI with d_x = NULL: the tuner does not need (or write) the caller's x
    g, v, t = ctypes.c_int(), ctypes.c_int(), ctypes.c_double()
    st = mf._lib.lib().mf_cg_autotune(A.h, ctypes.c_void_p(bt.data_ptr()), None, mf._lib.MF_PREC_MIXED
                                      if precision == "mixed" else mf._lib.MF_PREC_FP64, ctypes.byref(g),
                                      ctypes.byref(v), ctypes.byref(t), None)
    assert st == 0 and g.value > 0 and 1 <= v.value <= 4
    t_tuned, (x1, i1) = _median_time(lambda: A.cg(bt, tol=1e-10, precision=precision))
    print(f"cg autotune {precision}: default grid {i0['grid']} variant {i0['variant']} {t_def:.3f} ms; "
          f"tuned grid {i1['grid']} variant {i1['variant']} {t_tuned:.3f} ms", flush=True)
    assert i1["grid"] == g.value and i1["variant"] == v.value
    ref = orc.cg(rp, ci, va, b, tol=1e-10)
    assert abs(i1["iterations"] - ref["iterations"]) <= max(1, 0.02 * ref["iterations"])
    assert orc.true_relres(rp, ci, va, x1.cpu().numpy(), b) <= 1.2e-10
    assert t_tuned <= 1.10 * t_def
    # every variant keeps a grid of its own: the pipelined one (at most 2 rows per thread)
    # is not handed a grid it cannot run
    x3, i3 = A.cg(bt, tol=1e-10, precision=precision, variant=3)
    assert i3["status"] == 0 and i3["variant"] == 3
    A.close()


def test_batch_autotune_keeps_iterations_and_speed(mf, orc, torch):
    systems = mf.workloads.batched_poisson(count=48, seed=7, nmin=1000, nmax=40000)
    B = mf.Batch([(r, c, v) for r, c, v, _ in systems], class_policy="lossless")
    bp = B.pack([s[3] for s in systems])
    before = bp.clone()

    def solve():
        return B.cg(bp, use_graph=True, tol=1e-10, precision="mixed")

    solve()
    t_def, _ = _median_time(solve)
    rows, unroll, t_trial = B.autotune(bp, precision="mixed", iterations=30)
    assert rows in (1024, 2048, 4096, 8192) and unroll in (1, 2)
    assert torch.equal(bp, before)  # b is only read
    solve()
    t_tuned, (x, infos, ms) = _median_time(solve)
    print(f"batch autotune: default {t_def:.3f} ms; tuned rows {rows} unroll {unroll} {t_tuned:.3f} ms", flush=True)
    for (rp, ci, va, bs), xi, inf in zip(systems, B.unpack(x), infos):
        assert inf["status"] == 0
        assert orc.true_relres(rp, ci, va, xi, bs) <= 1.2e-10
    for s in range(0, len(systems), 6):
        rp, ci, va, bs = systems[s]
        ref = orc.cg(rp, ci, va, bs, tol=1e-10)
        assert abs(infos[s]["iterations"] - ref["iterations"]) <= max(1, 0.02 * ref["iterations"])
    assert t_tuned <= 1.10 * t_def
    B.close()
