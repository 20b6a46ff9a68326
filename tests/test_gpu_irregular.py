"""Irregular SPD systems: lognormal row lengths, banded random columns, symmetric D-scaling.

The structure of the irregular SPD workload (lognormal(3.0, 0.8) rows clipped to [3, 400], columns
uniform in a band around the diagonal, off-diagonals U[-1, 0], 30 % exactly -1, diagonal =
|row sum| + 1e-3, symmetric scaling by D = 10^u) at sizes and scaling ranges that unpreconditioned
CG solves to 1e-10: with u in [0, 2] the oracle needs about 2,500 iterations at n = 20,000, with
u in [0, 6] it does not reach 1e-10 within 200,000 (the textbook algorithm, no preconditioner).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def system(mf):
    rp, ci, va, b, xs = mf.workloads.irregular_spd(n=20000, band=400, logd_max=2.0)
    return rp, ci, va, b


def test_irregular_fp64_matches_oracle_iterations(mf, orc, torch, system):
    rp, ci, va, b = system
    ref = orc.cg(rp, ci, va, b, tol=1e-10, maxit=30000)
    assert ref["status"] == 0
    A = mf.TiledMatrix(rp, ci, va, class_policy="fp64")
    x, info = A.cg(b, tol=1e-10, precision="fp64", max_iter=30000)
    print(f"irregular fp64: gpu {info['iterations']} oracle {ref['iterations']} relres_true {info['relres_true']:.3e}",
          flush=True)
    assert info["status"] == 0
    assert abs(info["iterations"] - ref["iterations"]) <= max(2, 0.02 * ref["iterations"])
    assert orc.true_relres(rp, ci, va, x.cpu().numpy(), b) <= 1.2e-10
    A.close()


@pytest.mark.parametrize("policy", ["lossless", "rowscaled"])
def test_irregular_mixed_reaches_fp64_tolerance(mf, orc, torch, system, policy):
    rp, ci, va, b = system
    A = mf.TiledMatrix(rp, ci, va, class_policy=policy)
    st = A.stats()
    # the -1 entries give int8 tiles; the scaled random values binary64 tiles (lossless) or,
    # rounded within 2^-24 of the row maximum, binary32/binary16 tiles (rowscaled)
    assert st["tiles_by_class"][3] > 0
    assert sum(1 for c in st["tiles_by_class"] if c > 0) >= 2
    ref = orc.cg(rp, ci, va, b, tol=1e-10, maxit=30000)
    x, info = A.cg(b, tol=1e-10, precision="mixed", max_iter=30000)
    print(f"irregular mixed {policy}: classes {st['tiles_by_class']} gpu {info['iterations']} "
          f"(repl {info['replacements']}) oracle {ref['iterations']}", flush=True)
    assert info["status"] == 0
    assert orc.true_relres(rp, ci, va, x.cpu().numpy(), b) <= 1.2e-10
    # binary32 products perturb the Krylov recurrence; the count stays close to FP64's
    assert info["iterations"] <= 1.25 * ref["iterations"] + 10
    A.close()


def test_irregular_spmv_mixed_classes(mf, orc, torch, system):
    """SpMV over binary64 and int8 tiles (lossless classes: the exact -1 entries make int8 tiles,
    the scaled random values binary64 ones) against the oracle's emulation of the same
    arithmetic."""
    rp, ci, va, b = system
    n = rp.size - 1
    A = mf.TiledMatrix(rp, ci, va, class_policy="lossless")
    rng = np.random.default_rng(5)
    xv = rng.uniform(-1, 1, n)
    y64 = A.spmv(xv).cpu().numpy()
    scale = np.abs(va).max() * np.abs(xv).max()
    assert np.max(np.abs(y64 - orc.csr_spmv(rp, ci, va, xv))) <= 1e-14 * scale * 400
    d = orc.tile_create(n, n, rp, ci, va)
    cls = orc.tile_classes(d)
    assert np.any(cls == 0) and np.any(cls == 3)
    ym = A.spmv(xv, precision="mixed").cpu().numpy()
    want = orc.mixed_tile_spmv(d, xv, cls, "mixed")
    # same products; only the binary64 summation order differs
    assert np.max(np.abs(ym - want)) <= 1e-14 * scale * 400
    A.close()
