"""GPU parity of the single-kernel BiCGSTAB against the oracle."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch


def _nonsym(mf, name):
    w = mf.workloads
    if name == "convdiff12":
        rp, ci, va = w.convdiff3d(12)
    elif name == "convdiff_ragged":
        rp, ci, va = w.convdiff3d(9, ny=7, nz=5)     # n = 315, 315 % 16 = 11
    elif name == "add20_like":
        rng = np.random.default_rng(8)
        n = 777
        d = rng.random((n, n)) < 0.01
        np.fill_diagonal(d, False)
        r, c = np.nonzero(d)
        v = rng.standard_normal(r.size)
        import scipy.sparse as sp
        A = sp.csr_matrix((v, (r, c)), shape=(n, n)).toarray()
        A[np.arange(n), np.arange(n)] = np.abs(A).sum(1) + 1.0
        S = sp.csr_matrix(A)
        rp, ci, va = S.indptr.astype(np.int32), S.indices.astype(np.int32), S.data
    else:
        raise KeyError(name)
    b, xs = w.rhs_from_solution(rp, ci, va, 1)
    return rp, ci, va, b, xs


@pytest.mark.parametrize("variant", [2, 4])
@pytest.mark.parametrize("layout", ["auto", "chunks"])
@pytest.mark.parametrize("name", ["convdiff12", "convdiff_ragged", "add20_like"])
def test_bicgstab_fp64_iterations_and_residual(mf, orc, torch, name, layout, variant):
    rp, ci, va, b, xs = _nonsym(mf, name)
    A = mf.TiledMatrix(rp, ci, va)
    x, info = A.bicgstab(b, tol=1e-10, layout=layout, variant=variant)
    ref = orc.bicgstab(rp, ci, va, b, tol=1e-10)
    assert info["status"] == 0, info
    it, it_ref = info["iterations"], ref["iterations"]
    assert abs(it - it_ref) <= max(1, int(0.02 * it_ref)), (it, it_ref)
    rel = orc.true_relres(rp, ci, va, x.cpu().numpy(), b)
    assert rel <= 1.2e-10, rel
    A.close()


@pytest.mark.parametrize("variant", [2, 4])
@pytest.mark.parametrize("layout", ["auto", "chunks"])
@pytest.mark.parametrize("name", ["convdiff12", "add20_like"])
def test_bicgstab_mixed_reaches_fp64_tolerance(mf, orc, torch, name, layout, variant):
    rp, ci, va, b, xs = _nonsym(mf, name)
    A = mf.TiledMatrix(rp, ci, va, class_policy="rowscaled")
    x, info = A.bicgstab(b, tol=1e-10, precision="mixed", layout=layout, variant=variant)
    assert info["status"] == 0, info
    rel = orc.true_relres(rp, ci, va, x.cpu().numpy(), b)
    assert rel <= 1.2e-10, (rel, info)
    A.close()


@pytest.mark.parametrize("layout", ["auto", "chunks"])
def test_bicgstab_mixed_restarts_after_breakdown(mf, orc, torch, layout):
    """2D Poisson 256 x 256 (int8 tiles): with binary32 products the textbook recurrence
    reaches a zero rho or omega before 1e-10; the tile-mixed solve restarts from the binary64
    residual (rhat = p = r) and still reaches the FP64 tolerance."""
    rp, ci, va = mf.workloads.poisson2d(256)
    b, _ = mf.workloads.rhs_from_solution(rp, ci, va, 1)
    A = mf.TiledMatrix(rp, ci, va, class_policy="lossless")
    print(f"bicg restart poisson2d 256 {layout}", flush=True)
    x, info = A.bicgstab(b, tol=1e-10, precision="mixed", layout=layout, max_iter=20000)
    print(info, flush=True)
    assert info["status"] == 0 and info["replacements"] >= 1, info
    assert orc.true_relres(rp, ci, va, x.cpu().numpy(), b) <= 1.2e-10
    A.close()


@pytest.mark.parametrize("variant", [2, 4])
@pytest.mark.parametrize("layout", ["chunks", "sell"])
def test_bicgstab_fixed_iterations_match_oracle(mf, orc, torch, layout, variant):
    rp, ci, va, b, xs = _nonsym(mf, "convdiff12")
    A = mf.TiledMatrix(rp, ci, va)
    x, info = A.bicgstab(b, max_iter=6, fixed_iterations=True, layout=layout, variant=variant, grid=5)
    assert info["variant"] == variant
    ref = orc.bicgstab(rp, ci, va, b, tol=0.0, maxit=6)
    assert info["iterations"] == 6
    assert np.allclose(x.cpu().numpy(), ref["x"], rtol=1e-9, atol=1e-12)
    A.close()
