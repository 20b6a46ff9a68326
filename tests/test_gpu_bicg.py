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


@pytest.mark.parametrize("nx", [256, 512])
@pytest.mark.parametrize("layout", ["auto", "chunks"])
def test_bicgstab_mixed_follows_fp64(mf, orc, torch, layout, nx):
    """2D Poisson 256^2 and 512^2 (int8 tiles).  With binary32 products of p and s the
    recurrence needed 4x the iterations at 256^2 and diverged at 512^2; with the stored
    values widened and multiplied by the binary64 vectors the tile-mixed solve is the FP64
    recurrence on the same values, summed in another order (class streams).  On these
    symmetric systems BiCGSTAB's count is that sensitive to summation order alone: the
    oracle takes 515 iterations at 256^2 and 508 with every row's entries reversed (SciPy's
    bicgstab 548); at 512^2 1,099, 887 and 1,018 (24 % apart).  So the counts are held to
    30 % of the FP64 solve's (the 3D convection-diffusion system holds 2 %), with both at
    1e-10."""
    rp, ci, va = mf.workloads.poisson2d(nx)
    b, _ = mf.workloads.rhs_from_solution(rp, ci, va, 1)
    A = mf.TiledMatrix(rp, ci, va, class_policy="lossless")
    print(f"bicg poisson2d {nx} {layout}", flush=True)
    x64, i64 = A.bicgstab(b, tol=1e-10, precision="fp64", layout=layout, max_iter=20000)
    xm, im = A.bicgstab(b, tol=1e-10, precision="mixed", layout=layout, max_iter=20000)
    print(f"  fp64 {i64['iterations']} mixed {im['iterations']} (repl {im['replacements']})", flush=True)
    assert i64["status"] == 0 and im["status"] == 0, (i64, im)
    assert im["precision"] == 1
    assert abs(im["iterations"] - i64["iterations"]) <= 0.3 * i64["iterations"]
    assert orc.true_relres(rp, ci, va, x64.cpu().numpy(), b) <= 1.2e-10
    assert orc.true_relres(rp, ci, va, xm.cpu().numpy(), b) <= 1.2e-10
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
