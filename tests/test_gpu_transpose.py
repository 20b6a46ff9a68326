"""Transposed products and solves (mf_spmv transpose, mf_solve_options.transpose).

The library keeps A^T as a tiled matrix of its own, built from the host copy of A's CSR
on the first transposed call (the reference's CSR transpose, utils.h:172) and dropped
when the values change.  Products are checked against SciPy's A.T @ x (binary64) and the
oracle's tile-mixed emulation on the transposed CSR; solves against the oracle's CG and
BiCGSTAB (Mille-feuille-cg.h:1188-1346, Mille-feuille-bicg.h:562-822) run on A^T, with the
true residual ||b - A^T x|| / ||b|| recomputed on the host."""
import numpy as np
import pytest
import scipy.sparse as sp

pytestmark = pytest.mark.gpu


def _rect(seed, m, n, density):
    rng = np.random.default_rng(seed)
    M = sp.random(m, n, density=density, random_state=rng, format="csr")
    M.data = rng.integers(-4, 5, M.data.size).astype(np.float64) + 0.25 * (rng.random(M.data.size) < 0.3)
    M.eliminate_zeros()
    M.sort_indices()
    return M.indptr.astype(np.int32), M.indices.astype(np.int32), M.data, M


@pytest.mark.parametrize("shape", [(300, 517), (517, 300), (1000, 1000)])
def test_spmv_transpose_matches_scipy(mf, orc, torch, shape):
    m, n = shape
    rp, ci, va, M = _rect(11, m, n, 0.02)
    A = mf.TiledMatrix(rp, ci, va, n_cols=n)
    x = np.random.default_rng(3).uniform(-1, 1, m)
    y = A.spmv(x, precision="fp64", transpose=True).cpu().numpy()
    ref = M.T @ x
    scale = abs(M).T @ np.abs(x) + 1e-300
    err = float(np.max(np.abs(y - ref) / scale))
    print(f"spmv A^T {shape}: max |y - A^T x| / (|A^T||x|) = {err:.2e}", flush=True)
    assert y.shape == (n,) and err <= 1e-14
    # tile-mixed product of A^T: the classes of A^T's own tiles
    MT = M.T.tocsr()
    MT.sort_indices()
    trp, tci, tva = MT.indptr.astype(np.int32), MT.indices.astype(np.int32), MT.data
    ecls = orc.entry_tile_classes(trp, tci, tva, n_cols=m)
    ym = A.spmv(x, precision="mixed", transpose=True).cpu().numpy()
    refm = orc.mixed_csr_spmv(trp, tci, tva, x, ecls)
    errm = float(np.max(np.abs(ym - refm) / scale))
    assert errm <= 1e-14, errm
    # the forward product is untouched by the transposed copy
    xf = np.random.default_rng(4).uniform(-1, 1, n)
    yf = A.spmv(xf, precision="fp64").cpu().numpy()
    assert np.max(np.abs(yf - M @ xf) / (abs(M) @ np.abs(xf) + 1e-300)) <= 1e-14
    A.close()


def test_transposed_copy_follows_value_updates(mf, torch):
    rp, ci, va, M = _rect(12, 400, 400, 0.03)
    A = mf.TiledMatrix(rp, ci, va)
    x = np.random.default_rng(5).uniform(-1, 1, 400)
    y1 = A.spmv(x, transpose=True).cpu().numpy()
    A.update_values(2.0 * va)
    y2 = A.spmv(x, transpose=True).cpu().numpy()
    # the tile-chunk SpMV sums through LDS atomics: equal up to the summation order
    scale = abs(M).T @ np.abs(x)
    assert np.max(np.abs(y2 - 2.0 * y1) / scale) <= 1e-14
    assert np.max(np.abs(y2 - 2.0 * (M.T @ x)) / scale) <= 1e-14
    A.close()


def test_bicgstab_transposed_solve(mf, orc, torch):
    rp, ci, va = mf.workloads.convdiff3d(12)
    n = rp.size - 1
    M = sp.csr_matrix((va, ci, rp), shape=(n, n))
    MT = M.T.tocsr()
    MT.sort_indices()
    trp, tci, tva = MT.indptr.astype(np.int32), MT.indices.astype(np.int32), MT.data
    b = np.random.default_rng(6).uniform(-1, 1, n)
    A = mf.TiledMatrix(rp, ci, va)
    ref = orc.bicgstab(trp, tci, tva, b, tol=1e-10)
    for prec in ("fp64", "mixed"):
        x, info = A.bicgstab(b, tol=1e-10, precision=prec, transpose=True)
        xh = x.cpu().numpy()
        rel = float(np.linalg.norm(b - MT @ xh) / np.linalg.norm(b))
        print(f"bicgstab A^T {prec}: {info['iterations']} iterations (oracle {ref['iterations']}), "
              f"relres {rel:.2e}", flush=True)
        assert info["status"] == 0
        assert rel <= 1.01e-10
        assert abs(info["iterations"] - ref["iterations"]) <= max(1, 0.02 * ref["iterations"])
    # the untransposed solve still solves A x = b
    x, info = A.bicgstab(b, tol=1e-10, precision="fp64")
    assert np.linalg.norm(b - M @ x.cpu().numpy()) <= 1.01e-10 * np.linalg.norm(b)
    x, info = A.bicgstab_multikernel(b, tol=1e-10, precision="fp64", transpose=True)
    assert info["status"] == 0
    assert np.linalg.norm(b - MT @ x.cpu().numpy()) <= 1.01e-10 * np.linalg.norm(b)
    A.close()


def test_cg_transposed_solve_on_spd(mf, orc, torch):
    """For a symmetric matrix A^T = A: the transposed solve takes the oracle's iteration
    count and reaches the same tolerance."""
    rp, ci, va = mf.workloads.poisson3d(16)
    b, _ = mf.workloads.rhs_from_solution(rp, ci, va, seed=7)
    A = mf.TiledMatrix(rp, ci, va)
    ref = orc.cg(rp, ci, va, b, tol=1e-10)
    for prec in ("fp64", "mixed"):
        x, info = A.cg(b, tol=1e-10, precision=prec, transpose=True)
        assert info["status"] == 0
        assert abs(info["iterations"] - ref["iterations"]) <= max(1, 0.02 * ref["iterations"])
        assert orc.true_relres(rp, ci, va, x.cpu().numpy(), b) <= 1.2e-10
    A.close()
