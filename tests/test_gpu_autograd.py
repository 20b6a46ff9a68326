"""Differentiable solve: adjoint gradients vs a dense FP64 reference and central
finite differences (variable-coefficient 3D diffusion, k = exp(GRF))."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch


def _problem(mf, nx, seed=4):
    rp, ci, va, k = mf.workloads.varcoef_diffusion3d(nx, seed=seed)
    rng = np.random.default_rng(seed)
    n = rp.size - 1
    b = rng.uniform(-1, 1, n)
    xt = rng.uniform(-1, 1, n)
    return rp, ci, va, b, xt


def _adjoint(mf, torch, rp, ci, va, b, xt, **kw):
    from mille_feuille_amd.autograd import SparseOperator, sparse_solve
    op = SparseOperator(rp, ci, va, **kw)
    vals = torch.tensor(va, device="cuda", requires_grad=True)
    bt = torch.tensor(b, device="cuda", requires_grad=True)
    x = sparse_solve(vals, bt, op)
    loss = 0.5 * torch.sum((x - torch.as_tensor(xt, device="cuda")) ** 2)
    loss.backward()
    out = (x.detach().cpu().numpy(), vals.grad.cpu().numpy(), bt.grad.cpu().numpy(), float(loss), op)
    return out


@pytest.mark.parametrize("spd,precision", [(True, "mixed"), (True, "fp64"), (False, "mixed")])
def test_adjoint_matches_dense_reference_16cubed(mf, torch, spd, precision):
    import scipy.sparse as sp
    rp, ci, va, b, xt = _problem(mf, 16)
    n = rp.size - 1
    x, gv, gb, loss, op = _adjoint(mf, torch, rp, ci, va, b, xt, spd=spd, precision=precision)
    A = sp.csr_matrix((va, ci, rp), shape=(n, n)).toarray()
    xr = np.linalg.solve(A, b)
    lam = np.linalg.solve(A.T, xr - xt)
    rows = np.repeat(np.arange(n), np.diff(rp))
    gv_ref = -lam[rows] * xr[ci]
    assert np.linalg.norm(x - xr) <= 1e-8 * np.linalg.norm(xr)
    assert np.linalg.norm(gb - lam) <= 1e-6 * np.linalg.norm(lam)
    assert np.linalg.norm(gv - gv_ref) <= 1e-6 * np.linalg.norm(gv_ref)
    op.close()


def test_adjoint_matches_finite_differences(mf, torch):
    rp, ci, va, b, xt = _problem(mf, 16)
    x, gv, gb, loss, op = _adjoint(mf, torch, rp, ci, va, b, xt, spd=True, precision="fp64", tol=1e-13)
    rng = np.random.default_rng(0)
    picks = rng.choice(ci.size, 64, replace=False)
    h = 1e-6

    # a symmetric pattern entry perturbed alone makes A nonsymmetric, so the FD uses
    # the general adjoint convention checked above: dL/dA_ij = -lambda_i x_j
    op2_kw = dict(spd=False, precision="fp64", tol=1e-13)
    from mille_feuille_amd.autograd import SparseOperator
    op2 = SparseOperator(rp, ci, va, **op2_kw)

    def L2(v):
        op2.set_values(v)
        xx = op2.solve(b).cpu().numpy()
        return 0.5 * np.sum((xx - xt) ** 2)

    fd = []
    for k in picks:
        vp, vm = va.copy(), va.copy()
        vp[k] += h
        vm[k] -= h
        fd.append((L2(vp) - L2(vm)) / (2 * h))
    fd = np.array(fd)
    scale = np.abs(gv[picks]).max()
    err = np.max(np.abs(fd - gv[picks])) / scale
    print(f"finite-difference max relative error {err:.3e}")
    assert err <= 1e-4, err
    op.close()
    op2.close()


def test_adjoint_64cubed_mixed_vs_fp64(mf, torch):
    rp, ci, va, b, xt = _problem(mf, 64)
    x1, gv1, gb1, l1, op1 = _adjoint(mf, torch, rp, ci, va, b, xt, spd=True, precision="mixed")
    x2, gv2, gb2, l2, op2 = _adjoint(mf, torch, rp, ci, va, b, xt, spd=True, precision="fp64")
    assert op1.last_info["forward"]["relres_true"] <= 1.2e-10
    assert op1.last_info["adjoint"]["relres_true"] <= 1.2e-10
    assert np.linalg.norm(gv1 - gv2) <= 1e-6 * np.linalg.norm(gv2)
    assert np.linalg.norm(gb1 - gb2) <= 1e-6 * np.linalg.norm(gb2)
    op1.close()
    op2.close()


def test_two_forwards_then_backward_use_their_own_values(mf, torch):
    """Two solves with different values on one shared operator before either backward:
    each gradient must be that of its own A(vals) (dense FP64 reference)."""
    import scipy.sparse as sp
    from mille_feuille_amd.autograd import SparseOperator, sparse_solve
    rp, ci, va, b, xt = _problem(mf, 8)
    n = rp.size - 1
    op = SparseOperator(rp, ci, va, spd=True, precision="fp64")
    va2 = va * 1.5
    v1 = torch.tensor(va, device="cuda", requires_grad=True)
    v2 = torch.tensor(va2, device="cuda", requires_grad=True)
    bt = torch.as_tensor(b, device="cuda")
    xtt = torch.as_tensor(xt, device="cuda")
    x1 = sparse_solve(v1, bt, op)
    x2 = sparse_solve(v2, bt, op)   # overwrites the operator's values
    (0.5 * torch.sum((x1 - xtt) ** 2)).backward()
    (0.5 * torch.sum((x2 - xtt) ** 2)).backward()
    rows = np.repeat(np.arange(n), np.diff(rp))
    for vals, v in ((va, v1), (va2, v2)):
        A = sp.csr_matrix((vals, ci, rp), shape=(n, n)).toarray()
        xr = np.linalg.solve(A, b)
        lam = np.linalg.solve(A.T, xr - xt)
        ref = -lam[rows] * xr[ci]
        got = v.grad.cpu().numpy()
        assert np.linalg.norm(got - ref) <= 1e-6 * np.linalg.norm(ref)
    op.close()
