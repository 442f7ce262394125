"""The blocked FP64 Cholesky solve for K above one workgroup's LDS (tdi_cholesky_solve with
K > 112; Solver::solve, solver.h:70-105): the joint identity-stage size K = 1338 (199 + 199 +
10 x 94) and a batch of three K = 300 systems against numpy's FP64 solve of the same FP32 matrices
(relative error <= 1e-6 at condition numbers ~4e3; observed ~3e-8, the FP32 output's rounding), a
K that is not a multiple of the 32-column panel, and a non-positive-definite matrix in a batch
(TDI_NOT_PD for it alone).
"""
import importlib
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
tdi = importlib.import_module("3di_amd.tdi")
abi = importlib.import_module("3di_amd.abi")


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _spd(rng, n, K):
    A = rng.normal(size=(n, K, K)) / np.sqrt(K)
    return np.einsum("nij,nkj->nik", A, A) + 1e-3 * np.eye(K)


@pytest.mark.parametrize("n,K", [(1, 1338), (3, 300), (2, 113)])
def test_large_k_solve(n, K):
    _say(f"case: large-K cholesky n={n} K={K}")
    rng = np.random.default_rng(K)
    H = _spd(rng, n, K).astype(np.float32)
    g = rng.normal(size=(n, K)).astype(np.float32)
    d, st = tdi.cholesky_solve(torch.from_numpy(H).cuda(), torch.from_numpy(g).cuda(), 0)
    torch.cuda.synchronize()
    d = d.cpu().numpy().astype(np.float64)
    assert (st.cpu().numpy() == 0).all()
    for i in range(n):
        ref = np.linalg.solve(H[i].astype(np.float64), -g[i].astype(np.float64))
        err = np.linalg.norm(d[i] - ref) / np.linalg.norm(ref)
        _say(f"  matrix {i}: relative error {err:.1e}, condition {np.linalg.cond(H[i].astype(np.float64)):.1e}")
        assert err < 1e-6


def test_large_k_not_pd_isolated():
    _say("case: large-K cholesky, one non-PD matrix in a batch of three")
    rng = np.random.default_rng(9)
    n, K = 3, 200
    H = _spd(rng, n, K)
    H[1, 150, 150] = -1.0          # breaks positive definiteness late in the factorisation
    H = H.astype(np.float32)
    g = rng.normal(size=(n, K)).astype(np.float32)
    d, st = tdi.cholesky_solve(torch.from_numpy(H).cuda(), torch.from_numpy(g).cuda(), 0)
    torch.cuda.synchronize()
    st = st.cpu().numpy()
    assert st[1] == abi.TDI_NOT_PD and st[0] == 0 and st[2] == 0
    d = d.cpu().numpy().astype(np.float64)
    for i in (0, 2):
        ref = np.linalg.solve(H[i].astype(np.float64), -g[i].astype(np.float64))
        assert np.linalg.norm(d[i] - ref) / np.linalg.norm(ref) < 1e-6
