"""Pins of the oracle preconditioners: each against a dense restatement of the
factorisation it describes, and the converged published rows."""
import json

import numpy as np
import pytest
import scipy.sparse as sp

from oracle import precond, picard, tracy
from tests.test_oracle_golden import tracy_problem


def _matrix(n=23, seed=0):
    rng = np.random.default_rng(seed)
    A = sp.random(n, n, density=0.25, random_state=seed, format="csr")
    A = A + A.T + sp.diags(rng.uniform(3, 5, n))
    A = A.tolil()
    A[3, 4] = 0.7          # non-symmetric entries
    A[10, 2] = -0.4
    return A.tocsr()


def test_block_jacobi_is_inverse_of_2x2_blocks():
    A = _matrix(23)                     # odd n: trailing 1x1 block
    M = precond.block_jacobi(A)
    D = np.zeros((23, 23))
    Ad = A.toarray()
    for s in range(0, 23, 2):
        e = min(s + 2, 23)
        D[s:e, s:e] = Ad[s:e, s:e] + 1e-12 * np.eye(e - s)
    x = np.random.default_rng(1).normal(size=23)
    np.testing.assert_allclose(M(x), np.linalg.solve(D, x), rtol=1e-13)


def test_ssor_is_forward_gauss_seidel():
    A = _matrix()
    M = precond.ssor(A)
    Ad = A.toarray()
    DL = np.tril(Ad)
    x = np.random.default_rng(2).normal(size=A.shape[0])
    np.testing.assert_allclose(M(x), np.linalg.solve(DL, x), rtol=1e-13)


def test_ilu_is_two_triangular_factors():
    A = _matrix()
    M = precond.ilu(A)
    Ad = A.toarray()
    d = np.diag(Ad)
    Lu = np.eye(len(d)) + np.tril(Ad, -1) / d[None, :]
    DU = np.triu(Ad, 1) + np.diag(d)
    x = np.random.default_rng(3).normal(size=A.shape[0])
    np.testing.assert_allclose(M(x), np.linalg.solve(DU, np.linalg.solve(Lu, x)), rtol=1e-12)


@pytest.mark.parametrize("solver,pc", [("cg", "ilu"), ("bicgstab", "ilu"), ("gmres", "ilu")])
def test_published_converged_ilu_rows(golden, solver, pc):
    """resultsRE2D/benchmark_results_cpu_Test1_WA.csv, mesh 25, *,ilu rows:
    53 Picard iterations and the converged metrics."""
    p, pb = tracy_problem("Test1", 25, solver, pc)
    r = picard.run(pb)
    assert int(np.sum(r["iterations"])) == 53
    e = tracy.final_errors(p, r["pressure_head"][-1], r["theta"][-1], r["times"][-1], 1)
    assert f"{e['l2_pressure']:.4e}" == "1.7613e-01"
    assert f"{e['l2_saturation']:.4e}" == "5.0311e-03"
