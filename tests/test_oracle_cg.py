"""Oracle CG against Eigen's ConjugateGradient with the reference's
BlockJacobiPreconditioner3D (compiled probe)."""
import numpy as np

from oracle import oracle


def test_cg_matches_eigen(golden):
    sysd = golden["cg"]["system"]
    n = sysd["n"]
    A = np.asarray(sysd["A"]).reshape(n, n)
    for run in golden["cg"]["runs"]:
        x, it, err = oracle.cg_dense(A, sysd["b"], sysd["x0"], sysd["maxIter"], run["tol"])
        assert it == run["iterations"]
        np.testing.assert_allclose(x, run["x"], rtol=1e-9, atol=1e-12)
        # the residual estimate carries the rounding of the dot products
        assert abs(err - run["error"]) <= 1e-3 * run["error"]


def test_cg_solves_system(golden):
    sysd = golden["cg"]["system"]
    n = sysd["n"]
    A = np.asarray(sysd["A"]).reshape(n, n)
    x, it, err = oracle.cg_dense(A, sysd["b"], sysd["x0"], 200, 1e-12)
    np.testing.assert_allclose(x, np.linalg.solve(A, sysd["b"]), rtol=1e-8, atol=1e-10)
