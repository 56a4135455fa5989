"""The golden fixtures (tests/golden/*.npz, written by tests/golden/make_golden.py from the oracle):
the oracle must still give them bit for bit, and SciPy's dense factorizations and solves, which
share no code with the oracle, must agree with them to rounding. The GPU half of the check is in
test_gpu_golden.py."""
import glob
import os

import numpy as np
import pytest
import scipy.linalg

import oracle as O

GOLDEN = sorted(glob.glob(os.path.join(os.path.dirname(__file__), "golden", "*.npz")))


def load(path):
    with np.load(path, allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def matrix(g):
    n = int(g["n"])
    return O.Matrix(n, n, bool(g["a_row_major"]), g["a_ptr"], g["a_idx"], g["a_val"])


def dense_csc(n, ptr, idx, val):
    d = np.zeros((n, n))
    for j in range(n):
        for p in range(ptr[j], ptr[j + 1]):
            d[idx[p], j] = val[p]
    return d


def test_there_are_fixtures_of_both_kinds():
    kinds = {str(load(p)["kind"]) for p in GOLDEN}
    assert kinds == {"cholesky", "lu"} and len(GOLDEN) >= 4


@pytest.mark.parametrize("path", GOLDEN, ids=[os.path.basename(p) for p in GOLDEN])
def test_the_oracle_still_gives_the_fixture(path):
    g = load(path)
    a = matrix(g)
    if str(g["kind"]) == "cholesky":
        f = O.Cholesky(a, int(g["ordering"]))
        assert np.array_equal(f.perm.astype(np.int64), g["perm"])
        assert f.l.data.tobytes() == g["l_val"].tobytes() and np.array_equal(f.l.idx, g["l_idx"])
    else:
        f = O.Lu(a, int(g["ordering"]))
        assert np.array_equal(f.row_perm, g["row_perm"]) and np.array_equal(f.col_perm, g["col_perm"])
        assert f.l.data.tobytes() == g["l_val"].tobytes() and f.u.data.tobytes() == g["u_val"].tobytes()
        assert f.solve_transpose_many(g["rhs"]).tobytes() == g["xt"].tobytes()
    assert f.solve_many(g["rhs"]).tobytes() == g["x"].tobytes()


@pytest.mark.parametrize("path", GOLDEN, ids=[os.path.basename(p) for p in GOLDEN])
def test_scipy_agrees_with_the_fixture(path):
    g = load(path)
    n = int(g["n"])
    a = matrix(g).to_dense()
    rhs, x = g["rhs"], g["x"]
    cond = np.linalg.cond(a)
    tol = 64 * np.finfo(float).eps * cond
    if str(g["kind"]) == "cholesky":
        perm = g["perm"]
        l_dense = np.zeros((n, n))
        for i in range(n):
            for p in range(g["l_ptr"][i], g["l_ptr"][i + 1]):
                l_dense[i, g["l_idx"][p]] = g["l_val"][p]
        want_l = scipy.linalg.cholesky(a[np.ix_(perm, perm)], lower=True)
        assert np.max(np.abs(l_dense - want_l)) <= tol * np.max(np.abs(want_l))
        want_x = scipy.linalg.solve(a, rhs.T, assume_a="pos").T
    else:
        l_dense = dense_csc(n, g["l_ptr"], g["l_idx"], g["l_val"])
        u_dense = dense_csc(n, g["u_ptr"], g["u_idx"], g["u_val"])
        paq = a[np.ix_(g["row_perm"], g["col_perm"])]
        assert np.max(np.abs(l_dense @ u_dense - paq)) <= 1e-12 * np.max(np.abs(paq))
        want_x = scipy.linalg.solve(a, rhs.T).T
        want_xt = scipy.linalg.solve(a.T, rhs.T).T
        assert np.max(np.abs(g["xt"] - want_xt)) <= tol * np.max(np.abs(want_xt))
    assert np.max(np.abs(x - want_x)) <= tol * np.max(np.abs(want_x))
