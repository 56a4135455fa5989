"""Writes the golden fixtures in this folder from the oracle (the C++ restatement of ybus_sparse
and of ybus_powerflow's B' and Jacobian builders): for each case, the matrix, its factor
(ordering, triangles, permutations), a block of right-hand sides and their solutions (and the
transposed solutions for LU). Run from the repository root:

    python tests/golden/make_golden.py

The fixtures are data (numpy .npz, no pickles). tests/test_golden.py checks that the oracle still
gives them bit for bit, that SciPy's dense factorizations and solves agree with them to rounding,
and (on a GPU) that the device solves of the stored factors give their solutions bit for bit.
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

import oracle as O  # noqa: E402
from tests.fixtures import random_spd, random_square_cuda, right_hand_sides  # noqa: E402


def cases():
    """(name, kind, matrix, ordering)."""
    yield "spd_60", "cholesky", random_spd(60, 8, 3), O.AMD
    grid = O.make_grid(400, 20261210)
    bprime, _, _ = grid.dc()
    yield "dc_grid_400", "cholesky", bprime, O.AMD
    yield "unsym_50", "lu", random_square_cuda(50, 6, 4), O.AMD
    jac, _ = O.make_grid(120, 20261211).jacobian(20261211)
    yield "jacobian_120", "lu", jac, O.COLAMD


def write(name, kind, a, ordering):
    rhs = right_hand_sides(a.rows, 6, 77)
    out = dict(kind=np.array(kind), ordering=np.array(ordering), n=np.array(a.rows), a_row_major=np.array(a.row_major),
               a_ptr=a.ptr, a_idx=a.idx, a_val=a.data, rhs=rhs)
    if kind == "cholesky":
        f = O.Cholesky(a, ordering)
        out.update(perm=f.perm.astype(np.int64), l_ptr=f.l.ptr, l_idx=f.l.idx, l_val=f.l.data,
                   x=f.solve_many(rhs))
    else:
        f = O.Lu(a, ordering)
        out.update(row_perm=f.row_perm, col_perm=f.col_perm, l_ptr=f.l.ptr, l_idx=f.l.idx, l_val=f.l.data,
                   u_ptr=f.u.ptr, u_idx=f.u.idx, u_val=f.u.data, x=f.solve_many(rhs),
                   xt=f.solve_transpose_many(rhs))
    np.savez_compressed(os.path.join(HERE, f"{name}.npz"), **out)


if __name__ == "__main__":
    for case in cases():
        write(*case)
        print("wrote", case[0])
