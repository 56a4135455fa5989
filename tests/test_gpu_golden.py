"""The device solves of the golden fixtures' stored factors give the fixtures' solutions bit for
bit: the factor arrays are uploaded as stored, so the oracle takes no part in this check."""
import os

import pytest

import ybus_amd as Y
from tests.gpu_common import gpu, path  # noqa: F401
from tests.test_golden import GOLDEN, load

pytestmark = [pytest.mark.gpu, pytest.mark.usefixtures("path")]


@pytest.mark.parametrize("fixture", GOLDEN, ids=[os.path.basename(p) for p in GOLDEN])
def test_device_solves_give_the_fixture(gpu, fixture):
    g = load(fixture)
    n = int(g["n"])
    if str(g["kind"]) == "cholesky":
        factor = Y.GpuFactor.from_cholesky_arrays(gpu, n, g["l_ptr"], g["l_idx"], g["l_val"], g["perm"])
    else:
        factor = Y.GpuFactor.from_lu_arrays(gpu, n, g["l_ptr"], g["l_idx"], g["l_val"], g["u_ptr"], g["u_idx"],
                                            g["u_val"], g["row_perm"], g["col_perm"])
        assert factor.solve_transpose_many(g["rhs"]).tobytes() == g["xt"].tobytes()
    assert factor.solve_many(g["rhs"]).tobytes() == g["x"].tobytes()
