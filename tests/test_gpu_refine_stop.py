"""The FP32-refined solves' floor stop rule (STOP_AT_FLOOR): a column also stops as soon as its
ratio is at most 2 u || |A| |x| ||inf / ||b||inf, the rounding floor of an FP64 residual.
Checked against refine_normwise's rule on the same blocks. The reported steps are the best
iterate's corrections (refine_normwise's meaning), so a correction saved after the best one does
not show there; what the tests check is that the floor rule's results are the default rule's
iterates: where both report the same step the bits agree, where the floor rule reports an earlier
one its ratio had reached twice the returned x's own floor (computed on the CPU); and every
result meets the default's bar (1e-12 or twice the CPU solution's residual) or that floor. The time it saves is measured by the
benchmark."""
import numpy as np
import pytest

import oracle as O
import ybus_amd as Y
from tests.fixtures import complex_lu_case, complex_right_hand_sides, right_hand_sides
from tests.gpu_common import gpu, path  # noqa: F401

pytestmark = pytest.mark.gpu

U = 2.0**-53


def abs_product_norm(a, x):
    """|| |A| |x| ||inf for each row of x (A an oracle matrix)."""
    m = O.Matrix(a.rows, a.cols, a.row_major, a.ptr, a.idx, np.abs(a.data))
    return np.array([np.max(m.mul_vec(np.abs(v))) for v in x])


def check_floor_rule(steps_d, ratio_d, steps_f, ratio_f, x_f, x_d, rhs, floors_cpu, axn):
    """The floor rule's results against the default's (see the module); returns the columns
    where it reports an earlier iterate."""
    bn = np.max(np.abs(rhs), axis=1)
    stop_at = 2.0 * U * axn / bn
    assert np.all(steps_f <= steps_d)
    # the default's bar (1e-12 or twice the CPU solution's residual), or the floor the rule stops
    # at: u || |A| |x| || bounds the rounding of A x, and the CPU's residual is often below it
    assert np.all(ratio_f <= np.maximum(np.maximum(1e-12, 2.0 * floors_cpu), stop_at * (1 + 1e-12)))
    earlier = np.flatnonzero(steps_f < steps_d)
    for k in earlier:
        # the floor rule stopped this column: its ratio had reached twice the rounding floor of
        # the x it returned (|A| |x| summed in another order on the CPU: a relative 1e-12 slack)
        assert ratio_f[k] <= stop_at[k] * (1 + 1e-12), (k, ratio_f[k], stop_at[k])
        # (u || |A| |x| || bounds the rounding of A x, so the default's further corrections may
        # still bring the ratio below it, at the noise level: the floor rule gives that up)
    same = np.flatnonzero(steps_f == steps_d)
    for k in same:
        assert x_f[k].tobytes() == x_d[k].tobytes()
    return len(earlier)


@pytest.mark.parametrize("buses", [2500])
def test_floor_rule_on_a_jacobian(gpu, path, buses):
    g = O.make_grid(buses, 77)
    jac, _ = g.jacobian(77)
    lu = O.Lu(jac, O.COLAMD)
    rhs = O.normals(3, 9, 64 * jac.rows).reshape(64, jac.rows)
    factor = Y.GpuFactor.from_lu(gpu, lu)
    factor.set_matrix(jac)
    x_d, steps_d, ratio_d = factor.solve_many_refined(rhs, 1e-12, 10)
    factor.set_refine_stop(Y.STOP_AT_FLOOR)
    x_f, steps_f, ratio_f = factor.solve_many_refined(rhs, 1e-12, 10)
    cpu = lu.solve_many(rhs, threads=4)
    floors = np.array([jac.relative_residual(cpu[k], rhs[k]) for k in range(len(rhs))])
    for k in range(len(rhs)):
        assert ratio_f[k] == jac.relative_residual(x_f[k], rhs[k])
    check_floor_rule(steps_d, ratio_d, steps_f, ratio_f, x_f, x_d, rhs, floors, abs_product_norm(jac, x_f))
    # back to the default rule: the default's bits again
    factor.set_refine_stop(Y.STOP_NORMWISE)
    x_again, _, _ = factor.solve_many_refined(rhs, 1e-12, 10)
    assert x_again.tobytes() == x_d.tobytes()


def test_floor_rule_in_a_batch_and_its_graph(gpu):
    jac, _ = O.make_grid(1500, 77).jacobian(77)
    bp, _, _ = O.make_grid(3000, 91).dc()
    mats = [(jac, O.Lu(jac, O.COLAMD)), (bp, O.Cholesky(bp))]
    batch = Y.GpuBatch.new(gpu)
    for m, c in mats:
        f = Y.GpuFactor.from_lu(gpu, c) if isinstance(c, O.Lu) else Y.GpuFactor.from_cholesky(gpu, c)
        f.set_matrix(m)
        batch.add(f)
    rhs = [O.normals(5, 1, 9 * jac.rows).reshape(9, jac.rows), O.injections(5, 2, 3000, 6)]
    batch.stage(rhs)
    batch.solve_staged(Y.FP32_REFINED, 1e-12, 10)
    x_d = [x.copy() for x in batch.solutions()]
    ref_d = batch.refinement()
    batch.set_refine_stop(Y.STOP_AT_FLOOR)
    for replay in range(2):
        batch.solve_staged(Y.FP32_REFINED, 1e-12, 10)
        x_f = batch.solutions()
        for (m, c), b, xf, xd, (sd, rd), (sf, rf) in zip(mats, rhs, x_f, x_d, ref_d, batch.refinement()):
            floors = np.array([m.relative_residual(v, w) for v, w in zip(c.solve_many(b), b)])
            check_floor_rule(sd, rd, sf, rf, xf, xd, b, floors, abs_product_norm(m, xf))
    batch.set_refine_stop(Y.STOP_NORMWISE)
    batch.solve_staged(Y.FP32_REFINED, 1e-12, 10)
    for x, want in zip(batch.solutions(), x_d):
        assert x.tobytes() == want.tobytes()


def test_floor_rule_on_a_complex_factor(gpu):
    import scipy.sparse as sp

    lu, a = complex_lu_case(250, 4, 3)
    factor = Y.GpuFactor.from_lu_complex(gpu, lu)
    factor.set_matrix(sp.csr_matrix(a))
    b = complex_right_hand_sides(lu.n, 24, 8)
    x_d, steps_d, ratio_d = factor.solve_many_refined(b, 1e-30, 10)
    factor.set_refine_stop(Y.STOP_AT_FLOOR)
    x_f, steps_f, ratio_f = factor.solve_many_refined(b, 1e-30, 10)
    assert np.all(steps_f <= steps_d)
    same = steps_f == steps_d
    assert x_f[same].tobytes() == x_d[same].tobytes()
    axn = np.max(np.abs(a) @ np.abs(x_f.T), axis=0)
    bn = np.max(np.abs(b), axis=1)
    earlier = steps_f < steps_d
    stop_at = 2.0 * U * axn / bn
    assert np.all(ratio_f[earlier] <= stop_at[earlier] * (1 + 1e-12))
    assert np.all(ratio_f <= np.maximum(2.0 * ratio_d, stop_at * (1 + 1e-12)))


def test_unknown_rules_are_refused(gpu):
    factor = Y.GpuFactor.from_cholesky(gpu, O.Cholesky(O.make_grid(100, 3).dc()[0]))
    with pytest.raises(Y.InvalidArgument):
        factor.set_refine_stop(2)
    with pytest.raises(Y.InvalidArgument):
        Y.GpuBatch.new(gpu).set_refine_stop(-1)
