"""The C ABI library without a device: it loads, exports every entry point include/ybus_hip.h
declares, reports a missing device as the reference reports no CUDA (Gpu::new's Ok(None),
ybus_cuda/src/lib.rs:179-207), and its host-side analysis (level sets, supernodes) matches brute
force and the oracle exactly."""
import os
import re

import numpy as np
import pytest

import oracle as O
import ybus_amd as Y
from tests.fixtures import grid_laplacian, random_spd, random_square_cuda, two_by_two_blocks

HEADER = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "include", "ybus_hip.h")


def declared():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(ybus_hip_[a-z_0-9]+)\s*\(", text)))


def test_every_declared_entry_point_is_exported():
    names = declared()
    assert len(names) >= 38
    lib = Y.library()
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_no_device_is_none_and_errors_carry_messages():
    if os.path.exists("/dev/kfd"):
        pytest.skip("a device is present")
    assert Y.Gpu.new() is None
    with pytest.raises(Y.InvalidArgument) as e:
        Y.analyze_cholesky(2, np.array([0, 1, 2]), np.array([0, 0]))
    assert "diagonal" in str(e.value)


def brute_levels(n, deps_of):
    """Longest dependency path, by repeated relaxation (independent of the product's pass)."""
    level = [0] * n
    changed = True
    while changed:
        changed = False
        for i in range(n):
            want = max([level[j] + 1 for j in deps_of(i)], default=0)
            if want != level[i]:
                level[i] = want
                changed = True
    return level


@pytest.mark.parametrize("make", [lambda: random_spd(120, 4, 3), lambda: grid_laplacian(8),
                                  lambda: O.make_grid(300, 5).dc()[0]])
def test_cholesky_levels_and_supernodes(make):
    a = make()
    f = O.Cholesky(a)
    n = f.n
    got = Y.analyze_cholesky(n, f.l.ptr, f.l.idx)
    l = f.l
    lower = [[int(j) for j in l.idx[l.ptr[i]:l.ptr[i + 1]] if j < i] for i in range(n)]
    upper = [[] for _ in range(n)]
    for i in range(n):
        for j in lower[i]:
            upper[j].append(i)
    assert list(got["forward"]) == brute_levels(n, lambda i: lower[i])
    assert list(got["backward"]) == brute_levels(n, lambda i: upper[i])
    nodes, _ = f.supernodes()
    assert list(got["supernodes"]) == [s for s, _, _, _ in nodes] + [n]


@pytest.mark.parametrize("make", [lambda: random_square_cuda(90, 4, 3), lambda: two_by_two_blocks(20),
                                  lambda: O.make_grid(200, 9).jacobian(9)[0]])
def test_lu_levels(make):
    a = make()
    for ordering in [O.AMD, O.COLAMD]:
        f = O.Lu(a, ordering)
        n = f.n
        got = Y.analyze_lu(n, f.l.ptr, f.l.idx, f.u.ptr, f.u.idx)
        L, U = f.l, f.u
        lcol = [[int(i) for i in L.idx[L.ptr[j] + 1:L.ptr[j + 1]]] for j in range(n)]
        ucol = [[int(i) for i in U.idx[U.ptr[j]:U.ptr[j + 1] - 1]] for j in range(n)]
        lrow = [[] for _ in range(n)]
        urow = [[] for _ in range(n)]
        for j in range(n):
            for i in lcol[j]:
                lrow[i].append(j)
            for i in ucol[j]:
                urow[i].append(j)
        assert list(got["forward"]) == brute_levels(n, lambda i: lrow[i])
        assert list(got["backward"]) == brute_levels(n, lambda i: urow[i])
        assert list(got["transpose_forward"]) == brute_levels(n, lambda j: ucol[j])
        assert list(got["transpose_backward"]) == brute_levels(n, lambda j: lcol[j])
