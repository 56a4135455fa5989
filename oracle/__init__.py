"""CPU oracle for the GPU factor-solve path: test infrastructure only.

A Python face (ctypes + numpy) over ``liboracle.so``, the C++ restatement of ybus_sparse's
compressed matrices, orderings (AMD, COLAMD, RCM), Cholesky and LU factorizations, CPU solves,
residual and normwise refinement, supernodes, and of ybus_powerflow's B' and Jacobian
builders, plus the seeded synthetic-grid generator. Only ``tests/``, ``__graft_entry__.smoke()``
and ``bench.py``'s CPU-baseline leg may import this package; the product never does.

Parity status per function is recorded in the oracle's own tests (``tests/test_oracle_*.py``):
every function here is pinned by a golden literal from the reference's tests, a brute-force or
closed-form check, or an independent restatement.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
from dataclasses import dataclass

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")

NATURAL, RCM, AMD, COLAMD = 0, 1, 2, 3
# mistakes the tests can plant in AMD and COLAMD (oracle.hpp, Mutation)
FIFO_LISTS, LAST_REPRESENTATIVE, NO_AGGRESSIVE_ABSORPTION = 1, 2, 4

ERRORS = {
    1: "NotSquare", 2: "NotSymmetric", 3: "NotPositiveDefinite", 4: "SingularMatrix",
    5: "PatternMismatch", 6: "ShapeMismatch", 7: "OutOfBounds", 8: "InvalidPermutation",
    9: "UnsupportedOrdering", 10: "InvalidArgument", 99: "Internal",
}


class SparseError(Exception):
    """A ybus_sparse SparseError, by variant name (``kind``) and pivot where it has one."""

    def __init__(self, kind: str, message: str, pivot: int | None = None):
        super().__init__(f"{kind}: {message}" + (f" (pivot {pivot})" if pivot is not None else ""))
        self.kind = kind
        self.pivot = pivot


def build() -> str:
    """Compiles liboracle.so from its sources (make), returning its path."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _LIB_PATH


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        _lib = C.CDLL(_LIB_PATH)
        _declare(_lib)
    return _lib


P = C.c_void_p
U64 = C.c_uint64
DBL = C.c_double


def _declare(L):
    u64p = np.ctypeslib.ndpointer(np.uint64, flags="C")
    i64p = np.ctypeslib.ndpointer(np.int64, flags="C")
    f64p = np.ctypeslib.ndpointer(np.float64, flags="C")
    u8p = np.ctypeslib.ndpointer(np.uint8, flags="C")
    mat = [U64, U64, C.c_int, u64p, u64p, f64p]
    sig = {
        "yo_last_error": (C.c_size_t, [C.c_char_p, C.c_size_t]),
        "yo_free": (None, [P]),
        "yo_from_triplets": (C.c_int, [U64, U64, C.c_int, U64, u64p, u64p, f64p, C.POINTER(P)]),
        "yo_matrix_sizes": (None, [P, u64p]),
        "yo_matrix_copy": (None, [P, u64p, u64p, f64p]),
        "yo_reoriented": (C.c_int, mat + [C.c_int, C.POINTER(P)]),
        "yo_select": (C.c_int, mat + [U64, u64p, U64, u64p, C.POINTER(P)]),
        "yo_mul_vec": (C.c_int, mat + [f64p, f64p]),
        "yo_relative_residual": (C.c_int, mat + [f64p, f64p, C.POINTER(DBL)]),
        "yo_ordering": (C.c_int, mat + [C.c_int, u64p]),
        "yo_ordering_mutated": (C.c_int, mat + [C.c_int, C.c_uint, u64p]),
        "yo_cholesky": (C.c_int, mat + [C.c_int, P, C.c_int, C.POINTER(P), C.POINTER(U64)]),
        "yo_chol_sizes": (None, [P, u64p]),
        "yo_chol_copy_symbolic": (None, [P, u64p, i64p, u64p, u64p, u64p, u64p]),
        "yo_chol_copy_factor": (None, [P, u64p, u64p, f64p, u64p, u64p, f64p]),
        "yo_chol_solve_many": (C.c_int, [P, U64, f64p, f64p, C.c_int]),
        "yo_chol_refine": (C.c_int, [P] + mat + [f64p, U64, DBL, f64p, C.POINTER(U64), C.POINTER(DBL)]),
        "yo_chol_supernodes": (C.c_int, [P, C.POINTER(P)]),
        "yo_super_sizes": (None, [P, u64p]),
        "yo_super_copy": (None, [P, u64p, u64p, i64p, u64p, u64p]),
        "yo_super_front_tree": (None, [P, u64p]),
        "yo_lu": (C.c_int, mat + [C.c_int, DBL, P, C.POINTER(P), C.POINTER(U64)]),
        "yo_lu_sizes": (None, [P, u64p]),
        "yo_lu_copy": (None, [P, u64p, u64p, f64p, u64p, u64p, f64p, u64p, u64p]),
        "yo_lu_solve_many": (C.c_int, [P, U64, f64p, f64p, C.c_int]),
        "yo_lu_solve_transpose_many": (C.c_int, [P, U64, f64p, f64p, C.c_int]),
        "yo_lu_refine": (C.c_int, [P] + mat + [f64p, U64, DBL, f64p, C.POINTER(U64), C.POINTER(DBL)]),
        "yo_lu_refactor": (C.c_int, [P] + mat + [C.POINTER(C.c_int)]),
        "yo_lu_refactor_work": (U64, [P]),
        "yo_grid": (C.c_int, [U64, U64, C.POINTER(P)]),
        "yo_grid_sizes": (None, [P, u64p]),
        "yo_grid_copy": (None, [P, f64p, f64p, u64p, u64p, f64p]),
        "yo_grid_bridges": (None, [P, u8p]),
        "yo_dc": (C.c_int, [P, P, C.POINTER(P), C.POINTER(P), u64p]),
        "yo_jacobian": (C.c_int, [P, U64, DBL, U64, C.POINTER(P), C.POINTER(P)]),
        "yo_ac_sizes": (None, [P, u64p]),
        "yo_ac_copy": (None, [P, u64p, u64p, f64p, f64p, f64p, f64p]),
        "yo_uniforms": (None, [U64, U64, U64, f64p]),
        "yo_normals": (None, [U64, U64, U64, f64p]),
        "yo_integers": (None, [U64, U64, U64, u64p]),
        "yo_injections": (None, [U64, U64, U64, U64, f64p]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args


def _check(status: int, pivot: int | None = None):
    if status != 0:
        buf = C.create_string_buffer(512)
        lib().yo_last_error(buf, 512)
        raise SparseError(ERRORS.get(status, str(status)), buf.value.decode(),
                          None if pivot in (None, 2**64 - 1) else int(pivot))


def _u64(a):
    return np.ascontiguousarray(a, dtype=np.uint64)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


@dataclass
class Matrix:
    """CompressedMatrix<f64>: CSR when ``row_major`` else CSC."""

    rows: int
    cols: int
    row_major: bool
    ptr: np.ndarray
    idx: np.ndarray
    data: np.ndarray

    @property
    def nnz(self) -> int:
        return int(self.idx.shape[0])

    @property
    def shape(self):
        return (self.rows, self.cols)

    def args(self):
        return [self.rows, self.cols, int(self.row_major), _u64(self.ptr), _u64(self.idx), _f64(self.data)]

    @staticmethod
    def from_triplets(rows, cols, triplets, row_major=True) -> "Matrix":
        t = list(triplets)
        r = _u64([e[0] for e in t])
        c = _u64([e[1] for e in t])
        v = _f64([e[2] for e in t])
        return Matrix.from_coo(rows, cols, r, c, v, row_major)

    @staticmethod
    def from_coo(rows, cols, r, c, v, row_major=True) -> "Matrix":
        h = P()
        _check(lib().yo_from_triplets(rows, cols, int(row_major), len(r), _u64(r), _u64(c), _f64(v), C.byref(h)))
        return _take_matrix(h)

    def to_csr(self) -> "Matrix":
        return self._reoriented(True)

    def to_csc(self) -> "Matrix":
        return self._reoriented(False)

    def _reoriented(self, row_major):
        if self.row_major == row_major:
            return self
        h = P()
        _check(lib().yo_reoriented(*self.args(), int(row_major), C.byref(h)))
        return _take_matrix(h)

    def select(self, rows, cols) -> "Matrix":
        rows, cols = _u64(rows), _u64(cols)
        h = P()
        _check(lib().yo_select(*self.args(), len(rows), rows, len(cols), cols, C.byref(h)))
        return _take_matrix(h)

    def permute_symmetric(self, perm) -> "Matrix":
        return self.select(perm, perm)

    def mul_vec(self, x) -> np.ndarray:
        y = np.zeros(self.rows)
        _check(lib().yo_mul_vec(*self.args(), _f64(x), y))
        return y

    def relative_residual(self, x, b) -> float:
        r = DBL()
        _check(lib().yo_relative_residual(*self.args(), _f64(x), _f64(b), C.byref(r)))
        return r.value

    def ordering(self, kind: int, mutations: int = 0) -> np.ndarray:
        """The ordering's permutation (ordering.rs:46-63). `mutations` plants mistakes in AMD or
        COLAMD (FIFO_LISTS, LAST_REPRESENTATIVE, NO_AGGRESSIVE_ABSORPTION), for tests that show
        their literal permutations catch them."""
        perm = np.zeros(self.cols if kind == COLAMD else self.rows, dtype=np.uint64)
        if mutations:
            _check(lib().yo_ordering_mutated(*self.args(), kind, mutations, perm))
        else:
            _check(lib().yo_ordering(*self.args(), kind, perm))
        return perm.astype(np.int64)

    def to_dense(self) -> np.ndarray:
        d = np.zeros((self.rows, self.cols))
        for major in range(self.rows if self.row_major else self.cols):
            for p in range(int(self.ptr[major]), int(self.ptr[major + 1])):
                if self.row_major:
                    d[major, int(self.idx[p])] += self.data[p]
                else:
                    d[int(self.idx[p]), major] += self.data[p]
        return d


def _take_matrix(h) -> Matrix:
    L = lib()
    s = np.zeros(4, dtype=np.uint64)
    L.yo_matrix_sizes(h, s)
    rows, cols, rm, nnz = (int(v) for v in s)
    major = rows if rm else cols
    ptr = np.zeros(major + 1, dtype=np.uint64)
    idx = np.zeros(nnz, dtype=np.uint64)
    data = np.zeros(nnz)
    L.yo_matrix_copy(h, ptr, idx, data)
    L.yo_free(h)
    return Matrix(rows, cols, bool(rm), ptr.astype(np.int64), idx.astype(np.int64), data)


class Cholesky:
    """SymbolicCholesky + Cholesky<f64> (cholesky.rs:27-39, 321-328)."""

    def __init__(self, a: Matrix, ordering: int = AMD, perm=None, symbolic_only=False):
        h = P()
        piv = U64(2**64 - 1)
        permbuf = _u64(perm) if perm is not None else None
        st = lib().yo_cholesky(*a.args(), ordering, permbuf.ctypes.data if permbuf is not None else None,
                               int(symbolic_only), C.byref(h), C.byref(piv))
        _check(st, piv.value)
        self._h = h
        self.a = a
        L = lib()
        s = np.zeros(2, dtype=np.uint64)
        L.yo_chol_sizes(h, s)
        n, nnz = int(s[0]), int(s[1])
        self.n = n
        self.perm = np.zeros(n, dtype=np.uint64)
        self.parent = np.zeros(n, dtype=np.int64)
        self.postorder = np.zeros(n, dtype=np.uint64)
        self.pattern_ptr = np.zeros(n + 1, dtype=np.uint64)
        self.pattern_idx = np.zeros(nnz, dtype=np.uint64)
        self.column_counts = np.zeros(n, dtype=np.uint64)
        L.yo_chol_copy_symbolic(h, self.perm, self.parent, self.postorder, self.pattern_ptr, self.pattern_idx,
                                self.column_counts)
        for name in ("perm", "postorder", "pattern_ptr", "pattern_idx", "column_counts"):
            setattr(self, name, getattr(self, name).astype(np.int64))
        self.predicted_entries = nnz
        if not symbolic_only:
            lp, li, lv = np.zeros(n + 1, np.uint64), np.zeros(nnz, np.uint64), np.zeros(nnz)
            tp, ti, tv = np.zeros(n + 1, np.uint64), np.zeros(nnz, np.uint64), np.zeros(nnz)
            L.yo_chol_copy_factor(h, lp, li, lv, tp, ti, tv)
            self.l = Matrix(n, n, True, lp.astype(np.int64), li.astype(np.int64), lv)
            self.l_adjoint = Matrix(n, n, True, tp.astype(np.int64), ti.astype(np.int64), tv)

    def __del__(self):
        if getattr(self, "_h", None):
            lib().yo_free(self._h)
            self._h = None

    def elimination_tree(self):
        return [None if p < 0 else int(p) for p in self.parent]

    def solve_many(self, rhs: np.ndarray, threads: int = 1) -> np.ndarray:
        """rhs: (count, n) array; returns (count, n) solutions (factorization.rs:52-92)."""
        rhs = _f64(np.atleast_2d(rhs))
        out = np.zeros_like(rhs)
        _check(lib().yo_chol_solve_many(self._h, rhs.shape[0], rhs, out, threads))
        return out

    def solve(self, b) -> np.ndarray:
        return self.solve_many(np.asarray(b, dtype=np.float64)[None, :])[0]

    def solve_refined_normwise(self, a: Matrix, b, max_steps: int, tol: float):
        x = np.zeros(self.n)
        steps, ratio = U64(), DBL()
        _check(lib().yo_chol_refine(self._h, *a.args(), _f64(b), max_steps, tol, x, C.byref(steps), C.byref(ratio)))
        return x, int(steps.value), ratio.value

    def supernodes(self):
        """Fundamental supernodes (multifrontal.rs:94-138): list of (start, end, below, parent)."""
        L = lib()
        h = P()
        _check(L.yo_chol_supernodes(self._h, C.byref(h)))
        s = np.zeros(2, np.uint64)
        L.yo_super_sizes(h, s)
        k, tot = int(s[0]), int(s[1])
        start, end = np.zeros(k, np.uint64), np.zeros(k, np.uint64)
        parent = np.zeros(k, np.int64)
        bptr, below = np.zeros(k + 1, np.uint64), np.zeros(tot, np.uint64)
        L.yo_super_copy(h, start, end, parent, bptr, below)
        ft = np.zeros(7, np.uint64)
        L.yo_super_front_tree(h, ft)
        L.yo_free(h)
        nodes = [(int(start[i]), int(end[i]), [int(v) for v in below[int(bptr[i]):int(bptr[i + 1])]],
                  None if parent[i] < 0 else int(parent[i])) for i in range(k)]
        tree = dict(zip(["fronts", "roots", "height", "work", "critical_work", "solve_work",
                         "solve_critical_work"], (int(v) for v in ft)))
        return nodes, tree


class Lu:
    """Lu<f64> (lu.rs:101-132): L unit lower and U upper, both CSC; P A Q = L U."""

    def __init__(self, a_csc: Matrix, ordering: int = AMD, pivot_threshold: float = 0.1, col_perm=None):
        if a_csc.row_major:
            a_csc = a_csc.to_csc()
        h = P()
        piv = U64(2**64 - 1)
        qbuf = _u64(col_perm) if col_perm is not None else None
        st = lib().yo_lu(*a_csc.args(), ordering, pivot_threshold, qbuf.ctypes.data if qbuf is not None else None,
                         C.byref(h), C.byref(piv))
        _check(st, piv.value)
        self._h = h
        self._read()

    def _read(self):
        L, h = lib(), self._h
        s = np.zeros(3, np.uint64)
        L.yo_lu_sizes(h, s)
        n, nl, nu = (int(v) for v in s)
        self.n = n
        lp, li, lv = np.zeros(n + 1, np.uint64), np.zeros(nl, np.uint64), np.zeros(nl)
        up, ui, uv = np.zeros(n + 1, np.uint64), np.zeros(nu, np.uint64), np.zeros(nu)
        rp, cp = np.zeros(n, np.uint64), np.zeros(n, np.uint64)
        L.yo_lu_copy(h, lp, li, lv, up, ui, uv, rp, cp)
        self.l = Matrix(n, n, False, lp.astype(np.int64), li.astype(np.int64), lv)
        self.u = Matrix(n, n, False, up.astype(np.int64), ui.astype(np.int64), uv)
        self.row_perm = rp.astype(np.int64)
        self.col_perm = cp.astype(np.int64)

    def __del__(self):
        if getattr(self, "_h", None):
            lib().yo_free(self._h)
            self._h = None

    def permutation(self):
        return self.row_perm

    def refactor(self, a_csc: Matrix, read: bool = True) -> str:
        """Lu::refactor (lu.rs:662-757): "reused" (same pivots and patterns, new values) or
        "repivoted" (factored afresh under the same column ordering); raises SparseError
        (PatternMismatch, SingularMatrix) and leaves the factors unchanged on an error. With
        read=False the new factors stay in the C++ object (``l``, ``u`` and the permutations
        here are not updated until a later refactor with read=True)."""
        if a_csc.row_major:
            a_csc = a_csc.to_csc()
        outcome = C.c_int()
        _check(lib().yo_lu_refactor(self._h, *a_csc.args(), C.byref(outcome)))
        if read:
            self._read()
        return "reused" if outcome.value == 0 else "repivoted"

    def refactor_work(self) -> int:
        """Multiply-subtracts of one refactor (the sum of |L(:, j)| below the diagonal over the
        entries u_jk above U's diagonal)."""
        return int(lib().yo_lu_refactor_work(self._h))

    def column_permutation(self):
        return self.col_perm

    def solve_many(self, rhs, threads: int = 1):
        rhs = _f64(np.atleast_2d(rhs))
        out = np.zeros_like(rhs)
        _check(lib().yo_lu_solve_many(self._h, rhs.shape[0], rhs, out, threads))
        return out

    def solve(self, b):
        return self.solve_many(np.asarray(b, dtype=np.float64)[None, :])[0]

    def solve_transpose_many(self, rhs, threads: int = 1):
        rhs = _f64(np.atleast_2d(rhs))
        out = np.zeros_like(rhs)
        _check(lib().yo_lu_solve_transpose_many(self._h, rhs.shape[0], rhs, out, threads))
        return out

    def solve_transpose(self, b):
        return self.solve_transpose_many(np.asarray(b, dtype=np.float64)[None, :])[0]

    def solve_refined_normwise(self, a: Matrix, b, max_steps: int, tol: float):
        x = np.zeros(self.n)
        steps, ratio = U64(), DBL()
        _check(lib().yo_lu_refine(self._h, *a.args(), _f64(b), max_steps, tol, x, C.byref(steps), C.byref(ratio)))
        return x, int(steps.value), ratio.value


# ---- synthetic workloads ----

@dataclass
class Grid:
    buses: int
    bx: np.ndarray
    by: np.ndarray
    branch_from: np.ndarray
    branch_to: np.ndarray
    x: np.ndarray
    _h: object = None

    @property
    def branches(self) -> int:
        return int(self.branch_from.shape[0])

    def bridges(self) -> np.ndarray:
        out = np.zeros(self.branches, np.uint8)
        lib().yo_grid_bridges(self._h, out)
        return out.astype(bool)

    def dc(self, in_service=None):
        """(B', Bf, branch_rows): B' over buses 1..N-1 (CSR), Bf over the same unknowns."""
        L = lib()
        hb, hf = P(), P()
        rows = np.zeros(self.branches, np.uint64)
        mask = None
        if in_service is not None:
            mask = np.ascontiguousarray(in_service, dtype=np.uint8)
        _check(L.yo_dc(self._h, mask.ctypes.data if mask is not None else None, C.byref(hb), C.byref(hf), rows))
        bprime, bf = _take_matrix(hb), _take_matrix(hf)
        return bprime, bf, rows[: bf.rows].astype(np.int64)

    def jacobian(self, seed: int, shift: float = 0.0, shift_seed: int = 0):
        """The AC Newton Jacobian (CSC) and its operating point (pv, pq, vm, va). shift > 0 moves
        every |V| and non-slack angle by shift * U[-1, 1] (drawn from shift_seed), as a Newton
        iterate would, keeping the Jacobian's pattern."""
        L = lib()
        hj, ha = P(), P()
        _check(L.yo_jacobian(self._h, seed, shift, shift_seed, C.byref(hj), C.byref(ha)))
        jac = _take_matrix(hj)
        s = np.zeros(4, np.uint64)
        L.yo_ac_sizes(ha, s)
        pv, pq = np.zeros(int(s[0]), np.uint64), np.zeros(int(s[1]), np.uint64)
        vm, va = np.zeros(int(s[2])), np.zeros(int(s[2]))
        r, bc = np.zeros(int(s[3])), np.zeros(int(s[3]))
        L.yo_ac_copy(ha, pv, pq, vm, va, r, bc)
        L.yo_free(ha)
        return jac, dict(pv=pv.astype(np.int64), pq=pq.astype(np.int64), vm=vm, va=va, r=r, bc=bc)

    def __del__(self):
        if self._h:
            lib().yo_free(self._h)
            self._h = None


def make_grid(buses: int, seed: int) -> Grid:
    L = lib()
    h = P()
    _check(L.yo_grid(buses, seed, C.byref(h)))
    s = np.zeros(2, np.uint64)
    L.yo_grid_sizes(h, s)
    nb, nl = int(s[0]), int(s[1])
    bx, by = np.zeros(nb), np.zeros(nb)
    f, t, x = np.zeros(nl, np.uint64), np.zeros(nl, np.uint64), np.zeros(nl)
    L.yo_grid_copy(h, bx, by, f, t, x)
    return Grid(nb, bx, by, f.astype(np.int64), t.astype(np.int64), x, h)


def uniforms(seed: int, stream: int, count: int) -> np.ndarray:
    out = np.zeros(count)
    lib().yo_uniforms(seed, stream, count, out)
    return out


def normals(seed: int, stream: int, count: int) -> np.ndarray:
    out = np.zeros(count)
    lib().yo_normals(seed, stream, count, out)
    return out


def integers(seed: int, stream: int, count: int) -> np.ndarray:
    out = np.zeros(count, np.uint64)
    lib().yo_integers(seed, stream, count, out)
    return out


def injections(seed: int, stream: int, buses: int, count: int) -> np.ndarray:
    """``count`` zero-sum N(0,1) injection scenarios with the slack dropped: (count, buses-1)."""
    out = np.zeros(count * (buses - 1))
    lib().yo_injections(seed, stream, buses, count, out)
    return out.reshape(count, buses - 1)


# ---- the adjoint (no reference counterpart: the composition the GPU adjoint is checked against) ----

def adjoint(factor, a: Matrix, x_bar, x):
    """CPU composition of the adjoint of A x = b: b_bar_s = A^-T x_bar_s (Lu.solve_transpose_many,
    lu.rs:502-533; a Cholesky factor's A^T = A, so its solve_many), and a_bar = -(sum over s of
    b_bar_s x_s^T) on A's stored pattern, each stored entry summed over s in ascending order, one
    product and one addition rounded at a time. Returns (b_bar, a_bar), a_bar in A's storage
    order."""
    x_bar = np.asarray(x_bar, dtype=np.float64)
    x = np.asarray(x, dtype=np.float64)
    if isinstance(factor, Lu):
        b_bar = factor.solve_transpose_many(x_bar) if len(x_bar) else np.zeros_like(x_bar)
    else:
        b_bar = factor.solve_many(x_bar) if len(x_bar) else np.zeros_like(x_bar)
    major = np.repeat(np.arange(len(a.ptr) - 1), np.diff(np.asarray(a.ptr, dtype=np.int64)))
    minor = np.asarray(a.idx, dtype=np.int64)
    rows, cols = (major, minor) if a.row_major else (minor, major)
    acc = np.zeros(a.nnz)
    for s in range(len(x)):
        acc = acc + b_bar[s, rows] * x[s, cols]
    return b_bar, -acc
