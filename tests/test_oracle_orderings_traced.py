"""AMD's and COLAMD's exact permutations on tiny patterns, each derived by hand from the
reference's code and written out as the trace beside its literal, so that the tie rules the
permutations depend on are pinned, not only their properties:

- AMD (ordering.rs:236-492): the degree lists are last in, first out (:252-285); the initial
  lists are filled from node n-1 down to node 0 (:342-347), so node 0 heads its list, and the
  variables of a new element are put back in L_p order (:476-484), so the last one heads its
  list; supervariables merge into the first of their kind in L_p order (:449-470), whose
  members follow it (:387-388); an element whose variables all lie in L_p is absorbed
  (aggressive absorption, :409-413), so it drops out of the elements lists that decide which
  variables are indistinguishable.
- COLAMD (colamd.rs:35-261): score buckets are last in, first out (:35-87); a column merges into
  the first earlier column of the pivot row with the same rows (:224-250); a row with no column
  outside the pivot row is absorbed (:201-208).

Each mistake is then planted in the oracle (its Mutation flags) to show that these literals
catch it: with first-in first-out lists, the last representative, or no aggressive absorption,
at least one literal changes.

Notation in the traces: h[d] is degree (score) list d, head first; L_p the new element's
variables; w weights; "bound" the approximate degree min(remaining - w_i, d_i + |L_p| - w_i,
external_i + |L_p| - w_i) (ordering.rs:476-484)."""
import pytest

import oracle as O


def graph(n, edges):
    """The symmetric pattern of a graph's edges, with a diagonal (the diagonal is not an edge of
    AMD's adjacency, ordering.rs:75-117)."""
    t = [(i, i, 1.0) for i in range(n)]
    for i, j in edges:
        t += [(i, j, 1.0), (j, i, 1.0)]
    return O.Matrix.from_triplets(n, n, t)


def pattern(rows):
    """A CSC pattern from a dense 0/1 table."""
    t = [(i, j, 1.0) for i, row in enumerate(rows) for j, v in enumerate(row) if v]
    return O.Matrix.from_triplets(len(rows), len(rows[0]), t, row_major=False)


# A triangle 0-2-3 and an isolated node 1.
#   neighbours 0:[2,3] 1:[] 2:[0,3] 3:[0,2]; degrees 2,0,2,2; lists filled 3,2,1,0:
#   h[2] = 0,2,3 and h[0] = 1.
#   1. p = 1 (h[0]); L_p = {} .                                              order 1
#   2. p = 0 (h[2]); L_p = [2,3]. Both now have elements [0] and no variables (0 is an element,
#      the other in L_p): the same hash, so 3 merges into 2 (the first in L_p); w2 = 2.
#      bound(2) = min(remaining 2 - 2, ...) = 0: h[0] = 2.                   order 1,0
#   3. p = 2, then its member 3.                                             order 1,0,2,3
#   FIFO lists: h[2] = 3,2,0, so p = 3 at step 2 and 2 merges into 0: 1,3,0,2.
#   Last representative: 2 merges into 3 at step 2: 1,0,3,2.
TRIANGLE = (4, [(0, 3), (0, 2), (2, 3)], [1, 0, 2, 3])

# Five nodes, edges 0-4 1-3 1-2 2-4 1-4.
#   neighbours 0:[4] 1:[2,3,4] 2:[1,4] 3:[1] 4:[0,1,2]; degrees 1,3,2,1,3;
#   h[1] = 0,3; h[2] = 2; h[3] = 1,4.
#   1. p = 0; L_p = [4]: elements [0], variables [1,2] (external 2);
#      bound = min(4 - 1, 3, 2) = 2: h[2] = 4,2; |L_0| = 1.                  order 0
#   2. p = 3 (h[1]); L_p = [1]: elements [3], variables [2,4]; bound = min(3 - 1, 3, 2) = 2:
#      h[2] = 1,4,2.                                                         order 0,3
#   3. p = 1 (h[2]); L_p = [2,4] (1's variables), element 3 absorbed into 1. For 4, element 0
#      has |L_0 \ L_p| = 1 - 1 = 0, so it is absorbed (aggressive absorption): 2 and 4 both have
#      elements [1] and no variables, and 4 merges into 2; w2 = 2, bound 0. order 0,3,1
#   4. p = 2, then 4.                                                        order 0,3,1,2,4
#   No aggressive absorption: 4 keeps element 0 ([0,1] against 2's [1]), no merge; both get
#   bound min(2 - 1, ..., 0 + 1) = 1 and 4, put back last, heads h[1]: 0,3,1,4,2.
#   FIFO lists: h[1] = 3,0, so 3 goes first (1 joins h[2] behind 2), then 0 (4 joins h[2]
#   last), then 2 from h[2] = 2,1,4, whose L_p = [1,4] merge into 1: 3,0,2,1,4.
#   Last representative: 2 merges into 4 at step 3: 0,3,1,4,2.
ABSORPTION = (5, [(0, 4), (1, 3), (1, 2), (2, 4), (1, 4)], [0, 3, 1, 2, 4])

# The complete bipartite graph K(2,3): 0,1 against 2,3,4.
#   degrees 3,3,2,2,2; h[2] = 2,3,4; h[3] = 0,1.
#   1. p = 2; L_p = [0,1]: each has elements [2], variables [3,4]: 1 merges into 0; w0 = 2,
#      bound min(4 - 2, 3, 2) = 2: h[2] = 0,3,4.                             order 2
#   2. p = 0 (put back last, so at the head), then its member 1; L_p = [3,4] (element 2
#      absorbed): each has elements [0], no variables: 4 merges into 3, bound 0.
#                                                                            order 2,0,1
#   3. p = 3, then 4.                                                        order 2,0,1,3,4
#   FIFO lists: p = 4 first (h[2] = 4,3,2), then 3, 0 (with 1), 2: 4,3,0,1,2.
#   Last representative: 0 merges into 1, 3 into 4: 2,1,0,4,3.
BIPARTITE = (5, [(0, 2), (0, 3), (0, 4), (1, 2), (1, 3), (1, 4)], [2, 0, 1, 3, 4])

# The path 0-1-2-3-4: degrees 1,2,2,2,1; h[1] = 0,4; h[2] = 1,2,3.
#   Each step eliminates the head of h[1], and its one neighbour gets bound 1 and goes to the
#   head of h[1], ahead of 4: 0, 1, 2, 3, then 4.                            order 0,1,2,3,4
#   FIFO lists: h[1] = 4,0 takes 4, then 0; 3 and 1 join h[1] behind the others: 4,0,3,1,2.
PATH = (5, [(0, 1), (1, 2), (2, 3), (3, 4)], [0, 1, 2, 3, 4])

AMD_CASES = {"triangle": TRIANGLE, "absorption": ABSORPTION, "bipartite": BIPARTITE, "path": PATH}


@pytest.mark.parametrize("name", sorted(AMD_CASES))
def test_amd_permutations_traced_by_hand(name):
    n, edges, want = AMD_CASES[name]
    assert list(graph(n, edges).ordering(O.AMD)) == want


@pytest.mark.parametrize("mutation,changed", [
    (O.FIFO_LISTS, {"triangle": [1, 3, 0, 2], "absorption": [3, 0, 2, 1, 4], "bipartite": [4, 3, 0, 1, 2],
                    "path": [4, 0, 3, 1, 2]}),
    (O.LAST_REPRESENTATIVE, {"triangle": [1, 0, 3, 2], "absorption": [0, 3, 1, 4, 2], "bipartite": [2, 1, 0, 4, 3]}),
    (O.NO_AGGRESSIVE_ABSORPTION, {"absorption": [0, 3, 1, 4, 2]}),
])
def test_a_planted_amd_mistake_changes_a_traced_permutation(mutation, changed):
    """The mutated oracle gives the permutations the traces predict for that mistake, so each
    literal above fails under it."""
    for name, (n, edges, want) in AMD_CASES.items():
        got = list(graph(n, edges).ordering(O.AMD, mutation))
        assert got == changed.get(name, want), name
        assert (got != want) == (name in changed)


# COLAMD on the 3 x 3 identity: every column's score is 0 (its one row holds only it); the
# buckets are filled for columns 0, 1, 2 in turn, so h[0] = 2,1,0 and the order is 2,1,0.
# FIFO buckets: 0,1,2.
IDENTITY = ([[1, 0, 0], [0, 1, 0], [0, 0, 1]], [2, 1, 0])

# Rows: 0 = {0,1,2}, 1 = {1}, 2 = {2} (an arrow). Row sizes 3,1,1; scores (sum of row size - 1
# over a column's rows, at most live columns - 1 = 2): 2, 2, 2; h[2] = 2,1,0.
#   1. pivot 2; its rows 0 and 2 form pivot row 3 = {0,1} (size 2) and are dropped. Column 0's
#      rows are now [3]; column 1's [1,3], where row 1 has |1 \ 3| = 1 - 1 = 0 and is absorbed,
#      so [3]. Scores 2 - 1 = 1 each; 1 has 0's rows and merges into 0 (the first): w0 = 2,
#      score 1 - 1 = 0.                                                      order 2
#   2. pivot 0, then its member 1.                                           order 2,0,1
#   FIFO buckets: h[2] = 0,1,2: pivot 0, then 1 with member 2: 0,1,2.
#   Last representative: 0 merges into 1: 2,1,0. No aggressive absorption: 1 keeps row 1, no
#   merge; scores 1 and 1 (1 put back last): 2,1,0.
ARROW = ([[1, 1, 1], [0, 1, 0], [0, 0, 1]], [2, 0, 1])

# Rows: 0 = {0,3}, 1 = {1}, 2 = {0,1,2,3}, 3 = {3}. Row sizes 2,1,4,1; scores min(., 3):
# 3, 3, 3, 3; h[3] = 3,2,1,0.
#   1. pivot 3; rows 0, 2, 3 form pivot row 4 = {0,1,2}; row 1 (|1 \ 4| = 0) is absorbed; all
#      three columns have rows [4], score 3 - 1 = 2: 1 and 2 merge into 0 (w0 = 3, score 0).
#                                                                            order 3
#   2. pivot 0, then 1 and 2.                                                order 3,0,1,2
#   FIFO buckets: pivot 0 first (rows 0, 2 give pivot row {3,1,2}; 1 and 2 merge into 3): 0,3,1,2.
#   Last representative: 0 into 1, then 1 (with 0) into 2: 3,2,1,0.
#   No aggressive absorption: 1 keeps row 1, only 2 merges into 0 (w0 = 2, score 1, capped at
#   3 - 2 = 1) and 1 scores 2: pivot 0 (with 2), then 1: 3,0,2,1.
FOUR = ([[1, 0, 0, 1], [0, 1, 0, 0], [1, 1, 1, 1], [0, 0, 0, 1]], [3, 0, 1, 2])

COLAMD_CASES = {"identity": IDENTITY, "arrow": ARROW, "four": FOUR}


@pytest.mark.parametrize("name", sorted(COLAMD_CASES))
def test_colamd_permutations_traced_by_hand(name):
    rows, want = COLAMD_CASES[name]
    assert list(pattern(rows).ordering(O.COLAMD)) == want


@pytest.mark.parametrize("mutation,changed", [
    (O.FIFO_LISTS, {"identity": [0, 1, 2], "arrow": [0, 1, 2], "four": [0, 3, 1, 2]}),
    (O.LAST_REPRESENTATIVE, {"arrow": [2, 1, 0], "four": [3, 2, 1, 0]}),
    (O.NO_AGGRESSIVE_ABSORPTION, {"arrow": [2, 1, 0], "four": [3, 0, 2, 1]}),
])
def test_a_planted_colamd_mistake_changes_a_traced_permutation(mutation, changed):
    for name, (rows, want) in COLAMD_CASES.items():
        got = list(pattern(rows).ordering(O.COLAMD, mutation))
        assert got == changed.get(name, want), name
        assert (got != want) == (name in changed)


def test_the_mutations_change_the_large_orderings_too():
    """On the configurations' kind of matrix (a 2,000-bus B' for AMD, a Jacobian for COLAMD) each
    planted mistake gives another permutation, so the full-size factors the GPU tests solve
    depend on the rules pinned above."""
    bprime, _, _ = O.make_grid(2000, 20261203).dc()
    amd = list(bprime.ordering(O.AMD))
    jac, _ = O.make_grid(600, 5).jacobian(5)
    colamd = list(jac.ordering(O.COLAMD))
    for mutation in (O.FIFO_LISTS, O.LAST_REPRESENTATIVE, O.NO_AGGRESSIVE_ABSORPTION):
        assert list(bprime.ordering(O.AMD, mutation)) != amd, mutation
        assert list(jac.ordering(O.COLAMD, mutation)) != colamd, mutation
