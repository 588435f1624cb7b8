"""Pins the oracle's lattice geometry (src/lattice.rs) to the reference's doctests,
its iterator tests and exhaustive index identities."""
import itertools

import numpy as np
import pytest

import oracle as O

X, Y, Z, T = 0, 1, 2, 3


def test_link_canonical_doctest():
    # src/lattice.rs:62-82 (D=4, L=4, point [1,0,2,0])
    l = O.OracleLattice(4, 4)
    s = l.site_index([1, 0, 2, 0])
    lib = O.lib()
    assert lib.or_link_canonical(l.ref, s, X, 0) == l.site_index([0, 0, 2, 0]) * 4 + X
    assert lib.or_link_canonical(l.ref, s, X, 1) == l.site_index([1, 0, 2, 0]) * 4 + X
    assert lib.or_link_canonical(l.ref, s, Y, 0) == l.site_index([1, 3, 2, 0]) * 4 + Y


def test_link_position_wraps_doctest():
    # src/lattice.rs:105-118: D=3, L=4, point [1,2,5] -> [1,2,1]
    l = O.OracleLattice(3, 4)
    assert l.site_index([1, 2, 5]) == l.site_index([1, 2, 1])


def test_add_point_direction_doctests():
    # src/lattice.rs:256-270 and :290-305
    l = O.OracleLattice(4, 4)
    s = l.site_index([1, 0, 2, 0])
    assert l.site_coords(l.neighbor(s, X, True)) == [2, 0, 2, 0]
    assert l.site_coords(l.neighbor(s, Y, False)) == [1, 3, 2, 0]
    s2 = l.neighbor(l.neighbor(s, X, True), X, True)
    assert l.site_coords(s2) == [3, 0, 2, 0]
    s3 = s
    for _ in range(3):
        s3 = l.neighbor(s3, Y, False)
    assert l.site_coords(s3) == [1, 1, 2, 0]


@pytest.mark.parametrize("dim,L", [(4, 8), (4, 16), (3, 8)])
def test_number_of_points_doctest(dim, L):
    # src/lattice.rs:216-245
    l = O.OracleLattice(dim, L)
    assert l.nsite == L ** dim
    assert l.nlink == L ** dim * dim


@pytest.mark.parametrize("L", [2, 10, 26])
def test_iterator_length(L):
    # src/test/mod.rs:21-41; the iteration order is the storage order
    l = O.OracleLattice(4, L)
    sites, dirs = l.link_order()
    assert len(sites) == 4 * L ** 4
    idx = sites.astype(np.int64) * 4 + dirs
    assert np.array_equal(idx, np.arange(4 * L ** 4))


def test_to_index_formula_exhaustive():
    # src/lattice.rs:909-915: sum_i (x_i mod L) L^i, x_0 fastest
    for dim, ext in [(2, [3, 5]), (3, [4, 4, 4]), (4, [2, 3, 4, 6])]:
        l = O.OracleLattice(dim, ext)
        for x in itertools.product(*[range(e) for e in ext]):
            s = l.site_index(list(x))
            expect = 0
            stride = 1
            for i in range(dim):
                expect += x[i] * stride
                stride *= ext[i]
            assert s == expect
            assert l.site_coords(s) == list(x)


@pytest.mark.parametrize("dim,ext", [(4, [4, 4, 4, 4]), (3, [2, 4, 6]), (4, [2, 2, 4, 6])])
def test_neighbor_involution_and_shift(dim, ext):
    l = O.OracleLattice(dim, ext)
    for s in range(l.nsite):
        x = l.site_coords(s)
        for d in range(dim):
            p = l.neighbor(s, d, True)
            m = l.neighbor(s, d, False)
            assert l.neighbor(p, d, False) == s
            assert l.neighbor(m, d, True) == s
            xp = list(x)
            xp[d] = (xp[d] + 1) % ext[d]
            xm = list(x)
            xm[d] = (xm[d] - 1) % ext[d]
            assert l.site_coords(p) == xp
            assert l.site_coords(m) == xm
            if all(e % 2 == 0 for e in ext):
                assert l.parity(p) != l.parity(s)


def test_lattice_init_errors():
    # src/test/mod.rs:602-640 (LatticeInitializationError)
    with pytest.raises(ValueError):
        O.OracleLattice(3, 4, 0.0)
    with pytest.raises(ValueError):
        O.OracleLattice(4, 4, -1.0)
    with pytest.raises(ValueError):
        O.OracleLattice(4, 4, float("nan"))
    with pytest.raises(ValueError):
        O.OracleLattice(4, 4, -0.0)
    with pytest.raises(ValueError):
        O.OracleLattice(4, 4, float("inf"))
    with pytest.raises(ValueError):
        O.OracleLattice(3, 1, 1.0)
    O.OracleLattice(3, 2, 1.0)
