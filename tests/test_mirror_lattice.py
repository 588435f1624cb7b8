"""The host-side lattice geometry of the Python mirror (LatticeCyclic,
Direction, LatticeLink(Canonical); src/lattice.rs) against the reference's
doctests (lattice.rs:62-132, 256-323) and the oracle's C geometry (site
index, coordinates, neighbours, link order), exhaustively on small lattices.
Integer work: everything is exact."""
import numpy as np
import pytest

import oracle as O
import lqcd
from lqcd import Direction, LatticeLinkCanonical

X, Y = 0, 1


def test_reference_doctests():
    lat = lqcd.LatticeCyclic(1.0, 4, 4)
    point = (1, 0, 2, 0)
    # link_canonical, lattice.rs:66-79
    assert lat.link_canonical(point, Direction(X, False)) == LatticeLinkCanonical((0, 0, 2, 0), Direction(X, True))
    assert lat.link_canonical(point, Direction(X, True)) == LatticeLinkCanonical((1, 0, 2, 0), Direction(X, True))
    assert lat.link_canonical(point, Direction(Y, False)) == LatticeLinkCanonical((1, 3, 2, 0), Direction(Y, True))
    # link, lattice.rs:104-116 (D = 3)
    l3 = lqcd.LatticeCyclic(1.0, 4, 3)
    assert l3.link((1, 2, 5), Direction(0, True)).pos() == (1, 2, 1)
    # add_point_direction(_n), lattice.rs:256-302
    assert lat.add_point_direction(point, Direction(X, True)) == (2, 0, 2, 0)
    assert lat.add_point_direction(point, Direction(Y, False)) == (1, 3, 2, 0)
    assert lat.add_point_direction_n(point, Direction(X, True), 2) == (3, 0, 2, 0)
    assert lat.add_point_direction_n(point, Direction(Y, False), 3) == (1, 1, 2, 0)
    # LatticeLinkCanonical::new refuses a negative direction
    assert LatticeLinkCanonical.new(point, Direction(X, False)) is None
    assert Direction.directions_vec(2) == [Direction(0, True), Direction(0, False),
                                           Direction(1, True), Direction(1, False)]
    assert -Direction(1, True) == Direction(1, False)


@pytest.mark.parametrize("ext", [[4, 4, 4, 4], [3, 5, 2], [6, 4], [2, 3, 4, 5]])
def test_geometry_against_oracle(ext):
    d = len(ext)
    lat = lqcd.LatticeCyclic(1.0, ext, d)
    ol = O.OracleLattice(d, ext)
    points = list(lat.get_points())
    assert len(points) == lat.number_of_points() == ol.nsite
    for s, p in enumerate(points):
        assert list(p) == ol.site_coords(s)
        assert lat.index_of(p) == s == ol.site_index(p)
        for mu in range(d):
            for pos in (True, False):
                q = lat.add_point_direction(p, Direction(mu, pos))
                assert lat.index_of(q) == ol.neighbor(s, mu, pos)
                # n steps equal n single steps
                r = p
                for _ in range(3):
                    r = lat.add_point_direction(r, Direction(mu, pos))
                assert lat.add_point_direction_n(p, Direction(mu, pos), 3) == r
    sites, dirs = ol.link_order()
    links = list(lat.get_links())
    assert len(links) == lat.number_of_canonical_links_space()
    for i, l in enumerate(links):
        assert l.to_index(lat) == i
        assert lat.index_of(l.pos()) == int(sites[i]) and l.dir().index() == int(dirs[i])
    # a negative link is the positive link of the point behind it
    for s, p in enumerate(points[:50]):
        for mu in range(d):
            c = lat.into_canonical(lqcd.LatticeLink(p, Direction(mu, False)))
            assert lat.index_of(c.pos()) == ol.neighbor(s, mu, False) and c.dir() == Direction(mu, True)
    assert lat.has_compatible_length_links(np.zeros((lat.number_of_canonical_links_space(), 18)))
    assert not lat.has_compatible_length_links(np.zeros((3, 18)))
    assert lat.has_compatible_length_e_field(np.zeros((lat.number_of_canonical_links_space(), 8)))
