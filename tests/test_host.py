"""Host-side logic of the Python package that needs no GPU."""
import itnbp


def test_edge_coloring_is_a_proper_cover_and_cached():
    """simple_update_round_ applies one colour class per launch: every edge once,
    no two bonds of a class sharing a site, greedy first fit in edge order (the
    4-colouring of an open square lattice), and the same classes on every call."""
    for dims in ((16, 16), (5, 7), (1, 6)):
        g = itnbp.named_grid(dims)
        cls = itnbp.edge_coloring(g)
        flat = [e for c in cls for e in c]
        assert sorted(flat) == sorted(g.edges)
        for c in cls:
            sites = [v for e in c for v in e]
            assert len(sites) == len(set(sites))
        assert len(cls) <= 4
        assert itnbp.edge_coloring(g) == cls
    hh = itnbp.heavy_hex(3, 3)
    cls = itnbp.edge_coloring(hh)
    assert sorted(e for c in cls for e in c) == sorted(hh.edges)
