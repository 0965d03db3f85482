"""Graph semirings on the device against the oracle: R-MAT generation and the
graph build (bit-exact), BFS levels over lor-land (bit-exact with a textbook
queue BFS), k-hop neighbourhoods (bit-exact with the reference's knn fixtures)
and simple_pagerank (<= 1e-12 relative to the reference's fixtures)."""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def A():
    import alp_amd
    alp_amd.init(0)
    return alp_amd


def expected_crs(n, src, dst):
    I, J = O.symmetrize_dedup(n, src, dst)  # sorted by (I, J)
    off = np.zeros(n + 1, np.uint64)
    off[1:] = np.cumsum(np.bincount(I, minlength=n))
    return off, J, I


@pytest.mark.parametrize("scale,seed", [(6, 1), (10, 3), (14, 7)])
def test_rmat_generator_and_build_bitexact(A, scale, seed):
    n, m = 1 << scale, 16 << scale
    s, d = A.rmat_edges(scale, m, seed=seed)
    os_, od = O.rmat_edges(scale, m, seed=seed)
    assert np.array_equal(s.numpy(), os_) and np.array_equal(d.numpy(), od)
    G = A.graph_from_edges(n, s, d)
    off, idx, I = expected_crs(n, os_, od)
    goff, gidx, _ = G.download(0, with_values=False)
    assert np.array_equal(goff, off) and np.array_equal(gidx, idx)
    coff, cidx, _ = G.download(1, with_values=False)  # symmetric: CCS == CRS
    assert np.array_equal(coff, off) and np.array_equal(cidx, idx)


def test_build_directed_with_loops_and_duplicates(A):
    src = np.array([0, 0, 1, 2, 2, 3, 3], np.uint32)
    dst = np.array([1, 1, 1, 0, 3, 3, 0], np.uint32)
    G = A.graph_from_edges(5, A.Vector.from_numpy(src), A.Vector.from_numpy(dst), flags=0)
    off, idx, _ = G.download(0, with_values=False)
    assert list(off) == [0, 1, 2, 4, 6, 6]
    assert list(idx) == [1, 1, 0, 3, 0, 3]
    coff, cidx, _ = G.download(1, with_values=False)
    assert list(coff) == [0, 2, 4, 4, 6, 6] and list(cidx) == [2, 3, 0, 1, 2, 3]
    G2 = A.graph_from_edges(5, A.Vector.from_numpy(src), A.Vector.from_numpy(dst), flags=A.GRAPH_DROP_SELF_LOOPS)
    off2, idx2, _ = G2.download(0, with_values=False)
    # rows: 0 -> {1}, 1 -> {} (loop dropped), 2 -> {0, 3}, 3 -> {0} (loop dropped), 4 -> {}
    assert list(off2) == [0, 1, 1, 3, 4, 4] and list(idx2) == [1, 0, 3, 0]


@pytest.mark.parametrize("scale,seed", [(10, 1), (14, 5), (18, 2)])
def test_bfs_levels_bitexact(A, scale, seed):
    n, m = 1 << scale, 16 << scale
    s, d = A.rmat_edges(scale, m, seed=seed)
    G = A.graph_from_edges(n, s, d)
    goff, gidx, _ = G.download(0, with_values=False)
    ref = O.CSR(n, n, goff, gidx, None)
    src, _ = A.max_degree_vertex(G)
    assert src == int(np.argmax(np.diff(goff.astype(np.int64))))
    rng = np.random.default_rng(seed)
    for source in [src, int(rng.integers(0, n)), int(np.argmin(np.diff(goff.astype(np.int64))))]:
        lev, nl, reached = A.bfs(G, source)
        want = O.bfs_levels(ref, source)
        assert np.array_equal(lev.numpy(), want), source
        assert reached == int((want >= 0).sum()) and nl == int(want.max()) + 1


def test_bfs_on_a_path_forces_many_levels(A):
    n = 3000
    src = np.arange(n - 1, dtype=np.uint32)
    G = A.graph_from_edges(n, A.Vector.from_numpy(src), A.Vector.from_numpy(src + 1))
    lev, nl, reached = A.bfs(G, 0)
    assert np.array_equal(lev.numpy(), np.arange(n, dtype=np.int32)) and nl == n and reached == n


@pytest.mark.parametrize("scale", [9, 11])
def test_khop_and_pagerank_match_reference_fixtures(A, golden, scale):
    g = golden(f"graph_rmat_s{scale}")
    n = 1 << scale
    G = A.graph_from_edges(n, A.Vector.from_numpy(g["I"]), A.Vector.from_numpy(g["J"]), flags=0)
    src = int(g["source"][0])
    for k in (1, 2, 3):
        assert np.array_equal(A.khop(G, src, k).numpy(), g[f"knn_{k}"]), k
    rc, pr, it, q = A.pagerank(G, 0.85, 1e-9, 1000)
    assert rc == A.SUCCESS
    assert abs(it - int(g["iterations"][0])) <= 1
    want = g["pr"]
    assert np.max(np.abs(pr.numpy() - want)) <= 1e-12 * np.max(want)


def test_pagerank_smoke10(A, golden):
    g = golden("pagerank_smoke10")
    G = A.graph_from_edges(10, A.Vector.from_numpy(g["I"]), A.Vector.from_numpy(g["J"]), flags=0)
    rc, pr, it, q = A.pagerank(G, 0.85, 1e-7, 1000)
    assert rc == A.SUCCESS and abs(it - int(g["iterations"][0])) <= 1
    assert np.max(np.abs(pr.numpy() - g["pr"])) <= 1e-12
    assert A.pagerank(G, 0.85, 1e-30, 5)[0] == A.FAILED  # not converged: FAILED, as the reference


def test_rmat_scale22_bfs_and_pagerank_against_oracle(A):
    """The graph workload's own size: R-MAT scale 22, edge factor 16, seed 1,
    symmetrised. BFS levels bit-exact with the oracle's queue BFS; PageRank
    (0.85, 1e-9) against the oracle's restatement of simple_pagerank.hpp:146-533
    on the same CRS/CCS: iterations within 1, ranks within 1e-11 of the largest.
    The device folds a hub column (over 10^5 entries here) in chunks, the
    reference sequentially; recursive summation of L positive terms in two
    orders can differ by up to (L - 1) eps relative (1e-11 at L = 10^5), and
    45 iterations carry it. (The scale-9 and scale-11 graphs, whose columns are
    short, agree with the reference's fixtures to 1e-12.)"""
    scale = 22
    n, m = 1 << scale, 16 << scale
    s, d = A.rmat_edges(scale, m, seed=1)
    G = A.graph_from_edges(n, s, d)
    del s, d
    src, deg = A.max_degree_vertex(G)
    lev, nl, reached = A.bfs(G, src)
    off, idx, _ = G.download(0, with_values=False)
    crs = O.CSR(n, n, off, idx, None)
    want = O.bfs_levels(crs, src)
    assert np.array_equal(lev.numpy(), want)
    rc, pr, it, q = A.pagerank(G, 0.85, 1e-9, 1000)
    coff, cidx, _ = G.download(1, with_values=False)
    ccs = O.CSR(n, n, coff, cidx, None)
    orc, opr, oit, oq = O.pagerank(crs, ccs, 0.85, 1e-9, 1000)
    err = np.max(np.abs(pr.numpy() - opr)) / np.max(opr)
    print(f"scale 22: nnz={G.nnz} pagerank it={it} (oracle {oit}) q={q:.3e} (oracle {oq:.3e}) "
          f"max difference {err:.2e} of the largest rank", flush=True)
    assert rc == orc == A.SUCCESS
    assert abs(it - oit) <= 1
    assert err <= 1e-11


def test_bfs_directed_pull_walks_in_neighbours(A):
    """Directed pattern (no symmetrisation): a pull step must look at each
    vertex's in-neighbours. Edges 0->1, 1->2, 1->3, 4->2 from source 0: the
    source's level pushes (to 1); the next level pulls (the frontier's 2 edges
    exceed the 1 unexplored / 14) and finds 2 and 3 through their in-neighbour 1,
    which they do not point to; 4 is unreachable. With edges 2->0 and 3->0 added,
    the source's level reaches nothing new through them."""
    src = np.array([0, 1, 1, 4], np.uint32)
    dst = np.array([1, 2, 3, 2], np.uint32)
    G = A.graph_from_edges(5, A.Vector.from_numpy(src), A.Vector.from_numpy(dst), flags=0)
    lev, nl, reached = A.bfs(G, 0)
    assert list(lev.numpy()) == [0, 1, 2, 2, -1] and nl == 3 and reached == 4
    src = np.array([0, 1, 1, 4, 2, 3], np.uint32)
    dst = np.array([1, 2, 3, 2, 0, 0], np.uint32)
    G = A.graph_from_edges(5, A.Vector.from_numpy(src), A.Vector.from_numpy(dst), flags=0)
    lev, nl, reached = A.bfs(G, 0)
    assert list(lev.numpy()) == [0, 1, 2, 2, -1] and nl == 3 and reached == 4


def test_bfs_on_a_pattern_with_repeated_entries(A):
    """A COO pattern that stores positions more than once (create_coo keeps
    them, as buildMatrixUnique keeps what it is handed): every vertex is still
    claimed once per level, so the levels and the count of reached vertices
    match the oracle's queue BFS over the same CRS. Vertex 0's row repeats each
    of its 300 neighbours three times, so its first level may not take the
    plain-store claim of a duplicate-free single-vertex frontier."""
    rng = np.random.default_rng(5)
    n = 4096
    I = rng.integers(0, n, 40000)
    J = rng.integers(0, n, 40000)
    nb = rng.choice(np.arange(1, n), 300, replace=False)
    I = np.concatenate([I, np.zeros(900, np.int64)])
    J = np.concatenate([J, np.tile(nb, 3)])
    G = A.Matrix.from_coo(n, n, I, J)
    off, idx, _ = G.download(0, with_values=False)
    ref = O.CSR(n, n, off, idx, None)
    out = A.Vector(n, np.int32)
    for source in (0, 17, int(rng.integers(0, n))):
        lev, nl, reached = A.bfs(G, source)
        want = O.bfs_levels(ref, source)
        assert np.array_equal(lev.numpy(), want), source
        assert reached == int((want >= 0).sum()) and nl == int(want.max()) + 1
        # into a caller-owned vector, reused across sources
        lev2, nl2, reached2 = A.bfs(G, source, out=out)
        assert lev2 is out and np.array_equal(out.numpy(), want) and (nl2, reached2) == (nl, reached)
    with pytest.raises(ValueError):
        A.bfs(G, 0, out=A.Vector(n, np.uint8))


@pytest.mark.parametrize("scale,seed", [(12, 4), (16, 9)])
def test_bfs_directed_rmat_bitexact(A, scale, seed):
    """Directed R-MAT graphs (self loops dropped, no symmetrisation): levels
    follow out-edges, bit-exact with the oracle's queue BFS over the CRS, from
    the highest out-degree vertex (whose frontier forces the pull switch) and
    from a random one."""
    n, m = 1 << scale, 16 << scale
    s, d = A.rmat_edges(scale, m, seed=seed)
    G = A.graph_from_edges(n, s, d, flags=A.GRAPH_DROP_SELF_LOOPS)
    goff, gidx, _ = G.download(0, with_values=False)
    ref = O.CSR(n, n, goff, gidx, None)
    top, _ = A.max_degree_vertex(G)
    rng = np.random.default_rng(seed)
    for source in (top, int(rng.integers(0, n))):
        lev, nl, reached = A.bfs(G, source)
        want = O.bfs_levels(ref, source)
        assert np.array_equal(lev.numpy(), want), source
        assert reached == int((want >= 0).sum()) and nl == int(want.max()) + 1


@pytest.mark.parametrize("scale,seed", [(10, 3), (16, 7)])
def test_khop_on_symmetric_graph_matches_product_path(A, scale, seed):
    """A symmetrised graph answers k-hop with a BFS cut at depth k; the same
    pattern loaded as a plain CRS matrix takes the k boolean products.  Both
    must equal the oracle for every k, including 0 and a k past the diameter."""
    n, m = 1 << scale, 8 << scale
    s, d = A.rmat_edges(scale, m, seed=seed)
    G = A.graph_from_edges(n, s, d)
    goff, gidx, _ = G.download(0, with_values=False)
    P = A.Matrix.from_crs(n, n, goff, gidx[:int(goff[-1])])
    ref = O.CSR(n, n, goff, gidx, None)
    src, _ = A.max_degree_vertex(G)
    rng = np.random.default_rng(seed)
    for source in (src, int(rng.integers(0, n))):
        for k in (0, 1, 2, 3, 5, 1 << 40):
            if k <= 5:
                want = O.khop(ref, source, k)
            else:  # past the diameter: every vertex BFS reaches
                want = (O.bfs_levels(ref, source) >= 0).astype(np.uint8)
            got = A.khop(G, source, k).numpy()
            assert np.array_equal(got, want), (source, k)
            if k <= 5:
                assert np.array_equal(A.khop(P, source, k).numpy(), want), (source, k)


def test_edgeless_and_loop_only_graphs(A):
    """No edges, or only self loops that the build drops: every search stays at
    its source, and a source outside the graph is rejected."""
    n = 1000
    none = np.zeros(0, np.uint32)
    loops = np.arange(0, n, 7, dtype=np.uint32)
    for s, d in ((none, none), (loops, loops)):
        G = A.graph_from_edges(n, A.Vector.from_numpy(s), A.Vector.from_numpy(d))
        assert G.nnz == 0
        lev, nl, reached = A.bfs(G, 5)
        want = np.full(n, -1, np.int32)
        want[5] = 0
        assert np.array_equal(lev.numpy(), want) and nl == 1 and reached == 1
        one = np.zeros(n, np.uint8)
        one[5] = 1
        for k in (0, 1, 4):
            assert np.array_equal(A.khop(G, 5, k).numpy(), one), k
        with pytest.raises(A.AlpError):
            A.khop(G, n, 1)
        with pytest.raises(A.AlpError):
            A.bfs(G, n)
        # every vertex dangling: the teleport term alone, uniform ranks
        rc, pr, it, q = A.pagerank(G, 0.85, 1e-9, 1000)
        z = O.CSR(n, n, np.zeros(n + 1, np.uint64), np.zeros(1, np.uint32), None)
        orc, opr, oit, _ = O.pagerank(z, z, 0.85, 1e-9, 1000)
        assert rc == orc == A.SUCCESS and abs(it - oit) <= 1
        assert np.max(np.abs(pr.numpy() - 1.0 / n)) <= 1e-15 and np.max(np.abs(opr - 1.0 / n)) <= 1e-15
