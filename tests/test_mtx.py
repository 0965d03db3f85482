"""CPU checks of the MatrixMarket reader (alp_hip_mtx_read, host code only).

Content is pinned by scipy's own MatrixMarket reader (scipy.io.mmread), an
implementation independent of this one. The emission order (file order, an
off-diagonal entry of a symmetric file followed by its mirror) is the
reference iterator's (utils/parser/matrixFileIterator.hpp:411-436); it decides
the within-row order of the device build."""
import os

import numpy as np
import pytest
import scipy.io

import alp_amd as A


def write(tmp_path, name, text):
    p = os.path.join(str(tmp_path), name)
    with open(p, "w") as f:
        f.write(text)
    return p


GENERAL = """%%MatrixMarket matrix coordinate real general
% a comment
%
4 5 6
1 1 2.5
4 5 -1e-3
2 3 7
3 1 0.25
1 4 -3.0
4 1 1e2
"""

# the same file with what readers in the wild accept beyond the strict format:
# an upper-case banner, blank lines and comments between the entries
GENERAL_LOOSE = """%%MATRIXMARKET Matrix Coordinate Real General
% a comment
4 5 6

1 1 2.5
4 5 -1e-3
2 3 7
% comment between entries
3 1 0.25
1 4 -3.0

4 1 1e2
"""

SYMMETRIC = """%%MatrixMarket matrix coordinate real symmetric
5 5 7
1 1 4.0
2 1 -1.5
3 3 4.0
5 2 0.5
4 4 4.0
5 5 4.0
4 3 -2.25
"""

PATTERN_SYM = """%%MatrixMarket matrix coordinate pattern symmetric
4 4 4
2 1
3 3
4 2
4 1
"""

INTEGER = """%%MatrixMarket matrix coordinate integer general
3 3 3
1 2 5
3 3 -7
2 1 12
"""


def dense_of(nrows, ncols, I, J, V):
    D = np.zeros((nrows, ncols))
    np.add.at(D, (I, J), V)
    return D


@pytest.mark.parametrize("text", [GENERAL, SYMMETRIC, PATTERN_SYM, INTEGER])
def test_content_matches_scipy(tmp_path, text):
    p = write(tmp_path, "m.mtx", text)
    nrows, ncols, I, J, V, pat = A.mtx_read(p)
    ref = scipy.io.mmread(p)
    ref = ref.toarray() if hasattr(ref, "toarray") else np.asarray(ref)
    assert (nrows, ncols) == ref.shape
    assert pat == ("pattern" in text.splitlines()[0].lower())
    assert np.array_equal(dense_of(nrows, ncols, I, J, V), ref.astype(np.float64))
    # the expansion adds exactly the off-diagonal mirrors
    assert len(I) == np.count_nonzero(ref) or pat


def test_lenient_layout_reads_like_the_strict_file(tmp_path):
    strict = A.mtx_read(write(tmp_path, "g.mtx", GENERAL))
    loose = A.mtx_read(write(tmp_path, "l.mtx", GENERAL_LOOSE))
    assert strict[:2] == loose[:2] and strict[5] == loose[5]
    for a, b in zip(strict[2:5], loose[2:5]):
        assert np.array_equal(a, b)


def test_emission_order_is_the_iterators(tmp_path):
    p = write(tmp_path, "s.mtx", SYMMETRIC)
    _, _, I, J, V, _ = A.mtx_read(p)
    # file entries (1-based) 1 1, 2 1, 3 3, 5 2, 4 4, 5 5, 4 3: each off-diagonal one
    # followed by its mirror
    want = [(0, 0), (1, 0), (0, 1), (2, 2), (4, 1), (1, 4), (3, 3), (4, 4), (3, 2), (2, 3)]
    assert list(zip(I.tolist(), J.tolist())) == want
    assert V.tolist() == [4.0, -1.5, -1.5, 4.0, 0.5, 0.5, 4.0, 4.0, -2.25, -2.25]
    _, _, I, J, V, pat = A.mtx_read(write(tmp_path, "g.mtx", GENERAL))
    assert list(zip(I.tolist(), J.tolist())) == [(0, 0), (3, 4), (1, 2), (2, 0), (0, 3), (3, 0)]
    assert not pat


@pytest.mark.parametrize("banner", [
    "%%MatrixMarket matrix coordinate complex general",
    "%%MatrixMarket matrix coordinate complex hermitian",
    "%%MatrixMarket matrix coordinate real skew-symmetric",
    "%%MatrixMarket matrix array real general",
])
def test_unsupported_kinds(tmp_path, banner):
    p = write(tmp_path, "u.mtx", banner + "\n2 2 1\n1 1 1.0 0.0\n")
    with pytest.raises(RuntimeError, match="UNSUPPORTED"):
        A.mtx_read(p)


@pytest.mark.parametrize("text", [
    "%%NotMatrixMarket matrix coordinate real general\n2 2 1\n1 1 1.0\n",
    "%%MatrixMarket matrix coordinate real general\n2 2 1\n3 1 1.0\n",      # row out of range
    "%%MatrixMarket matrix coordinate real general\n2 2 1\n0 1 1.0\n",      # 0 is not 1-based
    "%%MatrixMarket matrix coordinate real general\n2 2 2\n1 1 1.0\n",      # too few entries
    "%%MatrixMarket matrix coordinate real general\n2 2 1\n1 1 1.0\n2 2 1\n",  # too many
    "%%MatrixMarket matrix coordinate real general\n2 2 1\n1 1\n",          # missing value
    "%%MatrixMarket matrix coordinate real symmetric\n2 3 1\n1 1 1.0\n",    # symmetric, not square
    "%%MatrixMarket matrix coordinate real general\n2 2 5\n1 1 1.0\n",      # more entries than positions
])
def test_malformed_files_are_rejected(tmp_path, text):
    with pytest.raises(RuntimeError, match="ILLEGAL"):
        A.mtx_read(write(tmp_path, "bad.mtx", text))


def test_huge_entry_count_is_rejected_before_reserving(tmp_path):
    """A header announcing 2^60 entries is refused from the header alone (no
    reservation of that size, no exception out of the C ABI)."""
    text = "%%MatrixMarket matrix coordinate real general\n4294967295 4294967295 1152921504606846976\n1 1 1.0\n"
    with pytest.raises(RuntimeError, match="OVERFLW"):
        A.mtx_read(write(tmp_path, "huge.mtx", text))


def test_missing_file(tmp_path):
    with pytest.raises(RuntimeError, match="ILLEGAL"):
        A.mtx_read(os.path.join(str(tmp_path), "absent.mtx"))
