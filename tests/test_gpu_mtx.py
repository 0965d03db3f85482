"""MatrixMarket files built on the device (alp_hip_matrix_read_mtx).

The device build from a file must equal the oracle's buildMatrixUnique restatement
of the same entries in the reader's emission order, bit for bit (CRS and CCS
arrays), and its product must equal scipy's on the same file."""
import os

import numpy as np
import pytest
import scipy.io

import oracle as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def A():
    import alp_amd
    alp_amd.init(0)
    return alp_amd


def random_file(path, n, density, symmetric, pattern, seed):
    rng = np.random.default_rng(seed)
    cells = set()
    want = int(density * n * n)
    while len(cells) < want:
        i, j = (int(v) for v in rng.integers(0, n, 2))
        if symmetric and j > i:
            i, j = j, i
        cells.add((i, j))
    cells = list(cells)
    rng.shuffle(cells)
    field = "pattern" if pattern else "real"
    sym = "symmetric" if symmetric else "general"
    with open(path, "w") as f:
        f.write(f"%%MatrixMarket matrix coordinate {field} {sym}\n% generated\n{n} {n} {len(cells)}\n")
        for i, j in cells:
            f.write(f"{i + 1} {j + 1}\n" if pattern else f"{i + 1} {j + 1} {rng.standard_normal():.17g}\n")


@pytest.mark.parametrize("symmetric", [False, True])
def test_device_build_equals_oracle_build(A, tmp_path, symmetric):
    p = os.path.join(str(tmp_path), "m.mtx")
    random_file(p, 300, 0.02, symmetric, False, 3 + symmetric)
    nrows, ncols, I, J, V, _ = A.mtx_read(p)
    M = A.Matrix.from_mtx(p, flags=A.MAT_WITH_CCS)
    for which in (0, 1):
        ref = O.coo_to_compressed(nrows, ncols, I, J, V, which=which)
        off, idx, val = M.download(which)
        assert np.array_equal(off, ref.off) and np.array_equal(idx, ref.idx) and np.array_equal(val, ref.val)
    S = scipy.io.mmread(p).tocsr()
    x = np.random.default_rng(9).standard_normal(ncols)
    y = A.Vector(nrows, fill=0.0)
    assert A.mxv(y, M, A.Vector.from_numpy(x)) == A.SUCCESS
    want = S @ x
    assert np.max(np.abs(y.numpy() - want)) <= 1e-12 * np.max(np.abs(want))


def test_pattern_file_as_pattern_and_as_values(A, tmp_path):
    p = os.path.join(str(tmp_path), "p.mtx")
    random_file(p, 200, 0.03, True, True, 5)
    nrows, ncols, I, J, _, pat = A.mtx_read(p)
    assert pat
    P = A.Matrix.from_mtx(p, value_type=A.VAL_PATTERN)
    ref = O.coo_to_compressed(nrows, ncols, I, J, None, which=0)
    off, idx, _ = P.download(0, with_values=False)
    assert np.array_equal(off, ref.off) and np.array_equal(idx, ref.idx)
    # as FP64 values, a pattern file's entries are 1 (the iterator's pattern value)
    M = A.Matrix.from_mtx(p)
    _, _, val = M.download(0)
    assert np.all(val == 1.0)
    y = A.Vector(nrows, fill=0.0)
    assert A.mxv(y, M, A.Vector(ncols, fill=1.0)) == A.SUCCESS
    assert np.array_equal(y.numpy(), np.diff(ref.off).astype(np.float64))
