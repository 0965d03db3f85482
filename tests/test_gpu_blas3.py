"""grb::mxm on the device (alp_hip_mxm) against ALP's reference grb::mxm fixtures
(reference/blas3.hpp:1148-1725; tests/golden/mxm.npz) and the oracle's
restatement of mxm_generic, which those fixtures pin bit for bit: each output
row's column order (first-touch), values (the reference's fold order) and the
ascending-row CCS."""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu

CASES = ["plus_times", "plus_times_inplace", "min_plus", "min_plus_inplace", "pattern",
         "transpose_left", "transpose_right", "wide"]


@pytest.fixture(scope="module")
def A():
    import alp_amd
    alp_amd.init(0)
    return alp_amd


def _operands(A, g, name):
    m, k, n, ring, tr, inplace, seed = (int(v) for v in g[f"{name}_meta"])
    ar, ac = (k, m) if tr == 1 else (m, k)
    br, bc = (n, k) if tr == 2 else (k, n)
    vt = A.VAL_PATTERN if ring == 2 else A.VAL_F64
    # the CCS built with the matrix follows the reference's order for a transposed operand
    Ma = A.Matrix.from_coo(ar, ac, g[f"{name}_IA"], g[f"{name}_JA"], None if ring == 2 else g[f"{name}_VA"],
                           value_type=vt, flags=A.MAT_WITH_CCS)
    Mb = A.Matrix.from_coo(br, bc, g[f"{name}_IB"], g[f"{name}_JB"], None if ring == 2 else g[f"{name}_VB"],
                           value_type=vt, flags=A.MAT_WITH_CCS)
    Mc = A.Matrix.from_coo(m, n, g[f"{name}_IC"], g[f"{name}_JC"], g[f"{name}_VC"], value_type=A.VAL_F64,
                           flags=A.MAT_WITH_CCS) if inplace else None
    descr = A.TRANSPOSE_LEFT if tr == 1 else A.TRANSPOSE_RIGHT if tr == 2 else 0
    ringid = A.MIN_PLUS_F64 if ring == 1 else A.PLUS_TIMES_F64
    return Ma, Mb, Mc, ringid, descr


@pytest.mark.parametrize("name", CASES)
def test_mxm_matches_reference_fixtures(A, golden, name):
    g = golden("mxm")
    Ma, Mb, Mc, ring, descr = _operands(A, g, name)
    C = A.mxm(Ma, Mb, ring, descr, C=Mc)
    off, idx, val = C.download(0)
    assert np.array_equal(off, g[f"{name}_crs_off"]) and np.array_equal(idx, g[f"{name}_crs_idx"])
    assert np.array_equal(val, g[f"{name}_crs_val"])
    coff, cidx, cval = C.download(1)
    assert np.array_equal(coff, g[f"{name}_ccs_off"]) and np.array_equal(cidx, g[f"{name}_ccs_idx"])
    assert np.array_equal(cval, g[f"{name}_ccs_val"])


def test_mxm_lorland_pattern_structure(A, golden):
    """lor-land over pattern operands: the pattern of the reference's product of
    the same void matrices, in the same order."""
    g = golden("mxm")
    Ma, Mb, _, _, _ = _operands(A, g, "pattern")
    C = A.mxm(Ma, Mb, A.LOR_LAND_BOOL)
    off, idx, _ = C.download(0, with_values=False)
    assert np.array_equal(off, g["pattern_crs_off"]) and np.array_equal(idx, g["pattern_crs_idx"])


def _rmat_csr(A, scale, ef, seed, values):
    n = 1 << scale
    s, d = A.rmat_edges(scale, ef * n, seed=seed)
    G = A.graph_from_edges(n, s, d, flags=A.GRAPH_DROP_SELF_LOOPS)
    off, idx, _ = G.download(0, with_values=False)
    val = np.random.default_rng(seed).standard_normal(len(idx)) if values else None
    return O.CSR(n, n, off, idx, val)


@pytest.mark.parametrize("ring", ["plus_times", "lor_land", "min_plus"])
def test_mxm_rmat_square_against_oracle(A, ring):
    """A directed R-MAT graph (scale 11, edge factor 8) squared: bit-exact with the
    oracle's mxm_generic, forced through many blocks of rows too (a small product
    budget) with identical results."""
    import os
    values = ring != "lor_land"
    R = _rmat_csr(A, 11, 8, 3, values)
    M = A.Matrix.from_crs(R.nrows, R.ncols, R.off, R.idx, R.val)
    rid = {"plus_times": A.PLUS_TIMES_F64, "lor_land": A.LOR_LAND_BOOL, "min_plus": A.MIN_PLUS_F64}[ring]
    want = O.mxm(R, R, {"plus_times": 0, "lor_land": 2, "min_plus": 1}[ring])
    outs = []
    for budget in (None, "4096"):
        if budget:
            os.environ["ALP_HIP_MXM_BUDGET"] = budget
        try:
            C = A.mxm(M, M, rid)
        finally:
            os.environ.pop("ALP_HIP_MXM_BUDGET", None)
        off, idx, val = C.download(0, with_values=values)
        assert np.array_equal(off, want.off) and np.array_equal(idx, want.idx), budget
        if values:
            assert np.array_equal(val, want.val), budget
        outs.append((off, idx, val))
    print(f"R-MAT^2 {ring}: {want.nnz} entries", flush=True)


def test_mxm_shifted_identities_and_errors(A):
    """tests/unit/mxm.cpp's closed forms: eye(n,n,1,1) eye(n,n,2,2) = eye(n,n,2,3),
    and the in-place update doubles it; dimension and type mismatches are
    rejected."""
    n = 100
    def eye(v, o):
        I = np.arange(n - o, dtype=np.uint32)
        return A.Matrix.from_coo(n, n, I, I + o, np.full(n - o, float(v)))
    C = A.mxm(eye(1, 1), eye(2, 2))
    off, idx, val = C.download(0)
    I = np.arange(n - 3)
    assert np.array_equal(idx, (I + 3).astype(np.uint32)) and np.all(val == 2.0)
    assert list(np.diff(off.astype(np.int64))) == [1] * (n - 3) + [0] * 3
    C = A.mxm(eye(1, 1), eye(2, 2), C=C)
    _, idx2, val2 = C.download(0)
    assert np.array_equal(idx2, idx) and np.all(val2 == 4.0)
    with pytest.raises(A.AlpError, match="MISMATCH"):
        A.mxm(A.Matrix.from_coo(3, 4, [0], [0], [1.0]), A.Matrix.from_coo(3, 4, [0], [0], [1.0]))
    P = A.Matrix.from_coo(4, 4, [0], [1])
    with pytest.raises(A.AlpError, match="MISMATCH"):
        A.mxm(A.Matrix.from_coo(4, 4, [0], [0], [1.0]), P, A.LOR_LAND_BOOL)
