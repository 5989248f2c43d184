"""The reference's own contraction identities (fricc/tests/test_sparse_tensors.py:48-269,
test_sparse_tensors_large.py:18-42) applied to the oracle's restatement of
its loop nests (src/sparse_tensor.cpp:11-346).  CPU only."""
import numpy as np
import pytest

from oracle import liboracle as lo

np.random.seed(20)


def create_sparse_ndarray(t_shape, nnz):
    size = np.prod(t_shape)
    tensor = np.zeros(size, order="C")
    idx = np.random.choice(np.arange(size), size=int(nnz))
    tensor[idx] = np.random.rand(int(nnz))
    return tensor.reshape(t_shape)


CASES = {
    "0101": (lambda no, nv: (no, no, no, no), "klij,klab->ijab", 1.0, lambda no, nv: (no, no, nv, nv)),
    "2323": (lambda no, nv: (nv, nv, nv, nv), "abcd,ijcd->ijab", 1.0, lambda no, nv: (no, no, nv, nv)),
    "1302": (lambda no, nv: (nv, no, no, nv), "akic,kjcb->ijab", 2.0, lambda no, nv: (no, no, nv, nv)),
    "1202": (lambda no, nv: (nv, no, nv, no), "akci,kjcb->ijab", -1.0, lambda no, nv: (no, no, nv, nv)),
    "1303": (lambda no, nv: (nv, no, no, nv), "akic,kjbc->ijab", 1.0, lambda no, nv: (no, no, nv, nv)),
    "1203": (lambda no, nv: (nv, no, nv, no), "bkci,kjac->ijab", 1.0, lambda no, nv: (no, no, nv, nv)),
    "1323": (lambda no, nv: (no, nv, no, nv), "kcld,ijcd->klij", 1.0, lambda no, nv: (no, no, no, no)),
    "0112": (lambda no, nv: (no, nv, no, nv), "ldkc,ilda->akic", -0.5, lambda no, nv: (nv, no, no, nv)),
    "0113": (lambda no, nv: (no, nv, no, nv), "ldkc,ilad->akic", 1.0, lambda no, nv: (nv, no, no, nv)),
    "0313": (lambda no, nv: (no, nv, no, nv), "lckd,ilad->akic", -0.5, lambda no, nv: (nv, no, no, nv)),
    "0312": (lambda no, nv: (no, nv, no, nv), "lckd,ilda->akci", -0.5, lambda no, nv: (nv, no, nv, no)),
}


@pytest.mark.parametrize("term", sorted(CASES))
@pytest.mark.parametrize("no,nv,frac", [(4, 8, 0.1), (5, 10, 0.01), (10, 20, 0.001)])
def test_oracle_contraction_identity(term, no, nv, frac):
    wshape, spec, alpha, oshape = CASES[term]
    w = np.ascontiguousarray(np.random.rand(*wshape(no, nv)))
    t_shape = (no, no, nv, nv)
    nnz = int(np.prod(t_shape) * frac)
    t2 = create_sparse_ndarray(t_shape, nnz)
    m = int(t2.size * frac)
    idx, val, _ = lo.compress(t2.ravel(), m, lo.LARGEST)
    ref = alpha * np.einsum(spec, w, t2, order="C")
    out = np.zeros(oshape(no, nv), order="C")
    lo.contract_ref(term, w, idx, val, t_shape, out)
    np.testing.assert_almost_equal(out, ref)


def test_oracle_strided_operands():
    """W and the output may be arbitrary strided views (src/sparse_tensor.cpp:353-354)."""
    no, nv = 5, 7
    t2 = create_sparse_ndarray((no, no, nv, nv), 120)
    idx, val, _ = lo.compress(t2.ravel(), 120, lo.LARGEST)
    w = np.random.rand(no, no, nv, nv).transpose(3, 1, 0, 2)  # (nv, no, no, nv) view
    out = np.zeros((nv, nv, no, no)).transpose(2, 3, 0, 1)  # (no, no, nv, nv) view
    lo.contract_ref("1302", w, idx, val, t2.shape, out)
    np.testing.assert_allclose(out, 2 * np.einsum("akic,kjcb->ijab", w, t2), rtol=1e-13, atol=1e-13)


def test_unknown_term_raises():
    w = np.zeros((2, 2, 2, 2))
    with pytest.raises(ValueError):
        lo.contract_ref("9999", w, np.zeros(1, dtype=np.int64), np.zeros(1), (2, 2, 2, 2), np.zeros((2, 2, 2, 2)))
