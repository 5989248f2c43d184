"""The reference's sparse-tensor test suite (fricc/tests/test_sparse_tensors.py,
test_sparse_tensors_large.py) run against this package's ``fricc.py_rccsd``
(GPU kernels through the C ABI).  Same shapes, seeds and assertions; numpy
in, numpy out, exactly as the reference's tests call its binding."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

npt = np.testing


@pytest.fixture(scope="module")
def py_rccsd():
    from fricc import py_rccsd as m

    return m


np.random.seed(20)


def create_sparse_ndarray(t_shape, nnz):
    size = np.prod(t_shape)
    tensor = np.zeros(size, order="C")
    idx = np.random.choice(np.arange(size), size=int(nnz))
    tensor[idx] = np.random.rand(int(nnz))
    return tensor.reshape(t_shape)


@pytest.mark.parametrize("no,nv,frac", [(4, 8, 0.1), (10, 15, 0.001), (100, 2, 0.001)])
def test_sparse_init(py_rccsd, no, nv, frac):
    a = np.ascontiguousarray(np.random.rand(no, no, nv, nv))
    m = int(a.size * frac)
    assert m <= 200
    a_compressed = py_rccsd.SparseTensor4d(a, a.shape, m, "largest")
    for mi in range(m):
        idx, value = a_compressed.get_element(mi)
        npt.assert_equal(value, a[tuple(idx)])


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


def _run(py_rccsd, term, no, nv, frac):
    wshape, spec, alpha, oshape = CASES[term]
    w = np.ascontiguousarray(np.random.rand(*wshape(no, nv)))
    t_shape = (no, no, nv, nv)
    nnz = int(np.prod(t_shape) * frac)
    t2 = create_sparse_ndarray(t_shape, nnz)
    m = int(t2.size * frac)
    t2_compressed = py_rccsd.SparseTensor4d(t2, t2.shape, m, "largest")
    ref = alpha * np.einsum(spec, w, t2, order="C")
    out = np.zeros(oshape(no, nv), order="C")
    py_rccsd.contract_DTSpT(w, t2_compressed, out, term)
    npt.assert_almost_equal(out, ref)
    return w, t2_compressed, t2


@pytest.mark.parametrize("term", sorted(CASES))
@pytest.mark.parametrize("no,nv,frac", [(4, 8, 0.1), (5, 10, 0.01), (10, 20, 0.001)])
def test_contraction(py_rccsd, term, no, nv, frac):
    _run(py_rccsd, term, no, nv, frac)


def test_2323_contraction_20_40(py_rccsd):
    _run(py_rccsd, "2323", 20, 40, 0.001)


@pytest.mark.parametrize("no,nv,frac", [(20, 40, 1e-4), (50, 90, 1e-6)])
def test_2323_contraction_large(py_rccsd, no, nv, frac):
    """fricc/tests/test_sparse_tensors_large.py:18-42 (90^4 FP64 W = 525 MB)."""
    _run(py_rccsd, "2323", no, nv, frac)


@pytest.mark.parametrize("term", sorted(CASES))
def test_contraction_dense_random_fp64_tight(py_rccsd, term):
    """Denser tensors than the reference's tests, checked to 1e-12 relative."""
    no, nv = 6, 13
    wshape, spec, alpha, oshape = CASES[term]
    rng = np.random.default_rng(hash(term) % 1000)
    w = rng.standard_normal(wshape(no, nv))
    t2 = rng.standard_normal((no, no, nv, nv))
    m = t2.size // 3
    T = py_rccsd.SparseTensor4d(t2, t2.shape, m, "largest")
    dense = np.zeros(t2.size)
    dense[T.indices()] = T.values()
    ref = alpha * np.einsum(spec, w, dense.reshape(t2.shape))
    out = rng.standard_normal(oshape(no, nv))
    out0 = out.copy()
    py_rccsd.contract_DTSpT(w, T, out, term)
    npt.assert_allclose(out - out0, ref, rtol=1e-12, atol=1e-12 * np.abs(ref).max())


@pytest.mark.parametrize("term", sorted(CASES))
@pytest.mark.parametrize("precision,tol", [("fp64", 1e-12), ("fp32", 2e-5)])
def test_contraction_lds_pipeline_shapes(py_rccsd, term, precision, tol):
    """o=8, v=40: every A width (64, 320, 1600 columns) is a whole number of
    16-byte chunks, so the FP64/FP32 products run the LDS-staged kernel, with
    partial last column blocks (1600 = 12.5 x 128 = 6.25 x 256) and many
    64-column tiles; checked against the einsum of the kept entries."""
    no, nv = 8, 40
    wshape, spec, alpha, oshape = CASES[term]
    rng = np.random.default_rng(hash(term) % 997 + 1)
    w = rng.standard_normal(wshape(no, nv))
    t2 = rng.standard_normal((no, no, nv, nv))
    T = py_rccsd.SparseTensor4d(t2, t2.shape, t2.size // 3, "largest")
    dense = np.zeros(t2.size)
    dense[T.indices()] = T.values()
    ref = alpha * np.einsum(spec, w, dense.reshape(t2.shape))
    out = rng.standard_normal(oshape(no, nv))
    out0 = out.copy()
    py_rccsd.contract_DTSpT(w, T, out, term, precision=precision)
    npt.assert_allclose(out - out0, ref, rtol=0, atol=tol * np.abs(ref).max())


@pytest.mark.parametrize("term", ["2323", "1302", "0312"])
def test_contraction_fp32(py_rccsd, term):
    no, nv = 6, 13
    wshape, spec, alpha, oshape = CASES[term]
    rng = np.random.default_rng(7)
    w = rng.standard_normal(wshape(no, nv))
    t2 = rng.standard_normal((no, no, nv, nv))
    T = py_rccsd.SparseTensor4d(t2, t2.shape, t2.size // 4, "largest")
    dense = np.zeros(t2.size)
    dense[T.indices()] = T.values()
    ref = alpha * np.einsum(spec, w, dense.reshape(t2.shape))
    out = np.zeros(oshape(no, nv))
    py_rccsd.contract_DTSpT(w, T, out, term, precision="fp32")
    # FP32 operands, FP32 accumulation over <= no*nv terms
    npt.assert_allclose(out, ref, rtol=0, atol=2e-5 * np.abs(ref).max())


@pytest.mark.parametrize("term", sorted(CASES))
def test_contraction_bf16_split(py_rccsd, term):
    """BF16 MFMA with hi/lo split operands: about 16 significant bits per
    product, FP32 accumulation."""
    no, nv = 7, 19
    wshape, spec, alpha, oshape = CASES[term]
    rng = np.random.default_rng(11)
    w = rng.standard_normal(wshape(no, nv))
    t2 = rng.standard_normal((no, no, nv, nv))
    T = py_rccsd.SparseTensor4d(t2, t2.shape, t2.size // 5, "largest")
    dense = np.zeros(t2.size)
    dense[T.indices()] = T.values()
    ref = alpha * np.einsum(spec, w, dense.reshape(t2.shape))
    out = np.zeros(oshape(no, nv))
    py_rccsd.contract_DTSpT(w, T, out, term, precision="bf16")
    npt.assert_allclose(out, ref, rtol=0, atol=2e-4 * np.abs(ref).max())


@pytest.mark.parametrize("term", sorted(CASES))
def test_contraction_bf16_single_pass(py_rccsd, term):
    """Single-pass BF16 (hi-only operand records): each operand rounded to
    BF16 (via FP32, round to nearest even) and the products accumulated in
    FP32, so the result equals the FP64 einsum of the rounded operands to
    FP32 accumulation error."""
    import torch

    no, nv = 7, 19
    wshape, spec, alpha, oshape = CASES[term]
    rng = np.random.default_rng(13)
    w = rng.standard_normal(wshape(no, nv))
    t2 = rng.standard_normal((no, no, nv, nv))
    T = py_rccsd.SparseTensor4d(t2, t2.shape, t2.size // 5, "largest")
    dense = np.zeros(t2.size)
    dense[T.indices()] = T.values()
    bf = lambda a: torch.from_numpy(a).float().to(torch.bfloat16).double().numpy()  # noqa: E731
    ref = alpha * np.einsum(spec, bf(w), bf(dense).reshape(t2.shape))
    out = np.zeros(oshape(no, nv))
    py_rccsd.contract_DTSpT(w, T, out, term, precision="bf16x1")
    npt.assert_allclose(out, ref, rtol=0, atol=2e-6 * np.abs(ref).max())
    # and it is a BF16 result: far from the FP64 one
    exact = alpha * np.einsum(spec, w, dense.reshape(t2.shape))
    assert np.abs(out - exact).max() > 1e-4 * np.abs(exact).max()


def test_strided_views(py_rccsd):
    """W and ouput may be arbitrary strided numpy views (src/sparse_tensor.cpp:353-354)."""
    no, nv = 5, 9
    rng = np.random.default_rng(3)
    t2 = rng.standard_normal((no, no, nv, nv))
    T = py_rccsd.SparseTensor4d(t2, t2.shape, 300, "largest")
    dense = np.zeros(t2.size)
    dense[T.indices()] = T.values()
    w = rng.standard_normal((no, no, nv, nv)).transpose(3, 1, 0, 2)
    out = np.zeros((nv, nv, no, no)).transpose(2, 3, 0, 1)
    py_rccsd.contract_DTSpT(w, T, out, "1302")
    npt.assert_allclose(out, 2 * np.einsum("akic,kjcb->ijab", w, dense.reshape(t2.shape)), rtol=1e-12, atol=1e-12)


def test_experimental_terms(py_rccsd):
    no, nv = 4, 8
    w = np.random.rand(nv, nv, nv, nv)
    t2 = create_sparse_ndarray((no, no, nv, nv), 50)
    T = py_rccsd.SparseTensor4d(t2, t2.shape, 50, "largest")
    for mode in ("taskloop", "atomic"):
        out = np.zeros(t2.shape)
        py_rccsd.contract_DTSpT_experimental(w, T, out, mode)
        npt.assert_almost_equal(out, np.einsum("abcd,ijcd->ijab", w, t2))


def test_errors_raise_instead_of_exit(py_rccsd):
    t2 = np.random.rand(3, 3, 4, 4)
    with pytest.raises(ValueError):
        py_rccsd.SparseTensor4d(t2, t2.shape, 10, "bogus")
    with pytest.raises(ValueError):
        py_rccsd.SparseTensor4d(t2, t2.shape, 10, "fri", "bogus")
    with pytest.raises(TypeError):
        py_rccsd.SparseTensor4d(t2.astype(np.float32), t2.shape, 10, "largest")
    with pytest.raises(TypeError):
        py_rccsd.SparseTensor4d(np.asfortranarray(t2), t2.shape, 10, "largest")
    T = py_rccsd.SparseTensor4d(t2, t2.shape, 10, "largest")
    with pytest.raises(ValueError):
        py_rccsd.contract_DTSpT(np.zeros((4, 4, 4, 4)), T, np.zeros(t2.shape), "9999")
    with pytest.raises(ValueError):
        py_rccsd.contract_DTSpT(np.zeros((3, 3, 3)), T, np.zeros(t2.shape), "2323")
    with pytest.raises(ValueError):  # W dims disagree with T
        py_rccsd.contract_DTSpT(np.zeros((5, 5, 5, 5)), T, np.zeros(t2.shape), "2323")
    with pytest.raises(IndexError):
        T.get_element(10)


def test_empty_constructor(py_rccsd):
    T = py_rccsd.SparseTensor4d([2, 3, 4, 5], 7)
    assert T.nnz == 7
    assert T.get_element(6) == ([0, 0, 0, 0], 0.0)


@pytest.mark.parametrize("precision", ["fp64", "fp32", "bf16"])
@pytest.mark.parametrize("term", ["2323", "1302", "0112"])
def test_repeated_indices_add_up(py_rccsd, term, precision):
    """Entries set by the caller may repeat a flat index; every contraction
    adds them (the reference loops over the entries, src/sparse_tensor.cpp:63-77),
    so the result is that of the dense tensor with repeats summed."""
    no, nv = 5, 11
    wshape, spec, alpha, oshape = CASES[term]
    rng = np.random.default_rng(5)
    w = rng.standard_normal(wshape(no, nv))
    size = no * no * nv * nv
    idx = rng.integers(0, size, 400)
    idx[:50] = idx[50:100]  # 50 repeats
    idx[100:110] = 0
    val = rng.standard_normal(idx.size)
    T = py_rccsd.SparseTensor4d([no, no, nv, nv], idx.size)
    T.set_entries(idx, val)
    dense = np.zeros(size)
    np.add.at(dense, idx, val)
    ref = alpha * np.einsum(spec, w, dense.reshape(no, no, nv, nv))
    out = np.zeros(oshape(no, nv))
    py_rccsd.contract_DTSpT(w, T, out, term, precision=precision)
    tol = {"fp64": 1e-12, "fp32": 2e-5, "bf16": 2e-4}[precision]
    npt.assert_allclose(out, ref, rtol=0, atol=tol * np.abs(ref).max())


@pytest.mark.parametrize("width", [128, 256, 512])
@pytest.mark.parametrize("term", ["2323", "1302", "0112"])
def test_bf16_view_widths(py_rccsd, term, width):
    """The BF16 kernels at every view (tile) width the automatic rule picks:
    128 (256-wide blocks of 128 columns), 256 and 512 (256x256 blocks, one or
    two per tile), on shapes with several tiles and a ragged last one."""
    from fricc import _capi

    no, nv = 23, 12
    wshape, spec, alpha, oshape = CASES[term]
    rng = np.random.default_rng(13)
    w = rng.standard_normal(wshape(no, nv))
    t2 = rng.standard_normal((no, no, nv, nv))
    T = py_rccsd.SparseTensor4d(t2, t2.shape, t2.size // 6, "largest")
    dense = np.zeros(t2.size)
    dense[T.indices()] = T.values()
    ref = alpha * np.einsum(spec, w, dense.reshape(t2.shape))
    out = np.zeros(oshape(no, nv))
    _capi.check(_capi.lib().fricc_set_bf16_view_width(width))
    try:
        py_rccsd.contract_DTSpT(w, T, out, term, precision="bf16")
    finally:
        _capi.check(_capi.lib().fricc_set_bf16_view_width(0))
    npt.assert_allclose(out, ref, rtol=0, atol=2e-4 * np.abs(ref).max())
