"""A determinant-space Hamiltonian built from integrals (ipt_hip_fci_build_csc),
stored and reloaded through the reference's CSC cache format, then solved by
the sparse block path on the GPU: the lowest pairs equal LAPACK's."""
import numpy as np
import pytest
import scipy.linalg as sl

from oracle import fci as ofci
from oracle import oracle as orc

pytestmark = pytest.mark.gpu

from block_ipt_amd import api  # noqa: E402


@pytest.mark.parametrize("norb,na,nb,k", [(8, 3, 3, 8), (9, 3, 2, 12)])
def test_fci_hamiltonian_solved_on_gpu(monkeypatch, tmp_path, norb, na, nb, k):
    monkeypatch.setenv("IPT_BLOCK_CLUSTER", "1")
    monkeypatch.setenv("IPT_DAVIDSON_ENRICH", "1")
    monkeypatch.setenv("IPT_DAVIDSON_STEPS", "40")
    monkeypatch.setenv("IPT_DAVIDSON_CONVERGED_TOL", "1e-11")
    for v in ("IPT_PRECISION", "IPT_BLOCK_CLUSTER_STALL", "IPT_DAVIDSON_MAX_BASIS", "IPT_DAVIDSON_RETAIN",
              "IPT_COUPLING_GAP_RATIO_TOL", "IPT_BLOCK_CLUSTER_MAXITER"):
        monkeypatch.delenv(v, raising=False)
    h1, eri = ofci.random_integrals(norb, 21)
    path = str(tmp_path / "FCIDUMP")
    api.fcidump_write(path, norb, na + nb, na - nb, h1, eri, -2.0)
    no, ne, ms2, h1r, erir, ec = api.fcidump_read(path)
    m = api.fci_build_csc(no, (ne + ms2) // 2, (ne - ms2) // 2, h1r, erir, ec)
    cache = str(tmp_path / "h.csc")
    api.csc_cache_write(cache, m)
    c = api.csc_cache_read(cache)
    n = c["n"]
    print("FCI n=%d nnz=%d" % (n, c["nnz"]), flush=True)
    vec, val, info = api.ipt_cuda_sparse_csc_tol(c["col_ptr"], c["row_ind"], c["values"], k, 1e-12, 200)
    H = np.zeros((n, n))
    cp, ri, v = c["col_ptr"], c["row_ind"], c["values"]
    for j in range(n):
        H[ri[cp[j]:cp[j + 1]], j] = v[cp[j]:cp[j + 1]]
    w = sl.eigh(H, eigvals_only=True, subset_by_index=[0, k - 1])
    order = np.argsort(val)
    assert np.abs(val[order] - w).max() <= 1e-10 * max(1.0, np.abs(w).max())
    assert orc.residuals(vec, val, dense=H).max() <= 1e-10


def test_comparator_all_methods(tmp_path):
    """The comparator harness on one built Hamiltonian: block IPT, Davidson
    alone and ARPACK reach the same pairs with residuals under the
    threshold, and the report names each method's time."""
    from block_ipt_amd import compare
    m = api.fci_build_csc(8, 3, 3, *ofci.random_integrals(8, 21), -2.0)
    res = compare.compare(m, 8, 1e-10)
    assert [r["method"] for r in res] == ["ipt", "davidson", "arpack"]
    for r in res:
        assert r["status"] == "success", r
    w0 = np.asarray(res[0]["eigenvalues"])
    for r in res[1:]:
        assert np.abs(np.asarray(r["eigenvalues"]) - w0).max() <= 1e-9
    text = open(compare.write_report(res, str(tmp_path), "fci", 8, 1e-10, "test")).read()
    assert "davidson_speedup_vs_ipt=" in text and "arpack_speedup_vs_ipt=" in text
