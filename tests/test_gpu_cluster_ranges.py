"""The block path's own cluster ranges, read back from the library after a
solve, against the ranges of the reference's ipt_target_degenerate_subspaces
(IPT/src/ipt_preparation.cu:501-645, compiled from source; recorded in
tests/golden/cluster_ranges.json).  Index work: compared exactly, through the
dense device-scan path, the host CSC path and the device CSR path."""
import numpy as np
import pytest

from conftest import load_golden
from oracle import oracle as orc

pytestmark = pytest.mark.gpu

from block_ipt_amd import api  # noqa: E402
from block_ipt_amd.abi import IptError  # noqa: E402

CASES = load_golden("cluster_ranges.json")["cases"]


@pytest.fixture
def block_env(monkeypatch):
    monkeypatch.setenv("IPT_BLOCK_CLUSTER", "1")
    for v in ("IPT_DAVIDSON_ENRICH", "IPT_DAVIDSON_STEPS", "IPT_BLOCK_CLUSTER_TOL", "IPT_PRECISION",
              "IPT_COUPLING_GAP_RATIO_TOL", "IPT_MAX_LIFT_BLOCK_SIZE", "IPT_BLOCK_CLUSTER_OVERSAMPLE",
              "IPT_BLOCK_CLUSTER_ADAPTIVE", "IPT_DEGENERACY_THRESHOLD"):
        monkeypatch.delenv(v, raising=False)
    # one block step is enough to read the ranges back
    monkeypatch.setenv("IPT_BLOCK_CLUSTER_MAXITER", "1")
    return monkeypatch


def _matrix(case):
    if case["kind"] == "eps":
        return orc.eps_dense(case["n"], case["vscale"], case["seed"])
    return orc.rand_dense(case["n"], case["seed"], case["lo"], case["hi"], case["vscale"])


def _csc(H):
    n = H.shape[0]
    cols, rows = np.nonzero(H.T)
    cp = np.zeros(n + 1, dtype=np.int32)
    cp[1:] = np.cumsum(np.bincount(cols, minlength=n))
    return cp, rows.astype(np.int32), np.ascontiguousarray(H.T[cols, rows])


def _check(case, got):
    n, k = case["n"], case["k"]
    multi = [r for r in got if r[1] > r[0]]
    assert multi == [tuple(r) for r in case["ranges"]], (case["name"], multi)
    if not case["too_large"]:
        cov = np.zeros(n, int)
        for a, b in got:
            cov[a:b + 1] += 1
        assert np.all(cov[:k] == 1) and np.all(cov <= 1)


def _solve(fn, case):
    print("cluster ranges: %s" % case["name"], flush=True)
    try:
        fn()
    except IptError as e:
        # a target block wider than the block-size limit stops the solve at
        # detection with "invalid argument", as in the reference
        assert case["too_large"] and e.status == 1, (case["name"], e.status)
        return
    assert not case["too_large"], case["name"]


@pytest.mark.parametrize("case", CASES, ids=lambda c: c["name"])
def test_dense_device_ranges(block_env, case):
    H = _matrix(case)
    _solve(lambda: api.ipt_cuda_tol(np.asfortranarray(H), case["k"], 1e-12, 1), case)
    _check(case, api.last_cluster_ranges())


@pytest.mark.parametrize("case", [c for c in CASES if c["n"] <= 4096], ids=lambda c: c["name"])
def test_csc_host_ranges(block_env, case):
    H = _matrix(case)
    cp, ri, v = _csc(H)
    _solve(lambda: api.ipt_cuda_sparse_csc_tol(cp, ri, v, case["k"], 1e-12, 1), case)
    _check(case, api.last_cluster_ranges())


@pytest.mark.parametrize("case", [c for c in CASES if c["n"] <= 4096], ids=lambda c: c["name"])
def test_csr_device_ranges(block_env, case):
    H = _matrix(case)
    cp, ri, v = _csc(H)
    hip = api.hip()
    bufs = [hip.to_device(cp), hip.to_device(ri), hip.to_device(v)]
    try:
        def run():
            dv, dl, _, _ = api.ipt_cuda_sparse_csc_device_tol(bufs[0], bufs[1], bufs[2], case["n"], case["k"],
                                                              len(v), 1e-12, 1)
            api.free_device_result(dv, dl)
        _solve(run, case)
    finally:
        for b in bufs:
            hip.free(b)
    _check(case, api.last_cluster_ranges())
