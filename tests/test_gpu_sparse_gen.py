"""Device generator of the orbital-sum sparse model against the oracle's
sequential restatement: identical CSR structure, values equal up to the
last-bit differences of the device log1p/cos."""
import numpy as np
import pytest

from oracle import oracle as orc

pytestmark = pytest.mark.gpu

from block_ipt_amd import api  # noqa: E402


def _device_csr(n, m, seed, **kw):
    hip = api.hip()
    cap = n * (2 * m + 1)
    drp, dci, dv = hip.malloc((n + 1) * 4), hip.malloc(cap * 4), hip.malloc(cap * 8)
    try:
        nnz = api.generate_orbital_csr(n, m, seed, drp, dci, dv, cap, **kw)
        rp = hip.to_host(drp, (n + 1,), dtype=np.int32)
        ci = hip.to_host(dci, (nnz,), dtype=np.int32)
        v = hip.to_host(dv, (nnz,))
    finally:
        for p in (drp, dci, dv):
            hip.free(p)
    return rp, ci, v


@pytest.mark.parametrize("n,m,seed,kw", [
    (3000, 20, 1, {}),
    (600, 400, 4, {}),                       # many collisions and mirrored duplicates
    (70, 69, 2, {"norb": 12, "nsub": 3}),    # every off-diagonal column drawn
    (2, 1, 0, {}),
    (500, 0, 9, {"sigma": 0.5}),             # diagonal only
    (20000, 200, 3, {"elo": -5.0, "ehi": 1.0, "sigma": 0.1}),
])
def test_device_generator_matches_oracle(n, m, seed, kw):
    rp, ci, v = _device_csr(n, m, seed, **kw)
    orp, oci, ov = orc.orbital_csr(n, m, seed, **kw)
    assert np.array_equal(rp, orp)
    assert np.array_equal(ci, oci)
    assert np.abs(v - ov).max() <= 4e-16 * max(1.0, np.abs(ov).max())


def test_invalid_arguments():
    with pytest.raises(api.IptError):
        api.generate_orbital_csr(10, 10, 0, 1, 1, 1, 100)  # m > n - 1
    with pytest.raises(api.IptError):
        api.generate_orbital_csr(1000, 600, 0, 1, 1, 1, 10 ** 7)  # m > 512
