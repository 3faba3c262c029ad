"""Determinant-space (full CI) Hamiltonian, FCIDUMP files and the binary CSC
cache (include/ipt_hip_fci.h; host code, no GPU).

The builder has no reference output to compare with (the reference takes
its integrals from PySCF at run time and ships no integral files): it is
checked against a restatement of the reference's own link-based algorithm
(oracle/fci.py) and against what the mathematics fixes -- invariance of the
spectrum under orbital rotations, one electron giving eig(h1), and the
closed form of minimal-basis H2.  The cache format is checked byte for byte
against the reference's header layout and round-tripped on the committed
H2O CSC fixture."""
import os
import struct

import numpy as np
import pytest
import scipy.linalg as sl

from conftest import load_golden
from oracle import fci as ofci

from block_ipt_amd import api
from block_ipt_amd.abi import IptError


def _dense(csc):
    n = csc["n"]
    H = np.zeros((n, n))
    cp, ri, v = csc["col_ptr"], csc["row_ind"], csc["values"]
    for c in range(n):
        H[ri[cp[c]:cp[c + 1]], c] = v[cp[c]:cp[c + 1]]
    return H


@pytest.mark.parametrize("norb,na,nb,seed", [(4, 2, 2, 1), (5, 2, 2, 2), (5, 3, 2, 3), (6, 3, 3, 4), (5, 1, 0, 5),
                                             (4, 4, 1, 6)])
def test_builder_equals_reference_algorithm(norb, na, nb, seed):
    h1, eri = ofci.random_integrals(norb, seed)
    m = api.fci_build_csc(norb, na, nb, h1, eri, -1.25)
    Href, perm = ofci.fci_dense(norb, na, nb, h1, eri, -1.25)
    assert m["n"] == Href.shape[0]
    H = _dense(m)
    assert np.abs(H - Href).max() <= 1e-13 * max(1.0, np.abs(Href).max())
    assert np.array_equal(m["perm"], perm)
    # CSC conventions: rows ascending, no stored zeros, diagonal field
    cp, ri = m["col_ptr"], m["row_ind"]
    for c in range(m["n"]):
        assert np.all(np.diff(ri[cp[c]:cp[c + 1]]) > 0)
    assert np.all(m["values"] != 0.0)
    assert np.array_equal(m["diagonal"], np.diag(H))
    assert np.all(np.diff(m["diagonal"]) >= 0)  # sorted
    assert abs(m["matrix_inf_norm"] - np.abs(H).sum(axis=1).max()) <= 1e-12 * m["matrix_inf_norm"]
    assert np.allclose(H, H.T, atol=1e-14)
    assert (m["na"], m["nb"], m["norb"], m["neleca"], m["nelecb"]) == (
        len(ofci.strings(norb, na)), len(ofci.strings(norb, nb)), norb, na, nb)


@pytest.mark.parametrize("norb,na,nb", [(5, 2, 2), (6, 3, 2), (6, 2, 2)])
def test_spectrum_invariant_under_orbital_rotation(norb, na, nb):
    """A rotation of the orbitals is a unitary change of the determinant
    basis: the FCI eigenvalues do not move.  A wrong phase, a missing
    exchange term or a transposed integral breaks this."""
    h1, eri = ofci.random_integrals(norb, 11, coupling=0.2)
    U, _ = np.linalg.qr(np.random.default_rng(3).standard_normal((norb, norb)))
    h1r, erir = ofci.rotate_integrals(h1, eri, U)
    w0 = np.linalg.eigvalsh(_dense(api.fci_build_csc(norb, na, nb, h1, eri, 0.3)))
    w1 = np.linalg.eigvalsh(_dense(api.fci_build_csc(norb, na, nb, h1r, erir, 0.3)))
    assert np.abs(w0 - w1).max() <= 1e-11


def test_one_electron_is_h1():
    h1, eri = ofci.random_integrals(7, 4, coupling=0.3)
    m = api.fci_build_csc(7, 1, 0, h1, eri, 0.0, sort_by_diagonal=False)
    assert np.allclose(_dense(m), h1, atol=1e-15)


def test_minimal_basis_h2_closed_form():
    """H2 in a minimal basis (gerade orbital 1, ungerade 2; Szabo & Ostlund,
    Modern Quantum Chemistry, chapters 3-4: h11 -1.2528, h22 -0.4756,
    (11|11) 0.6746, (22|22) 0.6975, (11|22) 0.6636, (12|12) 0.1813 at
    R = 1.4 bohr).  By symmetry the singlet ground state mixes only |1a1b>
    and |2a2b>: its energy is the lower root of
    [[2 h11 + J11, K12], [K12, 2 h22 + J22]]; the nuclear repulsion 1/R
    enters through ecore."""
    h11, h22, J11, J22, J12, K12 = -1.2528, -0.4756, 0.6746, 0.6975, 0.6636, 0.1813
    h1 = np.diag([h11, h22])
    eri = np.zeros((2, 2, 2, 2))
    eri[0, 0, 0, 0], eri[1, 1, 1, 1] = J11, J22
    for a, b, c, d in ((0, 0, 1, 1), (1, 1, 0, 0)):
        eri[a, b, c, d] = J12
    for a, b, c, d in ((0, 1, 0, 1), (1, 0, 1, 0), (0, 1, 1, 0), (1, 0, 0, 1)):
        eri[a, b, c, d] = K12
    ecore = 1.0 / 1.4
    m = api.fci_build_csc(2, 1, 1, h1, eri, ecore)
    w = np.linalg.eigvalsh(_dense(m))
    A = np.array([[2 * h11 + J11, K12], [K12, 2 * h22 + J22]])
    assert abs(w[0] - (np.linalg.eigvalsh(A)[0] + ecore)) <= 1e-14
    assert abs(w[0] - (-1.1373)) <= 2e-4  # the textbook FCI energy
    # the open-shell pair: h11 + h22 + J12 -/+ K12 (singlet / triplet m_s = 0)
    assert np.allclose(sorted(w[1:3]), sorted([h11 + h22 + J12 - K12 + ecore, h11 + h22 + J12 + K12 + ecore]),
                       atol=1e-14)


def test_fcidump_round_trip(tmp_path):
    norb = 6
    h1, eri = ofci.random_integrals(norb, 8)
    path = str(tmp_path / "FCIDUMP")
    api.fcidump_write(path, norb, 6, 0, h1, eri, -3.5)
    no, ne, ms2, h1r, erir, ec = api.fcidump_read(path)
    assert (no, ne, ms2) == (norb, 6, 0) and ec == -3.5
    assert np.array_equal(h1r, h1) and np.array_equal(erir, eri)
    a = api.fci_build_csc(norb, 3, 3, h1, eri, -3.5)
    b = api.fci_build_csc(norb, 3, 3, h1r, erir, ec)
    assert np.array_equal(a["values"], b["values"]) and np.array_equal(a["row_ind"], b["row_ind"])


def test_fcidump_header_forms(tmp_path):
    """Namelist forms of other writers: keys on several lines, '/' as the
    terminator, MS2 absent (0), orbital-energy lines ignored."""
    path = tmp_path / "FCIDUMP"
    path.write_text(" &FCI NORB=  2,\n  NELEC=2,\n  ORBSYM=1,1,\n  ISYM=1\n /\n"
                    "  0.5 1 1 1 1\n  0.25 2 2 1 1\n  0.7 2 2 2 2\n  0.1 2 1 2 1\n"
                    " -1.0 1 1 0 0\n -0.5 2 2 0 0\n  0.01 2 1 0 0\n  -1.0 1 0 0 0\n  0.75 0 0 0 0\n")
    no, ne, ms2, h1, eri, ec = api.fcidump_read(str(path))
    assert (no, ne, ms2, ec) == (2, 2, 0, 0.75)
    assert np.array_equal(h1, [[-1.0, 0.01], [0.01, -0.5]])
    assert eri[0, 0, 1, 1] == eri[1, 1, 0, 0] == 0.25 and eri[0, 1, 0, 1] == eri[1, 0, 1, 0] == 0.1
    assert eri[1, 0, 0, 1] == 0.1 and eri[1, 1, 1, 1] == 0.7
    bad = tmp_path / "BAD"
    bad.write_text("&FCI NORB=2 NELEC=2\n")  # no terminator
    with pytest.raises(IptError):
        api.fcidump_read(str(bad))


def test_cache_round_trip_on_h2o_fixture(tmp_path):
    """The committed 10x10 H2O CSC (IPT/test/sparse/h20_w_csc.h) through the
    reference's cache format; header bytes at the reference's offsets."""
    g = load_golden("h2o_w_csc.json")
    n = g["n"]
    cp = np.asarray(g["col_ptr"], dtype=np.int32)
    ri = np.asarray(g["row_ind"], dtype=np.int32)
    v = np.asarray(g["values"])
    H = np.zeros((n, n))
    for c in range(n):
        H[ri[cp[c]:cp[c + 1]], c] = v[cp[c]:cp[c + 1]]
    csc = {"col_ptr": cp, "row_ind": ri, "values": v, "diagonal": np.diag(H).copy(),
           "matrix_inf_norm": np.abs(H).sum(axis=1).max(), "norb": 6, "nelec": 8, "neleca": 4, "nelecb": 4,
           "na": 15, "nb": 15}
    path = str(tmp_path / "h2o.csc")
    api.csc_cache_write(path, csc)
    raw = open(path, "rb").read()
    assert len(raw) == 56 + 4 * (n + 1) + 12 * len(v) + 8 * n
    magic, version, nn, nnz, norm, norb, nelec, na_, nb_, sa, sb = struct.unpack("<8sIii4xd6i", raw[:56])
    assert magic == b"IPTCH4C1" and version == 1 and (nn, nnz) == (n, len(v))
    assert norm == csc["matrix_inf_norm"] and (norb, nelec, na_, nb_, sa, sb) == (6, 8, 4, 4, 15, 15)
    off = 56
    assert np.array_equal(np.frombuffer(raw, np.int32, n + 1, off), cp)
    off += 4 * (n + 1)
    assert np.array_equal(np.frombuffer(raw, np.int32, len(v), off), ri)
    off += 4 * len(v)
    assert np.array_equal(np.frombuffer(raw, np.float64, len(v), off), v)
    back = api.csc_cache_read(path)
    for key in ("col_ptr", "row_ind", "values", "diagonal"):
        assert np.array_equal(back[key], csc[key])
    assert back["matrix_inf_norm"] == csc["matrix_inf_norm"] and back["perm"] is None
    assert not os.path.exists(path + ".tmp")


def test_cache_rejects_invalid_files(tmp_path):
    m = api.fci_build_csc(4, 2, 2, *ofci.random_integrals(4, 1), 0.0)
    path = tmp_path / "m.csc"
    api.csc_cache_write(str(path), m)
    raw = bytearray(path.read_bytes())
    for mutate in (lambda b: b.__setitem__(0, ord("X")),               # magic
                   lambda b: b.__setitem__(8, 2),                      # version
                   lambda b: b.__setitem__(slice(56, 60), b"\xff\xff\xff\x7f"),  # col_ptr[0]
                   lambda b: b.__delitem__(slice(len(b) - 8, len(b)))):  # truncated
        bad = bytearray(raw)
        mutate(bad)
        p = tmp_path / "bad.csc"
        p.write_bytes(bytes(bad))
        with pytest.raises(IptError):
            api.csc_cache_read(str(p))
    with pytest.raises(IptError):
        api.csc_cache_read(str(tmp_path / "missing.csc"))


def test_builder_rejects_invalid_sizes():
    import ctypes
    from block_ipt_amd import abi
    h1, eri = np.zeros(25 * 25), np.zeros(25 ** 4)
    for norb, na, nb in ((3, 4, 0), (3, -1, 1), (0, 0, 0), (25, 1, 1)):
        m = abi.IPTFciCsc()
        st = abi.lib().ipt_hip_fci_build_csc(norb, na, nb, h1.ctypes.data_as(abi._dp), eri.ctypes.data_as(abi._dp),
                                             0.0, 1, ctypes.byref(m))
        assert st == 1 and m.n == 0 and not m.col_ptr


def test_comparator_report_on_host(tmp_path):
    """The comparator harness (block_ipt_amd/compare.py) with the host
    method only: ARPACK's pairs of a built Hamiltonian equal LAPACK's and the
    report has the reference's key=value layout."""
    from block_ipt_amd import compare
    m = api.fci_build_csc(6, 3, 2, *ofci.random_integrals(6, 9), -1.0)
    res = compare.compare(m, 6, 1e-10, methods=("arpack",))
    w = np.linalg.eigvalsh(_dense(m))[:6]
    assert res[0]["status"] == "success" and np.abs(np.asarray(res[0]["eigenvalues"]) - w).max() <= 1e-10
    path = compare.write_report(res, str(tmp_path), "m", 6, 1e-10, "test")
    text = open(path).read()
    assert "convergence_metric=max_relative_eigen_residual" in text
    assert "method=arpack status=success returned_k=6" in text
    assert os.path.exists(str(tmp_path / "m_compare.csv"))
