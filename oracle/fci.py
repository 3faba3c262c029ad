"""CPU oracle for the determinant-space (full CI) Hamiltonian -- TEST
INFRASTRUCTURE ONLY (imported by tests/, never by the library).

A restatement of the reference's generate_hamiltonian_csc
(IPT/test/sparse/CH4/benchmark_ch4_sto6g_fci_15876_ipt.cu:613-903): alpha /
beta occupation strings in the order of generate_strings (:637-661), the
one-body links E_pq (:671-690, apply_one_body :582-598) and the two-body
links a+_p a+_r a_s a_q (:692-713, apply_two_body :600-622) of every string,
and per column the sum of ecore, h1 over one-body links, 0.5 (pq|rs) over
same-spin two-body links and (pq|rs) over pairs of alpha and beta one-body
links (:814-860); the diagonal of :694-730 orders rows and columns by a
stable sort.  Pure Python: small active spaces only.

Pinned in tests/test_fci.py by what the mathematics fixes, independently of
this restatement: the spectrum is invariant under orbital rotations (a wrong
phase or a dropped exchange term breaks that), one electron gives eig(h1),
and the minimal-basis H2 FCI energy equals its closed 2x2 form (Szabo &
Ostlund, Modern Quantum Chemistry, chapters 3-4).
"""
import numpy as np


def strings(norb, nel):
    """Occupation bit strings, occupied orbitals in lexicographic order."""
    out = []

    def rec(start, left, bits):
        if left == 0:
            out.append(bits)
            return
        for o in range(start, norb - left + 1):
            rec(o + 1, left - 1, bits | (1 << o))

    rec(0, nel, 0)
    return out


def _annihilate(bits, o, sign):
    m = 1 << o
    if not bits & m:
        return None, sign
    if bin(bits & (m - 1)).count("1") & 1:
        sign = -sign
    return bits ^ m, sign


def _create(bits, o, sign):
    m = 1 << o
    if bits & m:
        return None, sign
    if bin(bits & (m - 1)).count("1") & 1:
        sign = -sign
    return bits | m, sign


def one_links(bits, norb, addr):
    """[(p, q, target, sign)] of E_pq |bits> (q occupied)."""
    out = []
    for q in range(norb):
        if not bits >> q & 1:
            continue
        for p in range(norb):
            b, s = _annihilate(bits, q, 1)
            b, s = _create(b, p, s)
            if b is not None and b in addr:
                out.append((p, q, addr[b], s))
    return out


def two_links(bits, norb, addr):
    """[(p, q, r, s, target, sign)] of a+_p a+_r a_s a_q |bits>."""
    out = []
    occ = [o for o in range(norb) if bits >> o & 1]
    for q in occ:
        for s_ in occ:
            for r in range(norb):
                for p in range(norb):
                    b, sg = _annihilate(bits, q, 1)
                    b, sg = _annihilate(b, s_, sg) if b is not None else (None, sg)
                    b, sg = _create(b, r, sg) if b is not None else (None, sg)
                    b, sg = _create(b, p, sg) if b is not None else (None, sg)
                    if b is not None and b in addr:
                        out.append((p, q, r, s_, addr[b], sg))
    return out


def determinant_diagonal(h1, eri, ecore, abits, bbits, norb):
    """<D|H|D>, terms added in the order of the reference's
    determinant_diagonal (:694-730); the sort key of the row order."""
    v = ecore
    oa = [o for o in range(norb) if abits >> o & 1]
    ob = [o for o in range(norb) if bbits >> o & 1]
    for o in range(norb):  # orbital by orbital, alpha before beta (:700-712)
        if abits >> o & 1:
            v += h1[o, o]
        if bbits >> o & 1:
            v += h1[o, o]
    for occ in (oa, ob):
        for i in occ:
            for j in occ:
                v += 0.5 * (eri[i, i, j, j] - eri[i, j, j, i])
    for i in oa:
        for j in ob:
            v += eri[i, i, j, j]
    return v


def fci_dense(norb, neleca, nelecb, h1, eri, ecore, sort=True):
    """Dense FCI Hamiltonian in the sorted-diagonal order; returns (H, perm)
    with perm[sorted] = ia * nb + ib."""
    h1 = np.asarray(h1, dtype=np.float64).reshape(norb, norb)
    eri = np.asarray(eri, dtype=np.float64).reshape(norb, norb, norb, norb)
    sa, sb = strings(norb, neleca), strings(norb, nelecb)
    aa = {b: i for i, b in enumerate(sa)}
    ab = {b: i for i, b in enumerate(sb)}
    na, nb = len(sa), len(sb)
    n = na * nb
    a1 = [one_links(b, norb, aa) for b in sa]
    b1 = [one_links(b, norb, ab) for b in sb]
    a2 = [two_links(b, norb, aa) for b in sa]
    b2 = [two_links(b, norb, ab) for b in sb]
    diag = np.array([determinant_diagonal(h1, eri, ecore, sa[d // nb], sb[d % nb], norb) for d in range(n)])
    perm = np.argsort(diag, kind="stable") if sort else np.arange(n)
    inv = np.empty(n, dtype=np.int64)
    inv[perm] = np.arange(n)
    H = np.zeros((n, n))
    for c in range(n):
        old = perm[c]
        ia, ib = old // nb, old % nb
        H[inv[old], c] += ecore
        for p, q, t, s in a1[ia]:
            H[inv[t * nb + ib], c] += h1[p, q] * s
        for p, q, t, s in b1[ib]:
            H[inv[ia * nb + t], c] += h1[p, q] * s
        for p, q, r, s_, t, s in a2[ia]:
            H[inv[t * nb + ib], c] += 0.5 * eri[p, q, r, s_] * s
        for p, q, r, s_, t, s in b2[ib]:
            H[inv[ia * nb + t], c] += 0.5 * eri[p, q, r, s_] * s
        for pa, qa, ta, sa_ in a1[ia]:
            for pb, qb, tb, sb_ in b1[ib]:
                H[inv[ta * nb + tb], c] += eri[pa, qa, pb, qb] * sa_ * sb_
    return H, perm


def random_integrals(norb, seed, spread=2.0, coupling=0.05):
    """Real integrals with the 8-fold symmetry of a molecular basis:
    orbital energies spread over [-spread, spread] on the diagonal of h1,
    Coulomb-like (pp|qq), exchange-like (pq|qp) and small random rest."""
    rng = np.random.default_rng(seed)
    h1 = np.diag(np.linspace(-spread, spread, norb)) + coupling * rng.standard_normal((norb, norb))
    h1 = 0.5 * (h1 + h1.T)
    eri = np.zeros((norb,) * 4)
    for p in range(norb):
        for q in range(norb):
            for r in range(norb):
                for s in range(norb):
                    eri[p, q, r, s] = np.nan
    vals = {}
    for p in range(norb):
        for q in range(p + 1):
            for r in range(norb):
                for s in range(r + 1):
                    pq, rs = p * (p + 1) // 2 + q, r * (r + 1) // 2 + s
                    if pq < rs:
                        continue
                    if p == q and r == s:
                        v = 0.4 + 0.1 * rng.uniform()
                    elif (p, q) == (r, s):
                        v = 0.05 + 0.05 * rng.uniform()
                    else:
                        v = coupling * 0.3 * rng.standard_normal()
                    vals[(p, q, r, s)] = v
    for (p, q, r, s), v in vals.items():
        for a, b, c, d in ((p, q, r, s), (q, p, r, s), (p, q, s, r), (q, p, s, r),
                           (r, s, p, q), (s, r, p, q), (r, s, q, p), (s, r, q, p)):
            eri[a, b, c, d] = v
    assert not np.isnan(eri).any()
    return h1, eri


def rotate_integrals(h1, eri, U):
    """Integrals in the orbitals U (columns): h' = U^T h U,
    (pq|rs)' = sum U_ap U_bq U_cr U_ds (ab|cd)."""
    h = U.T @ h1 @ U
    g = np.einsum("abcd,ap,bq,cr,ds->pqrs", eri, U, U, U, U, optimize=True)
    return h, g
