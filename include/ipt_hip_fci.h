/*
 * ipt_hip_fci.h -- inputs for the sparse path from quantum chemistry: the
 * determinant-space (full CI) Hamiltonian of an active space built from
 * one- and two-electron integrals, the FCIDUMP integral file, and the binary
 * CSC matrix cache of the reference's sparse benchmarks.  Host code only (no
 * device work).  Callers: drivers that feed real-chemistry Hamiltonians to
 * ipt_cuda_sparse_csc* (the reference's CH4 / NH3 / CO / N2 benchmarks).
 */
#ifndef IPT_HIP_FCI_H
#define IPT_HIP_FCI_H

#ifdef __cplusplus
extern "C" {
#endif

/* A CSC matrix with the active-space metadata of the cache header.
 * Arrays are malloc-owned by the library when filled by it and released by
 * ipt_hip_fci_free_csc. */
typedef struct {
    int n;
    int nnz;
    int *col_ptr;      /* n + 1 */
    int *row_ind;      /* nnz, ascending within a column */
    double *values;    /* nnz */
    double *diagonal;  /* n: the matrix diagonal (first entry of row == column) */
    int *perm;         /* n: perm[sorted] = determinant index ia * nb + ib (NULL when read from a cache) */
    double matrix_inf_norm;  /* max row sum of |values| */
    int norb;
    int nelec;
    int neleca;
    int nelecb;
    int na;            /* alpha strings: C(norb, neleca) */
    int nb;            /* beta strings: C(norb, nelecb) */
} IPTFciCsc;

/* Full-CI Hamiltonian of `neleca` alpha and `nelecb` beta electrons in
 * `norb` spatial orbitals (norb <= 24), real integrals in chemists' notation:
 * h1[p*norb + q] = h_pq, eri[((p*norb + q)*norb + r)*norb + s] = (pq|rs)
 * (8-fold symmetric), constant ecore.  Determinant index ia * nb + ib over
 * alpha / beta occupation strings enumerated with their occupied orbitals in
 * lexicographic order; with sort_by_diagonal the rows and columns are
 * ordered by a stable sort of the diagonal.  Entries summed per position,
 * exact zeros dropped, rows ascending.  The reference builds the same matrix
 * in generate_hamiltonian_csc
 * (IPT/test/sparse/CH4/benchmark_ch4_sto6g_fci_15876_ipt.cu:743-903) from
 * integrals it obtains through an embedded PySCF; here the integrals come
 * from the caller (ipt_hip_fcidump_read).  Status 1 for invalid sizes, 2 on
 * allocation failure. */
int ipt_hip_fci_build_csc(int norb, int neleca, int nelecb, const double *h1, const double *eri, double ecore,
                          int sort_by_diagonal, IPTFciCsc *out);
void ipt_hip_fci_free_csc(IPTFciCsc *m);

/* FCIDUMP integral file (Knowles & Handy, Comput. Phys. Commun. 54, 75
 * (1989)): namelist header &FCI NORB=, NELEC=, MS2= ... &END (or /), then
 * lines "value i j k l" with 1-based orbital indices: (ij|kl) for i,j,k,l > 0
 * (8-fold symmetry applied), h_ij for k = l = 0, the core energy for all
 * zero; orbital energies (j = k = l = 0) are ignored.  read_header gives the
 * sizes, read fills h1 (norb^2), eri (norb^4) and ecore.  Status 1 for a
 * missing or malformed file. */
int ipt_hip_fcidump_read_header(const char *path, int *norb, int *nelec, int *ms2);
int ipt_hip_fcidump_read(const char *path, int norb, double *h1, double *eri, double *ecore);
/* Writes the unique (8-fold reduced) nonzero integrals of h1 / eri. */
int ipt_hip_fcidump_write(const char *path, int norb, int nelec, int ms2, const double *h1, const double *eri,
                          double ecore);

/* Binary CSC cache, the reference's format
 * (benchmark_ch4_sto6g_fci_15876_ipt.cu:182-194, 1015-1121): a 56-byte
 * header {char magic[8] = "IPTCH4C1"; uint32 version = 1; int32 n, nnz;
 * double matrix_inf_norm; int32 norb, nelec, neleca, nelecb, na, nb} in the
 * host's byte order (with 4 bytes of padding before the double), then
 * col_ptr (n + 1 int32), row_ind (nnz int32), values (nnz double), diagonal
 * (n double).  write goes through a temporary file renamed into place;
 * read validates the magic, version and CSC structure and fills a
 * library-owned IPTFciCsc (perm NULL).  Status 1 for an invalid file. */
int ipt_hip_csc_cache_write(const char *path, const IPTFciCsc *m);
int ipt_hip_csc_cache_read(const char *path, IPTFciCsc *out);

#ifdef __cplusplus
}
#endif

#endif /* IPT_HIP_FCI_H */
