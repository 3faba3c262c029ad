"""Write the golden fixtures under tests/golden/ from data files the reference
checkout holds (result tables and the H2O CSC fixture), and cluster ranges
from the reference's own host code built from source in oracle/_ref.

Run in the container where /root/reference exists:
    python tests/golden/make_golden.py
Only numbers are written (JSON); no reference source text is copied.
"""
import csv
import json
import os
import re
import sys

REF = "/root/reference/IPT_C_GPU"
HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))


def _write(name, obj):
    path = os.path.join(HERE, name)
    with open(path, "w") as f:
        json.dump(obj, f, indent=1, sort_keys=True)
    print("wrote", path)


def runtests_kat():
    """Dense KAT table: IPT/results/runtests_summary.csv (cases of
    IPT/test_temp/runtests.cu:304-313, 20 iterations, tolerance 5e-9)."""
    rows = []
    with open(os.path.join(REF, "results/runtests_summary.csv")) as f:
        for r in csv.DictReader(f):
            rows.append({
                "case": r["case"], "n": int(r["n"]), "k": int(r["k"]),
                "symmetric": int(r["symmetric"]),
                "vector_max_abs_diff": float(r["vector_max_abs_diff"]),
                "value_max_abs_diff": float(r["value_max_abs_diff"]),
                "max_relative_residual": float(r["max_relative_residual"]),
                "status": r["status"],
            })
    _write("dense_kat.json", {"iterations": 20, "diff_tol": 5e-9, "residual_tol": 5e-9,
                              "source": "IPT/results/runtests_summary.csv", "cases": rows})


def sparse_kat():
    """Sparse KAT table: IPT/results/sparse/sparse_runtests_summary.csv."""
    rows = []
    with open(os.path.join(REF, "results/sparse/sparse_runtests_summary.csv")) as f:
        for r in csv.DictReader(f):
            rows.append({k: (r[k] if k in ("case", "status") else float(r[k])) for k in r})
    _write("sparse_kat.json", {"source": "IPT/results/sparse/sparse_runtests_summary.csv",
                               "cases": rows})


def dense_eps_sweep():
    """splitmix64 sweep: IPT/results/gpu_cusolver_vs_ipt_results.csv with the
    configuration of IPT/results/gpu_cusolver_vs_ipt_summary.txt (seed
    20260603+N, eps 0.01, tol 1e-12, k = N)."""
    rows = []
    with open(os.path.join(REF, "results/gpu_cusolver_vs_ipt_results.csv")) as f:
        for r in csv.DictReader(f):
            rows.append({"n": int(r["N"]), "epsilon": float(r["epsilon"]),
                         "ipt_iterations": int(r["ipt_iterations"]),
                         "ipt_fixed_point_residual": float(r["ipt_fixed_point_residual"]),
                         "residual_ipt_frobenius": float(r["residual_ipt"])})
    _write("dense_eps_sweep.json", {"seed_base": 20260603, "tol": 1e-12, "maxiter": 1000,
                                    "source": "IPT/results/gpu_cusolver_vs_ipt_results.csv",
                                    "rows": rows})


def h2o_csc():
    """10x10 H2O FCI matrix W (IPT/test/sparse/h20_w_csc.h), numbers only."""
    text = open(os.path.join(REF, "test/sparse/h20_w_csc.h")).read()

    def arr(name, conv):
        m = re.search(name + r"\[\d+\]\s*=\s*\{([^}]*)\}", text)
        return [conv(t) for t in m.group(1).replace("\n", " ").split(",") if t.strip()]

    n = int(re.search(r"H20_W_CSC_N\s*=\s*(\d+)", text).group(1))
    obj = {"n": n, "col_ptr": arr("H20_W_CSC_COL_PTR", int),
           "row_ind": arr("H20_W_CSC_ROW_IND", int),
           "values": arr("H20_W_CSC_VALUES", float),
           "k": 1, "iterations": 80, "diff_tol": 1e-8, "residual_tol": 1e-10,
           "source": "IPT/test/sparse/h20_w_csc.h"}
    assert len(obj["values"]) == obj["col_ptr"][-1]
    _write("h2o_w_csc.json", obj)


def cluster_ranges():
    """Cluster ranges from the reference's ipt_target_degenerate_subspaces
    compiled from source (oracle/_ref) on seeded inputs of this repo's
    generators."""
    sys.path.insert(0, ROOT)
    import numpy as np
    from oracle import oracle as orc

    cases = []
    specs = [
        ("rand_dense_n512_k16_s0", 512, 16, 0, 0.0, 10.0, 0.05),
        ("rand_dense_n1024_k32_s3", 1024, 32, 3, 0.0, 50.0, 0.1),
        ("rand_dense_n2048_k16_s0", 2048, 16, 0, 0.0, 10.0, 0.05),
        ("rand_dense_n2048_k64_s1", 2048, 64, 1, 0.0, 6.25, 0.1),
        ("eps_dense_n256_k256", 256, 256, 20260603 + 256, None, None, 0.05),
        # ragged sizes (not multiples of the scan tiles) and the C2 recipe
        ("rand_dense_n777_k20_s5", 777, 20, 5, 0.0, 10.0, 0.05),
        ("rand_dense_n1000_k24_s7", 1000, 24, 7, 0.0, 12.0, 0.08),
        ("rand_dense_n1531_k40_s2", 1531, 40, 2, 0.0, 30.0, 0.1),
        ("rand_dense_n3001_k48_s9", 3001, 48, 9, 0.0, 40.0, 0.1),
        ("rand_dense_n4096_k64_s4", 4096, 64, 4, 0.0, 50.0, 0.1),
        ("rand_dense_n8192_k64_s0", 8192, 64, 0, 0.0, 50.0, 0.1),
        ("rand_dense_n16384_k64_s0", 16384, 64, 0, 0.0, 50.0, 0.1),
    ]
    for name, n, k, seed, lo, hi, vs in specs:
        if lo is None:
            H = orc.eps_dense(n, vs, seed)
        else:
            H = orc.rand_dense(n, seed, lo, hi, vs)
        # CSC of the nonzeros, column by column (rows ascending)
        cols, rows = np.nonzero(H.T)
        cp = np.zeros(n + 1, dtype=np.int32)
        cp[1:] = np.cumsum(np.bincount(cols, minlength=n))
        ri = rows.astype(np.int32)
        v = np.ascontiguousarray(H.T[cols, rows])
        del cols, rows
        got = orc.ref_target_clusters(cp, ri, v, k)
        if got is None:
            raise SystemExit("oracle/_ref not built: run make -C oracle ref")
        ranges, too_large = got
        cases.append({"name": name, "kind": "eps" if lo is None else "rand", "n": n, "k": k,
                      "seed": seed, "lo": lo, "hi": hi, "vscale": vs,
                      "ranges": ranges, "too_large": too_large})
    _write("cluster_ranges.json", {"source": "IPT/src/ipt_preparation.cu:501-645 via oracle/_ref",
                                   "cases": cases})


def acx_sweep():
    """ACX vs plain IPT: IPT/results/ACX/acx_vs_original_ipt_summary.csv, the
    configuration of IPT/test/ACX/benchmark_acx_cuda_sweep.cu:20-25 (k = N,
    tol 1e-12, maxiter 1000, seed 20260615 + N + a per-epsilon offset)."""
    rows = []
    with open(os.path.join(REF, "results/ACX/acx_vs_original_ipt_summary.csv")) as f:
        for r in csv.DictReader(f):
            rows.append({"n": int(r["N"]), "epsilon": float(r["epsilon"]),
                         "acx_iterations": int(r["acx_iterations"]), "acx_f_calls": int(r["acx_f_calls"]),
                         "ipt_iterations": int(r["original_ipt_iterations"]),
                         "acx_fixed_point_residual": float(r["acx_fixed_point_residual"]),
                         "ipt_fixed_point_residual": float(r["original_ipt_fixed_point_residual"]),
                         "acx_time_sec": float(r["acx_time_sec"]),
                         "ipt_time_sec": float(r["original_ipt_time_sec"]),
                         "status": r["acx_status"] + "/" + r["original_ipt_status"]})
    _write("acx_sweep.json", {"seed_base": 20260615, "tol": 1e-12, "maxiter": 1000,
                              "gpu": "NVIDIA A100-SXM4-80GB",
                              "source": "IPT/results/ACX/acx_vs_original_ipt_summary.csv", "rows": rows})


def acx_mixed_sweep():
    """Two-phase ACX (TF32 then FP64): IPT/results/ACX/ACX_MIXED/
    acx_mixed_tf32_cuda_sweep_epsilon_*.csv, the configuration of
    IPT/test/ACX/ACX_MIXED/benchmark_acx_mixed_tf32_cuda_sweep.cu:17-26 (k = N,
    the ACX sweep's matrices and seeds)."""
    import glob

    rows = []
    for f in sorted(glob.glob(os.path.join(REF, "results/ACX/ACX_MIXED/acx_mixed_tf32_cuda_sweep_epsilon_0_*.csv"))):
        with open(f) as fh:
            for r in csv.DictReader(fh):
                rows.append({"n": int(r["N"]), "epsilon": float(r["epsilon"]),
                             "tf32_iterations": int(r["tf32_iterations"]), "tf32_f_calls": int(r["tf32_f_calls"]),
                             "fp64_iterations": int(r["fp64_iterations"]), "fp64_f_calls": int(r["fp64_f_calls"]),
                             "tf32_converged": int(r["tf32_converged"]), "fp64_converged": int(r["fp64_converged"]),
                             "tf32_fixed_point_residual": float(r["tf32_fixed_point_residual"]),
                             "fp64_fixed_point_residual": float(r["fp64_fixed_point_residual"]),
                             "time_mixed_total_sec": float(r["time_mixed_total_sec"]), "status": r["status"]})
    _write("acx_mixed_sweep.json", {"seed_base": 20260615, "tf32_tol": 1e-6, "fp64_tol": 1e-12, "maxiter": 1000,
                                    "source": "IPT/results/ACX/ACX_MIXED/acx_mixed_tf32_cuda_sweep_epsilon_*.csv",
                                    "rows": rows})


def c2_lowest():
    """The 65 lowest eigenvalues (LAPACK dsyevr) of the dense N=16384 matrix
    of the headline workload (d ~ U[0,50] sorted, V ~ N(0,1) 0.1/sqrt(N),
    seed 0; the oracle's generator, which the device generator equals bit
    for bit).  About two minutes on 8 cores."""
    import numpy as np
    import scipy.linalg as sl
    sys.path.insert(0, ROOT)
    from oracle import oracle as orc
    n, k = 16384, 64
    H = orc.rand_dense(n, 0, 0.0, 50.0, 0.1)
    w = sl.eigh(H, eigvals_only=True, subset_by_index=[0, k], driver="evr", overwrite_a=True, check_finite=False)
    _write("c2_lowest.json", {"n": n, "seed": 0, "lo": 0.0, "hi": 50.0, "vscale": 0.1,
                              "source": "scipy.linalg.eigh (LAPACK dsyevr) of oracle.rand_dense(16384, 0, 0, 50, 0.1)",
                              "values": [float(x) for x in w]})


def c5_lowest():
    """The 256 lowest eigenvalues (LAPACK dsyevr) of the dense k=256
    workload's recipe at N=8192: d ~ U[0,50] sorted, V ~ N(0,1) vscale/sqrt(N),
    seed 0, with vscale 0.1 (the recipe as written) and 0.1 sqrt(8) (the
    coupling-to-spacing ratio of N=65536), the latter also with the
    near-degenerate gaps (1e-4 after diagonal positions 10, 50, 120, 200).
    About a minute on 8 cores."""
    import numpy as np
    import scipy.linalg as sl
    sys.path.insert(0, ROOT)
    from oracle import oracle as orc
    n, k = 8192, 256
    cases = []
    for name, vs, gaps in [("recipe", 0.1, False), ("regime", 0.1 * np.sqrt(8.0), False),
                           ("regime_near_degenerate", 0.1 * np.sqrt(8.0), True)]:
        H = orc.rand_dense(n, 0, 0.0, 50.0, vs)
        if gaps:
            orc.set_diagonal_gaps(H)
        w = sl.eigh(H, eigvals_only=True, subset_by_index=[0, k], driver="evr", overwrite_a=True,
                    check_finite=False)
        cases.append({"name": name, "vscale": float(vs), "gaps": gaps, "values": [float(x) for x in w]})
    _write("c5_lowest.json", {"n": n, "k": k, "seed": 0, "lo": 0.0, "hi": 50.0,
                              "gap_positions": [10, 50, 120, 200], "gap": 1e-4,
                              "source": "scipy.linalg.eigh (LAPACK dsyevr) of oracle.rand_dense(8192, 0, 0, 50, vscale)",
                              "cases": cases})


if __name__ == "__main__":
    if len(sys.argv) > 1:
        for name in sys.argv[1:]:
            globals()[name]()
        sys.exit(0)
    runtests_kat()
    sparse_kat()
    dense_eps_sweep()
    h2o_csc()
    cluster_ranges()
    acx_sweep()
    acx_mixed_sweep()
    c2_lowest()
    c5_lowest()
