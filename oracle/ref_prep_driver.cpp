// Driver (ours) that compiles the reference's host-only preparation source
// IPT/src/ipt_preparation.cu in place and exposes its cluster detection
// through a C entry point, so tests can compare the oracle's restatement
// with the reference's own code.  Built only where the reference checkout
// exists; the output goes to oracle/_ref/.  Test infrastructure only.
#include IPT_REF_SRC

extern "C" int ref_target_clusters(const int *col_ptr, const int *row_ind,
                                   const double *values, int n, int k,
                                   double degeneracy_threshold,
                                   int *ranges_out, int *too_large)
{
    std::vector<int> cp(col_ptr, col_ptr + n + 1);
    std::vector<int> ri(row_ind, row_ind + cp[n]);
    std::vector<double> v(values, values + cp[n]);
    std::vector<double> diag;
    ipt_extract_csc_diagonal_vectors(cp, ri, v, n, &diag);
    IPTDegeneracyOptions options = ipt_degeneracy_options(degeneracy_threshold);
    std::vector<std::pair<int, int>> ranges =
        ipt_target_degenerate_subspaces(cp, ri, v, diag, k, options, too_large);
    for (size_t i = 0; i < ranges.size(); ++i) {
        ranges_out[2 * i] = ranges[i].first;
        ranges_out[2 * i + 1] = ranges[i].second;
    }
    return (int)ranges.size();
}
