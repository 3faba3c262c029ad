// Test driver for block_ipt_amd/csrc/ipt_small_eig.h: reads m, G, P
// (column-major doubles) from a binary file, writes theta then C.  With a
// third argument `need` it calls the rank-revealing variant.
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "../../block_ipt_amd/csrc/ipt_small_eig.h"

int main(int argc, char **argv)
{
    if (argc != 3 && argc != 4) {
        return 2;
    }
    FILE *f = fopen(argv[1], "rb");
    if (!f) {
        return 2;
    }
    int m = 0;
    if (fread(&m, sizeof(int), 1, f) != 1) {
        return 2;
    }
    std::vector<double> G((size_t)m * m), P((size_t)m * m), theta((size_t)m), C((size_t)m * m);
    if (fread(G.data(), sizeof(double), G.size(), f) != G.size() || fread(P.data(), sizeof(double), P.size(), f) != P.size()) {
        return 2;
    }
    fclose(f);
    const int st = argc == 4 ? ipt::sym_generalized_eig_robust(m, G.data(), P.data(), theta.data(), C.data(),
                                                               atoi(argv[3]))
                             : ipt::sym_generalized_eig(m, G.data(), P.data(), theta.data(), C.data());
    FILE *o = fopen(argv[2], "wb");
    fwrite(&st, sizeof(int), 1, o);
    fwrite(theta.data(), sizeof(double), theta.size(), o);
    fwrite(C.data(), sizeof(double), C.size(), o);
    fclose(o);
    return 0;
}
