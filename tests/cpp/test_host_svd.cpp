// Driver for tests/test_host_svd.py: reads matrices (int64 rows, cols, then
// column-major FP64 values) from argv[1] and writes, for each and for both
// builds of the host SVD (0 baseline, 1 AVX2 + FMA), a flag byte (1 if the
// build ran) followed by S (min(m, n)), U (m x p) and V (n x p), p = min(m, n),
// column-major, to argv[2].
#include <cstdint>
#include <cstdio>
#include <vector>

#include "host_svd.hpp"

using irlba_hip::HostMat;

int main(int argc, char** argv) {
    if (argc != 3) return 2;
    FILE* in = std::fopen(argv[1], "rb");
    FILE* out = std::fopen(argv[2], "wb");
    if (!in || !out) return 2;
    while (true) {
        std::int64_t dims[2];
        if (std::fread(dims, sizeof(dims), 1, in) != 1) break;
        HostMat A(dims[0], dims[1]);
        if (std::fread(A.a.data(), sizeof(double), A.a.size(), in) != A.a.size()) return 3;
        for (int variant = 0; variant < 2; ++variant) {
            HostMat U, V;
            std::vector<double> S;
            const unsigned char ran = irlba_hip::host_svd_variant(A, U, S, V, variant) ? 1 : 0;
            std::fwrite(&ran, 1, 1, out);
            if (!ran) continue;
            std::fwrite(S.data(), sizeof(double), S.size(), out);
            std::fwrite(U.a.data(), sizeof(double), U.a.size(), out);
            std::fwrite(V.a.data(), sizeof(double), V.a.size(), out);
        }
    }
    std::fclose(in);
    std::fclose(out);
    return 0;
}
