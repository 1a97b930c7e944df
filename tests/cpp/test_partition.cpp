// CPU check of the thread-partition getters of HipSparseMatrix: reads a
// compressed pattern (binary: n_primary, n_secondary, nnz, T as int64, then
// nnz+1... see tests/test_partition.py) and prints the partition that
// irlba::detail::sparse_partition computes, one line per array.  Uses only the
// header's host code: no device, no library.
#include <cstdint>
#include <cstdio>
#include <vector>

#include "irlba/Matrix/hip.hpp"

int main(int argc, char** argv) {
    if (argc != 2) return 2;
    std::FILE* f = std::fopen(argv[1], "rb");
    if (!f) return 2;
    std::int64_t hdr[4];
    if (std::fread(hdr, sizeof(std::int64_t), 4, f) != 4) return 2;
    const std::int64_t np = hdr[0], ns = hdr[1], nnz = hdr[2];
    const int T = static_cast<int>(hdr[3]);
    std::vector<std::size_t> ptr(static_cast<std::size_t>(np) + 1);
    std::vector<int> idx(static_cast<std::size_t>(nnz));
    std::vector<std::int64_t> p64(ptr.size());
    if (std::fread(p64.data(), sizeof(std::int64_t), p64.size(), f) != p64.size()) return 2;
    if (nnz && std::fread(idx.data(), sizeof(int), idx.size(), f) != idx.size()) return 2;
    std::fclose(f);
    for (std::size_t i = 0; i < ptr.size(); ++i) ptr[i] = static_cast<std::size_t>(p64[i]);
    const auto part = irlba::detail::sparse_partition<std::size_t>(np, ns, ptr, idx, T);
    for (auto v : part.primary) std::printf("%lld ", static_cast<long long>(v));
    std::printf("\n");
    for (auto v : part.secondary) std::printf("%lld ", static_cast<long long>(v));
    std::printf("\n");
    for (const auto& row : part.nonzero)
        for (auto v : row) std::printf("%lld ", static_cast<long long>(v));
    std::printf("\n");
    return 0;
}
