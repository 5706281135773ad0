// Test infrastructure only: a stable host sort standing in for rocPRIM's device radix
// sort in the CPU emulation build (same call shape: size query with a null buffer,
// then the sort of (key, value) pairs on bits [begin_bit, end_bit) of the key).
#pragma once
#include <algorithm>
#include <numeric>
#include <vector>

namespace rocprim {

template <class K, class V>
hipError_t radix_sort_pairs(void* temp, size_t& bytes, const K* keysIn, K* keysOut, const V* valsIn, V* valsOut,
                            size_t n, unsigned beginBit = 0, unsigned endBit = 8 * sizeof(K),
                            hipStream_t = nullptr, bool = false) {
    if (!temp) {
        bytes = 16;
        return hipSuccess;
    }
    const unsigned width = endBit - beginBit;
    const K mask = width >= 8 * sizeof(K) ? ~K(0) : ((K(1) << width) - 1);
    auto bitsOf = [&](const K& k) { return (k >> beginBit) & mask; };
    std::vector<size_t> idx(n);
    std::iota(idx.begin(), idx.end(), size_t(0));
    std::stable_sort(idx.begin(), idx.end(), [&](size_t a, size_t b) { return bitsOf(keysIn[a]) < bitsOf(keysIn[b]); });
    std::vector<K> k2(n);
    std::vector<V> v2(n);
    for (size_t i = 0; i < n; ++i) { k2[i] = keysIn[idx[i]]; v2[i] = valsIn[idx[i]]; }
    std::copy(k2.begin(), k2.end(), keysOut);
    std::copy(v2.begin(), v2.end(), valsOut);
    return hipSuccess;
}

}  // namespace rocprim
