// Serial stand-ins for the hipcub calls the library makes -- TEST ONLY.
#pragma once
#include <algorithm>
#include <cstddef>
#include <cstdint>
#include <numeric>
#include <type_traits>
#include <vector>

namespace hipcub {
struct DeviceScan {
    template <class I, class O, class N>
    static int ExclusiveSum(void *tmp, size_t &bytes, I in, O out, N n,
                            void * = nullptr)
    {
        if (!tmp) {
            bytes = 16;
            return 0;
        }
        typename std::remove_reference<decltype(*out)>::type acc = 0;
        for (N i = 0; i < n; ++i) {
            auto v = in[i];
            out[i] = acc;
            acc += v;
        }
        return 0;
    }
    template <class I, class O, class N>
    static int InclusiveSum(void *tmp, size_t &bytes, I in, O out, N n,
                            void * = nullptr)
    {
        if (!tmp) {
            bytes = 16;
            return 0;
        }
        typename std::remove_reference<decltype(*out)>::type acc = 0;
        for (N i = 0; i < n; ++i) {
            acc += in[i];
            out[i] = acc;
        }
        return 0;
    }
};

// stable sort of (key, value) by key bits [begin_bit, end_bit)
struct DeviceRadixSort {
    template <class K, class V, class N>
    static int SortPairs(void *tmp, size_t &bytes, const K *kin, K *kout,
                         const V *vin, V *vout, N n, int begin_bit = 0,
                         int end_bit = sizeof(K) * 8, void * = nullptr)
    {
        if (!tmp) {
            bytes = 16;
            return 0;
        }
        const int w = end_bit - begin_bit;
        const uint64_t mask = w >= 64 ? ~0ull : ((1ull << w) - 1ull);
        auto key = [&](size_t i) {
            return ((uint64_t)kin[i] >> begin_bit) & mask;
        };
        std::vector<size_t> ord((size_t)n);
        std::iota(ord.begin(), ord.end(), 0);
        std::stable_sort(ord.begin(), ord.end(),
                         [&](size_t a, size_t b) { return key(a) < key(b); });
        std::vector<K> ks((size_t)n);
        std::vector<V> vs((size_t)n);
        for (size_t i = 0; i < (size_t)n; ++i) {
            ks[i] = kin[ord[i]];
            vs[i] = vin[ord[i]];
        }
        std::copy(ks.begin(), ks.end(), kout);
        std::copy(vs.begin(), vs.end(), vout);
        return 0;
    }
};
}  // namespace hipcub
