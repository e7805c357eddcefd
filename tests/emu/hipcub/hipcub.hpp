// Serial stand-in for the one hipcub call the library makes -- TEST ONLY.
#pragma once
#include <cstddef>
#include <type_traits>

namespace hipcub {
struct DeviceScan {
    template <class I, class O>
    static int ExclusiveSum(void *tmp, size_t &bytes, I in, O out, int n,
                            void * = nullptr)
    {
        if (!tmp) {
            bytes = 16;
            return 0;
        }
        typename std::remove_reference<decltype(*out)>::type acc = 0;
        for (int i = 0; i < n; ++i) {
            auto v = in[i];
            out[i] = acc;
            acc += v;
        }
        return 0;
    }
};
}  // namespace hipcub
