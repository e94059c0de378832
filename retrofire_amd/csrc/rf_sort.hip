// Device radix sorts of the depth-sort keys (64-bit keys, 32-bit values; and
// 32-bit draw keys), through hipCUB's DeviceRadixSort (rocPRIM's radix sort on
// gfx950, stable).
#include <hip/hip_runtime.h>
#include <hipcub/device/device_radix_sort.hpp>

#include <cstddef>
#include <cstdint>

namespace rf {

// With tmp == nullptr only sets tmp_bytes (the scratch the sort needs).
hipError_t sort_pairs_u64(void* tmp, size_t& tmp_bytes, const unsigned long long* keys_in,
                          unsigned long long* keys_out, const uint32_t* vals_in, uint32_t* vals_out, uint32_t n,
                          hipStream_t s) {
    return hipcub::DeviceRadixSort::SortPairs(tmp, tmp_bytes, keys_in, keys_out, vals_in, vals_out, (int)n, 0, 64,
                                              s);
}

// 32-bit keys (only bits [0, end_bit) compared), 32-bit values; stable.
hipError_t sort_pairs_u32(void* tmp, size_t& tmp_bytes, const uint32_t* keys_in, uint32_t* keys_out,
                          const uint32_t* vals_in, uint32_t* vals_out, uint32_t n, uint32_t end_bit, hipStream_t s) {
    return hipcub::DeviceRadixSort::SortPairs(tmp, tmp_bytes, keys_in, keys_out, vals_in, vals_out, (int)n, 0,
                                              (int)end_bit, s);
}

}  // namespace rf
