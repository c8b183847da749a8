// Device-wide exclusive scan and radix sort (hipCUB) used by the row pipeline.
#include <hipcub/hipcub.hpp>

#include "tcg_device.hpp"

namespace tcg {

size_t scan_temp_bytes(int n) {
    size_t bytes = 0;
    (void)hipcub::DeviceScan::ExclusiveSum(nullptr, bytes, (const long long *)nullptr, (long long *)nullptr, n);
    return bytes;
}

hipError_t scan_exclusive(void *tmp, size_t bytes, const long long *in, long long *out, int n, hipStream_t st) {
    return hipcub::DeviceScan::ExclusiveSum(tmp, bytes, in, out, n, st);
}

size_t sort_temp_bytes(int n, int end_bit) {
    size_t bytes = 0;
    (void)hipcub::DeviceRadixSort::SortPairs(nullptr, bytes, (const unsigned *)nullptr, (unsigned *)nullptr,
                                             (const int *)nullptr, (int *)nullptr, n, 0, end_bit);
    return bytes;
}

hipError_t sort_pairs(void *tmp, size_t bytes, const unsigned *kin, unsigned *kout, const int *vin, int *vout, int n,
                      int end_bit, hipStream_t st) {
    return hipcub::DeviceRadixSort::SortPairs(tmp, bytes, kin, kout, vin, vout, n, 0, end_bit, st);
}

// indices g (0..n-1) with flags[g] != 0, in increasing order; *nsel receives the count
size_t select_temp_bytes(long long n) {
    size_t bytes = 0;
    hipcub::CountingInputIterator<long long> it(0);
    (void)hipcub::DeviceSelect::Flagged(nullptr, bytes, it, (const unsigned char *)nullptr, (long long *)nullptr,
                                        (long long *)nullptr, (int)n);
    return bytes;
}

hipError_t select_flagged(void *tmp, size_t bytes, const unsigned char *flags, long long *out, long long *nsel,
                          long long n, hipStream_t st) {
    hipcub::CountingInputIterator<long long> it(0);
    return hipcub::DeviceSelect::Flagged(tmp, bytes, it, flags, out, nsel, (int)n, st);
}

size_t sort_ll_temp_bytes(long long n, int end_bit) {
    size_t bytes = 0;
    (void)hipcub::DeviceRadixSort::SortPairs(nullptr, bytes, (const unsigned *)nullptr, (unsigned *)nullptr,
                                             (const long long *)nullptr, (long long *)nullptr, (int)n, 0, end_bit);
    return bytes;
}

hipError_t sort_pairs_ll(void *tmp, size_t bytes, const unsigned *kin, unsigned *kout, const long long *vin,
                         long long *vout, long long n, int end_bit, hipStream_t st) {
    return hipcub::DeviceRadixSort::SortPairs(tmp, bytes, kin, kout, vin, vout, (int)n, 0, end_bit, st);
}

size_t scan_int_temp_bytes(int n) {
    size_t bytes = 0;
    (void)hipcub::DeviceScan::ExclusiveSum(nullptr, bytes, (const int *)nullptr, (long long *)nullptr, n);
    return bytes;
}

hipError_t scan_int_exclusive(void *tmp, size_t bytes, const int *in, long long *out, int n, hipStream_t st) {
    return hipcub::DeviceScan::ExclusiveSum(tmp, bytes, in, out, n, st);
}

}  // namespace tcg
