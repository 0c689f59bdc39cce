// Stable sort of point indices by bin (rocPRIM's LSD radix sort), used by the
// NUFFT plan's setpts.  A stable order keeps the points of one bin in index
// order: the layout (and with it every sum over a bin's points) is the same
// on every call with the same points, and consecutive points of one image
// that share a bin sit next to each other.
#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>
#include <rocprim/iterator/counting_iterator.hpp>

#include "common.h"

namespace cov3d {

// Sorts (key = bin of point j, value = j) for j < n by the low nbits of the key.  With tmp ==
// nullptr only tmp_bytes is set.  keys_out receives the sorted bins, idx_out the point indices.
hipError_t sort_points_by_bin(void* tmp, size_t& tmp_bytes, const unsigned* keys_in, unsigned* keys_out,
                              int* idx_out, int n, int nbits, hipStream_t st) {
  return rocprim::radix_sort_pairs(tmp, tmp_bytes, keys_in, keys_out, rocprim::counting_iterator<int>(0), idx_out,
                                   (unsigned)n, 0u, (unsigned)nbits, st);
}

// Exclusive prefix sum of n ints (device-wide).  With tmp == nullptr only tmp_bytes is set.
hipError_t exclusive_scan_ints(void* tmp, size_t& tmp_bytes, const int* in, int* out, int n, hipStream_t st) {
  return rocprim::exclusive_scan(tmp, tmp_bytes, in, out, 0, (size_t)n, rocprim::plus<int>(), st);
}

}  // namespace cov3d
