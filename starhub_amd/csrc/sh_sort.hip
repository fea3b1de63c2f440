// Device-wide stable radix sorts and scans for the index build (one over every
// solve of a group: each solve's keys carry its own base, so its entries sort
// into one contiguous run in the same relative order as alone).
// rocPRIM's LSD radix sort is stable, which is exactly the semantics the
// reference's CUB DeviceRadixSort::SortPairs calls rely on
// (libs/star/common/algorithm_types.cuh:26-124), so the sorted node->term and
// node-pair->term lists match element for element.
#include <cstring>

#include <rocprim/rocprim.hpp>

#include "sh_kernels.hpp"

namespace sh
{
// Merge-sort limit 0: 32-bit keys always take the onesweep LSD path (stable;
// for a few hundred thousand 15..24-bit keys it is a fraction of the
// block-sort + merge-pass path rocPRIM picks below 1M items by default).
using OnesweepOnly = rocprim::radix_sort_config<rocprim::default_config, rocprim::default_config,
                                                rocprim::default_config, 0>;

// Ten bits a pass for 9..20-bit keys: two passes (three for a batch's 17..20-bit
// block keys at rocPRIM's eight), and for the 16-bit node keys the same two
// passes run faster than rocPRIM's default 8-bit kernel, which keeps a
// per-thread array in scratch (two passes 511 -> 439 us on a 32-solve batch).
// The sort kernel's shape is rocPRIM's own gfx950 choice for 32-bit keys and
// values.
using Onesweep10 = rocprim::radix_sort_config<
    rocprim::default_config, rocprim::default_config,
    rocprim::radix_sort_onesweep_config<rocprim::kernel_config<1024, 16>, rocprim::kernel_config<1024, 16>, 10,
                                        rocprim::block_radix_rank_algorithm::match>,
    0>;

hipError_t sort_pairs_u32_u32(void *temp, size_t &temp_bytes, const uint32_t *kin, uint32_t *kout,
                              const uint32_t *vin, uint32_t *vout, uint32_t n, int end_bit, hipStream_t st)
{
    if (end_bit > 8 && end_bit <= 20)
        return rocprim::radix_sort_pairs<Onesweep10>(temp, temp_bytes, kin, kout, vin, vout, n, 0u,
                                                     unsigned(end_bit), st);
    return rocprim::radix_sort_pairs<OnesweepOnly>(temp, temp_bytes, kin, kout, vin, vout, n, 0u, unsigned(end_bit),
                                                   st);
}

// The flag stage's scans cover every pixel of a group (16 M on a 32-solve
// batch). rocPRIM has no tuned scan for gfx950 (its fallback is 256 threads x
// 16 items); 32 items a thread halves the blocks the decoupled look-back
// chains through: 72 -> 48 us a scan on that batch.
using ScanConfig =
    rocprim::scan_config<256, 32, rocprim::block_load_method::block_load_transpose,
                         rocprim::block_store_method::block_store_transpose,
                         rocprim::block_scan_algorithm::using_warp_scan>;

hipError_t exclusive_scan_u32(void *temp, size_t &temp_bytes, const uint32_t *in, uint32_t *out, uint32_t n,
                              hipStream_t st)
{
    return rocprim::exclusive_scan<ScanConfig>(temp, temp_bytes, in, out, 0u, n, rocprim::plus<uint32_t>(), st);
}
} // namespace sh
