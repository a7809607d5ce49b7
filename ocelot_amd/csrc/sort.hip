// Stable key/value radix sort used to order a resident bunch by mesh cell.
// rocPRIM's device radix sort is deterministic (stable, fixed decomposition),
// so the particle order, and with it every reduction order of the kick, is
// reproducible run to run.  Kept in its own translation unit: the rocPRIM
// templates are slow to compile.
#include <hip/hip_runtime.h>
#include <rocprim/device/device_radix_sort.hpp>

#include <cstddef>
#include <cstdint>

namespace ocsc {

// temp == nullptr: *temp_bytes receives the scratch size.
hipError_t sort_pairs_u32(void* temp, size_t* temp_bytes, const uint32_t* keys_in, uint32_t* keys_out,
                          const uint32_t* vals_in, uint32_t* vals_out, int64_t n, int end_bit, hipStream_t st) {
  return rocprim::radix_sort_pairs(temp, *temp_bytes, keys_in, keys_out, vals_in, vals_out, (size_t)n, 0, end_bit,
                                   st);
}

}  // namespace ocsc
