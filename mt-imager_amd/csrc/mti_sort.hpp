// mti_sort.hpp -- stable radix sort shared by the row order and the tile bucketing.
#pragma once
#include "mti_internal.hpp"

namespace mti {
// Sorts (keys, vals) by the low key_bits bits of keys, stably.  Ping-pongs
// between (keys, vals) and (keys_tmp, vals_tmp); *result_in_tmp tells which
// pair holds the result.  Scratch slot SLOT_RADIX holds the digit counts.
mti_status radix_sort_pairs(mti_ctx *ctx, uint32_t *keys, int32_t *vals, uint32_t *keys_tmp,
                            int32_t *vals_tmp, int64_t n, int key_bits, hipStream_t st,
                            bool *result_in_tmp);
void plan_radix_sort(int64_t n, ScratchPlan &pl);

// Baseline-major order of nrows rows into index_out (rows with a1 == a2 or an
// antenna out of range sort behind the kept rows) and the kept count into
// *kept (device).  tmp: 3 * nrows 32-bit words.  No host synchronisation.
mti_status sort_rows_device(mti_ctx *ctx, const int32_t *ant1, const int32_t *ant2, int64_t nrows, int nant,
                            int32_t *index_out, uint32_t *tmp, unsigned long long *kept, hipStream_t st);
void plan_sort_rows_device(int64_t nrows, ScratchPlan &pl);
}  // namespace mti
