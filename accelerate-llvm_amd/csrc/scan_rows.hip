// Entry point of the rank >= 2 scans (kernels in scan_rows.hpp): the
// general kernels (seeds, exclusive values, row totals, unaligned arrays).
#include "scan_rows.hpp"

// All rows of a rank >= 2 scan in one launch (the output layouts are listed at
// the top of scan_rows.hpp).  exclusive: emit exclusive values; sums: scan' totals;
// seed_at_end: scanr z (the seed closes each output row).
int acc_scan_rows(acc_stream_t s, acc_type t, acc_op op, bool reverse, bool exclusive, const void* seed,
                  const void* in, int64_t outer, int64_t inner, void* out, int64_t out_row, void* sums,
                  bool seed_at_end) {
  using namespace acc;
  return with_type(t, [&](auto z) {
    using T = decltype(z);
    return with_op<T>(op, [&](auto f) {
      using Op = decltype(f);
      if ((uintptr_t)in % sizeof(T) || (uintptr_t)out % sizeof(T))
        return fail(ACC_ERROR_INVALID_VALUE, "scan: arrays not element-aligned");
      const bool aligned = (uintptr_t)in % 16 == 0 && (uintptr_t)out % 16 == 0;
      if (aligned && !exclusive && !seed && !sums && !seed_at_end && out_row == inner)
        return scan_rows_plain(s, t, op, reverse, in, outer, inner, out);
      RowScanArgs<T> a{};
      a.in = (const T*)in;
      a.out = (T*)out;
      a.n = outer * inner;
      a.inner = inner;
      a.out_row = out_row;
      a.exclusive = exclusive ? 1u : 0u;
      a.has_seed = seed ? 1u : 0u;
      a.seed = seed ? *(const T*)seed : T();
      if (sums) {
        a.total_out = (T*)sums;
        a.total_stride = 1;
      } else if (exclusive && !reverse) {  // scanl z: the total closes each output row
        a.total_out = (T*)out + inner;
        a.total_stride = out_row;
      }
      if (seed_at_end) {
        a.seed_out = (T*)out + inner;
        a.seed_stride = out_row;
      }
      if (reverse)
        return aligned ? scan_rows_launch<T, Op, true, true, false>(s, a)
                       : scan_rows_launch<T, Op, true, false, false>(s, a);
      return aligned ? scan_rows_launch<T, Op, false, true, false>(s, a)
                     : scan_rows_launch<T, Op, false, false, false>(s, a);
    });
  });
}
