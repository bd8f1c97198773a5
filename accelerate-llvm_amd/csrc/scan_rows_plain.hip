// scanl1 / scanr1 along the rows of a rank >= 2 array (kernels in
// scan_rows.hpp) for 16-byte-aligned arrays: inclusive values only, so the
// kernels carry no seed, exclusive or row-total code.  Built apart from
// scan_rows.hip so the two compile in parallel.
#include "scan_rows.hpp"

namespace acc {

int scan_rows_plain(acc_stream_t s, acc_type t, acc_op op, bool reverse, const void* in, int64_t outer,
                    int64_t inner, void* out) {
  return with_type(t, [&](auto z) {
    using T = decltype(z);
    return with_op<T>(op, [&](auto f) {
      using Op = decltype(f);
      RowScanArgs<T> a{};
      a.in = (const T*)in;
      a.out = (T*)out;
      a.n = outer * inner;
      a.inner = inner;
      a.out_row = inner;
      return reverse ? scan_rows_launch<T, Op, true, true, true>(s, a)
                     : scan_rows_launch<T, Op, false, true, true>(s, a);
    });
  });
}

}  // namespace acc
