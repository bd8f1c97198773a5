// Kernel argument block of the run-time compiled skeletons.  The same
// definition is compiled into the library (host) and into every kernel
// compiled at run time (device), so its layout is fixed and independent of
// the element type: arrays travel as untyped field pointers (one per tuple
// field, struct of arrays, LLVM/Execute/Marshal.hs:71-100) and the generated
// code gives them their types.
#pragma once
#include <stdint.h>

namespace accjit {

constexpr int kMaxFields = 4;  // tuple fields of an element
constexpr int kMaxFree = 8;    // free arrays a function may read

// Arrays a function reads by index, `(read k i)` (the reference's free array
// environment Gamma, LLVM/CodeGen/Environment.hs:65-88).
struct Free {
  const void* p[kMaxFree];
  int64_t n[kMaxFree];
};

struct Args {
  const void* in[kMaxFields];   // manifest input fields
  void* out[kMaxFields];        // output fields
  void* sums[kMaxFields];       // scan': totals
  const void* offsets;          // foldSeg: segment offsets
  int64_t outer, inner;         // input extents (rows x innermost)
  int64_t nseg;                 // foldSeg: segments per row; permute: output size
  int64_t chunk;                // fold: elements per workgroup
  int64_t out_row;              // scan rows: output elements per row (inner, or inner + 1)
  uint32_t ntiles;              // scan: tiles; fold: workgroups
  uint32_t vec;                 // every manifest field pointer allows 4-element vector accesses
  uint32_t has_total;           // scan: 1 = write the total, 2 = write the seed, through Tot
  uint32_t exclusive;           // scan: emit exclusive values
  void* ctl;                    // SyncCtl of the stream workspace
  uint64_t* status;             // tile status words
  uint32_t* locks;              // permute: one lock word per output slot
  Free fv;
};

}  // namespace accjit
