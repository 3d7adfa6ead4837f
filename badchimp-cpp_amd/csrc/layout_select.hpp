// Which address layout a launch over some number of population fields uses.
//
// Pure host arithmetic (no HIP types), so that a CPU test can check the
// selection directly.
#pragma once
#include <cstdint>

namespace lbamd {

// The row windows address the population buffer with 32-bit byte offsets
// from its base, so every element of all n_fields fields of one buffer must
// lie below 4 GiB.
inline bool rowwin_fits(int Q, int64_t NP, int n_fields, int64_t elem_bytes) {
  return static_cast<uint64_t>(n_fields) * static_cast<uint64_t>(Q) * static_cast<uint64_t>(NP) *
             static_cast<uint64_t>(elem_bytes) <
         (1ull << 32);
}

// Layout of a launch over n_fields fields: the context's layout (0 link
// table, 1 rows, 2 row windows), with the row windows replaced by the rows
// (64-bit addresses, same sources) when the fields do not fit below 4 GiB.
inline int launch_layout(int layout, int Q, int64_t NP, int n_fields, int64_t elem_bytes) {
  return layout == 2 && !rowwin_fits(Q, NP, n_fields, elem_bytes) ? 1 : layout;
}

}  // namespace lbamd
