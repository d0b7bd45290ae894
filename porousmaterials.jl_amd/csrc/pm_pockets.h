// pm_pockets.h -- pocket blocking of an accessibility grid (host code of
// libpmhip.so, behind pm_block_pockets).
#pragma once

#include <cstdint>
#include <string>

namespace pmhost {

// The steps of compute_accessibility_grid after the energy grid
// (src/grid.jl:383-646, 704-764).  Returns "" or an error message.
//   n:       grid points per axis (data i fastest)
//   block:   false -> out_acc = data < tol; true -> flood fill, segment
//            graph, simple cycles, blocked segments
//   out_seg: nullable; the flood fill's labels (block only)
//   counts:  nullable; segments, segments blocked, simple cycles (block only,
//            zeros otherwise)
std::string block_pockets(const int32_t n[3], const double* data, double tol, bool block,
                          int64_t max_cycles, uint8_t* out_acc, int32_t* out_seg, int64_t* counts);

}  // namespace pmhost
