// The vertex-key packing plan shared by the host planner (gt_plan.h) and the
// kernels.  Plain C++ (no HIP) so host-only code and tests can include it.
#pragma once

#include <stdint.h>

namespace gt {

constexpr int kMaxDH = 16;

// How a vertex key is packed.  Packed format: coordinate i is stored as
// (c_i - lo[i]) in the bits span[i] at shift[i] of one 64-bit word
// (coordinate 0 most significant, bit 63 clear), so the word itself sorts
// lexicographically.  Wide format: int16 fields, re-packed by the same
// lo/shift/span into `sort_words` words for the radix sort.
struct KeyPlan {
    int lo[kMaxDH];
    uint32_t span[kMaxDH];
    int shift[kMaxDH];     // bit offset in the packed big integer (word = shift/64 for the sort)
    int packed;            // 1: table keys are the packed word
    int bits;              // total packed bits (coordinates + image index)
    int coord_bits;        // bits of the coordinate fields (bits [0, coord_bits))
    int img_bits;          // batched lattices: image index in bits [coord_bits, bits)
};

} // namespace gt
