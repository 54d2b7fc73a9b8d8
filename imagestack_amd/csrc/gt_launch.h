// Shared helpers of the kernel launchers (gt_launch_*.hip).
#pragma once
#include "gt_kernels.h"

namespace gt {

inline unsigned blocks_for(int64_t n, int b) { return (unsigned)((n + b - 1) / b); }

// Waves-per-CU knob of the pixel and neighbour kernels: dynamic LDS (unused by
// the kernel) sized so that at most waves_per_cu / 4 workgroups of 256
// threads fit the CU's 160 KiB of LDS.  A workgroup holds at most 64 KiB, so
// the cap starts at 3 workgroups (12 waves); 0 leaves occupancy to registers.
constexpr int64_t kLdsPerCU = 160 * 1024;
inline unsigned occupancy_pad(size_t static_lds, int waves_per_cu) {
    if (waves_per_cu <= 0) return 0;
    int wgs = waves_per_cu / (kBlock / kWave);
    if (wgs < 3) wgs = 3;
    int64_t pad = kLdsPerCU / wgs - (int64_t)static_lds - 2048;
    if (pad > 65536 - (int64_t)static_lds) pad = 65536 - (int64_t)static_lds;
    return pad > 0 ? (unsigned)pad : 0u;
}

#define GT_DISPATCH_D(d, MACRO)                                                      \
    switch (d) {                                                                     \
    case 1: MACRO(1); break;  case 2: MACRO(2); break;  case 3: MACRO(3); break;     \
    case 4: MACRO(4); break;  case 5: MACRO(5); break;  case 6: MACRO(6); break;     \
    case 7: MACRO(7); break;  case 8: MACRO(8); break;  case 9: MACRO(9); break;     \
    case 10: MACRO(10); break; case 11: MACRO(11); break; case 12: MACRO(12); break; \
    case 13: MACRO(13); break; case 14: MACRO(14); break; case 15: MACRO(15); break; \
    case 16: MACRO(16); break;                                                       \
    default: return hipErrorInvalidValue;                                            \
    }

#define GT_DISPATCH_DT(d, MACRO)                                                     \
    switch (d) {                                                                     \
    case 1: MACRO(1); break;  case 2: MACRO(2); break;  case 3: MACRO(3); break;     \
    case 4: MACRO(4); break;  case 5: MACRO(5); break;  case 6: MACRO(6); break;     \
    default: return hipErrorInvalidValue;                                            \
    }

inline int64_t pixel_blocks(int n, int row_w, int &tile_w) {
    static_assert(kBlock == kTilePixels, "a tile is one workgroup's pixels");
    return pixel_tiles(n, row_w, tile_w);
}


// Splat launcher for d = 7..10 with the LDS tile table (gt_launch_mid.hip).
hipError_t launch_splat_tile_mid(int d, bool packed, bool bilat, const PixSrc &src, int n, int row_w, int tile_w,
                                 int vd, const LatConst &lc, const KeyPlan &kp, const ValScale *vsc,
                                 const TableArgs &ta, const Replay &rp, Counters *ctr, bool count, int wpc,
                                 hipStream_t st);

// Launchers for the lattice dimensions without an LDS tile table (d > 6).
hipError_t launch_splat_direct_hi(int d, bool packed, bool bilat, const PixSrc &src, int n, int row_w, int tile_w,
                                  int vd, const LatConst &lc, const KeyPlan &kp, const ValScale *vsc,
                                  const TableArgs &ta, Counters *ctr, bool count, hipStream_t st);
hipError_t launch_slice_direct_hi(int d, bool packed, bool bilat, int storage, const PixSrc &src, int n, int vd,
                                  int vs, const LatConst &lc, const KeyPlan &kp, const TableArgs &ta, const void *vals,
                                  const unsigned *sbits, const OutDst &dst, Counters *ctr, bool count, hipStream_t st);
hipError_t launch_slice_direct_top(int d, bool packed, bool bilat, int storage, const PixSrc &src, int n, int vd,
                                   int vs, const LatConst &lc, const KeyPlan &kp, const TableArgs &ta,
                                   const void *vals, const unsigned *sbits, const OutDst &dst, Counters *ctr, bool count,
                                   hipStream_t st);

} // namespace gt
