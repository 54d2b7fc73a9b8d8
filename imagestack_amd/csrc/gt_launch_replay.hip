// Slice launcher from the tiled splat's replay record (d <= 10, vd <= 4,
// single image or batch).  The kernel builds no keys, so it is templated on
// the lattice dimension, the position form, the storage type, the channel
// group and the batch form only.
#include "gt_launch.h"

namespace gt {

hipError_t launch_slice_replay(int d, bool bilat, int storage, const PixSrc &src, int n, int row_w, int tile_w,
                               const BatchArgs &ba, int ntiles, int vd, const LatConst &lc, const KeyPlan &kp,
                               const TableArgs &ta, const Replay &rp, const void *vals, const unsigned *sbits,
                               const OutDst &dst, Counters *ctr, Counters *hctr, int wpc, hipStream_t st) {
    const bool batch = ba.count > 0;
    if (vd > 4 || d < 1 || d > 10 || !rp.rec) return hipErrorInvalidValue;
    int64_t tiles = ntiles;
    if (!batch) tiles = pixel_blocks(n, row_w, tile_w);
    if (tiles <= 0) return hipSuccess;
    if (tiles * kBlock > rp.stride) return hipErrorInvalidValue;   // not the splat's decomposition
#define GT_RP(D, P, T, LVD, B)                                                                               \
    hipLaunchKernelGGL((k_slice_replay<D, P, T, LVD, kTileKeys, B>), dim3((unsigned)tiles), dim3(kBlock),     \
                       occupancy_pad((size_t)kTileKeys * LVD * 4, wpc), st, src, n, row_w, tile_w, ba, vd, lc, \
                       kp, ta, rp, (const T *)vals, sbits, dst, ctr, hctr)
#define GT_RP_B(D, T, LVD)                          \
    {                                               \
        if (batch) GT_RP(D, true, T, LVD, true);    \
        else if (bilat) GT_RP(D, true, T, LVD, false); \
        else GT_RP(D, false, T, LVD, false);        \
    }
#define GT_RP_V(D, T)                    \
    {                                    \
        if (vd <= 2) GT_RP_B(D, T, 2)    \
        else GT_RP_B(D, T, 4)            \
    }
#define GT_RP_S(D)                                    \
    {                                                 \
        if (storage == 0) GT_RP_V(D, float)           \
        else if (storage == 1) GT_RP_V(D, __hip_bfloat16) \
        else GT_RP_V(D, __half)                       \
    }
    switch (d) {
    case 1: GT_RP_S(1); break;  case 2: GT_RP_S(2); break;  case 3: GT_RP_S(3); break;
    case 4: GT_RP_S(4); break;  case 5: GT_RP_S(5); break;  case 6: GT_RP_S(6); break;
    case 7: GT_RP_S(7); break;  case 8: GT_RP_S(8); break;  case 9: GT_RP_S(9); break;
    case 10: GT_RP_S(10); break;
    default: return hipErrorInvalidValue;
    }
#undef GT_RP
#undef GT_RP_B
#undef GT_RP_V
#undef GT_RP_S
    return hipGetLastError();
}

} // namespace gt
