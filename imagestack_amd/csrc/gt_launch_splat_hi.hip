// Direct splat for lattice dimensions 7..16.
#include "gt_launch.h"

namespace gt {

hipError_t launch_splat_direct_hi(int d, bool packed, bool bilat, const PixSrc &src, int n, int row_w, int tile_w,
                                  int vd, const LatConst &lc, const KeyPlan &kp, const ValScale *vsc,
                                  const TableArgs &ta, Counters *ctr, bool count, hipStream_t st) {
    const int64_t nblk = pixel_blocks(n, row_w, tile_w);
    if (nblk == 0) return hipSuccess;
    const int cnt = count ? 1 : 0;
#define GT_SPLAT_D(D, P, B) hipLaunchKernelGGL((k_splat_direct<D, P, B>), dim3(nblk), dim3(kBlock), 0, st, src, n, \
                                               row_w, tile_w, vd, lc, kp, vsc, ta, ctr, cnt)
#define GT_SPLAT_DIRECT(D)                                                                     \
    {                                                                                          \
        if (packed) { if (bilat) GT_SPLAT_D(D, true, true); else GT_SPLAT_D(D, true, false); } \
        else { if (bilat) GT_SPLAT_D(D, false, true); else GT_SPLAT_D(D, false, false); }      \
    }
    switch (d) {
    case 7: GT_SPLAT_DIRECT(7); break;    case 8: GT_SPLAT_DIRECT(8); break;
    case 9: GT_SPLAT_DIRECT(9); break;    case 10: GT_SPLAT_DIRECT(10); break;
    case 11: GT_SPLAT_DIRECT(11); break;  case 12: GT_SPLAT_DIRECT(12); break;
    case 13: GT_SPLAT_DIRECT(13); break;  case 14: GT_SPLAT_DIRECT(14); break;
    case 15: GT_SPLAT_DIRECT(15); break;  case 16: GT_SPLAT_DIRECT(16); break;
    default: return hipErrorInvalidValue;
    }
#undef GT_SPLAT_D
#undef GT_SPLAT_DIRECT
    return hipGetLastError();
}

} // namespace gt
