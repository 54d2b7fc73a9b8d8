// LDS-tiled splat for lattice dimensions 7..10 (the range where
// JointBilateral's AUTO picks the lattice above the plain bilateral, e.g.
// non-local means on 6 patch dimensions + x, y): the same kernel as d <= 6,
// in its own translation unit to keep compile times apart.  (The tiled slice
// does not pay at these dimensions: at d = 8 on 2048^2 it took 0.71 ms
// against 0.53 ms for the direct slice, so a slice at other positions stays
// direct; at the splat's own positions it reads the splat's replay record.)
#include "gt_launch.h"

namespace gt {

hipError_t launch_splat_tile_mid(int d, bool packed, bool bilat, const PixSrc &src, int n, int row_w, int tile_w,
                                 int vd, const LatConst &lc, const KeyPlan &kp, const ValScale *vsc,
                                 const TableArgs &ta, const Replay &rp, Counters *ctr, bool count, int wpc,
                                 hipStream_t st) {
    const int64_t ntiles = pixel_blocks(n, row_w, tile_w);
    if (ntiles == 0) return hipSuccess;
    if (vd > 4) return hipErrorInvalidValue;
    const int cnt = count ? 1 : 0;
    const BatchArgs none{nullptr, 0};
#define GT_SPLAT_T(D, P, B)                                                                                     \
    {                                                                                                           \
        constexpr int W_ = KeyOps<D, P>::W;                                                                     \
        if (vd <= 2) hipLaunchKernelGGL((k_splat_tile<D, P, B, 2, kTileKeys, false>), dim3(ntiles),             \
                                        dim3(kBlock), occupancy_pad(splat_tile_lds<W_, 2, kTileKeys>(), wpc),    \
                                        st, src, n, row_w, tile_w, none, vd, lc, kp, vsc, ta, rp, ctr, cnt);         \
        else hipLaunchKernelGGL((k_splat_tile<D, P, B, 4, kTileKeys, false>), dim3(ntiles), dim3(kBlock),        \
                                occupancy_pad(splat_tile_lds<W_, 4, kTileKeys>(), wpc), st, src, n, row_w,       \
                                tile_w, none, vd, lc, kp, vsc, ta, rp, ctr, cnt);                                    \
    }
#define GT_SPLAT_TILE(D)                                                                    \
    {                                                                                       \
        if (packed) { if (bilat) GT_SPLAT_T(D, true, true) else GT_SPLAT_T(D, true, false) } \
        else { if (bilat) GT_SPLAT_T(D, false, true) else GT_SPLAT_T(D, false, false) }      \
    }
    switch (d) {
    case 7: GT_SPLAT_TILE(7); break;  case 8: GT_SPLAT_TILE(8); break;
    case 9: GT_SPLAT_TILE(9); break;  case 10: GT_SPLAT_TILE(10); break;
    default: return hipErrorInvalidValue;
    }
#undef GT_SPLAT_T
#undef GT_SPLAT_TILE
    return hipGetLastError();
}

} // namespace gt
