// Splat launchers: LDS-tiled splat for d <= 6 and vd <= 4 (single image or a
// batch; d = 7..10 in gt_launch_mid.hip), direct splat otherwise (d > 6 in
// gt_launch_splat_hi.hip).
#include "gt_launch.h"

namespace gt {

hipError_t launch_splat(int d, bool packed, bool bilat, bool dedup, const PixSrc &src, int n, int row_w, int tile_w,
                        int vd, const LatConst &lc, const KeyPlan &kp, const ValScale *vsc, const TableArgs &ta,
                        const Replay &rp, Counters *ctr, bool count, int wpc, hipStream_t st) {
    if (d > 6) {
        if (d <= 10 && dedup && vd <= 4)
            return launch_splat_tile_mid(d, packed, bilat, src, n, row_w, tile_w, vd, lc, kp, vsc, ta, rp, ctr,
                                         count, wpc, st);
        return launch_splat_direct_hi(d, packed, bilat, src, n, row_w, tile_w, vd, lc, kp, vsc, ta, ctr, count, st);
    }
    const int64_t ntiles = pixel_blocks(n, row_w, tile_w);
    if (ntiles == 0) return hipSuccess;
    const int cnt = count ? 1 : 0;
    const bool tiled = dedup && vd <= 4;
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
#define GT_SPLAT_D(D, P, B) hipLaunchKernelGGL((k_splat_direct<D, P, B>), dim3(ntiles), dim3(kBlock), 0, st, src, n, \
                                               row_w, tile_w, vd, lc, kp, vsc, ta, ctr, cnt)
#define GT_SPLAT_DIRECT(D)                                                                     \
    {                                                                                          \
        if (packed) { if (bilat) GT_SPLAT_D(D, true, true); else GT_SPLAT_D(D, true, false); } \
        else { if (bilat) GT_SPLAT_D(D, false, true); else GT_SPLAT_D(D, false, false); }      \
    }
    if (tiled) {
        GT_DISPATCH_DT(d, GT_SPLAT_TILE)
    } else {
        GT_DISPATCH_DT(d, GT_SPLAT_DIRECT)
    }
#undef GT_SPLAT_T
#undef GT_SPLAT_TILE
#undef GT_SPLAT_D
#undef GT_SPLAT_DIRECT
    return hipGetLastError();
}

hipError_t launch_splat_batch(int d, const PixSrc &tmpl, const BatchArgs &ba, int ntiles, int tile_w, int vd,
                              const LatConst &lc, const KeyPlan &kp, const ValScale *vsc, const TableArgs &ta,
                              const Replay &rp, Counters *ctr, bool count, int wpc, hipStream_t st) {
    if (ntiles <= 0) return hipSuccess;
    if (vd > 4 || !kp.packed) return hipErrorInvalidValue;
    const int cnt = count ? 1 : 0;
#define GT_SPLAT_B(D)                                                                                           \
    {                                                                                                           \
        if (vd <= 2) hipLaunchKernelGGL((k_splat_tile<D, true, true, 2, kTileKeys, true>), dim3(ntiles),         \
                                        dim3(kBlock), occupancy_pad(splat_tile_lds<1, 2, kTileKeys>(), wpc), st, \
                                        tmpl, 0, 0, tile_w, ba, vd, lc, kp, vsc, ta, rp, ctr, cnt);                  \
        else hipLaunchKernelGGL((k_splat_tile<D, true, true, 4, kTileKeys, true>), dim3(ntiles), dim3(kBlock),   \
                                occupancy_pad(splat_tile_lds<1, 4, kTileKeys>(), wpc), st, tmpl, 0, 0, tile_w,   \
                                ba, vd, lc, kp, vsc, ta, rp, ctr, cnt);                                              \
    }
    GT_DISPATCH_DT(d, GT_SPLAT_B)
#undef GT_SPLAT_B
    return hipGetLastError();
}

} // namespace gt
