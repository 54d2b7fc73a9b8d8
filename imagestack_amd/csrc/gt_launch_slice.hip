// Slice launchers: LDS-tiled slice for d <= 6 and vd <= 4 (single image or a
// batch), direct slice otherwise (d > 6 in gt_launch_slice_hi.hip).
#include "gt_launch.h"

namespace gt {

hipError_t launch_slice(int d, bool packed, bool bilat, bool tiled, int storage, const PixSrc &src, int n, int row_w,
                        int tile_w, int vd, int vs, const LatConst &lc, const KeyPlan &kp, const TableArgs &ta,
                        const void *vals, const unsigned *sbits, const OutDst &dst, Counters *ctr, bool count,
                        int wpc, hipStream_t st) {
    if (n <= 0) return hipSuccess;
    if (d > 6) return launch_slice_direct_hi(d, packed, bilat, storage, src, n, vd, vs, lc, kp, ta, vals, sbits, dst,
                                             ctr, count, st);
    const int cnt = count ? 1 : 0;
    tiled = tiled && vs <= 4;
    if (tiled) {
        const int64_t ntiles = pixel_blocks(n, row_w, tile_w);
        const BatchArgs none{nullptr, 0};
#define GT_SLT(D, P, B, T)                                                                                      \
    {                                                                                                           \
        constexpr int W_ = KeyOps<D, P>::W;                                                                     \
        if (vs <= 2) hipLaunchKernelGGL((k_slice_tile<D, P, B, T, 2, kTileKeys, false>), dim3(ntiles),           \
                                        dim3(kBlock), occupancy_pad(slice_tile_lds<W_, 2, kTileKeys>(), wpc),    \
                                        st, src, n, row_w, tile_w, none, vd, lc, kp, ta, (const T *)vals, sbits, \
                                        dst, ctr, cnt);                                                          \
        else hipLaunchKernelGGL((k_slice_tile<D, P, B, T, 4, kTileKeys, false>), dim3(ntiles), dim3(kBlock),     \
                                occupancy_pad(slice_tile_lds<W_, 4, kTileKeys>(), wpc), st, src, n, row_w,       \
                                tile_w, none, vd, lc, kp, ta, (const T *)vals, sbits, dst, ctr, cnt);            \
    }
#define GT_SLT_S(D, P, B)                                              \
    {                                                                  \
        if (storage == 0) GT_SLT(D, P, B, float)                       \
        else if (storage == 1) GT_SLT(D, P, B, __hip_bfloat16)         \
        else GT_SLT(D, P, B, __half)                                   \
    }
#define GT_SLICE_TILE(D)                                                                 \
    {                                                                                    \
        if (packed) { if (bilat) GT_SLT_S(D, true, true) else GT_SLT_S(D, true, false) } \
        else { if (bilat) GT_SLT_S(D, false, true) else GT_SLT_S(D, false, false) }      \
    }
        GT_DISPATCH_DT(d, GT_SLICE_TILE)
#undef GT_SLT
#undef GT_SLT_S
#undef GT_SLICE_TILE
    } else {
        const dim3 g(blocks_for(n, kBlock)), b(kBlock);
#define GT_SLD(D, P, B, T) hipLaunchKernelGGL((k_slice_direct<D, P, B, T>), g, b, 0, st, src, n, vd, vs, lc, kp, ta, \
                                              (const T *)vals, sbits, dst, ctr, cnt)
#define GT_SLD_S(D, P, B)                                              \
    {                                                                  \
        if (storage == 0) GT_SLD(D, P, B, float);                      \
        else if (storage == 1) GT_SLD(D, P, B, __hip_bfloat16);        \
        else GT_SLD(D, P, B, __half);                                  \
    }
#define GT_SLICE_DIRECT(D)                                                               \
    {                                                                                    \
        if (packed) { if (bilat) GT_SLD_S(D, true, true) else GT_SLD_S(D, true, false) } \
        else { if (bilat) GT_SLD_S(D, false, true) else GT_SLD_S(D, false, false) }      \
    }
        GT_DISPATCH_DT(d, GT_SLICE_DIRECT)
#undef GT_SLD
#undef GT_SLD_S
#undef GT_SLICE_DIRECT
    }
    return hipGetLastError();
}

hipError_t launch_slice_batch(int d, int storage, const PixSrc &tmpl, const BatchArgs &ba, int ntiles, int tile_w,
                              int vd, const LatConst &lc, const KeyPlan &kp, const TableArgs &ta, const void *vals,
                              const unsigned *sbits, const OutDst &tmpl_dst, Counters *ctr, bool count, int wpc,
                              hipStream_t st) {
    if (ntiles <= 0) return hipSuccess;
    if (vd > 4 || !kp.packed) return hipErrorInvalidValue;
    const int cnt = count ? 1 : 0;
#define GT_SLB(D, T)                                                                                            \
    {                                                                                                           \
        if (vd <= 2) hipLaunchKernelGGL((k_slice_tile<D, true, true, T, 2, kTileKeys, true>), dim3(ntiles),      \
                                        dim3(kBlock), occupancy_pad(slice_tile_lds<1, 2, kTileKeys>(), wpc), st, \
                                        tmpl, 0, 0, tile_w, ba, vd, lc, kp, ta, (const T *)vals, sbits,          \
                                        tmpl_dst, ctr, cnt);                                                     \
        else hipLaunchKernelGGL((k_slice_tile<D, true, true, T, 4, kTileKeys, true>), dim3(ntiles),              \
                                dim3(kBlock), occupancy_pad(slice_tile_lds<1, 4, kTileKeys>(), wpc), st, tmpl,   \
                                0, 0, tile_w, ba, vd, lc, kp, ta, (const T *)vals, sbits, tmpl_dst, ctr, cnt);   \
    }
#define GT_SLB_S(D)                                          \
    {                                                        \
        if (storage == 0) GT_SLB(D, float)                   \
        else if (storage == 1) GT_SLB(D, __hip_bfloat16)     \
        else GT_SLB(D, __half)                               \
    }
    GT_DISPATCH_DT(d, GT_SLB_S)
#undef GT_SLB
#undef GT_SLB_S
    return hipGetLastError();
}

} // namespace gt
