// Direct slice for lattice dimensions 7..11 (12..16 in gt_launch_slice_hi2.hip:
// the two halves compile in parallel).
#include "gt_launch.h"

namespace gt {

hipError_t launch_slice_direct_hi(int d, bool packed, bool bilat, int storage, const PixSrc &src, int n, int vd,
                                  int vs, const LatConst &lc, const KeyPlan &kp, const TableArgs &ta, const void *vals,
                                  const unsigned *sbits, const OutDst &dst, Counters *ctr, bool count, hipStream_t st) {
    const int cnt = count ? 1 : 0;
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
    switch (d) {
    case 7: GT_SLICE_DIRECT(7); break;    case 8: GT_SLICE_DIRECT(8); break;
    case 9: GT_SLICE_DIRECT(9); break;    case 10: GT_SLICE_DIRECT(10); break;
    case 11: GT_SLICE_DIRECT(11); break;
    default:
        return launch_slice_direct_top(d, packed, bilat, storage, src, n, vd, vs, lc, kp, ta, vals, sbits, dst, ctr,
                                       count, st);
    }
#undef GT_SLD
#undef GT_SLD_S
#undef GT_SLICE_DIRECT
    return hipGetLastError();
}

} // namespace gt
