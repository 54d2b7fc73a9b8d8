// Direct slice for lattice dimensions 12..16 (7..11 in gt_launch_slice_hi.hip).
#include "gt_launch.h"

namespace gt {

hipError_t launch_slice_direct_top(int d, bool packed, bool bilat, int storage, const PixSrc &src, int n, int vd,
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
    case 12: GT_SLICE_DIRECT(12); break;
    case 13: GT_SLICE_DIRECT(13); break;  case 14: GT_SLICE_DIRECT(14); break;
    case 15: GT_SLICE_DIRECT(15); break;  case 16: GT_SLICE_DIRECT(16); break;
    default: return hipErrorInvalidValue;
    }
#undef GT_SLD
#undef GT_SLD_S
#undef GT_SLICE_DIRECT
    return hipGetLastError();
}

} // namespace gt
