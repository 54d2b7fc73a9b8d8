// Host-side key packing plan.  Header-only plain C++ (no HIP), so the CPU
// tests compile it with g++.
//
// The lattice key of a point is within 3(d+1) of its elevation, coordinate by
// coordinate (rounding to the nearest multiple of d+1, the walk back onto the
// hyperplane and the canonical offset, src/Permutohedral.h:305-369), and the
// elevation is linear in the position (:293-298).  Interval arithmetic over
// per-channel position bounds therefore bounds every key coordinate; each
// coordinate gets just the bits its range needs, coordinate 0 most
// significant.  If the total fits 63 bits a key is one word that sorts as an
// integer.
#pragma once

#include <cmath>
#include <cstdint>

#include "gt_keyplan.h"

namespace gt {

// src/Permutohedral.h:257-274, same FP32 expressions.
inline void plan_scale_factors(int d, float *sf) {
    const float s23 = sqrtf((float)(2.0 / 3));
    for (int i = 0; i < d; i++) {
        float v = 1.0f / (sqrtf((float)(i + 1) * (float)(i + 2)));
        v *= (float)(d + 1) * s23;
        sf[i] = v;
    }
}

inline int plan_bitlen(uint32_t v) {
    int b = 0;
    while (v) { b++; v >>= 1; }
    return b;
}

// lo/hi: per-channel bounds of the scaled positions (d values each).
// images > 1: a batch of independent lattices, the image index stored above
// the coordinates (most significant, so a batch sorts image by image).
inline void plan_from_bounds(int d, const float *sf, const float *lo, const float *hi, KeyPlan &kp,
                             int images = 1) {
    double A[kMaxDH + 1][kMaxDH];
    for (int i = 0; i <= d; i++)
        for (int j = 0; j < d; j++) A[i][j] = 0.0;
    // Rows of the elevation e = A p, following the recurrence of :293-298.
    A[d][d - 1] = -(double)d * sf[d - 1];
    for (int i = d - 1; i > 0; i--) {
        for (int j = 0; j < d; j++) A[i][j] = A[i + 1][j];
        A[i][i - 1] -= (double)i * sf[i - 1];
        A[i][i] += (double)(i + 2) * sf[i];
    }
    for (int j = 0; j < d; j++) A[0][j] = A[1][j];
    A[0][0] += 2.0 * sf[0];

    bool finite = true;
    for (int j = 0; j < d; j++) finite = finite && std::isfinite(lo[j]) && std::isfinite(hi[j]);
    const int lim = 32767 - 2 * (d + 1);
    int total = 0, bits[kMaxDH];
    for (int i = 0; i < d; i++) {
        int klo = -lim, khi = lim;
        if (finite) {
            double elo = 0, ehi = 0, mag = 0;
            for (int j = 0; j < d; j++) {
                const double a = A[i][j] * lo[j], b = A[i][j] * hi[j];
                elo += std::fmin(a, b);
                ehi += std::fmax(a, b);
                mag += std::fabs(A[i][j]) * std::fmax(std::fabs((double)lo[j]), std::fabs((double)hi[j]));
            }
            const double m = 1e-4 * mag + 1.0;   // far above the FP32 rounding of the elevation
            const double flo = std::floor(elo - m) - 3.0 * (d + 1), fhi = std::ceil(ehi + m) + 3.0 * (d + 1);
            klo = (int)std::fmax(flo, (double)-lim);
            khi = (int)std::fmin(fhi, (double)lim);
            if (khi < klo) khi = klo;
        }
        kp.lo[i] = klo;
        bits[i] = plan_bitlen((uint32_t)(khi - klo));
        kp.span[i] = bits[i] >= 32 ? 0xffffffffu : ((1u << bits[i]) - 1u);
    }
    for (int i = d - 1; i >= 0; i--) {
        kp.shift[i] = total;
        total += bits[i];
    }
    for (int i = d; i < kMaxDH; i++) { kp.lo[i] = 0; kp.span[i] = 0; kp.shift[i] = 0; }
    kp.coord_bits = total;
    kp.img_bits = images > 1 ? plan_bitlen((uint32_t)(images - 1)) : 0;
    kp.bits = total + kp.img_bits;
    kp.packed = kp.bits <= 63 ? 1 : 0;
}

} // namespace gt
