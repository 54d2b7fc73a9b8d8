// CPU restatement of ImageStack's bilateral grid (src/DenseGrid.h) as
// GaussTransform::apply drives it in its GRID case
// (src/GaussTransform.cpp:227-297, taps = 5).
//
// TEST INFRASTRUCTURE ONLY: the oracle of the GPU grid path (tests/, the smoke
// check and bench.py's CPU leg call it to check results); the product library
// never links it.  Included by lattice_oracle.cpp and compiled with the same
// strict flags (-O2 -ffp-contract=off, no fast math).
//
// T = float: the reference's own arithmetic, step for step, including the
// double-precision intermediates of its blur (`0.5 * (a + b)` with a double
// literal); in raster order its output is bit-identical to the reference's
// (pinned by tests against oracle/_ref).  T = double: the FP64 oracle of the
// same grid: cell indices and fractions in FP32 exactly as the reference
// computes them, weights, sums, blur and slice in FP64.
#pragma once

#include <cmath>
#include <cstdint>
#include <vector>

namespace grid_oracle {

template <typename T>
struct DenseGrid {
    int d, vd;
    float sf;                       // sqrtf(1/3 + (taps-1)/4): src/DenseGrid.h:104-122
    std::vector<float> mn, mx;      // bounds of position * sf (:138-156)
    bool have = false;
    std::vector<int> sizes, stride; // stride[0] = vd, stride[i+1] = stride[i] * sizes[i] (:159-169)
    std::vector<T> g;

    DenseGrid(int d_, int vd_) : d(d_), vd(vd_), mn(d_), mx(d_), sizes(d_), stride(d_ + 1) {
        const int taps = 5;
        sf = sqrtf((float)(1.0 / 3 + (taps - 1) * 0.25));
    }

    void preview(const float *p) {
        for (int i = 0; i < d; i++) {
            const float v = p[i] * sf;
            if (!have) {
                mn[i] = v;
                mx[i] = v;
            } else {
                if (v < mn[i]) mn[i] = v;
                if (v > mx[i]) mx[i] = v;
            }
        }
        have = true;
    }

    void allocate() {
        stride[0] = vd;
        for (int i = 0; i < d; i++) {
            sizes[i] = (int)(std::ceil(mx[i] - mn[i]) + 1);
            stride[i + 1] = stride[i] * sizes[i];
        }
        g.assign((size_t)stride[d], (T)0);
    }

    // src/DenseGrid.h:297-338.  Corners whose index passes the end of a
    // dimension carry weight exactly 0 (the fraction is 0 there); the
    // reference touches the neighbouring memory with that zero weight, this
    // restatement skips them, which leaves every sum unchanged.
    void query(bool splatting, const float *p, T *val) {
        int I[16];
        float F[16], Fi[16];
        for (int i = 0; i < d; i++) {
            const float f = p[i] * sf - mn[i];
            I[i] = (short)floorf(f);
            F[i] = f - I[i];
            Fi[i] = 1 - F[i];
        }
        if (!splatting)
            for (int k = 0; k < vd; k++) val[k] = 0;
        int64_t base = 0;
        for (int i = 0; i < d; i++) base += (int64_t)I[i] * stride[i];
        for (int c = 0; c < (1 << d); c++) {
            T w = 1;
            int64_t idx = base;
            bool inside = true;
            for (int j = 0; j < d; j++) {
                if (c & (1 << j)) {
                    idx += stride[j];
                    w *= (T)F[j];
                    inside = inside && I[j] + 1 < sizes[j];
                } else {
                    w *= (T)Fi[j];
                }
            }
            if (!inside) continue;
            if (splatting)
                for (int k = 0; k < vd; k++) g[(size_t)idx + k] += w * val[k];
            else
                for (int k = 0; k < vd; k++) val[k] += w * g[(size_t)idx + k];
        }
    }

    // src/DenseGrid.h:214-295: per dimension, taps/2 = 2 rounds of an
    // in-place [1/4 1/2 1/4] along every row (zero outside the grid).
    void blur() {
        const int64_t total = stride[d];
        std::vector<T> t1(vd), t2(vd);
        for (int j = 0; j < d; j++) {
            const int64_t s = stride[j], n = sizes[j];
            const int64_t rows = total / ((int64_t)vd * n);
            for (int64_t row = 0; row < rows; row++) {
                // the row's start: the digits of `row` over the other dimensions
                int64_t rem = row, start = 0;
                for (int i = 0; i < d; i++) {
                    if (i == j) continue;
                    start += (rem % sizes[i]) * stride[i];
                    rem /= sizes[i];
                }
                for (int it = 0; it < 2; it++) {
                    T *r = g.data() + start;
                    for (int k = 0; k < vd; k++) t1[k] = r[k] / 2;
                    for (int64_t i = 0; i < n - 1; i++) {
                        for (int k = 0; k < vd; k++) t2[k] = 0.5 * (r[k] + r[k + s]);
                        for (int k = 0; k < vd; k++) r[k] = 0.5 * (t1[k] + t2[k]);
                        std::swap(t1, t2);
                        r += s;
                    }
                    for (int k = 0; k < vd; k++) r[k] = 0.5 * (t1[k] + 0.5 * r[k]);
                }
            }
        }
    }
};

// GaussTransform::apply(slice, splat, values, sigma, GRID) on dense planar
// inputs of n (splat) and n_slice (slice) pixels in raster order; slice ==
// nullptr slices at the splat positions.  Returns the number of grid cells.
template <typename T>
int64_t transform(int n, int d, int vd, const float *splat, int64_t splat_cs, const float *values,
                  int64_t values_cs, int n_slice, const float *slice, int64_t slice_cs, const float *inv_sigma,
                  T *out, int64_t out_cs) {
    DenseGrid<T> grid(d, vd);
    float pos[16];
    T val[16];
    auto load = [&](const float *src, int64_t cs, int i) {
        for (int c = 0; c < d; c++) pos[c] = src[c * cs + i] * inv_sigma[c];
    };
    for (int i = 0; i < n; i++) {
        load(splat, splat_cs, i);
        grid.preview(pos);
    }
    if (slice)
        for (int i = 0; i < n_slice; i++) {
            load(slice, slice_cs, i);
            grid.preview(pos);
        }
    grid.allocate();
    for (int i = 0; i < n; i++) {
        load(splat, splat_cs, i);
        for (int c = 0; c < vd; c++) val[c] = (T)values[c * values_cs + i];
        grid.query(true, pos, val);
    }
    grid.blur();
    const int ns = slice ? n_slice : n;
    const float *src = slice ? slice : splat;
    const int64_t scs = slice ? slice_cs : splat_cs;
    for (int i = 0; i < ns; i++) {
        load(src, scs, i);
        grid.query(false, pos, val);
        for (int c = 0; c < vd; c++) out[c * out_cs + i] = val[c];
    }
    return grid.stride[d] / vd;
}

} // namespace grid_oracle
