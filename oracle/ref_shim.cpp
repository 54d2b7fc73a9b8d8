// C entry points into the reference ImageStack library, compiled from the
// reference's own sources by oracle/Makefile into oracle/_ref/.  TEST
// INFRASTRUCTURE ONLY (checker and CPU-baseline timing); never part of the
// product.  Each function builds ImageStack::Image objects from dense planar
// float buffers and calls the reference code unchanged.

#include "main.h"
#include "Image.h"
#include "GaussTransform.h"
#include "Permutohedral.h"
#include "Convolve.h"
#include "Arithmetic.h"
#include "Statistics.h"

using namespace ImageStack;

static Image planar_to_image(int W, int H, int F, int C, const float *src) {
    Image im(W, H, F, C);
    const size_t n = (size_t)W * H * F;
    for (int c = 0; c < C; c++)
        for (int t = 0; t < F; t++)
            for (int y = 0; y < H; y++)
                for (int x = 0; x < W; x++)
                    im(x, y, t, c) = src[c * n + ((size_t)t * H + y) * W + x];
    return im;
}

static void image_to_planar(const Image &im, float *dst) {
    const size_t n = (size_t)im.width * im.height * im.frames;
    for (int c = 0; c < im.channels; c++)
        for (int t = 0; t < im.frames; t++)
            for (int y = 0; y < im.height; y++)
                for (int x = 0; x < im.width; x++)
                    dst[c * n + ((size_t)t * im.height + y) * im.width + x] = im(x, y, t, c);
}

extern "C" {

// GaussTransform::apply(slice, splat, values, sigmas, PERMUTOHEDRAL)
// (src/GaussTransform.cpp:115-226).  slice == NULL passes the splat image as
// the slice image, which selects the reference's replay slice.
int ref_gauss_transform(int W, int H, int F, int d, const float *splat, int vd,
                        const float *values, const float *slice, const float *sigmas,
                        float *out) {
    try {
        Image sp = planar_to_image(W, H, F, d, splat);
        Image va = planar_to_image(W, H, F, vd, values);
        Image sl = slice ? planar_to_image(W, H, F, d, slice) : sp;
        vector<float> s(sigmas, sigmas + d);
        Image o = GaussTransform::apply(sl, sp, va, s, GaussTransform::PERMUTOHEDRAL);
        image_to_planar(o, out);
        return 0;
    } catch (...) {
        return -1;
    }
}

// GaussTransform::apply(slice, splat, values, sigmas, GRID)
// (src/GaussTransform.cpp:227-297).  slice == NULL passes the splat image.
int ref_grid_transform(int W, int H, int F, int d, const float *splat, int vd,
                       const float *values, const float *slice, const float *sigmas,
                       float *out) {
    try {
        Image sp = planar_to_image(W, H, F, d, splat);
        Image va = planar_to_image(W, H, F, vd, values);
        Image sl = slice ? planar_to_image(W, H, F, d, slice) : sp;
        vector<float> s(sigmas, sigmas + d);
        Image o = GaussTransform::apply(sl, sp, va, s, GaussTransform::GRID);
        image_to_planar(o, out);
        return 0;
    } catch (...) {
        return -1;
    }
}

// JointBilateral::apply(im, ref, fw, fh, ff, colorSigma, GRID).
int ref_joint_bilateral_grid(int W, int H, int F, int C, const float *im, int Cr,
                             const float *ref, float fw, float fh, float ff, float colorSigma,
                             float *out) {
    try {
        Image a = planar_to_image(W, H, F, C, im);
        Image b = planar_to_image(W, H, F, Cr, ref);
        JointBilateral::apply(a, b, fw, fh, ff, colorSigma, GaussTransform::GRID);
        image_to_planar(a, out);
        return 0;
    } catch (...) {
        return -1;
    }
}

// JointBilateral::apply(im, ref, fw, fh, ff, colorSigma, PERMUTOHEDRAL)
// (src/GaussTransform.cpp:504-678); the filtered image is written to out.
int ref_joint_bilateral(int W, int H, int F, int C, const float *im, int Cr,
                        const float *ref, float fw, float fh, float ff, float colorSigma,
                        float *out) {
    try {
        Image a = planar_to_image(W, H, F, C, im);
        Image b = planar_to_image(W, H, F, Cr, ref);
        JointBilateral::apply(a, b, fw, fh, ff, colorSigma, GaussTransform::PERMUTOHEDRAL);
        image_to_planar(a, out);
        return 0;
    } catch (...) {
        return -1;
    }
}

// Drives PermutohedralLattice exactly as src/GaussTransform.cpp:164-189 does
// (splat in raster order with pos = splat * invSigma, then optionally blur)
// and exports the insertion-ordered key array (M x d int16) and the value
// array (M x vd float).  Call with keys == NULL to get M only.
int ref_lattice(int n, int d, int vd, const float *splat, const float *values,
                const float *inv_sigma, int blur, int max_m, short *keys, float *vals) {
    PermutohedralLattice lat(d, vd, n);
    vector<float> pos(d), val(vd);
    for (int i = 0; i < n; i++) {
        for (int c = 0; c < d; c++) pos[c] = splat[(size_t)c * n + i] * inv_sigma[c];
        for (int c = 0; c < vd; c++) val[c] = values[(size_t)c * n + i];
        lat.splat(&pos[0], &val[0]);
    }
    if (blur) lat.blur();
    const int M = lat.hashTable.size();
    if (keys && M <= max_m) {
        memcpy(keys, lat.hashTable.getKeys(), sizeof(short) * (size_t)M * d);
        if (vals) memcpy(vals, lat.hashTable.getValues(), sizeof(float) * (size_t)M * vd);
    }
    return M;
}

// Convolve::apply(im, filters, Zero, Multiply::Inner) (src/Convolve.cpp:248-270)
// with a filter bank of Kc planes of fw x fh x ft; out holds Kc / C planes.
int ref_convolve_inner(int W, int H, int F, int C, const float *im, int fw, int fh, int ft, int Kc,
                       const float *filt, float *out) {
    try {
        Image a = planar_to_image(W, H, F, C, im);
        Image f = planar_to_image(fw, fh, ft, Kc, filt);
        Image o = Convolve::apply(a, f, Convolve::Zero, Multiply::Inner);
        image_to_planar(o, out);
        return o.channels;
    } catch (...) {
        return -1;
    }
}

// PatchPCA::apply(im, sigma, dims) (src/Statistics.cpp:1710-1770), or
// PatchPCA3D::apply when three_d, after srand(seed): the random patch choice
// is the reference's own rand() stream.  Returns the patch side; filters
// (side^2 or side^3 floats per plane, C*dims planes) are written when
// capacity allows.
int ref_patch_pca(unsigned seed, int three_d, int W, int H, int F, int C, const float *im, float sigma,
                  int dims, long long capacity, float *filters) {
    try {
        Image a = planar_to_image(W, H, F, C, im);
        srand(seed);
        Image f = three_d ? PatchPCA3D::apply(a, sigma, dims) : PatchPCA::apply(a, sigma, dims);
        if ((long long)f.width * f.height * f.frames * f.channels <= capacity) image_to_planar(f, filters);
        return f.width;
    } catch (...) {
        return -1;
    }
}

// NLMeans::apply(im, sigma, dims, spatialSigma, patchSigma) (src/GaussTransform.cpp:867-874),
// or NLMeans3D::apply when three_d, after srand(seed); the filtered image is
// written to out.  With the same seed it runs on the filters ref_patch_pca returns.
int ref_nlmeans(unsigned seed, int three_d, int W, int H, int F, int C, const float *im, float sigma, int dims,
                float spatialSigma, float patchSigma, float *out) {
    try {
        Image a = planar_to_image(W, H, F, C, im);
        srand(seed);
        if (three_d) NLMeans3D::apply(a, sigma, dims, spatialSigma, patchSigma);
        else NLMeans::apply(a, sigma, dims, spatialSigma, patchSigma);
        image_to_planar(a, out);
        return 0;
    } catch (...) {
        return -1;
    }
}

}
