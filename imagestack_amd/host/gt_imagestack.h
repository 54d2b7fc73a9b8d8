// gt_imagestack.h - ImageStack-side adapter for libgtlattice.so.
//
// Include it from ImageStack code (after main.h, which ImageStack's build
// force-includes) to run the permutohedral lattice on the GPU with the
// reference's own types and error convention:
//
//   ImageStack::GPU::gaussTransform(slice, splat, values, sigmas)
//       drop-in for GaussTransform::apply(..., PERMUTOHEDRAL)
//       (src/GaussTransform.cpp:115-226): same size checks, same 1.0f/sigma,
//       replay slice when slice == splat (Image identity, :196), returns the
//       unnormalised W x H x F x vd image.
//   ImageStack::GPU::jointBilateral(im, ref, fw, fh, ff, colorSigma)
//       drop-in for JointBilateral::apply(..., PERMUTOHEDRAL)
//       (src/GaussTransform.cpp:504-678): filters im in place.
//   ImageStack::GPU::gaussTransformGrid / GPU::jointBilateralGrid
//       the same for the GRID method (the bilateral grid,
//       src/DenseGrid.h; at most 6 position channels).
//   ImageStack::GPU::nlmeans(im, filters, spatialSigma, patchSigma)
//       NLMeans::apply / NLMeans3D::apply after their PatchPCA call
//       (src/GaussTransform.cpp:867-874, 934-941): filters in place.
//
// Library failures become ImageStack::panic() with the library's message,
// exactly how the reference reports errors (src/Exception.cpp:18-33).
// Images may be non-dense views: they are packed into dense planar host
// buffers before the copy to the device.
#pragma once

#include <vector>

#include "gtlattice.h"

namespace ImageStack {
namespace GPU {

// One library context per host thread, created on first use (device 0, or
// the device named by IMAGESTACK_GPU_DEVICE).
inline gt_ctx *context() {
    struct Holder {
        gt_ctx *c = nullptr;
        ~Holder() { if (c) gt_ctx_destroy(c); }
    };
    static thread_local Holder h;
    if (!h.c) {
        const char *dev = getenv("IMAGESTACK_GPU_DEVICE");
        if (gt_ctx_create(&h.c, dev ? atoi(dev) : 0) != GT_OK)
            panic("GPU permutohedral lattice unavailable: %s\n", gt_last_error());
    }
    return h.c;
}

inline std::vector<float> packPlanar(const Image &im) {
    const size_t n = (size_t)im.width * im.height * im.frames;
    std::vector<float> buf(n * im.channels);
    for (int c = 0; c < im.channels; c++)
        for (int t = 0; t < im.frames; t++)
            for (int y = 0; y < im.height; y++)
                for (int x = 0; x < im.width; x++)
                    buf[c * n + ((size_t)t * im.height + y) * im.width + x] = im(x, y, t, c);
    return buf;
}

inline void unpackPlanar(const std::vector<float> &buf, Image im) {
    const size_t n = (size_t)im.width * im.height * im.frames;
    for (int c = 0; c < im.channels; c++)
        for (int t = 0; t < im.frames; t++)
            for (int y = 0; y < im.height; y++)
                for (int x = 0; x < im.width; x++)
                    im(x, y, t, c) = buf[c * n + ((size_t)t * im.height + y) * im.width + x];
}

inline Image gaussTransform(Image slice, Image splat, Image values, std::vector<float> sigmas) {
    assert(splat.width == values.width && splat.height == values.height && splat.frames == values.frames,
           "Weights and locations of the Gaussians must be the same size\n");
    assert(slice.channels == splat.channels,
           "The evaluation locations and the locations of the Gaussians must have"
           " the same number of channels.\n");
    assert((int)sigmas.size() == splat.channels, "One standard deviation per position channel is required\n");
    std::vector<float> invSigma(sigmas.size());
    for (size_t i = 0; i < sigmas.size(); i++) invSigma[i] = 1.0f / sigmas[i];
    const int n = splat.width * splat.height * splat.frames;
    const bool replay = (slice == splat);
    std::vector<float> sp = packPlanar(splat), va = packPlanar(values), sl;
    if (!replay) sl = packPlanar(slice);
    const int ns = slice.width * slice.height * slice.frames;
    Image out(slice.width, slice.height, slice.frames, values.channels);
    std::vector<float> ob((size_t)ns * values.channels);
    if (gt_gauss_transform_host(context(), n, splat.width, splat.channels, values.channels, sp.data(), va.data(),
                                ns, replay ? nullptr : sl.data(), invSigma.data(), ob.data(), GT_STORE_F32,
                                nullptr, nullptr) != GT_OK)
        panic("%s\n", gt_last_error());
    unpackPlanar(ob, out);
    return out;
}

// GaussTransform::apply(..., GRID) (src/GaussTransform.cpp:227-297) on the GPU.
inline Image gaussTransformGrid(Image slice, Image splat, Image values, std::vector<float> sigmas) {
    assert(splat.width == values.width && splat.height == values.height && splat.frames == values.frames,
           "Weights and locations of the Gaussians must be the same size\n");
    assert(slice.channels == splat.channels,
           "The evaluation locations and the locations of the Gaussians must have"
           " the same number of channels.\n");
    assert((int)sigmas.size() == splat.channels, "One standard deviation per position channel is required\n");
    std::vector<float> invSigma(sigmas.size());
    for (size_t i = 0; i < sigmas.size(); i++) invSigma[i] = 1.0f / sigmas[i];
    const int n = splat.width * splat.height * splat.frames;
    const bool replay = (slice == splat);
    std::vector<float> sp = packPlanar(splat), va = packPlanar(values), sl;
    if (!replay) sl = packPlanar(slice);
    const int ns = slice.width * slice.height * slice.frames;
    Image out(slice.width, slice.height, slice.frames, values.channels);
    std::vector<float> ob((size_t)ns * values.channels);
    if (gt_gauss_transform_grid_host(context(), n, splat.channels, values.channels, sp.data(), va.data(), ns,
                                     replay ? nullptr : sl.data(), invSigma.data(), ob.data(), nullptr) != GT_OK)
        panic("%s\n", gt_last_error());
    unpackPlanar(ob, out);
    return out;
}

inline void jointBilateral(Image im, Image ref, float filterWidth, float filterHeight, float filterFrames,
                           float colorSigma) {
    assert(im.width == ref.width && im.height == ref.height && im.frames == ref.frames,
           "Image and reference must be the same size\n");
    std::vector<float> a = packPlanar(im), b = packPlanar(ref), o(a.size());
    if (gt_joint_bilateral_host(context(), im.width, im.height, im.frames, im.channels, a.data(), ref.channels,
                                b.data(), filterWidth, filterHeight, filterFrames, colorSigma, o.data(),
                                GT_STORE_F32, nullptr, nullptr) != GT_OK)
        panic("%s\n", gt_last_error());
    unpackPlanar(o, im);
}

// JointBilateral::apply(..., GRID) on the GPU: filters im in place.
inline void jointBilateralGrid(Image im, Image ref, float filterWidth, float filterHeight, float filterFrames,
                           float colorSigma) {
    assert(im.width == ref.width && im.height == ref.height && im.frames == ref.frames,
           "Image and reference must be the same size\n");
    std::vector<float> a = packPlanar(im), b = packPlanar(ref), o(a.size());
    if (gt_joint_bilateral_grid_host(context(), im.width, im.height, im.frames, im.channels, a.data(),
                                     ref.channels, b.data(), filterWidth, filterHeight, filterFrames, colorSigma,
                                     o.data(), nullptr) != GT_OK)
        panic("%s\n", gt_last_error());
    unpackPlanar(o, im);
}

// NLMeans::apply after PatchPCA on the GPU: the patch projection
// Convolve::apply(im, filters, Zero, Inner) and JointBilateral::apply(im,
// projection, s, s, INF, patchSigma) with AUTO's choice.  Returns false, with
// im untouched, when that choice (EXACT, GKDTREE) is not on the GPU.
inline bool nlmeans(Image im, Image filters, float spatialSigma, float patchSigma) {
    assert(filters.channels % im.channels == 0,
           "To perform an inner or matrix product, the channel count of either the image or the filter must be a"
           " multiple of the channel count of the other.");
    std::vector<float> a = packPlanar(im), f = packPlanar(filters), o(a.size());
    const int rc = gt_nlmeans_host(context(), im.width, im.height, im.frames, im.channels, a.data(), filters.width,
                                   filters.height, filters.frames, filters.channels / im.channels, f.data(),
                                   spatialSigma, patchSigma, o.data(), GT_STORE_F32, nullptr, nullptr);
    if (rc == GT_EUNSUPPORTED) return false;
    if (rc != GT_OK) panic("%s\n", gt_last_error());
    unpackPlanar(o, im);
    return true;
}

} // namespace GPU
} // namespace ImageStack
