// Route 1 check: ImageStack with the in-process GPU method case
// (GaussTransform::PERMUTOHEDRAL_GPU, src/GaussTransform.h:8 and
// src/GaussTransform.cpp:134, :566, patched by the Makefile's sed recipe) calls
// the library from inside GaussTransform::apply and JointBilateral::apply.
// This driver runs both methods on the same images -- the reference's own
// PERMUTOHEDRAL on the CPU and the patched PERMUTOHEDRAL_GPU -- and prints
// the largest relative difference (floor 1e-3 of the largest magnitude).
// Exit status 0 when every difference is below 1e-5.
//
// Built by `make -C imagestack_amd route1` against the reference sources and
// objects (linked with libgtlattice.so); needs a GPU to run.
#include "GaussTransform.h"

#include <cmath>
#include <cstdio>
#include <vector>

using namespace ImageStack;

namespace {

// Seeded RGB image: axis-aligned constant rectangles plus uniform noise.
Image rect_image(int w, int h, unsigned seed) {
    Image im(w, h, 1, 3);
    unsigned s = seed;
    auto next = [&]() {
        s = s * 1664525u + 1013904223u;
        return (float)(s >> 8) / 16777216.0f;
    };
    for (int k = 0; k < 12; k++) {
        const int x0 = (int)(next() * w), y0 = (int)(next() * h);
        const int x1 = x0 + 1 + (int)(next() * w / 2), y1 = y0 + 1 + (int)(next() * h / 2);
        const float col[3] = {next(), next(), next()};
        for (int y = y0; y < std::min(h, y1); y++)
            for (int x = x0; x < std::min(w, x1); x++)
                for (int c = 0; c < 3; c++) im(x, y, 0, c) = col[c];
    }
    for (int y = 0; y < h; y++)
        for (int x = 0; x < w; x++)
            for (int c = 0; c < 3; c++) im(x, y, 0, c) += 0.05f * (next() - 0.5f);
    return im;
}

double max_rel(const Image &a, const Image &b) {
    double mx = 0, err = 0;
    for (int c = 0; c < a.channels; c++)
        for (int y = 0; y < a.height; y++)
            for (int x = 0; x < a.width; x++) mx = std::max(mx, (double)std::fabs(a(x, y, 0, c)));
    for (int c = 0; c < a.channels; c++)
        for (int y = 0; y < a.height; y++)
            for (int x = 0; x < a.width; x++) {
                const double d = std::fabs((double)a(x, y, 0, c) - (double)b(x, y, 0, c));
                err = std::max(err, d / std::max((double)std::fabs(a(x, y, 0, c)), 1e-3 * mx));
            }
    return err;
}

}  // namespace

int main() {
    const int W = 96, H = 64;
    Image im = rect_image(W, H, 7);

    // Gauss transform: positions (x, y, r, g, b), values (r, g, b, 1).
    Image pos(W, H, 1, 5), val(W, H, 1, 4);
    for (int y = 0; y < H; y++)
        for (int x = 0; x < W; x++) {
            pos(x, y, 0, 0) = (float)x;
            pos(x, y, 0, 1) = (float)y;
            for (int c = 0; c < 3; c++) {
                pos(x, y, 0, 2 + c) = im(x, y, 0, c);
                val(x, y, 0, c) = im(x, y, 0, c);
            }
            val(x, y, 0, 3) = 1.0f;
        }
    const std::vector<float> sigmas = {6.0f, 6.0f, 0.15f, 0.15f, 0.15f};
    Image cpu = GaussTransform::apply(pos, pos, val, sigmas, GaussTransform::PERMUTOHEDRAL);
    Image gpu = GaussTransform::apply(pos, pos, val, sigmas, GaussTransform::PERMUTOHEDRAL_GPU);
    const double e1 = max_rel(cpu, gpu);

    // Joint bilateral, in place, including a zero spatial sigma (rows only).
    Image a = im.copy(), b = im.copy();
    JointBilateral::apply(a, im, 5.0f, 5.0f, 0.0f, 0.1f, GaussTransform::PERMUTOHEDRAL);
    JointBilateral::apply(b, im, 5.0f, 5.0f, 0.0f, 0.1f, GaussTransform::PERMUTOHEDRAL_GPU);
    const double e2 = max_rel(a, b);
    Image c = im.copy(), d = im.copy();
    JointBilateral::apply(c, im, 4.0f, 0.0f, 0.0f, 0.1f, GaussTransform::PERMUTOHEDRAL);
    JointBilateral::apply(d, im, 4.0f, 0.0f, 0.0f, 0.1f, GaussTransform::PERMUTOHEDRAL_GPU);
    const double e3 = max_rel(c, d);

    printf("route1 gausstransform %.3e jointbilateral %.3e jointbilateral_rows %.3e\n", e1, e2, e3);
    return (e1 < 1e-5 && e2 < 1e-5 && e3 < 1e-5) ? 0 : 1;
}
