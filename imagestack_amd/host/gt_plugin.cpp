// ImageStack plugin: GPU permutohedral lattice behind the reference's own
// operations.  Load it into an unmodified ImageStack binary with
//
//     ImageStack -plugin libgt_plugin.so -load im.tmp -bilateral 0.1 8 8 0 permutohedral -save out.tmp
//
// init_imagestack_plugin (src/Plugin.cpp:28-35, examples/plugin.cpp:73-77)
// replaces -gausstransform, -jointbilateral, -bilateral, -nlmeans and
// -nlmeans3d with subclasses that keep the reference's argument grammar
// (src/GaussTransform.cpp:79-113, :473-502, :699-728, :840-865, :907-932) and
// its AUTO method choice (:540-564).  Whenever the
// method resolves to PERMUTOHEDRAL the lattice, and to GRID (at most 6
// position channels) the bilateral grid, runs in libgtlattice.so; the other
// methods (exact, gkdtree) run the reference's own code.
//
// Built against the reference's headers (main.h is force-included, as the
// reference's own build does), so Image and Operation match the host binary.
#include "Operation.h"
#include "GaussTransform.h"
#include "Statistics.h"
#include "Convolve.h"
#include "Arithmetic.h"
#include "gt_imagestack.h"

namespace ImageStack {
namespace GPU {

GaussTransform::Method methodByName(const string &s) {
    if (s == "exact") return GaussTransform::EXACT;
    if (s == "grid") return GaussTransform::GRID;
    if (s == "permutohedral") return GaussTransform::PERMUTOHEDRAL;
    if (s == "gkdtree") return GaussTransform::GKDTREE;
    panic("Unknown method %s\n", s.c_str());
    return GaussTransform::AUTO;
}

// The shape of the pieces JointBilateral::apply filters: a zero spatial sigma
// on an axis longer than 1 splits the image into columns, rows or frames
// (:512-538), every piece of the same shape; pw/ph/pf receive it and the
// sigmas the pieces see.
void pieceShape(const Image &im, float &fw, float &fh, float &ff, int &pw, int &ph, int &pf) {
    pw = im.width;
    ph = im.height;
    pf = im.frames;
    if (pw == 1) fw = 0;
    if (ph == 1) fh = 0;
    if (pf == 1) ff = 0;
    if (fw == 0) pw = 1;
    if (fh == 0) ph = 1;
    if (ff == 0) pf = 1;
}

// Position channels of the lattice each piece builds (:540-546).
int pieceDims(const Image &im, const Image &ref, float fw, float fh, float ff) {
    int pw, ph, pf;
    pieceShape(im, fw, fh, ff, pw, ph, pf);
    return ref.channels + (pw > 1 && fw < 10 * pw ? 1 : 0) + (ph > 1 && fh < 10 * ph ? 1 : 0) +
           (pf > 1 && ff < 10 * pf ? 1 : 0);
}

// The method JointBilateral::apply uses for this call: m, or for AUTO the
// choice each piece makes (:553-564) -- all pieces alike, so one choice.
GaussTransform::Method resolveMethod(Image im, Image ref, float fw, float fh, float ff, GaussTransform::Method m) {
    if (m != GaussTransform::AUTO) return m;
    int pw, ph, pf;
    pieceShape(im, fw, fh, ff, pw, ph, pf);
    bool fx = pw > 1 && fw < 10 * pw;
    bool fy = ph > 1 && fh < 10 * ph;
    bool ft = pf > 1 && ff < 10 * pf;
    int posChannels = ref.channels + (fx ? 1 : 0) + (fy ? 1 : 0) + (ft ? 1 : 0);
    int sx = fx ? ((int)(fw * 6 + 1) | 1) : 1;
    int sy = fy ? ((int)(fh * 6 + 1) | 1) : 1;
    int st = ft ? ((int)(ff * 6 + 1) | 1) : 1;
    if (st * sx * sy < 16) return GaussTransform::EXACT;
    if (posChannels <= 4) return GaussTransform::GRID;
    if (posChannels <= 10) return GaussTransform::PERMUTOHEDRAL;
    return GaussTransform::GKDTREE;
}

// The library filters the zero-sigma pieces itself.  Methods it does not run
// (EXACT, GKDTREE) and lattices beyond GT_MAX_D position channels (grids
// beyond 6) go to the reference's own code.
void jointBilateralDispatch(Image im, Image ref, float fw, float fh, float ff, float cs, GaussTransform::Method m) {
    const GaussTransform::Method r = resolveMethod(im, ref, fw, fh, ff, m);
    const int d = pieceDims(im, ref, fw, fh, ff);
    if (r == GaussTransform::PERMUTOHEDRAL && d <= GT_MAX_D)
        jointBilateral(im, ref, fw, fh, ff, cs);
    else if (r == GaussTransform::GRID && d <= 6)
        jointBilateralGrid(im, ref, fw, fh, ff, cs);
    else
        JointBilateral::apply(im, ref, fw, fh, ff, cs, r);
}

// Relative difference below tol everywhere (floor of 1e-3 of the largest
// magnitude, so near-zero homogeneous weights do not dominate).
bool closeTo(Image a, Image b, float tol) {
    float mx = 0;
    for (int c = 0; c < b.channels; c++)
        for (int t = 0; t < b.frames; t++)
            for (int y = 0; y < b.height; y++)
                for (int x = 0; x < b.width; x++) mx = std::max(mx, fabsf(b(x, y, t, c)));
    for (int c = 0; c < b.channels; c++)
        for (int t = 0; t < b.frames; t++)
            for (int y = 0; y < b.height; y++)
                for (int x = 0; x < b.width; x++) {
                    float den = std::max(fabsf(b(x, y, t, c)), 1e-3f * mx);
                    if (fabsf(a(x, y, t, c) - b(x, y, t, c)) > tol * den) return false;
                }
    return true;
}

class GaussTransformGPU : public GaussTransform {
public:
    void help() {
        GaussTransform::help();
        pprintf("(GPU plugin: the permutohedral method runs on the GPU.)\n");
    }

    // GaussTransform::test (src/GaussTransform.cpp:39-77) with the GPU lattice:
    // agreement with the exact transform in the homogeneous sense
    // (nearlyEqual), and with the reference's own CPU lattice to 1e-4.
    bool test() {
        Image splat(56, 34, 2, 3);
        Image slice(56, 34, 2, 3);
        Image values(56, 34, 2, 2);
        Noise::apply(splat, 0, 1);
        Noise::apply(slice, 0, 1);
        Noise::apply(values, 0, 1);
        values.channel(1).set(1.0f);
        vector<float> sigma(3);
        sigma[0] = 0.2f;
        sigma[1] = 0.3f;
        sigma[2] = 0.4f;
        Image correct = GaussTransform::apply(slice, splat, values, sigma, EXACT);
        correct = correct.channel(0) / correct.channel(1);
        Image cpu = GaussTransform::apply(slice, splat, values, sigma, PERMUTOHEDRAL);
        Image gpu = gaussTransform(slice, splat, values, sigma);
        if (!closeTo(gpu, cpu, 1e-4f)) return false;
        Image out = gpu.channel(0) / gpu.channel(1);
        return nearlyEqual(out, correct);
    }

    void parse(vector<string> args) {
        Method m = EXACT;
        assert(args.size() > 0, "-gausstransform takes at least one argument");
        m = methodByName(args[0]);
        vector<float> sigmas;
        if (args.size() == 1) {
            sigmas = vector<float>(stack(0).channels, 1.0);
        } else if (args.size() == 2) {
            sigmas = vector<float>(stack(0).channels, readFloat(args[1]));
        } else if ((int)args.size() == 1 + stack(1).channels) {
            for (int i = 0; i < stack(0).channels; i++) sigmas.push_back(readFloat(args[i + 1]));
        } else {
            panic("-gausstransform takes one argument, two arguments, or one plus the"
                  " number of channels in the second image on the stack arguments\n");
        }
        // lattices of more than GT_MAX_D position channels (grids of more
        // than 6) and the other methods run the reference's own code
        const int d = stack(1).channels;
        Image im = (m == PERMUTOHEDRAL && d <= GT_MAX_D) ? gaussTransform(stack(0), stack(1), stack(2), sigmas)
                   : (m == GRID && d <= 6)               ? gaussTransformGrid(stack(0), stack(1), stack(2), sigmas)
                                                         : GaussTransform::apply(stack(0), stack(1), stack(2), sigmas, m);
        pop();
        push(im);
    }
};

static void parseBilateralArgs(const vector<string> &args, const char *name, float &cs, float &fw, float &fh,
                               float &ff, GaussTransform::Method &m) {
    if (args.size() < 2 || args.size() > 5) panic("%s takes from two to five arguments\n", name);
    ff = 0;
    m = GaussTransform::AUTO;
    cs = readFloat(args[0]);
    fw = fh = readFloat(args[1]);
    if (args.size() > 2) fh = readFloat(args[2]);
    if (args.size() > 3) ff = readFloat(args[3]);
    if (args.size() > 4) m = methodByName(args[4]);
}

class JointBilateralGPU : public JointBilateral {
public:
    void help() {
        JointBilateral::help();
        pprintf("(GPU plugin: the permutohedral method runs on the GPU.)\n");
    }

    // JointBilateral::test (src/GaussTransform.cpp:442-471) with the GPU
    // lattice: nearlyEqual to EXACT, and within 1e-4 of the CPU lattice.
    bool test() {
        Image im(60, 50, 3, 3);
        Image ref(im.width, im.height, im.frames, im.channels);
        Noise::apply(im, 0, 1);
        Noise::apply(ref, 0, 1);
        Image correct = im.copy();
        JointBilateral::apply(correct, ref, 3, 4, 1, 0.25, GaussTransform::EXACT);
        Image cpu = im.copy();
        JointBilateral::apply(cpu, ref, 3, 4, 1, 0.25, GaussTransform::PERMUTOHEDRAL);
        Image out = im.copy();
        jointBilateral(out, ref, 3, 4, 1, 0.25);
        return closeTo(out, cpu, 1e-4f) && nearlyEqual(out, correct);
    }

    void parse(vector<string> args) {
        float cs, fw, fh, ff;
        GaussTransform::Method m;
        parseBilateralArgs(args, "-jointbilateral", cs, fw, fh, ff, m);
        jointBilateralDispatch(stack(0), stack(1), fw, fh, ff, cs, m);
    }
};

class BilateralGPU : public Bilateral {
public:
    void help() {
        Bilateral::help();
        pprintf("(GPU plugin: the permutohedral method runs on the GPU.)\n");
    }
    bool test() { return true; }  // as the reference: it calls joint bilateral
    void parse(vector<string> args) {
        float cs, fw, fh, ff;
        GaussTransform::Method m;
        parseBilateralArgs(args, "-bilateral", cs, fw, fh, ff, m);
        jointBilateralDispatch(stack(0), stack(0), fw, fh, ff, cs, m);
    }
};

// NLMeans::parse (src/GaussTransform.cpp:840-865): the same arguments; the
// method is parsed and, as in the reference, not passed on (:873).
static void parseNLMeansArgs(const vector<string> &args, float &patchSize, int &dims, float &spatialSigma,
                             float &patchSigma) {
    assert(args.size() == 4 || args.size() == 5, "-nlmeans takes four or five arguments\n");
    patchSize = readFloat(args[0]);
    dims = readInt(args[1]);
    spatialSigma = readFloat(args[2]);
    patchSigma = readFloat(args[3]);
    if (args.size() > 4) methodByName(args[4]);
}

// The steps of NLMeans::apply / NLMeans3D::apply (:867-874, :934-941) with
// the host's own PatchPCA (so the random patches are the ones the reference
// draws), then the projection and joint bilateral on the GPU; where AUTO
// picks a method the GPU does not run, the reference's own Convolve and
// JointBilateral finish the job with the same filters.
static void nlmeansSteps(Image image, Image filters, float spatialSigma, float patchSigma) {
    if (nlmeans(image, filters, spatialSigma, patchSigma)) return;
    Image pca = Convolve::apply(image, filters, Convolve::Zero, Multiply::Inner);
    JointBilateral::apply(image, pca, spatialSigma, spatialSigma, INF, patchSigma);
}

// NLMeans::test's check (src/GaussTransform.cpp:824-838) on a synthetic image
// (the reference's uses pics/dog1.jpg): noise is removed, and the GPU result
// is within 1e-4 of the reference's NLMeans::apply on the same random stream.
static bool nlmeansTest(bool threeD) {
    Image clean(threeD ? 24 : 80, threeD ? 20 : 60, threeD ? 8 : 1, 3);
    for (int c = 0; c < 3; c++)
        for (int t = 0; t < clean.frames; t++)
            for (int y = 0; y < clean.height; y++)
                for (int x = 0; x < clean.width; x++)
                    clean(x, y, t, c) = ((x / 10 + y / 12 + c) % 3) * 0.4f + 0.1f;
    Image noisy = clean.copy();
    Noise::apply(noisy, -0.2, 0.2);
    const float sigma = threeD ? 0.5f : 1.0f;
    const unsigned seed = (unsigned)rand();
    Image cpu = noisy.copy();
    srand(seed);
    if (threeD) NLMeans3D::apply(cpu, sigma, 4, 4, 0.4);
    else NLMeans::apply(cpu, sigma, 6, 16, 0.4);
    Image gpu = noisy.copy();
    srand(seed);
    Image filters = threeD ? PatchPCA3D::apply(gpu, sigma, 4) : PatchPCA::apply(gpu, sigma, 6);
    if (!nlmeans(gpu, filters, threeD ? 4 : 16, 0.4)) return false;
    double errNoisy = 0, errOut = 0;
    for (int c = 0; c < 3; c++)
        for (int t = 0; t < clean.frames; t++)
            for (int y = 0; y < clean.height; y++)
                for (int x = 0; x < clean.width; x++) {
                    errNoisy += fabs(noisy(x, y, t, c) - clean(x, y, t, c));
                    errOut += fabs(gpu(x, y, t, c) - clean(x, y, t, c));
                }
    return closeTo(gpu, cpu, 1e-4f) && errOut < 0.5 * errNoisy;
}

class NLMeansGPU : public NLMeans {
public:
    void help() {
        NLMeans::help();
        pprintf("(GPU plugin: the patch projection and the joint bilateral run on the GPU.)\n");
    }
    bool test() { return nlmeansTest(false); }
    void parse(vector<string> args) {
        float patchSize, spatialSigma, patchSigma;
        int dims;
        parseNLMeansArgs(args, patchSize, dims, spatialSigma, patchSigma);
        Image filters = PatchPCA::apply(stack(0), patchSize, dims);
        nlmeansSteps(stack(0), filters, spatialSigma, patchSigma);
    }
};

class NLMeans3DGPU : public NLMeans3D {
public:
    void help() {
        NLMeans3D::help();
        pprintf("(GPU plugin: the patch projection and the joint bilateral run on the GPU.)\n");
    }
    bool test() { return nlmeansTest(true); }
    void parse(vector<string> args) {
        float patchSize, spatialSigma, patchSigma;
        int dims;
        parseNLMeansArgs(args, patchSize, dims, spatialSigma, patchSigma);
        Image filters = PatchPCA3D::apply(stack(0), patchSize, dims);
        nlmeansSteps(stack(0), filters, spatialSigma, patchSigma);
    }
};

} // namespace GPU
} // namespace ImageStack

extern "C" void init_imagestack_plugin(map<string, ImageStack::Operation *> &operationMap) {
    operationMap["-gausstransform"] = new ImageStack::GPU::GaussTransformGPU();
    operationMap["-jointbilateral"] = new ImageStack::GPU::JointBilateralGPU();
    operationMap["-bilateral"] = new ImageStack::GPU::BilateralGPU();
    operationMap["-nlmeans"] = new ImageStack::GPU::NLMeansGPU();
    operationMap["-nlmeans3d"] = new ImageStack::GPU::NLMeans3DGPU();
}
