// gafw_check.cpp -- drives the C++ host layer (include/mtimager_amd_gafw.hpp)
// the way mt-imager drives its GAFW operators, and checks it:
//   gafw_check validate   shape/type validation, factory names, property
//                         parsing and the exception split (no GPU needed);
//   gafw_check gpu        one imaging job in the reference's WImager chunk
//                         layout through the WImager module (two accumulated
//                         chunks) and the ImageFinalizer module, and the same
//                         grid through the reference's finalizer operators one
//                         by one (PreIFFTUVCorrection, IFFT2D,
//                         ZeroToCentre2DShift, StokesConverter, casagridsf,
//                         SincCorrector), against the FP64 oracle.
// Test code: it links the oracle as the checker.  Exit status 0 = pass.
#include <hip/hip_runtime_api.h>

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/mtimager_amd_gafw.hpp"
#include "../../oracle/mti_oracle.h"

using namespace mtimager_amd::gafw;

static int failures = 0;
#define CHECK(cond, msg)                                         \
    do {                                                         \
        if (cond) {                                              \
            std::printf("ok   %s\n", msg);                       \
        } else {                                                 \
            std::printf("FAIL %s  (%s:%d)\n", msg, __FILE__, __LINE__); \
            failures++;                                          \
        }                                                        \
    } while (0)

template <class E, class F>
static bool throws(F f)
{
    try {
        f();
    } catch (const E &) {
        return true;
    } catch (...) {
        return false;
    }
    return false;
}

static Properties image_props()
{
    Properties p;
    p.setProperty("image.nx", "256");
    p.setProperty("image.ny", "256");
    p.setProperty("image.lintervalsize", "8");
    p.setProperty("image.mintervalsize", "8");
    p.setProperty("mode", "wproj");
    p.setProperty("wproj.wplanes", "16");
    p.setProperty("wproj.convFunction.sampling", "4");
    return p;
}

static std::vector<ArrayDescriptor> chunk_desc(long long N, int P)
{
    std::vector<ArrayDescriptor> in(5);
    in[0].dim = ArrayDimensions({3, N});
    in[0].type = real_float;
    in[1].dim = ArrayDimensions({N, 2 * P});
    in[1].type = real_float;
    in[2].dim = ArrayDimensions({N});
    in[2].type = real_float;
    in[3].dim = ArrayDimensions({N, P});
    in[3].type = real_float;
    in[4].dim = ArrayDimensions({N, P});
    in[4].type = real_int;
    return in;
}

static int run_validate()
{
    ImagerFactoryHelper f(0);
    CHECK(f.createOperator("x", "NoSuchOperator") == nullptr, "unknown operator name gives nullptr");
    for (const char *n : {"PreProcessUVW", "PreIFFTUVCorrection", "IFFT2D", "ZeroToCentre2DShift", "StokesConverter",
                          "casagridsf", "SincCorrector", "SortIndex"}) {
        auto op = f.createOperator(std::string("nick_") + n, n);
        CHECK(op && op->name() == n, (std::string("factory creates ") + n).c_str());
    }
    // reference names whose contracts consume the gridding plan or the
    // convolution-function chain: not offered (nullptr, as the reference's
    // factory gives for names it does not know)
    for (const char *n : {"ConvolutionGridder", "CreateIndexAndReorder", "NotSameSupportManipulate",
                          "CreateBlockDataIndex", "CreateCompressionPlan", "CompressVisibilities",
                          "WProjPreFFTConvFunc", "ConvFunctionReOrganize", "ConvFunctionSupportFind",
                          "SumConvolutionFunctions", "multiFFT2D", "AccumulatedSum", "Sum"})
        CHECK(f.createOperator("n", n) == nullptr, (std::string("no operator with another contract under ") + n).c_str());
    Properties p = image_props();
    mti_image_params ip = image_params_from(p);
    CHECK(ip.nx == 256 && ip.wplanes == 16 && ip.sampling == 4 && ip.mode == MTI_MODE_WPROJ,
          "image.* / wproj.* properties parsed");
    CHECK(std::fabs(ip.cell_l - 8.0 * M_PI / 180.0 / 3600.0) < 1e-18, "lintervalsize converted from arcsec");
    Properties bad;
    bad.setProperty("image.nx", "12x");
    CHECK(throws<ValidationException>([&] { bad.getIntProperty("image.nx"); }), "malformed integer property throws");
    CHECK(throws<ValidationException>([&] { bad.getIntProperty("image.ny"); }), "missing property throws");
    std::vector<ArrayDescriptor> out;
    ValidationData vd;

    // PreProcessUVW: six inputs, seven outputs
    auto pre = f.createOperator("pre", "PreProcessUVW");
    std::vector<ArrayDescriptor> pin = chunk_desc(1000, 4);
    pin.erase(pin.begin() + 1);  // no visibilities
    pin.insert(pin.begin(), ArrayDescriptor());
    pin.insert(pin.begin(), ArrayDescriptor());
    pin[0].dim = ArrayDimensions({16, 2});
    pin[0].type = real_int;
    pin[1].dim = ArrayDimensions({16, 25});
    pin[1].type = real_float;
    Properties pp;
    pp.setProperty("ConvolutionFunction.sampling", "4");
    pp.setProperty("ConvolutionFunction.wsquareincrement", 0.01);
    pp.setProperty("uvImage.u_increment", 50.0);
    pp.setProperty("uvImage.v_increment", 50.0);
    pp.setProperty("uvImage.rows", "256");
    pp.setProperty("uvImage.columns", "256");
    pre->validate(pin, out, pp, vd);
    CHECK(out.size() == 7 && out[0].dim == ArrayDimensions({1000, 4}) && out[0].type == real_int &&
              out[5].type == real_float && out[6].dim == ArrayDimensions({4, 1000}) && out[6].type == real_float,
          "PreProcessUVW outputs: GriddingData int [N][4] ... ConvolutionFunctionSum real_float [P][N]");
    std::vector<ArrayDescriptor> five(pin.begin(), pin.begin() + 5);
    CHECK(throws<ValidationException>([&] { pre->validate(five, out, pp, vd); }), "PreProcessUVW needs six inputs");
    Properties pq = pp;
    pq = Properties();
    CHECK(throws<ValidationException>([&] { pre->validate(pin, out, pq, vd); }), "PreProcessUVW needs its parameters");

    // IFFT2D shapes: any size
    auto fft = f.createOperator("f", "IFFT2D");
    std::vector<ArrayDescriptor> fi(1);
    fi[0].dim = ArrayDimensions({3, 128, 64});
    fi[0].type = complex_float;
    out.clear();
    fft->validate(fi, out, p, vd);
    CHECK(out.size() == 1 && out[0].dim == fi[0].dim && out[0].type == complex_float, "IFFT2D output shape");
    fi[0].dim = ArrayDimensions({100, 75});
    out.clear();
    fft->validate(fi, out, p, vd);
    CHECK(out.size() == 1 && out[0].dim == fi[0].dim, "IFFT2D takes sizes that are not powers of two");
    fi[0].type = real_float;
    CHECK(throws<ValidationException>([&] { fft->validate(fi, out, p, vd); }), "IFFT2D rejects real input");

    // the finalizer operators
    auto pre2 = f.createOperator("u", "PreIFFTUVCorrection");
    std::vector<ArrayDescriptor> g1(1);
    g1[0].dim = ArrayDimensions({256, 128, 4});
    g1[0].type = complex_float;
    pre2->validate(g1, out, p, vd);
    CHECK(out[0].dim == ArrayDimensions({4, 256, 128}) && out[0].type == complex_float,
          "PreIFFTUVCorrection: [v][u][P] -> [P][v][u]");
    auto stk = f.createOperator("s", "StokesConverter");
    std::vector<ArrayDescriptor> si(2);
    si[0].dim = ArrayDimensions({4, 256, 128});
    si[0].type = complex_float;
    si[1].dim = ArrayDimensions({4});
    si[1].type = real_float;
    Properties ps;
    ps.setProperty("PolarizationType", "1");
    stk->validate(si, out, ps, vd);
    CHECK(out[0].dim == si[0].dim && out[0].type == real_float, "StokesConverter output real_float [P][v][u]");
    si[1].type = real_double;
    CHECK(throws<ValidationException>([&] { stk->validate(si, out, ps, vd); }), "StokesConverter wants a real_float normaliser");
    si[1].type = real_float;
    ps.setProperty("PolarizationType", "2");
    CHECK(throws<ValidationException>([&] { stk->validate(si, out, ps, vd); }), "bad PolarizationType rejected");
    auto csf = f.createOperator("c", "casagridsf");
    Properties pc;
    pc.setProperty("support", "7");
    pc.setProperty("sampling", "4");
    pc.setProperty("lmPlanePostCorrection", "false");
    pc.setProperty("generate", "true");
    pc.setProperty("complexOutput", "true");
    std::vector<ArrayDescriptor> none;
    csf->validate(none, out, pc, vd);
    CHECK(out[0].dim == ArrayDimensions({29, 29}) && out[0].type == complex_float,
          "casagridsf generate: (S s + 1)^2 complex_float for even sampling");
    pc.setProperty("support", "6");
    CHECK(throws<ValidationException>([&] { csf->validate(none, out, pc, vd); }), "even support rejected");
    auto sinc = f.createOperator("n", "SincCorrector");
    std::vector<ArrayDescriptor> sc(1);
    sc[0].dim = ArrayDimensions({4, 64, 64});
    sc[0].type = real_float;
    Properties pn;
    pn.setProperty("sampling", "4");
    sinc->validate(sc, out, pn, vd);
    CHECK(out[0].dim == sc[0].dim && out[0].type == real_float, "SincCorrector keeps shape");
    sc[0].type = real_double;
    CHECK(throws<ValidationException>([&] { sinc->validate(sc, out, pn, vd); }), "SincCorrector works only on real_floats");

    // SortIndex shapes
    auto so = f.createOperator("s", "SortIndex");
    std::vector<ArrayDescriptor> sx(2);
    sx[0].dim = sx[1].dim = ArrayDimensions({500});
    sx[0].type = sx[1].type = real_int;
    Properties pa;
    pa.setProperty("nant", "27");
    out.clear();
    so->validate(sx, out, pa, vd);
    CHECK(out.size() == 2 && out[0].dim == sx[0].dim && out[1].type == real_longint, "SortIndex outputs");

    // modules without a GPU call: configuration checks
    CHECK(throws<ValidationException>([&] { WImager w(nullptr, p, 4); }), "WImager needs a generator");
    Properties pf = p;
    pf.setProperty("PolarizationType", "elliptical");
    CHECK(throws<ValidationException>([&] { ImageFinalizer fz(pf); }), "ImageFinalizer rejects a bad PolarizationType");
    std::printf("%s\n", failures ? "FAILED" : "PASSED");
    return failures ? 1 : 0;
}

// ---- GPU mode ----------------------------------------------------------------

static uint64_t lcg = 12345;
static double urand()
{
    lcg = lcg * 6364136223846793005ull + 1442695040888963407ull;
    return (double)(lcg >> 11) / 9007199254740992.0;
}

static void hipok(hipError_t e, const char *w)
{
    if (e != hipSuccess) {
        std::printf("HIP error %s: %s\n", w, hipGetErrorString(e));
        std::exit(2);
    }
}

template <class T>
static T *dev_copy(const std::vector<T> &h)
{
    T *d = nullptr;
    hipok(hipMalloc(&d, std::max<size_t>(1, h.size()) * sizeof(T)), "hipMalloc");
    hipok(hipMemcpy(d, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice), "H2D");
    return d;
}

static double rel_rms(const std::vector<double> &a, const std::vector<double> &b)
{
    double n = 0, d = 0;
    for (size_t i = 0; i < a.size(); i++) {
        n += (a[i] - b[i]) * (a[i] - b[i]);
        d += b[i] * b[i];
    }
    return std::sqrt(n / d);
}

static int run_gpu()
{
    const int P = 4;
    const long long N = 20000;  // two chunks of N records
    const double C = 299792458.0;
    const float invl = (float)(1.42e9 / C);
    Properties p = image_props();
    std::printf("gafw_check gpu: 2 x %lld records, P=%d, 256^2, 16 w-planes, s=4\n", N, P);
    std::fflush(stdout);

    // two chunks in the reference's WImager layout
    std::vector<float> uvw[2], vis[2], freq[2], wt[2];
    std::vector<int32_t> fl[2];
    for (int c = 0; c < 2; c++) {
        uvw[c].resize(3 * N);
        vis[c].resize(N * 2 * P);
        freq[c].assign(N, invl);
        wt[c].resize(N * P);
        fl[c].resize(N * P);
        for (long long i = 0; i < N; i++) {
            uvw[c][i] = (float)(3000.0 * (urand() - 0.5));
            uvw[c][N + i] = (float)(3000.0 * (urand() - 0.5));
            uvw[c][2 * N + i] = (float)(600.0 * (urand() - 0.5));
            for (int q = 0; q < P; q++) {
                vis[c][(i * P + q) * 2] = (float)(urand() - 0.5);
                vis[c][(i * P + q) * 2 + 1] = (float)(urand() - 0.5);
                wt[c][i * P + q] = (float)(0.5 + urand());
                fl[c][i * P + q] = urand() < 0.05 ? 1 : 0;
            }
        }
    }

    auto conv = std::make_shared<ConvolutionFunctionGenerator>(0, p);
    conv->calculateConvFunction();
    std::printf("convolution functions: max support %d, sampling %d, w increment %.6g\n", conv->getMaxSupport(),
                conv->getSampling(), conv->getWSquareIncrement());

    hipStream_t st;
    hipok(hipStreamCreate(&st), "stream");
    WImager wim(conv, p, P);
    wim.setStream(st);
    std::vector<void *> owned;
    for (int c = 0; c < 2; c++) {
        std::vector<ArrayDescriptor> in = chunk_desc(N, P);
        in[0].pointer = dev_copy(uvw[c]);
        in[1].pointer = dev_copy(vis[c]);
        in[2].pointer = dev_copy(freq[c]);
        in[3].pointer = dev_copy(wt[c]);
        in[4].pointer = dev_copy(fl[c]);
        for (int i = 0; i < 5; i++) {
            wim.setInput(i, in[i]);
            owned.push_back(in[i].pointer);
        }
        wim.calculate();  // the second chunk adds to the first
    }
    const ArrayDescriptor &gout = wim.getOutput(0), &nout = wim.getOutput(1);
    CHECK(gout.dim == ArrayDimensions({256, 256, P}) && gout.type == complex_float, "WImager Grid complex_float [v][u][P]");
    CHECK(nout.dim == ArrayDimensions({P}) && nout.type == real_float, "WImager NormalizerValue real_float [P]");
    ImageFinalizer fin(p);
    fin.setStream(st);
    fin.setInput(0, gout);
    fin.setInput(1, nout);
    fin.calculate();
    const ArrayDescriptor &fout = fin.getOutput(0);
    CHECK(fout.dim == ArrayDimensions({P, 256, 256}) && fout.type == real_float, "ImageFinalizer real_float [P][v][u]");

    // the same grid through the reference's finalizer operators one by one
    // (comp/ImageFinalizer.cpp:31-44)
    ImagerFactoryHelper f(0);
    std::vector<std::unique_ptr<ArrayOperator>> chain;
    const char *names[] = {"PreIFFTUVCorrection", "IFFT2D", "ZeroToCentre2DShift", "StokesConverter", "casagridsf",
                           "SincCorrector"};
    for (const char *n : names) chain.push_back(f.createOperator(std::string("fin.") + n, n));
    Properties pk = p;
    pk.setProperty("PolarizationType", "0");
    pk.setProperty("support", "7");
    pk.setProperty("sampling", "4");
    pk.setProperty("lmPlanePostCorrection", "true");
    pk.setProperty("generate", "false");
    std::vector<ArrayDescriptor> cur{gout};
    std::vector<void *> chain_bufs;
    for (size_t k = 0; k < chain.size(); k++) {
        std::vector<ArrayDescriptor> in = cur, out;
        if (k == 3) in.push_back(nout);  // StokesConverter takes the normaliser
        ValidationData vd;
        chain[k]->validate(in, out, pk, vd);
        const size_t esz = out[0].type == complex_float ? 8 : out[0].type == complex_double ? 16 : 4;
        out[0].pointer = nullptr;
        hipok(hipMalloc(&out[0].pointer, (size_t)out[0].dim.total() * esz), "chain output");
        chain_bufs.push_back(out[0].pointer);
        void *buf = nullptr;
        if (vd.bufferSize) {
            hipok(hipMalloc(&buf, vd.bufferSize), "chain buffer");
            chain_bufs.push_back(buf);
        }
        SubmissionData sd;
        sd.inputs = in;
        sd.outputs = out;
        sd.stream = st;
        sd.params = pk;
        sd.bufferGPUPointer = buf;
        sd.buffersize = vd.bufferSize;
        chain[k]->submitToGPU(sd);
        cur = out;
    }
    hipok(hipStreamSynchronize(st), "sync");

    std::vector<float> g_h((size_t)gout.dim.total() * 2), n_h(P), img_h((size_t)fout.dim.total()), chn_h(img_h.size());
    hipok(hipMemcpy(g_h.data(), gout.pointer, g_h.size() * 4, hipMemcpyDeviceToHost), "D2H grid");
    hipok(hipMemcpy(n_h.data(), nout.pointer, P * 4, hipMemcpyDeviceToHost), "D2H norm");
    hipok(hipMemcpy(img_h.data(), fout.pointer, img_h.size() * 4, hipMemcpyDeviceToHost), "D2H image");
    hipok(hipMemcpy(chn_h.data(), cur[0].pointer, chn_h.size() * 4, hipMemcpyDeviceToHost), "D2H chain image");

    // the oracle on the same records (FP64 throughout)
    orc_params op{256, 256, 8.0 * M_PI / 180.0 / 3600.0, 8.0 * M_PI / 180.0 / 3600.0, 1, 16, 4, 7};
    orc_kernels K;
    std::memset(&K, 0, sizeof(K));
    if (orc_kernels_build(&op, &K) != 0) {
        std::printf("oracle kernels failed\n");
        return 2;
    }
    std::vector<double> G((size_t)256 * 256 * P * 2, 0.0), NR(P, 0.0), IMG((size_t)P * 256 * 256);
    const double il = (double)invl;
    for (int c = 0; c < 2; c++) {
        std::vector<double> u(N * 3);
        std::vector<uint8_t> fu(N * P);
        for (long long i = 0; i < N; i++) {
            u[i * 3] = uvw[c][i];
            u[i * 3 + 1] = uvw[c][N + i];
            u[i * 3 + 2] = uvw[c][2 * N + i];
            for (int q = 0; q < P; q++) fu[i * P + q] = (uint8_t)fl[c][i * P + q];
        }
        int64_t ng = 0;
        orc_grid(&op, &K, N, 1, P, u.data(), &il, vis[c].data(), fu.data(), wt[c].data(), nullptr, 0, 0, G.data(),
                 NR.data(), &ng);
    }
    orc_finalize(&op, P, 0, G.data(), NR.data(), IMG.data());
    orc_kernels_free(&K);

    std::vector<double> gd(g_h.begin(), g_h.end()), id(img_h.begin(), img_h.end()), cd(chn_h.begin(), chn_h.end());
    double nerr = 0;
    for (int q = 0; q < P; q++) nerr = std::max(nerr, std::fabs(n_h[q] - NR[q]) / NR[q]);
    const double ge = rel_rms(gd, G), ie = rel_rms(id, IMG), ce = rel_rms(cd, IMG);
    std::printf("normaliser max rel err %.3e, grid rel RMS %.3e, image rel RMS %.3e, operator chain %.3e\n", nerr, ge,
                ie, ce);
    CHECK(nerr < 1e-6, "WImager NormalizerValue (real_float) matches the oracle");
    CHECK(ge < 1e-5, "WImager Grid (complex_float [v][u][P], two accumulated chunks) matches the oracle");
    CHECK(ie < 1e-4, "ImageFinalizer module Stokes images match the oracle");
    CHECK(ce < 1e-4, "the reference's finalizer operators, one by one, match the oracle");
    for (void *q : owned) hipok(hipFree(q), "free");
    for (void *q : chain_bufs) hipok(hipFree(q), "free");
    hipok(hipStreamDestroy(st), "stream destroy");
    std::printf("%s\n", failures ? "FAILED" : "PASSED");
    return failures ? 1 : 0;
}

int main(int argc, char **argv)
{
    const std::string mode = argc > 1 ? argv[1] : "validate";
    try {
        if (mode == "validate") return run_validate();
        if (mode == "gpu") return run_gpu();
    } catch (const std::exception &e) {
        std::printf("exception: %s\n", e.what());
        return 3;
    }
    std::printf("usage: gafw_check validate|gpu\n");
    return 2;
}
