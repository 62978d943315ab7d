// gafw_check.cpp -- drives the C++ host layer (include/mtimager_amd_gafw.hpp)
// the way mt-imager drives its GAFW operators, and checks it:
//   gafw_check validate   shape/type validation, factory names, property
//                         parsing and the exception split (no GPU needed);
//   gafw_check gpu        one imaging job through SortIndex-free chunk input
//                         (the reference's WImager chunk layout), the
//                         ConvolutionGridder and ImageFinalizer operators, in
//                         two accumulated chunks, against the FP64 oracle.
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
    for (const char *n : {"SortIndex", "PreProcessUVW", "ConvolutionGridder", "IFFT2D", "ImageFinalizer"}) {
        auto op = f.createOperator(std::string("nick_") + n, n);
        CHECK(op && op->name() == n, (std::string("factory creates ") + n).c_str());
    }
    Properties p = image_props();
    mti_image_params ip = image_params_from(p);
    CHECK(ip.nx == 256 && ip.wplanes == 16 && ip.sampling == 4 && ip.mode == MTI_MODE_WPROJ,
          "image.* / wproj.* properties parsed");
    CHECK(std::fabs(ip.cell_l - 8.0 * M_PI / 180.0 / 3600.0) < 1e-18, "lintervalsize converted from arcsec");
    Properties bad;
    bad.setProperty("image.nx", "12x");
    CHECK(throws<ValidationException>([&] { bad.getIntProperty("image.nx"); }), "malformed integer property throws");
    CHECK(throws<ValidationException>([&] { bad.getIntProperty("image.ny"); }), "missing property throws");

    // ConvolutionGridder without convolution functions
    auto g = f.createOperator("g", "ConvolutionGridder");
    std::vector<ArrayDescriptor> in = chunk_desc(1000, 4), out;
    ValidationData vd;
    CHECK(throws<ValidationException>([&] { g->validate(in, out, p, vd); }),
          "gridder without a convolution function is rejected");

    // IFFT2D shapes
    auto fft = f.createOperator("f", "IFFT2D");
    std::vector<ArrayDescriptor> fi(1);
    fi[0].dim = ArrayDimensions({3, 128, 64});
    fi[0].type = complex_float;
    out.clear();
    fft->validate(fi, out, p, vd);
    CHECK(out.size() == 1 && out[0].dim == fi[0].dim && out[0].type == complex_float, "IFFT2D output shape");
    fi[0].dim = ArrayDimensions({100, 64});
    CHECK(throws<ValidationException>([&] { fft->validate(fi, out, p, vd); }), "IFFT2D rejects non powers of two");
    fi[0].dim = ArrayDimensions({64, 64});
    fi[0].type = real_float;
    CHECK(throws<ValidationException>([&] { fft->validate(fi, out, p, vd); }), "IFFT2D rejects real input");

    // ImageFinalizer shapes and buffer request
    auto fin = f.createOperator("i", "ImageFinalizer");
    std::vector<ArrayDescriptor> ii(2);
    ii[0].dim = ArrayDimensions({256, 128, 4});
    ii[0].type = complex_float;
    ii[1].dim = ArrayDimensions({4});
    ii[1].type = real_float;
    out.clear();
    vd = ValidationData();
    fin->validate(ii, out, p, vd);
    CHECK(out.size() == 1 && out[0].dim == ArrayDimensions({4, 256, 128}) && out[0].type == real_float,
          "ImageFinalizer output is real_float [P][v][u]");
    {
        mti_image_params fp = image_params_from(p);
        fp.ny = 256;
        fp.nx = 128;
        size_t ws = 0;
        CHECK(mti_finalize_workspace_size(&fp, 4, 1, &ws) == MTI_OK && ws > 0 && vd.bufferSize == 256 + ws,
              "ImageFinalizer asks for the normaliser staging and the finalize workspace");
    }
    ii[1].dim = ArrayDimensions({2});
    CHECK(throws<ValidationException>([&] { fin->validate(ii, out, p, vd); }), "normaliser length must be P");
    Properties pc = p;
    pc.setProperty("PolarizationType", "elliptical");
    ii[1].dim = ArrayDimensions({4});
    CHECK(throws<ValidationException>([&] { fin->validate(ii, out, pc, vd); }), "bad PolarizationType rejected");

    // SortIndex shapes
    auto so = f.createOperator("s", "SortIndex");
    std::vector<ArrayDescriptor> si(2);
    si[0].dim = si[1].dim = ArrayDimensions({500});
    si[0].type = si[1].type = real_int;
    Properties ps;
    ps.setProperty("nant", "27");
    out.clear();
    so->validate(si, out, ps, vd);
    CHECK(out.size() == 2 && out[0].dim == si[0].dim && out[1].type == real_longint, "SortIndex outputs");
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

    ImagerFactoryHelper f(0);
    auto grid_op = f.createOperator("WImager.gridder", "ConvolutionGridder");
    auto fin_op = f.createOperator("ImageFinalizer", "ImageFinalizer");
    grid_op->setConvolutionFunction(conv);
    fin_op->setConvolutionFunction(conv);

    std::vector<ArrayDescriptor> gout;
    ValidationData gvd;
    grid_op->validate(chunk_desc(N, P), gout, p, gvd);
    CHECK(gout[0].dim == ArrayDimensions({256, 256, P}) && gout[0].type == complex_float, "gridder output shape");
    float *d_grid = nullptr, *d_norm = nullptr;
    void *d_buf = nullptr;
    hipok(hipMalloc(&d_grid, (size_t)gout[0].dim.total() * 8), "grid");
    hipok(hipMalloc(&d_norm, P * sizeof(float)), "norm");
    hipok(hipMalloc(&d_buf, gvd.bufferSize), "buffer");
    hipStream_t st;
    hipok(hipStreamCreate(&st), "stream");

    std::vector<void *> owned;
    for (int c = 0; c < 2; c++) {
        SubmissionData sd;
        sd.inputs = chunk_desc(N, P);
        sd.inputs[0].pointer = dev_copy(uvw[c]);
        sd.inputs[1].pointer = dev_copy(vis[c]);
        sd.inputs[2].pointer = dev_copy(freq[c]);
        sd.inputs[3].pointer = dev_copy(wt[c]);
        sd.inputs[4].pointer = dev_copy(fl[c]);
        for (auto &d : sd.inputs) owned.push_back(d.pointer);
        sd.outputs = gout;
        sd.outputs[0].pointer = d_grid;
        sd.outputs[1].pointer = d_norm;
        sd.stream = st;
        sd.params = p;
        sd.params.setProperty("initialize_output", c == 0 ? "true" : "false");
        sd.bufferGPUPointer = d_buf;
        sd.buffersize = gvd.bufferSize;
        grid_op->submitToGPU(sd);
    }
    std::vector<ArrayDescriptor> fin_in(2), fin_out;
    fin_in[0] = gout[0];
    fin_in[1] = gout[1];
    ValidationData fvd;
    fin_op->validate(fin_in, fin_out, p, fvd);
    float *d_img = nullptr;
    void *d_fbuf = nullptr;
    hipok(hipMalloc(&d_img, (size_t)fin_out[0].dim.total() * 4), "image");
    hipok(hipMalloc(&d_fbuf, fvd.bufferSize), "fbuf");
    SubmissionData fd;
    fd.inputs = fin_in;
    fd.inputs[0].pointer = d_grid;
    fd.inputs[1].pointer = d_norm;
    fd.outputs = fin_out;
    fd.outputs[0].pointer = d_img;
    fd.stream = st;
    fd.params = p;
    fd.bufferGPUPointer = d_fbuf;
    fd.buffersize = fvd.bufferSize;
    fin_op->submitToGPU(fd);
    hipok(hipStreamSynchronize(st), "sync");

    std::vector<float> g_h((size_t)gout[0].dim.total() * 2), n_h(P), img_h((size_t)fin_out[0].dim.total());
    hipok(hipMemcpy(g_h.data(), d_grid, g_h.size() * 4, hipMemcpyDeviceToHost), "D2H grid");
    hipok(hipMemcpy(n_h.data(), d_norm, P * 4, hipMemcpyDeviceToHost), "D2H norm");
    hipok(hipMemcpy(img_h.data(), d_img, img_h.size() * 4, hipMemcpyDeviceToHost), "D2H image");

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

    std::vector<double> gd(g_h.begin(), g_h.end()), id(img_h.begin(), img_h.end());
    double nerr = 0;
    for (int q = 0; q < P; q++) nerr = std::max(nerr, std::fabs(n_h[q] - NR[q]) / NR[q]);
    const double ge = rel_rms(gd, G), ie = rel_rms(id, IMG);
    std::printf("normaliser max rel err %.3e, grid rel RMS %.3e, image rel RMS %.3e\n", nerr, ge, ie);
    CHECK(nerr < 1e-6, "NormalizerValue (f32 output) matches the oracle");
    CHECK(ge < 1e-5, "Grid (complex_float [v][u][P], two accumulated chunks) matches the oracle");
    CHECK(ie < 1e-4, "ImageFinalizer Stokes images match the oracle");
    for (void *q : owned) hipok(hipFree(q), "free");
    for (void *q : {(void *)d_grid, (void *)d_norm, d_buf, (void *)d_img, d_fbuf}) hipok(hipFree(q), "free");
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
