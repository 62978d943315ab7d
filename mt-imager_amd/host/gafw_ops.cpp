// gafw_ops.cpp -- the GAFW-shaped C++ host layer (include/mtimager_amd_gafw.hpp)
// over the C-ABI.  Operators validate shapes the way the reference's
// validate() methods do and enqueue the C-ABI calls on the caller's stream.
#include "../../include/mtimager_amd_gafw.hpp"

#include <hip/hip_runtime.h>

#include <cmath>
#include <mutex>
#include <sstream>

namespace mtimager_amd {
namespace gafw {

// ---- small utilities ----------------------------------------------------------

long long ArrayDimensions::total() const
{
    long long t = 1;
    for (long long v : d_) t *= v;
    return d_.empty() ? 0 : t;
}

std::string ArrayDimensions::str() const
{
    std::ostringstream o;
    o << "[";
    for (size_t i = 0; i < d_.size(); i++) o << (i ? "][" : "") << d_[i];
    o << "]";
    return o.str();
}

void Properties::setProperty(const std::string &k, double v)
{
    std::ostringstream o;
    o.precision(17);
    o << v;
    m_[k] = o.str();
}

std::string Properties::getStringProperty(const std::string &k) const
{
    auto it = m_.find(k);
    if (it == m_.end()) throw ValidationException("property " + k + " is not set");
    return it->second;
}

int Properties::getIntProperty(const std::string &k) const
{
    const std::string v = getStringProperty(k);
    try {
        size_t pos = 0;
        const int r = std::stoi(v, &pos);
        if (pos != v.size()) throw std::invalid_argument(v);
        return r;
    } catch (const std::exception &) {
        throw ValidationException("property " + k + " is not an integer: " + v);
    }
}

double Properties::getDoubleProperty(const std::string &k) const
{
    const std::string v = getStringProperty(k);
    try {
        size_t pos = 0;
        const double r = std::stod(v, &pos);
        if (pos != v.size()) throw std::invalid_argument(v);
        return r;
    } catch (const std::exception &) {
        throw ValidationException("property " + k + " is not a number: " + v);
    }
}

bool Properties::getBoolProperty(const std::string &k) const
{
    const std::string v = getStringProperty(k);
    if (v == "true" || v == "1" || v == "yes") return true;
    if (v == "false" || v == "0" || v == "no") return false;
    throw ValidationException("property " + k + " is not a boolean: " + v);
}

namespace {

constexpr double ARCSEC = 3.14159265358979323846 / 180.0 / 3600.0;

std::mutex g_ctx_mutex;
std::map<int, mti_ctx *> g_ctx;

void throw_status(mti_status s, mti_ctx *ctx, const std::string &what)
{
    if (s == MTI_OK) return;
    const std::string msg = what + ": " + (ctx ? mti_last_error(ctx) : "no context");
    if (s == MTI_ERR_INVALID || s == MTI_ERR_UNSUPPORTED) throw ValidationException(msg);
    throw GpuException(msg);
}

void hip_check(hipError_t e, const char *what)
{
    if (e != hipSuccess) throw GpuException(std::string(what) + ": " + hipGetErrorString(e));
}

void expect(bool ok, const std::string &op, const std::string &msg)
{
    if (!ok) throw ValidationException(op + ": " + msg);
}

int current_device()
{
    int d = 0;
    hip_check(hipGetDevice(&d), "hipGetDevice");
    return d;
}

__global__ void k_f32_to_f64(const float *a, double *b, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) b[i] = a[i];
}

__global__ void k_f64_to_f32(const double *a, float *b, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) b[i] = (float)a[i];
}

// WImager inputs (app/mtimager.cpp:156-160): 0 UVW [3][N] f32, 1 VIS [N][2P]
// f32, 2 FREQUENCY [N] f32 (1/lambda), 3 WEIGHT [N][P] f32, 4 FLAGS [N][P] int.
struct ChunkShape {
    long long N;
    int P;
};

ChunkShape check_chunk_inputs(const std::vector<ArrayDescriptor> &in, const std::string &op, bool need_vis)
{
    expect(in.size() >= 5, op, "needs 5 inputs (UVW, VIS, FREQUENCY, WEIGHT, FLAGS)");
    const ArrayDimensions &uvw = in[0].dim;
    expect(uvw.getNoOfDimensions() == 2 && uvw.getY() == 3, op, "UVW must be [3][N], got " + uvw.str());
    expect(in[0].type == real_float, op, "UVW must be real_float");
    const long long N = uvw.getX();
    const ArrayDimensions &w = in[3].dim;
    expect(w.getNoOfDimensions() == 2 && w.getY() == N, op, "WEIGHT must be [N][P], got " + w.str());
    const int P = (int)w.getX();
    expect(P == 1 || P == 2 || P == 4, op, "1, 2 or 4 polarisations");
    expect(in[3].type == real_float, op, "WEIGHT must be real_float");
    if (need_vis) {
        expect(in[1].dim.getNoOfDimensions() == 2 && in[1].dim.getY() == N && in[1].dim.getX() == 2 * P, op,
               "VIS must be [N][2P], got " + in[1].dim.str());
        expect(in[1].type == real_float, op, "VIS must be real_float");
    }
    expect(in[2].dim.getNoOfDimensions() == 1 && in[2].dim.getX() == N, op, "FREQUENCY must be [N]");
    expect(in[2].type == real_float, op, "FREQUENCY must be real_float (1/lambda)");
    expect(in[4].dim.getNoOfDimensions() == 2 && in[4].dim.getY() == N && in[4].dim.getX() == P, op,
           "FLAGS must be [N][P], got " + in[4].dim.str());
    expect(in[4].type == real_int, op, "FLAGS must be real_int");
    return {N, P};
}

mti_vis chunk_vis(const std::vector<ArrayDescriptor> &in, const ChunkShape &c)
{
    mti_vis v{};
    v.nrows = c.N;
    v.nchan = 1;
    v.npol = c.P;
    v.uvw = in[0].pointer;
    v.uvw_f32 = 1;
    v.uvw_row_stride = 1;
    v.uvw_comp_stride = c.N;
    v.inv_lambda = in[2].pointer;
    v.inv_lambda_f32 = 1;
    v.inv_lambda_per_row = 1;
    v.vis = (const float *)in[1].pointer;
    v.vis_row_stride = c.P;
    v.vis_chan_stride = 0;
    v.flags = in[4].pointer;
    v.flags_i32 = 1;
    v.flags_row_stride = c.P;
    v.flags_chan_stride = 0;
    v.weight = (const float *)in[3].pointer;
    v.weight_row_stride = c.P;
    return v;
}

int weighting_of(const Properties &p)
{
    const std::string w = p.getStringProperty("weighting", "natural");
    if (w == "natural") return MTI_WEIGHT_NATURAL;
    if (w == "uniform") return MTI_WEIGHT_UNIFORM;
    if (w == "briggs" || w == "robust") return MTI_WEIGHT_BRIGGS;
    throw ValidationException("weighting must be natural, uniform or briggs: " + w);
}

// The engine's operator buffer (sized by validate()) serves one call: bytes
// [0, NORM_STAGE) hold the FP64 normaliser staging, the rest is the library's
// workspace for the call, so the library allocates nothing itself.
constexpr size_t NORM_STAGE = 256;

class WorkspaceScope {
public:
    WorkspaceScope(mti_ctx *ctx, const SubmissionData &d, size_t need, const std::string &op) : ctx_(ctx)
    {
        if (!d.bufferGPUPointer || d.buffersize < NORM_STAGE + need)
            throw GpuException(op + ": the operator buffer is missing or smaller than validate() asked for");
        throw_status(mti_set_workspace(ctx, (char *)d.bufferGPUPointer + NORM_STAGE, d.buffersize - NORM_STAGE), ctx,
                     op);
    }
    ~WorkspaceScope() { (void)mti_set_workspace(ctx_, nullptr, 0); }
    WorkspaceScope(const WorkspaceScope &) = delete;
    WorkspaceScope &operator=(const WorkspaceScope &) = delete;

private:
    mti_ctx *ctx_;
};

// Sizes-only description of a chunk for the workspace queries.
mti_vis chunk_sizes(const ChunkShape &c)
{
    mti_vis v{};
    v.nrows = c.N;
    v.nchan = 1;
    v.npol = c.P;
    v.inv_lambda_per_row = 1;
    return v;
}

mti_kernel_info info_of(const mti_kernels *k)
{
    mti_kernel_info i{};
    (void)mti_kernels_info(k, &i);
    return i;
}

size_t query(mti_status s, size_t bytes, const std::string &op)
{
    if (s != MTI_OK) throw ValidationException(op + ": arguments rejected by the workspace query");
    return bytes;
}

// ---- operators ----------------------------------------------------------------

// SortIndex (vis/SortIndex.cpp:46-87): inputs ANTENNA1, ANTENNA2 int [N]; output 0
// int [N] (the first `kept` entries valid), output 1 real_longint [1] = kept;
// property nant.  Blocks for the kept count like the reference's host sort.
class SortIndexOp : public ArrayOperator {
public:
    using ArrayOperator::ArrayOperator;
    void validate(const std::vector<ArrayDescriptor> &in, std::vector<ArrayDescriptor> &out, const Properties &p,
                  ValidationData &vd) override
    {
        expect(in.size() == 2, name_, "needs ANTENNA1 and ANTENNA2");
        expect(in[0].dim.getNoOfDimensions() == 1 && in[0].dim == in[1].dim, name_, "antenna columns must be [N]");
        expect(in[0].type == real_int && in[1].type == real_int, name_, "antenna columns must be real_int");
        expect(p.getIntProperty("nant") >= 1, name_, "nant >= 1");
        out.resize(2);
        out[0].dim = in[0].dim;
        out[0].type = real_int;
        out[1].dim = ArrayDimensions({1});
        out[1].type = real_longint;
        size_t ws = 0;
        vd.bufferSize = NORM_STAGE + query(mti_sort_rows_workspace_size(in[0].dim.getX(), &ws), ws, name_);
    }
    void submitToGPU(SubmissionData &d) override
    {
        mti_ctx *ctx = device_context(current_device());
        size_t ws = 0;
        query(mti_sort_rows_workspace_size(d.inputs[0].dim.getX(), &ws), ws, name_);
        WorkspaceScope wsc(ctx, d, ws, name_);
        throw_status(mti_sort_rows(ctx, (const int32_t *)d.inputs[0].pointer, (const int32_t *)d.inputs[1].pointer,
                                   d.inputs[0].dim.getX(), d.params.getIntProperty("nant"),
                                   (int32_t *)d.outputs[0].pointer, &kept_, d.stream),
                     ctx, name_);
        if (d.outputs.size() > 1 && d.outputs[1].pointer)
            hip_check(hipMemcpyAsync(d.outputs[1].pointer, &kept_, sizeof(kept_), hipMemcpyHostToDevice,
                                     (hipStream_t)d.stream),
                      "SortIndex kept count");
    }

private:
    int64_t kept_ = 0;
};

// PreProcessUVW (op/WProjPreProcessUVW.cu:111-324): WImager inputs -> int [N][8]
// {to_grid, plane, support, corner_x, corner_y, off_x, off_y, conj}.
class PreProcessUVWOp : public ArrayOperator {
public:
    using ArrayOperator::ArrayOperator;
    void validate(const std::vector<ArrayDescriptor> &in, std::vector<ArrayDescriptor> &out, const Properties &,
                  ValidationData &) override
    {
        const ChunkShape c = check_chunk_inputs(in, name_, false);
        expect(conv_ != nullptr, name_, "no convolution function set");
        out.resize(1);
        out[0].dim = ArrayDimensions({c.N, 8});
        out[0].type = real_int;
    }
    void submitToGPU(SubmissionData &d) override
    {
        const ChunkShape c{d.inputs[0].dim.getX(), (int)d.inputs[3].dim.getX()};
        mti_vis v = chunk_vis(d.inputs, c);
        mti_ctx *ctx = device_context(current_device());
        throw_status(mti_preprocess(ctx, conv_->kernels(), &conv_->imageParams(), &v, 0, (int32_t *)d.outputs[0].pointer,
                                    d.stream),
                     ctx, name_);
    }
};

// The WImager gridding DAG (comp/WImager.cpp:167-234).
// Output 0 Grid complex_float [v][u][P] (or complex_double), output 1
// NormalizerValue [P] real_float (or real_double).  Properties: psf,
// initialize_output (default true; false adds to the outputs like the
// reference's overwrite() chunks), weighting, robust.
class ConvolutionGridderOp : public ArrayOperator {
public:
    using ArrayOperator::ArrayOperator;
    void validate(const std::vector<ArrayDescriptor> &in, std::vector<ArrayDescriptor> &out, const Properties &p,
                  ValidationData &vd) override
    {
        const ChunkShape c = check_chunk_inputs(in, name_, true);
        expect(conv_ != nullptr, name_, "no convolution function set");
        const mti_image_params ip = conv_->imageParams();
        weighting_of(p);
        out.resize(2);
        const StoreType gt = out[0].type == complex_double ? complex_double : complex_float;
        out[0].dim = ArrayDimensions({ip.ny, ip.nx, c.P});
        out[0].type = gt;
        const StoreType nt = out[1].type == real_double ? real_double : real_float;
        out[1].dim = ArrayDimensions({c.P});
        out[1].type = nt;
        vd.bufferMode = BufferMode::Normal;
        // FP64 normaliser staging + the gridding workspace
        expect(conv_->kernels() != nullptr, name_, "convolution function not calculated");
        const mti_grid_opts o = opts(p, gt);
        const mti_vis v = chunk_sizes(c);
        size_t ws = 0;
        const mti_kernel_info ki = info_of(conv_->kernels());
        vd.bufferSize = NORM_STAGE + query(mti_grid_workspace_size(&ki, &ip, &v, &o, &ws), ws, name_);
    }
    void submitToGPU(SubmissionData &d) override
    {
        const ChunkShape c{d.inputs[0].dim.getX(), (int)d.inputs[3].dim.getX()};
        mti_vis v = chunk_vis(d.inputs, c);
        const mti_grid_opts o = opts(d.params, d.outputs[0].type);
        hipStream_t st = (hipStream_t)d.stream;
        mti_ctx *ctx = device_context(current_device());
        size_t ws = 0;
        const mti_kernel_info ki = info_of(conv_->kernels());
        query(mti_grid_workspace_size(&ki, &conv_->imageParams(), &v, &o, &ws), ws, name_);
        WorkspaceScope wsc(ctx, d, ws, name_);
        double *norm = (double *)d.outputs[1].pointer;
        const bool f32_norm = d.outputs[1].type != real_double;
        if (f32_norm) {
            norm = (double *)d.bufferGPUPointer;
            if (o.accumulate)
                hipLaunchKernelGGL(k_f32_to_f64, dim3(1), dim3(64), 0, st, (const float *)d.outputs[1].pointer, norm,
                                   c.P);
        }
        throw_status(mti_grid(ctx, conv_->kernels(), &conv_->imageParams(), &v, &o, d.outputs[0].pointer, norm,
                              nullptr, d.stream),
                     ctx, name_);
        if (f32_norm)
            hipLaunchKernelGGL(k_f64_to_f32, dim3(1), dim3(64), 0, st, norm, (float *)d.outputs[1].pointer, c.P);
        hip_check(hipGetLastError(), "normaliser conversion");
    }

private:
    static mti_grid_opts opts(const Properties &p, StoreType grid_type)
    {
        mti_grid_opts o{};
        o.chan0 = 0;
        o.nchan = 1;
        o.accumulate = p.getBoolProperty("initialize_output", true) ? 0 : 1;
        o.psf = p.getBoolProperty("psf", false) ? 1 : 0;
        o.weighting = weighting_of(p);
        o.robust = p.getDoubleProperty("robust", 0.0);
        o.grid_layout = MTI_GRID_VUP;
        o.grid_precision = grid_type == complex_double ? MTI_GRID_C128 : MTI_GRID_C64;
        return o;
    }
};

// IFFT2D (std/IFFT2D.cpp:86-111): unnormalised inverse C2C over the last two
// dimensions of a complex_float array; out-of-place calls copy first.
class IFFT2DOp : public ArrayOperator {
public:
    using ArrayOperator::ArrayOperator;
    void validate(const std::vector<ArrayDescriptor> &in, std::vector<ArrayDescriptor> &out, const Properties &,
                  ValidationData &vd) override
    {
        expect(in.size() == 1, name_, "one input");
        expect(in[0].type == complex_float, name_, "input must be complex_float");
        expect(in[0].dim.getNoOfDimensions() >= 2, name_, "input must be at least 2-D");
        const long long nx = in[0].dim.getX(), ny = in[0].dim.getY();
        auto pow2 = [](long long n) { return n >= 2 && n <= 8192 && (n & (n - 1)) == 0; };
        expect(pow2(nx) && pow2(ny), name_, "dimensions must be powers of two in [2, 8192]");
        out.resize(1);
        out[0].dim = in[0].dim;
        out[0].type = complex_float;
        size_t ws = 0;
        vd.bufferSize = NORM_STAGE + query(mti_fft2d_workspace_size((int)nx, (int)ny,
                                                                    (int)(in[0].dim.total() / (nx * ny)), &ws),
                                           ws, name_);
    }
    void submitToGPU(SubmissionData &d) override
    {
        const ArrayDimensions &dim = d.inputs[0].dim;
        const long long nx = dim.getX(), ny = dim.getY(), nb = dim.total() / (nx * ny);
        mti_ctx *ctx = device_context(current_device());
        size_t ws = 0;
        query(mti_fft2d_workspace_size((int)nx, (int)ny, (int)nb, &ws), ws, name_);
        WorkspaceScope wsc(ctx, d, ws, name_);
        if (d.outputs[0].pointer != d.inputs[0].pointer)
            hip_check(hipMemcpyAsync(d.outputs[0].pointer, d.inputs[0].pointer, (size_t)dim.total() * 8,
                                     hipMemcpyDeviceToDevice, (hipStream_t)d.stream),
                      "IFFT2D copy");
        throw_status(mti_fft2d_inverse(ctx, (float *)d.outputs[0].pointer, (int)nx, (int)ny, (int)nb, d.stream), ctx,
                     name_);
    }
};

// ImageFinalizer (comp/ImageFinalizer.cpp:31-44): UVImage [v][u][P] (complex_float or
// complex_double) + NormalizerValue [P] (real_float or real_double) ->
// real_float [P][v][u]; property PolarizationType (linear | circular).
class ImageFinalizerOp : public ArrayOperator {
public:
    using ArrayOperator::ArrayOperator;
    void validate(const std::vector<ArrayDescriptor> &in, std::vector<ArrayDescriptor> &out, const Properties &p,
                  ValidationData &vd) override
    {
        expect(in.size() == 2, name_, "needs UVImage and NormalizerValue");
        expect(in[0].type == complex_float || in[0].type == complex_double, name_, "UVImage must be complex");
        expect(in[0].dim.getNoOfDimensions() == 3, name_, "UVImage must be [v][u][P], got " + in[0].dim.str());
        const long long P = in[0].dim.getX();
        expect(P == 1 || P == 2 || P == 4, name_, "1, 2 or 4 polarisations");
        expect(in[1].dim.getNoOfDimensions() == 1 && in[1].dim.getX() == P, name_, "NormalizerValue must be [P]");
        expect(in[1].type == real_float || in[1].type == real_double, name_, "NormalizerValue must be real");
        const std::string pt = p.getStringProperty("PolarizationType", "linear");
        expect(pt == "linear" || pt == "circular", name_, "PolarizationType must be linear or circular");
        out.resize(1);
        out[0].dim = ArrayDimensions({P, in[0].dim.getZ(), in[0].dim.getY()});
        out[0].type = real_float;
        mti_image_params ip = conv_ ? conv_->imageParams() : image_params_from(p);
        ip.ny = (int)in[0].dim.getZ();
        ip.nx = (int)in[0].dim.getY();
        size_t ws = 0;
        vd.bufferSize = NORM_STAGE + query(mti_finalize_workspace_size(&ip, (int)P, 1, &ws), ws, name_);
    }
    void submitToGPU(SubmissionData &d) override
    {
        mti_image_params ip = conv_ ? conv_->imageParams() : image_params_from(d.params);
        const int P = (int)d.inputs[0].dim.getX();
        ip.ny = (int)d.inputs[0].dim.getZ();
        ip.nx = (int)d.inputs[0].dim.getY();
        mti_ctx *ctx = device_context(current_device());
        size_t ws = 0;
        query(mti_finalize_workspace_size(&ip, P, 1, &ws), ws, name_);
        WorkspaceScope wsc(ctx, d, ws, name_);
        const double *norm = (const double *)d.inputs[1].pointer;
        if (d.inputs[1].type == real_float) {
            hipLaunchKernelGGL(k_f32_to_f64, dim3(1), dim3(64), 0, (hipStream_t)d.stream,
                               (const float *)d.inputs[1].pointer, (double *)d.bufferGPUPointer, P);
            hip_check(hipGetLastError(), "normaliser conversion");
            norm = (const double *)d.bufferGPUPointer;
        }
        const int pt = d.params.getStringProperty("PolarizationType", "linear") == "circular" ? MTI_POL_CIRCULAR
                                                                                              : MTI_POL_LINEAR;
        throw_status(mti_finalize(ctx, &ip, P, pt, 1, MTI_GRID_VUP,
                                  d.inputs[0].type == complex_double ? MTI_GRID_C128 : MTI_GRID_C64,
                                  d.inputs[0].pointer, norm, (float *)d.outputs[0].pointer, d.stream),
                     ctx, name_);
    }
};

}  // namespace

mti_image_params image_params_from(const Properties &p)
{
    mti_image_params ip{};
    ip.nx = p.getIntProperty("image.nx");
    ip.ny = p.getIntProperty("image.ny");
    ip.cell_l = p.getDoubleProperty("image.lintervalsize") * ARCSEC;
    ip.cell_m = p.getDoubleProperty("image.mintervalsize") * ARCSEC;
    const std::string mode = p.getStringProperty("mode", "wproj");
    if (mode == "wproj") {
        ip.mode = MTI_MODE_WPROJ;
        ip.wplanes = p.getIntProperty("wproj.wplanes");
        ip.sampling = p.getIntProperty("wproj.convFunction.sampling");
        ip.taper_support = p.getIntProperty("taper.support", 7);
    } else if (mode == "normal") {
        ip.mode = MTI_MODE_NORMAL;
        ip.wplanes = 1;
        ip.sampling = p.getIntProperty("taper.sampling");
        ip.taper_support = p.getIntProperty("taper.support");
    } else {
        throw ValidationException("mode must be normal or wproj: " + mode);
    }
    return ip;
}

mti_ctx *device_context(int device)
{
    std::lock_guard<std::mutex> lock(g_ctx_mutex);
    auto it = g_ctx.find(device);
    if (it != g_ctx.end()) return it->second;
    mti_ctx *c = nullptr;
    const mti_status s = mti_create(device, &c);
    if (s != MTI_OK) throw GpuException("mti_create failed on device " + std::to_string(device));
    g_ctx[device] = c;
    return c;
}

ConvolutionFunctionGenerator::ConvolutionFunctionGenerator(int device, const Properties &params, void *stream)
    : device_(device), p_(image_params_from(params)), stream_(stream)
{
    const std::string prec = params.getStringProperty("convFunction.precision", "f32");
    if (prec == "f32") precision_ = MTI_TABLE_F32;
    else if (prec == "f16") precision_ = MTI_TABLE_F16;
    else if (prec == "bf16") precision_ = MTI_TABLE_BF16;
    else throw ValidationException("convFunction.precision must be f32, f16 or bf16: " + prec);
}

ConvolutionFunctionGenerator::~ConvolutionFunctionGenerator()
{
    if (k_) mti_kernels_free(k_);
}

void ConvolutionFunctionGenerator::calculateConvFunction()
{
    if (k_) {
        mti_kernels_free(k_);
        k_ = nullptr;
    }
    hip_check(hipSetDevice(device_), "hipSetDevice");
    mti_ctx *ctx = device_context(device_);
    throw_status(mti_kernels_build(ctx, &p_, precision_, &k_, stream_), ctx, "calculateConvFunction");
}

static const mti_kernels *need(const mti_kernels *k)
{
    if (!k) throw ValidationException("calculateConvFunction() has not been called");
    return k;
}

const void *ConvolutionFunctionGenerator::getConvFunction() const { return mti_kernels_table(need(k_)); }
const float *ConvolutionFunctionGenerator::getConvFunctionSumData() const { return mti_kernels_sums(need(k_)); }

std::vector<std::pair<int, long long>> ConvolutionFunctionGenerator::getConvFunctionPositionData() const
{
    mti_kernel_info info{};
    mti_kernels_info(need(k_), &info);
    std::vector<int32_t> s(info.nplanes);
    std::vector<int64_t> b(info.nplanes);
    mti_kernels_supports(k_, s.data(), b.data());
    std::vector<std::pair<int, long long>> r(info.nplanes);
    for (int i = 0; i < info.nplanes; i++) r[i] = {s[i], b[i]};
    return r;
}

int ConvolutionFunctionGenerator::getMaxSupport() const
{
    mti_kernel_info info{};
    mti_kernels_info(need(k_), &info);
    return info.max_support;
}

int ConvolutionFunctionGenerator::getSampling() const { return p_.sampling; }

float ConvolutionFunctionGenerator::getWSquareIncrement() const
{
    mti_kernel_info info{};
    mti_kernels_info(need(k_), &info);
    return (float)info.w_increment;
}

std::unique_ptr<ArrayOperator> ImagerFactoryHelper::createOperator(const std::string &nickname,
                                                                  const std::string &name) const
{
    (void)device_;
    if (name == "SortIndex") return std::unique_ptr<ArrayOperator>(new SortIndexOp(nickname, name));
    if (name == "PreProcessUVW") return std::unique_ptr<ArrayOperator>(new PreProcessUVWOp(nickname, name));
    if (name == "ConvolutionGridder") return std::unique_ptr<ArrayOperator>(new ConvolutionGridderOp(nickname, name));
    if (name == "IFFT2D") return std::unique_ptr<ArrayOperator>(new IFFT2DOp(nickname, name));
    if (name == "ImageFinalizer") return std::unique_ptr<ArrayOperator>(new ImageFinalizerOp(nickname, name));
    return nullptr;
}

}  // namespace gafw
}  // namespace mtimager_amd
