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

// PreProcessUVW (op/WProjPreProcessUVW.cpp:39-86): inputs 0 convolution
// position data int [W][2] {support, base}, 1 sums real_float [W][T*T],
// 2 UVW [3][N], 3 FREQUENCY [N], 4 WEIGHT [N][P], 5 FLAGS [N][P]; outputs
// 0 GriddingData int [N][4], 1 ConvolutionIndex, 2 ToGrid, 3 ToCompress,
// 4 Support int [N], 5 TakeConjugate real_float [N], 6 ConvolutionFunctionSum
// real_float [P][N]; parameters ConvolutionFunction.sampling,
// ConvolutionFunction.wsquareincrement, uvImage.{u_increment, v_increment,
// rows, columns}.
class PreProcessUVWOp : public ArrayOperator {
public:
    using ArrayOperator::ArrayOperator;
    void validate(const std::vector<ArrayDescriptor> &in, std::vector<ArrayDescriptor> &out, const Properties &p,
                  ValidationData &) override
    {
        expect(in.size() == 6, name_,
               "This operator expects six inputs which are Convolution Function support data and UVW entries");
        for (const char *k : {"ConvolutionFunction.sampling", "uvImage.rows", "uvImage.columns"}) p.getIntProperty(k);
        for (const char *k : {"ConvolutionFunction.wsquareincrement", "uvImage.u_increment", "uvImage.v_increment"})
            p.getDoubleProperty(k);
        expect(in[0].dim.getNoOfDimensions() == 2 && in[0].dim.getX() == 2 && in[0].type == real_int, name_,
               "input 0 (convolution position data) must be int [W][2]");
        expect(in[1].type == real_float && in[1].dim.getNoOfDimensions() == 2 && in[1].dim.getY() == in[0].dim.getY(),
               name_, "input 1 (convolution sums) must be real_float [W][T*T]");
        const ChunkShape c = check_chunk_inputs(reorder(in), name_, false);
        out.resize(7);
        out[0].dim = ArrayDimensions({c.N, 4});
        out[0].type = real_int;
        for (int i = 1; i <= 4; i++) {
            out[i].dim = ArrayDimensions({c.N});
            out[i].type = real_int;
        }
        out[5].dim = ArrayDimensions({c.N});
        out[5].type = real_float;
        out[6].dim = ArrayDimensions({(long long)c.P, c.N});
        out[6].type = real_float;
    }
    void submitToGPU(SubmissionData &d) override
    {
        const std::vector<ArrayDescriptor> ch = reorder(d.inputs);
        const ChunkShape c{ch[0].dim.getX(), (int)ch[3].dim.getX()};
        mti_vis v = chunk_vis(ch, c);
        mti_preprocess_ref_params rp{};
        rp.rows = d.params.getIntProperty("uvImage.rows");
        rp.columns = d.params.getIntProperty("uvImage.columns");
        rp.u_increment = d.params.getDoubleProperty("uvImage.u_increment");
        rp.v_increment = d.params.getDoubleProperty("uvImage.v_increment");
        rp.sampling = d.params.getIntProperty("ConvolutionFunction.sampling");
        rp.wsquareincrement = d.params.getDoubleProperty("ConvolutionFunction.wsquareincrement");
        rp.wplanes = (int)d.inputs[0].dim.getY();
        mti_preprocess_ref_outputs o{(int32_t *)d.outputs[0].pointer, (int32_t *)d.outputs[1].pointer,
                                     (int32_t *)d.outputs[2].pointer, (int32_t *)d.outputs[3].pointer,
                                     (int32_t *)d.outputs[4].pointer, (float *)d.outputs[5].pointer,
                                     (float *)d.outputs[6].pointer};
        mti_ctx *ctx = device_context(current_device());
        throw_status(mti_preprocess_reference(ctx, &rp, (const int32_t *)d.inputs[0].pointer,
                                              (const float *)d.inputs[1].pointer, &v, 0, &o, d.stream),
                     ctx, name_);
    }

private:
    // the WImager chunk order (UVW, VIS, FREQUENCY, WEIGHT, FLAGS) of inputs 2-5
    static std::vector<ArrayDescriptor> reorder(const std::vector<ArrayDescriptor> &in)
    {
        std::vector<ArrayDescriptor> r(5);
        r[0] = in[2];
        r[2] = in[3];
        r[3] = in[4];
        r[4] = in[5];
        return r;
    }
};

// IFFT2D (std/IFFT2D.cpp:86-111): unnormalised inverse C2C over the last two
// dimensions of a complex_float array, any size; out-of-place calls copy first.
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
        expect(nx >= 1 && ny >= 1 && nx <= 32768 && ny <= 32768, name_, "dimensions must be in [1, 32768]");
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

void op_status(mti_status s, const std::string &what)
{
    if (s == MTI_ERR_INVALID) throw ValidationException(what + ": invalid arguments");
    if (s != MTI_OK) throw GpuException(what + ": launch failed");
}

// PreIFFTUVCorrection (op/PreIFFTUVCorrection.cpp:36-62): one 3-D complex input
// [v][u][P] -> [P][v][u] of the same type.
class PreIFFTUVCorrectionOp : public ArrayOperator {
public:
    using ArrayOperator::ArrayOperator;
    void validate(const std::vector<ArrayDescriptor> &in, std::vector<ArrayDescriptor> &out, const Properties &,
                  ValidationData &) override
    {
        expect(in.size() == 1, name_, "Only one input is supported");
        expect(in[0].dim.getNoOfDimensions() == 3, name_, "Expected a 3D Array (lengthxwidthxpolarization");
        expect(in[0].type == complex_float || in[0].type == complex_double, name_, "input must be complex");
        out.resize(1);
        out[0].dim = ArrayDimensions({in[0].dim.getX(), in[0].dim.getZ(), in[0].dim.getY()});
        out[0].type = in[0].type;
    }
    void submitToGPU(SubmissionData &d) override
    {
        const ArrayDimensions &dim = d.inputs[0].dim;
        op_status(mti_uv_reorder(d.inputs[0].pointer, d.outputs[0].pointer, (int)dim.getZ(), (int)dim.getY(),
                                 (int)dim.getX(), d.inputs[0].type == complex_double, d.stream),
                  name_);
    }
};

// ZeroToCentre2DShift (StandardOperators/ZeroToCentre2DShift.cpp:36-62): a
// complex array of two or more dimensions, every [y][x] plane shifted.
class ZeroToCentreOp : public ArrayOperator {
public:
    using ArrayOperator::ArrayOperator;
    void validate(const std::vector<ArrayDescriptor> &in, std::vector<ArrayDescriptor> &out, const Properties &,
                  ValidationData &) override
    {
        expect(in.size() == 1, name_, "Only 1 input is supported");
        expect(in[0].type == complex_float || in[0].type == complex_double, name_,
               in[0].type == real_float ? "At current FFT2DShift of a real array is not supported"
                                        : "Unsupported store type of input");
        expect(in[0].dim.getNoOfDimensions() >= 2, name_, "FFT2DShift can work on arrays of 2 dimensions or above!");
        out.resize(1);
        out[0].dim = in[0].dim;
        out[0].type = in[0].type;
    }
    void submitToGPU(SubmissionData &d) override
    {
        const ArrayDimensions &dim = d.inputs[0].dim;
        const long long ny = dim.getY(), nx = dim.getX();
        op_status(mti_centre_shift(d.inputs[0].pointer, d.outputs[0].pointer, dim.total() / (nx * ny), (int)ny,
                                   (int)nx, d.inputs[0].type == complex_double, d.stream),
                  name_);
    }
};

// StokesConverter (op/PostIFFTStokesConverter.cpp:37-67): input 0 complex_float
// [P][v][u], input 1 the real_float normaliser (P elements); parameter
// PolarizationType 0 (linear) or 1 (circular); output real_float [P][v][u].
class StokesConverterOp : public ArrayOperator {
public:
    using ArrayOperator::ArrayOperator;
    void validate(const std::vector<ArrayDescriptor> &in, std::vector<ArrayDescriptor> &out, const Properties &p,
                  ValidationData &) override
    {
        expect(in.size() == 2, name_, "Two inputs are expected");
        const int pt = p.getIntProperty("PolarizationType");
        expect(pt == 0 || pt == 1, name_,
               "PolarizationType parameter is expected to be 0 (linear Polarisation) or 1(Circular Polarisation)");
        expect(in[0].dim.getNoOfDimensions() == 3, name_, "Expected a 3D Array (polarizationxlengthxwidth)");
        expect(in[0].type == complex_float, name_, "input 0 must be complex_float");
        expect(in[0].dim.getZ() == in[1].dim.total(), name_,
               "The No Of polarizations in imager and normalizer are not the same");
        expect(in[1].type == real_float, name_, "Normaliser input type expected to be real_float");
        const long long P = in[0].dim.getZ();
        expect(P == 1 || P == 2 || P == 4, name_, "1, 2 or 4 polarisations");
        out.resize(1);
        out[0].dim = in[0].dim;
        out[0].type = real_float;
    }
    void submitToGPU(SubmissionData &d) override
    {
        const ArrayDimensions &dim = d.inputs[0].dim;
        op_status(mti_stokes((const float *)d.inputs[0].pointer, (const float *)d.inputs[1].pointer,
                             (float *)d.outputs[0].pointer, (int)dim.getZ(), (int)dim.getY(), (int)dim.getX(),
                             d.params.getIntProperty("PolarizationType") == 1 ? MTI_POL_CIRCULAR : MTI_POL_LINEAR,
                             d.stream),
                  name_);
    }
};

// casagridsf (op/CasaGridSF.cpp:38-128).  lmPlanePostCorrection = false,
// generate = true: no input, the taper-only gridding function [dim][dim]
// (complexOutput, double-precision select the type); lmPlanePostCorrection
// = true, generate = false: divide the real input by the taper;
// lmPlanePostCorrection = true, generate = true: the taper grid [ny][nx].
class CasaGridSFOp : public ArrayOperator {
public:
    using ArrayOperator::ArrayOperator;
    void validate(const std::vector<ArrayDescriptor> &in, std::vector<ArrayDescriptor> &out, const Properties &p,
                  ValidationData &) override
    {
        const int support = p.getIntProperty("support"), sampling = p.getIntProperty("sampling");
        const bool post = p.getBoolProperty("lmPlanePostCorrection"), gen = p.getBoolProperty("generate");
        const bool dbl = p.getBoolProperty("double-precision", false);
        expect(support >= 1, name_, "support requested is less then 1");
        expect(support % 2 == 1, name_, "Support is expected to be odd");
        expect(sampling >= 1, name_, "Sampling must be set to at least 1");
        out.resize(1);
        if (!post) {
            expect(gen, name_, "Only generate=true is supported for lmPlanePostCorrection=false");
            expect(in.empty(), name_, "This operator does not take any inputs but validation found inputs");
            const bool cplx = p.getBoolProperty("complexOutput", false);
            const long long dim = (sampling % 2) ? (long long)support * sampling : (long long)support * sampling + 1;
            out[0].dim = ArrayDimensions({dim, dim});
            out[0].type = cplx ? (dbl ? complex_double : complex_float) : (dbl ? real_double : real_float);
        } else if (!gen) {
            expect(in.size() == 1, name_, "An input is expected as to correct");
            expect(in[0].dim.getNoOfDimensions() >= 2, name_, "The input is expected to be at least 2 dimensional");
            expect(in[0].type == real_float || in[0].type == real_double, name_,
                   "Input is expected to be of real_float or real_double type");
            out[0].dim = in[0].dim;
            out[0].type = in[0].type;
        } else {
            expect(in.empty(), name_, "No input is expected");
            const int nx = p.getIntProperty("nx"), ny = p.getIntProperty("ny");
            expect(nx >= 1 && ny >= 1, name_, "Parameters nx or ny must be at least 1");
            out[0].dim = ArrayDimensions({ny, nx});
            out[0].type = dbl ? real_double : real_float;
        }
    }
    void submitToGPU(SubmissionData &d) override
    {
        const bool post = d.params.getBoolProperty("lmPlanePostCorrection"), gen = d.params.getBoolProperty("generate");
        const ArrayDescriptor &o = d.outputs[0];
        const bool f64 = o.type == real_double || o.type == complex_double;
        if (!post) {
            op_status(mti_taper_kernel(o.pointer, d.params.getIntProperty("support"),
                                       d.params.getIntProperty("sampling"),
                                       o.type == complex_float || o.type == complex_double, f64, d.stream),
                      name_);
            return;
        }
        const long long ny = o.dim.getY(), nx = o.dim.getX();
        op_status(mti_taper_correct(gen ? nullptr : d.inputs[0].pointer, o.pointer, o.dim.total() / (nx * ny),
                                    (int)ny, (int)nx, f64, d.stream),
                  name_);
    }
};

// SincCorrector (op/SincCorrector.cpp:37-50): real_float input of 2 or 3
// dimensions, parameter sampling.
class SincCorrectorOp : public ArrayOperator {
public:
    using ArrayOperator::ArrayOperator;
    void validate(const std::vector<ArrayDescriptor> &in, std::vector<ArrayDescriptor> &out, const Properties &p,
                  ValidationData &) override
    {
        expect(in.size() == 1, name_, "This operator expects one input");
        const int nd = in[0].dim.getNoOfDimensions();
        expect(nd == 2 || nd == 3, name_, "Input 0 is expected to be of 2 or 3  dimensions");
        expect(in[0].type == real_float, name_, "This operator works only on real_floats");
        p.getIntProperty("sampling");
        out.resize(1);
        out[0].dim = in[0].dim;
        out[0].type = real_float;
    }
    void submitToGPU(SubmissionData &d) override
    {
        const ArrayDimensions &dim = d.inputs[0].dim;
        const long long ny = dim.getY(), nx = dim.getX();
        op_status(mti_sinc_correct((const float *)d.inputs[0].pointer, (float *)d.outputs[0].pointer,
                                   dim.total() / (nx * ny), (int)ny, (int)nx, d.params.getIntProperty("sampling"),
                                   d.stream),
                  name_);
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
const int32_t *ConvolutionFunctionGenerator::getConvFunctionPositionDevice() const
{
    return mti_kernels_position(need(k_));
}

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
    auto make = [&](ArrayOperator *o) { return std::unique_ptr<ArrayOperator>(o); };
    if (name == "PreProcessUVW") return make(new PreProcessUVWOp(nickname, name));
    if (name == "PreIFFTUVCorrection") return make(new PreIFFTUVCorrectionOp(nickname, name));
    if (name == "IFFT2D") return make(new IFFT2DOp(nickname, name));
    if (name == "ZeroToCentre2DShift") return make(new ZeroToCentreOp(nickname, name));
    if (name == "StokesConverter") return make(new StokesConverterOp(nickname, name));
    if (name == "casagridsf") return make(new CasaGridSFOp(nickname, name));
    if (name == "SincCorrector") return make(new SincCorrectorOp(nickname, name));
    if (name == "SortIndex") return make(new SortIndexOp(nickname, name));
    return nullptr;
}

// ---- modules --------------------------------------------------------------------

namespace {

void *device_alloc(size_t bytes, const char *what)
{
    void *p = nullptr;
    hip_check(hipMalloc(&p, std::max<size_t>(bytes, 256)), what);
    return p;
}

}  // namespace

WImager::WImager(std::shared_ptr<ConvolutionFunctionGenerator> gen, const Properties &conf, int npol)
    : gen_(std::move(gen)), conf_(conf), npol_(npol)
{
    if (!gen_) throw ValidationException("WImager: no convolution-function generator");
    if (!(npol == 1 || npol == 2 || npol == 4))
        throw ValidationException("No of polarisations requested is not supported, it must be 1,2 or 4");
    const mti_image_params &ip = gen_->imageParams();
    if (ip.nx < 1 || ip.ny < 1) throw ValidationException("v or u pixels do not make sense");
    weighting_of(conf_);
    const std::string gt = conf_.getStringProperty("Grid.type", "complex_float");
    if (gt != "complex_float" && gt != "complex_double")
        throw ValidationException("Grid.type must be complex_float or complex_double: " + gt);
    hip_check(hipSetDevice(gen_->device()), "hipSetDevice");
    out_[0].dim = ArrayDimensions({ip.ny, ip.nx, npol});
    out_[0].type = gt == "complex_double" ? complex_double : complex_float;
    out_[0].pointer = device_alloc((size_t)out_[0].dim.total() * (gt == "complex_double" ? 16 : 8), "WImager Grid");
    out_[1].dim = ArrayDimensions({npol});
    out_[1].type = real_float;
    out_[1].pointer = device_alloc(npol * sizeof(float), "WImager NormalizerValue");
    norm64_ = (double *)device_alloc(npol * sizeof(double), "WImager normaliser");
}

WImager::~WImager()
{
    for (void *p : {out_[0].pointer, out_[1].pointer, (void *)norm64_})
        if (p) (void)hipFree(p);
}

void WImager::setInput(int inputNo, const ArrayDescriptor &a)
{
    if (inputNo < 0 || inputNo > 4) throw ValidationException("WImager: inputs are 0-4 (UVW, VISIBILITY, FREQUENCY, WEIGHT, FLAGS)");
    in_[inputNo] = a;
}

const ArrayDescriptor &WImager::getOutput(int outputNo) const
{
    if (outputNo < 0 || outputNo > 1) throw ValidationException("WImager: outputs are 0 (Grid) and 1 (NormalizerValue)");
    return out_[outputNo];
}

void WImager::calculate()
{
    if (!gen_->kernels()) throw ValidationException("WImager: calculateConvFunction() has not been called");
    const std::vector<ArrayDescriptor> in(in_, in_ + 5);
    const ChunkShape c = check_chunk_inputs(in, "WImager", true);
    if (c.P != npol_) throw ValidationException("WImager: chunk has " + std::to_string(c.P) + " polarisations, module " +
                                                std::to_string(npol_));
    hip_check(hipSetDevice(gen_->device()), "hipSetDevice");
    mti_vis v = chunk_vis(in, c);
    mti_grid_opts o{};
    o.chan0 = 0;
    o.nchan = 1;
    o.accumulate = calls_ > 0 ? 1 : 0;
    o.psf = conf_.getBoolProperty("psf", false) ? 1 : 0;
    o.weighting = weighting_of(conf_);
    o.robust = conf_.getDoubleProperty("robust", 0.0);
    o.grid_layout = MTI_GRID_VUP;
    o.grid_precision = out_[0].type == complex_double ? MTI_GRID_C128 : MTI_GRID_C64;
    mti_ctx *ctx = device_context(gen_->device());
    mti_grid_stats st{};
    throw_status(mti_grid(ctx, gen_->kernels(), &gen_->imageParams(), &v, &o, out_[0].pointer, norm64_, &st, stream_),
                 ctx, "WImager");
    throw_status(mti_normalizer_f32(norm64_, (float *)out_[1].pointer, npol_, stream_) == MTI_OK ? MTI_OK : MTI_ERR_HIP,
                 ctx, "WImager normaliser");
    gridded_ += st.gridded_records;
    calls_++;
}

ImageFinalizer::ImageFinalizer(const Properties &conf) : conf_(conf)
{
    image_params_from(conf_);
    const std::string pt = conf_.getStringProperty("PolarizationType", "linear");
    if (pt != "linear" && pt != "circular") throw ValidationException("PolarizationType must be linear or circular");
    norm64_ = (double *)device_alloc(4 * sizeof(double), "ImageFinalizer normaliser");
}

ImageFinalizer::~ImageFinalizer()
{
    for (void *p : {out_.pointer, (void *)norm64_})
        if (p) (void)hipFree(p);
}

void ImageFinalizer::setInput(int inputNo, const ArrayDescriptor &a)
{
    if (inputNo < 0 || inputNo > 1) throw ValidationException("ImageFinalizer: inputs are 0 (UVImage) and 1 (Normalizer)");
    in_[inputNo] = a;
}

const ArrayDescriptor &ImageFinalizer::getOutput(int outputNo) const
{
    if (outputNo != 0) throw ValidationException("ImageFinalizer: one output");
    return out_;
}

void ImageFinalizer::calculate()
{
    const ArrayDescriptor &g = in_[0], &n = in_[1];
    expect(g.type == complex_float || g.type == complex_double, "ImageFinalizer", "UVImage must be complex");
    expect(g.dim.getNoOfDimensions() == 3, "ImageFinalizer", "UVImage must be [v][u][P], got " + g.dim.str());
    const int P = (int)g.dim.getX();
    expect(P == 1 || P == 2 || P == 4, "ImageFinalizer", "1, 2 or 4 polarisations");
    expect(n.dim.total() == P && (n.type == real_float || n.type == real_double), "ImageFinalizer",
           "Normalizer must be real [P]");
    mti_image_params ip = image_params_from(conf_);
    ip.ny = (int)g.dim.getZ();
    ip.nx = (int)g.dim.getY();
    const ArrayDimensions od({P, ip.ny, ip.nx});
    if (!out_.pointer || !(out_.dim == od)) {
        if (out_.pointer) hip_check(hipFree(out_.pointer), "hipFree");
        out_.pointer = device_alloc((size_t)od.total() * sizeof(float), "ImageFinalizer image");
        out_.dim = od;
        out_.type = real_float;
    }
    hipStream_t st = (hipStream_t)stream_;
    const double *norm = (const double *)n.pointer;
    if (n.type == real_float) {
        hipLaunchKernelGGL(k_f32_to_f64, dim3(1), dim3(64), 0, st, (const float *)n.pointer, norm64_, P);
        hip_check(hipGetLastError(), "normaliser conversion");
        norm = norm64_;
    }
    int dev = 0;
    hip_check(hipGetDevice(&dev), "hipGetDevice");
    mti_ctx *ctx = device_context(dev);
    const int pt = conf_.getStringProperty("PolarizationType", "linear") == "circular" ? MTI_POL_CIRCULAR : MTI_POL_LINEAR;
    throw_status(mti_finalize(ctx, &ip, P, pt, 1, MTI_GRID_VUP, g.type == complex_double ? MTI_GRID_C128 : MTI_GRID_C64,
                              g.pointer, norm, (float *)out_.pointer, stream_),
                 ctx, "ImageFinalizer");
}

}  // namespace gafw
}  // namespace mtimager_amd
