// mtimager_amd_gafw.hpp -- C++ host layer over the C-ABI, shaped like the
// reference's GAFW operator interface so that mt-imager code written against
// GAFW operators can drive the MI355X path with the same names, argument
// meaning and error behaviour.
//
// Mirrors (reference paths relative to code/gafw/):
//   StoreType              Implimentation/GeneralImplimentation/gafw-enums.h:22-34
//   ArrayDimensions        pil/ArrayDimensions.h:25-83 (row-major, last dim fastest;
//                          getX = last, getY = second last, getZ = third last)
//   ArrayDescriptor        GPUArrayDescriptor, Implimentation/GPUImplimentation/GPUafw.h:62-68
//   ValidationData         Implimentation/GPUImplimentation/ValidationData.h:25-34
//   SubmissionData         GPUSubmissionData, Implimentation/GPUImplimentation/GPUSubmissionData.h:27-47
//   ArrayOperator          GPUArrayOperator, Implimentation/GPUImplimentation/GPUArrayOperator.h:27-43
//   ImagerFactoryHelper    mt-imager lib/operators/ImagerFactoryHelper.cpp:45-87
//   ConvolutionFunctionGenerator
//                          mt-imager lib/components/ConvolutionFunctionsGenerator/
//                          ConvolutionFunctionGenerator.h:23-32 (WProjectionConvFunction,
//                          OnlyTaperConvFunction)
//   Module, WImager, ImageFinalizer
//                          pil/Module.h:28-46; mt-imager lib/components/WImager.cpp:150-313,
//                          ImageFinalizer.cpp:31-86
//   ValidationException / GpuException: the reference throws ValidationException from
//   validate() and CudaException from submitToGPU (GPUafw.h:28-32).
//
// There is no engine here: the caller passes the arrays' descriptors to
// validate() and the device pointers plus a hipStream_t to submitToGPU(),
// exactly the data a GAFW engine would hand its operators.
#ifndef MTIMAGER_AMD_GAFW_HPP
#define MTIMAGER_AMD_GAFW_HPP

#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "mtimager_amd.h"

namespace mtimager_amd {
namespace gafw {

enum StoreType {
    complex_float = 0,
    complex_double,
    real_float,
    real_double,
    real_int,
    real_uint,
    real_longint,
    real_shortint,
    real_ulongint,
    real_ushortint,
    StoreTypeUnknown
};

class ArrayDimensions {
public:
    ArrayDimensions() = default;
    explicit ArrayDimensions(std::vector<long long> dims) : d_(std::move(dims)) {}
    int getNoOfDimensions() const { return (int)d_.size(); }
    long long getDimension(int i) const { return d_.at(i); }
    void setDimension(int i, long long v) { d_.at(i) = v; }
    long long getX() const { return from_end(1); }
    long long getY() const { return from_end(2); }
    long long getZ() const { return from_end(3); }
    long long total() const;
    bool operator==(const ArrayDimensions &o) const { return d_ == o.d_; }
    std::string str() const;

private:
    long long from_end(int k) const { return (int)d_.size() >= k ? d_[d_.size() - k] : 1; }
    std::vector<long long> d_;
};

struct ArrayDescriptor {
    void *pointer = nullptr;
    ArrayDimensions dim;
    StoreType type = StoreTypeUnknown;
};

// String-valued properties with typed getters (the reference's CppProperties).
class Properties {
public:
    void setProperty(const std::string &k, const std::string &v) { m_[k] = v; }
    void setProperty(const std::string &k, double v);
    bool isPropertySet(const std::string &k) const { return m_.count(k) != 0; }
    std::string getStringProperty(const std::string &k) const;
    int getIntProperty(const std::string &k) const;
    double getDoubleProperty(const std::string &k) const;
    float getFloatProperty(const std::string &k) const { return (float)getDoubleProperty(k); }
    bool getBoolProperty(const std::string &k) const;
    // with a default when unset
    std::string getStringProperty(const std::string &k, const std::string &d) const { return isPropertySet(k) ? getStringProperty(k) : d; }
    int getIntProperty(const std::string &k, int d) const { return isPropertySet(k) ? getIntProperty(k) : d; }
    double getDoubleProperty(const std::string &k, double d) const { return isPropertySet(k) ? getDoubleProperty(k) : d; }
    bool getBoolProperty(const std::string &k, bool d) const { return isPropertySet(k) ? getBoolProperty(k) : d; }

private:
    std::map<std::string, std::string> m_;
};

enum class BufferMode { Normal, DeallocBeforeSubmit, UnknownSize };

struct ValidationData {
    BufferMode bufferMode = BufferMode::Normal;
    size_t bufferSize = 0;
};

struct SubmissionData {
    std::vector<ArrayDescriptor> inputs;
    std::vector<ArrayDescriptor> outputs;
    void *stream = nullptr;  // hipStream_t
    Properties params;
    void *bufferGPUPointer = nullptr;
    size_t buffersize = 0;
};

class ValidationException : public std::runtime_error {
public:
    using std::runtime_error::runtime_error;
};

class GpuException : public std::runtime_error {
public:
    using std::runtime_error::runtime_error;
};

// Imaging parameters from the mt-imager property keys
// (comp/ConfigurationManager.cpp:33-62): image.nx, image.ny,
// image.lintervalsize / image.mintervalsize (arcsec), mode (normal | wproj),
// wproj.wplanes, wproj.convFunction.sampling, taper.support, taper.sampling.
mti_image_params image_params_from(const Properties &p);

class ConvolutionFunctionGenerator;

// The GPU operator contract: validate() fixes the output dimensions/types
// from the inputs (throws ValidationException), submitToGPU() enqueues the
// work on data.stream (throws GpuException).
class ArrayOperator {
public:
    ArrayOperator(std::string nickname, std::string name) : nickname_(std::move(nickname)), name_(std::move(name)) {}
    virtual ~ArrayOperator() = default;
    virtual void validate(const std::vector<ArrayDescriptor> &inputs, std::vector<ArrayDescriptor> &outputs,
                          const Properties &params, ValidationData &data) = 0;
    virtual void submitToGPU(SubmissionData &data) = 0;
    virtual void postRunExecute(void *) {}
    const std::string &nickname() const { return nickname_; }
    const std::string &name() const { return name_; }
    // convolution functions used by the gridding operators
    void setConvolutionFunction(std::shared_ptr<ConvolutionFunctionGenerator> g) { conv_ = std::move(g); }

protected:
    std::string nickname_, name_;
    std::shared_ptr<ConvolutionFunctionGenerator> conv_;
};

// W-projection / taper-only convolution functions on the device
// (WProjectionConvFunction / OnlyTaperConvFunction).
class ConvolutionFunctionGenerator {
public:
    ConvolutionFunctionGenerator(int device, const Properties &params, void *stream = nullptr);
    ~ConvolutionFunctionGenerator();
    ConvolutionFunctionGenerator(const ConvolutionFunctionGenerator &) = delete;
    ConvolutionFunctionGenerator &operator=(const ConvolutionFunctionGenerator &) = delete;
    void calculateConvFunction();
    const void *getConvFunction() const;                       // device table
    std::vector<std::pair<int, long long>> getConvFunctionPositionData() const;  // {support, base}
    const int32_t *getConvFunctionPositionDevice() const;      // device int [planes][2] {support, base}
    const float *getConvFunctionSumData() const;               // device [planes][T*T]
    int getMaxSupport() const;
    int getSampling() const;
    float getWSquareIncrement() const;
    const mti_kernels *kernels() const { return k_; }
    const mti_image_params &imageParams() const { return p_; }
    int device() const { return device_; }

private:
    int device_;
    mti_image_params p_;
    int precision_;
    void *stream_;
    mti_kernels *k_ = nullptr;
};

// Creates operators by the reference's names (ImagerFactoryHelper::createOperator,
// op/ImagerFactoryHelper.cpp:45-87, and the GAFW standard operators of
// StandardOperators/GPUOperatorsFactoryHelper.cu:79-131), each with the
// reference's own inputs, outputs, types and parameters:
//   "PreProcessUVW"        6 inputs (convolution position data int [W][2],
//                          sums [W][T*T], UVW [3][N], FREQUENCY [N], WEIGHT [N][P],
//                          FLAGS [N][P]) -> 7 outputs (op/WProjPreProcessUVW.cpp:39-86)
//   "PreIFFTUVCorrection"  complex [v][u][P] -> complex [P][v][u]
//   "IFFT2D"               complex_float [..][ny][nx], unnormalised inverse, any size
//   "ZeroToCentre2DShift"  complex [..][ny][nx] -> same
//   "StokesConverter"      complex_float [P][v][u] + real_float [P] -> real_float [P][v][u]
//   "casagridsf"           its three modes (lmPlanePostCorrection, generate)
//   "SincCorrector"        real_float [..][v][u] -> same
//   "SortIndex"            (vis/SortIndex.cpp:46-87; not a reference operator)
//                          antenna1/2 int [N] -> int index [N], real_longint [1] kept
// Names whose reference contracts consume arrays this build does not
// produce -- the gridding plan and its consumer (CreateIndexAndReorder,
// NotSameSupportManipulate, CreateBlockDataIndex, CreateCompressionPlan,
// CompressVisibilities, ConvolutionGridder, AccumulatedSum, ValueChangeDetect,
// Sum, GriddedPointsCalculator) and the convolution-function chain
// (WProjPreFFTConvFunc, multiFFT2D, ConvFunctionSupportFind,
// ConvFunctionReOrganize, SumConvolutionFunctions, ArraysToVector,
// SingleMultiply) -- give nullptr, as unknown names do: their work is done
// by the WImager module and the convolution-function generator below.
class ImagerFactoryHelper {
public:
    explicit ImagerFactoryHelper(int device = 0) : device_(device) {}
    std::unique_ptr<ArrayOperator> createOperator(const std::string &nickname, const std::string &name) const;

private:
    int device_;
};

// The reference's names for the two convolution-function generators.
typedef ConvolutionFunctionGenerator WProjectionConvFunction;
typedef ConvolutionFunctionGenerator OnlyTaperConvFunction;

// A module: inputs bound with setInput, calculate() enqueues the work on the
// module's stream, getOutput(i) describes output i (device memory owned by
// the module, valid until the module is destroyed).  GAFW::Module,
// pil/Module.h:28-46, without an engine: inputs are array descriptors.
class Module {
public:
    virtual ~Module() = default;
    virtual void reset() = 0;
    virtual void calculate() = 0;
    virtual void setInput(int inputNo, const ArrayDescriptor &a) = 0;
    virtual const ArrayDescriptor &getOutput(int outputNo) const = 0;
    void setStream(void *stream) { stream_ = stream; }

protected:
    void *stream_ = nullptr;  // hipStream_t
};

// The W-projection imager of one channel (comp/WImager.cpp:150-313).
// Inputs 0 UVW [3][N] real_float, 1 VISIBILITY [N][2P] real_float,
// 2 FREQUENCY [N] real_float (1/lambda), 3 WEIGHT [N][P] real_float,
// 4 FLAGS [N][P] real_int.  Outputs 0 Grid [v][u][P] complex_float
// (conf "Grid.type" = complex_double for complex128), 1 NormalizerValue [P]
// real_float.  Every calculate() grids the current chunk and adds it to the
// outputs (the reference's overwrite() chunks, WImager.cpp:271-290), the first
// one after construction or reset() starting from zero.  conf: psf,
// weighting, robust.
class WImager : public Module {
public:
    WImager(std::shared_ptr<ConvolutionFunctionGenerator> gen, const Properties &conf, int npol);
    ~WImager() override;
    WImager(const WImager &) = delete;
    WImager &operator=(const WImager &) = delete;
    void reset() override { calls_ = 0; }
    void calculate() override;
    void setInput(int inputNo, const ArrayDescriptor &a) override;
    const ArrayDescriptor &getOutput(int outputNo) const override;
    long long griddedRecords() const { return gridded_; }  // Statistic:TotalGriddedRecords

private:
    std::shared_ptr<ConvolutionFunctionGenerator> gen_;
    Properties conf_;
    int npol_;
    ArrayDescriptor in_[5], out_[2];
    double *norm64_ = nullptr;  // FP64 normaliser the chunks add into
    int calls_ = 0;
    long long gridded_ = 0;
};

// The image finalizer (comp/ImageFinalizer.cpp:31-86): input 0 UVImage
// [v][u][P] complex_float or complex_double, input 1 Normalizer [P]
// real_float or real_double; output 0 real_float [P][v][u] Stokes.  conf:
// the image.* / wproj.* / taper.* keys (the correction's sampling) and
// PolarizationType (linear | circular).
class ImageFinalizer : public Module {
public:
    explicit ImageFinalizer(const Properties &conf);
    ~ImageFinalizer() override;
    ImageFinalizer(const ImageFinalizer &) = delete;
    ImageFinalizer &operator=(const ImageFinalizer &) = delete;
    void reset() override {}
    void calculate() override;
    void setInput(int inputNo, const ArrayDescriptor &a) override;
    const ArrayDescriptor &getOutput(int outputNo) const override;

private:
    Properties conf_;
    ArrayDescriptor in_[2], out_;
    double *norm64_ = nullptr;
};

// The per-device context the operators share (one submit thread per device).
mti_ctx *device_context(int device);

}  // namespace gafw
}  // namespace mtimager_amd

#endif
