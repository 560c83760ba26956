// MuhRec plugin modules backed by the gfx950 kernels of libmuhrec_hip.so.
//
// Module names, parameter keys and the plugin quartet follow the reference
// libraries they replace:
//   ProjectionFilterSingle  frameworks/tomography/Preprocessing/StdPreprocModules/src/ProjectionFilter.cpp:78-375
//   FullLogNorm             .../StdPreprocModules/src/NormPlugins.cpp:20-558 (log and LUT paths)
//   MultiProjBPparallel, MultiProjBP, NearestNeighborBP
//                           frameworks/tomography/Backprojectors/StdBackProjectors/src/StdBackProjectorBase.cpp:21-342
//   FDKbp_single            frameworks/tomography/Backprojectors/FDKBackProjectors/src/fdkreconbase.cpp:20-330
//   FDKbp                   .../FDKBackProjectors/src/fdkbp.cpp, fdkbackproj.cpp:29-30
//   GetModule/Destroy/LibVersion/GetModuleList
//                           .../StdBackProjectors/src/StdBackProjectors.cpp:25-88,
//                           .../StdPreprocModules/src/StdPreprocModules.cpp:40-170
//
// Each module keeps its data in host TImages (the API contract) and moves it
// through device buffers it owns; the arithmetic runs in the HIP kernels.
// Kernel/HIP failures surface as ReconException, as the reference's do.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstring>
#include <limits>
#include <mutex>
#include <sstream>
#include <string>
#include <thread>
#include <vector>

#include "mr/module_api.h"
#include "muhrec_hip.h"
#include "muhrec_modules.h"

namespace {

// volume = 0.0f after the Resize of SetROI (StdBackProjectorBase.cpp:188),
// on up to 16 threads: a 2048^2 x 256 volume is 4 GiB
void zero_volume(kipl::base::TImage<float, 3> &v)
{
    float *d = v.GetDataPtr();
    const size_t n = v.Size();
    const size_t T = n < (size_t(1) << 22) ? 1 : std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    std::vector<std::thread> th;
    for (size_t t = 1; t < T; ++t) th.emplace_back([=] { std::fill(d + n * t / T, d + n * (t + 1) / T, 0.0f); });
    std::fill(d, d + n / T, 0.0f);
    for (auto &x : th) x.join();
}

void check_mr(int rc, const char *what, const char *file, int line)
{
    if (rc != MR_OK) throw ReconException(std::string(what) + ": " + mr_last_error(), file, line);
}
void check_hip(hipError_t e, const char *what, const char *file, int line)
{
    if (e != hipSuccess) throw ReconException(std::string(what) + ": " + hipGetErrorString(e), file, line);
}
#define MRM_CHECK(x) check_mr((x), #x, __FILE__, __LINE__)
#define MRM_HIP(x) check_hip((x), #x, __FILE__, __LINE__)

template <typename T>
std::string value2string(T v)
{
    std::ostringstream s;
    s << v;
    return s.str();
}

std::string lower(std::string s)
{
    for (auto &c : s) c = (char)std::tolower((unsigned char)c);
    return s;
}

bool string2bool(const std::string &str)
{
    const std::string s = lower(str);
    if (s == "true" || s == "yes") return true;
    if (s == "false" || s == "no") return false;
    throw ReconException("string2bool could not parse string", __FILE__, __LINE__);
}

bool optional_bool(std::map<std::string, std::string> &p, const std::string &key, bool dflt)
{
    auto it = p.find(key);
    return (it == p.end() || it->second.empty()) ? dflt : string2bool(it->second);
}

// Axis-tilt settings of a block's plan: bCorrectTilt, fTiltAngle,
// fTiltPivotPosition, and the projection height whose tan|angle| multiple
// shrinks the circle mask (BackProjectorModuleBase.cpp:171-180).
void set_tilt(const ReconConfig &c, const size_t *roi, int &on, float &angle, float &pivot, size_t &rows)
{
    on = c.ProjectionInfo.bCorrectTilt ? 1 : 0;
    angle = c.ProjectionInfo.fTiltAngle;
    pivot = c.ProjectionInfo.fTiltPivotPosition;
    rows = c.ProjectionInfo.imagetype == ReconConfig::cProjections::ImageType_Proj_RepeatSinogram
               ? roi[3]
               : c.ProjectionInfo.nDims[1];
}

// Growable device buffer of floats.
struct DevBuf {
    float *p = nullptr;
    size_t n = 0;
    void reserve(size_t m)
    {
        if (m <= n) return;
        release();
        MRM_HIP(hipMalloc(&p, m * sizeof(float)));
        n = m;
    }
    void release()
    {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
    ~DevBuf() { release(); }
};

// one device-to-host copy on a thread of its own; finish() joins it and
// raises its error (the destructor only joins, for an exception thrown
// meanwhile)
struct Download {
    std::thread t;
    int rc = MR_OK;
    std::string err;
    void start(float *h, const float *d, size_t bytes, hipStream_t s)
    {
        t = std::thread([=] {
            rc = mr_copy_d2h(h, d, bytes, s);
            if (rc != MR_OK) err = mr_last_error();
        });
    }
    void finish()
    {
        if (t.joinable()) t.join();
        if (rc != MR_OK) throw ReconException("mr_copy_d2h: " + err, __FILE__, __LINE__);
    }
    ~Download()
    {
        if (t.joinable()) t.join();
    }
};

// Device hand-off of a filtered stack.  MuhRec's engine passes the
// projection TImage its pre-processing modules changed in place straight on
// to the back-projector's Process(3D)
// (frameworks/tomography/Framework/ReconFramework/src/ReconEngine.cpp:1499-1511,1663-1714).
// With keepondevice=true this library's ProjectionFilterSingle keeps the
// filtered stack in a device buffer of its own after writing it back to the
// host image, and records here which host image that buffer mirrors; the
// parallel-beam back-projectors of this library then read it from there
// instead of copying the host stack to the device again.  A record is taken
// at most once, and only for the same host buffer (address, length, device)
// whose sampled values still equal what the filter wrote, so a stack that
// another module changed in between is copied as usual.  keepondevice is for
// chains whose last pre-processing module is this filter.
class DeviceStack {
public:
    static DeviceStack &get()
    {
        static DeviceStack *s = new DeviceStack();   // never destroyed: no HIP call at exit
        return *s;
    }
    void publish(const void *owner, const float *dev, const float *host, size_t n)
    {
        std::lock_guard<std::mutex> lk(mu_);
        int d = 0;
        MRM_HIP(hipGetDevice(&d));
        owner_ = owner;
        dev_ = dev;
        host_ = host;
        n_ = n;
        device_ = d;
        probe_.resize(std::min(n, kProbes));
        for (size_t k = 0; k < probe_.size(); ++k) probe_[k] = host[index(k)];
    }
    // the owner's buffer is about to change or go away
    void retract(const void *owner)
    {
        std::lock_guard<std::mutex> lk(mu_);
        if (owner_ == owner) clear();
    }
    void count_taken() { taken_.fetch_add(1); }
    size_t taken() const { return taken_.load(); }
    // device copy of host[0, n) or nullptr
    const float *take(const float *host, size_t n)
    {
        std::lock_guard<std::mutex> lk(mu_);
        if (host_ == nullptr || host != host_ || n != n_) return nullptr;
        int d = 0;
        MRM_HIP(hipGetDevice(&d));
        const float *dev = dev_;
        bool same = d == device_;
        for (size_t k = 0; same && k < probe_.size(); ++k)
            same = std::memcmp(&probe_[k], &host[index(k)], sizeof(float)) == 0;
        clear();
        return same ? dev : nullptr;
    }

private:
    static constexpr size_t kProbes = 4096;
    size_t index(size_t k) const { return probe_.size() < 2 ? 0 : k * (n_ - 1) / (probe_.size() - 1); }
    void clear()
    {
        owner_ = nullptr;
        dev_ = host_ = nullptr;
        n_ = 0;
        device_ = -1;
        probe_.clear();
    }
    std::mutex mu_;
    const void *owner_ = nullptr;
    const float *dev_ = nullptr, *host_ = nullptr;
    size_t n_ = 0;
    int device_ = -1;
    std::vector<float> probe_;
    std::atomic<size_t> taken_{0};
};

const char *kFilterNames[] = {"Ram-Lak", "Shepp-Logan", "Hanning", "Hamming", "Butterworth"};

int filter_enum(const std::string &str)
{
    const std::string s = lower(str);
    const char *keys[] = {"ram-lak", "shepp-logan", "hanning", "hamming", "butterworth"};
    for (int i = 0; i < 5; ++i)
        if (s == keys[i]) return i;
    throw ReconException("Could not convert filter string to enum.", __FILE__, __LINE__);
}

// -------------------------------------------------------------------------
// ProjectionFilterSingle
// -------------------------------------------------------------------------
class HipProjectionFilterSingle : public PreprocModuleBase {
public:
    explicit HipProjectionFilterSingle(kipl::interactors::InteractionBase *interactor)
        : PreprocModuleBase("ProjectionFilterSingle", interactor)
    {
    }
    ~HipProjectionFilterSingle() override
    {
        DeviceStack::get().retract(this);
        drop_plan();
        if (up_) (void)hipStreamDestroy(up_);
        if (down_) (void)hipStreamDestroy(down_);
    }

    int Configure(ReconConfig config, std::map<std::string, std::string> parameters) override
    {
        const size_t N = config.ProjectionInfo.roi[2] - config.ProjectionInfo.roi[0];
        m_FilterType = filter_enum(GetStringParameter(parameters, "filtertype"));
        m_fCutOff = GetFloatParameter(parameters, "cutoff");
        m_fOrder = GetFloatParameter(parameters, "order");
        m_bUseBias = string2bool(GetStringParameter(parameters, "usebias"));
        m_fBiasWeight = GetFloatParameter(parameters, "biasweight");
        m_nPaddingDoubler = (size_t)GetIntParameter(parameters, "paddingdoubler");
        // optional, HIP-only: keep the filtered stack on the device for this
        // library's back-projector (see DeviceStack)
        keep_ = optional_bool(parameters, "keepondevice", false);
        nImageSize = N;
        if (!((0 < m_fCutOff) && (m_fCutOff <= 0.5f)))
            throw ReconException("Cut off frequency is not in valid range", __FILE__, __LINE__);
        DeviceStack::get().retract(this);
        drop_plan();
        return 0;
    }

    std::map<std::string, std::string> GetParameters() override
    {
        std::map<std::string, std::string> p;
        p["filtertype"] = kFilterNames[m_FilterType];
        p["cutoff"] = value2string(m_fCutOff);
        p["order"] = value2string(m_fOrder);
        p["usebias"] = m_bUseBias ? "true" : "false";
        p["biasweight"] = value2string(m_fBiasWeight);
        p["paddingdoubler"] = value2string(m_nPaddingDoubler);
        return p;
    }

protected:
    int ProcessCore(kipl::base::TImage<float, 2> &img, std::map<std::string, std::string> &) override
    {
        if (img.Size() == 0) throw ReconException("Empty projection image", __FILE__, __LINE__);
        DeviceStack::get().retract(this);
        filter_rows(img.GetDataPtr(), img.Size(1), img.Size(0));
        return 0;
    }

    // ProjectionFilterBase::ProcessCore(3D) (reference :103-118) filters the
    // projections one by one and polls the interactor before each; here the
    // stack goes through the GPU in chunks of 64 projections with the poll
    // between chunks (an abort leaves the later projections unfiltered, as
    // the reference's loop does).  The chunks are pipelined: while chunk k is
    // uploaded and filtered on one stream, a second host thread downloads
    // chunk k - 1 on another (two device buffers take turns unless the whole
    // stack is kept on the device).
    int ProcessCore(kipl::base::TImage<float, 3> &img, std::map<std::string, std::string> &) override
    {
        if (img.Size() == 0) throw ReconException("Empty projection image", __FILE__, __LINE__);
        DeviceStack::get().retract(this);
        if (img.Size(0) != nImageSize)
            throw ReconException("Projection width does not match the configured ROI", __FILE__, __LINE__);
        make_plan();
        if (!up_) MRM_HIP(hipStreamCreateWithFlags(&up_, hipStreamNonBlocking));
        if (!down_) MRM_HIP(hipStreamCreateWithFlags(&down_, hipStreamNonBlocking));
        // the previous module's device copy of this very stack (FullLogNorm
        // with keepondevice), else the host stack is copied in
        const float *src = DeviceStack::get().take(img.GetDataPtr(), img.Size());
        if (src) DeviceStack::get().count_taken();
        const size_t nProj = img.Size(2), per = img.Size(0) * img.Size(1);
        constexpr size_t kChunk = 64;
        if (keep_) stack_.reserve(img.Size());
        else
            for (DevBuf &b : pipe_) b.reserve(std::min(kChunk, nProj) * per);
        Download down;
        for (size_t p0 = 0, k = 0; p0 < nProj; p0 += kChunk, ++k) {
            if (UpdateStatus((float)p0 / nProj, m_sModuleName)) {
                down.finish();
                return 0;
            }
            const size_t n = std::min(kChunk, nProj - p0), bytes = n * per * sizeof(float);
            float *rows = img.GetDataPtr() + p0 * per;
            float *dev = keep_ ? stack_.p + p0 * per : pipe_[k & 1].p;
            if (src) MRM_HIP(hipMemcpyAsync(dev, src + p0 * per, bytes, hipMemcpyDeviceToDevice, up_));
            else MRM_CHECK(mr_copy_h2d(dev, rows, bytes, up_));
            MRM_CHECK(mr_filter_rows(plan_, dev, img.Size(1) * n, up_));
            MRM_HIP(hipStreamSynchronize(up_));
            // chunk k - 1's download read the buffer chunk k + 1 fills
            down.finish();
            down.start(rows, dev, bytes, down_);
        }
        down.finish();
        if (keep_) DeviceStack::get().publish(this, stack_.p, img.GetDataPtr(), img.Size());
        UpdateStatus(1.0f, m_sModuleName);
        return 0;
    }

private:
    void make_plan()
    {
        if (!plan_)
            MRM_CHECK(mr_filter_plan_create(&plan_, nImageSize, m_FilterType, m_fCutOff, m_fOrder,
                                            m_bUseBias ? 1 : 0, m_fBiasWeight, m_nPaddingDoubler));
    }

    void drop_plan()
    {
        if (plan_) mr_filter_plan_destroy(plan_);
        plan_ = nullptr;
    }

    // filters one projection's host rows through the device buffer buf_
    void filter_rows(float *rows, size_t n_rows, size_t width)
    {
        if (width != nImageSize)
            throw ReconException("Projection width does not match the configured ROI", __FILE__, __LINE__);
        make_plan();
        const size_t n = n_rows * width;
        buf_.reserve(n);
        MRM_CHECK(mr_copy_h2d(buf_.p, rows, n * sizeof(float), nullptr));
        MRM_CHECK(mr_filter_rows(plan_, buf_.p, n_rows, nullptr));
        MRM_CHECK(mr_copy_d2h(rows, buf_.p, n * sizeof(float), nullptr));
    }

    int m_FilterType = 3;  // Hamming
    float m_fCutOff = 0.5f, m_fOrder = 1.0f;
    bool m_bUseBias = true;
    float m_fBiasWeight = 0.1f;
    size_t m_nPaddingDoubler = 2, nImageSize = 0;
    bool keep_ = false;
    mr_filter_plan *plan_ = nullptr;
    DevBuf buf_, stack_, pipe_[2];
    hipStream_t up_ = nullptr, down_ = nullptr;   // chunk uploads + filtering, chunk downloads
};

// -------------------------------------------------------------------------
// FullLogNorm (NormPlugins.cpp:20-90, 232-558), log and LUT paths.  Loading and
// averaging the open-beam / dark images from files (NormBase::ReferenceLoader)
// is I/O: the caller hands the averaged images and their doses to
// SetReferenceImages, the public entry the reference's loader itself ends in.
// -------------------------------------------------------------------------
class HipFullLogNorm : public PreprocModuleBase {
public:
    explicit HipFullLogNorm(kipl::interactors::InteractionBase *interactor)
        : PreprocModuleBase("FullLogNorm", interactor)
    {
    }
    ~HipFullLogNorm() override
    {
        DeviceStack::get().retract(this);
        if (up_) (void)hipStreamDestroy(up_);
        if (down_) (void)hipStreamDestroy(down_);
    }

    int Configure(ReconConfig config, std::map<std::string, std::string> parameters) override
    {
        m_Config = config;
        nOBCount = config.ProjectionInfo.nOBCount;
        nDCCount = config.ProjectionInfo.nDCCount;
        bUseNormROI = string2bool(GetStringParameter(parameters, "usenormregion"));
        bUseLUT = optional_bool(parameters, "uselut", false);
        auto it = parameters.find("referenceaverage");
        sReferenceAverage = it == parameters.end() ? "ImageWeightedAverage" : it->second;
        // optional, HIP-only: keep the normalised stack on the device for this
        // library's ProjectionFilterSingle (see DeviceStack)
        keep_ = optional_bool(parameters, "keepondevice", false);
        DeviceStack::get().retract(this);
        return 0;
    }

    std::map<std::string, std::string> GetParameters() override
    {
        std::map<std::string, std::string> p;
        p["usenormregion"] = bUseNormROI ? "true" : "false";
        p["uselut"] = bUseLUT ? "true" : "false";
        p["referenceaverage"] = sReferenceAverage;
        return p;
    }

    // NormBase::SetROI (NormPlugins.cpp:41-53) loads the references of the
    // block's ROI; here the references given to SetReferenceImages (those of
    // the configured projection ROI, or of one block) are cut to the block's
    // rows.
    bool SetROI(size_t *roi) override
    {
        if ((nOBCount != 0 || nDCCount != 0) && !refs_set_)
            throw ReconException("FullLogNorm: reference images must be given with SetReferenceImages "
                                 "(file loading is not part of the HIP module)",
                                 __FILE__, __LINE__);
        row_off_ = 0;
        block_pix_ = npix_;
        if (refs_set_ && npix_ != 0) {
            const size_t w = roi[2] - roi[0], rows = roi[3] - roi[1];
            if (w != ref_width_) throw ReconException("Reference images do not match the ROI width", __FILE__, __LINE__);
            const size_t ref_rows = npix_ / ref_width_;
            if (rows != ref_rows) {
                const size_t r0 = mConfigRoi1();
                if (roi[1] < r0 || roi[1] - r0 + rows > ref_rows)
                    throw ReconException("Reference images do not cover the ROI", __FILE__, __LINE__);
                row_off_ = (roi[1] - r0) * ref_width_;
            }
            block_pix_ = rows * ref_width_;
        }
        return true;
    }

    // The averaged references and their doses (NormBase::ReferenceLoader's
    // results; flat_dose already has dark_dose removed, as there).
    void SetReferenceImages(kipl::base::TImage<float, 2> dark, kipl::base::TImage<float, 2> flat, float dark_dose,
                            float flat_dose)
    {
        fDarkDose = dark_dose;
        fFlatDose = flat_dose;
        npix_ = 0;
        if (nOBCount != 0) {
            npix_ = flat.Size();
            ref_width_ = flat.Size(0);
            dflat_.reserve(npix_);
            MRM_HIP(hipMemcpy(dflat_.p, flat.GetDataPtr(), npix_ * sizeof(float), hipMemcpyHostToDevice));
        }
        if (nDCCount != 0) {
            if (npix_ && dark.Size() != npix_)
                throw ReconException("Dark and flat images differ in size", __FILE__, __LINE__);
            npix_ = dark.Size();
            ref_width_ = dark.Size(0);
            ddark_.reserve(npix_);
            MRM_HIP(hipMemcpy(ddark_.p, dark.GetDataPtr(), npix_ * sizeof(float), hipMemcpyHostToDevice));
        }
        if (nOBCount != 0)
            MRM_CHECK(mr_lognorm_reference(dflat_.p, nDCCount != 0 ? ddark_.p : nullptr, npix_, fFlatDose, nullptr));
        refs_set_ = true;
        row_off_ = 0;
        block_pix_ = npix_;
    }

protected:
    int ProcessCore(kipl::base::TImage<float, 2> &img, std::map<std::string, std::string> &coeff) override
    {
        // the 2-D path replaces a dose <= 0 by 1 (:343-357), the 3-D one a dose < 1
        float dose = 1.0f;
        if (bUseNormROI) {
            dose = GetFloatParameter(coeff, "dose") - fDarkDose;
            dose = dose <= 0.0f ? 1.0f : dose;
        }
        if (dose != dose) throw ReconException("The projection dose is a NaN", __FILE__, __LINE__);
        DeviceStack::get().retract(this);
        const float cLogDose = std::log(dose);
        const size_t N = img.Size();
        if ((nOBCount != 0 || nDCCount != 0) && block_pix_ != N)
            throw ReconException("Flat field does not match projection", __FILE__, __LINE__);
        buf_.reserve(N);
        MRM_CHECK(mr_copy_h2d(buf_.p, img.GetDataPtr(), N * sizeof(float), nullptr));
        const float *fl = nOBCount != 0 ? dflat_.p + row_off_ : nullptr;
        const float *dk = nDCCount != 0 ? ddark_.p + row_off_ : nullptr;
        if (bUseLUT) MRM_CHECK(mr_lognorm_lut_term(buf_.p, 1, N, fl, dk, &cLogDose, nullptr));
        else MRM_CHECK(mr_lognorm_logdose(buf_.p, 1, N, fl, dk, &cLogDose, nullptr));
        MRM_CHECK(mr_copy_d2h(img.GetDataPtr(), buf_.p, N * sizeof(float), nullptr));
        return 0;
    }

    int ProcessCore(kipl::base::TImage<float, 3> &img, std::map<std::string, std::string> &coeff) override
    {
        const size_t N = img.Size(0) * img.Size(1), nProj = img.Size(2);
        if (nOBCount != 0 && block_pix_ != N) {
            std::ostringstream msg;
            msg << "Flat field (" << block_pix_ << " pixels) does not match projection (" << img.Size(0) << ", "
                << img.Size(1) << ")";
            throw ReconException(msg.str(), __FILE__, __LINE__);
        }
        if (nDCCount != 0 && block_pix_ != N)
            throw ReconException("Dark field does not match projection", __FILE__, __LINE__);
        std::vector<float> dose(nProj);
        if (bUseNormROI) GetFloatParameterVector(coeff, "dose", dose.data(), (int)nProj);
        DeviceStack::get().retract(this);
        if (UpdateStatus(0.0f, m_sModuleName)) return 0;
        buf_.reserve(img.Size());
        if (!up_) MRM_HIP(hipStreamCreateWithFlags(&up_, hipStreamNonBlocking));
        if (!down_) MRM_HIP(hipStreamCreateWithFlags(&down_, hipStreamNonBlocking));
        // chunks of 64 projections: chunk k is uploaded and normalised on one
        // stream while a second host thread downloads chunk k - 1 on another
        constexpr size_t kChunk = 64;
        Download down;
        for (size_t p0 = 0; p0 < nProj; p0 += kChunk) {
            const size_t n = std::min(kChunk, nProj - p0), bytes = n * N * sizeof(float);
            float *d = buf_.p + p0 * N, *h = img.GetDataPtr() + p0 * N;
            MRM_CHECK(mr_copy_h2d(d, h, bytes, up_));
            // the LUT path (NormPlugins.cpp:455-475) or the log path (:476-555)
            MRM_CHECK((bUseLUT ? mr_lognorm_lut : mr_lognorm)(
                d, n, N, nOBCount != 0 ? dflat_.p + row_off_ : nullptr, nDCCount != 0 ? ddark_.p + row_off_ : nullptr,
                bUseNormROI ? dose.data() + p0 : nullptr, fDarkDose, up_));
            MRM_HIP(hipStreamSynchronize(up_));
            down.finish();
            down.start(h, d, bytes, down_);
        }
        down.finish();
        if (keep_) DeviceStack::get().publish(this, buf_.p, img.GetDataPtr(), img.Size());
        UpdateStatus(1.0f, m_sModuleName);
        return 0;
    }

private:
    ReconConfig m_Config;
    size_t mConfigRoi1() const { return m_Config.ProjectionInfo.roi[1]; }
    size_t nOBCount = 0, nDCCount = 0, npix_ = 0;
    size_t ref_width_ = 0, row_off_ = 0, block_pix_ = 0;   // the current block's rows of the references
    bool bUseNormROI = true, bUseLUT = false, refs_set_ = false, keep_ = false;
    std::string sReferenceAverage = "ImageWeightedAverage";
    float fDarkDose = 0.0f, fFlatDose = 1.0f;
    DevBuf dflat_, ddark_, buf_;
    hipStream_t up_ = nullptr, down_ = nullptr;   // chunk uploads + normalisation, chunk downloads
};

// -------------------------------------------------------------------------
// MultiProjBPparallel / MultiProjBP / NearestNeighborBP
//
// The volume is kept on the device in the kernels' ZXY layout and handed to
// MuhRec as a MatrixXYZ TImage (transposed on the GPU before the copy back):
// MuhRec reads slices through the non-virtual GetSlice, which copies an XYZ
// slice row by row instead of gathering it with a stride
// (BackProjectorModuleBase.cpp:62-124); MatrixAlignment tells every reader
// which layout the volume has (the FDK modules use XYZ as well).
// -------------------------------------------------------------------------
class HipStdBackProjector : public BackProjectorModuleBase {
public:
    HipStdBackProjector(const std::string &name, int interp, int multiprojbp,
                        kipl::interactors::InteractionBase *interactor)
        : BackProjectorModuleBase("muhrec", name, MatrixXYZ, interactor), interp_(interp),
          multiprojbp_(multiprojbp)
    {
        nSubVolume[0] = nSubVolume[1] = 1;
    }
    ~HipStdBackProjector() override
    {
        if (plan_) mr_bp_plan_destroy(plan_);
        if (up_) (void)hipStreamDestroy(up_);
    }

    int Configure(ReconConfig config, std::map<std::string, std::string> parameters) override
    {
        mConfig = config;
        nProjectionBufferSize = (size_t)GetIntParameter(parameters, "ProjectionBufferSize");
        nSliceBlock = (size_t)GetIntParameter(parameters, "SliceBlock");
        GetUIntParameterVector(parameters, "SubVolume", nSubVolume, 2);
        bufferDrop_ = optional_bool(parameters, "BufferDrop", true);
        fused_ = optional_bool(parameters, "fused", false);
        return 0;
    }
    int Initialize() override { return 0; }

    std::map<std::string, std::string> GetParameters() override
    {
        std::map<std::string, std::string> p;
        p["ProjectionBufferSize"] = value2string(nProjectionBufferSize);
        p["SliceBlock"] = value2string(nSliceBlock);
        p["SubVolume"] = value2string(nSubVolume[0]) + " " + value2string(nSubVolume[1]);
        p["BufferDrop"] = bufferDrop_ ? "true" : "false";
        p["fused"] = fused_ ? "true" : "false";
        return p;
    }

    void SetROI(size_t *roi) override
    {
        ClearAll();
        SizeU = roi[2] - roi[0];
        SizeV = mConfig.ProjectionInfo.imagetype == ReconConfig::cProjections::ImageType_Proj_RepeatSinogram
                    ? roi[3]
                    : roi[3] - roi[1];
        for (int i = 0; i < 4; ++i) mConfig.ProjectionInfo.roi[i] = roi[i];
        MatrixDims[0] = SizeU;
        MatrixDims[1] = SizeU;
        MatrixDims[2] = SizeV;
        volume.Resize(MatrixDims);
        zero_volume(volume);
        BuildCircleMask();
        if (plan_) mr_bp_plan_destroy(plan_);
        plan_ = nullptr;
        mr_bp_params p{};
        p.width = SizeU;
        p.slices = SizeV;
        p.center = mConfig.ProjectionInfo.fCenter;
        p.direction = (int)mConfig.ProjectionInfo.eDirection;
        p.interp = interp_;
        // a stack handed to Process(3D) from an empty buffer is back-projected
        // in one go with the drop done by the kernels' used-projection list:
        // every ProjectionBufferSize-th projection (sizes 0 and 1 drop every
        // projection, as the reference's counter does); see Process
        p.buffer_size = bufferDrop_ ? (int)std::max<size_t>(nProjectionBufferSize, 1) : 0;
        p.multiprojbp = multiprojbp_;
        p.use_roi = mConfig.MatrixInfo.bUseROI ? 1 : 0;
        for (int i = 0; i < 4; ++i) p.roi[i] = mConfig.MatrixInfo.roi[i];
        p.fused = fused_ ? 1 : 0;
        set_tilt(mConfig, roi, p.correct_tilt, p.tilt_angle, p.tilt_pivot, p.tilt_rows);
        p.tilt_row0 = roi[1];
        MRM_CHECK(mr_bp_plan_create(&plan_, &p));
        dvol_.reserve(volume.Size());
        dxyz_.reserve(volume.Size());
        MRM_HIP(hipMemset(dvol_.p, 0, volume.Size() * sizeof(float)));
        pending_.clear();
        pending_angle_.clear();
        pending_weight_.clear();
        nProjCounter = 0;
    }

    // StdBackProjectorBase::Process(2D) (reference :64-118): buffer the
    // projection; when the last flag is set or the buffer is full, the
    // buffered projections are back-projected, without the newest one when
    // the buffer was full (unless BufferDrop is false).
    size_t Process(kipl::base::TImage<float, 2> proj, float angle, float weight, bool bLast) override
    {
        if (volume.Size() == 0) throw ReconException("The target matrix is not allocated.", __FILE__, __LINE__);
        if (proj.Size(0) != SizeU || proj.Size(1) != SizeV)
            throw ReconException("Projection size does not match the ROI", __FILE__, __LINE__);
        pending_.insert(pending_.end(), proj.GetDataPtr(), proj.GetDataPtr() + proj.Size());
        pending_angle_.push_back(angle);
        pending_weight_.push_back(weight);
        ++nProjCounter;
        if (bLast || nProjectionBufferSize <= nProjCounter) {
            if (nProjectionBufferSize <= nProjCounter && bufferDrop_) {
                pending_.resize(pending_.size() - proj.Size());
                pending_angle_.pop_back();
                pending_weight_.pop_back();
            }
            flush();
        }
        return nProjCounter;
    }

    // StdBackProjectorBase::Process(3D) (reference :120-149): the reference
    // feeds the projections to Process(2D) one by one with the last flag on
    // the final one.  From an empty buffer that back-projects every
    // projection except each ProjectionBufferSize-th (all of them with
    // BufferDrop false), in order, which is exactly the plan's used list: the
    // stack is copied to the device once and back-projected in launches of
    // whole buffers, with the progress/abort poll between launches (the
    // reference polls per projection; here every 256 projections or so,
    // about 20 ms of GPU work at 2048^2 x 256).  Per voxel the additions
    // happen in the same order as one launch, so the result is bit-identical
    // to it.  A stack arriving while Process(2D) has projections buffered
    // continues that buffer projection by projection.
    size_t Process(kipl::base::TImage<float, 3> projections, std::map<std::string, std::string> parameters) override
    {
        if (volume.Size() == 0) throw ReconException("The target matrix is not allocated.", __FILE__, __LINE__);
        const size_t nProj = projections.Size(2);
        if (projections.Size(0) != SizeU || projections.Size(1) != SizeV)
            throw ReconException("Projection size does not match the ROI", __FILE__, __LINE__);
        std::vector<float> weights(nProj + 16), angles(nProj + 16);
        GetFloatParameterVector(parameters, "weights", weights.data(), (int)nProj);
        GetFloatParameterVector(parameters, "angles", angles.data(), (int)nProj);
        const size_t psz = SizeU * SizeV;
        if (nProjCounter != 0 || !pending_angle_.empty()) {
            for (size_t i = 0; i < nProj; ++i) {
                if (UpdateStatus((float)i / nProj, "Back-projecting")) {
                    discard();
                    return 0;
                }
                kipl::base::TImage<float, 2> img(projections.Dims());
                std::memcpy(img.GetDataPtr(), projections.GetLinePtr(0, i), sizeof(float) * psz);
                Process(img, angles[i], weights[i], i == nProj - 1);
            }
            return 0;
        }
        if (UpdateStatus(0.0f, "Back-projecting")) return 0;
        // the filter's device copy of this very stack, else a copy of the host stack
        const float *src = DeviceStack::get().take(projections.GetDataPtr(), nProj * psz);
        if (src) DeviceStack::get().count_taken();
        const bool upload = src == nullptr;
        if (upload) {
            dproj_.reserve(nProj * psz);
            if (!up_) MRM_HIP(hipStreamCreateWithFlags(&up_, hipStreamNonBlocking));
            src = dproj_.p;
        }
        // launches of whole buffers, so that the drop pattern restarts where
        // the reference's buffer does
        const size_t B = std::max<size_t>(nProjectionBufferSize, 1);
        const size_t chunk = std::max<size_t>(B, (256 + B - 1) / B * B);
        for (size_t p0 = 0; p0 < nProj; p0 += chunk) {
            if (p0 > 0 && UpdateStatus((float)p0 / nProj, "Back-projecting")) {
                MRM_HIP(hipDeviceSynchronize());
                download();   // what was back-projected so far, as the reference's volume holds it
                return 0;
            }
            const size_t n = std::min(chunk, nProj - p0);
            // the host stack goes up chunk by chunk on a stream of its own
            // (the copy returns once its bytes are on the device), so chunk
            // k + 1's upload overlaps chunk k's back-projection
            if (upload)
                MRM_CHECK(mr_copy_h2d(dproj_.p + p0 * psz, projections.GetLinePtr(0, p0), n * psz * sizeof(float),
                                      up_));
            MRM_CHECK(mr_bp_parallel(plan_, src + p0 * psz, n, angles.data() + p0, weights.data() + p0,
                                     dvol_.p, 1, nullptr));
        }
        download();
        UpdateStatus(1.0f, "Back-projecting");
        return 0;
    }

    void GetHistogram(float *axis, size_t *hist, size_t nBins) override
    {
        // Min()/Max() of the reference read the first slice of each masked
        // column (StdBackProjectorBase.cpp:282-320); the volume is XYZ here:
        // voxel (x, y, z) at volume.GetLinePtr(y, z)[x]
        float mn = std::numeric_limits<float>::max(), mx = -std::numeric_limits<float>::max();
        for (size_t y = 0; y < mask.size(); ++y) {
            const float *row = volume.GetLinePtr(y, 0);
            for (size_t x = mask[y].first; x < mask[y].second; ++x) {
                mn = std::min(mn, row[x]);
                mx = std::max(mx, row[x]);
            }
        }
        std::memset(axis, 0, sizeof(float) * nBins);
        std::memset(hist, 0, sizeof(size_t) * nBins);
        const float scale = (mx - mn) / (nBins + 1);
        const float inv = 1.0f / scale;
        for (size_t z = 0; z < volume.Size(2); ++z)
            for (size_t y = 0; y < mask.size(); ++y) {
                const float *row = volume.GetLinePtr(y, z);
                for (size_t x = mask[y].first; x < mask[y].second; ++x) {
                    size_t idx = static_cast<size_t>(inv * (row[x] - mn));
                    if (row[x] < mn) idx = 0;
                    else if (nBins <= idx) idx = nBins - 1;
                    hist[idx]++;
                }
            }
        axis[0] = mn + scale / 2.0f;
        for (size_t i = 1; i < nBins; ++i) axis[i] = axis[i - 1] + scale;
    }

private:
    // Back-project every buffered projection (the buffer holds at most
    // ProjectionBufferSize - 1 after the drop, so the plan's drop rule never
    // fires) into the device volume and refresh the host volume.
    void flush()
    {
        const size_t n = pending_angle_.size();
        if (n > 0) {
            dproj_.reserve(pending_.size());
            MRM_CHECK(mr_copy_h2d(dproj_.p, pending_.data(), pending_.size() * sizeof(float), nullptr));
            MRM_CHECK(mr_bp_parallel(plan_, dproj_.p, n, pending_angle_.data(), pending_weight_.data(), dvol_.p,
                                     1, nullptr));
            download();
        }
        discard();
    }

    // host volume (XYZ) from the device volume (ZXY)
    void download()
    {
        MRM_CHECK(mr_volume_zxy_to_xyz(dvol_.p, dxyz_.p, SizeU, SizeV, nullptr));
        MRM_CHECK(mr_copy_d2h(volume.GetDataPtr(), dxyz_.p, volume.Size() * sizeof(float), nullptr));
    }

    void discard()
    {
        pending_.clear();
        pending_angle_.clear();
        pending_weight_.clear();
        nProjCounter = 0;
    }

    int interp_, multiprojbp_;
    size_t SizeU = 0, SizeV = 0, nProjCounter = 0;
    size_t nProjectionBufferSize = 16, nSliceBlock = 32, nSubVolume[2];
    bool bufferDrop_ = true, fused_ = false;
    std::vector<float> pending_, pending_angle_, pending_weight_;
    mr_bp_plan *plan_ = nullptr;
    DevBuf dproj_, dvol_, dxyz_;
    hipStream_t up_ = nullptr;   // chunk uploads of a host stack in Process(3D)
};

// -------------------------------------------------------------------------
// FDKbp_single (float geometry) and FDKbp (double geometry)
// -------------------------------------------------------------------------
class HipFdk : public BackProjectorModuleBase {
public:
    HipFdk(const std::string &name, int module, kipl::interactors::InteractionBase *interactor)
        : BackProjectorModuleBase("muhrec", name, MatrixXYZ, interactor), module_(module)
    {
        nSubVolume[0] = nSubVolume[1] = 1;
    }
    ~HipFdk() override
    {
        if (plan_) mr_fdk_plan_destroy(plan_);
    }

    int Configure(ReconConfig config, std::map<std::string, std::string> parameters) override
    {
        mConfig = config;
        nProjectionBufferSize = (size_t)GetIntParameter(parameters, "ProjectionBufferSize");
        nSliceBlock = (size_t)GetIntParameter(parameters, "SliceBlock");
        GetUIntParameterVector(parameters, "SubVolume", nSubVolume, 2);
        // optional, HIP-only: FMA form of the sampling (FDKbp_single)
        fused_ = module_ == MR_FDK_SINGLE && optional_bool(parameters, "fused", false);
        // optional, HIP-only filter extension: "filter" = "muhrec" (default) or
        // "shepp-logan"; "cosineweight" = Feldkamp cosine pre-weighting
        const auto f = parameters.find("filter");
        if (f == parameters.end() || f->second == "muhrec")
            filter_ = 0;
        else if (f->second == "shepp-logan")
            filter_ = 1;
        else
            throw ReconException("FDK filter must be muhrec or shepp-logan", __FILE__, __LINE__);
        cosine_ = optional_bool(parameters, "cosineweight", false);
        return 0;
    }
    int Initialize() override { return 0; }

    std::map<std::string, std::string> GetParameters() override
    {
        std::map<std::string, std::string> p;
        p["ProjectionBufferSize"] = value2string(nProjectionBufferSize);
        p["SliceBlock"] = value2string(nSliceBlock);
        p["SubVolume"] = value2string(nSubVolume[0]) + " " + value2string(nSubVolume[1]);
        return p;
    }

    void SetROI(size_t *roi) override
    {
        ClearAll();
        SizeU = roi[2] - roi[0];
        SizeV = mConfig.ProjectionInfo.imagetype == ReconConfig::cProjections::ImageType_Proj_RepeatSinogram
                    ? roi[3]
                    : roi[3] - roi[1];
        for (int i = 0; i < 4; ++i) mConfig.ProjectionInfo.roi[i] = roi[i];
        MatrixDims[0] = SizeU;
        MatrixDims[1] = SizeU;
        MatrixDims[2] = SizeV;
        volume.Resize(MatrixDims);
        zero_volume(volume);
        BuildCircleMask();
        if (plan_) mr_fdk_plan_destroy(plan_);
        plan_ = nullptr;
        const auto &P = mConfig.ProjectionInfo;
        mr_fdk_params q{};
        q.sod = P.fSOD;
        q.sdd = P.fSDD;
        q.pixel[0] = P.fResolution[0];
        q.pixel[1] = P.fResolution[1];
        for (int i = 0; i < 3; ++i) q.voxel[i] = mConfig.MatrixInfo.fVoxelSize[i];
        q.ppoint[0] = P.fpPoint[0];
        q.ppoint[1] = P.fpPoint[1];
        q.center = P.fCenter;
        q.direction = (int)P.eDirection;
        for (int i = 0; i < 4; ++i) q.roi[i] = roi[i];
        q.projection_roi3 = P.projection_roi[3];
        q.use_roi = mConfig.MatrixInfo.bUseROI ? 1 : 0;
        for (int i = 0; i < 4; ++i) q.matrix_roi[i] = mConfig.MatrixInfo.roi[i];
        q.module = module_;
        q.fused = fused_ ? 1 : 0;
        q.filter = filter_;
        q.cosine_weight = cosine_ ? 1 : 0;
        set_tilt(mConfig, roi, q.correct_tilt, q.tilt_angle, q.tilt_pivot, q.tilt_rows);
        MRM_CHECK(mr_fdk_plan_create(&plan_, &q));
    }

    // FdkReconBase::Process(3D) (reference :217-277): every projection is
    // weighted, filtered and back-projected; the volume is mirrored in x and
    // flipped in z.
    size_t Process(kipl::base::TImage<float, 3> projections, std::map<std::string, std::string> parameters) override
    {
        if (volume.Size() == 0) throw ReconException("The target matrix is not allocated.", __FILE__, __LINE__);
        const size_t nProj = projections.Size(2);
        if (projections.Size(0) != SizeU)
            throw ReconException("Projection width does not match the ROI", __FILE__, __LINE__);
        std::vector<float> weights(nProj + 16), angles(nProj + 16);
        GetFloatParameterVector(parameters, "weights", weights.data(), (int)nProj);
        GetFloatParameterVector(parameters, "angles", angles.data(), (int)nProj);
        if (UpdateStatus(0.0f, "FDK back-projection")) return 0;
        dproj_.reserve(projections.Size());
        dvol_.reserve(volume.Size());
        MRM_CHECK(mr_copy_h2d(dproj_.p, projections.GetDataPtr(), projections.Size() * sizeof(float), nullptr));
        MRM_CHECK(mr_fdk_bp(plan_, dproj_.p, nProj, projections.Size(1), angles.data(), weights.data(), dvol_.p,
                            nullptr));
        MRM_CHECK(mr_copy_d2h(volume.GetDataPtr(), dvol_.p, volume.Size() * sizeof(float), nullptr));
        UpdateStatus(1.0f, "FDK back-projection");
        return 0;
    }

    void GetHistogram(float *axis, size_t *hist, size_t nBins) override
    {
        float mn = std::numeric_limits<float>::max(), mx = -std::numeric_limits<float>::max();
        const float *d = volume.GetDataPtr();
        for (size_t i = 0; i < volume.Size(); ++i) {
            mn = std::min(mn, d[i]);
            mx = std::max(mx, d[i]);
        }
        std::memset(axis, 0, sizeof(float) * nBins);
        std::memset(hist, 0, sizeof(size_t) * nBins);
        const float scale = (mx - mn) / (nBins + 1);
        const float inv = 1.0f / scale;
        for (size_t i = 0; i < volume.Size(); ++i) {
            size_t idx = static_cast<size_t>(inv * (d[i] - mn));
            if (d[i] < mn) idx = 0;
            else if (nBins <= idx) idx = nBins - 1;
            hist[idx]++;
        }
        axis[0] = mn + scale / 2.0f;
        for (size_t i = 1; i < nBins; ++i) axis[i] = axis[i - 1] + scale;
    }

    float GetVoxelSize() override { return mConfig.MatrixInfo.fVoxelSize[0]; }

private:
    size_t SizeU = 0, SizeV = 0;
    size_t nProjectionBufferSize = 2000, nSliceBlock = 32, nSubVolume[2];
    int module_ = MR_FDK_SINGLE, filter_ = 0;
    bool fused_ = false, cosine_ = false;
    mr_fdk_plan *plan_ = nullptr;
    DevBuf dproj_, dvol_;
};

const char *kBpNames[] = {"MultiProjBP", "MultiProjBPparallel", "NearestNeighborBP", "FDKbp_single", "FDKbp"};

}  // namespace

// ---------------------------------------------------------------------------
// Plugin quartet
// ---------------------------------------------------------------------------
extern "C" void *GetModule(const char *application, const char *name, void *vinteractor)
{
    if (application == nullptr || name == nullptr) return nullptr;
    auto *interactor = reinterpret_cast<kipl::interactors::InteractionBase *>(vinteractor);
    const std::string app = application, n = name;
    try {
        if (app == "muhrec") {
            if (n == "ProjectionFilterSingle") return static_cast<PreprocModuleBase *>(new HipProjectionFilterSingle(interactor));
            if (n == "FullLogNorm") return static_cast<PreprocModuleBase *>(new HipFullLogNorm(interactor));
            return nullptr;
        }
        if (app == "muhrecbp") {
            if (n == "MultiProjBP") return static_cast<BackProjectorModuleBase *>(new HipStdBackProjector(n, 0, 1, interactor));
            if (n == "MultiProjBPparallel") return static_cast<BackProjectorModuleBase *>(new HipStdBackProjector(n, 0, 0, interactor));
            if (n == "NearestNeighborBP") return static_cast<BackProjectorModuleBase *>(new HipStdBackProjector(n, 1, 0, interactor));
            if (n == "FDKbp_single") return static_cast<BackProjectorModuleBase *>(new HipFdk(n, MR_FDK_SINGLE, interactor));
            if (n == "FDKbp") return static_cast<BackProjectorModuleBase *>(new HipFdk(n, MR_FDK_DOUBLE, interactor));
        }
    } catch (...) {
        return nullptr;
    }
    return nullptr;
}

extern "C" int Destroy(const char *application, void *obj)
{
    if (application == nullptr) return -1;
    const std::string app = application;
    if (app == "muhrec") {
        delete reinterpret_cast<PreprocModuleBase *>(obj);
        return 0;
    }
    if (app == "muhrecbp") {
        delete reinterpret_cast<BackProjectorModuleBase *>(obj);
        return 0;
    }
    return -1;
}

extern "C" int LibVersion() { return -1; }

extern "C" int GetModuleList(const char *application, void *listptr)
{
    if (application == nullptr || listptr == nullptr) return -1;
    auto *list = reinterpret_cast<std::map<std::string, std::map<std::string, std::string>> *>(listptr);
    const std::string app = application;
    if (app == "muhrec") {
        HipProjectionFilterSingle f(nullptr);
        (*list)["ProjectionFilterSingle"] = f.GetParameters();
        HipFullLogNorm g(nullptr);
        (*list)["FullLogNorm"] = g.GetParameters();
        return 0;
    }
    if (app == "muhrecbp") {
        for (const char *n : kBpNames) {
            BackProjectorModuleBase *m = reinterpret_cast<BackProjectorModuleBase *>(GetModule("muhrecbp", n, nullptr));
            if (!m) continue;
            (*list)[n] = m->GetParameters();
            delete m;
        }
        return 0;
    }
    return -1;
}

extern "C" size_t mr_modules_device_handoffs(void) { return DeviceStack::get().taken(); }

// FullLogNorm's averaged references for a module made by GetModule("muhrec",
// "FullLogNorm"): the call the reference's module makes itself after loading
// the files (FullLogNorm::LoadReferenceImages, NormPlugins.cpp:260-297).
// dark / flat: rows x width host images, NULL when their count is 0.
// Returns 0, -1 for another module, -2 on an error (mr_modules_last_error).
namespace {
thread_local std::string g_mod_err;
}
extern "C" int mr_modules_set_references(void *module, const float *dark, const float *flat, size_t rows,
                                         size_t width, float dark_dose, float flat_dose)
{
    g_mod_err.clear();
    auto *norm = dynamic_cast<HipFullLogNorm *>(reinterpret_cast<PreprocModuleBase *>(module));
    if (!norm) return -1;
    try {
        size_t dims[2] = {width, rows};
        kipl::base::TImage<float, 2> d(dims), f(dims);
        if (dark) std::memcpy(d.GetDataPtr(), dark, sizeof(float) * rows * width);
        if (flat) std::memcpy(f.GetDataPtr(), flat, sizeof(float) * rows * width);
        norm->SetReferenceImages(d, f, dark_dose, flat_dose);
    } catch (const std::exception &e) {
        g_mod_err = e.what();
        return -2;
    }
    return 0;
}

extern "C" const char *mr_modules_last_error(void) { return g_mod_err.c_str(); }
