// MuhRec module API used by the HIP modules, restated from the reference
// interfaces it must stay source-compatible with:
//   kipl::interactors::InteractionBase  core/kipl/kipl/include/interactors/interactionbase.h:19-73
//   kipl::logging::Logger               core/kipl/kipl/include/logging/logger.h:47-53
//   ReconException / ModuleException     frameworks/tomography/Framework/ReconFramework/include/ReconException.h,
//                                        core/modules/ModuleConfig/include/ModuleException.h
//   ParameterHandling                    core/modules/ModuleConfig/src/ParameterHandling.cpp:25-77
//   ReconConfig (hot-path fields)        frameworks/tomography/Framework/ReconFramework/include/ReconConfig.h:23-216
//   ProcessModuleBase                    core/modules/ModuleConfig/include/ProcessModuleBase.h:20-94
//   PreprocModuleBase                    frameworks/tomography/Framework/ReconFramework/include/PreprocModuleBase.h:32-76
//   BackProjectorModuleBase              frameworks/tomography/Framework/ReconFramework/include/BackProjectorModuleBase.h:28-136
#pragma once

#include <chrono>
#include <cstddef>
#include <iostream>
#include <map>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "timage.h"

namespace kipl {
namespace base {
enum eRotationDirection { RotationDirCW = 0, RotationDirCCW = 1 };
}  // namespace base

namespace interactors {
// Progress/abort link between modules and a caller (GUI, CLI, test).
class InteractionBase {
public:
    explicit InteractionBase(std::string name = "InteractionBase") : name_(std::move(name)) {}
    virtual ~InteractionBase() = default;
    virtual void Reset() { abort_ = false; finished_ = false; progress_ = 0.0f; }
    virtual void Abort() { abort_ = true; }
    virtual void Done() { finished_ = true; }
    // Returns true when the caller asked to abort.
    virtual bool SetProgress(float progress, std::string msg = "")
    {
        progress_ = progress;
        msg_ = std::move(msg);
        return abort_;
    }
    virtual bool SetOverallProgress(float progress)
    {
        overall_ = progress;
        return abort_;
    }
    virtual bool Aborted() { return abort_; }
    virtual bool Finished() { return finished_; }
    virtual float CurrentProgress() { return progress_; }
    virtual float CurrentOverallProgress() { return overall_; }
    virtual std::string CurrentMessage() { return msg_; }

protected:
    std::string name_, msg_;
    bool abort_ = false, finished_ = false;
    float progress_ = 0.0f, overall_ = 0.0f;
};
}  // namespace interactors

namespace logging {
class Logger {
public:
    enum LogLevel { LogError = 0, LogWarning, LogMessage, LogVerbose, LogDebug };
    explicit Logger(std::string name) : name_(std::move(name)) {}
    void operator()(LogLevel level, const std::string &msg) const
    {
        if (level <= s_level) std::clog << "[" << name_ << "] " << msg << std::endl;
    }
    void message(const std::string &msg) const { (*this)(LogMessage, msg); }
    static LogLevel s_level;

private:
    std::string name_;
};
}  // namespace logging

namespace profile {
// Wall-clock accumulator with the reference Timer's Tic/Toc/ElapsedSeconds.
class Timer {
public:
    void Tic() { t0_ = std::chrono::steady_clock::now(); }
    void Toc() { total_ += std::chrono::duration<double>(std::chrono::steady_clock::now() - t0_).count(); }
    void Reset() { total_ = 0.0; }
    double ElapsedSeconds() const { return total_; }

private:
    std::chrono::steady_clock::time_point t0_{};
    double total_ = 0.0;
};
}  // namespace profile
}  // namespace kipl

// Exceptions carry the throwing file and line, as the reference's do.
class ReconException : public std::runtime_error {
public:
    ReconException(const std::string &msg, const std::string &file = "", size_t line = 0)
        : std::runtime_error(msg + (file.empty() ? "" : " (" + file + ":" + std::to_string(line) + ")")) {}
};
class ModuleException : public std::runtime_error {
public:
    ModuleException(const std::string &msg, const std::string &file = "", size_t line = 0)
        : std::runtime_error(msg + (file.empty() ? "" : " (" + file + ":" + std::to_string(line) + ")")) {}
};

// Parameter helpers: a missing or empty key throws ModuleException.
std::string GetStringParameter(std::map<std::string, std::string> parameters, std::string parameter);
float GetFloatParameter(std::map<std::string, std::string> parameters, std::string parameter);
int GetIntParameter(std::map<std::string, std::string> parameters, std::string parameter);
int GetFloatParameterVector(std::map<std::string, std::string> parameters, std::string parameter,
                            float *arr, int N);
int GetUIntParameterVector(std::map<std::string, std::string> parameters, std::string parameter,
                           size_t *arr, int N);

// The ReconConfig fields the hot path reads (ReconConfig.h:75-156).
class ReconConfig {
public:
    struct cProjections {
        enum eScanType { SequentialScan, GoldenSectionScan };
        enum eImageType { ImageType_Projections = 0, ImageType_Sinograms, ImageType_Proj_RepeatProjection,
                          ImageType_Proj_RepeatSinogram };
        enum eBeamGeometry { BeamGeometry_Parallel = 0, BeamGeometry_Cone, BeamGeometry_Helix };
        size_t nDims[2] = {2048, 2048};
        size_t roi[4] = {0, 0, 2047, 2047};
        size_t projection_roi[4] = {0, 2047, 0, 2047};
        float fCenter = 0.0f;
        kipl::base::eRotationDirection eDirection = kipl::base::RotationDirCW;
        float fResolution[2] = {0.01f, 0.01f};
        float fScanArc[2] = {0.0f, 360.0f};
        size_t nFirstIndex = 1, nLastIndex = 625, nProjectionStep = 1, nMargin = 2;
        eScanType scantype = SequentialScan;
        eImageType imagetype = ImageType_Projections;
        eBeamGeometry beamgeometry = BeamGeometry_Parallel;
        float fSOD = 100.0f, fSDD = 100.0f;
        float fpPoint[2] = {500.0f, 500.0f};
        bool bCorrectTilt = false;
        float fTiltAngle = 0.0f, fTiltPivotPosition = 0.0f;
        // reference images (ReconConfig.h:103-113)
        std::string sReferencePath, sOBFileMask, sDCFileMask;
        size_t nOBFirstIndex = 1, nOBCount = 0, nDCFirstIndex = 1, nDCCount = 0;
        size_t dose_roi[4] = {0, 0, 0, 0};
    } ProjectionInfo;
    struct cMatrix {
        float fRotation = 0.0f;
        bool bUseROI = false;
        size_t roi[4] = {0, 0, 0, 0};
        size_t nDims[3] = {0, 0, 0};
        float fVoxelSize[3] = {0.1f, 0.1f, 0.1f};
    } MatrixInfo;
    struct cModule {
        std::string m_sSharedObject, m_sModule;
        std::map<std::string, std::string> parameters;
    } backprojector;
};

class ProcessModuleBase {
public:
    explicit ProcessModuleBase(std::string name = "ProcessModuleBase",
                               kipl::interactors::InteractionBase *interactor = nullptr);
    virtual ~ProcessModuleBase() = default;
    // In place; times ProcessCore.
    virtual int Process(kipl::base::TImage<float, 2> &img, std::map<std::string, std::string> &parameters);
    virtual int Process(kipl::base::TImage<float, 3> &img, std::map<std::string, std::string> &parameters);
    virtual int Configure(std::map<std::string, std::string> parameters) = 0;
    virtual int Initialize() { timer.Reset(); return 0; }
    virtual std::map<std::string, std::string> GetParameters() = 0;
    std::string ModuleName() { return m_sModuleName; }
    double ExecTime() { return timer.ElapsedSeconds(); }
    virtual std::string Version() { return ""; }

protected:
    virtual int ProcessCore(kipl::base::TImage<float, 2> &img, std::map<std::string, std::string> &parameters);
    virtual int ProcessCore(kipl::base::TImage<float, 3> &img, std::map<std::string, std::string> &parameters);
    bool UpdateStatus(float val, std::string msg);
    kipl::logging::Logger logger;
    std::string m_sModuleName;
    kipl::profile::Timer timer;
    kipl::interactors::InteractionBase *m_Interactor;
};

class PreprocModuleBase : public ProcessModuleBase {
public:
    explicit PreprocModuleBase(std::string name = "PreprocModuleBase",
                               kipl::interactors::InteractionBase *interactor = nullptr)
        : ProcessModuleBase(std::move(name), interactor) {}
    virtual int Configure(ReconConfig config, std::map<std::string, std::string> parameters) = 0;
    virtual bool SetROI(size_t *roi) { (void)roi; return false; }
    virtual ~PreprocModuleBase() = default;

protected:
    int Configure(std::map<std::string, std::string> parameters) override
    {
        return Configure(ReconConfig(), parameters);
    }
};

class BackProjectorModuleBase {
public:
    enum eMatrixAlignment { MatrixXYZ, MatrixZXY };
    const eMatrixAlignment MatrixAlignment;

    BackProjectorModuleBase(std::string application, std::string name, eMatrixAlignment alignment,
                            kipl::interactors::InteractionBase *interactor = nullptr);
    virtual size_t Process(kipl::base::TImage<float, 2> proj, float angle, float weight, bool bLastProjection);
    virtual size_t Process(kipl::base::TImage<float, 3> proj, std::map<std::string, std::string> parameters);
    virtual int Configure(ReconConfig config, std::map<std::string, std::string> parameters) = 0;
    virtual int Initialize() = 0;
    virtual std::map<std::string, std::string> GetParameters() = 0;
    virtual void SetROI(size_t *roi) = 0;
    kipl::base::TImage<float, 3> GetVolume() { return volume; }
    kipl::base::TImage<float, 2> GetSlice(size_t idx);
    size_t GetNSlices();
    std::string Name() { return m_sModuleName; }
    virtual void GetMatrixDims(size_t *dims)
    {
        dims[0] = volume.Size(0);
        dims[1] = volume.Size(1);
        dims[2] = volume.Size(2);
    }
    virtual void GetHistogram(float *x, size_t *y, size_t N) = 0;
    virtual float GetVoxelSize() { return mConfig.ProjectionInfo.fResolution[0]; }
    virtual ~BackProjectorModuleBase() = default;
    double ExecTime() { return timer.ElapsedSeconds(); }

protected:
    virtual void ClearAll();
    void BuildCircleMask();
    bool UpdateStatus(float val, std::string msg);

    kipl::logging::Logger logger;
    ReconConfig mConfig;
    std::string m_sModuleName;
    kipl::base::TImage<float, 3> volume;
    std::vector<std::pair<size_t, size_t>> mask;
    kipl::profile::Timer timer;
    size_t MatrixDims[3] = {0, 0, 0};
    std::string m_sApplication;
    kipl::interactors::InteractionBase *m_Interactor;
};
