// Implementation of the MuhRec module API declared in include/mr/module_api.h:
// parameter helpers (core/modules/ModuleConfig/src/ParameterHandling.cpp:25-77),
// ProcessModuleBase (core/modules/ModuleConfig/src/ProcessModuleBase.cpp:32-86)
// and BackProjectorModuleBase
// (frameworks/tomography/Framework/ReconFramework/src/BackProjectorModuleBase.cpp:20-218).
#include "mr/module_api.h"

#include <cmath>
#include <cstdlib>
#include <cstring>
#include <sstream>

kipl::logging::Logger::LogLevel kipl::logging::Logger::s_level = kipl::logging::Logger::LogWarning;

std::string GetStringParameter(std::map<std::string, std::string> parameters, std::string parameter)
{
    std::string str = parameters[parameter];
    if (str.empty()) throw ModuleException("Could not find parameter " + parameter, __FILE__, __LINE__);
    return str;
}

float GetFloatParameter(std::map<std::string, std::string> parameters, std::string parameter)
{
    return static_cast<float>(std::atof(GetStringParameter(parameters, parameter).c_str()));
}

int GetIntParameter(std::map<std::string, std::string> parameters, std::string parameter)
{
    return std::atoi(GetStringParameter(parameters, parameter).c_str());
}

namespace {
// Whitespace-separated values; when the list is short the last parsed value
// is the last one written (kipl::strings::String2Array behaviour).
template <typename T, typename Conv>
void parse_list(const std::string &s, T *arr, int N, Conv conv)
{
    std::istringstream is(s);
    std::string tok;
    for (int i = 0; i < N; ++i) {
        if (!(is >> tok)) break;
        arr[i] = conv(tok);
    }
}
}  // namespace

int GetFloatParameterVector(std::map<std::string, std::string> parameters, std::string parameter, float *arr,
                            int N)
{
    parse_list(GetStringParameter(parameters, parameter), arr, N,
               [](const std::string &t) { return static_cast<float>(std::atof(t.c_str())); });
    return N;
}

int GetUIntParameterVector(std::map<std::string, std::string> parameters, std::string parameter, size_t *arr,
                           int N)
{
    parse_list(GetStringParameter(parameters, parameter), arr, N,
               [](const std::string &t) { return static_cast<size_t>(std::atol(t.c_str())); });
    return N;
}

// ---------------------------------------------------------------- ProcessModuleBase
ProcessModuleBase::ProcessModuleBase(std::string name, kipl::interactors::InteractionBase *interactor)
    : logger(name), m_sModuleName(name), m_Interactor(interactor)
{
}

int ProcessModuleBase::Process(kipl::base::TImage<float, 2> &img, std::map<std::string, std::string> &parameters)
{
    timer.Tic();
    int res = ProcessCore(img, parameters);
    timer.Toc();
    return res;
}

int ProcessModuleBase::Process(kipl::base::TImage<float, 3> &img, std::map<std::string, std::string> &parameters)
{
    timer.Tic();
    int res = ProcessCore(img, parameters);
    timer.Toc();
    return res;
}

int ProcessModuleBase::ProcessCore(kipl::base::TImage<float, 2> &, std::map<std::string, std::string> &)
{
    throw ModuleException("Processing module " + m_sModuleName + " does not support 2D processing.", __FILE__,
                          __LINE__);
}

int ProcessModuleBase::ProcessCore(kipl::base::TImage<float, 3> &, std::map<std::string, std::string> &)
{
    throw ModuleException("Processing module " + m_sModuleName + " does not support 3D processing.", __FILE__,
                          __LINE__);
}

bool ProcessModuleBase::UpdateStatus(float val, std::string msg)
{
    return m_Interactor != nullptr ? m_Interactor->SetProgress(val, msg) : false;
}

// ---------------------------------------------------------------- BackProjectorModuleBase
BackProjectorModuleBase::BackProjectorModuleBase(std::string application, std::string name,
                                                 eMatrixAlignment alignment,
                                                 kipl::interactors::InteractionBase *interactor)
    : MatrixAlignment(alignment), logger(name), m_sModuleName(name), m_sApplication(application),
      m_Interactor(interactor)
{
}

size_t BackProjectorModuleBase::Process(kipl::base::TImage<float, 2>, float, float, bool)
{
    throw ReconException("Backprojector module " + m_sModuleName + " does not support 2D processing.",
                         __FILE__, __LINE__);
}

size_t BackProjectorModuleBase::Process(kipl::base::TImage<float, 3>, std::map<std::string, std::string>)
{
    throw ReconException("Backprojector module " + m_sModuleName + " does not support 3D processing.",
                         __FILE__, __LINE__);
}

kipl::base::TImage<float, 2> BackProjectorModuleBase::GetSlice(size_t idx)
{
    size_t origin[2] = {0, 0};
    size_t dims[2] = {0, 0};
    if (mConfig.MatrixInfo.bUseROI) {
        dims[0] = mConfig.MatrixInfo.roi[2] - mConfig.MatrixInfo.roi[0] + 1;
        dims[1] = mConfig.MatrixInfo.roi[3] - mConfig.MatrixInfo.roi[1] + 1;
        origin[0] = mConfig.MatrixInfo.roi[0];
        origin[1] = mConfig.MatrixInfo.roi[1];
    } else if (MatrixAlignment == MatrixXYZ) {
        dims[0] = volume.Size(0);
        dims[1] = volume.Size(1);
    } else {
        dims[0] = volume.Size(1);
        dims[1] = volume.Size(2);
    }
    kipl::base::TImage<float, 2> slice(dims);
    const size_t sxy = volume.Size(0) * volume.Size(1);
    const size_t sx = volume.Size(0);
    if (MatrixAlignment == MatrixXYZ) {
        if (volume.Size(2) <= idx) {
            std::ostringstream msg;
            msg << "Slice index greater than matrix size (XYZ matrix). Slice index is " << idx
                << " and Matrix Sixe is: " << volume.Size(2);
            throw ReconException(msg.str(), __FILE__, __LINE__);
        }
        for (size_t i = 0; i < slice.Size(1); ++i)
            std::memcpy(slice.GetLinePtr(i), volume.GetLinePtr(i + origin[1], idx) + origin[0],
                        slice.Size(0) * sizeof(float));
    } else {
        if (volume.Size(0) <= idx) {
            std::ostringstream msg;
            msg << "Slice index greater than matrix size (ZXY matrix). Slice index is " << idx
                << " and Matrix Sixe is: " << volume.Size(0);
            throw ReconException(msg.str(), __FILE__, __LINE__);
        }
        const float *pv = volume.GetDataPtr() + idx + origin[0] * sx + origin[1] * sxy;
        for (size_t y = 0; y < dims[1]; ++y) {
            float *ps = slice.GetLinePtr(y);
            for (size_t x = 0; x < dims[0]; ++x) ps[x] = pv[x * sx];
            pv += sxy;
        }
    }
    return slice;
}

size_t BackProjectorModuleBase::GetNSlices()
{
    return MatrixAlignment == MatrixXYZ ? volume.Size(2) : volume.Size(0);
}

bool BackProjectorModuleBase::UpdateStatus(float val, std::string msg)
{
    return m_Interactor != nullptr ? m_Interactor->SetProgress(val, msg) : false;
}

void BackProjectorModuleBase::ClearAll()
{
    mask.clear();
    std::memset(MatrixDims, 0, sizeof(MatrixDims));
}

void BackProjectorModuleBase::BuildCircleMask()
{
    size_t nx = 0, ny = 0;
    if (MatrixAlignment == MatrixXYZ) {
        nx = MatrixDims[0];
        ny = MatrixDims[1];
    } else {
        nx = MatrixDims[1];
        ny = MatrixDims[2];
    }
    const float cx = static_cast<float>(nx >> 1);
    const float cy = static_cast<float>(ny >> 1);
    mask.resize(ny);
    float R = cx - 1;
    if (mConfig.ProjectionInfo.bCorrectTilt) {
        // the tilted axis leaves the detector in the outer voxels: the radius
        // shrinks by floor(tan|angle| * projection height)
        const float slices = static_cast<float>(
            mConfig.ProjectionInfo.imagetype == ReconConfig::cProjections::ImageType_Proj_RepeatSinogram
                ? mConfig.ProjectionInfo.roi[3]
                : mConfig.ProjectionInfo.nDims[1]);
        R -= std::floor(std::tan(std::fabs(mConfig.ProjectionInfo.fTiltAngle) * 3.1415926535897932384626f / 180.0f) *
                        slices);
    }
    const float R2 = R * R;
    for (size_t i = 0; i < ny; ++i) {
        const float y = static_cast<float>(i) - cy;
        const float y2 = y * y;
        if (y2 <= R2) {
            const float x = std::sqrt(R2 - y2);
            mask[i].first = static_cast<size_t>(std::ceil(cx - x));
            mask[i].second = std::min(nx - 1u, static_cast<size_t>(std::floor(cx + x)));
        } else {
            mask[i].first = 0u;
            mask[i].second = 0u;
        }
        if (mConfig.MatrixInfo.bUseROI) {
            if (mConfig.MatrixInfo.roi[1] <= i && i <= mConfig.MatrixInfo.roi[3]) {
                if (mConfig.ProjectionInfo.beamgeometry == ReconConfig::cProjections::BeamGeometry_Parallel) {
                    mask[i].first = std::max(mask[i].first, mConfig.MatrixInfo.roi[0]);
                    mask[i].second = std::min(mask[i].second, mConfig.MatrixInfo.roi[2]);
                } else if (mConfig.ProjectionInfo.beamgeometry == ReconConfig::cProjections::BeamGeometry_Cone) {
                    mask[i].first = std::max(mask[i].first, nx - mConfig.MatrixInfo.roi[2]);
                    mask[i].second = std::min(mask[i].second, nx - mConfig.MatrixInfo.roi[0]);
                }
            } else {
                mask[i].first = 0u;
                mask[i].second = 0u;
            }
        }
    }
}
