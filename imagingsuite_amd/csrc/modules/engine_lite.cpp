// Block driver over the MuhRec plugin boundary: loads modules with
// dlopen/dlsym exactly as the reference loader does
// (frameworks/tomography/Framework/ReconFramework/src/ModuleItem.cpp:82-168)
// and runs the parallel-beam and cone-beam block loops of
// ReconEngine::Run3DFull / Process3D / BackProject3D / TransferMatrix
// (frameworks/tomography/Framework/ReconFramework/src/ReconEngine.cpp:531-574, 838-1185, 1314-1714).
#include <dlfcn.h>

#include <algorithm>
#include <chrono>
#include <cstring>
#include <exception>
#include <sstream>
#include <string>
#include <thread>
#include <vector>

#include "mr/module_api.h"
#include "muhrec_hip.h"
#include "muhrec_modules.h"

namespace {

thread_local std::string g_err;

typedef void *(*GetModule_t)(const char *, const char *, void *);
typedef int (*Destroy_t)(const char *, void *);
typedef int (*GetModuleList_t)(const char *, void *);
typedef int (*SetReferences_t)(void *, const float *, const float *, size_t, size_t, float, float);

std::map<std::string, std::string> parse_kv(const char *s)
{
    std::map<std::string, std::string> m;
    if (!s) return m;
    std::string str(s);
    size_t pos = 0;
    while (pos <= str.size()) {
        size_t end = str.find(';', pos);
        if (end == std::string::npos) end = str.size();
        std::string item = str.substr(pos, end - pos);
        size_t eq = item.find('=');
        if (eq != std::string::npos) {
            auto trim = [](std::string t) {
                size_t a = t.find_first_not_of(" \t"), b = t.find_last_not_of(" \t");
                return a == std::string::npos ? std::string() : t.substr(a, b - a + 1);
            };
            m[trim(item.substr(0, eq))] = trim(item.substr(eq + 1));
        }
        pos = end + 1;
    }
    return m;
}

template <typename T>
void parse_n(const std::string &s, T *out, int n)
{
    std::istringstream is(s);
    for (int i = 0; i < n; ++i) is >> out[i];
}

ReconConfig make_config(const std::map<std::string, std::string> &kv)
{
    ReconConfig c;
    auto get = [&](const char *k) -> const std::string * {
        auto it = kv.find(k);
        return it == kv.end() ? nullptr : &it->second;
    };
    if (auto v = get("geometry"))
        c.ProjectionInfo.beamgeometry = (*v == "cone") ? ReconConfig::cProjections::BeamGeometry_Cone
                                                       : ReconConfig::cProjections::BeamGeometry_Parallel;
    if (auto v = get("center")) c.ProjectionInfo.fCenter = std::stof(*v);
    if (auto v = get("direction"))
        c.ProjectionInfo.eDirection = (*v == "ccw") ? kipl::base::RotationDirCCW : kipl::base::RotationDirCW;
    if (auto v = get("resolution")) parse_n(*v, c.ProjectionInfo.fResolution, 2);
    if (auto v = get("voxelsize")) parse_n(*v, c.MatrixInfo.fVoxelSize, 3);
    if (auto v = get("sod")) c.ProjectionInfo.fSOD = std::stof(*v);
    if (auto v = get("sdd")) c.ProjectionInfo.fSDD = std::stof(*v);
    if (auto v = get("ppoint")) parse_n(*v, c.ProjectionInfo.fpPoint, 2);
    if (auto v = get("roi")) parse_n(*v, c.ProjectionInfo.roi, 4);
    if (auto v = get("projection_roi3")) c.ProjectionInfo.projection_roi[3] = std::stoul(*v);
    if (auto v = get("usematrixroi")) c.MatrixInfo.bUseROI = (*v == "true");
    if (auto v = get("ob_count")) c.ProjectionInfo.nOBCount = std::stoul(*v);
    if (auto v = get("dc_count")) c.ProjectionInfo.nDCCount = std::stoul(*v);
    if (auto v = get("matrixroi")) parse_n(*v, c.MatrixInfo.roi, 4);
    // ProjectionInfo's "dims", "correcttilt", "tiltangle", "tiltpivot" (ReconConfig.cpp:151,163-165)
    if (auto v = get("dims")) parse_n(*v, c.ProjectionInfo.nDims, 2);
    if (auto v = get("correcttilt")) c.ProjectionInfo.bCorrectTilt = (*v == "true");
    if (auto v = get("tiltangle")) c.ProjectionInfo.fTiltAngle = std::stof(*v);
    if (auto v = get("tiltpivot")) c.ProjectionInfo.fTiltPivotPosition = std::stof(*v);
    return c;
}

// ProjectionReader::Read sends angles and weights as ostream text (6 digits).
std::string list_text(const float *v, size_t n)
{
    std::ostringstream s;
    for (size_t i = 0; i < n; ++i) s << v[i] << " ";
    return s.str();
}

}  // namespace

// Interactor that asks for an abort from its `limit`-th progress report on
// (the GUI's cancel button, pressed at a known point).
class AbortAfter : public kipl::interactors::InteractionBase {
public:
    explicit AbortAfter(int limit) : limit_(limit) {}
    bool SetProgress(float, std::string) override { return ++calls_ >= limit_; }

private:
    int limit_, calls_ = 0;
};

struct mre_engine {
    void *so = nullptr;
    GetModule_t get = nullptr;
    Destroy_t destroy = nullptr;
    // the pre-processing chain, in order
    std::vector<PreprocModuleBase *> pres;
    std::vector<std::string> pre_names;
    std::vector<std::map<std::string, std::string>> pre_params;
    BackProjectorModuleBase *bp = nullptr;
    std::map<std::string, std::string> bp_params;
    std::vector<float> doses;   // per-projection doses ("dose" list); empty: 1
    ReconConfig cfg;
    size_t slice_block = 32;
    // 0: Run3DFull's SliceBlock loop (Process(3D)); 1: ReconEngine::Run, one
    // projection at a time through Process(2D) with its last flag
    // (never set, ReconEngine.cpp:505); 2: the same with the last flag on the
    // final projection
    int projection_at_a_time = 0;
    AbortAfter *abort_bp = nullptr;   // bp_abort_after: the back-projector's interactor
    AbortAfter *abort_filter = nullptr;   // filter_abort_after: ProjectionFilterSingle's interactor
    float rescale_between = 1.0f;     // stack multiplied by this between the two modules
    // wall time of the last run's phases, ms: slab copy-in, pre-processing
    // Process, back-projector SetROI, back-projector Process, GetSlice copy-out
    double phase_ms[5] = {0, 0, 0, 0, 0};
    // the SliceBlock stack, kept between blocks and runs of the same shape so
    // that its pages are allocated and zeroed once (held until destroy)
    kipl::base::TImage<float, 3> stack;
};

namespace {

class PhaseClock {
public:
    explicit PhaseClock(double *acc) : acc_(acc), t_(std::chrono::steady_clock::now()) {}
    // adds the time since the previous lap to phase i
    void lap(int i)
    {
        const auto now = std::chrono::steady_clock::now();
        acc_[i] += std::chrono::duration<double, std::milli>(now - t_).count();
        t_ = now;
    }

private:
    double *acc_;
    std::chrono::steady_clock::time_point t_;
};

// Runs f(t, nt) for t = 0..nt-1 on nt threads (t = 0 on the calling thread),
// nt = the hardware threads, at most 16; rethrows the first exception raised.
template <typename F>
void on_threads(F f)
{
    const int nt = (int)std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    std::vector<std::exception_ptr> err(nt);
    auto run = [&](int t) {
        try {
            f(t, nt);
        } catch (...) {
            err[t] = std::current_exception();
        }
    };
    std::vector<std::thread> th;
    for (int t = 1; t < nt; ++t) th.emplace_back(run, t);
    run(0);
    for (auto &x : th) x.join();
    for (auto &x : err)
        if (x) std::rethrow_exception(x);
}

// [a, b) of n items for thread t of nt
inline void share(size_t n, int t, int nt, size_t &a, size_t &b)
{
    a = n * (size_t)t / (size_t)nt;
    b = n * (size_t)(t + 1) / (size_t)nt;
}

}  // namespace

extern "C" const char *mre_last_error(void) { return g_err.c_str(); }

extern "C" int mre_engine_destroy(mre_engine *e)
{
    if (!e) return 0;
    if (e->destroy) {
        for (auto *m : e->pres)
            if (m) e->destroy("muhrec", m);
        if (e->bp) e->destroy("muhrecbp", e->bp);
    }
    if (e->so) dlclose(e->so);
    delete e->abort_bp;
    delete e->abort_filter;
    delete e;
    return 0;
}

extern "C" int mre_engine_create(mre_engine **engine, const char *module_so, const char *preproc_module,
                                 const char *preproc_params, const char *bp_module, const char *bp_params,
                                 const char *config)
{
    g_err.clear();
    if (!engine || !module_so || !bp_module) {
        g_err = "NULL argument";
        return 1;
    }
    *engine = nullptr;
    auto *e = new mre_engine();
    try {
        e->so = dlopen(module_so, RTLD_LAZY);
        if (!e->so) throw ReconException(std::string("Could not load module library: ") + dlerror(), __FILE__, __LINE__);
        e->get = reinterpret_cast<GetModule_t>(dlsym(e->so, "GetModule"));
        if (!e->get) throw ReconException("Could not find symbol GetModule", __FILE__, __LINE__);
        e->destroy = reinterpret_cast<Destroy_t>(dlsym(e->so, "Destroy"));
        if (!e->destroy) throw ReconException("Could not find symbol Destroy", __FILE__, __LINE__);
        e->cfg = make_config(parse_kv(config));
        // pre-processing chain: module names and their parameter strings separated by '|'
        auto split = [](const char *str) {
            std::vector<std::string> out;
            std::string v = str ? str : "";
            size_t pos = 0;
            for (;;) {
                const size_t end = v.find('|', pos);
                out.push_back(v.substr(pos, end == std::string::npos ? std::string::npos : end - pos));
                if (end == std::string::npos) break;
                pos = end + 1;
            }
            return out;
        };
        if (preproc_module && *preproc_module) {
            e->pre_names = split(preproc_module);
            const std::vector<std::string> ps = split(preproc_params);
            for (size_t i = 0; i < e->pre_names.size(); ++i)
                e->pre_params.push_back(parse_kv(i < ps.size() ? ps[i].c_str() : ""));
        }
        e->bp_params = parse_kv(bp_params);
        e->cfg.backprojector.m_sModule = bp_module;
        e->cfg.backprojector.parameters = e->bp_params;
        const auto kvc = parse_kv(config);
        // each module of this library hands its device copy of the stack to
        // the next one when that can take it (FullLogNorm ->
        // ProjectionFilterSingle -> a parallel-beam back-projector) unless the
        // caller says otherwise
        const std::string bpn = bp_module;
        const auto dh = kvc.find("device_handoff");
        if (dh != kvc.end() && dh->second != "true" && dh->second != "false")
            throw ReconException("device_handoff must be true or false", __FILE__, __LINE__);
        const bool handoff = dh == kvc.end() || dh->second == "true";
        const bool bp_takes = bpn == "MultiProjBPparallel" || bpn == "MultiProjBP" || bpn == "NearestNeighborBP";
        for (size_t i = 0; i < e->pre_names.size(); ++i) {
            const std::string &n = e->pre_names[i];
            const bool next_takes = i + 1 < e->pre_names.size() ? e->pre_names[i + 1] == "ProjectionFilterSingle"
                                                                  : bp_takes;
            const bool gives = (n == "FullLogNorm" && i + 1 < e->pre_names.size() &&
                                e->pre_names[i + 1] == "ProjectionFilterSingle") ||
                               (n == "ProjectionFilterSingle" && next_takes && i + 1 == e->pre_names.size());
            if (gives && e->pre_params[i].find("keepondevice") == e->pre_params[i].end())
                e->pre_params[i]["keepondevice"] = handoff ? "true" : "false";
        }
        const auto rb = kvc.find("rescale_between");
        if (rb != kvc.end()) e->rescale_between = std::stof(rb->second);
        const auto fa = kvc.find("filter_abort_after");
        if (fa != kvc.end()) e->abort_filter = new AbortAfter(std::stoi(fa->second));
        for (size_t i = 0; i < e->pre_names.size(); ++i) {
            kipl::interactors::InteractionBase *ia =
                e->pre_names[i] == "ProjectionFilterSingle" ? e->abort_filter : nullptr;
            auto *m = static_cast<PreprocModuleBase *>(e->get("muhrec", e->pre_names[i].c_str(), ia));
            if (!m) throw ReconException("Failed to create module " + e->pre_names[i], __FILE__, __LINE__);
            e->pres.push_back(m);
            m->Configure(e->cfg, e->pre_params[i]);
        }
        const auto ab = kvc.find("bp_abort_after");
        if (ab != kvc.end()) e->abort_bp = new AbortAfter(std::stoi(ab->second));
        e->bp = static_cast<BackProjectorModuleBase *>(e->get("muhrecbp", bp_module, e->abort_bp));
        if (!e->bp) throw ReconException(std::string("Failed to create module ") + bp_module, __FILE__, __LINE__);
        e->bp->Configure(e->cfg, e->bp_params);
        e->slice_block = (size_t)GetIntParameter(e->bp_params, "SliceBlock");
        if (e->slice_block == 0) throw ReconException("SliceBlock must be positive", __FILE__, __LINE__);
        const auto kv = parse_kv(config);
        const auto pa = kv.find("projection_at_a_time");
        if (pa != kv.end()) {
            if (pa->second == "reference") e->projection_at_a_time = 1;
            else if (pa->second == "final") e->projection_at_a_time = 2;
            else if (pa->second != "false") throw ReconException("projection_at_a_time must be false, reference or final",
                                                                 __FILE__, __LINE__);
        }
    } catch (const std::exception &ex) {
        g_err = ex.what();
        mre_engine_destroy(e);
        return 2;
    }
    *engine = e;
    return 0;
}

extern "C" int mre_engine_run(mre_engine *e, const float *projections, size_t n_proj, size_t rows, size_t width,
                              const float *angles, const float *weights, float *volume)
{
    g_err.clear();
    if (!e || !projections || !angles || !weights || !volume) {
        g_err = "NULL argument";
        return 1;
    }
    try {
        for (double &t : e->phase_ms) t = 0.0;
        PhaseClock clock(e->phase_ms);
        const auto &P = e->cfg.ProjectionInfo;
        const size_t *roi = P.roi;
        const bool cone = P.beamgeometry == ReconConfig::cProjections::BeamGeometry_Cone;
        if (roi[2] <= roi[0] || roi[3] <= roi[1]) throw ReconException("Invalid projection ROI", __FILE__, __LINE__);
        if (width != roi[2] - roi[0]) throw ReconException("width does not match roi", __FILE__, __LINE__);
        if (!cone && rows != roi[3] - roi[1]) throw ReconException("rows do not match roi", __FILE__, __LINE__);
        const size_t total = roi[3] - roi[1];
        if (e->projection_at_a_time) {
            // ReconEngine::Run (ReconEngine.cpp:455-520): the whole ROI at once,
            // per projection the pre-processing modules' Process(2D) and the
            // back-projector's Process(2D) with the float angle and weight
            if (cone) throw ReconException("projection-at-a-time runs parallel beam only", __FILE__, __LINE__);
            size_t block[4] = {roi[0], roi[1], roi[2], roi[3]};
            for (auto *m : e->pres) m->SetROI(block);   // ReconEngine.cpp:420-430
            e->bp->SetROI(block);
            std::map<std::string, std::string> params;
            params["dose"] = "1";
            size_t dims[2] = {width, rows};
            for (size_t p = 0; p < n_proj; ++p) {
                kipl::base::TImage<float, 2> img(dims);
                std::memcpy(img.GetDataPtr(), projections + p * rows * width, sizeof(float) * rows * width);
                if (!e->doses.empty()) params["dose"] = list_text(&e->doses[p], 1);
                for (auto *m : e->pres) m->Process(img, params);
                const bool last = e->projection_at_a_time == 2 ? p + 1 == n_proj : n_proj < p + 1;
                e->bp->Process(img, angles[p], weights[p], last);
            }
            for (size_t i = 0; i < total; ++i) {
                kipl::base::TImage<float, 2> slice = e->bp->GetSlice(i);
                std::memcpy(volume + i * width * width, slice.GetDataPtr(),
                            sizeof(float) * std::min(slice.Size(), width * width));
            }
            return 0;
        }
        const size_t sb = e->slice_block;
        const size_t nblocks = total / sb + (total % sb ? 1 : 0);
        std::map<std::string, std::string> params;
        params["angles"] = list_text(angles, n_proj);
        params["weights"] = list_text(weights, n_proj);
        if (!e->doses.empty()) {
            if (e->doses.size() != n_proj) throw ReconException("dose list does not match the projections", __FILE__, __LINE__);
            params["dose"] = list_text(e->doses.data(), n_proj);
        } else {
            std::string dose;
            for (size_t i = 0; i < n_proj; ++i) dose += "1 ";
            params["dose"] = dose;
        }
        for (size_t b = 0; b < nblocks; ++b) {
            size_t block[4] = {roi[0], roi[1] + b * sb, roi[2], std::min(roi[1] + (b + 1) * sb, roi[3])};
            size_t r0 = block[1] - roi[1], r1 = block[3] - roi[1];   // rows of the stack used
            if (cone) {
                mr_fdk_params q{};
                q.sod = P.fSOD; q.sdd = P.fSDD;
                q.pixel[0] = P.fResolution[0]; q.pixel[1] = P.fResolution[1];
                for (int i = 0; i < 3; ++i) q.voxel[i] = e->cfg.MatrixInfo.fVoxelSize[i];
                q.ppoint[0] = P.fpPoint[0]; q.ppoint[1] = P.fpPoint[1];
                q.center = P.fCenter; q.direction = (int)P.eDirection;
                for (int i = 0; i < 4; ++i) q.roi[i] = block[i];
                q.projection_roi3 = P.projection_roi[3];
                size_t cb[4];
                if (mr_fdk_window(&q, cb) != MR_OK) throw ReconException(mr_last_error(), __FILE__, __LINE__);
                r0 = cb[1];
                r1 = std::min(cb[3], rows);
                if (r1 <= r0) throw ReconException("empty detector window", __FILE__, __LINE__);
            }
            const size_t nr = r1 - r0;
            size_t dims[3] = {width, nr, n_proj};
            // the block's stack (Resize keeps the buffer when the shape is
            // unchanged; Clone detaches it from a module still holding the
            // previous block), filled projection by projection on all threads
            e->stack.Resize(dims);
            e->stack.Clone();
            kipl::base::TImage<float, 3> &img = e->stack;
            on_threads([&](int t, int nt) {
                size_t a, b;
                share(n_proj, t, nt, a, b);
                for (size_t p = a; p < b; ++p)
                    std::memcpy(img.GetLinePtr(0, p), projections + (p * rows + r0) * width,
                                sizeof(float) * nr * width);
            });
            clock.lap(0);
            // ReconEngine::Process3D (ReconEngine.cpp:1204-1214): SetROI of every
            // pre-processing module for the block, then its Process
            // (for cone beam the projections' rows are the detector window)
            size_t proj_roi[4] = {block[0], r0 + (cone ? 0 : roi[1]), block[2], r1 + (cone ? 0 : roi[1])};
            for (auto *m : e->pres) m->SetROI(proj_roi);
            for (auto *m : e->pres) m->Process(img, params);
            if (e->rescale_between != 1.0f) {
                float *d = img.GetDataPtr();
                const float k = e->rescale_between;
                on_threads([&](int t, int nt) {
                    size_t a, b;
                    share(img.Size(), t, nt, a, b);
                    for (size_t i = a; i < b; ++i) d[i] *= k;
                });
            }
            clock.lap(1);
            e->bp->SetROI(block);
            clock.lap(2);
            e->bp->Process(img, params);
            clock.lap(3);
            // GetSlice only reads the module's volume: slices in parallel
            const size_t ns = block[3] - block[1];
            on_threads([&](int t, int nt) {
                size_t a, b;
                share(ns, t, nt, a, b);
                for (size_t i = a; i < b; ++i) {
                    kipl::base::TImage<float, 2> slice = e->bp->GetSlice(i);
                    std::memcpy(volume + (block[1] - roi[1] + i) * width * width, slice.GetDataPtr(),
                                sizeof(float) * std::min(slice.Size(), width * width));
                }
            });
            clock.lap(4);
        }
    } catch (const std::exception &ex) {
        g_err = ex.what();
        return 2;
    }
    return 0;
}

extern "C" int mre_engine_set_references(mre_engine *e, const float *dark, const float *flat, size_t rows,
                                         size_t width, float dark_dose, float flat_dose)
{
    g_err.clear();
    if (!e) {
        g_err = "NULL argument";
        return 1;
    }
    auto set = reinterpret_cast<SetReferences_t>(dlsym(e->so, "mr_modules_set_references"));
    if (!set) {
        g_err = "the module library has no mr_modules_set_references";
        return 2;
    }
    int found = 0;
    for (size_t i = 0; i < e->pres.size(); ++i) {
        if (e->pre_names[i] != "FullLogNorm") continue;
        const int rc = set(e->pres[i], dark, flat, rows, width, dark_dose, flat_dose);
        if (rc != 0) {
            auto err = reinterpret_cast<const char *(*)(void)>(dlsym(e->so, "mr_modules_last_error"));
            g_err = std::string("SetReferenceImages failed: ") + (err ? err() : "");
            return 2;
        }
        ++found;
    }
    if (!found) {
        g_err = "no FullLogNorm module in the chain";
        return 2;
    }
    return 0;
}

extern "C" int mre_engine_set_doses(mre_engine *e, const float *dose, size_t n)
{
    g_err.clear();
    if (!e || (n && !dose)) {
        g_err = "NULL argument";
        return 1;
    }
    e->doses.assign(dose, dose + n);
    return 0;
}

extern "C" int mre_engine_phase_times(const mre_engine *e, double *ms, size_t n)
{
    if (!e || !ms) return 1;
    for (size_t i = 0; i < n && i < 5; ++i) ms[i] = e->phase_ms[i];
    return 0;
}

// Serialises GetModuleList of a module library: "name:key=value,key=value;..."
extern "C" int mre_module_list(const char *module_so, const char *application, char *buf, size_t len)
{
    g_err.clear();
    void *so = dlopen(module_so, RTLD_LAZY);
    if (!so) {
        g_err = std::string("Could not load module library: ") + dlerror();
        return 2;
    }
    auto f = reinterpret_cast<GetModuleList_t>(dlsym(so, "GetModuleList"));
    if (!f) {
        g_err = "Could not find symbol GetModuleList";
        dlclose(so);
        return 2;
    }
    std::map<std::string, std::map<std::string, std::string>> list;
    int rc = f(application, &list);
    std::string out;
    for (auto &m : list) {
        out += m.first + ":";
        bool first = true;
        for (auto &kv : m.second) {
            if (!first) out += ",";
            out += kv.first + "=" + kv.second;
            first = false;
        }
        out += ";";
    }
    dlclose(so);
    if (buf && len) {
        std::strncpy(buf, out.c_str(), len - 1);
        buf[len - 1] = '\0';
    }
    return rc;
}
