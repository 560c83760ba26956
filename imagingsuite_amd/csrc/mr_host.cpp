// Host-side set-up of the ramp filters, the circle mask, the projection
// tables and the FDK projection matrices.  See mr_host.h for the reference
// functions each routine restates.
#include "mr_host.h"

#include <algorithm>
#include <cmath>
#include <string>

#include "mr_common.h"

namespace mr {

namespace {
const float kFPi = 3.1415926535897932384626f;
const double kDPi = 3.1415926535897932384626433832795028841971693993751;
}  // namespace

int build_pb_filter(size_t width, int filter_type, float cutoff, float order, bool use_bias,
                    float bias_weight, size_t padding_doubler, size_t &P, size_t &n_insert,
                    std::vector<float> &H, std::vector<float> &taper)
{
    const double e = std::log(static_cast<double>(width)) / std::log(2.0);
    P = static_cast<size_t>(1) << (static_cast<size_t>(std::ceil(e)) + padding_doubler);
    const size_t N2 = P / 2;
    if (!((0 < cutoff) && (cutoff <= 0.5f))) {
        set_error("Cut off frequency is not in valid range");
        return MR_ERR_INVALID;
    }
    const float filter_order = filter_type == MR_FILTER_BUTTERWORTH ? order : 0.0f;
    const size_t c = static_cast<size_t>(2 * cutoff * static_cast<float>(N2));
    H.assign(N2, 0.0f);
    const float rampstep = 1.0f / ((static_cast<float>(N2) - 1.0f));
    for (size_t i = 0; i < N2; ++i) H[i] = static_cast<float>(i) * rampstep;
    const float fstep = 1.0f / ((static_cast<float>(c) - 1.0f));
    for (size_t i = 0; i < c && i < N2; ++i) {
        const float fi = static_cast<float>(i);
        switch (filter_type) {
        case MR_FILTER_SHEPPLOGAN:
            H[i] = static_cast<float>(std::sin(kDPi * static_cast<double>(i) * fstep * cutoff)) /
                   static_cast<float>(N2);
            break;
        case MR_FILTER_HANNING:
            H[i] *= static_cast<float>(0.5f + 0.5f * std::cos(fi * fstep * kFPi * cutoff));
            break;
        case MR_FILTER_HAMMING:
            H[i] *= static_cast<float>(0.54 + 0.46 * std::cos(fi * fstep * kDPi * cutoff));
            break;
        case MR_FILTER_BUTTERWORTH:
            H[i] /= static_cast<float>(1 + std::pow(fi * fstep * cutoff, filter_order));
            break;
        default: break;
        }
    }
    for (size_t i = c; i < N2; ++i) H[i] = 0.0f;
    if (use_bias) H[0] = bias_weight * H[1];
    n_insert = (P / 2) - (width / 2);
    taper.resize(n_insert);
    for (size_t i = 0; i < n_insert; ++i) {
        const float x = static_cast<float>(i) / static_cast<float>(n_insert) - 0.5f;
        taper[i] = static_cast<float>(1.0 / (1.0 + std::exp(-(x - 0.0f) / 0.07f)));
    }
    return MR_OK;
}

void circle_mask(size_t nx, size_t ny, bool use_roi, const size_t *roi, bool cone,
                 std::vector<size_t> &first, std::vector<size_t> &second, float shrink)
{
    const float cx = static_cast<float>(nx >> 1);
    const float cy = static_cast<float>(ny >> 1);
    float R = cx - 1;
    R -= shrink;
    const float R2 = R * R;
    first.assign(ny, 0);
    second.assign(ny, 0);
    for (size_t i = 0; i < ny; ++i) {
        const float y = static_cast<float>(i) - cy;
        const float y2 = y * y;
        if (y2 <= R2) {
            const float x = std::sqrt(R2 - y2);
            first[i] = static_cast<size_t>(std::ceil(cx - x));
            second[i] = std::min(nx - 1u, static_cast<size_t>(std::floor(cx + x)));
        }
        if (use_roi) {
            if (roi[1] <= i && i <= roi[3]) {
                if (!cone) {
                    first[i] = std::max(first[i], roi[0]);
                    second[i] = std::min(second[i], roi[2]);
                } else {
                    first[i] = std::max(first[i], nx - roi[2]);
                    second[i] = std::min(second[i], nx - roi[0]);
                }
            } else {
                first[i] = 0;
                second[i] = 0;
            }
        }
    }
}

float tilt_mask_shrink(float tilt_angle, size_t rows)
{
    return std::floor(std::tan(std::fabs(tilt_angle) * kFPi / 180.0f) * static_cast<float>(rows));
}

void bp_tilt_terms(float tilt_angle, float tilt_pivot, size_t row0, float &centeroffset, float &centerinc)
{
    const float t = std::tan(tilt_angle * kFPi / 180.0f);
    centeroffset = (static_cast<float>(row0) - tilt_pivot) * t;
    centerinc = 4.0f * t;
}

std::vector<size_t> used_projections(size_t n_proj, size_t buffer_size)
{
    std::vector<size_t> used;
    used.reserve(n_proj);
    if (buffer_size == 0) {
        for (size_t i = 0; i < n_proj; ++i) used.push_back(i);
        return used;
    }
    // Slot buffer_size-1 of every full chunk is discarded before BackProject.
    for (size_t i = 0; i < n_proj; ++i)
        if (i % buffer_size != buffer_size - 1) used.push_back(i);
    return used;
}

void bp_tables(const float *angles, const size_t *idx, size_t n, int direction, float center,
               size_t matrix_center_x, std::vector<float> &tab4)
{
    tab4.assign(4 * n, 0.0f);
    const float dir_weight = 2.0f * (static_cast<float>(direction) - 0.5f);
    for (size_t k = 0; k < n; ++k) {
        const float angle = angles[idx ? idx[k] : k];
        const float s = std::sin(dir_weight * angle * kFPi / 180.0f);
        const float c = std::cos(dir_weight * angle * kFPi / 180.0f);
        tab4[4 * k + 0] = static_cast<float>(matrix_center_x) * (s - c) + center;
        tab4[4 * k + 1] = c;
        tab4[4 * k + 2] = s;
    }
}

void fdk_proj_matrix(const FdkGeometry &g, float angle, float pm[12])
{
    float ext[16] = {0}, intr[12] = {0};
    const float sad = g.sod, sid = g.sdd;
    const float off_center = (g.ppoint[0] - static_cast<float>(g.roi[0]) - g.center) * g.voxel[0];
    float cam[3] = {0.0f, 0.0f, 0.0f};
    if (g.direction == 0) {
        cam[0] = sad * std::cos(-angle * kFPi / 180.0f);
        cam[1] = sad * std::sin(-angle * kFPi / 180.0f);
    } else {
        cam[0] = sad * std::cos(angle * kFPi / 180.0f);
        cam[1] = sad * std::sin(angle * kFPi / 180.0f);
    }
    const float vup[3] = {0.0f, 0.0f, 1.0f}, tgt[3] = {0.0f, 0.0f, 0.0f};
    float nrm[3], plt[3], pup[3];
    for (int i = 0; i < 3; ++i) nrm[i] = cam[i] - tgt[i];
    const float nn = std::sqrt(nrm[0] * nrm[0] + nrm[1] * nrm[1] + nrm[2] * nrm[2]);
    for (int i = 0; i < 3; ++i) nrm[i] /= nn;
    plt[0] = nrm[1] * vup[2] - nrm[2] * vup[1];
    plt[1] = nrm[2] * vup[0] - nrm[0] * vup[2];
    plt[2] = nrm[0] * vup[1] - nrm[1] * vup[0];
    const float np_ = std::sqrt(plt[0] * plt[0] + plt[1] * plt[1] + plt[2] * plt[2]);
    for (int i = 0; i < 3; ++i) plt[i] /= np_;
    pup[0] = plt[1] * nrm[2] - plt[2] * nrm[1];
    pup[1] = plt[2] * nrm[0] - plt[0] * nrm[2];
    pup[2] = plt[0] * nrm[1] - plt[1] * nrm[0];
    const float nu_ = std::sqrt(pup[0] * pup[0] + pup[1] * pup[1] + pup[2] * pup[2]);
    for (int i = 0; i < 3; ++i) pup[i] /= nu_;
    for (int i = 0; i < 3; ++i) {
        ext[i] = -1.0f * plt[i];
        ext[4 + i] = -1.0f * pup[i];
        ext[8 + i] = -1.0f * nrm[i];
    }
    ext[15] = 1.0f;
    ext[3] = plt[0] * tgt[0] + plt[1] * tgt[1] + plt[2] * tgt[2] - off_center;
    ext[7] = pup[0] * tgt[0] + pup[1] * tgt[1] + pup[2] * tgt[2];
    ext[11] = nrm[0] * tgt[0] + nrm[1] * tgt[1] + nrm[2] * tgt[2] + sad;
    intr[0] = 1.0f / g.pixel[0];
    intr[5] = 1.0f / g.pixel[1];
    intr[10] = 1.0f / sid;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 4; ++j) {
            float acc = 0.0f;
            for (int k = 0; k < 4; ++k) acc += intr[i * 4 + k] * ext[k * 4 + j];
            pm[i * 4 + j] = acc;
        }
}

void fdkd_proj_matrix(const FdkGeometry &g, float angle, double pm[12])
{
    double ext[16] = {0}, intr[12] = {0};
    const double sad = g.sod, sid = g.sdd;
    // a float expression in the reference (float operands), then widened
    const double off_center = (g.ppoint[0] - static_cast<float>(g.roi[0]) - g.center) * g.voxel[0];
    constexpr double kDPi = 3.14159265358979323846264338327950288;
    double cam[3] = {0.0, 0.0, 0.0};
    if (g.direction == 0) {
        cam[0] = sad * std::cos(-angle * kDPi / 180);
        cam[1] = sad * std::sin(-angle * kDPi / 180);
    } else {
        cam[0] = sad * std::cos(angle * kDPi / 180);
        cam[1] = sad * std::sin(angle * kDPi / 180);
    }
    const double vup[3] = {0.0, 0.0, 1.0}, tgt[3] = {0.0, 0.0, 0.0};
    double nrm[3], plt[3], pup[3];
    for (int i = 0; i < 3; ++i) nrm[i] = cam[i] - tgt[i];
    const double nn = std::sqrt(nrm[0] * nrm[0] + nrm[1] * nrm[1] + nrm[2] * nrm[2]);
    for (int i = 0; i < 3; ++i) nrm[i] /= nn;
    plt[0] = nrm[1] * vup[2] - nrm[2] * vup[1];
    plt[1] = nrm[2] * vup[0] - nrm[0] * vup[2];
    plt[2] = nrm[0] * vup[1] - nrm[1] * vup[0];
    const double np_ = std::sqrt(plt[0] * plt[0] + plt[1] * plt[1] + plt[2] * plt[2]);
    for (int i = 0; i < 3; ++i) plt[i] /= np_;
    pup[0] = plt[1] * nrm[2] - plt[2] * nrm[1];
    pup[1] = plt[2] * nrm[0] - plt[0] * nrm[2];
    pup[2] = plt[0] * nrm[1] - plt[1] * nrm[0];
    const double nu_ = std::sqrt(pup[0] * pup[0] + pup[1] * pup[1] + pup[2] * pup[2]);
    for (int i = 0; i < 3; ++i) pup[i] /= nu_;
    for (int i = 0; i < 3; ++i) {
        ext[i] = -1.0 * plt[i];
        ext[4 + i] = -1.0 * pup[i];
        ext[8 + i] = -1.0 * nrm[i];
    }
    ext[15] = 1.0;
    ext[3] = plt[0] * tgt[0] + plt[1] * tgt[1] + plt[2] * tgt[2] - off_center;
    ext[7] = pup[0] * tgt[0] + pup[1] * tgt[1] + pup[2] * tgt[2];
    ext[11] = nrm[0] * tgt[0] + nrm[1] * tgt[1] + nrm[2] * tgt[2] + sad;
    intr[0] = 1 / g.pixel[0];   // int / float: a float quotient
    intr[5] = 1 / g.pixel[1];
    intr[10] = 1 / sid;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 4; ++j) {
            float acc = 0.0f;
            for (int k = 0; k < 4; ++k) acc = static_cast<float>(acc + intr[i * 4 + k] * ext[k * 4 + j]);
            pm[i * 4 + j] = acc;
        }
}

void fdk_cbct_roi(const FdkGeometry &g, size_t nvol, size_t cb[4], bool fdkbp_variant)
{
    const float radius = static_cast<float>(nvol) * g.voxel[0] / 2;
    const float pp1 = g.ppoint[1];
    const float r1 = static_cast<float>(g.roi[1]), r3 = static_cast<float>(g.roi[3]);
    const float vs = g.voxel[0], sdd = g.sdd, sod = g.sod, res = g.pixel[0];
    cb[0] = g.roi[0];
    cb[2] = g.roi[2];
    cb[1] = 0;
    cb[3] = 0;
    if (pp1 >= r1 && pp1 >= r3) {
        cb[3] = static_cast<size_t>(pp1 - ((pp1 - r3) * vs * sdd / (sod + radius)) / res);
        const float value = pp1 - ((pp1 - r1) * vs * sdd / (sod - radius)) / res;
        cb[1] = value <= 0 ? 0 : static_cast<size_t>(pp1 - ((pp1 - r1) * vs * sdd / (sod - radius)) / res);
    }
    if (pp1 < r1 && pp1 < r3) {
        const float value = pp1 + ((r1 - pp1) * vs * sdd / (sod + radius)) / res;
        cb[1] = static_cast<size_t>(value);
        const float value2 = pp1 + ((r3 - pp1) * vs * sdd / (sod - radius)) / res;
        cb[3] = value2 >= static_cast<float>(g.projection_roi3) ? g.projection_roi3
                                                                 : static_cast<size_t>(value2);
    }
    if (pp1 >= r1 && pp1 < r3) {
        const float value = pp1 - ((pp1 - r1) * vs * sdd / (sod - radius)) / res;
        cb[1] = value <= 0 ? 0 : static_cast<size_t>(pp1 - ((pp1 - r1) * vs * sdd / (sod - radius)) / res);
        const float value2 = pp1 + ((r3 - pp1) * vs * sdd / (sod - radius)) / res;
        cb[3] = value2 >= static_cast<float>(g.projection_roi3) ? g.projection_roi3
                                                                 : static_cast<size_t>(value2);
    }
    if (fdkbp_variant || cb[1] != 0) cb[1] -= 8;
    if (cb[3] + 8 <= g.projection_roi3) cb[3] += 8;
}

size_t fdk_padwidth(size_t width)
{
    const unsigned int w = static_cast<unsigned int>(width);
    return static_cast<unsigned int>(2 * std::pow(2, std::ceil(std::log(w) / std::log(2))));
}

}  // namespace mr
