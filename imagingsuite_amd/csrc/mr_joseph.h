// Internal launch interface of the matched Joseph pair (mr_joseph.hip), used
// by the iterative solvers to run A and A^T with resident tables.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <vector>

#include "muhrec_hip.h"

namespace mr {

// Per angle (su, cos, sin, 1/m) as four floats.
std::vector<float> joseph_tables(const mr_joseph_geom &g, const float *angles, size_t n_ang);

// Host tables of one angle set: the per-angle entries, the forward's angle
// chunks (first, count, stepping class, 0) and the angle order they index.
struct JosephPacked {
    std::vector<float> tab;
    std::vector<int4> chunks;
    std::vector<int> order;
    size_t bytes() const;
};
struct JosephDevTables {
    const float4 *tab = nullptr;
    const int4 *chunks = nullptr;
    const int *order = nullptr;
    int n_chunks = 0;
};
JosephPacked joseph_pack(const mr_joseph_geom &g, const float *angles, size_t n_ang);
// Copies the packed tables into d_base (pk.bytes() bytes) on s; the host
// tables must outlive the copies.
int joseph_upload(const JosephPacked &pk, void *d_base, hipStream_t s, JosephDevTables &dt);
int joseph_check(const mr_joseph_geom *g, size_t n_ang);
// Enqueues A (forward: volume -> sinogram) or A^T on s with device tables.
// Does not synchronise.
int joseph_launch(const mr_joseph_geom *g, const float *src, float *dst, const JosephDevTables &dt, size_t n_ang,
                  bool forward, hipStream_t s);

}  // namespace mr
