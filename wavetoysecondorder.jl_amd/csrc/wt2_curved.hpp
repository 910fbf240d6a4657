// wt2_curved.hpp -- curved (warped mesh and/or ADM background) operator of
// libwt2: static FP64 geometry/background coefficients and the element
// kernels that apply the variable-coefficient SIPG operator plus the u_t
// advection terms.  KO on curved meshes reuses the flat pencil passes.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

#include "../../include/wt2.h"

namespace wt2 {

struct CurvedData {
    int N = 0;
    int M[3] = {0, 0, 0};
    int periodic[3] = {0, 0, 0};
    int64_t Ne = 0, npts = 0;
    double tau = 0.0;
    int som = 0;                // bit f: outer face f (0..5) is a Sommerfeld face
    double inv_R = 0.0;         // 1/R of the Sommerfeld faces' u/R term
    double w[16] = {0};
    double G[256] = {0};
    float G32[256] = {0};       // FP32 G, the same matrix as the FP32 even/odd table
    double h_min = 0.0;
    double* coef64 = nullptr;   // 13 per node, SoA: coef64[q*npts + n]
    float* coef32 = nullptr;    // same, FP32 copy (mixed precision)
    double* wu = nullptr;       // energy weight of u*Lu: H3 (per node)
    double* wv = nullptr;       // energy weight of v^2: H3 / wgt
    double* fbuf = nullptr;     // face flux buffer: 6 N^2 per element
    // face-node coefficients of the face terms, 5 x 6 N^2 per element in the
    // order of the face loop (K^{a0}, K^{a1}, K^{a2}, kappa, c^a; a = the
    // face's axis), FP64 and FP32
    double* fc64 = nullptr;
    float* fc32 = nullptr;
    bool pending = false;       // allocated; coefficients not computed yet
    const void* op64 = nullptr; // even/odd operator tables (owned by the context)
    const void* op32 = nullptr;
    // general connectivity (wt2_ctx_create_mesh); empty / nullptr on uniform meshes
    std::vector<double> xyz;    // node coordinates (3,N,N,N,Ne)
    int2* conn = nullptr;       // per (e, f): (neighbour or -1, fn | ori << 3 | tag << 6)
};

struct CurvedCall {
    const void* u;      // stage u (double, or float when in32)
    const void* v;      // stage u_t; nullptr: no advection term
    int in32;           // stage input is FP32
    int mixed;          // FP32 arithmetic
    int weighted;       // 1: multiply by wgt (RHS); 0: unweighted Ltilde u (energy)
    double bv[6];       // Dirichlet values per tag
    double* out;        // FP64 output (drop-in / energy), or nullptr
    void* Pv;           // partial sum of the u_t equation (FP64 or FP32), written when out == nullptr
};

// Batched curved operator: one (face-flux, volume) launch pair over every
// element of every instance in the group (same N and precision).
struct CurvedBatchPlan {
    int N = 0, kind = 0, nwork = 0, nwork2 = 0;
    int gen = 0;                // some instance has a connectivity table
    void* tab = nullptr;
    int2* work = nullptr;       // (instance, element): face-flux kernel
    int2* work2 = nullptr;      // (instance, first element of a block): element kernel
    size_t tab_cap = 0, work_cap = 0, work2_cap = 0;
};
hipError_t curved_batch_build(const CurvedData* const* cds, const CurvedCall* ccs, int n, CurvedBatchPlan& plan);
hipError_t curved_batch_launch(const CurvedBatchPlan& plan, hipStream_t s);
void curved_batch_free(CurvedBatchPlan& plan);

int curved_init(CurvedData& cd, const wt2_mesh_desc& d, const double* xi, const double* w,
                const double* G, const double* bg10, hipStream_t s, int64_t* bytes);
int curved_set_background(CurvedData& cd, const wt2_mesh_desc& d, const double* xi,
                          const double* w, const double* G, const double* bg10, hipStream_t s);
void curved_free(CurvedData& cd);
// The same for a mesh given by node coordinates xyz (3,N,N,N,Ne) and the
// per-face connectivity table conn (Ne*6 entries, the kernels' encoding).
int curved_init_mesh(CurvedData& cd, const wt2_mesh_desc& d, const double* xi, const double* w,
                     const double* G, const double* xyz, const std::vector<int2>& conn, hipStream_t s,
                     int64_t* bytes);
hipError_t curved_operator(const CurvedData& cd, const CurvedCall& cc, hipStream_t s);

// Node coordinates (3,N,N,N,Ne) of the (possibly warped) uniform hex mesh.
void mesh_coords(const wt2_mesh_desc& d, const double* xi, double* xyz);

}  // namespace wt2
