// wt2_internal.hpp -- shared declarations of libwt2 (not part of the C ABI).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>

#include "../../include/wt2.h"

namespace wt2 {

constexpr int NMIN = 3;
constexpr int NMAX = 10;
constexpr int MMAX_AXIS = 64;   // one element line per axis fits one wavefront
constexpr int SOMMERFELD_TAG = 7;   // face_kind of a Sommerfeld face (SOMMERFELD_BDRY_TAG, src/wave.jl:23)
constexpr int EXCISION_TAG = 8;     // excision / pure-outflow outer face: no SAT (src/wave_strong_rhs.jl:188-213)

// Epilogue of a pencil pass: what is done with the axis contribution c.
enum Epi : int {
    EPI_FIRST = 0,    // P  = c
    EPI_MID = 1,      // P += c
    EPI_OUT_RHS = 2,  // (du, dv) = (u_t, P + c)
    EPI_OUT_LAP = 3,  // out = P + c
    EPI_RK1 = 4,      // RK4 stage updates (Horner form) with k = (u_t, P + c)
    EPI_RK2 = 5,
    EPI_RK3 = 6,
    EPI_RK4 = 7,
};

// Precision variant of a pass kernel.
enum Variant : int {
    VAR_F64 = 0,      // compute, stage input, partial: double
    VAR_MIX_IN64 = 1, // compute float, stage input double, partial float
    VAR_MIX_IN32 = 2, // compute float, stage input float, partial float
};

// Per-axis 1D operator scalars in FP64 (converted to the compute type at
// launch); the matrix itself is the device table PassDesc::op64 / op32.
struct AxisOps {
    double kh;               // KO scale kos / h
    double koL, koR;         // KO interface SAT: -kos/(2 w_0 h), +kos/(2 w_{N-1} h)
    double tau, ih2;         // penalty, 1/h^2
    double gL, gR;           // Dirichlet values at the low / high outer face
    double winv0, winvN;     // 1/w_0, 1/w_{N-1}
    double hv, hiR;          // Sommerfeld faces: h and h/R (the drag's scales)
    int somL, somR;          // 1: the low / high outer face is a Sommerfeld face
};

// Even/odd operator table of the GLL differentiation matrix (see wt2_flat.hip).
int op_table_size(int N);
void build_op_table(int N, const double* G, const double* w, double* tab);
void build_op_table32(int N, const double* G, const double* w, float* tab, float* G32);

// Everything one pencil-pass launch needs (type-erased pointers).
struct PassDesc {
    int N, variant;
    int axis, periodic;
    int M[3];
    int nb;             // lines-block size: element planes per tile (KT / JT)
    int direct;         // 1: LDS-free kernel (strided segment loads through L2)
    int epi;
    int lap;            // add the SIPG Laplacian of u along this axis
    int ko;             // add eps mu^-5 D^6 u_t along this axis
    int needv;          // load u_t (KO input or kept for the final epilogue)
    int keepv;          // final epilogue: u' = u_t needs u_t after the compute
    int fold;           // RK epilogue: also write P = this axis's contribution of the next stage
    const void* Su; const void* Sv;   // stage input (variant input type)
    void* P;                          // partial sum of the u_t equation (variant partial type)
    void* So_u; void* So_v;           // next stage input (partial type)
    double* Yu; double* Yv;           // RK state
    float* Yfu; float* Yfv;           // mixed: FP32 mirror of Y (read at stages 1-3, rewritten at stage 4)
    double* Ou; double* Ov;           // outputs
    double dt;
    AxisOps ops;
    const void* op64; const void* op32;   // device operator tables
};

// Launch one pencil pass on stream s.  Returns hipError_t.
hipError_t launch_flat_pass(const PassDesc& p, hipStream_t s);

// Mixed contexts: the FP32 mirror (fu, fv) = float(u, v) of the RK state.
hipError_t launch_mirror32(const double* u, const double* v, float* fu, float* fv, int64_t n, hipStream_t s);

// The flat SIPG Laplacian out = L u in one launch (no KO term): a workgroup
// per block of elements, the element staged in LDS, the face neighbours'
// lines read through the caches.  variant: VAR_F64 or VAR_MIX_IN64.
struct Lap1Desc {
    int N, variant;
    int M[3], periodic[3];
    AxisOps ops[3];
    const double* u;
    const double* v;    // u_t at the Sommerfeld faces' nodes; nullptr = zero
    double* out;
    const void* op64; const void* op32;
};
hipError_t launch_flat_lap1(const Lap1Desc& d, hipStream_t s);

// The first-order (Phi, Pi) system (wt2_fo.hip).  launch_fo_rhs: the six
// launches of one right-hand side (seven with fuse = 0) (with grad_only: the three that fill D);
// launch_fo_energy: the energy of (Phi, Pi) from D; launch_fo_coef: the
// coefficient fields from bg10.
struct FoDesc {
    int N;
    int M[3], periodic[3];
    double h[3], w[16];
    const double* Phi; const double* Pi;
    const double* coef;          // 11 fields
    double* D;                   // 3 fields
    double* dPhi; double* dPi;
    const double* gPhi; const double* gPi;
    const void* op64;
    double koT;
    int kinds[6];
    double sigma;
    double* partial; double* Eout;
    int grad_only;               // phase-0 passes only, D alone (for the energy)
    int fuse;                    // the z pass of phase 0 does the pointwise step
};
hipError_t launch_fo_rhs(const FoDesc& d, hipStream_t s);
hipError_t launch_fo_energy(const FoDesc& d, hipStream_t s);
hipError_t launch_fo_coef(const double* bg10, double* coef11, int64_t npts, hipStream_t s);
// Classical RK4 stage update of (Phi, Pi) = (Y0, Y1); registers S, A, k (two
// fields each) at R + f * stride, f = 0..5.
hipError_t launch_fo_rk_update(double* Y0, double* Y1, double* R, int64_t stride, int64_t n, int stage, double dt,
                               hipStream_t s);

// A batch of pencil passes: one launch over every instance of a group (same N,
// precision variant, axis, kernel kind and epilogue), driven by a device table
// of per-instance descriptors and a packed (instance, tile) work list.
struct FlatBatchPlan {
    int N = 0, variant = 0, direct = 0, lean = 0, ntiles = 0, nthr = 0;
    size_t shmem = 0;
    void* tab = nullptr;     // device descriptors
    int2* work = nullptr;    // device work list
    size_t tab_cap = 0, work_cap = 0;
};
hipError_t flat_batch_build(const PassDesc* descs, int n, FlatBatchPlan& plan);
hipError_t flat_batch_launch(const FlatBatchPlan& plan, hipStream_t s);
void flat_batch_free(FlatBatchPlan& plan);

// Deterministic FP64 energy reduction: E = 1/2 sum H (v^2 - u Lu) over a flat
// or curved mesh.  partial must hold red_blocks doubles; result written to
// out (device).
constexpr int RED_BLOCKS = 512;
constexpr int RED_THREADS = 256;
struct EnergyDesc {
    int N; int64_t Ne;
    const double* u; const double* v; const double* Lu;
    const double* wv;   // per-node weight of v^2 (nullptr: same as the u weight)
    double hvol;        // u weight is w_i w_j w_k hvol
    double w[NMAX];
    double* partial; double* out;
};
hipError_t launch_energy(const EnergyDesc& e, hipStream_t s);

void set_error(const std::string& msg);

}  // namespace wt2
