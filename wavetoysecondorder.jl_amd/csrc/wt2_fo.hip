// wt2_fo.hip -- the first-order (Phi, Pi) ADM right-hand side of
// wave3d_curved_rhs! (src/wave3d_curved.jl:205-258) on axis-aligned affine
// meshes: the per-axis SAT-coupled derivative apply_D! (one-sided at outer
// faces), the flux F^i = b^i Pi + alpha sqrtg g^{ij} DPhi_j and its divergence,
// eps mu^-5 D_d^6 on both fields (_ko_axis3!, :182-192), and the affine
// outer-boundary pass (_bc3d_affine_kernel!, src/boundaries3d.jl:76-126).
//
// Six launches per call.  Pencil pass (axis d, phase 0): DPhi_d = D_d Phi,
// kept in a context buffer, and the KO terms of both fields along d; the last
// of them (z) then does the pointwise step on its own nodes: the fluxes F^i
// over that buffer (DPhi_3 still in registers), b . DPhi + (alpha/sqrtg) Pi
// into dPhi and the boundary terms.  Pencil pass (axis d, phase 1):
// dPi += D_d F^d.  A pass workgroup owns whole element
// lines along its axis (LT lines x Ma elements, work-item = (element, line)
// with adjacent lines in adjacent lanes), and the face values of neighbouring
// elements go through LDS, as in the second-order pencil passes.
#include "wt2_internal.hpp"
#include "wt2_ops.hpp"

#include <algorithm>
#include <cstdlib>

namespace wt2 {

namespace {

constexpr int FO_MAXT = 512;

struct FoArgs {
    const double* Phi; const double* Pi;
    const double* coef;      // 11 fields of npts (alpha, sqrtg, b1..3, gu11, gu12, gu13, gu22, gu23, gu33)
    double* D;               // 3 fields: DPhi_1..3
    double* dPhi; double* dPi;
    const double* gPhi; const double* gPi;
    const double* op;        // OpLayout<N> table
    int64_t npts;
    int M0, M1, M2;
    int per0, per1, per2;
    int axis, first, anybc;
    int fuse;                // z pass of phase 0 also does the pointwise flux step
    int Ma, T0, slo, shi, estride, LT, nbk;
    double ih, sat0, satN;   // 1/h; SAT -1/(2 w_0 h), +1/(2 w_{N-1} h) along the axis
    double koT;              // eps mu^-5 (0: no KO)
    int kinds[6];
    double sigma;
    double bw[3][2];         // boundary weight 1/(h_d w_face) per axis, low / high face
};

// Face values of the line's neighbouring elements through LDS (two
// alternating buffers; every work-item of the workgroup makes each call).
struct FoX {
    double* buf;
    int me, idxL, idxR, k;
    __device__ __forceinline__ void operator()(double last, double first, double& fromL, double& fromR)
    {
        double* b = buf + (k & 1) * 2 * FO_MAXT;
        b[me] = last;
        b[FO_MAXT + me] = first;
        __syncthreads();
        fromL = b[idxL];
        fromR = b[FO_MAXT + idxR];
        ++k;
    }
};

// y = apply_D!(x) on this segment: (1/h) G x plus the centred-flux SAT at
// interior faces only (one-sided at outer faces, METHODS.md:256-264).
template <int N>
__device__ __forceinline__ void fo_D(const double* __restrict__ op, const FoArgs& a, FoX& xch, bool hasL, bool hasR,
                                     const double (&x)[N], double (&y)[N])
{
    double xl, xr;
    xch(x[N - 1], x[0], xl, xr);
    double g[N];
    apply_G<N, double>(op, x, g);
#pragma unroll
    for (int i = 0; i < N; ++i) y[i] = g[i] * a.ih;
    if (hasL) y[0] = fma(a.sat0, xl - x[0], y[0]);
    if (hasR) y[N - 1] = fma(a.satN, xr - x[N - 1], y[N - 1]);
}

// N values of a segment at node stride GS (compile-time: contiguous x-line
// segments load as 16-byte vectors).
template <int N, int GS>
__device__ __forceinline__ void fo_load(const double* __restrict__ p, bool on, double (&x)[N])
{
    if (!on) {
#pragma unroll
        for (int r = 0; r < N; ++r) x[r] = 0.0;
        return;
    }
    if constexpr (GS == 1 && N % 2 == 0) {
#pragma unroll
        for (int r = 0; r < N; r += 2) {
            const double2 t = *reinterpret_cast<const double2*>(p + r);
            x[r] = t.x;
            x[r + 1] = t.y;
        }
    } else {
#pragma unroll
        for (int r = 0; r < N; ++r) x[r] = p[r * GS];
    }
}

template <int N, int GS>
__device__ __forceinline__ void fo_store(double* __restrict__ p, bool on, const double (&x)[N])
{
    if (!on) return;
    if constexpr (GS == 1 && N % 2 == 0) {
#pragma unroll
        for (int r = 0; r < N; r += 2) *reinterpret_cast<double2*>(p + r) = make_double2(x[r], x[r + 1]);
    } else {
#pragma unroll
        for (int r = 0; r < N; ++r) p[r * GS] = x[r];
    }
}

// The pointwise step at node q (element e, node (i, j, k)): F^i = b^i Pi +
// alpha sqrtg g^{ij} DPhi_j, dF += b . DPhi + (alpha / sqrtg) Pi, and the
// affine outer-boundary terms (src/boundaries3d.jl:76-126) into dF and dP.
// ec: the element's coordinates (ex, ey, ez) in the element grid.
template <int N>
__device__ __forceinline__ void fo_point(const FoArgs& a, int64_t q, const int (&ec)[3], int i, int j, int k,
                                         double D1, double D2, double D3, double Pq,
                                         double& dF, double& dP, double& F1, double& F2, double& F3)
{
    const int64_t n = a.npts;
    const double* __restrict__ c = a.coef;
    const double al = c[q], sg = c[n + q];
    const double b1 = c[2 * n + q], b2 = c[3 * n + q], b3 = c[4 * n + q];
    const double g11 = c[5 * n + q], g12 = c[6 * n + q], g13 = c[7 * n + q];
    const double g22 = c[8 * n + q], g23 = c[9 * n + q], g33 = c[10 * n + q];
    const double as = al * sg;
    F1 = b1 * Pq + as * (g11 * D1 + g12 * D2 + g13 * D3);
    F2 = b2 * Pq + as * (g12 * D1 + g22 * D2 + g23 * D3);
    F3 = b3 * Pq + as * (g13 * D1 + g23 * D2 + g33 * D3);
    dF = dF + (b1 * D1 + b2 * D2 + b3 * D3 + (al / sg) * Pq);
    if (!a.anybc) return;
    const int nd[3] = {i, j, k};
    const int Mv[3] = {a.M0, a.M1, a.M2};
    const int pv[3] = {a.per0, a.per1, a.per2};
    const double Dv[3] = {D1, D2, D3};
    const double bv[3] = {b1, b2, b3};
    const double gnn[3] = {g11, g22, g33};
    for (int f = 0; f < 6; ++f) {
        const int dn = f >> 1, hi = f & 1;
        if (pv[dn] || ec[dn] != (hi ? Mv[dn] - 1 : 0) || nd[dn] != (hi ? N - 1 : 0)) continue;
        const int kind = a.kinds[f];
        if (kind == WT2_BC_EXCISION) continue;
        const double nh = hi ? 1.0 : -1.0;
        const double a_n = al * sqrt(gnn[dn]);
        const double bax = bv[dn];
        const double bn = nh * bax;
        const double wt = a.bw[dn][hi];
        if (kind == WT2_BC_FULL_DIRICHLET) {
            const double tau = a.sigma * (fabs(a_n - bax) + fabs(a_n + bax)) * wt;
            const double gF = a.gPhi ? a.gPhi[q] : 0.0, gP = a.gPi ? a.gPi[q] : 0.0;
            dF += -tau * (a.Phi[q] - gF);
            dP += -tau * (Pq - gP);
        } else {
            const double rr = Pq + ((bn + a_n) / (al / sg)) * (nh * Dv[dn]);
            const double gv = (kind == WT2_BC_SOMMERFELD || !a.gPhi) ? 0.0 : a.gPhi[q];
            dP += -a.sigma * fabs(a_n + bn) * wt * (rr - gv);
        }
    }
}

// One pencil pass along axis AX.  PHASE 0: D[AX] = D_AX Phi and the KO terms
// of both fields along AX into dPhi, dPi (the first pass writes them).
// PHASE 1: dPi += D_AX F^AX, F^AX held in D[AX] (written by fo_flux_kernel).
template <int N, int AX, int PHASE>
__global__ void __launch_bounds__(FO_MAXT) fo_pass_kernel(FoArgs a)
{
    constexpr int N3 = N * N * N;
    constexpr int GS = AX == 0 ? 1 : (AX == 1 ? N : N * N);
    constexpr int LA = AX == 2 ? N : N * N;          // line -> (line / N) * LA + (line % N) * LB
    constexpr int LB = AX == 0 ? N : 1;
    __shared__ double op[OpLayout<N>::SIZE];
    __shared__ double xbuf[4 * FO_MAXT];
    for (int q = threadIdx.x; q < OpLayout<N>::SIZE; q += blockDim.x) op[q] = a.op[q];
    __syncthreads();

    int t = blockIdx.x;
    const int blk = t % a.nbk;
    t /= a.nbk;
    const int ebase = (t % a.T0) * a.slo + (t / a.T0) * a.shi;
    const int tid = threadIdx.x;
    const int m = tid / a.LT;
    const int lane_line = tid - m * a.LT;
    const int line = blk * a.LT + lane_line;
    const int Ma = a.Ma;
    const bool active = m < Ma && line < N * N;
    const int per = AX == 0 ? a.per0 : (AX == 1 ? a.per1 : a.per2);
    const bool hasL = active && (m > 0 || per);
    const bool hasR = active && (m < Ma - 1 || per);
    FoX xch;
    xch.buf = xbuf; xch.me = tid; xch.k = 0;
    xch.idxL = hasL ? (m > 0 ? m - 1 : Ma - 1) * a.LT + lane_line : tid;
    xch.idxR = hasR ? (m < Ma - 1 ? m + 1 : 0) * a.LT + lane_line : tid;
    const int mm = active ? m : 0, ll = active ? line : 0;
    const int e = ebase + mm * a.estride;
    const int64_t g0 = int64_t(e) * N3 + (ll / N) * LA + (ll % N) * LB;
    const int64_t n = a.npts;

    if constexpr (PHASE == 0 && AX == 2) {
        if (a.fuse) {
            // DPhi_3 stays in registers; KO sums are kept for the pointwise
            // step.  Pi is read twice (for its KO chain and per node in the
            // pointwise step) rather than held, and the pointwise step runs
            // two nodes at a time: together they hold the pass to three
            // wavefronts per SIMD (161 VGPRs at N = 8, against 219 and two).
            // Forcing four (128 VGPRs) spills and measured 14 % slower.
            double x[N], d3[N], acc[2][N];
            fo_load<N, GS>(a.Phi + g0, active, x);
            fo_D<N>(op, a, xch, hasL, hasR, x, d3);
            for (int fld = 0; fld < 2; ++fld) {
                if (a.koT != 0.0) {
                    double y[N];
                    if (fld == 1) fo_load<N, GS>(a.Pi + g0, active, x);
#pragma unroll 1
                    for (int rep = 0; rep < 3; ++rep) {
                        fo_D<N>(op, a, xch, hasL, hasR, x, y);
                        fo_D<N>(op, a, xch, hasL, hasR, y, x);
                    }
                    fo_load<N, GS>((fld == 0 ? a.dPhi : a.dPi) + g0, active, acc[fld]);
#pragma unroll
                    for (int r = 0; r < N; ++r) acc[fld][r] = fma(a.koT, x[r], acc[fld][r]);
                } else {
                    fo_load<N, GS>((fld == 0 ? a.dPhi : a.dPi) + g0, active, acc[fld]);
                }
            }
            if (!active) return;
            const int i = ll % N, j = ll / N;
            // the element's grid coordinates, once per work-item (32-bit: the
            // element count fits an int)
            const int exy = e % (a.M0 * a.M1);
            const int ec[3] = {exy % a.M0, exy / a.M0, e / (a.M0 * a.M1)};
#pragma unroll 2
            for (int r = 0; r < N; ++r) {
                const int64_t q = g0 + int64_t(r) * GS;
                double dP = 0.0, F1, F2, F3;
                fo_point<N>(a, q, ec, i, j, r, a.D[q], a.D[n + q], d3[r], a.Pi[q], acc[0][r], dP, F1, F2, F3);
                a.D[q] = F1; a.D[n + q] = F2; a.D[2 * n + q] = F3;
                if (dP != 0.0) acc[1][r] += dP;
            }
            fo_store<N, GS>(a.dPhi + g0, true, acc[0]);
            fo_store<N, GS>(a.dPi + g0, true, acc[1]);
            return;
        }
    }
    if constexpr (PHASE == 0) {
        double x[N], y[N];
        fo_load<N, GS>(a.Phi + g0, active, x);
        fo_D<N>(op, a, xch, hasL, hasR, x, y);
        fo_store<N, GS>(a.D + AX * n + g0, active, y);
        for (int fld = 0; fld < 2; ++fld) {
            double* out = (fld == 0 ? a.dPhi : a.dPi) + g0;
            if (a.koT != 0.0) {
                if (fld == 1) fo_load<N, GS>(a.Pi + g0, active, x);
#pragma unroll 1
                for (int rep = 0; rep < 3; ++rep) {
                    fo_D<N>(op, a, xch, hasL, hasR, x, y);
                    fo_D<N>(op, a, xch, hasL, hasR, y, x);
                }
                double acc[N];
                if (a.first) {
#pragma unroll
                    for (int r = 0; r < N; ++r) acc[r] = a.koT * x[r];
                } else {
                    fo_load<N, GS>(out, active, acc);
#pragma unroll
                    for (int r = 0; r < N; ++r) acc[r] = fma(a.koT, x[r], acc[r]);
                }
                fo_store<N, GS>(out, active, acc);
            } else if (a.first) {
                double z[N];
#pragma unroll
                for (int r = 0; r < N; ++r) z[r] = 0.0;
                fo_store<N, GS>(out, active, z);
            }
        }
    } else {
        double F[N], DF[N], acc[N];
        fo_load<N, GS>(a.D + AX * n + g0, active, F);
        fo_D<N>(op, a, xch, hasL, hasR, F, DF);
        fo_load<N, GS>(a.dPi + g0, active, acc);
#pragma unroll
        for (int r = 0; r < N; ++r) acc[r] += DF[r];
        fo_store<N, GS>(a.dPi + g0, active, acc);
    }
}

// The pointwise step as a kernel of its own, between the two phases (the
// unfused schedule, WT2_FO_FUSE=0): each work-item reads and writes only its
// own node.
template <int N>
__global__ void __launch_bounds__(256) fo_flux_kernel(FoArgs a)
{
    constexpr int N3 = N * N * N;
    const int64_t n = a.npts;
    for (int64_t q = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; q < n; q += int64_t(gridDim.x) * blockDim.x) {
        const int64_t e = q / N3;
        const int r = int(q - e * N3);
        const int ei = int(e), exy = ei % (a.M0 * a.M1);
        const int ec[3] = {exy % a.M0, exy / a.M0, ei / (a.M0 * a.M1)};
        double dF = a.dPhi[q], dP = 0.0, F1, F2, F3;
        fo_point<N>(a, q, ec, r % N, (r / N) % N, r / (N * N), a.D[q], a.D[n + q], a.D[2 * n + q], a.Pi[q],
                    dF, dP, F1, F2, F3);
        a.D[q] = F1; a.D[n + q] = F2; a.D[2 * n + q] = F3;
        a.dPhi[q] = dF;
        if (dP != 0.0) a.dPi[q] += dP;
    }
}

template <int N, int AX, int PHASE>
hipError_t fo_launch_pass(FoArgs& a, hipStream_t s)
{
    const int Mx = a.M0, My = a.M1, Mz = a.M2;
    int rows;
    if (AX == 0) {        // rows (ey, ez), line (j, k): base N j + N^2 k
        a.Ma = Mx; a.T0 = My; a.slo = Mx; a.shi = Mx * My; a.estride = 1; rows = My * Mz;
    } else if (AX == 1) { // rows (ex, ez), line (i, k): base i + N^2 k
        a.Ma = My; a.T0 = Mx; a.slo = 1; a.shi = Mx * My; a.estride = Mx; rows = Mx * Mz;
    } else {              // rows (ex, ey), line (i, j): base i + N j
        a.Ma = Mz; a.T0 = Mx; a.slo = 1; a.shi = Mx; a.estride = Mx * My; rows = Mx * My;
    }
    a.LT = std::min(N * N, FO_MAXT / a.Ma);
    a.nbk = (N * N + a.LT - 1) / a.LT;
    const int nthr = (a.LT * a.Ma + 63) / 64 * 64;
    if (a.Ma > 64 || a.LT < 1 || nthr > FO_MAXT) return hipErrorInvalidConfiguration;
    (void)hipGetLastError();
    hipLaunchKernelGGL((fo_pass_kernel<N, AX, PHASE>), dim3(rows * a.nbk), dim3(nthr), 0, s, a);
    return hipGetLastError();
}

template <int N>
hipError_t launch_fo_n(const FoDesc& fd, hipStream_t s)
{
    FoArgs a;
    a.Phi = fd.Phi; a.Pi = fd.Pi; a.coef = fd.coef; a.D = fd.D;
    a.dPhi = fd.dPhi; a.dPi = fd.dPi; a.gPhi = fd.gPhi; a.gPi = fd.gPi;
    a.op = static_cast<const double*>(fd.op64);
    a.M0 = fd.M[0]; a.M1 = fd.M[1]; a.M2 = fd.M[2];
    a.per0 = fd.periodic[0]; a.per1 = fd.periodic[1]; a.per2 = fd.periodic[2];
    a.npts = int64_t(N) * N * N * fd.M[0] * fd.M[1] * fd.M[2];
    a.koT = fd.grad_only ? 0.0 : fd.koT;
    a.anybc = 0;
    for (int f = 0; f < 6; ++f) {
        a.kinds[f] = fd.kinds[f];
        if (fd.kinds[f] != WT2_BC_EXCISION && !fd.periodic[f / 2]) a.anybc = 1;
    }
    a.sigma = fd.sigma;
    for (int dd = 0; dd < 3; ++dd) {
        a.bw[dd][0] = 1.0 / (fd.h[dd] * fd.w[0]);
        a.bw[dd][1] = 1.0 / (fd.h[dd] * fd.w[N - 1]);
    }
    auto axis_consts = [&](int d) {
        a.axis = d;
        a.ih = 1.0 / fd.h[d];
        a.sat0 = -1.0 / (2.0 * fd.w[0] * fd.h[d]);
        a.satN = 1.0 / (2.0 * fd.w[N - 1] * fd.h[d]);
    };
    hipError_t e;
    a.fuse = 0;
    a.first = !fd.grad_only;
    axis_consts(0);
    if ((e = fo_launch_pass<N, 0, 0>(a, s)) != hipSuccess) return e;
    a.first = 0;
    axis_consts(1);
    if ((e = fo_launch_pass<N, 1, 0>(a, s)) != hipSuccess) return e;
    axis_consts(2);
    a.fuse = fd.grad_only ? 0 : fd.fuse;   // the z pass does the pointwise step
    if ((e = fo_launch_pass<N, 2, 0>(a, s)) != hipSuccess) return e;
    if (fd.grad_only) return hipSuccess;
    if (!a.fuse) {
        const int64_t blocks = std::min<int64_t>((a.npts + 255) / 256, 8192);
        (void)hipGetLastError();
        hipLaunchKernelGGL(fo_flux_kernel<N>, dim3(unsigned(std::max<int64_t>(blocks, 1))), dim3(256), 0, s, a);
        if ((e = hipGetLastError()) != hipSuccess) return e;
    }
    axis_consts(0);
    if ((e = fo_launch_pass<N, 0, 1>(a, s)) != hipSuccess) return e;
    axis_consts(1);
    if ((e = fo_launch_pass<N, 1, 1>(a, s)) != hipSuccess) return e;
    axis_consts(2);
    return fo_launch_pass<N, 2, 1>(a, s);
}

// coef11 from 10 background values per node (sample_background3d!'s kernel,
// src/wave3d_curved.jl:75-96): lapse, sqrt(det gamma), shift, cofactor inverse.
__global__ void fo_coef_kernel(const double* __restrict__ bg, double* __restrict__ c, int64_t n)
{
    for (int64_t q = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; q < n; q += int64_t(gridDim.x) * blockDim.x) {
        const double* b = bg + 10 * q;
        const double g11 = b[4], g12 = b[5], g13 = b[6], g22 = b[7], g23 = b[8], g33 = b[9];
        const double det = g11 * (g22 * g33 - g23 * g23) - g12 * (g12 * g33 - g23 * g13) + g13 * (g12 * g23 - g22 * g13);
        c[q] = b[0];
        c[n + q] = sqrt(det);
        c[2 * n + q] = b[1];
        c[3 * n + q] = b[2];
        c[4 * n + q] = b[3];
        c[5 * n + q] = (g22 * g33 - g23 * g23) / det;
        c[6 * n + q] = (g13 * g23 - g12 * g33) / det;
        c[7 * n + q] = (g12 * g23 - g13 * g22) / det;
        c[8 * n + q] = (g11 * g33 - g13 * g13) / det;
        c[9 * n + q] = (g12 * g13 - g11 * g23) / det;
        c[10 * n + q] = (g11 * g22 - g12 * g12) / det;
    }
}

// Per-node energy density of wave3d_energy (src/wave3d_curved.jl:268-287),
// 1/2 [Pi^2 / sqrtg + sqrtg g^{ij} DPhi_i DPhi_j] Hphys, summed per block in a
// fixed order; the blocks' partial sums are added in a fixed order after.
struct FoWeights { double w[16]; };

template <int N>
__global__ void __launch_bounds__(RED_THREADS) fo_energy_partial(const double* __restrict__ Pi,
                                                                const double* __restrict__ c,
                                                                const double* __restrict__ D, double hvol,
                                                                FoWeights w, int64_t n,
                                                                double* __restrict__ partial)
{
    constexpr int N3 = N * N * N;
    __shared__ double wl[N];
    __shared__ double red[RED_THREADS];
    if (threadIdx.x < N) wl[threadIdx.x] = w.w[threadIdx.x];
    __syncthreads();
    double s = 0.0;
    for (int64_t q = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; q < n; q += int64_t(gridDim.x) * blockDim.x) {
        const int r = int(q % N3);
        const double H = hvol * wl[r % N] * wl[(r / N) % N] * wl[r / (N * N)];
        const double d1 = D[q], d2 = D[n + q], d3 = D[2 * n + q];
        const double sg = c[n + q];
        s += (Pi[q] * Pi[q] / sg +
              sg * (c[5 * n + q] * d1 * d1 + c[8 * n + q] * d2 * d2 + c[10 * n + q] * d3 * d3 + 2 * c[6 * n + q] * d1 * d2 +
                    2 * c[7 * n + q] * d1 * d3 + 2 * c[9 * n + q] * d2 * d3)) * H / 2;
    }
    red[threadIdx.x] = s;
    __syncthreads();
    for (int k = RED_THREADS / 2; k > 0; k >>= 1) {
        if (int(threadIdx.x) < k) red[threadIdx.x] += red[threadIdx.x + k];
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[blockIdx.x] = red[0];
}

__global__ void fo_energy_final(const double* __restrict__ partial, int nb, double* __restrict__ out)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        double s = 0.0;
        for (int b = 0; b < nb; ++b) s += partial[b];
        *out = s;
    }
}

template <int N>
hipError_t launch_fo_energy_n(const FoDesc& fd, hipStream_t s)
{
    const int64_t n = int64_t(N) * N * N * fd.M[0] * fd.M[1] * fd.M[2];
    FoWeights w;
    for (int i = 0; i < 16; ++i) w.w[i] = i < N ? fd.w[i] : 0.0;
    (void)hipGetLastError();
    hipLaunchKernelGGL(fo_energy_partial<N>, dim3(RED_BLOCKS), dim3(RED_THREADS), 0, s, fd.Pi, fd.coef, fd.D,
                       fd.h[0] * fd.h[1] * fd.h[2], w, n, fd.partial);
    hipLaunchKernelGGL(fo_energy_final, dim3(1), dim3(64), 0, s, fd.partial, RED_BLOCKS, fd.Eout);
    return hipGetLastError();
}

}  // namespace

hipError_t launch_fo_rhs(const FoDesc& fd, hipStream_t s)
{
    switch (fd.N) {
    case 3: return launch_fo_n<3>(fd, s);
    case 4: return launch_fo_n<4>(fd, s);
    case 5: return launch_fo_n<5>(fd, s);
    case 6: return launch_fo_n<6>(fd, s);
    case 7: return launch_fo_n<7>(fd, s);
    case 8: return launch_fo_n<8>(fd, s);
    case 9: return launch_fo_n<9>(fd, s);
    case 10: return launch_fo_n<10>(fd, s);
    default: return hipErrorInvalidValue;
    }
}

hipError_t launch_fo_energy(const FoDesc& fd, hipStream_t s)
{
    switch (fd.N) {
    case 3: return launch_fo_energy_n<3>(fd, s);
    case 4: return launch_fo_energy_n<4>(fd, s);
    case 5: return launch_fo_energy_n<5>(fd, s);
    case 6: return launch_fo_energy_n<6>(fd, s);
    case 7: return launch_fo_energy_n<7>(fd, s);
    case 8: return launch_fo_energy_n<8>(fd, s);
    case 9: return launch_fo_energy_n<9>(fd, s);
    case 10: return launch_fo_energy_n<10>(fd, s);
    default: return hipErrorInvalidValue;
    }
}

// Classical RK4 stage update of the first-order system, both fields in one
// pass, k = (kPhi, kPi) the stage's right-hand side (test/wave3d_tests.jl:37-42):
//   stage 1: A = Y + dt/6 k, S = Y + dt/2 k;  stage 2: A += dt/3 k, S = Y + dt/2 k;
//   stage 3: A += dt/3 k, S = Y + dt k;       stage 4: Y = A + dt/6 k.
// Y: the caller's (Phi, Pi); S, A, k: the context's registers, field f of
// each at base + f * stride.
__global__ void fo_rk_update_kernel(double* __restrict__ Y0, double* __restrict__ Y1, double* __restrict__ R,
                                    int64_t stride, int64_t n, int stage, double dt)
{
    for (int64_t g = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; g < 2 * n; g += int64_t(gridDim.x) * blockDim.x) {
        const int f = g >= n;
        const int64_t q = g - f * n;
        double* Y = f ? Y1 : Y0;
        double* S = R + f * stride;
        double* A = R + (2 + f) * stride;
        const double kq = R[(4 + f) * stride + q];
        if (stage == 1) {
            const double y = Y[q];
            A[q] = y + dt / 6.0 * kq;
            S[q] = y + 0.5 * dt * kq;
        } else if (stage == 4) {
            Y[q] = A[q] + dt / 6.0 * kq;
        } else {
            A[q] += dt / 3.0 * kq;
            S[q] = Y[q] + (stage == 3 ? dt : 0.5 * dt) * kq;
        }
    }
}

hipError_t launch_fo_rk_update(double* Y0, double* Y1, double* R, int64_t stride, int64_t n, int stage, double dt,
                               hipStream_t s)
{
    const int64_t blocks = std::min<int64_t>((2 * n + 255) / 256, 65536);
    hipLaunchKernelGGL(fo_rk_update_kernel, dim3(unsigned(std::max<int64_t>(blocks, 1))), dim3(256), 0, s, Y0, Y1, R,
                       stride, n, stage, dt);
    return hipGetLastError();
}

hipError_t launch_fo_coef(const double* bg10, double* coef11, int64_t npts, hipStream_t s)
{
    const int64_t blocks = std::min<int64_t>((npts + 255) / 256, 4096);
    (void)hipGetLastError();
    hipLaunchKernelGGL(fo_coef_kernel, dim3(unsigned(std::max<int64_t>(blocks, 1))), dim3(256), 0, s, bg10, coef11, npts);
    return hipGetLastError();
}

}  // namespace wt2
