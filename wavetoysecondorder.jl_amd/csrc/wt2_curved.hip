// wt2_curved.hip -- curved (warped mesh and/or static ADM background) operator.
//
// Second-order form of the reference's conservative ADM system
// (src/wave3d_curved.jl:1-12): for a static background
//   u_tt = wgt [ d_a (K^ab d_b u) + c^a d_a u_t + d_a (c^a u_t) ]
// in unit-element coordinates, with per-node FP64 coefficients precomputed once
// (the reference resamples them every stage, src/evolve.jl:1289):
//   K^ab = |J| Jinv_ai K^ij Jinv_bj,  K^ij = sqrt(g)(alpha g^ij - b^i b^j / alpha),
//   c^a  = |J| Jinv_ai sqrt(g) b^i / alpha,   wgt = alpha / (sqrt(g) |J|),
// and the SIPG interface treatment of README.md:57-81 generalised to K:
//   vol   sum_a G_a (F^a + c^a v),   F^a = K^ab G_b u,   plus c^a G_a v
//   face  (1/w_r)[ -1/2 n dF^a - tau kappa du - n c^a (v - v_nbr) ]
//         + alpha n sum_b (H_b^-1 G_b^T H_b)(K^ab du)   (symmetrising term)
// kappa = K^aa + sum_{b!=a} |K^ab| averaged over the two sides.  Two kernels
// per evaluation: one writes the normal flux F^a of every element face node
// into a compact face buffer (6 N^2 per element), the second does everything
// else for one element per workgroup with the element staged in LDS and the
// neighbours' face data read from the face buffer and the field arrays.
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <type_traits>
#include <vector>

#include "wt2_curved.hpp"
#include "wt2_internal.hpp"
#include "wt2_ops.hpp"

namespace wt2 {

// ------------------------------------------------------------ host set-up

void mesh_coords(const wt2_mesh_desc& d, const double* xi_n, double* xyz)
{
    const int N = d.N;
    const int* M = d.M;
    const int64_t Ne = int64_t(M[0]) * M[1] * M[2];
    for (int64_t e = 0; e < Ne; ++e) {
        const int ec[3] = {int(e % M[0]), int((e / M[0]) % M[1]), int(e / (int64_t(M[0]) * M[1]))};
        for (int k = 0; k < N; ++k)
            for (int j = 0; j < N; ++j)
                for (int i = 0; i < N; ++i) {
                    const int ii[3] = {i, j, k};
                    double xi[3];
                    for (int a = 0; a < 3; ++a) xi[a] = (ec[a] + xi_n[ii[a]]) / M[a];
                    double s = 0.0;
                    if (d.warp_A != 0.0) {
                        const double tp = 2.0 * M_PI;
                        s = d.warp_A * std::sin(tp * xi[0]) * std::sin(tp * xi[1]) * std::sin(tp * xi[2]);
                    }
                    const int64_t n = i + int64_t(N) * (j + int64_t(N) * (k + int64_t(N) * e));
                    for (int a = 0; a < 3; ++a) xyz[3 * n + a] = d.x0[a] + (d.x1[a] - d.x0[a]) * (xi[a] + s);
                }
    }
}

// Per-node coefficients (SoA, 13 x npts, FP64 and an FP32 copy) and the
// energy weight H3/wgt, computed on the device once per context, from the
// element map and the background (bg10 == nullptr: Minkowski):
//   J = dx/dxi (the analytic warp map, or G applied along each reference axis
//   to the node coordinates xyz of a general mesh), sqrt(det gamma) and the
//   cofactor inverse gamma^ij (src/wave3d_curved.jl:85-94), then
//   K^ab = |J| Jinv_ai K^ij Jinv_bj, c^a = |J| Jinv_ai c^i, wgt = alpha/(sqrt(g)|J|),
//   kappa_a = K^aa + sum_{b != a} |K^ab|.
// emin[e]: the smallest distance between reference-axis-adjacent nodes of
// element e (src/evolve.jl:60-69).  One workgroup per element.
struct CoefArgs {
    int N, M0, M1, M2;
    double x0[3], L[3];
    double warpA;
    const double* xi;      // GLL nodes, weights and G (device copies)
    const double* w;
    const double* G;
    const double* xyz;     // general mesh: node coordinates (3,N,N,N,Ne); nullptr: warp map
    const double* bg10;    // node-major background, nullptr: Minkowski
    double* coef64;
    float* coef32;
    double* wv;
    double* emin;
    int64_t npts;
};

__global__ void __launch_bounds__(256) curved_coef_kernel(CoefArgs a)
{
    __shared__ double sG[256], sxi[16], sw[16];
    __shared__ double X[3 * 1000];
    __shared__ double red[256];
    const int N = a.N, N3 = N * N * N, tid = threadIdx.x;
    const int64_t e = blockIdx.x;
    const int ec[3] = {int(e % a.M0), int((e / a.M0) % a.M1), int(e / (int64_t(a.M0) * a.M1))};
    const int Ms[3] = {a.M0, a.M1, a.M2};
    for (int q = tid; q < N * N; q += blockDim.x) sG[q] = a.G[q];
    for (int q = tid; q < N; q += blockDim.x) { sxi[q] = a.xi[q]; sw[q] = a.w[q]; }
    __syncthreads();
    const double tp = 2.0 * M_PI;
    for (int q = tid; q < N3; q += blockDim.x) {
        const int ii[3] = {q % N, (q / N) % N, q / (N * N)};
        if (a.xyz) {
            for (int c = 0; c < 3; ++c) X[3 * q + c] = a.xyz[3 * (e * N3 + q) + c];
        } else {
            double xr[3];
            for (int c = 0; c < 3; ++c) xr[c] = (ec[c] + sxi[ii[c]]) / Ms[c];
            const double sv = a.warpA * sin(tp * xr[0]) * sin(tp * xr[1]) * sin(tp * xr[2]);
            for (int c = 0; c < 3; ++c) X[3 * q + c] = a.x0[c] + a.L[c] * (xr[c] + sv);
        }
    }
    __syncthreads();
    double hmin = 1e300;
    for (int q = tid; q < N3; q += blockDim.x) {
        const int i = q % N, j = (q / N) % N, k = q / (N * N);
        const int ii[3] = {i, j, k};
        double J[3][3];
        if (a.xyz) {
            for (int c = 0; c < 3; ++c)
                for (int b = 0; b < 3; ++b) {
                    double acc = 0.0;
                    for (int p = 0; p < N; ++p) {
                        int jj[3] = {i, j, k};
                        jj[b] = p;
                        acc += sG[ii[b] * N + p] * X[3 * (jj[0] + N * (jj[1] + N * jj[2])) + c];
                    }
                    J[c][b] = acc;
                }
        } else {
            double xr[3], sn[3], cs[3];
            for (int c = 0; c < 3; ++c) {
                xr[c] = (ec[c] + sxi[ii[c]]) / Ms[c];
                sn[c] = sin(tp * xr[c]);
                cs[c] = cos(tp * xr[c]);
            }
            const double A = a.warpA;
            const double ds[3] = {A * tp * cs[0] * sn[1] * sn[2], A * tp * sn[0] * cs[1] * sn[2],
                                  A * tp * sn[0] * sn[1] * cs[2]};
            for (int c = 0; c < 3; ++c)
                for (int b = 0; b < 3; ++b) J[c][b] = a.L[c] * ((c == b ? 1.0 : 0.0) + ds[b]) / Ms[b];
        }
        const double det = J[0][0] * (J[1][1] * J[2][2] - J[1][2] * J[2][1]) -
                           J[0][1] * (J[1][0] * J[2][2] - J[1][2] * J[2][0]) +
                           J[0][2] * (J[1][0] * J[2][1] - J[1][1] * J[2][0]);
        double Ji[3][3];
        Ji[0][0] = (J[1][1] * J[2][2] - J[1][2] * J[2][1]) / det;
        Ji[0][1] = (J[0][2] * J[2][1] - J[0][1] * J[2][2]) / det;
        Ji[0][2] = (J[0][1] * J[1][2] - J[0][2] * J[1][1]) / det;
        Ji[1][0] = (J[1][2] * J[2][0] - J[1][0] * J[2][2]) / det;
        Ji[1][1] = (J[0][0] * J[2][2] - J[0][2] * J[2][0]) / det;
        Ji[1][2] = (J[0][2] * J[1][0] - J[0][0] * J[1][2]) / det;
        Ji[2][0] = (J[1][0] * J[2][1] - J[1][1] * J[2][0]) / det;
        Ji[2][1] = (J[0][1] * J[2][0] - J[0][0] * J[2][1]) / det;
        Ji[2][2] = (J[0][0] * J[1][1] - J[0][1] * J[1][0]) / det;
        const double aJ = fabs(det);
        const int64_t n = e * N3 + q;
        double al = 1.0, be[3] = {0, 0, 0}, g[6] = {1, 0, 0, 1, 0, 1};
        if (a.bg10) {
            const double* b = a.bg10 + 10 * n;
            al = b[0];
            for (int c = 0; c < 3; ++c) be[c] = b[1 + c];
            for (int c = 0; c < 6; ++c) g[c] = b[4 + c];
        }
        const double g11 = g[0], g12 = g[1], g13 = g[2], g22 = g[3], g23 = g[4], g33 = g[5];
        const double dg = g11 * (g22 * g33 - g23 * g23) - g12 * (g12 * g33 - g23 * g13) + g13 * (g12 * g23 - g22 * g13);
        const double sg = sqrt(dg);
        double gu[3][3];
        gu[0][0] = (g22 * g33 - g23 * g23) / dg;
        gu[0][1] = gu[1][0] = (g13 * g23 - g12 * g33) / dg;
        gu[0][2] = gu[2][0] = (g12 * g23 - g13 * g22) / dg;
        gu[1][1] = (g11 * g33 - g13 * g13) / dg;
        gu[1][2] = gu[2][1] = (g12 * g13 - g11 * g23) / dg;
        gu[2][2] = (g11 * g22 - g12 * g12) / dg;
        double K[3][3], cv[3];
        for (int p = 0; p < 3; ++p) {
            cv[p] = sg * be[p] / al;
            for (int r = 0; r < 3; ++r) K[p][r] = sg * (al * gu[p][r] - be[p] * be[r] / al);
        }
        double Kh[3][3], ch[3];
        for (int c = 0; c < 3; ++c) {
            ch[c] = 0.0;
            for (int p = 0; p < 3; ++p) ch[c] += aJ * Ji[c][p] * cv[p];
            for (int b = 0; b < 3; ++b) {
                double acc = 0.0;
                for (int p = 0; p < 3; ++p)
                    for (int r = 0; r < 3; ++r) acc += Ji[c][p] * K[p][r] * Ji[b][r];
                Kh[c][b] = aJ * acc;
            }
        }
        const double vals[13] = {Kh[0][0], Kh[0][1], Kh[0][2], Kh[1][1], Kh[1][2], Kh[2][2],
                                 ch[0], ch[1], ch[2], al / (sg * aJ),
                                 Kh[0][0] + fabs(Kh[0][1]) + fabs(Kh[0][2]),
                                 Kh[1][1] + fabs(Kh[0][1]) + fabs(Kh[1][2]),
                                 Kh[2][2] + fabs(Kh[0][2]) + fabs(Kh[1][2])};
        for (int c = 0; c < 13; ++c) {
            a.coef64[c * a.npts + n] = vals[c];
            a.coef32[c * a.npts + n] = float(vals[c]);
        }
        a.wv[n] = sw[i] * sw[j] * sw[k] / vals[9];
        for (int c = 0; c < 3; ++c) {
            if (ii[c] == 0) continue;
            int jj[3] = {i, j, k};
            jj[c] -= 1;
            const double* P = X + 3 * q;
            const double* Q = X + 3 * (jj[0] + N * (jj[1] + N * jj[2]));
            const double dd = sqrt((P[0] - Q[0]) * (P[0] - Q[0]) + (P[1] - Q[1]) * (P[1] - Q[1]) +
                                   (P[2] - Q[2]) * (P[2] - Q[2]));
            hmin = fmin(hmin, dd);
        }
    }
    red[tid] = hmin;
    __syncthreads();
    for (int st = blockDim.x / 2; st > 0; st >>= 1) {
        if (tid < st) red[tid] = fmin(red[tid], red[tid + st]);
        __syncthreads();
    }
    if (tid == 0) a.emin[e] = red[0];
}

// neighbour element across face (a, side) of the element with grid
// coordinates ec, -1 at a boundary (element counts fit an int)
__device__ __forceinline__ int64_t nbr_elem(const int (&ec)[3], int a, int side, int M0, int M1, int M2,
                                            int p0, int p1, int p2)
{
    const int Ms[3] = {M0, M1, M2};
    const int ps[3] = {p0, p1, p2};
    int c[3] = {ec[0], ec[1], ec[2]};
    int t = c[a] + (side ? 1 : -1);
    if (t < 0 || t >= Ms[a]) {
        if (!ps[a]) return -1;
        t = (t + Ms[a]) % Ms[a];
    }
    c[a] = t;
    return c[0] + int64_t(M0) * (c[1] + int64_t(M1) * c[2]);
}

// symmetric K^ab stored as (11, 12, 13, 22, 23, 33)
__device__ __forceinline__ int kidx(int a, int b)
{
    if (a == b) return a == 0 ? 0 : (a == 1 ? 3 : 5);
    const int s = a + b;
    return s == 1 ? 1 : (s == 2 ? 2 : 4);
}

// Face-node coefficients of the face terms (static, like the node
// coefficients): for face node q = f N^2 + p + N qq of element e (face f on
// axis a = f / 2), K^{a0}, K^{a1}, K^{a2} and c^a at the node, and kappa_a --
// the average of the two sides' values where the face has a neighbour, the
// own value at an outer face.  Stored per element in the order of the face
// loops, so the operator reads them with adjacent work-items at adjacent
// addresses instead of gathering them at x-face nodes (node stride N) from
// the node fields, which the L2 no longer holds by then.  The FP32 copy
// averages the FP32 node values in FP32, as the FP32 operator did.
struct FaceCoefArgs {
    int N, M0, M1, M2, per0, per1, per2;
    const double* coef64;
    const float* coef32;
    const int2* conn;          // general connectivity, or nullptr (uniform mesh)
    double* fc64;
    float* fc32;
    int64_t npts, Ne;
};

__device__ __forceinline__ int fidx_n(int N, int a, int r, int p, int q)
{
    if (a == 0) return r + N * (p + N * q);
    if (a == 1) return p + N * (r + N * q);
    return p + N * (q + N * r);
}

__global__ void __launch_bounds__(256) curved_facecoef_kernel(FaceCoefArgs a)
{
    const int N = a.N, N2 = N * N, N3 = N2 * N, F2 = 6 * N2;
    const int64_t t = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t >= a.Ne * F2) return;
    const int64_t e = t / F2;
    const int q = int(t - e * F2);
    const int f = q / N2, pq = q - f * N2, p = pq % N, qq = pq / N;
    const int ax = f >> 1, side = f & 1;
    const int64_t gn = e * N3 + fidx_n(N, ax, side ? N - 1 : 0, p, qq);
    int64_t en;
    int fn, pn = p, qn = qq;
    if (a.conn) {
        const int2 cf = a.conn[e * 6 + f];
        en = cf.x;
        fn = cf.y & 7;
        const int ori = (cf.y >> 3) & 7;
        if (ori & 1) { pn = qq; qn = p; }
        if (ori & 2) pn = N - 1 - pn;
        if (ori & 4) qn = N - 1 - qn;
    } else {
        const int ei = int(e), exy = ei % (a.M0 * a.M1);
        const int ec[3] = {exy % a.M0, exy / a.M0, ei / (a.M0 * a.M1)};
        en = nbr_elem(ec, ax, side, a.M0, a.M1, a.M2, a.per0, a.per1, a.per2);
        fn = (ax << 1) | (1 - side);
    }
    const int64_t np = a.npts;
    double* o64 = a.fc64 + e * 5 * F2 + q;
    float* o32 = a.fc32 + e * 5 * F2 + q;
    for (int b = 0; b < 3; ++b) {
        o64[b * F2] = a.coef64[kidx(ax, b) * np + gn];
        o32[b * F2] = a.coef32[kidx(ax, b) * np + gn];
    }
    double k64 = a.coef64[(10 + ax) * np + gn];
    float k32 = a.coef32[(10 + ax) * np + gn];
    if (en >= 0) {
        const int axn = fn >> 1;
        const int64_t gnn = en * N3 + fidx_n(N, axn, (fn & 1) ? N - 1 : 0, pn, qn);
        k64 = 0.5 * (k64 + a.coef64[(10 + axn) * np + gnn]);
        k32 = 0.5f * (k32 + a.coef32[(10 + axn) * np + gnn]);
    }
    o64[3 * F2] = k64;
    o32[3 * F2] = k32;
    o64[4 * F2] = a.coef64[(6 + ax) * np + gn];
    o32[4 * F2] = a.coef32[(6 + ax) * np + gn];
}

// Compute cd's coefficients on the device (set-up path: the background and
// general-mesh coordinates are staged through temporary device buffers).
static int compute_coefs_device(CurvedData& cd, const wt2_mesh_desc& d, const double* xi, const double* w,
                                const double* G, const double* bg10, hipStream_t s)
{
    const int N = d.N;
    const size_t n = size_t(cd.npts), Ne = size_t(cd.Ne);
    double* tmp[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};   // ops, bg, xyz, emin
    auto cleanup = [&]() { for (double* p : tmp) if (p) (void)hipFree(p); };
    const size_t ops_n = 16 + 16 + 256;
    if (hipMalloc(&tmp[0], ops_n * sizeof(double)) != hipSuccess ||
        (bg10 && hipMalloc(&tmp[1], 10 * n * sizeof(double)) != hipSuccess) ||
        (!cd.xyz.empty() && hipMalloc(&tmp[2], 3 * n * sizeof(double)) != hipSuccess) ||
        hipMalloc(&tmp[3], Ne * sizeof(double)) != hipSuccess) {
        cleanup();
        set_error("curved set-up: device allocation failed");
        return WT2_ERR_OOM;
    }
    std::vector<double> ops(ops_n, 0.0);
    for (int i = 0; i < N; ++i) { ops[i] = xi[i]; ops[16 + i] = w[i]; }
    for (int i = 0; i < N * N; ++i) ops[32 + i] = G[i];
    CoefArgs a;
    a.N = N; a.M0 = d.M[0]; a.M1 = d.M[1]; a.M2 = d.M[2];
    for (int c = 0; c < 3; ++c) { a.x0[c] = d.x0[c]; a.L[c] = d.x1[c] - d.x0[c]; }
    a.warpA = d.warp_A;
    a.xi = tmp[0]; a.w = tmp[0] + 16; a.G = tmp[0] + 32;
    a.xyz = tmp[2]; a.bg10 = tmp[1];
    a.coef64 = cd.coef64; a.coef32 = cd.coef32; a.wv = cd.wv; a.emin = tmp[3];
    a.npts = cd.npts;
    std::vector<double> emin(Ne);
    hipError_t err = hipMemcpyAsync(tmp[0], ops.data(), ops_n * sizeof(double), hipMemcpyHostToDevice, s);
    if (err == hipSuccess && bg10) err = hipMemcpyAsync(tmp[1], bg10, 10 * n * sizeof(double), hipMemcpyHostToDevice, s);
    if (err == hipSuccess && tmp[2]) err = hipMemcpyAsync(tmp[2], cd.xyz.data(), 3 * n * sizeof(double), hipMemcpyHostToDevice, s);
    if (err == hipSuccess) {
        hipLaunchKernelGGL(curved_coef_kernel, dim3(unsigned(Ne)), dim3(256), 0, s, a);
        err = hipGetLastError();
    }
    if (err == hipSuccess) {
        FaceCoefArgs fa;
        fa.N = N; fa.M0 = d.M[0]; fa.M1 = d.M[1]; fa.M2 = d.M[2];
        fa.per0 = cd.periodic[0]; fa.per1 = cd.periodic[1]; fa.per2 = cd.periodic[2];
        fa.coef64 = cd.coef64; fa.coef32 = cd.coef32; fa.conn = cd.conn;
        fa.fc64 = cd.fc64; fa.fc32 = cd.fc32;
        fa.npts = cd.npts; fa.Ne = cd.Ne;
        const int64_t nt = cd.Ne * 6 * N * N;
        hipLaunchKernelGGL(curved_facecoef_kernel, dim3(unsigned((nt + 255) / 256)), dim3(256), 0, s, fa);
        err = hipGetLastError();
    }
    if (err == hipSuccess) err = hipMemcpyAsync(emin.data(), tmp[3], Ne * sizeof(double), hipMemcpyDeviceToHost, s);
    if (err == hipSuccess) err = hipStreamSynchronize(s);
    cleanup();
    if (err != hipSuccess) {
        set_error(std::string("curved set-up (coefficients): ") + hipGetErrorString(err));
        return WT2_ERR_HIP;
    }
    cd.h_min = *std::min_element(emin.begin(), emin.end());
    return WT2_OK;
}

int curved_init(CurvedData& cd, const wt2_mesh_desc& d, const double* xi, const double* w,
                const double* G, const double* bg10, hipStream_t s, int64_t* bytes)
{
    cd = CurvedData();
    cd.N = d.N;
    for (int a = 0; a < 3; ++a) { cd.M[a] = d.M[a]; cd.periodic[a] = d.periodic[a]; }
    cd.Ne = int64_t(d.M[0]) * d.M[1] * d.M[2];
    cd.npts = cd.Ne * d.N * d.N * d.N;
    cd.tau = d.tau;
    for (int f = 0; f < 6; ++f)
        if (d.face_kind[f] == SOMMERFELD_TAG && !d.periodic[f / 2]) cd.som |= 1 << f;
    cd.inv_R = d.sommerfeld_inv_R;
    for (int i = 0; i < d.N; ++i) cd.w[i] = w[i];
    for (int i = 0; i < d.N * d.N; ++i) cd.G[i] = G[i];
    const size_t n = size_t(cd.npts);
    const size_t fb = size_t(cd.Ne) * 6 * d.N * d.N;
    if (hipMalloc(&cd.coef64, 13 * n * sizeof(double)) != hipSuccess ||
        hipMalloc(&cd.coef32, 13 * n * sizeof(float)) != hipSuccess ||
        hipMalloc(&cd.wv, n * sizeof(double)) != hipSuccess ||
        hipMalloc(&cd.fbuf, fb * sizeof(double)) != hipSuccess ||
        hipMalloc(&cd.fc64, 5 * fb * sizeof(double)) != hipSuccess ||
        hipMalloc(&cd.fc32, 5 * fb * sizeof(float)) != hipSuccess) {
        curved_free(cd);
        set_error("curved set-up: device allocation failed");
        return WT2_ERR_OOM;
    }
    *bytes += int64_t(13 * n * (sizeof(double) + sizeof(float)) + n * sizeof(double) + fb * sizeof(double) +
                      5 * fb * (sizeof(double) + sizeof(float)));
    if (!bg10) {
        // Minkowski coefficients are computed on first use, so that a
        // background set right after creation is computed only once
        cd.pending = true;
        return WT2_OK;
    }
    return compute_coefs_device(cd, d, xi, w, G, bg10, s);
}

int curved_set_background(CurvedData& cd, const wt2_mesh_desc& d, const double* xi,
                          const double* w, const double* G, const double* bg10, hipStream_t s)
{
    const int rc = compute_coefs_device(cd, d, xi, w, G, bg10, s);
    if (rc == WT2_OK) cd.pending = false;
    return rc;
}

int curved_init_mesh(CurvedData& cd, const wt2_mesh_desc& d, const double* xi, const double* w,
                     const double* G, const double* xyz, const std::vector<int2>& conn, hipStream_t s,
                     int64_t* bytes)
{
    int rc = curved_init(cd, d, xi, w, G, nullptr, s, bytes);
    if (rc != WT2_OK) return rc;
    cd.xyz.assign(xyz, xyz + size_t(3 * cd.npts));
    cd.som = 0;    // Sommerfeld faces come from the per-face tags
    if (hipMalloc(&cd.conn, conn.size() * sizeof(int2)) != hipSuccess) {
        curved_free(cd);
        set_error("curved set-up: device allocation failed");
        return WT2_ERR_OOM;
    }
    *bytes += int64_t(conn.size() * sizeof(int2));
    if (hipMemcpy(cd.conn, conn.data(), conn.size() * sizeof(int2), hipMemcpyHostToDevice) != hipSuccess) {
        curved_free(cd);
        set_error("curved set-up: connectivity upload failed");
        return WT2_ERR_HIP;
    }
    return WT2_OK;
}

void curved_free(CurvedData& cd)
{
    (void)hipFree(cd.coef64); (void)hipFree(cd.coef32); (void)hipFree(cd.wu); (void)hipFree(cd.wv);
    (void)hipFree(cd.fbuf); (void)hipFree(cd.conn); (void)hipFree(cd.fc64); (void)hipFree(cd.fc32);
    cd = CurvedData();
}

// ---------------------------------------------------------------- kernels

template <int N, typename TC>
struct CConst {
    TC G[N * N];
    TC w[N];
    TC winv[N];
    TC tau;
    TC bv[6];
    TC iR;      // 1/R of the Sommerfeld faces
    int som;    // bit f: outer face f is a Sommerfeld face
};

template <int N, typename TC>
struct CArgs {
    CConst<N, TC> k;
    const void* u; const void* v;   // stage fields (TS)
    const TC* coef;                 // 13 x npts SoA
    TC* fbuf;                       // 6 N^2 per element
    const TC* fcoef;                // 5 x 6 N^2 per element: face-node coefficients
    int fcnode;                     // element kernels: the face node's own K row and c^a from
                                    // the node coefficients (cache hits), not the face-node table
    int xcd;                        // element kernel (one context): contiguous element ranges per XCD
    const TC* op;                   // even/odd operator table (OpLayout<N>)
    void* out;                      // output (TO): wgt (...) or unweighted
    int64_t npts;
    int M0, M1, M2, per0, per1, per2;
    int adv, weighted;
    const int2* conn;               // general connectivity, or nullptr (uniform mesh)
};

template <int N>
__device__ __forceinline__ int fidx(int a, int r, int p, int q)
{
    if (a == 0) return r + N * (p + N * q);
    if (a == 1) return p + N * (r + N * q);
    return p + N * (q + N * r);
}

template <int N, typename TC>
__device__ __forceinline__ TC kab(const TC* coef, int64_t npts, int64_t n, int a, int b)
{
    return coef[kidx(a, b) * npts + n];
}

// reference gradient of the element field in LDS at node (i,j,k)
template <int N, typename TC>
__device__ __forceinline__ void grad_lds(const TC* f, const CConst<N, TC>& k, int i, int j, int l, TC g[3])
{
    TC s0 = 0, s1 = 0, s2 = 0;
#pragma unroll
    for (int p = 0; p < N; ++p) {
        s0 = fma(k.G[i * N + p], f[p + N * (j + N * l)], s0);
        s1 = fma(k.G[j * N + p], f[i + N * (p + N * l)], s1);
        s2 = fma(k.G[l * N + p], f[i + N * (j + N * p)], s2);
    }
    g[0] = s0; g[1] = s1; g[2] = s2;
}

// Kernel A: normal flux F^a = K^ab G_b u at every face node of every element.
template <int N, typename TC, typename TS>
__device__ __forceinline__ void faceflux_body(const CArgs<N, TC>& a, const int64_t e)
{
    // launched with FT work-items; work-item t takes face nodes t, t + FT, ...
    constexpr int FT = 256, F2 = 6 * N * N, NQ = (F2 + FT - 1) / FT;
    __shared__ TC su[N * N * N];
    __shared__ CConst<N, TC> kl;     // G indexed per lane: from LDS, not kernel arguments
    const TS* u = static_cast<const TS*>(a.u);
    for (int q = threadIdx.x; q < N * N * N; q += FT) su[q] = TC(u[e * N * N * N + q]);
    for (int q = threadIdx.x; q < N * N; q += FT) kl.G[q] = a.k.G[q];
    // this work-item's K^{a0}, K^{a1}, K^{a2} are loaded before the barrier, so
    // their latency overlaps the element's u
    const TC* fc = a.fcoef + e * 5 * F2;
    TC K0[NQ], K1[NQ], K2[NQ];
#pragma unroll
    for (int t = 0; t < NQ; ++t) {
        const int q = int(threadIdx.x) + FT * t;
        K0[t] = K1[t] = K2[t] = TC(0);
        if (q < F2) { K0[t] = fc[q]; K1[t] = fc[F2 + q]; K2[t] = fc[2 * F2 + q]; }
    }
    __syncthreads();
#pragma unroll
    for (int t = 0; t < NQ; ++t) {
        const int q = int(threadIdx.x) + FT * t;
        if (q >= F2) break;
        const int f = q / (N * N), pq = q - f * N * N;
        const int p = pq % N, qq = pq / N;
        const int ax = f >> 1, r = (f & 1) ? N - 1 : 0;
        const int n = fidx<N>(ax, r, p, qq);
        const int i = n % N, j = (n / N) % N, l = n / (N * N);
        TC g[3];
        grad_lds<N, TC>(su, kl, i, j, l, g);
        TC F = K0[t] * g[0];
        F = fma(K1[t], g[1], F);
        F = fma(K2[t], g[2], F);
        a.fbuf[e * F2 + q] = F;
    }
}

// Kernel B: everything else, for a block of EW elements per workgroup.  A
// work-item owns one z-line (i, j, all k) of an element:
//   P1  u, u_t of the elements -> LDS (A, B);
//   P2  gradients: d/dz by the even/odd operator on the line in registers,
//       d/dx and d/dy from LDS with the work-item's rows G[i], G[j] in
//       registers; fluxes W^a = K^ab g_b + c^a u_t and c^a d_a u_t per node;
//       face terms per face node (face-parallel over the element's 6 N^2
//       face nodes) into LDS: S^b = alpha n K^ab du and the point term;
//   P3  W^0, W^1 -> LDS (over A, B); tangential lifts of S at the face nodes
//       (face-parallel, added to the point term);
//   P4  sum_a G_a W^a (x, y from LDS, z in registers) + c^a G_a u_t + the
//       normal lifts of S^a along each face-normal line + the face terms,
//       times wgt, to the output (coalesced over the work-items of a k).
template <int N>
struct VolShape {
    static constexpr int L = N * N;                      // work-items per element
    static constexpr int EW = (128 / L) > 0 ? 128 / L : 1;
    static constexpr int NT = EW * L;
    static constexpr int N3 = N * N * N;
    static constexpr int F2 = 6 * N * N;                 // face nodes per element
    static constexpr int PER = 2 * N3 + 24 * N * N;      // LDS values per element
};

template <int N, typename TC>
struct VolTables {
    TC G[N * N];          // G[i][p]
    TC GW[N * N];         // G[l][t] * w[l] (tangential lifts)
    TC op[OpLayout<N>::SIZE];
    TC w[N], winv[N];
};

// Face terms of face node q (face f = q / N^2) of element e: S^b = alpha n K^ab du
// into Sf and the point term into Pt (both in LDS, per element); A, B: the
// element's u, u_t in LDS; ec: the element's grid coordinates (uniform meshes).
template <int N, typename TC, typename TS, bool GEN>
__device__ __forceinline__ void curved_face_node(const CArgs<N, TC>& a, const int64_t e, const int (&ec)[3],
                                                 const int q, const TC* A, const TC* B, TC* Sf, TC* Pt,
                                                 const TC* __restrict__ winv)
{
    constexpr int N2 = N * N, N3 = N * N * N, F2 = 6 * N * N;
    const TS* u = static_cast<const TS*>(a.u);
    const TS* v = static_cast<const TS*>(a.v);
    const bool adv = a.adv != 0;
    const int64_t np = a.npts;
    const int64_t eb = e * N3;
    const int f = q / N2, pq = q - f * N2;
    const int p = pq % N, qq = pq / N;
    const int ax = f >> 1, side = f & 1;
    const int r = side ? N - 1 : 0;
    const TC nrm = side ? TC(1) : TC(-1);
    const int n = fidx<N>(ax, r, p, qq);
    const int64_t gn = eb + n;
    const TC* fc = a.fcoef + e * 5 * F2 + q;   // K^{a0}, K^{a1}, K^{a2}, kappa, c^a
    // this side's K row and c^a: equal copies sit in the face-node table (read
    // in face-loop order) and in the node coefficients (which this workgroup's
    // node work-items have just read, so these gathers are cache hits)
    TC kr[3], ca;
    if (a.fcnode) {
#pragma unroll
        for (int b = 0; b < 3; ++b) kr[b] = a.coef[kidx(ax, b) * np + gn];
        ca = a.coef[(6 + ax) * np + gn];
    } else {
#pragma unroll
        for (int b = 0; b < 3; ++b) kr[b] = fc[b * F2];
        ca = fc[4 * F2];
    }
    // neighbour element en (-1: outer face), its face fn, the neighbour
    // face's node (pn, qn) that meets this face's (p, qq), and whether
    // this outer face is a Sommerfeld face; bvf: its Dirichlet value
    int64_t en;
    int fn, pn = p, qn = qq;
    bool som, exc = false;
    TC bvf;
    if (GEN && a.conn) {
        const int2 cf = a.conn[e * 6 + f];
        en = cf.x;
        fn = cf.y & 7;
        const int ori = (cf.y >> 3) & 7, tag = cf.y >> 6;
        if (ori & 1) { pn = qq; qn = p; }
        if (ori & 2) pn = N - 1 - pn;
        if (ori & 4) qn = N - 1 - qn;
        som = tag == 7;
        exc = tag == 8;
        bvf = (tag >= 1 && tag <= 6) ? a.k.bv[tag - 1] : TC(0);
    } else {
        en = nbr_elem(ec, ax, side, a.M0, a.M1, a.M2, a.per0, a.per1, a.per2);
        fn = (ax << 1) | (1 - side);
        som = (a.k.som >> f) & 1;
        bvf = a.k.bv[f];
    }
    if (en < 0 && exc) {
        // excision face (src/wave_strong_rhs.jl:188-213): no SAT at all
#pragma unroll
        for (int b = 0; b < 3; ++b) Sf[(b * 6 + f) * N2 + pq] = TC(0);
        Pt[q] = TC(0);
        return;
    }
    if (en < 0 && som) {
        // Sommerfeld face (src/wave.jl:444-518): no SIPG terms; the drag
        // -(wF/Hphys)(u_t + d_n u + u/R), here unweighted (the output
        // is multiplied by wgt = 1/|J|): |n_u| d_n u = n F^a with this
        // element's own normal flux, |n_u| = sqrt(K^aa / wgt), and
        // wF/Hphys = |n_u| / (w_r |J|).  B holds u_t (zero without it).
#pragma unroll
        for (int b = 0; b < 3; ++b) Sf[(b * 6 + f) * N2 + pq] = TC(0);
        const TC nu = sqrt(kr[ax] / a.coef[9 * np + gn]);
        const TC drag = fma(nrm, a.fbuf[e * F2 + q], nu * fma(a.k.iR, A[n], B[n]));
        Pt[q] = -drag * winv[r];
        return;
    }
    TC du, dF = TC(0), dv = TC(0), alpha;
    const TC kap = fc[3 * F2];
    if (en >= 0) {
        // outward normal fluxes of the two sides: nrm F (own) and
        // nrm_n F_n (neighbour); dF = F + (nrm nrm_n) F_n
        const int axn = fn >> 1;
        const int64_t gnn = en * N3 + fidx<N>(axn, (fn & 1) ? N - 1 : 0, pn, qn);
        du = A[n] - TC(u[gnn]);
        const TC Fn = a.fbuf[en * F2 + fn * N2 + pn + N * qn];
        const TC Fo = a.fbuf[e * F2 + q];
        dF = ((fn & 1) == side) ? Fo + Fn : Fo - Fn;
        if (adv) dv = B[n] - TC(v[gnn]);
        alpha = TC(0.5);
    } else {
        du = A[n] - bvf;
        if (adv) dv = B[n];
        alpha = TC(1);
    }
    const TC an = alpha * nrm;
#pragma unroll
    for (int b = 0; b < 3; ++b) Sf[(b * 6 + f) * N2 + pq] = an * kr[b] * du;
    TC pt = TC(-0.5) * nrm * dF - a.k.tau * kap * du;
    if (adv) pt -= nrm * ca * dv;
    Pt[q] = pt * winv[r];
}

// Tangential lifts of S at face node q of an element: added to its point term.
template <int N, typename TC>
__device__ __forceinline__ void curved_face_lift(const int q, const TC* Sf, TC* Pt, const TC* __restrict__ GW,
                                                 const TC* __restrict__ winv)
{
    constexpr int N2 = N * N;
    const int f = q / N2, pq = q - f * N2;
    const int tp = pq % N, tq = pq / N;
    const int ax = f >> 1, r = (f & 1) ? N - 1 : 0;
    const int tb0 = ax == 0 ? 1 : 0, tb1 = ax == 2 ? 1 : 2;
    TC t0 = TC(0), t1 = TC(0);
#pragma unroll
    for (int l = 0; l < N; ++l) {
        t0 = fma(GW[l * N + tp], Sf[(tb0 * 6 + f) * N2 + l + N * tq], t0);
        t1 = fma(GW[l * N + tq], Sf[(tb1 * 6 + f) * N2 + tp + N * l], t1);
    }
    Pt[q] = fma(winv[r], fma(winv[tp], t0, winv[tq] * t1), Pt[q]);
}

// GEN: the mesh may carry a connectivity table (a.conn); false compiles the
// uniform mesh's neighbour arithmetic alone
template <int N, typename TC, typename TS, typename TO, bool GEN>
__device__ __forceinline__ void volume_body(const CArgs<N, TC>& a, const int64_t e0, const int nel)
{
    using S = VolShape<N>;
    constexpr int N3 = S::N3, N2 = N * N;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    __shared__ VolTables<N, TC> T;
    const int tid = threadIdx.x;
    for (int q = tid; q < N2; q += blockDim.x) {
        T.G[q] = a.k.G[q];
        T.GW[q] = a.k.G[q] * a.k.w[q / N];
    }
    for (int q = tid; q < OpLayout<N>::SIZE; q += blockDim.x) T.op[q] = a.op[q];
    for (int q = tid; q < N; q += blockDim.x) { T.w[q] = a.k.w[q]; T.winv[q] = a.k.winv[q]; }

    const TS* u = static_cast<const TS*>(a.u);
    const TS* v = static_cast<const TS*>(a.v);
    const bool adv = a.adv != 0;
    TC* lds = reinterpret_cast<TC*>(smem_raw);
    // P1
    for (int q = tid; q < nel * N3; q += blockDim.x) {
        const int el = q / N3, n = q - el * N3;
        const int64_t g = e0 * N3 + q;
        lds[el * S::PER + n] = TC(u[g]);
        lds[el * S::PER + N3 + n] = adv ? TC(v[g]) : TC(0);
    }
    __syncthreads();

    const int slot = tid / S::L, lane = tid - slot * S::L;
    const bool act = slot < nel;
    const int sl = act ? slot : 0;
    const int64_t e = e0 + sl;
    const int64_t eb = e * N3;
    TC* A = lds + sl * S::PER;          // u, then W^0
    TC* B = A + N3;                     // u_t, then W^1
    TC* Sf = B + N3;                    // 3 x 6 N^2: S^b per face node
    TC* Pt = Sf + 18 * N2;              // 6 N^2: point (+ tangential) term per face node
    const int i = lane % N, j = lane / N;
    const int64_t np = a.npts;
    // this element's grid coordinates (uniform meshes), once per work-item
    int ec[3];
    {
        const int ei = int(e), exy = ei % (a.M0 * a.M1);
        ec[0] = exy % a.M0; ec[1] = exy / a.M0; ec[2] = ei / (a.M0 * a.M1);
    }

    TC W0[N], W1[N], W2[N], cgv[N];
    if (act) {
        // ---- P2 (face-parallel): face-node terms of this element, first, while
        // the fluxes of the z-line are not yet live in registers (more of the
        // face loop's gathers in flight)
        for (int q = lane; q < S::F2; q += S::L)
            curved_face_node<N, TC, TS, GEN>(a, e, ec, q, A, B, Sf, Pt, T.winv);
        // ---- P2: gradients and fluxes on this z-line
        TC Gi[N], Gj[N];
#pragma unroll
        for (int p = 0; p < N; ++p) { Gi[p] = T.G[i * N + p]; Gj[p] = T.G[j * N + p]; }
        TC g2[N], gv2[N];
        {
            TC x[N];
#pragma unroll
            for (int k = 0; k < N; ++k) x[k] = A[i + N * j + N2 * k];
            apply_G<N, TC>(T.op, x, g2);
#pragma unroll
            for (int k = 0; k < N; ++k) x[k] = B[i + N * j + N2 * k];
            if (adv) apply_G<N, TC>(T.op, x, gv2);
            else {
#pragma unroll
                for (int k = 0; k < N; ++k) gv2[k] = TC(0);
            }
        }
#pragma unroll
        for (int k = 0; k < N; ++k) {
            const int n = i + N * j + N2 * k;
            TC g0 = TC(0), g1 = TC(0);
#pragma unroll
            for (int p = 0; p < N; ++p) {
                g0 = fma(Gi[p], A[p + N * j + N2 * k], g0);
                g1 = fma(Gj[p], A[i + N * p + N2 * k], g1);
            }
            const int64_t gn = eb + n;
            const TC K00 = a.coef[0 * np + gn], K01 = a.coef[1 * np + gn], K02 = a.coef[2 * np + gn];
            const TC K11 = a.coef[3 * np + gn], K12 = a.coef[4 * np + gn], K22 = a.coef[5 * np + gn];
            TC F0 = K00 * g0, F1 = K01 * g0, F2 = K02 * g0;
            F0 = fma(K01, g1, F0); F1 = fma(K11, g1, F1); F2 = fma(K12, g1, F2);
            F0 = fma(K02, g2[k], F0); F1 = fma(K12, g2[k], F1); F2 = fma(K22, g2[k], F2);
            TC cg = TC(0);
            if (adv) {
                const TC c0 = a.coef[6 * np + gn], c1 = a.coef[7 * np + gn], c2 = a.coef[8 * np + gn];
                const TC vn = B[n];
                F0 = fma(c0, vn, F0); F1 = fma(c1, vn, F1); F2 = fma(c2, vn, F2);
                TC gv0 = TC(0), gv1 = TC(0);
#pragma unroll
                for (int p = 0; p < N; ++p) {
                    gv0 = fma(Gi[p], B[p + N * j + N2 * k], gv0);
                    gv1 = fma(Gj[p], B[i + N * p + N2 * k], gv1);
                }
                cg = c0 * gv0 + c1 * gv1 + c2 * gv2[k];
            }
            W0[k] = F0; W1[k] = F1; W2[k] = F2; cgv[k] = cg;
        }
    }
    __syncthreads();
    if (act) {
        // ---- P3: fluxes to LDS; tangential lifts of S at the face nodes
#pragma unroll
        for (int k = 0; k < N; ++k) {
            A[i + N * j + N2 * k] = W0[k];
            B[i + N * j + N2 * k] = W1[k];
        }
        for (int q = lane; q < S::F2; q += S::L) curved_face_lift<N, TC>(q, Sf, Pt, T.GW, T.winv);
    }
    __syncthreads();
    if (!act) return;
    // ---- P4: divergence, lifts, face terms, output
    TC Gi[N], Gj[N];
#pragma unroll
    for (int p = 0; p < N; ++p) { Gi[p] = T.G[i * N + p]; Gj[p] = T.G[j * N + p]; }
    TC d2[N];
    apply_G<N, TC>(T.op, W2, d2);
    // normal-line lift coefficients G[r][m] / w[m] of this work-item's i and j
    const TC cx0 = T.G[i] * T.winv[i], cxN = T.G[(N - 1) * N + i] * T.winv[i];
    const TC cy0 = T.G[j] * T.winv[j], cyN = T.G[(N - 1) * N + j] * T.winv[j];
    const TC sz0 = Sf[(2 * 6 + 4) * N2 + i + N * j], szN = Sf[(2 * 6 + 5) * N2 + i + N * j];
    TO* out = static_cast<TO*>(a.out);
#pragma unroll
    for (int k = 0; k < N; ++k) {
        TC s = cgv[k] + d2[k];
#pragma unroll
        for (int p = 0; p < N; ++p) {
            s = fma(Gi[p], A[p + N * j + N2 * k], s);
            s = fma(Gj[p], B[i + N * p + N2 * k], s);
        }
        // lifts along the normal lines: x faces (0, 1) at (j, k), y faces (2, 3) at (i, k),
        // z faces (4, 5) at (i, j)
        s = fma(cx0, Sf[(0 * 6 + 0) * N2 + j + N * k], s);
        s = fma(cxN, Sf[(0 * 6 + 1) * N2 + j + N * k], s);
        s = fma(cy0, Sf[(1 * 6 + 2) * N2 + i + N * k], s);
        s = fma(cyN, Sf[(1 * 6 + 3) * N2 + i + N * k], s);
        s = fma(T.G[k] * T.winv[k], sz0, s);
        s = fma(T.G[(N - 1) * N + k] * T.winv[k], szN, s);
        // point and tangential terms at the face nodes
        if (i == 0) s += Pt[0 * N2 + j + N * k];
        if (i == N - 1) s += Pt[1 * N2 + j + N * k];
        if (j == 0) s += Pt[2 * N2 + i + N * k];
        if (j == N - 1) s += Pt[3 * N2 + i + N * k];
        if (k == 0) s += Pt[4 * N2 + i + N * j];
        if (k == N - 1) s += Pt[5 * N2 + i + N * j];
        const int64_t gn = eb + i + N * j + N2 * k;
        if (a.weighted) s *= a.coef[9 * np + gn];
        out[gn] = TO(s);
    }
}

// Element kernel, one work-item per element node (the form every launch
// uses; volume_body above is the one-work-item-per-z-line form it replaced).
// Each work-item issues all of its node's coefficient loads at the start, so
// a workgroup keeps 10 N^3 loads in flight (the z-line form issued them node
// by node, two round trips per node, at two wavefronts per SIMD), and holds
// few registers (N = 10 FP64: one element of 1024 work-items).
//   P1  u, u_t -> LDS (A, B); the node's K^ab, c^a (and wgt) -> registers;
//   P2  face terms (face-parallel over the 6 N^2 face nodes, into Sf, Pt);
//       the gradients of u and u_t at the node from LDS; the fluxes
//       W^a = K^ab g_b + c^a u_t and c^a d_a u_t;
//   P3  W^0, W^1, W^2 -> LDS (A, B, C); tangential lifts at the face nodes;
//   P4  sum_a G_a W^a + c^a d_a u_t + normal lifts + face terms, times wgt.
template <int N>
struct VolNodeShape {
    static constexpr int N2 = N * N, N3 = N * N * N, F2 = 6 * N * N;
    static constexpr int EW = N3 >= 256 ? 1 : 256 / N3;      // elements per workgroup
    static constexpr int NT = ((EW * N3 + 63) / 64) * 64;    // work-items per workgroup
    static constexpr int PER = 3 * N3 + 24 * N2;             // LDS values per element
};

template <int N, typename TC, typename TS, typename TO, bool GEN>
__device__ __forceinline__ void volume_node_body(const CArgs<N, TC>& a, const int64_t e0, const int nel)
{
    using S = VolNodeShape<N>;
    constexpr int N3 = S::N3, N2 = S::N2;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    __shared__ VolTables<N, TC> T;
    const int tid = threadIdx.x;
    for (int q = tid; q < N2; q += blockDim.x) {
        T.G[q] = a.k.G[q];
        T.GW[q] = a.k.G[q] * a.k.w[q / N];
    }
    for (int q = tid; q < N; q += blockDim.x) { T.w[q] = a.k.w[q]; T.winv[q] = a.k.winv[q]; }

    const TS* u = static_cast<const TS*>(a.u);
    const TS* v = static_cast<const TS*>(a.v);
    const bool adv = a.adv != 0;
    const int64_t np = a.npts;
    const int slot = tid / N3, n = tid - slot * N3;
    const bool act = slot < nel;
    const int sl = act ? slot : 0;
    const int64_t e = e0 + sl;
    const int64_t gn = e * N3 + n;
    TC* lds = reinterpret_cast<TC*>(smem_raw);
    TC* A = lds + sl * S::PER;          // u, then W^0
    TC* B = A + N3;                     // u_t, then W^1
    TC* C = B + N3;                     // W^2
    TC* Sf = C + N3;                    // 3 x 6 N^2: S^b per face node
    TC* Pt = Sf + 18 * N2;              // 6 N^2: point (+ tangential) term per face node
    const int i = n % N, j = (n / N) % N, k = n / N2;

    // P1
    TC K00 = TC(0), K01 = TC(0), K02 = TC(0), K11 = TC(0), K12 = TC(0), K22 = TC(0);
    TC c0 = TC(0), c1 = TC(0), c2 = TC(0), wg = TC(1), vn = TC(0);
    if (act) {
        K00 = a.coef[0 * np + gn]; K01 = a.coef[1 * np + gn]; K02 = a.coef[2 * np + gn];
        K11 = a.coef[3 * np + gn]; K12 = a.coef[4 * np + gn]; K22 = a.coef[5 * np + gn];
        if (adv) { c0 = a.coef[6 * np + gn]; c1 = a.coef[7 * np + gn]; c2 = a.coef[8 * np + gn]; }
        if (a.weighted) wg = a.coef[9 * np + gn];
        A[n] = TC(u[gn]);
        vn = adv ? TC(v[gn]) : TC(0);
        B[n] = vn;
    }
    int ec[3];
    {
        const int ei = int(e), exy = ei % (a.M0 * a.M1);
        ec[0] = exy % a.M0; ec[1] = exy / a.M0; ec[2] = ei / (a.M0 * a.M1);
    }
    __syncthreads();

    // P2
    TC W0 = TC(0), W1 = TC(0), W2 = TC(0), cg = TC(0);
    if (act) {
        for (int q = n; q < S::F2; q += N3) curved_face_node<N, TC, TS, GEN>(a, e, ec, q, A, B, Sf, Pt, T.winv);
        // gradients of u (g) and u_t (h) in one loop, not unrolled: each G
        // value is used for both as soon as it is read (unrolled, or as two
        // loops, the compiler kept the loop's 3N G and 6N field values live:
        // N = 10 FP64 118 VGPRs, one workgroup per CU, against 64 and two)
        // (FP32: unrolled by two, which fits the same 64 VGPRs and issues the
        // address arithmetic once per two steps)
        constexpr int GUNROLL = sizeof(TC) == 4 ? 2 : 1;
        TC g0 = TC(0), g1 = TC(0), g2 = TC(0), h0 = TC(0), h1 = TC(0), h2 = TC(0);
#pragma unroll GUNROLL
        for (int p = 0; p < N; ++p) {
            const TC gi = T.G[i * N + p], gj = T.G[j * N + p], gk = T.G[k * N + p];
            g0 = fma(gi, A[p + N * j + N2 * k], g0);
            g1 = fma(gj, A[i + N * p + N2 * k], g1);
            g2 = fma(gk, A[i + N * j + N2 * p], g2);
            if (adv) {
                h0 = fma(gi, B[p + N * j + N2 * k], h0);
                h1 = fma(gj, B[i + N * p + N2 * k], h1);
                h2 = fma(gk, B[i + N * j + N2 * p], h2);
            }
        }
        W0 = K00 * g0; W1 = K01 * g0; W2 = K02 * g0;
        W0 = fma(K01, g1, W0); W1 = fma(K11, g1, W1); W2 = fma(K12, g1, W2);
        W0 = fma(K02, g2, W0); W1 = fma(K12, g2, W1); W2 = fma(K22, g2, W2);
        if (adv) {
            W0 = fma(c0, vn, W0); W1 = fma(c1, vn, W1); W2 = fma(c2, vn, W2);
            cg = c0 * h0 + c1 * h1 + c2 * h2;
        }
    }
    __syncthreads();
    // P3
    if (act) {
        A[n] = W0; B[n] = W1; C[n] = W2;
        for (int q = n; q < S::F2; q += N3) curved_face_lift<N, TC>(q, Sf, Pt, T.GW, T.winv);
    }
    __syncthreads();
    if (!act) return;
    // P4
    TC d2 = TC(0);
#pragma unroll
    for (int p = 0; p < N; ++p) d2 = fma(T.G[k * N + p], C[i + N * j + N2 * p], d2);
    TC s = cg + d2;
#pragma unroll
    for (int p = 0; p < N; ++p) {
        s = fma(T.G[i * N + p], A[p + N * j + N2 * k], s);
        s = fma(T.G[j * N + p], B[i + N * p + N2 * k], s);
    }
    // lifts along the normal lines: x faces (0, 1) at (j, k), y faces (2, 3)
    // at (i, k), z faces (4, 5) at (i, j)
    s = fma(T.G[i] * T.winv[i], Sf[(0 * 6 + 0) * N2 + j + N * k], s);
    s = fma(T.G[(N - 1) * N + i] * T.winv[i], Sf[(0 * 6 + 1) * N2 + j + N * k], s);
    s = fma(T.G[j] * T.winv[j], Sf[(1 * 6 + 2) * N2 + i + N * k], s);
    s = fma(T.G[(N - 1) * N + j] * T.winv[j], Sf[(1 * 6 + 3) * N2 + i + N * k], s);
    s = fma(T.G[k] * T.winv[k], Sf[(2 * 6 + 4) * N2 + i + N * j], s);
    s = fma(T.G[(N - 1) * N + k] * T.winv[k], Sf[(2 * 6 + 5) * N2 + i + N * j], s);
    // point and tangential terms at the face nodes
    if (i == 0) s += Pt[0 * N2 + j + N * k];
    if (i == N - 1) s += Pt[1 * N2 + j + N * k];
    if (j == 0) s += Pt[2 * N2 + i + N * k];
    if (j == N - 1) s += Pt[3 * N2 + i + N * k];
    if (k == 0) s += Pt[4 * N2 + i + N * j];
    if (k == N - 1) s += Pt[5 * N2 + i + N * j];
    if (a.weighted) s *= wg;
    static_cast<TO*>(a.out)[gn] = TO(s);
}

template <int N, typename TC, typename TS>
__global__ void __launch_bounds__(256) curved_faceflux_kernel(CArgs<N, TC> a)
{
    faceflux_body<N, TC, TS>(a, blockIdx.x);
}

// FP32 variants are held to 128 VGPRs (4 wavefronts per SIMD); FP64 ones get
// what they need (2 per SIMD at N = 10).
template <typename TC>
struct VolWaves { static constexpr int value = std::is_same<TC, float>::value ? 4 : 1; };

template <int N, typename TC, typename TS, typename TO, bool GEN>
__global__ void __launch_bounds__(128) __attribute__((amdgpu_waves_per_eu(VolWaves<TC>::value)))
curved_volume_kernel(CArgs<N, TC> a, int64_t Ne)
{
    const int64_t e0 = int64_t(blockIdx.x) * VolShape<N>::EW;
    volume_body<N, TC, TS, TO, GEN>(a, e0, int(min<int64_t>(VolShape<N>::EW, Ne - e0)));
}

// Two workgroups per CU at N = 8 and 10: NT / 128 wavefronts per SIMD (N = 10:
// 8, so at most 64 VGPRs, reached without spills).  Other N need no bound
// (N = 9 would spill at its 80).
template <int N>
struct VolNodeWaves {
    static constexpr int value = (VolNodeShape<N>::NT == 512 || VolNodeShape<N>::NT == 1024) ? VolNodeShape<N>::NT / 128 : 1;
};

template <int N, typename TC, typename TS, typename TO, bool GEN>
__global__ void __launch_bounds__(VolNodeShape<N>::NT) __attribute__((amdgpu_waves_per_eu(VolNodeWaves<N>::value)))
curved_volnode_kernel(CArgs<N, TC> a, int64_t Ne)
{
    // workgroups are dispatched round-robin over the 8 XCDs: with xcd set, XCD x
    // takes a contiguous range of element blocks, so the face neighbours whose
    // u, u_t and face fluxes an element gathers (e +- 1 along x: node stride N)
    // are read by the same XCD and meet in its L2
    int64_t wg = blockIdx.x;
    const int64_t nwg = gridDim.x, full = (nwg / 8) * 8;
    if (a.xcd && wg < full) wg = (wg % 8) * (nwg / 8) + wg / 8;
    const int64_t e0 = wg * VolNodeShape<N>::EW;
    volume_node_body<N, TC, TS, TO, GEN>(a, e0, int(min<int64_t>(VolNodeShape<N>::EW, Ne - e0)));
}

template <int N, typename TC, typename TS, typename TO, bool GEN>
__global__ void __launch_bounds__(VolNodeShape<N>::NT) __attribute__((amdgpu_waves_per_eu(VolNodeWaves<N>::value)))
curved_volnode_batch_kernel(const CArgs<N, TC>* __restrict__ tab, const int2* __restrict__ work)
{
    const int2 w = work[blockIdx.x];
    const CArgs<N, TC>& a = tab[w.x];
    const int64_t Ne = a.npts / (N * N * N);
    volume_node_body<N, TC, TS, TO, GEN>(a, w.y, int(min<int64_t>(VolNodeShape<N>::EW, Ne - w.y)));
}

template <int N, typename TC, typename TS>
__global__ void __launch_bounds__(256) curved_faceflux_batch_kernel(const CArgs<N, TC>* __restrict__ tab,
                                                                    const int2* __restrict__ work)
{
    const int2 w = work[blockIdx.x];
    faceflux_body<N, TC, TS>(tab[w.x], w.y);
}

template <int N, typename TC, typename TS, typename TO, bool GEN>
__global__ void __launch_bounds__(128) __attribute__((amdgpu_waves_per_eu(VolWaves<TC>::value)))
curved_volume_batch_kernel(const CArgs<N, TC>* __restrict__ tab, const int2* __restrict__ work)
{
    // work: (instance, first element); the block ends at the instance's last element
    const int2 w = work[blockIdx.x];
    const CArgs<N, TC>& a = tab[w.x];
    const int64_t Ne = a.npts / (N * N * N);
    volume_body<N, TC, TS, TO, GEN>(a, w.y, int(min<int64_t>(VolShape<N>::EW, Ne - w.y)));
}

// Element-kernel form: one work-item per node (default) or per z-line
// (WT2_CURVED_VOL=line, the earlier form, kept for comparison).
static bool vol_node()
{
    static const bool node = [] {
        const char* ev = std::getenv("WT2_CURVED_VOL");
        return !(ev && std::strcmp(ev, "line") == 0);
    }();
    return node;
}

template <int N>
static int vol_ew()
{
    return vol_node() ? VolNodeShape<N>::EW : VolShape<N>::EW;
}

template <int N, typename TC, typename TS, typename TO, bool GEN>
static hipError_t launch_vol(const CArgs<N, TC>* a, const CArgs<N, TC>* tab, const int2* work, int64_t nblk,
                             int64_t Ne, hipStream_t st)
{
    const bool node = vol_node();
    const size_t shmem = node ? size_t(VolNodeShape<N>::EW) * VolNodeShape<N>::PER * sizeof(TC)
                              : size_t(VolShape<N>::EW) * VolShape<N>::PER * sizeof(TC);
    const int nt = node ? VolNodeShape<N>::NT : VolShape<N>::NT;
    const void* kern;
    if (node)
        kern = a ? reinterpret_cast<const void*>(curved_volnode_kernel<N, TC, TS, TO, GEN>)
                 : reinterpret_cast<const void*>(curved_volnode_batch_kernel<N, TC, TS, TO, GEN>);
    else
        kern = a ? reinterpret_cast<const void*>(curved_volume_kernel<N, TC, TS, TO, GEN>)
                 : reinterpret_cast<const void*>(curved_volume_batch_kernel<N, TC, TS, TO, GEN>);
    // raise the dynamic-LDS limit once per kernel (the first launch of a
    // configuration is always eager, outside any capture)
    static bool raised[4] = {false, false, false, false};
    const int slot = (node ? 2 : 0) + (a ? 1 : 0);
    hipError_t err = hipSuccess;
    if (shmem + sizeof(VolTables<N, TC>) + 256 > 64 * 1024 && !raised[slot]) {
        err = hipFuncSetAttribute(kern, hipFuncAttributeMaxDynamicSharedMemorySize, int(shmem));
        if (err != hipSuccess) return err;
        raised[slot] = true;
    }
    (void)hipGetLastError();
    if (a) {
        if (node)
            hipLaunchKernelGGL((curved_volnode_kernel<N, TC, TS, TO, GEN>), dim3(unsigned(nblk)), dim3(nt), shmem, st, *a, Ne);
        else
            hipLaunchKernelGGL((curved_volume_kernel<N, TC, TS, TO, GEN>), dim3(unsigned(nblk)), dim3(nt), shmem, st, *a, Ne);
    } else {
        if (node)
            hipLaunchKernelGGL((curved_volnode_batch_kernel<N, TC, TS, TO, GEN>), dim3(unsigned(nblk)), dim3(nt), shmem, st,
                               tab, work);
        else
            hipLaunchKernelGGL((curved_volume_batch_kernel<N, TC, TS, TO, GEN>), dim3(unsigned(nblk)), dim3(nt), shmem, st,
                               tab, work);
    }
    return hipGetLastError();
}

template <int N, typename TC>
static void make_cargs(const CurvedData& cd, const CurvedCall& cc, CArgs<N, TC>& a)
{
    // FP32: the matrix the FP32 even/odd table applies (z derivatives), so
    // that all three axes and the lifts use one G
    for (int i = 0; i < N * N; ++i) a.k.G[i] = std::is_same<TC, float>::value ? TC(cd.G32[i]) : TC(cd.G[i]);
    for (int i = 0; i < N; ++i) { a.k.w[i] = TC(cd.w[i]); a.k.winv[i] = TC(1.0 / cd.w[i]); }
    a.k.tau = TC(cd.tau);
    for (int f = 0; f < 6; ++f) a.k.bv[f] = TC(cc.bv[f]);
    a.k.iR = TC(cd.inv_R);
    a.k.som = cd.som;
    a.u = cc.u; a.v = cc.v;
    a.coef = reinterpret_cast<const TC*>(std::is_same<TC, float>::value ? static_cast<const void*>(cd.coef32)
                                                                          : static_cast<const void*>(cd.coef64));
    a.fbuf = reinterpret_cast<TC*>(cd.fbuf);
    a.fcoef = reinterpret_cast<const TC*>(std::is_same<TC, float>::value ? static_cast<const void*>(cd.fc32)
                                                                           : static_cast<const void*>(cd.fc64));
    // WT2_CURVED_FCNODE = 0: the element kernels read the face node's own K
    // row and c^a from the face-node table (the earlier form) instead
    static const int fcnode = [] {
        const char* ev = std::getenv("WT2_CURVED_FCNODE");
        return ev ? std::atoi(ev) : 1;
    }();
    a.fcnode = fcnode;
    // WT2_CURVED_XCD = 0: the element kernel takes element blocks in dispatch order
    static const int xcd = [] {
        const char* ev = std::getenv("WT2_CURVED_XCD");
        return ev ? std::atoi(ev) : 1;
    }();
    a.xcd = xcd;
    a.op = static_cast<const TC*>(std::is_same<TC, float>::value ? cd.op32 : cd.op64);
    a.out = cc.out ? static_cast<void*>(cc.out) : cc.Pv;
    a.npts = cd.npts;
    a.M0 = cd.M[0]; a.M1 = cd.M[1]; a.M2 = cd.M[2];
    a.per0 = cd.periodic[0]; a.per1 = cd.periodic[1]; a.per2 = cd.periodic[2];
    a.adv = cc.v != nullptr;
    a.weighted = cc.weighted;
    a.conn = cd.conn;
}

template <int N, typename TC, typename TS, typename TO>
static hipError_t launch_c(const CurvedData& cd, const CurvedCall& cc, hipStream_t st)
{
    CArgs<N, TC> a;
    make_cargs<N, TC>(cd, cc, a);
    (void)hipGetLastError();   // launch errors below are this launch's own
    hipLaunchKernelGGL((curved_faceflux_kernel<N, TC, TS>), dim3(unsigned(cd.Ne)), dim3(256), 0, st, a);
    hipError_t err = hipGetLastError();
    if (err != hipSuccess) return err;
    const int ew = vol_ew<N>();
    const int64_t nblk = (cd.Ne + ew - 1) / ew;
    if (cd.conn != nullptr) return launch_vol<N, TC, TS, TO, true>(&a, nullptr, nullptr, nblk, cd.Ne, st);
    return launch_vol<N, TC, TS, TO, false>(&a, nullptr, nullptr, nblk, cd.Ne, st);
}

template <int N>
static hipError_t launch_cn(const CurvedData& cd, const CurvedCall& cc, hipStream_t st)
{
    if (!cc.mixed) {
        if (cc.out) return launch_c<N, double, double, double>(cd, cc, st);
        return launch_c<N, double, double, double>(cd, cc, st);
    }
    if (cc.out) return launch_c<N, float, double, double>(cd, cc, st);
    if (cc.in32) return launch_c<N, float, float, float>(cd, cc, st);
    return launch_c<N, float, double, float>(cd, cc, st);
}

// ------------------------------------------------------------------ batches

template <int N, typename TC, typename TS, typename TO>
static hipError_t cbatch_build_t(const CurvedData* const* cds, const CurvedCall* ccs, int n,
                                 CurvedBatchPlan& plan)
{
    std::vector<CArgs<N, TC>> tab(n);
    std::vector<int2> work, work2;
    plan.gen = 0;
    for (int i = 0; i < n; ++i) {
        make_cargs<N, TC>(*cds[i], ccs[i], tab[i]);
        plan.gen |= cds[i]->conn != nullptr;
        for (int64_t e = 0; e < cds[i]->Ne; ++e) work.push_back(make_int2(i, int(e)));
        for (int64_t e = 0; e < cds[i]->Ne; e += vol_ew<N>()) work2.push_back(make_int2(i, int(e)));
    }
    const size_t wb2 = sizeof(int2) * work2.size();
    if (wb2 > plan.work2_cap) {
        (void)hipFree(plan.work2);
        if (hipMalloc(reinterpret_cast<void**>(&plan.work2), wb2) != hipSuccess) {
            plan.work2 = nullptr; plan.work2_cap = 0;
            return hipErrorOutOfMemory;
        }
        plan.work2_cap = wb2;
    }
    const size_t tb = sizeof(CArgs<N, TC>) * size_t(n), wb = sizeof(int2) * work.size();
    if (tb > plan.tab_cap) {
        (void)hipFree(plan.tab);
        if (hipMalloc(&plan.tab, tb) != hipSuccess) { plan.tab = nullptr; plan.tab_cap = 0; return hipErrorOutOfMemory; }
        plan.tab_cap = tb;
    }
    if (wb > plan.work_cap) {
        (void)hipFree(plan.work);
        if (hipMalloc(reinterpret_cast<void**>(&plan.work), wb) != hipSuccess) {
            plan.work = nullptr; plan.work_cap = 0;
            return hipErrorOutOfMemory;
        }
        plan.work_cap = wb;
    }
    hipError_t e = hipMemcpy(plan.tab, tab.data(), tb, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(plan.work, work.data(), wb, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(plan.work2, work2.data(), wb2, hipMemcpyHostToDevice);
    plan.N = N;
    plan.nwork = int(work.size());
    plan.nwork2 = int(work2.size());
    return e;
}

template <int N, typename TC, typename TS, typename TO>
static hipError_t cbatch_launch_t(const CurvedBatchPlan& plan, hipStream_t st)
{
    if (plan.nwork == 0) return hipSuccess;
    const CArgs<N, TC>* tab = static_cast<const CArgs<N, TC>*>(plan.tab);
    (void)hipGetLastError();
    hipLaunchKernelGGL((curved_faceflux_batch_kernel<N, TC, TS>), dim3(plan.nwork), dim3(256), 0, st, tab, plan.work);
    hipError_t err = hipGetLastError();
    if (err != hipSuccess) return err;
    if (plan.gen) return launch_vol<N, TC, TS, TO, true>(nullptr, tab, plan.work2, plan.nwork2, 0, st);
    return launch_vol<N, TC, TS, TO, false>(nullptr, tab, plan.work2, plan.nwork2, 0, st);
}

template <int N>
static hipError_t cbatch_n(const CurvedData* const* cds, const CurvedCall* ccs, int n, CurvedBatchPlan* plan,
                           const CurvedBatchPlan* lp, hipStream_t st)
{
    const int kind = plan ? (ccs[0].mixed ? (ccs[0].in32 ? 2 : 1) : 0) : lp->kind;
    if (plan) plan->kind = kind;
    switch (kind) {
    case 0: return plan ? cbatch_build_t<N, double, double, double>(cds, ccs, n, *plan)
                        : cbatch_launch_t<N, double, double, double>(*lp, st);
    case 1: return plan ? cbatch_build_t<N, float, double, float>(cds, ccs, n, *plan)
                        : cbatch_launch_t<N, float, double, float>(*lp, st);
    case 2: return plan ? cbatch_build_t<N, float, float, float>(cds, ccs, n, *plan)
                        : cbatch_launch_t<N, float, float, float>(*lp, st);
    default: return hipErrorInvalidValue;
    }
}

static hipError_t cbatch_dispatch(int N, const CurvedData* const* cds, const CurvedCall* ccs, int n,
                                  CurvedBatchPlan* plan, const CurvedBatchPlan* lp, hipStream_t st)
{
    switch (N) {
    case 3: return cbatch_n<3>(cds, ccs, n, plan, lp, st);
    case 4: return cbatch_n<4>(cds, ccs, n, plan, lp, st);
    case 5: return cbatch_n<5>(cds, ccs, n, plan, lp, st);
    case 6: return cbatch_n<6>(cds, ccs, n, plan, lp, st);
    case 7: return cbatch_n<7>(cds, ccs, n, plan, lp, st);
    case 8: return cbatch_n<8>(cds, ccs, n, plan, lp, st);
    case 9: return cbatch_n<9>(cds, ccs, n, plan, lp, st);
    case 10: return cbatch_n<10>(cds, ccs, n, plan, lp, st);
    default: return hipErrorInvalidValue;
    }
}

hipError_t curved_batch_build(const CurvedData* const* cds, const CurvedCall* ccs, int n, CurvedBatchPlan& plan)
{
    if (n <= 0) { plan.nwork = 0; return hipSuccess; }
    for (int i = 0; i < n; ++i)
        if (cds[i]->N != cds[0]->N || ccs[i].mixed != ccs[0].mixed || ccs[i].in32 != ccs[0].in32 || ccs[i].out)
            return hipErrorInvalidValue;
    return cbatch_dispatch(cds[0]->N, cds, ccs, n, &plan, nullptr, nullptr);
}

hipError_t curved_batch_launch(const CurvedBatchPlan& plan, hipStream_t s)
{
    return cbatch_dispatch(plan.N, nullptr, nullptr, 0, nullptr, &plan, s);
}

void curved_batch_free(CurvedBatchPlan& plan)
{
    (void)hipFree(plan.tab);
    (void)hipFree(plan.work);
    (void)hipFree(plan.work2);
    plan = CurvedBatchPlan();
}

hipError_t curved_operator(const CurvedData& cd, const CurvedCall& cc, hipStream_t s)
{
    if (!cd.coef64) return hipErrorNotInitialized;
    switch (cd.N) {
    case 3: return launch_cn<3>(cd, cc, s);
    case 4: return launch_cn<4>(cd, cc, s);
    case 5: return launch_cn<5>(cd, cc, s);
    case 6: return launch_cn<6>(cd, cc, s);
    case 7: return launch_cn<7>(cd, cc, s);
    case 8: return launch_cn<8>(cd, cc, s);
    case 9: return launch_cn<9>(cd, cc, s);
    case 10: return launch_cn<10>(cd, cc, s);
    default: return hipErrorInvalidValue;
    }
}

}  // namespace wt2
