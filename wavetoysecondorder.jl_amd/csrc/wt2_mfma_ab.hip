// wt2_mfma_ab.hip -- A/B of the per-axis operator contractions of the
// drop-in RHS at N = 8, FP64, on VALU and on MFMA.  A standalone program
// (built next to libwt2.so, not part of it): for every element of a
// 32^3-element field it forms, along each of the three axes, g = G u and
// gg = G G u on all 64 lines of the element and accumulates gg + g per node,
// the contraction work of flat_lap1_kernel without its face terms.
//
//   valu: one work-item per line, the line in registers, G applied twice by
//         the even/odd split (apply_G of wt2_ops.hpp, the kernels' own code).
//   mfma: v_mfma_f64_16x16x4f64 with the 16 x 8 operator [G G; G] as the A
//         operand (so one product gives gg and g) and 16 lines of the element
//         from LDS as the B operand; four line tiles and two k-steps per axis.
//         Lane l holds A[l mod 16][l / 16], B[l / 16][l mod 16] and
//         D[4r + l / 16][l mod 16] in register r (checked on the device).
//
// Both stage the element in padded LDS and keep the running sum there, as the
// RHS kernel does.  Prints one JSON line: microseconds per launch (HIP events,
// 50 launches after 5), the bytes moved, and the largest difference of the two
// results relative to the largest value.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "wt2.h"
#include "wt2_ops.hpp"

namespace {

constexpr int N = 8, N3 = N * N * N, RS = N + 1, ES = RS * N * N;
typedef double d4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ int lidx(int i, int j, int k) { return i + RS * (j + N * k); }

// LDS base and stride of line (la, lb) along axis ax
__device__ __forceinline__ void line_of(int ax, int la, int lb, int& base, int& ls)
{
    if (ax == 0) { base = lidx(0, la, lb); ls = 1; }
    else if (ax == 1) { base = lidx(la, 0, lb); ls = RS; }
    else { base = lidx(la, lb, 0); ls = RS * N; }
}

__device__ __forceinline__ void stage_in(const double* __restrict__ u, double* U, int64_t e)
{
    for (int q = threadIdx.x; q < N3; q += 64)
        U[lidx(q % N, (q / N) % N, q / (N * N))] = u[e * N3 + q];
}

__device__ __forceinline__ void stage_out(double* __restrict__ out, const double* R, int64_t e)
{
    for (int q = threadIdx.x; q < N3; q += 64)
        out[e * N3 + q] = R[lidx(q % N, (q / N) % N, q / (N * N))];
}

__global__ void __launch_bounds__(64) ab_valu(const double* __restrict__ u, double* __restrict__ out,
                                              const double* __restrict__ optab)
{
    __shared__ double U[ES], R[ES], op[wt2::OpLayout<N>::SIZE];
    const int64_t e = blockIdx.x;
    for (int q = threadIdx.x; q < wt2::OpLayout<N>::SIZE; q += 64) op[q] = optab[q];
    stage_in(u, U, e);
    __syncthreads();
    const int la = threadIdx.x % N, lb = threadIdx.x / N;
    for (int ax = 2; ax >= 0; --ax) {
        int base, ls;
        line_of(ax, la, lb, base, ls);
        double x[N], g[N], gg[N];
#pragma unroll
        for (int m = 0; m < N; ++m) x[m] = U[base + m * ls];
        wt2::apply_G<N, double>(op, x, g);
        wt2::apply_G<N, double>(op, g, gg);
#pragma unroll
        for (int m = 0; m < N; ++m) {
            const double v = gg[m] + g[m];
            R[base + m * ls] = ax == 2 ? v : R[base + m * ls] + v;
        }
        __syncthreads();
    }
    stage_out(out, R, e);
}

__global__ void __launch_bounds__(64) ab_mfma(const double* __restrict__ u, double* __restrict__ out,
                                              const double* __restrict__ Acat)
{
    __shared__ double U[ES], R[ES];
    const int64_t e = blockIdx.x;
    stage_in(u, U, e);
    const int l = threadIdx.x, c = l % 16, q = l / 16;
    // A operand of the two k-steps: rows 0..7 G G, rows 8..15 G
    const double a0 = Acat[c * N + q], a1 = Acat[c * N + 4 + q];
    __syncthreads();
    for (int ax = 2; ax >= 0; --ax) {
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            const int L = 16 * t + c;
            int base, ls;
            line_of(ax, L % N, L / N, base, ls);
            const double b0 = U[base + q * ls], b1 = U[base + (4 + q) * ls];
            d4 d = {0.0, 0.0, 0.0, 0.0};
            d = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b0, d, 0, 0, 0);
            d = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b1, d, 0, 0, 0);
            // d[r] = row 4r + q of line L: gg at nodes q, q + 4 (r = 0, 1),
            // g at nodes q, q + 4 (r = 2, 3)
            const int n0 = base + q * ls, n1 = base + (q + 4) * ls;
            const double v0 = d[0] + d[2], v1 = d[1] + d[3];
            R[n0] = ax == 2 ? v0 : R[n0] + v0;
            R[n1] = ax == 2 ? v1 : R[n1] + v1;
        }
        __syncthreads();
    }
    stage_out(out, R, e);
}

int check(hipError_t e, const char* what)
{
    if (e != hipSuccess) {
        std::fprintf(stderr, "wt2_mfma_ab: %s: %s\n", what, hipGetErrorString(e));
        std::exit(1);
    }
    return 0;
}

}  // namespace

int main(int argc, char** argv)
{
    const int M = argc > 1 ? std::atoi(argv[1]) : 32;
    if (M < 1 || M > 64) { std::fprintf(stderr, "wt2_mfma_ab: elements per axis in 1..64\n"); return 2; }
    const int64_t Ne = int64_t(M) * M * M, n = Ne * N3;
    std::vector<double> xi(N), w(N), G(N * N);
    if (wt2_gll(N, xi.data(), w.data(), G.data()) != 0) { std::fprintf(stderr, "wt2_gll failed\n"); return 1; }
    // even/odd table of apply_G: Ae[i][p] = (G[i][p] + G[i][N-1-p]) / 2 (i < N/2),
    // Ao[i][p] = (G[i][p] - G[i][N-1-p]) / 2 (p < N/2)
    using L = wt2::OpLayout<N>;
    std::vector<double> tab(L::SIZE, 0.0);
    for (int i = 0; i < L::H; ++i)
        for (int p = 0; p < L::E; ++p) tab[L::AE + i * L::E + p] = 0.5 * (G[i * N + p] + G[i * N + N - 1 - p]);
    for (int i = 0; i < L::E; ++i)
        for (int p = 0; p < L::H; ++p) tab[L::AO + i * L::H + p] = 0.5 * (G[i * N + p] - G[i * N + N - 1 - p]);
    // [G G; G], 16 x 8, row-major
    std::vector<double> A(16 * N);
    for (int i = 0; i < N; ++i)
        for (int k = 0; k < N; ++k) {
            double s = 0.0;
            for (int m = 0; m < N; ++m) s += G[i * N + m] * G[m * N + k];
            A[i * N + k] = s;
            A[(N + i) * N + k] = G[i * N + k];
        }
    std::vector<double> hu(n);
    uint64_t st = 88172645463325252ull;
    for (auto& v : hu) {
        st ^= st << 13; st ^= st >> 7; st ^= st << 17;
        v = double(st >> 11) / double(1ull << 53) - 0.5;
    }
    double *du, *dv, *dm, *dt, *dA;
    check(hipMalloc(&du, n * sizeof(double)), "alloc");
    check(hipMalloc(&dv, n * sizeof(double)), "alloc");
    check(hipMalloc(&dm, n * sizeof(double)), "alloc");
    check(hipMalloc(&dt, tab.size() * sizeof(double)), "alloc");
    check(hipMalloc(&dA, A.size() * sizeof(double)), "alloc");
    check(hipMemcpy(du, hu.data(), n * sizeof(double), hipMemcpyHostToDevice), "copy");
    check(hipMemcpy(dt, tab.data(), tab.size() * sizeof(double), hipMemcpyHostToDevice), "copy");
    check(hipMemcpy(dA, A.data(), A.size() * sizeof(double), hipMemcpyHostToDevice), "copy");
    hipEvent_t e0, e1;
    check(hipEventCreate(&e0), "event");
    check(hipEventCreate(&e1), "event");
    const int warm = 5, reps = 50;
    float tv = 0.f, tm = 0.f;
    for (int round = 0; round < 2; ++round) {
        for (int r = 0; r < warm; ++r) {
            hipLaunchKernelGGL(ab_valu, dim3(unsigned(Ne)), dim3(64), 0, 0, du, dv, dt);
            hipLaunchKernelGGL(ab_mfma, dim3(unsigned(Ne)), dim3(64), 0, 0, du, dm, dA);
        }
        check(hipGetLastError(), "launch");
        check(hipEventRecord(e0, 0), "event");
        for (int r = 0; r < reps; ++r) hipLaunchKernelGGL(ab_valu, dim3(unsigned(Ne)), dim3(64), 0, 0, du, dv, dt);
        check(hipEventRecord(e1, 0), "event");
        check(hipEventSynchronize(e1), "sync");
        float t = 0.f;
        check(hipEventElapsedTime(&t, e0, e1), "elapsed");
        tv = round ? std::min(tv, t) : t;
        check(hipEventRecord(e0, 0), "event");
        for (int r = 0; r < reps; ++r) hipLaunchKernelGGL(ab_mfma, dim3(unsigned(Ne)), dim3(64), 0, 0, du, dm, dA);
        check(hipEventRecord(e1, 0), "event");
        check(hipEventSynchronize(e1), "sync");
        check(hipEventElapsedTime(&t, e0, e1), "elapsed");
        tm = round ? std::min(tm, t) : t;
    }
    check(hipGetLastError(), "launch");
    std::vector<double> ov(n), om(n);
    check(hipMemcpy(ov.data(), dv, n * sizeof(double), hipMemcpyDeviceToHost), "copy");
    check(hipMemcpy(om.data(), dm, n * sizeof(double), hipMemcpyDeviceToHost), "copy");
    double mx = 0.0, md = 0.0;
    for (int64_t i = 0; i < n; ++i) {
        mx = std::fmax(mx, std::fabs(ov[i]));
        md = std::fmax(md, std::fabs(ov[i] - om[i]));
    }
    const double us_v = 1e3 * tv / reps, us_m = 1e3 * tm / reps, bytes = 2.0 * n * sizeof(double);
    std::printf("{\"N\": %d, \"elements\": %lld, \"points\": %lld, \"valu_us\": %.2f, \"mfma_us\": %.2f, "
                "\"bytes_per_launch\": %.0f, \"valu_GBs\": %.1f, \"mfma_GBs\": %.1f, \"mfma_over_valu\": %.3f, "
                "\"max_rel_diff\": %.3e}\n",
                N, (long long)Ne, (long long)n, us_v, us_m, bytes, bytes / (us_v * 1e3), bytes / (us_m * 1e3),
                us_m / us_v, md / mx);
    (void)hipFree(du); (void)hipFree(dv); (void)hipFree(dm); (void)hipFree(dt); (void)hipFree(dA);
    return md <= 1e-12 * mx ? 0 : 3;
}
