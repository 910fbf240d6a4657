// wt2_ops.hpp -- the GLL differentiation matrix applied to one element line
// (N values in registers), shared by the flat pencil passes, the flat
// single-pass Laplacian and the curved element operator.
#pragma once

#include <hip/hip_runtime.h>

namespace wt2 {

// Operator table layout (TC values), built by build_op_table on the host:
//   [0, E*H)          Ae[i][p] (i < H, p < E)  even part
//   [E*H, 2*E*H)      Ao[i][p] (i < E, p < H)  odd part
//   [2EH, 2EH+N)      G[0][i] / w_i
//   [2EH+N, 2EH+2N)   G[N-1][i] / w_i
// with H = floor(N/2), E = ceil(N/2), Ae[i][p] = (G[i][p] + G[i][N-1-p]) / 2,
// Ao[i][p] = (G[i][p] - G[i][N-1-p]) / 2 and, for odd N, Ae[i][H] = G[i][H].
template <int N>
struct OpLayout {
    static constexpr int H = N / 2;
    static constexpr int E = N - N / 2;
    static constexpr int AE = 0;
    static constexpr int AO = E * H;
    static constexpr int C0 = 2 * E * H;
    static constexpr int CN = 2 * E * H + N;
    static constexpr int SIZE = 2 * E * H + 2 * N;
};

template <typename T>
__device__ __forceinline__ T shfl64(T v, int src)
{
    return __shfl(v, src, 64);
}

// y = G x on one element segment by the even/odd split of the GLL
// differentiation matrix (G[N-1-i][p] = -G[i][N-1-p]): with e_p = x_p +
// x_{N-1-p}, o_p = x_p - x_{N-1-p} (e_H = x_H for odd N), y_i = a_i + b_i and
// y_{N-1-i} = b_i - a_i, a = Ae e, b = Ao o -- half the multiply-adds of G x.
// The halves sit in the table, so no product by 1/2 is formed here; the
// result is bit for bit that of halving e and o (both halvings are exact).
template <int N, typename TC>
__device__ __forceinline__ void apply_G(const TC* __restrict__ op, const TC (&x)[N], TC (&y)[N])
{
    using L = OpLayout<N>;
    constexpr int H = L::H, E = L::E;
    TC e[E], o[H];
#pragma unroll
    for (int p = 0; p < H; ++p) {
        e[p] = x[p] + x[N - 1 - p];
        o[p] = x[p] - x[N - 1 - p];
    }
    if (E > H) e[E - 1] = x[H];
#pragma unroll
    for (int i = 0; i < E; ++i) {
        TC b = TC(0);
#pragma unroll
        for (int p = 0; p < H; ++p) b = fma(op[L::AO + i * H + p], o[p], b);
        if (i < H) {
            TC av = TC(0);
#pragma unroll
            for (int p = 0; p < E; ++p) av = fma(op[L::AE + i * E + p], e[p], av);
            y[i] = b + av;
            y[N - 1 - i] = b - av;
        } else {
            y[i] = b;   // middle node of an odd N
        }
    }
}

}  // namespace wt2
