// wt2_pass.hpp -- device pieces shared by the pencil-pass kernels of
// wt2_flat.hip and wt2_flat_lean.hip: per-axis constants and kernel
// arguments, the neighbour exchanges, the KO chain, the 1D SIPG operator on a
// register-resident element segment, and segment loads / stores.
#pragma once

#include "wt2_internal.hpp"
#include "wt2_ops.hpp"

#include <type_traits>

namespace wt2 {

// Per-axis scalars (kernel arguments).  The differentiation matrix itself is
// not a kernel argument (as kernel arguments the matrices took more SGPRs than
// exist and were spilled): it lives in a small device table, which a kernel
// either copies to LDS or reads with wave-uniform addresses.
template <int N, typename TC>
struct AxisConst {
    TC kh;                   // KO scale kos / h
    TC koL, koR, tau, ih2, gL, gR;
    TC winv0, winvN;         // 1 / w at the two face nodes
    TC hv, hiR;              // Sommerfeld faces: h and h / R
    int somL, somR;          // 1: that outer face is a Sommerfeld face
};

template <int N, typename TC>
struct KArgs {
    AxisConst<N, TC> k;
    const TC* op;      // operator table (OpLayout<N>)
    const void* Su; const void* Sv;
    void* P;
    void* So_u; void* So_v;
    double* Yu; double* Yv;
    float* Yfu; float* Yfv;   // mixed: FP32 mirror of Y
    double* Ou; double* Ov;
    double dt;
    // per-axis geometry, precomputed on the host (the kernel has no branch on
    // the axis; see launch_t for the meaning of each field)
    int Ma;            // elements along the line
    int T0, slo, shi;  // tile -> (t % T0) * slo + (t / T0) * shi = first element
    int estride;       // element-index stride along the line
    int nbk;           // line blocks per element column
    int CPE, CO, clen; // chunks per element, chunk offset per block, chunk length
    int ES, CS;        // LDS strides: per element, per chunk within an element
    int SR, LA, LB;    // segment: node stride, line -> (line/N)*LA + (line%N)*LB
    int LT, lpw;       // lines per tile, lines per wavefront
    int GS;            // node stride of a segment in memory (LDS-free kernel)
    int periodic, epi;
    int lap, ko, needv, keepv, fold;
    int xcd;           // 1: each XCD takes a contiguous range of tiles
    int ypre;          // RK passes: Y (or its FP32 mirror) goes straight to LDS at the start
};

// Workgroups are dispatched round-robin over the 8 XCDs; with xcd set, XCD x
// runs the contiguous tile range [x T/8, (x+1) T/8) instead of every 8th tile,
// so the line blocks of one element column share that XCD's L2.
__device__ __forceinline__ int pass_tile(int xcd)
{
    int wg = blockIdx.x;
    const int nwg = gridDim.x, full = (nwg / 8) * 8;
    if (xcd && wg < full) wg = (wg % 8) * (nwg / 8) + wg / 8;
    return wg;
}

// Classical RK4 on the linear autonomous system y' = A y + b (static
// background, constant boundary data) is the polynomial
// y + dt f + dt^2/2 A f + dt^3/6 A^2 f + dt^4/24 A^3 f, f = A y + b, which the
// stages evaluate in Horner form with one register besides y:
//   S1 = y + dt/4 f(y), S2 = y + dt/3 f(S1), S3 = y + dt/2 f(S2),
//   y <- y + dt f(S3).
// The stage-4 coefficient is dt; the others come from here.
__device__ __forceinline__ double rk_horner_c(int epi, double dt)
{
    return epi == EPI_RK1 ? 0.25 * dt : (epi == EPI_RK2 ? dt / 3.0 : (epi == EPI_RK3 ? 0.5 * dt : dt));
}

// Face values of the neighbouring elements on the same line: the low
// neighbour's last value and the high neighbour's first value.
// ShflX: the line's elements sit in lanes of this wavefront (shuffles).
template <typename TC>
struct ShflX {
    int srcL, srcR;
    __device__ __forceinline__ void operator()(TC last, TC first, TC& fromL, TC& fromR)
    {
        fromL = shfl64(last, srcL);
        fromR = shfl64(first, srcR);
    }
};

// LdsX: the line's elements are spread over the workgroup; values go through
// LDS, one barrier per exchange (two alternating buffers, so a buffer is only
// rewritten after every work-item has passed the barrier that follows its
// reads).  Every work-item of the workgroup must make the same calls.
template <typename TC>
struct LdsX {
    TC* buf;           // 2 buffers x (last, first) x n
    int n, me, idxL, idxR, k;
    __device__ __forceinline__ void operator()(TC last, TC first, TC& fromL, TC& fromR)
    {
        TC* b = buf + (k & 1) * 2 * n;
        b[me] = last;
        b[n + me] = first;
        __syncthreads();
        fromL = b[idxL];
        fromR = b[n + idxR];
        ++k;
    }
};

// x <- (kos D)^6 x on this lane's element segment; D = (1/h) G plus the
// centred-flux SAT (apply_D!, METHODS.md:256-264); an outer face takes the same
// SAT against a zero ghost value, which keeps H D skew on a bounded line.
template <int N, typename TC, typename X>
__device__ __forceinline__ void ko_chain(TC (&x)[N], const TC* __restrict__ op,
                                         const AxisConst<N, TC>& k, X& xch,
                                         bool hasL, bool hasR)
{
#pragma unroll 1
    for (int rep = 0; rep < 6; ++rep) {
        TC xl, xr;
        xch(x[N - 1], x[0], xl, xr);
        if (!hasL) xl = TC(0);
        if (!hasR) xr = TC(0);
        TC y[N];
        apply_G<N, TC>(op, x, y);
        const TC x0 = x[0], xN = x[N - 1];
#pragma unroll
        for (int i = 0; i < N; ++i) x[i] = k.kh * y[i];
        x[0] = fma(k.koL, xl - x0, x[0]);
        x[N - 1] = fma(k.koR, xr - xN, x[N - 1]);
    }
}

// The two end values y_0, y_{N-1} of y = G x, with the same arithmetic as
// apply_G (a neighbour's face derivative).
template <int N, typename TC>
__device__ __forceinline__ void apply_G_ends(const TC* __restrict__ op, const TC (&x)[N], TC& y0, TC& yN)
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
    TC b = TC(0), av = TC(0);
#pragma unroll
    for (int p = 0; p < H; ++p) b = fma(op[L::AO + p], o[p], b);
#pragma unroll
    for (int p = 0; p < E; ++p) av = fma(op[L::AE + p], e[p], av);
    y0 = b + av;
    yN = b - av;
}

// out += 1D SIPG Laplacian of u along this axis (README.md:57-81):
//   (1/h^2){ G G u + H^-1 [ alpha G^T B du - 1/2 B d(Gu) - tau |B| du ] },
// given g = G u and the neighbours' face values (uL, gnL: value and G u at the
// last node of the low neighbour; uR, gnR: at the first node of the high one).
// An outer Sommerfeld face (src/wave.jl:444-518) has no SIPG terms and instead
// the drag -(u_t + d_n u + u/R) / (h w) on its node, d_n u = -+(1/h) G u at the
// low / high face; v0, vN: u_t at the segment's two end nodes.  The drag rides
// in the dg slot of the face term, (1/h^2)(1/w)(+-1/2 dg - tau du) with du = 0
// and dg = 2 (+-G u - h u_t - (h/R) u) (the factor 2 and 1/2 are exact).
template <int N, typename TC>
__device__ __forceinline__ void sipg_seg(const TC (&u)[N], const TC (&g)[N], TC (&out)[N],
                                         const TC* __restrict__ op, const AxisConst<N, TC>& k,
                                         TC uL, TC gnL, TC uR, TC gnR, bool hasL, bool hasR,
                                         TC v0, TC vN)
{
    using L = OpLayout<N>;
    TC duL, dgL, aL, duR, dgR, aR;
    if (hasL)        { duL = u[0] - uL; dgL = g[0] - gnL; aL = TC(0.5); }
    else if (k.somL) { duL = TC(0); dgL = TC(2) * (g[0] - fma(k.hv, v0, k.hiR * u[0])); aL = TC(0); }
    else             { duL = u[0] - k.gL; dgL = TC(0); aL = TC(1); }
    if (hasR)        { duR = u[N - 1] - uR; dgR = g[N - 1] - gnR; aR = TC(0.5); }
    else if (k.somR) { duR = TC(0); dgR = TC(2) * (g[N - 1] + fma(k.hv, vN, k.hiR * u[N - 1])); aR = TC(0); }
    else             { duR = u[N - 1] - k.gR; dgR = TC(0); aR = TC(1); }
    const TC tL = -aL * duL, tR = aR * duR;
    TC gg[N];
    apply_G<N, TC>(op, g, gg);
#pragma unroll
    for (int i = 0; i < N; ++i) {
        TC s = fma(op[L::C0 + i], tL, gg[i]);
        s = fma(op[L::CN + i], tR, s);
        if (i == 0) s = fma(k.winv0, TC(0.5) * dgL - k.tau * duL, s);
        if (i == N - 1) s = fma(k.winvN, TC(-0.5) * dgR - k.tau * duR, s);
        out[i] = fma(s, k.ih2, out[i]);
    }
}

// The same with the neighbour values exchanged along the line.
template <int N, typename TC, typename X>
__device__ __forceinline__ void sipg_add(const TC (&u)[N], TC (&out)[N], const TC* __restrict__ op,
                                         const AxisConst<N, TC>& k, X& xch,
                                         bool hasL, bool hasR, TC v0, TC vN)
{
    TC g[N];
    apply_G<N, TC>(op, u, g);
    TC uL, gnL, uR, gnR;
    xch(u[N - 1], u[0], uL, uR);
    xch(g[N - 1], g[0], gnL, gnR);
    sipg_seg<N, TC>(u, g, out, op, k, uL, gnL, uR, gnR, hasL, hasR, v0, vN);
}

// Dynamic LDS of the LDS-free RK-pass kernel (flat_xpass_*): the neighbour
// exchange buffers (kind 2: 4 values per work-item) and, with ypre, the
// stage's Y segments of both fields (FP64, or the FP32 mirror at the mixed
// stages 1-3), each work-item's values in slots of its own.
template <int N, typename TC>
__host__ __device__ constexpr size_t xpass_xbuf_bytes(bool xch, int nthr)
{
    return xch ? (size_t(4) * size_t(nthr) * sizeof(TC) + 15) / 16 * 16 : 0;
}
template <int N>
__host__ __device__ constexpr size_t xpass_y_bytes(bool y32, int nthr)
{
    return size_t(2) * N * (y32 ? 4 : 8) * size_t(nthr);
}

typedef __attribute__((address_space(3))) void* lds_void_ptr;

// One field's Y segment (N values of type TY at stride gs from p) loaded
// straight from memory into LDS (no registers), chunk q of the field at
// ybase + q * nthr * W, work-item t's W dwords at + t * W: 16-byte chunks
// where the segment is contiguous and 16-byte aligned, else one dword per
// chunk (an FP64 value as two).  The LDS address of these loads is the
// wavefront's base (each lane's data lands at base + lane * size).
template <int N, typename TY>
__device__ __forceinline__ void y_to_lds(const TY* p, int gs, uint32_t* ybase, int nthr, int tid)
{
    const int wb = tid & ~63;
    constexpr int PER16 = 16 / int(sizeof(TY));
    if (gs == 1 && N % PER16 == 0) {
#pragma unroll
        for (int c = 0; c < N / PER16; ++c)
            __builtin_amdgcn_global_load_lds(static_cast<const void*>(p + PER16 * c),
                                             (lds_void_ptr)(ybase + (c * nthr + wb) * 4), 16, 0, 0);
    } else if constexpr (sizeof(TY) == 8) {
#pragma unroll
        for (int r = 0; r < N; ++r)
#pragma unroll
            for (int h = 0; h < 2; ++h)
                __builtin_amdgcn_global_load_lds(static_cast<const void*>(reinterpret_cast<const uint32_t*>(p + r * gs) + h),
                                                 (lds_void_ptr)(ybase + (2 * r + h) * nthr + wb), 4, 0, 0);
    } else {
#pragma unroll
        for (int r = 0; r < N; ++r)
            __builtin_amdgcn_global_load_lds(static_cast<const void*>(p + r * gs),
                                             (lds_void_ptr)(ybase + r * nthr + wb), 4, 0, 0);
    }
}

// This work-item's values of a segment placed by y_to_lds (after the loads
// have completed: s_waitcnt vmcnt(0)).
template <int N, typename TY>
__device__ __forceinline__ void y_from_lds(const uint32_t* ybase, int gs, int nthr, int tid, double (&y)[N])
{
    constexpr int PER16 = 16 / int(sizeof(TY));
    if (gs == 1 && N % PER16 == 0) {
#pragma unroll
        for (int c = 0; c < N / PER16; ++c) {
            const uint4 q = *reinterpret_cast<const uint4*>(ybase + (c * nthr + tid) * 4);
            if constexpr (sizeof(TY) == 8) {
                y[2 * c] = __hiloint2double(int(q.y), int(q.x));
                y[2 * c + 1] = __hiloint2double(int(q.w), int(q.z));
            } else {
                y[4 * c] = double(__uint_as_float(q.x)); y[4 * c + 1] = double(__uint_as_float(q.y));
                y[4 * c + 2] = double(__uint_as_float(q.z)); y[4 * c + 3] = double(__uint_as_float(q.w));
            }
        }
    } else if constexpr (sizeof(TY) == 8) {
#pragma unroll
        for (int r = 0; r < N; ++r)
            y[r] = __hiloint2double(int(ybase[(2 * r + 1) * nthr + tid]), int(ybase[2 * r * nthr + tid]));
    } else {
#pragma unroll
        for (int r = 0; r < N; ++r) y[r] = double(__uint_as_float(ybase[r * nthr + tid]));
    }
}

// FP64 Y of a kind-2 pass with strided segments (y, z) and even N: adjacent
// work-items t, t ^ 1 hold adjacent lines of one element, whose values at a
// node are adjacent in memory, so one 16-byte load covers both.  Load q of
// the field covers nodes 2q (even work-items load their pair's two values at
// node 2q) and 2q + 1 (odd work-items, at node 2q + 1); work-item t's 16
// bytes land at ybase + (q * nthr + t) * 4.  pbase: the pair's first value
// (this work-item's line minus t & 1), any in-bounds line for inactive pairs.
template <int N>
__device__ __forceinline__ void y_to_lds_pair(const double* pbase, int gs, uint32_t* ybase, int nthr, int tid)
{
    const int wb = tid & ~63, h = tid & 1;
#pragma unroll
    for (int q = 0; q < N / 2; ++q)
        __builtin_amdgcn_global_load_lds(static_cast<const void*>(pbase + (2 * q + h) * gs),
                                         (lds_void_ptr)(ybase + (q * nthr + wb) * 4), 16, 0, 0);
}

template <int N>
__device__ __forceinline__ void y_from_lds_pair(const uint32_t* ybase, int nthr, int tid, double (&y)[N])
{
    const int pb = tid & ~1, h = tid & 1;
#pragma unroll
    for (int q = 0; q < N / 2; ++q)
#pragma unroll
        for (int hr = 0; hr < 2; ++hr) {
            const uint32_t* w = ybase + (q * nthr + pb + hr) * 4 + 2 * h;
            y[2 * q + hr] = __hiloint2double(int(w[1]), int(w[0]));
        }
}

// Segment load / store: N values at stride gs (gs == 1: widest aligned vector).
template <int N, typename TS, typename TC>
__device__ __forceinline__ void load_seg_s(const TS* __restrict__ p, int gs, TC (&x)[N]);
template <int N, typename TD, typename TC>
__device__ __forceinline__ void store_seg_s(TD* __restrict__ p, int gs, const TC (&x)[N]);

template <int N, typename TS, typename TC>
__device__ __forceinline__ void load_seg(const TS* __restrict__ p, TC (&x)[N])
{
    if constexpr (std::is_same<TS, double>::value && N % 2 == 0) {
        const double2* q = reinterpret_cast<const double2*>(p);
#pragma unroll
        for (int r = 0; r < N / 2; ++r) {
            const double2 t = q[r];
            x[2 * r] = TC(t.x);
            x[2 * r + 1] = TC(t.y);
        }
    } else if constexpr (std::is_same<TS, float>::value && N % 4 == 0) {
        const float4* q = reinterpret_cast<const float4*>(p);
#pragma unroll
        for (int r = 0; r < N / 4; ++r) {
            const float4 t = q[r];
            x[4 * r] = TC(t.x); x[4 * r + 1] = TC(t.y); x[4 * r + 2] = TC(t.z); x[4 * r + 3] = TC(t.w);
        }
    } else {
#pragma unroll
        for (int r = 0; r < N; ++r) x[r] = TC(p[r]);
    }
}

template <int N, typename TD, typename TC>
__device__ __forceinline__ void store_seg(TD* __restrict__ p, const TC (&x)[N])
{
    if constexpr (std::is_same<TD, double>::value && N % 2 == 0) {
        double2* q = reinterpret_cast<double2*>(p);
#pragma unroll
        for (int r = 0; r < N / 2; ++r) q[r] = make_double2(double(x[2 * r]), double(x[2 * r + 1]));
    } else if constexpr (std::is_same<TD, float>::value && N % 4 == 0) {
        float4* q = reinterpret_cast<float4*>(p);
#pragma unroll
        for (int r = 0; r < N / 4; ++r)
            q[r] = make_float4(float(x[4 * r]), float(x[4 * r + 1]), float(x[4 * r + 2]), float(x[4 * r + 3]));
    } else {
#pragma unroll
        for (int r = 0; r < N; ++r) p[r] = TD(x[r]);
    }
}

template <int N, typename TS, typename TC>
__device__ __forceinline__ void load_seg_s(const TS* __restrict__ p, int gs, TC (&x)[N])
{
    if (gs == 1) { load_seg<N, TS, TC>(p, x); return; }
#pragma unroll
    for (int r = 0; r < N; ++r) x[r] = TC(p[r * gs]);
}

template <int N, typename TD, typename TC>
__device__ __forceinline__ void store_seg_s(TD* __restrict__ p, int gs, const TC (&x)[N])
{
    if (gs == 1) { store_seg<N, TD, TC>(p, x); return; }
#pragma unroll
    for (int r = 0; r < N; ++r) p[r * gs] = TD(x[r]);
}

// The pass of the partial-sum and output epilogues (EPI_FIRST .. EPI_OUT_RHS)
// with the operator table in SGPRs (wt2_flat_lean.hip): one context, or a batch
// group's descriptor table and packed work list.  xch: neighbour values
// through LDS (kernel kind 2), else by shuffles (kind 1).
template <int N, typename TC, typename TS, typename TP>
hipError_t launch_xpass_lean(const KArgs<N, TC>& a, int tiles, int nthr, bool xch, hipStream_t s);
template <int N, typename TC, typename TS, typename TP>
hipError_t launch_xpass_lean_batch(const KArgs<N, TC>* tab, const int2* work, int ntiles, int nthr, bool xch,
                                   hipStream_t s);

}  // namespace wt2
