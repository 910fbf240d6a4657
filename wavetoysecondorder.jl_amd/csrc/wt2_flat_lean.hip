// wt2_flat_lean.hip -- the LDS-free pencil pass for the epilogues that add to
// the partial sum or write an output (EPI_FIRST .. EPI_OUT_RHS; every z pass of
// a folded RK4 step), in a translation unit of its own.
//
// The pass is flat_xpass_body's (wt2_flat.hip) without the RK update and the
// fold, and it reads the operator table with wave-uniform addresses straight
// from its device table, so the table sits in SGPRs (scalar loads), where the
// LDS copy of flat_xpass_body is held in VGPRs over the KO chain: N = 8 FP64,
// 93 VGPRs (5 wavefronts per SIMD) against 154 (3); z pass 129 -> 121 us at
// 32^3 x 8^3, same results bit for bit.  The RK passes keep flat_xpass_body:
// with the update and the fold the table no longer fits the SGPRs (spilled to
// VGPR lanes), and a form compiled for the RK epilogues alone took more VGPRs;
// both measured slower.  A separate translation unit because compiling these
// kernels beside flat_xpass_body's changed that kernel's register allocation
// (154 -> 176 VGPRs at N = 8 FP64).
#include "wt2_pass.hpp"

namespace wt2 {

template <int N, typename TC, typename TS, typename TP, bool XCH>
__device__ __forceinline__ void flat_xpass_lean_body(const KArgs<N, TC>& a, const int tile)
{
    constexpr int N3 = N * N * N;
    const TC* __restrict__ op = a.op;
    __shared__ TC xbuf[XCH ? 4 * 512 : 1];

    int t = tile;
    const int blk = t % a.nbk;
    t /= a.nbk;
    const int ebase = (t % a.T0) * a.slo + (t / a.T0) * a.shi;
    const int Ma = a.Ma;
    const int tid = threadIdx.x;
    int m, line, lidxL = 0, lidxR = 0;
    if (XCH) {
        m = tid / a.LT;
        line = tid - m * a.LT;
    } else {
        const int wave = tid >> 6, lane = tid & 63;
        m = lane / a.lpw;
        line = wave * a.lpw + (lane - m * a.lpw);
    }
    const bool active = (m < Ma) && (line < a.LT);
    const bool per = a.periodic != 0;
    const bool hasL = active && (m > 0 || per);
    const bool hasR = active && (m < Ma - 1 || per);
    if (XCH) {
        lidxL = hasL ? (m > 0 ? m - 1 : Ma - 1) * a.LT + line : tid;
        lidxR = hasR ? (m < Ma - 1 ? m + 1 : 0) * a.LT + line : tid;
    } else {
        const int lane = tid & 63, lw = lane - (lane / a.lpw) * a.lpw;
        lidxL = hasL ? ((m > 0 ? m - 1 : Ma - 1) * a.lpw + lw) : lane;
        lidxR = hasR ? ((m < Ma - 1 ? m + 1 : 0) * a.lpw + lw) : lane;
    }
    typename std::conditional<XCH, LdsX<TC>, ShflX<TC>>::type xch;
    if constexpr (XCH) {
        xch.buf = xbuf; xch.n = 512; xch.me = tid; xch.idxL = lidxL; xch.idxR = lidxR; xch.k = 0;
    } else {
        xch.srcL = lidxL; xch.srcR = lidxR;
    }
    const int mm = active ? m : 0;
    const int ll = active ? line : 0;
    const int64_t g0 = int64_t(ebase + mm * a.estride) * N3 + int64_t(blk) * a.CO +
                       int64_t(ll / N) * a.LA + int64_t(ll % N) * a.LB;
    const int gs = a.GS;

    TC u[N], cv[N];
    TS v[N];   // u_t exactly as held (u' of an output epilogue)
#pragma unroll
    for (int r = 0; r < N; ++r) { u[r] = TC(0); v[r] = TS(0); cv[r] = TC(0); }
    if (a.lap) load_seg_s<N, TS, TC>(static_cast<const TS*>(a.Su) + g0, gs, u);
    if (a.needv) load_seg_s<N, TS, TS>(static_cast<const TS*>(a.Sv) + g0, gs, v);
    TP* P = static_cast<TP*>(a.P);
    const int epi = a.epi;
    TC pv[N];
    if (epi != EPI_FIRST) load_seg_s<N, TP, TC>(P + g0, gs, pv);
    if (a.ko) {
        TC x[N];
#pragma unroll
        for (int r = 0; r < N; ++r) x[r] = TC(v[r]);
        ko_chain<N, TC>(x, op, a.k, xch, hasL, hasR);
#pragma unroll
        for (int r = 0; r < N; ++r) cv[r] = x[r];
    }
    if (a.lap) sipg_add<N, TC>(u, cv, op, a.k, xch, hasL, hasR, TC(v[0]), TC(v[N - 1]));
    if (!active) return;
    if (epi == EPI_FIRST) { store_seg_s<N, TP, TC>(P + g0, gs, cv); return; }
#pragma unroll
    for (int r = 0; r < N; ++r) cv[r] += pv[r];          // u_t' = P + c
    if (epi == EPI_MID) { store_seg_s<N, TP, TC>(P + g0, gs, cv); return; }
    if (epi == EPI_OUT_LAP) { store_seg_s<N, double, TC>(a.Ou + g0, gs, cv); return; }
    store_seg_s<N, double, TS>(a.Ou + g0, gs, v);       // EPI_OUT_RHS: u' = u_t
    store_seg_s<N, double, TC>(a.Ov + g0, gs, cv);
}

template <int N, typename TC, typename TS, typename TP, bool XCH>
__global__ void __launch_bounds__(512) flat_lean_kernel(KArgs<N, TC> a)
{
    flat_xpass_lean_body<N, TC, TS, TP, XCH>(a, pass_tile(a.xcd));
}

template <int N, typename TC, typename TS, typename TP, bool XCH>
__global__ void __launch_bounds__(512) flat_lean_batch_kernel(const KArgs<N, TC>* __restrict__ tab,
                                                              const int2* __restrict__ work)
{
    const int2 w = work[blockIdx.x];
    flat_xpass_lean_body<N, TC, TS, TP, XCH>(tab[w.x], w.y);
}

template <int N, typename TC, typename TS, typename TP>
hipError_t launch_xpass_lean(const KArgs<N, TC>& a, int tiles, int nthr, bool xch, hipStream_t s)
{
    (void)hipGetLastError();   // launch errors below are this launch's own
    if (xch)
        hipLaunchKernelGGL((flat_lean_kernel<N, TC, TS, TP, true>), dim3(tiles), dim3(nthr), 0, s, a);
    else
        hipLaunchKernelGGL((flat_lean_kernel<N, TC, TS, TP, false>), dim3(tiles), dim3(nthr), 0, s, a);
    return hipGetLastError();
}

template <int N, typename TC, typename TS, typename TP>
hipError_t launch_xpass_lean_batch(const KArgs<N, TC>* tab, const int2* work, int ntiles, int nthr, bool xch,
                                   hipStream_t s)
{
    (void)hipGetLastError();
    if (xch)
        hipLaunchKernelGGL((flat_lean_batch_kernel<N, TC, TS, TP, true>), dim3(ntiles), dim3(nthr), 0, s, tab, work);
    else
        hipLaunchKernelGGL((flat_lean_batch_kernel<N, TC, TS, TP, false>), dim3(ntiles), dim3(nthr), 0, s, tab,
                           work);
    return hipGetLastError();
}

#define WT2_LEAN_INST(N, TC, TS, TP)                                                                      \
    template hipError_t launch_xpass_lean<N, TC, TS, TP>(const KArgs<N, TC>&, int, int, bool, hipStream_t); \
    template hipError_t launch_xpass_lean_batch<N, TC, TS, TP>(const KArgs<N, TC>*, const int2*, int, int, bool, \
                                                               hipStream_t);
#define WT2_LEAN_INST_N(N)                 \
    WT2_LEAN_INST(N, double, double, double) \
    WT2_LEAN_INST(N, float, double, float)   \
    WT2_LEAN_INST(N, float, float, float)
WT2_LEAN_INST_N(3)
WT2_LEAN_INST_N(4)
WT2_LEAN_INST_N(5)
WT2_LEAN_INST_N(6)
WT2_LEAN_INST_N(7)
WT2_LEAN_INST_N(8)
WT2_LEAN_INST_N(9)
WT2_LEAN_INST_N(10)

}  // namespace wt2
