// wt2_focurv.hip -- the first-order (Phi, Pi) system on general (multi-patch,
// curvilinear) meshes: wave3d_curved_rhs! with `metric` given
// (src/wave3d_curved.jl:218-256), its energy (:268-287) and the curvilinear
// boundary pass _apply_bc3d_curv! (src/boundaries3d.jl:131-280).
//
// Metric terms in the conservative-curl form, from the node coordinates with
// the element's own differentiation matrix G:
//   J a^i_n = (grad_xi x C^n)_i,  C^n = 1/2 (X_l grad_xi X_m - X_m grad_xi X_l),
// ((n, l, m) and (i, j, k) cyclic), which makes sum_i D_i (J a^i_n) vanish
// discretely (free stream).  They are stored with the element's orientation
// divided out (JAp = sign(J) J a, and |J|), so that left-handed elements need
// no special case.  Gradient and divergence take the split form
// 1/2 (conservative + advective) -- a skew-adjoint pair in the mass
// w_i w_j w_k |J| -- with the centred-flux SAT across element faces (the
// neighbour's node through the D4 face map) and nothing at outer faces.
//
// One right-hand side: a gradient kernel (D = grad Phi, the flux F^i and Phi'),
// a divergence kernel (Pi' = div F plus the boundary pass), and with KO six
// launches of the per-axis derivative on the six (axis, field) channels and a
// pointwise accumulation.  One workgroup per element; the element's fields in
// LDS.  FP64.
#include "wt2_focurv.hpp"
#include "wt2_internal.hpp"

#include <algorithm>
#include <vector>

namespace wt2 {

namespace {

constexpr int FC_THREADS = 256;

template <int N>
struct FcShape {
    static constexpr int N3 = N * N * N;
    static constexpr int PER = (N3 + FC_THREADS - 1) / FC_THREADS;   // nodes per thread
};

__device__ __forceinline__ int fc_fidx(int N, int a, int r, int p, int q)
{
    if (a == 0) return r + N * (p + N * q);
    if (a == 1) return p + N * (r + N * q);
    return p + N * (q + N * r);
}

// D4 map of face node (p, q) to the neighbour face's node (swap, then reverse
// p', then reverse q')
__device__ __forceinline__ void fc_orient(int ori, int N, int p, int q, int& pn, int& qn)
{
    const int a = (ori & 1) ? q : p, b = (ori & 1) ? p : q;
    pn = (ori & 2) ? N - 1 - a : a;
    qn = (ori & 4) ? N - 1 - b : b;
}

// reference derivative along axis b at node (i, j, k) of an element field in LDS
template <int N>
__device__ __forceinline__ double fc_dref(const double* G, const double* f, int b, int i, int j, int k)
{
    double s = 0.0;
    if (b == 0) {
#pragma unroll
        for (int p = 0; p < N; ++p) s = fma(G[i * N + p], f[p + N * (j + N * k)], s);
    } else if (b == 1) {
#pragma unroll
        for (int p = 0; p < N; ++p) s = fma(G[j * N + p], f[i + N * (p + N * k)], s);
    } else {
#pragma unroll
        for (int p = 0; p < N; ++p) s = fma(G[k * N + p], f[i + N * (j + N * p)], s);
    }
    return s;
}

// ---- metric terms (once per context)
template <int N>
__global__ void __launch_bounds__(FC_THREADS) fc_metric_kernel(FcDesc d)
{
    using S = FcShape<N>;
    constexpr int N3 = S::N3;
    __shared__ double G[N * N];
    __shared__ double X[3 * N3];
    __shared__ double C[3 * N3];
    const int tid = threadIdx.x;
    const int64_t e = blockIdx.x;
    for (int q = tid; q < N * N; q += FC_THREADS) G[q] = d.G[q];
    for (int q = tid; q < 3 * N3; q += FC_THREADS) X[q] = d.xyz[3 * (e * N3 + q / 3) + q % 3];
    __syncthreads();
    double JA[S::PER][9], sg[S::PER];
    for (int r = 0; r < S::PER; ++r) {
        const int q = tid + r * FC_THREADS;
        sg[r] = 1.0;
        if (q >= N3) continue;
        const int i = q % N, j = (q / N) % N, k = q / (N * N);
        double dX[3][3];
        for (int m = 0; m < 3; ++m)
            for (int b = 0; b < 3; ++b) {
                double s = 0.0;
                for (int p = 0; p < N; ++p) {
                    const int ii = b == 0 ? p : i, jj = b == 1 ? p : j, kk = b == 2 ? p : k;
                    const int iv = b == 0 ? i : (b == 1 ? j : k);
                    s = fma(G[iv * N + p], X[3 * (ii + N * (jj + N * kk)) + m], s);
                }
                dX[m][b] = s;
            }
        const double det = dX[0][0] * (dX[1][1] * dX[2][2] - dX[1][2] * dX[2][1]) -
                           dX[0][1] * (dX[1][0] * dX[2][2] - dX[1][2] * dX[2][0]) +
                           dX[0][2] * (dX[1][0] * dX[2][1] - dX[1][1] * dX[2][0]);
        sg[r] = det < 0 ? -1.0 : 1.0;
        d.Jp[e * N3 + q] = fabs(det);
    }
    for (int n = 0; n < 3; ++n) {
        const int l = (n + 1) % 3, m = (n + 2) % 3;
        __syncthreads();
        for (int q = tid; q < N3; q += FC_THREADS) {
            const int i = q % N, j = (q / N) % N, k = q / (N * N);
            for (int b = 0; b < 3; ++b) {
                double dl = 0.0, dm = 0.0;
                for (int p = 0; p < N; ++p) {
                    const int ii = b == 0 ? p : i, jj = b == 1 ? p : j, kk = b == 2 ? p : k;
                    const int iv = b == 0 ? i : (b == 1 ? j : k);
                    const int nn = ii + N * (jj + N * kk);
                    dl = fma(G[iv * N + p], X[3 * nn + l], dl);
                    dm = fma(G[iv * N + p], X[3 * nn + m], dm);
                }
                C[b * N3 + q] = 0.5 * (X[3 * q + l] * dm - X[3 * q + m] * dl);
            }
        }
        __syncthreads();
        for (int r = 0; r < S::PER; ++r) {
            const int q = tid + r * FC_THREADS;
            if (q >= N3) continue;
            const int i = q % N, j = (q / N) % N, k = q / (N * N);
            for (int ia = 0; ia < 3; ++ia) {
                const int ja = (ia + 1) % 3, ka = (ia + 2) % 3;
                JA[r][ia * 3 + n] = fc_dref<N>(G, C + ka * N3, ja, i, j, k) - fc_dref<N>(G, C + ja * N3, ka, i, j, k);
            }
        }
    }
    for (int r = 0; r < S::PER; ++r) {
        const int q = tid + r * FC_THREADS;
        if (q >= N3) continue;
        for (int c = 0; c < 9; ++c) d.JAp[int64_t(c) * d.npts + e * N3 + q] = sg[r] * JA[r][c];
    }
    // element extent along each reference axis: distance of the face centroids
    __syncthreads();
    if (tid == 0)
        for (int a = 0; a < 3; ++a) {
            double d2 = 0.0;
            for (int m = 0; m < 3; ++m) {
                double acc = 0.0;
                for (int pq = 0; pq < N * N; ++pq) {
                    const int p = pq % N, qq = pq / N;
                    acc += X[3 * fc_fidx(N, a, N - 1, p, qq) + m] - X[3 * fc_fidx(N, a, 0, p, qq) + m];
                }
                d2 += acc * acc;
            }
            d.ell[3 * e + a] = sqrt(d2) / (N * N);
        }
}

// ---- gradient, flux F^i = b^i Pi + alpha sqrtg g^ij D_j and Phi' (grad_only: D alone)
template <int N>
__global__ void __launch_bounds__(FC_THREADS) fc_grad_kernel(FcDesc d)
{
    using S = FcShape<N>;
    constexpr int N3 = S::N3;
    __shared__ double G[N * N];
    __shared__ double Ph[N3];
    __shared__ double W[3 * N3];
    const int tid = threadIdx.x;
    const int64_t e = blockIdx.x;
    const int64_t np = d.npts;
    for (int q = tid; q < N * N; q += FC_THREADS) G[q] = d.G[q];
    for (int q = tid; q < N3; q += FC_THREADS) Ph[q] = d.Phi[e * N3 + q];
    double out[S::PER][3];
    for (int n = 0; n < 3; ++n) {
        // the work-item's nodes' metric terms of this component, read once
        // for the flux form and the advective form
        double ja[S::PER][3];
        __syncthreads();
        for (int r = 0; r < S::PER; ++r) {
            const int q = tid + r * FC_THREADS;
            if (q >= N3) continue;
            for (int a = 0; a < 3; ++a) {
                ja[r][a] = d.JAp[int64_t(a * 3 + n) * np + e * N3 + q];
                W[a * N3 + q] = ja[r][a] * Ph[q];
            }
        }
        __syncthreads();
        for (int r = 0; r < S::PER; ++r) {
            const int q = tid + r * FC_THREADS;
            if (q >= N3) continue;
            const int i = q % N, j = (q / N) % N, k = q / (N * N);
            double cons = 0.0, adv = 0.0;
            for (int a = 0; a < 3; ++a) {
                cons += fc_dref<N>(G, W + a * N3, a, i, j, k);
                adv = fma(ja[r][a], fc_dref<N>(G, Ph, a, i, j, k), adv);
            }
            out[r][n] = 0.5 * (cons + adv) / d.Jp[e * N3 + q];
        }
    }
    for (int r = 0; r < S::PER; ++r) {
        const int q = tid + r * FC_THREADS;
        if (q >= N3) continue;
        const int ijk[3] = {q % N, (q / N) % N, q / (N * N)};
        const int64_t gq = e * N3 + q;
        // centred SAT: 1/2 s JAp^a (Phi_nbr - Phi) / (w_r |J|) on each face the node lies on
        bool outer = false;
        for (int f = 0; f < 6; ++f) {
            const int a = f >> 1, sd = f & 1, rr = sd ? N - 1 : 0;
            if (ijk[a] != rr) continue;
            const int2 cf = d.conn[e * 6 + f];
            if (cf.x < 0) { outer = true; continue; }
            const int fn = cf.y & 7, ori = (cf.y >> 3) & 7;
            const int p = a == 0 ? ijk[1] : ijk[0], qq = a == 2 ? ijk[1] : ijk[2];
            int pn, qn;
            fc_orient(ori, N, p, qq, pn, qn);
            const int64_t gn = int64_t(cf.x) * N3 + fc_fidx(N, fn >> 1, (fn & 1) ? N - 1 : 0, pn, qn);
            const double sc = 0.5 * (sd ? 1.0 : -1.0) * (d.Phi[gn] - Ph[q]) / (d.w[rr] * d.Jp[gq]);
            for (int n = 0; n < 3; ++n) out[r][n] = fma(sc, d.JAp[int64_t(a * 3 + n) * np + gq], out[r][n]);
        }
        // grad Phi is read again only by the energy (grad_only) and by the
        // boundary pass at outer-face nodes
        if (d.grad_only || outer)
            for (int n = 0; n < 3; ++n) d.D[n * np + gq] = out[r][n];
        if (d.grad_only) continue;
        const double* cf = d.coef;
        const double al = cf[gq], sgm = cf[np + gq];
        const double b0 = cf[2 * np + gq], b1 = cf[3 * np + gq], b2 = cf[4 * np + gq];
        const double g11 = cf[5 * np + gq], g12 = cf[6 * np + gq], g13 = cf[7 * np + gq];
        const double g22 = cf[8 * np + gq], g23 = cf[9 * np + gq], g33 = cf[10 * np + gq];
        const double Pi = d.Pi[gq];
        const double D0 = out[r][0], D1 = out[r][1], D2 = out[r][2];
        const double as = al * sgm;
        const double Fv[3] = {b0 * Pi + as * (g11 * D0 + g12 * D1 + g13 * D2),
                              b1 * Pi + as * (g12 * D0 + g22 * D1 + g23 * D2),
                              b2 * Pi + as * (g13 * D0 + g23 * D1 + g33 * D2)};
        for (int m = 0; m < 3; ++m) d.F[m * np + gq] = Fv[m];
        d.dPhi[gq] = b0 * D0 + b1 * D1 + b2 * D2 + (al / sgm) * Pi;
        // the face-normal flux of each face this node lies on, for the
        // neighbour's SAT (a contiguous face buffer instead of six gathers
        // of its metric terms and flux per face node)
        for (int f = 0; f < 6; ++f) {
            const int a = f >> 1, rr = (f & 1) ? N - 1 : 0;
            if (ijk[a] != rr) continue;
            const int p = a == 0 ? ijk[1] : ijk[0], qq = a == 2 ? ijk[1] : ijk[2];
            double Wf = 0.0;
            for (int m = 0; m < 3; ++m) Wf = fma(d.JAp[int64_t(a * 3 + m) * np + gq], Fv[m], Wf);
            d.fb[(e * 6 + f) * N * N + p + N * qq] = Wf;
        }
    }
}

// ---- divergence of F into Pi', then the curvilinear boundary pass
template <int N>
__global__ void __launch_bounds__(FC_THREADS) fc_div_kernel(FcDesc d)
{
    using S = FcShape<N>;
    constexpr int N3 = S::N3;
    __shared__ double G[N * N];
    __shared__ double Fl[3 * N3];
    __shared__ double W[3 * N3];
    const int tid = threadIdx.x;
    const int64_t e = blockIdx.x;
    const int64_t np = d.npts;
    for (int q = tid; q < N * N; q += FC_THREADS) G[q] = d.G[q];
    for (int q = tid; q < 3 * N3; q += FC_THREADS) Fl[q] = d.F[int64_t(q / N3) * np + e * N3 + q % N3];
    __syncthreads();
    // the work-item's nodes' metric terms, read once for both forms
    double ja[S::PER][9];
    for (int r = 0; r < S::PER; ++r) {
        const int q = tid + r * FC_THREADS;
        if (q >= N3) continue;
        for (int c = 0; c < 9; ++c) ja[r][c] = d.JAp[int64_t(c) * np + e * N3 + q];
        for (int a = 0; a < 3; ++a) {
            double s = 0.0;
            for (int m = 0; m < 3; ++m) s = fma(ja[r][a * 3 + m], Fl[m * N3 + q], s);
            W[a * N3 + q] = s;
        }
    }
    __syncthreads();
    for (int r = 0; r < S::PER; ++r) {
        const int q = tid + r * FC_THREADS;
        if (q >= N3) continue;
        const int ijk[3] = {q % N, (q / N) % N, q / (N * N)};
        const int i = ijk[0], j = ijk[1], k = ijk[2];
        const int64_t gq = e * N3 + q;
        double cons = 0.0, adv = 0.0;
        for (int a = 0; a < 3; ++a) {
            cons += fc_dref<N>(G, W + a * N3, a, i, j, k);
            for (int m = 0; m < 3; ++m) adv = fma(ja[r][a * 3 + m], fc_dref<N>(G, Fl + m * N3, a, i, j, k), adv);
        }
        const double Jq = d.Jp[gq];
        double out = 0.5 * (cons + adv) / Jq;
        double dF = 0.0;
        for (int f = 0; f < 6; ++f) {
            const int a = f >> 1, sd = f & 1, rr = sd ? N - 1 : 0;
            if (ijk[a] != rr) continue;
            const double s = sd ? 1.0 : -1.0;
            const int2 cf = d.conn[e * 6 + f];
            if (cf.x >= 0) {
                const int fn = cf.y & 7, ori = (cf.y >> 3) & 7, an = fn >> 1;
                const int p = a == 0 ? j : i, qq = a == 2 ? j : k;
                int pn, qn;
                fc_orient(ori, N, p, qq, pn, qn);
                const double Wn = d.fb[(int64_t(cf.x) * 6 + fn) * N * N + pn + N * qn];
                const double qo = s * W[a * N3 + q], qb = ((fn & 1) ? 1.0 : -1.0) * Wn;
                out -= 0.5 * (qo + qb) / (d.w[rr] * Jq);
                continue;
            }
            // outer face: one kind everywhere except the excision tag
            const int tag = cf.y >> 6;
            if (!d.has_bc || d.kind == WT2_BC_EXCISION || tag == d.excision_tag) continue;
            double nf[3];
            for (int m = 0; m < 3; ++m) nf[m] = s * d.JAp[int64_t(a * 3 + m) * np + gq];
            const double JF = sqrt(nf[0] * nf[0] + nf[1] * nf[1] + nf[2] * nf[2]);
            const double nx = nf[0] / JF, ny = nf[1] / JF, nz = nf[2] / JF;
            const double* cf11 = d.coef;
            const double al = cf11[gq], sgm = cf11[np + gq];
            const double g11 = cf11[5 * np + gq], g12 = cf11[6 * np + gq], g13 = cf11[7 * np + gq];
            const double g22 = cf11[8 * np + gq], g23 = cf11[9 * np + gq], g33 = cf11[10 * np + gq];
            const double av = al / sgm;
            const double a_n = al * sqrt(g11 * nx * nx + g22 * ny * ny + g33 * nz * nz + 2 * g12 * nx * ny +
                                         2 * g13 * nx * nz + 2 * g23 * ny * nz);
            const double bn = cf11[2 * np + gq] * nx + cf11[3 * np + gq] * ny + cf11[4 * np + gq] * nz;
            const double wt = JF / (d.w[rr] * Jq);
            const double Pi = d.Pi[gq];
            if (d.kind == WT2_BC_FULL_DIRICHLET) {
                const double tau = d.sigma * (fabs(a_n - bn) + fabs(a_n + bn)) * wt;
                const double gF = d.gPhi ? d.gPhi[gq] : 0.0, gP = d.gPi ? d.gPi[gq] : 0.0;
                dF -= tau * (d.Phi[gq] - gF);
                out -= tau * (Pi - gP);
            } else {
                const double qn = nx * d.D[gq] + ny * d.D[np + gq] + nz * d.D[2 * np + gq];
                const double rv = Pi + ((bn + a_n) / av) * qn;
                double g = 0.0;
                if (d.kind == WT2_BC_DIRICHLET) {
                    const double gP = d.gPi ? d.gPi[gq] : 0.0;
                    const double gd = d.gD ? nx * d.gD[gq] + ny * d.gD[np + gq] + nz * d.gD[2 * np + gq] : 0.0;
                    g = gP + ((bn + a_n) / av) * gd;
                }
                out -= d.sigma * fabs(a_n + bn) * wt * (rv - g);
            }
        }
        d.dPi[gq] = out;
        if (dF != 0.0) d.dPhi[gq] += dF;
    }
}

// ---- the KO derivative on six channels: out[c] = D_{c % 3}(in[c]) along each
// element's reference axis, (1/ell) [G u + s/(2 w_r) (u_nbr - u)] at faces with
// a neighbour, one-sided at outer faces.  Channel inputs: in0 for c < 3 and in1
// for c >= 3 when `first`, else the previous launch's outputs.
template <int N>
__global__ void __launch_bounds__(FC_THREADS) fc_D_kernel(FcDesc d, const double* in0, const double* in1,
                                                           const double* prev, double* out, int nch, int first)
{
    constexpr int N3 = FcShape<N>::N3;
    __shared__ double G[N * N];
    const int tid = threadIdx.x;
    const int64_t e = blockIdx.x / nch;
    const int c = blockIdx.x % nch;
    const int axis = c % 3;
    const int64_t np = d.npts;
    for (int q = tid; q < N * N; q += FC_THREADS) G[q] = d.G[q];
    __syncthreads();
    const double* src = first ? (c < 3 ? in0 : in1) : prev + int64_t(c) * np;
    const double il = 1.0 / d.ell[3 * e + axis];
    for (int q = tid; q < N3; q += FC_THREADS) {
        const int ijk[3] = {q % N, (q / N) % N, q / (N * N)};
        const double* fe = src + e * N3;
        double s = 0.0;
        const int iv = ijk[axis];
#pragma unroll
        for (int p = 0; p < N; ++p) {
            int t[3] = {ijk[0], ijk[1], ijk[2]};
            t[axis] = p;
            s = fma(G[iv * N + p], fe[t[0] + N * (t[1] + N * t[2])], s);
        }
        s *= il;
        for (int sd = 0; sd < 2; ++sd) {
            const int rr = sd ? N - 1 : 0;
            if (iv != rr) continue;
            const int2 cf = d.conn[e * 6 + 2 * axis + sd];
            if (cf.x < 0) continue;
            const int fn = cf.y & 7, ori = (cf.y >> 3) & 7;
            const int p = axis == 0 ? ijk[1] : ijk[0], qq = axis == 2 ? ijk[1] : ijk[2];
            int pn, qn;
            fc_orient(ori, N, p, qq, pn, qn);
            const int64_t gn = int64_t(cf.x) * N3 + fc_fidx(N, fn >> 1, (fn & 1) ? N - 1 : 0, pn, qn);
            s = fma(il * (sd ? 1.0 : -1.0) / (2.0 * d.w[rr]), src[gn] - fe[q], s);
        }
        out[int64_t(c) * np + e * N3 + q] = s;
    }
}

// Phi' += koT (sum of channels 0..2), Pi' += koT (sum of channels 3..5)
__global__ void __launch_bounds__(FC_THREADS) fc_ko_add_kernel(FcDesc d, const double* ko)
{
    const int64_t np = d.npts;
    for (int64_t q = int64_t(blockIdx.x) * FC_THREADS + threadIdx.x; q < np; q += int64_t(gridDim.x) * FC_THREADS) {
        d.dPhi[q] += d.koT * ((ko[q] + ko[np + q]) + ko[2 * np + q]);
        d.dPi[q] += d.koT * ((ko[3 * np + q] + ko[4 * np + q]) + ko[5 * np + q]);
    }
}

// energy partial sums: 1/2 [Pi^2/sqrtg + sqrtg g^ij D_i D_j] w_i w_j w_k |J|
__global__ void __launch_bounds__(RED_THREADS) fc_energy_partial_kernel(FcDesc d)
{
    __shared__ double sh[RED_THREADS];
    const int64_t np = d.npts;
    const int N = d.N;
    double acc = 0.0;
    for (int64_t q = int64_t(blockIdx.x) * RED_THREADS + threadIdx.x; q < np; q += int64_t(RED_BLOCKS) * RED_THREADS) {
        const int r = int(q % (N * N * N));
        const double Hd = d.w[r % N] * d.w[(r / N) % N] * d.w[r / (N * N)] * d.Jp[q];
        const double* c = d.coef;
        const double sg = c[np + q];
        const double d1 = d.D[q], d2 = d.D[np + q], d3 = d.D[2 * np + q], Pi = d.Pi[q];
        acc += (Pi * Pi / sg + sg * (c[5 * np + q] * d1 * d1 + c[8 * np + q] * d2 * d2 + c[10 * np + q] * d3 * d3 +
                                     2 * c[6 * np + q] * d1 * d2 + 2 * c[7 * np + q] * d1 * d3 +
                                     2 * c[9 * np + q] * d2 * d3)) * Hd / 2;
    }
    sh[threadIdx.x] = acc;
    __syncthreads();
    for (int st = RED_THREADS / 2; st > 0; st >>= 1) {
        if (int(threadIdx.x) < st) sh[threadIdx.x] += sh[threadIdx.x + st];
        __syncthreads();
    }
    if (threadIdx.x == 0) d.partial[blockIdx.x] = sh[0];
}

// partial sums of x^2 (the power iteration's norm), fixed order
__global__ void __launch_bounds__(RED_THREADS) fc_norm2_partial_kernel(const double* x, int64_t n, double* partial)
{
    __shared__ double sh[RED_THREADS];
    double acc = 0.0;
    for (int64_t q = int64_t(blockIdx.x) * RED_THREADS + threadIdx.x; q < n; q += int64_t(RED_BLOCKS) * RED_THREADS)
        acc = fma(x[q], x[q], acc);
    sh[threadIdx.x] = acc;
    __syncthreads();
    for (int st = RED_THREADS / 2; st > 0; st >>= 1) {
        if (int(threadIdx.x) < st) sh[threadIdx.x] += sh[threadIdx.x + st];
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[blockIdx.x] = sh[0];
}

__global__ void __launch_bounds__(FC_THREADS) fc_scale_kernel(double* x, int64_t n, const double* partial)
{
    __shared__ double tot;
    if (threadIdx.x == 0) {
        double s = 0.0;
        for (int q = 0; q < RED_BLOCKS; ++q) s += partial[q];
        tot = sqrt(s);
    }
    __syncthreads();
    const double inv = tot > 0 ? 1.0 / tot : 0.0;
    for (int64_t q = int64_t(blockIdx.x) * FC_THREADS + threadIdx.x; q < n; q += int64_t(gridDim.x) * FC_THREADS)
        x[q] *= inv;
}

// start vector of the power iteration: isodd(i+j+k+m) ? 1 : -1 (1-based)
__global__ void __launch_bounds__(FC_THREADS) fc_alt_kernel(double* x, int64_t n, int N)
{
    for (int64_t q = int64_t(blockIdx.x) * FC_THREADS + threadIdx.x; q < n; q += int64_t(gridDim.x) * FC_THREADS) {
        const int64_t e = q / (N * N * N);
        const int r = int(q % (N * N * N));
        const int64_t s = (r % N) + ((r / N) % N) + (r / (N * N)) + e + 4;
        x[q] = (s & 1) ? 1.0 : -1.0;
    }
}

template <int N>
hipError_t fc_launch_n(const FcDesc& d, int what, hipStream_t s)
{
    const unsigned Ne = unsigned(d.npts / (N * N * N));
    if (what == 0) {
        hipLaunchKernelGGL(fc_metric_kernel<N>, dim3(Ne), dim3(FC_THREADS), 0, s, d);
        return hipGetLastError();
    }
    hipLaunchKernelGGL(fc_grad_kernel<N>, dim3(Ne), dim3(FC_THREADS), 0, s, d);
    hipError_t err = hipGetLastError();
    if (err != hipSuccess || what == 2) {
        if (err == hipSuccess && what == 2) {
            hipLaunchKernelGGL(fc_energy_partial_kernel, dim3(RED_BLOCKS), dim3(RED_THREADS), 0, s, d);
            err = hipGetLastError();
        }
        return err;
    }
    hipLaunchKernelGGL(fc_div_kernel<N>, dim3(Ne), dim3(FC_THREADS), 0, s, d);
    err = hipGetLastError();
    if (err != hipSuccess || d.koT == 0.0) return err;
    for (int rep = 0; rep < 6; ++rep) {
        const double* prev = (rep & 1) ? d.ko : d.ko + 6 * d.npts;
        double* out = (rep & 1) ? d.ko + 6 * d.npts : d.ko;
        hipLaunchKernelGGL(fc_D_kernel<N>, dim3(Ne * 6), dim3(FC_THREADS), 0, s, d, d.Phi, d.Pi, prev, out, 6, rep == 0);
        err = hipGetLastError();
        if (err != hipSuccess) return err;
    }
    // after six launches the result is in the second half (rep 5 wrote it)
    const unsigned nb = unsigned(std::min<int64_t>((d.npts + FC_THREADS - 1) / FC_THREADS, 65536));
    hipLaunchKernelGGL(fc_ko_add_kernel, dim3(nb), dim3(FC_THREADS), 0, s, d, d.ko + 6 * d.npts);
    return hipGetLastError();
}

template <int N>
hipError_t fc_mu_n(const FcDesc& d, double* v, double* w, hipStream_t s, double* mu)
{
    const unsigned Ne = unsigned(d.npts / (N * N * N));
    const unsigned nb = unsigned(std::min<int64_t>((d.npts + FC_THREADS - 1) / FC_THREADS, 65536));
    hipLaunchKernelGGL(fc_alt_kernel, dim3(nb), dim3(FC_THREADS), 0, s, v, d.npts, N);
    double mu2 = 0.0;
    std::vector<double> part(RED_BLOCKS);
    for (int it = 0; it < 30; ++it) {
        hipLaunchKernelGGL(fc_D_kernel<N>, dim3(Ne), dim3(FC_THREADS), 0, s, d, v, v, v, w, 1, 1);
        hipLaunchKernelGGL(fc_D_kernel<N>, dim3(Ne), dim3(FC_THREADS), 0, s, d, w, w, w, v, 1, 1);
        hipLaunchKernelGGL(fc_norm2_partial_kernel, dim3(RED_BLOCKS), dim3(RED_THREADS), 0, s, v, d.npts, d.partial);
        hipError_t err = hipMemcpyAsync(part.data(), d.partial, RED_BLOCKS * sizeof(double), hipMemcpyDeviceToHost, s);
        if (err == hipSuccess) err = hipStreamSynchronize(s);
        if (err != hipSuccess) return err;
        double ss = 0.0;
        for (double p : part) ss += p;
        mu2 = std::sqrt(ss);
        if (!(mu2 > 0)) break;
        hipLaunchKernelGGL(fc_scale_kernel, dim3(nb), dim3(FC_THREADS), 0, s, v, d.npts, d.partial);
        err = hipGetLastError();
        if (err != hipSuccess) return err;
    }
    *mu = std::sqrt(mu2);
    return hipSuccess;
}

}  // namespace

hipError_t launch_fc(const FcDesc& d, int what, hipStream_t s)
{
    switch (d.N) {
    case 3: return fc_launch_n<3>(d, what, s);
    case 4: return fc_launch_n<4>(d, what, s);
    case 5: return fc_launch_n<5>(d, what, s);
    case 6: return fc_launch_n<6>(d, what, s);
    case 7: return fc_launch_n<7>(d, what, s);
    case 8: return fc_launch_n<8>(d, what, s);
    case 9: return fc_launch_n<9>(d, what, s);
    case 10: return fc_launch_n<10>(d, what, s);
    default: return hipErrorInvalidValue;
    }
}

hipError_t fc_mu(const FcDesc& d, double* v, double* w, hipStream_t s, double* mu)
{
    switch (d.N) {
    case 3: return fc_mu_n<3>(d, v, w, s, mu);
    case 4: return fc_mu_n<4>(d, v, w, s, mu);
    case 5: return fc_mu_n<5>(d, v, w, s, mu);
    case 6: return fc_mu_n<6>(d, v, w, s, mu);
    case 7: return fc_mu_n<7>(d, v, w, s, mu);
    case 8: return fc_mu_n<8>(d, v, w, s, mu);
    case 9: return fc_mu_n<9>(d, v, w, s, mu);
    case 10: return fc_mu_n<10>(d, v, w, s, mu);
    default: return hipErrorInvalidValue;
    }
}

}  // namespace wt2
