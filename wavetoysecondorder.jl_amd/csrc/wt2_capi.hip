// wt2_capi.hip -- the extern "C" entry points of libwt2.so (declared in
// include/wt2.h): context set-up (operators, mu, coercivity check, device
// workspace) and the host-side sequencing of the pencil passes.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <map>
#include <mutex>
#include <cstdlib>
#include <cstring>
#include <initializer_list>
#include <string>
#include <vector>

#include "wt2_curved.hpp"
#include "wt2_focurv.hpp"
#include "wt2_internal.hpp"

namespace wt2 {

static thread_local std::string g_err;
void set_error(const std::string& msg) { g_err = msg; }

// ----------------------------------------------------------- GLL operators

static void legendre(int n, double x, double* pn, double* pnm1)
{
    double p0 = 1.0, p1 = x;
    if (n == 0) { *pn = 1.0; *pnm1 = 0.0; return; }
    for (int m = 2; m <= n; ++m) {
        const double p2 = ((2.0 * m - 1.0) * x * p1 - (m - 1.0) * p0) / m;
        p0 = p1;
        p1 = p2;
    }
    *pn = p1;
    *pnm1 = p0;
}

// GLL nodes on [0,1] (Newton on the Lobatto polynomial), weights summing to 1,
// differentiation matrix by the barycentric formula with negative-sum diagonal.
static int gll(int N, double* xi, double* w, double* G)
{
    if (N < 2 || N > 16) return -1;
    const int n = N - 1;
    double x[16];
    for (int k = 0; k < N; ++k) {
        double xk = -std::cos(M_PI * k / n);
        for (int it = 0; it < 100; ++it) {
            double pn, pnm1;
            legendre(n, xk, &pn, &pnm1);
            const double dx = (xk * pn - pnm1) / ((n + 1.0) * pn);
            xk -= dx;
            if (std::fabs(dx) < 1e-16) break;
        }
        x[k] = xk;
    }
    x[0] = -1.0;
    x[n] = 1.0;
    for (int k = 0; k <= n / 2; ++k) {
        const double s = 0.5 * (x[k] - x[n - k]);
        x[k] = s;
        x[n - k] = -s;
    }
    if (N % 2 == 1) x[n / 2] = 0.0;
    for (int k = 0; k < N; ++k) {
        double pn, pnm1;
        legendre(n, x[k], &pn, &pnm1);
        w[k] = 1.0 / (n * (n + 1.0) * pn * pn);
        xi[k] = 0.5 * (x[k] + 1.0);
    }
    double lam[16];
    for (int j = 0; j < N; ++j) {
        double p = 1.0;
        for (int k = 0; k < N; ++k)
            if (k != j) p *= (xi[j] - xi[k]);
        lam[j] = 1.0 / p;
    }
    for (int i = 0; i < N; ++i) {
        double diag = 0.0;
        for (int j = 0; j < N; ++j) {
            if (i == j) continue;
            const double v = (lam[j] / lam[i]) / (xi[i] - xi[j]);
            G[i * N + j] = v;
            diag -= v;
        }
        G[i * N + i] = diag;
    }
    return 0;
}

// 1D SAT-coupled derivative on a line of M elements (host, for mu).
static void apply_D_1d(int N, int M, bool periodic, double h, const double* G, const double* w,
                       const double* x, double* y)
{
    for (int e = 0; e < M; ++e) {
        const double* xe = x + e * N;
        double* ye = y + e * N;
        for (int i = 0; i < N; ++i) {
            double s = 0.0;
            for (int p = 0; p < N; ++p) s += G[i * N + p] * xe[p];
            ye[i] = s / h;
        }
        const bool hl = e > 0 || periodic, hr = e < M - 1 || periodic;
        if (hl) {
            const double xl = x[((e - 1 + M) % M) * N + N - 1];
            ye[0] += -1.0 / (2.0 * w[0] * h) * (xl - xe[0]);
        }
        if (hr) {
            const double xr = x[((e + 1) % M) * N];
            ye[N - 1] += 1.0 / (2.0 * w[N - 1] * h) * (xr - xe[N - 1]);
        }
    }
}

// Largest eigenvalue of sym(H L) for the 1D SIPG operator (homogeneous
// Dirichlet or periodic ends), in unit-element units, by bisection on the
// inertia of sym(H L) - s I (LDL^T without pivoting on a symmetric matrix is
// the Sylvester count).  Returns the bracketed max eigenvalue.
static double sipg_max_sym_eig(int N, int M, bool periodic, double tau, const double* G,
                               const double* w)
{
    const int L = M * N;
    std::vector<double> A(size_t(L) * L, 0.0), x(L), y(L);
    auto seg = [&](const double* u, const double* ul, const double* ur, double* out) {
        double g[16];
        for (int i = 0; i < N; ++i) {
            double s = 0.0;
            for (int p = 0; p < N; ++p) s += G[i * N + p] * u[p];
            g[i] = s;
        }
        double duL, dgL, aL, duR, dgR, aR;
        if (ul) {
            double gn = 0.0;
            for (int p = 0; p < N; ++p) gn += G[(N - 1) * N + p] * ul[p];
            duL = u[0] - ul[N - 1]; dgL = g[0] - gn; aL = 0.5;
        } else { duL = u[0]; dgL = 0.0; aL = 1.0; }
        if (ur) {
            double gn = 0.0;
            for (int p = 0; p < N; ++p) gn += G[p] * ur[p];
            duR = u[N - 1] - ur[0]; dgR = g[N - 1] - gn; aR = 0.5;
        } else { duR = u[N - 1]; dgR = 0.0; aR = 1.0; }
        for (int i = 0; i < N; ++i) {
            double s = 0.0;
            for (int p = 0; p < N; ++p) s += G[i * N + p] * g[p];
            s += (aL * G[i] * (-duL) + aR * G[(N - 1) * N + i] * duR) / w[i];
            if (i == 0) s += (0.5 * dgL - tau * duL) / w[0];
            if (i == N - 1) s += (-0.5 * dgR - tau * duR) / w[N - 1];
            out[i] = s;
        }
    };
    for (int col = 0; col < L; ++col) {
        std::fill(x.begin(), x.end(), 0.0);
        x[col] = 1.0;
        for (int e = 0; e < M; ++e) {
            const double* ul = e > 0 ? &x[(e - 1) * N] : (periodic ? &x[(M - 1) * N] : nullptr);
            const double* ur = e < M - 1 ? &x[(e + 1) * N] : (periodic ? &x[0] : nullptr);
            seg(&x[e * N], ul, ur, &y[e * N]);
        }
        for (int row = 0; row < L; ++row) A[size_t(row) * L + col] = y[row] * w[row % N];
    }
    for (int r = 0; r < L; ++r)
        for (int c = r + 1; c < L; ++c) {
            const double s = 0.5 * (A[size_t(r) * L + c] + A[size_t(c) * L + r]);
            A[size_t(r) * L + c] = A[size_t(c) * L + r] = s;
        }
    double bound = 0.0;
    for (int r = 0; r < L; ++r) {
        double s = 0.0;
        for (int c = 0; c < L; ++c) s += std::fabs(A[size_t(r) * L + c]);
        bound = std::max(bound, s);
    }
    // number of eigenvalues > s via LDL^T of (A - s I)
    std::vector<double> B(size_t(L) * L);
    auto count_above = [&](double s) {
        B = A;
        for (int r = 0; r < L; ++r) B[size_t(r) * L + r] -= s;
        int pos = 0;
        for (int k = 0; k < L; ++k) {
            double d = B[size_t(k) * L + k];
            if (d == 0.0) d = 1e-300;
            if (d > 0) ++pos;
            for (int i = k + 1; i < L; ++i) {
                const double f = B[size_t(i) * L + k] / d;
                if (f == 0.0) continue;
                for (int j = k + 1; j <= i; ++j) B[size_t(i) * L + j] -= f * B[size_t(j) * L + k];
            }
        }
        return pos;
    };
    double lo = -bound - 1.0, hi = bound + 1.0;
    for (int it = 0; it < 60; ++it) {
        const double mid = 0.5 * (lo + hi);
        if (count_above(mid) > 0) lo = mid; else hi = mid;
        if (hi - lo < 1e-9 * (1.0 + std::fabs(hi))) break;
    }
    return hi;
}

}  // namespace wt2

using namespace wt2;


struct wt2_ctx {
    wt2_mesh_desc d;
    int N, Ne;
    int64_t npts;
    double xi[16], w[16], G[256];
    float G32[256];   // the FP32 G of the FP32 operator table (build_op_table32)
    double h[3];
    double mu, inv_mu5, kos, max_sym_eig, h_min;
    int curved;
    bool mesh_som;  // general mesh with Sommerfeld-tagged outer faces
    int fold;       // RK4 schedule with the first pass of each stage folded into the previous stage
    int fmid;       // folded schedule: the axis of each stage's middle pass
    int fe[4];      // folded schedule: the axis of stage s's last (RK + fold) pass
    int nb[3];
    int direct[3];  // per axis: LDS-free pass kernel
    int device;
    hipStream_t own_stream, stream;
    // workspace: each block holds two fields of npts elements (FP64, or FP32
    // in a mixed context, where a block is also one FP64 field of scratch)
    void* Sb;       // stage input
    void* Pb;       // partial sums
    float* Yf;      // mixed: FP32 mirror of the RK state (2 fields), else null
    double* Y[2];   // device staging for host-pointer calls (lazy)
    double* O[2];   // device staging for host-pointer outputs (lazy)
    double* Lu;     // energy scratch
    double* red;    // reduction partials + result
    double* Edev;   // energy history (device)
    void* op64;     // operator table (FP64)
    void* op32;     // operator table (FP32, each entry rounded)
    void* op32c;    // FP32 table of the curved operator: the same matrix as G32
    int Ecap;
    CurvedData cd;
    double* fo_D;   // first-order path: DPhi_1..3 (lazy)
    // first-order path on general meshes (lazy): metric terms, |J|, element
    // extents, flux, KO channels, G and w on the device, mu of the KO derivative
    double* fc_JAp; double* fc_Jp; double* fc_ell; double* fc_F; double* fc_ko; double* fc_fb;
    double* fc_ops;
    double fc_mu;
    double* fo_rk;  // first-order RK4 registers: S, A, k (2 fields each, lazy)
    int64_t bytes;
    // captured RK4 step
    hipGraphExec_t gexec;
    bool gvalid;
    double g_dt;
    const void* g_u; const void* g_v;
};

#define HIPCHK(x)                                                                     \
    do {                                                                              \
        hipError_t _e = (x);                                                          \
        if (_e != hipSuccess) {                                                       \
            set_error(std::string("HIP error: ") + hipGetErrorString(_e) + " at " #x); \
            return WT2_ERR_HIP;                                                       \
        }                                                                             \
    } while (0)

static int alloc_dev(wt2_ctx* c, void** p, size_t bytes)
{
    if (hipMalloc(p, bytes) != hipSuccess) {
        *p = nullptr;
        set_error("device allocation of " + std::to_string(bytes) + " bytes failed");
        return WT2_ERR_OOM;
    }
    c->bytes += int64_t(bytes);
    return WT2_OK;
}

static bool is_mixed(const wt2_ctx* c) { return c->d.precision == WT2_MIXED; }
static size_t elem_size(const wt2_ctx* c) { return is_mixed(c) ? sizeof(float) : sizeof(double); }
static void* fld(const wt2_ctx* c, void* blk, int f)
{
    return static_cast<char*>(blk) + size_t(f) * size_t(c->npts) * elem_size(c);
}

static void fill_axis_ops(const wt2_ctx* c, int axis, const double bv[6], AxisOps& o)
{
    const int N = c->N;
    const double h = c->h[axis];
    o.kh = c->kos / h;
    o.winv0 = 1.0 / c->w[0];
    o.winvN = 1.0 / c->w[N - 1];
    o.koL = -c->kos / (2.0 * c->w[0] * h);
    o.koR = c->kos / (2.0 * c->w[N - 1] * h);
    o.tau = c->d.tau;
    o.ih2 = 1.0 / (h * h);
    o.gL = bv[2 * axis];
    o.gR = bv[2 * axis + 1];
    o.hv = h;
    o.hiR = h * c->d.sommerfeld_inv_R;
    o.somL = !c->d.periodic[axis] && c->d.face_kind[2 * axis] == SOMMERFELD_TAG;
    o.somR = !c->d.periodic[axis] && c->d.face_kind[2 * axis + 1] == SOMMERFELD_TAG;
}

// A lap pass along this axis reads u_t at the axis's Sommerfeld faces.
static bool som_axis(const wt2_ctx* c, int axis)
{
    return !c->d.periodic[axis] &&
           (c->d.face_kind[2 * axis] == SOMMERFELD_TAG || c->d.face_kind[2 * axis + 1] == SOMMERFELD_TAG);
}

// A pass-kernel choice for axis a: kind 0 = LDS-staged, 1 = LDS-free with
// shuffles, 2 = LDS-free with the neighbour exchange through LDS.
static bool tiles_ok(const wt2_ctx* c, int a, int nb, int kind)
{
    const int N = c->N, Ma = c->d.M[a];
    if (nb < 1 || N % nb != 0 || kind < 0 || kind > 2) return false;
    if (kind == 2) return N * nb * Ma <= 512;
    const int lpw = 64 / Ma;
    return (N * nb + lpw - 1) / lpw <= (kind == 1 || !is_mixed(c) ? 8 : 16);
}

// Default element planes per tile: the smallest divisor nb of N that gives a
// workgroup of at least `want` wavefronts.
static int default_nb(int N, int Ma, int want)
{
    const int lpw = 64 / Ma;
    int best = 1;
    for (int nb = 1; nb <= N; ++nb) {
        if (N % nb) continue;
        const int waves = (N * nb + lpw - 1) / lpw;
        if (waves > 8) break;       // at most 512 threads per workgroup
        best = nb;
        if (waves >= want) return nb;
    }
    return best;
}

static PassDesc base_pass(const wt2_ctx* c, int axis, const double bv[6])
{
    PassDesc p;
    std::memset(&p, 0, sizeof(p));
    p.N = c->N;
    p.axis = axis;
    p.periodic = c->d.periodic[axis];
    for (int a = 0; a < 3; ++a) p.M[a] = c->d.M[a];
    p.nb = c->nb[axis];
    p.direct = c->direct[axis];
    p.op64 = c->op64;
    p.op32 = c->op32;
    fill_axis_ops(c, axis, bv, p.ops);
    return p;
}

static int pass_rc(hipError_t e, const char* what)
{
    if (e == hipSuccess) return WT2_OK;
    set_error(std::string(what) + ": " + hipGetErrorString(e));
    return WT2_ERR_HIP;
}

// ------------------------------------------------------------ sequencing

// Drop-in Laplacian out = L u (flat) or wgt (Ltilde u + A v) (curved).
// fp64 forces FP64 arithmetic in a mixed context (energy).
static int run_lap(wt2_ctx* c, const double* u, const double* v, double* out, const double bv[6],
                   bool fp64, bool with_adv)
{
    const bool mixed = is_mixed(c) && !fp64;
    if (c->curved) {
        CurvedCall cc;
        std::memset(&cc, 0, sizeof(cc));
        cc.u = u; cc.v = with_adv ? v : nullptr; cc.out = out;
        cc.mixed = mixed; cc.weighted = 1;
        for (int f = 0; f < 6; ++f) cc.bv[f] = bv[f];
        return pass_rc(curved_operator(c->cd, cc, c->stream), "curved operator");
    }
    // one launch: every element with its face neighbours' lines
    Lap1Desc d;
    std::memset(&d, 0, sizeof(d));
    d.N = c->N;
    d.variant = mixed ? VAR_MIX_IN64 : VAR_F64;
    for (int a = 0; a < 3; ++a) {
        d.M[a] = c->d.M[a];
        d.periodic[a] = c->d.periodic[a];
        fill_axis_ops(c, a, bv, d.ops[a]);
    }
    d.u = u;
    d.v = with_adv ? v : nullptr;
    d.out = out;
    d.op64 = c->op64;
    d.op32 = c->op32;
    return pass_rc(launch_flat_lap1(d, c->stream), "flat Laplacian");
}

// The launches of one full RHS evaluation of u' = u_t,
// u_t' = lap(u, u_t) + KO(u_t): an optional curved operator (phase 0) and the
// pencil passes z, y, x (phases 1..3; the x pass carries the final epilogue).
// stage 0: output (du, dv) to (Ou, Ov); stage 1..4: classical RK4 stage on the
// context's registers with state Y (stage 1 reads Y, later stages S).
struct StageItems {
    bool has_curved = false;
    CurvedCall cc;
    bool has_pass[3] = {false, false, false};
    PassDesc pass[3];
};

// The folded RK4 schedule (flat contexts whose e-axis passes are LDS-free
// kernels): stage s runs the middle axis m (adding to P, which already holds
// the folded contribution of axis f_s) and then e_s with the RK update and
// the fold: e_s's contribution of stage s+1, computed on the new stage input
// in registers.  f_s = e_(s-1), f_1 = e_4 (the previous step's stage 4, or
// the prologue at the start of a run).  8 passes per step instead of 12.
// Orders named m e_1 e_2 (e alternates): measured at 32^3 x 8^3 with the
// current kernels (300 steps, same box): FP64 zyx 1.219, zxy 1.227, xzy
// 1.266, xyz 1.268-1.284, yzx 1.303, yxz 1.318 ms per step; mixed xzy 0.769,
// xyz 0.784, zxy 0.789, yxz 0.797, yzx 0.802, zyx 0.803.  So FP64 contexts
// take zyx (m = z, e = y, x, y, x) and mixed ones xzy (m = x, e = z, y, z,
// y); WT2_FOLD_ORDER = one of the six names forces that order for both.
static void fold_order(wt2_ctx* c)
{
    static const int forced = [] {
        const char* ev = std::getenv("WT2_FOLD_ORDER");
        if (ev && std::strcmp(ev, "zyx") == 0) return 1;
        if (ev && std::strcmp(ev, "xyz") == 0) return 0;
        if (ev && std::strcmp(ev, "yxz") == 0) return 2;
        if (ev && std::strcmp(ev, "yzx") == 0) return 3;
        if (ev && std::strcmp(ev, "zxy") == 0) return 4;
        if (ev && std::strcmp(ev, "xzy") == 0) return 5;
        return -1;
    }();
    const int order = forced >= 0 ? forced : (is_mixed(c) ? 5 : 1);
    // m, then e_1..e_4 (e alternates between the two other axes)
    static const int tab[6][5] = {{0, 1, 2, 1, 2}, {2, 1, 0, 1, 0}, {1, 0, 2, 0, 2},
                                  {1, 2, 0, 2, 0}, {2, 0, 1, 0, 1}, {0, 2, 1, 2, 1}};
    c->fmid = tab[order][0];
    for (int s = 0; s < 4; ++s) c->fe[s] = tab[order][1 + s];
}

static bool fold_on(const wt2_ctx* c)
{
    return c->fold && !c->curved && c->direct[c->fe[0]] != 0 && c->direct[c->fe[1]] != 0;
}

static bool fold_axis(const wt2_ctx* c, int a)
{
    return a == c->fe[0] || a == c->fe[1];
}

static void stage_items(wt2_ctx* c, int stage, const double* u, const double* v, double* Ou, double* Ov,
                        double* Yu, double* Yv, double dt, StageItems& it, bool allow_fold = true)
{
    if (stage >= 1 && allow_fold && fold_on(c)) {
        const bool mixed = is_mixed(c);
        const bool ko = c->d.eps_ko > 0 && c->kos > 0;
        const double* bv = c->d.bdry_values;
        // stage 1 reads Y (mixed: its FP32 mirror)
        const void* Su = stage != 1 ? fld(c, c->Sb, 0) : mixed ? static_cast<const void*>(c->Yf) : Yu;
        const void* Sv = stage != 1 ? fld(c, c->Sb, 1) : mixed ? static_cast<const void*>(c->Yf + c->npts) : Yv;
        const int variant = !mixed ? VAR_F64 : VAR_MIX_IN32;
        it = StageItems();
        PassDesc z = base_pass(c, c->fmid, bv);
        z.variant = variant;
        z.lap = 1; z.ko = ko; z.needv = ko || som_axis(c, c->fmid); z.keepv = 0;
        z.Su = Su; z.Sv = Sv; z.P = c->Pb;
        z.epi = EPI_MID;
        it.pass[0] = z;
        it.has_pass[0] = true;
        PassDesc e = base_pass(c, c->fe[stage - 1], bv);
        e.variant = variant;
        e.lap = 1; e.ko = ko; e.needv = 1; e.keepv = 1; e.fold = 1;
        e.Su = Su; e.Sv = Sv; e.P = c->Pb;
        e.epi = EPI_RK1 + (stage - 1);
        e.So_u = fld(c, c->Sb, 0); e.So_v = fld(c, c->Sb, 1);
        e.Yu = Yu; e.Yv = Yv;
        e.Yfu = mixed ? c->Yf : nullptr; e.Yfv = mixed ? c->Yf + c->npts : nullptr;
        e.dt = dt;
        it.pass[1] = e;
        it.has_pass[1] = true;
        return;
    }
    const bool mixed = is_mixed(c);
    const bool ko = c->d.eps_ko > 0 && c->kos > 0;
    const double* bv = c->d.bdry_values;
    // stage 0 reads (u, v); stage 1 reads Y (mixed: its FP32 mirror); later stages S
    const bool mirror = mixed && stage == 1;
    const void* Su = stage == 0 ? static_cast<const void*>(u) : stage > 1 ? fld(c, c->Sb, 0)
                   : mirror ? static_cast<const void*>(c->Yf) : static_cast<const void*>(Yu);
    const void* Sv = stage == 0 ? static_cast<const void*>(v) : stage > 1 ? fld(c, c->Sb, 1)
                   : mirror ? static_cast<const void*>(c->Yf + c->npts) : static_cast<const void*>(Yv);
    const int variant = !mixed ? VAR_F64 : (stage == 0 ? VAR_MIX_IN64 : VAR_MIX_IN32);
    const int final_epi = stage == 0 ? EPI_OUT_RHS : EPI_RK1 + (stage - 1);
    void* P = c->Pb;
    bool have_p = false;
    it = StageItems();
    if (c->curved) {
        CurvedCall& cc = it.cc;
        std::memset(&cc, 0, sizeof(cc));
        cc.u = Su; cc.v = Sv; cc.in32 = (variant == VAR_MIX_IN32);
        cc.Pv = P;
        cc.mixed = mixed; cc.weighted = 1;
        for (int f = 0; f < 6; ++f) cc.bv[f] = bv[f];
        it.has_curved = true;
        have_p = true;
    }
    for (int q = 0; q < 3; ++q) {
        const int axis = 2 - q;
        const bool last = q == 2;
        const bool lap = !c->curved;
        if (!last && !lap && !ko) continue;   // nothing to add along this axis
        PassDesc p = base_pass(c, axis, bv);
        p.variant = variant;
        p.lap = lap;
        p.ko = ko;
        p.needv = ko || last || (lap && som_axis(c, axis));
        p.keepv = last;
        p.Su = Su; p.Sv = Sv;
        p.P = P;
        p.epi = last ? final_epi : (have_p ? EPI_MID : EPI_FIRST);
        p.So_u = fld(c, c->Sb, 0); p.So_v = fld(c, c->Sb, 1);
        p.Yu = Yu; p.Yv = Yv;
        p.Yfu = mixed ? c->Yf : nullptr; p.Yfv = mixed ? c->Yf + c->npts : nullptr;
        p.Ou = Ou; p.Ov = Ov;
        p.dt = dt;
        it.pass[q] = p;
        it.has_pass[q] = true;
        have_p = true;
    }
}

// only: -2 = everything, -1 = the curved operator alone, 0..2 = that axis's pass alone
static int run_stage(wt2_ctx* c, int stage, const double* u, const double* v, double* Ou,
                     double* Ov, double* Yu, double* Yv, double dt, int only = -2, bool allow_fold = true)
{
    StageItems it;
    stage_items(c, stage, u, v, Ou, Ov, Yu, Yv, dt, it, allow_fold);
    if (it.has_curved && (only == -2 || only == -1)) {
        int rc = pass_rc(curved_operator(c->cd, it.cc, c->stream), "curved operator");
        if (rc) return rc;
    }
    for (int q = 0; q < 3; ++q) {
        if (!it.has_pass[q]) continue;
        if (only != -2 && only != it.pass[q].axis) continue;
        const PassDesc& pd = it.pass[q];
        const hipError_t e = launch_flat_pass(pd, c->stream);
        if (e != hipSuccess) {
            char what[160];
            std::snprintf(what, sizeof(what), "flat pass (N=%d M=%dx%dx%d axis=%d stage=%d nb=%d lds=%d variant=%d)",
                          pd.N, pd.M[0], pd.M[1], pd.M[2], pd.axis, stage, pd.nb, pd.direct ? 0 : 1, int(pd.variant));
            return pass_rc(e, what);
        }
    }
    return WT2_OK;
}

// partial: the reduction's RED_BLOCKS partial sums (nullptr: the context's own).
namespace {
bool fo_general(const wt2_ctx* c);
int fc_mu_ready(wt2_ctx* c);
}  // namespace

static int run_energy(wt2_ctx* c, const double* u, const double* v, double* Edst, double* partial = nullptr)
{
    static const double zero6[6] = {0, 0, 0, 0, 0, 0};
    if (c->curved) {
        CurvedCall cc;
        std::memset(&cc, 0, sizeof(cc));
        cc.u = u; cc.v = nullptr; cc.out = c->Lu;
        cc.mixed = false; cc.weighted = 0;
        int rc = pass_rc(curved_operator(c->cd, cc, c->stream), "curved operator (energy)");
        if (rc) return rc;
    } else {
        int rc = run_lap(c, u, nullptr, c->Lu, zero6, true, false);
        if (rc) return rc;
    }
    EnergyDesc e;
    std::memset(&e, 0, sizeof(e));
    e.N = c->N; e.Ne = c->Ne;
    e.u = u; e.v = v; e.Lu = c->Lu;
    e.wv = c->curved ? c->cd.wv : nullptr;
    e.hvol = c->curved ? 1.0 : c->h[0] * c->h[1] * c->h[2];
    for (int i = 0; i < c->N; ++i) e.w[i] = c->w[i];
    e.partial = partial ? partial : c->red;
    e.out = Edst;
    return pass_rc(launch_energy(e, c->stream), "energy");
}

namespace {
bool misaligned(const void* p) { return (reinterpret_cast<uintptr_t>(p) & 15u) != 0; }

int stage_in(wt2_ctx* c, int nf, const double* const* src, double** dst, bool host)
{
    const size_t bytes = size_t(c->npts) * sizeof(double);
    for (int f = 0; f < nf; ++f) {
        if (!src[f]) { dst[f] = nullptr; continue; }
        if (!host) {
            if (misaligned(src[f])) { set_error("device field pointers must be 16-byte aligned"); return WT2_ERR_ARG; }
            dst[f] = const_cast<double*>(src[f]);
            continue;
        }
        if (!c->Y[f]) {
            int rc = alloc_dev(c, reinterpret_cast<void**>(&c->Y[f]), bytes);
            if (rc) return rc;
        }
        HIPCHK(hipMemcpyAsync(c->Y[f], src[f], bytes, hipMemcpyHostToDevice, c->stream));
        dst[f] = c->Y[f];
    }
    return WT2_OK;
}

int out_buf(wt2_ctx* c, int nf, double* const* user, double** dst, bool host)
{
    const size_t bytes = size_t(c->npts) * sizeof(double);
    for (int f = 0; f < nf; ++f) {
        if (!host) {
            if (misaligned(user[f])) { set_error("device field pointers must be 16-byte aligned"); return WT2_ERR_ARG; }
            dst[f] = user[f];
            continue;
        }
        if (!c->O[f]) {
            int rc = alloc_dev(c, reinterpret_cast<void**>(&c->O[f]), bytes);
            if (rc) return rc;
        }
        dst[f] = c->O[f];
    }
    return WT2_OK;
}

int copy_out(wt2_ctx* c, int nf, double* const* user, double* const* dev)
{
    const size_t bytes = size_t(c->npts) * sizeof(double);
    for (int f = 0; f < nf; ++f)
        HIPCHK(hipMemcpyAsync(user[f], dev[f], bytes, hipMemcpyDeviceToHost, c->stream));
    return WT2_OK;
}

int sync(wt2_ctx* c)
{
    HIPCHK(hipStreamSynchronize(c->stream));
    return WT2_OK;
}
}  // namespace

namespace {
struct TuneKey {
    int v[9];
    bool operator<(const TuneKey& o) const { return std::lexicographical_compare(v, v + 9, o.v, o.v + 9); }
};
struct TuneVal {
    int nb[3], direct[3];
    float ms[3];
};
std::mutex g_tune_mu;
std::map<TuneKey, TuneVal> g_tune;

TuneKey tune_key(const wt2_ctx* c)
{
    TuneKey k;
    k.v[0] = c->N; k.v[1] = c->d.M[0]; k.v[2] = c->d.M[1]; k.v[3] = c->d.M[2];
    k.v[4] = c->d.precision; k.v[5] = c->curved; k.v[6] = c->d.eps_ko > 0;
    k.v[7] = c->d.periodic[0] + 2 * c->d.periodic[1] + 4 * c->d.periodic[2];
    k.v[8] = int(som_axis(c, 0)) + 2 * int(som_axis(c, 1)) + 4 * int(som_axis(c, 2)) + 8 * (c->fold != 0) +
             16 * c->fmid;
    return k;
}
}  // namespace

extern "C" {

const char* wt2_last_error(void) { return g_err.c_str(); }
const char* wt2_version(void) { return "wt2 1.0 (gfx950)"; }

int wt2_gll(int32_t N, double* xi, double* w, double* G)
{
    if (!xi || !w || !G) { set_error("wt2_gll: null pointer"); return WT2_ERR_ARG; }
    if (N < NMIN || N > NMAX) { set_error("wt2_gll: N out of range 3..10"); return WT2_ERR_SHAPE; }
    gll(N, xi, w, G);
    return WT2_OK;
}

// Curved contexts compute their Minkowski coefficients on first use.
static int ensure_coefs(const wt2_ctx* cc)
{
    wt2_ctx* c = const_cast<wt2_ctx*>(cc);
    if (!c->curved || !c->cd.pending) return WT2_OK;
    return curved_set_background(c->cd, c->d, c->xi, c->w, c->G, nullptr, c->stream);
}

int wt2_ctx_footprint(const wt2_mesh_desc* desc, int32_t with_background, int64_t* bytes)
{
    if (!desc || !bytes) { set_error("wt2_ctx_footprint: null pointer"); return WT2_ERR_ARG; }
    if (desc->N < NMIN || desc->N > NMAX || desc->M[0] < 1 || desc->M[1] < 1 || desc->M[2] < 1 ||
        (desc->precision != WT2_FP64 && desc->precision != WT2_MIXED)) {
        set_error("wt2_ctx_footprint: bad N, M or precision");
        return WT2_ERR_SHAPE;
    }
    const int N = desc->N;
    const int64_t Ne = int64_t(desc->M[0]) * desc->M[1] * desc->M[2];
    const int64_t n = Ne * N * N * N;
    const int64_t es = desc->precision == WT2_MIXED ? int64_t(sizeof(float)) : int64_t(sizeof(double));
    const int64_t ts = op_table_size(N);
    // the same allocations as wt2_ctx_create: S, P register blocks, Lu,
    // reduction partials, operator tables; plus the curved coefficients
    int64_t b = 2 * (2 * n * es) + n * int64_t(sizeof(double)) + (RED_BLOCKS + 8) * int64_t(sizeof(double)) +
                ts * int64_t(sizeof(double) + 2 * sizeof(float));
    if (desc->precision == WT2_MIXED) b += 2 * n * int64_t(sizeof(float));   // FP32 mirror of the state
    if (desc->warp_A != 0.0 || with_background)
        b += 13 * n * int64_t(sizeof(double) + sizeof(float)) + n * int64_t(sizeof(double)) +
             Ne * 6 * N * N * int64_t(sizeof(double)) + 5 * Ne * 6 * N * N * int64_t(sizeof(double) + sizeof(float));
    *bytes = b;
    return WT2_OK;
}

// General-mesh input of wt2_ctx_create_mesh: node coordinates and the
// kernels' connectivity table.
struct MeshIn {
    const double* xyz;
    std::vector<int2> conn;
    bool som;           // some outer face carries the Sommerfeld tag
};

static int ctx_create_impl(const wt2_mesh_desc* desc, const MeshIn* mi, wt2_ctx** out);

int wt2_ctx_create(const wt2_mesh_desc* desc, wt2_ctx** out)
{
    return ctx_create_impl(desc, nullptr, out);
}

// D4 map of a face's node (p, q) to the neighbour face's node (the encoding of
// wt2_ctx_create_mesh's orientation argument)
static void orient_pq(int ori, int N, int p, int q, int* pn, int* qn)
{
    int a = p, b = q;
    if (ori & 1) { a = q; b = p; }
    if (ori & 2) a = N - 1 - a;
    if (ori & 4) b = N - 1 - b;
    *pn = a; *qn = b;
}

// element-local node index of face f's node (p, q)
static int face_node(int N, int f, int p, int q)
{
    const int ax = f >> 1, r = (f & 1) ? N - 1 : 0;
    if (ax == 0) return r + N * (p + N * q);
    if (ax == 1) return p + N * (r + N * q);
    return p + N * (q + N * r);
}

int wt2_ctx_create_mesh(const wt2_mesh_desc* desc, int64_t Ne, const double* coords,
                        const int32_t* neighbour, const int32_t* neighbour_face,
                        const int32_t* orientation, const int32_t* bdry, wt2_ctx** out)
{
    if (!desc || !coords || !neighbour || !neighbour_face || !bdry || !out) {
        set_error("wt2_ctx_create_mesh: null pointer");
        return WT2_ERR_ARG;
    }
    *out = nullptr;
    const int N = desc->N;
    if (N < NMIN || N > NMAX) { set_error("wt2_ctx_create_mesh: N must be in 3..10"); return WT2_ERR_SHAPE; }
    if (Ne < 1 || Ne > (int64_t(1) << 30)) { set_error("wt2_ctx_create_mesh: Ne out of range"); return WT2_ERR_SHAPE; }
    if (desc->eps_ko != 0.0) {
        set_error("wt2_ctx_create_mesh: Kreiss-Oliger dissipation is built for uniform meshes only (eps_ko must be 0)");
        return WT2_ERR_UNSUPPORTED;
    }
    // the element count as M0*M1*M2 within the per-axis limits (the RK4
    // update kernels tile the element array by that shape)
    const int mmax = N >= 9 ? 32 : MMAX_AXIS;
    int Mf[3] = {0, 0, 0};
    for (int m2 = 1; m2 <= mmax && !Mf[0]; ++m2) {
        if (Ne % m2) continue;
        const int64_t r = Ne / m2;
        for (int m1 = 1; m1 <= mmax; ++m1)
            if (r % m1 == 0 && r / m1 <= mmax) { Mf[0] = int(r / m1); Mf[1] = m1; Mf[2] = m2; break; }
    }
    if (!Mf[0]) {
        set_error("wt2_ctx_create_mesh: Ne does not factor as M0*M1*M2 within the per-axis element limits");
        return WT2_ERR_UNSUPPORTED;
    }
    const int N2 = N * N;
    const int64_t N3 = int64_t(N2) * N;
    auto X = [&](int64_t e, int n, int a) { return coords[size_t(3 * (e * N3 + n) + a)]; };
    MeshIn mi;
    mi.xyz = coords;
    mi.som = false;
    mi.conn.resize(size_t(Ne) * 6);
    char buf[256];
    for (int64_t e = 0; e < Ne; ++e) {
        // element size for the coordinate tolerance: its longest corner-to-corner span
        double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
        for (int n = 0; n < N3; ++n)
            for (int a = 0; a < 3; ++a) { lo[a] = std::min(lo[a], X(e, n, a)); hi[a] = std::max(hi[a], X(e, n, a)); }
        const double size = std::max(hi[0] - lo[0], std::max(hi[1] - lo[1], hi[2] - lo[2]));
        if (!(size > 0) || !std::isfinite(size)) {
            std::snprintf(buf, sizeof(buf), "wt2_ctx_create_mesh: element %lld has degenerate or non-finite coordinates",
                          (long long)(e + 1));
            set_error(buf);
            return WT2_ERR_ARG;
        }
        for (int f = 0; f < 6; ++f) {
            const size_t k = size_t(6 * e + f);
            const int32_t nb = neighbour[k];
            int2 cf;
            if (nb == 0) {
                const int tag = bdry[k];
                if (tag < 1 || tag > EXCISION_TAG) {
                    std::snprintf(buf, sizeof(buf),
                                  "wt2_ctx_create_mesh: outer face %d of element %lld has tag %d (1..6 Dirichlet, 7 Sommerfeld, "
                                  "8 excision)",
                                  f + 1, (long long)(e + 1), tag);
                    set_error(buf);
                    return WT2_ERR_ARG;
                }
                mi.som |= tag == SOMMERFELD_TAG;
                cf.x = -1;
                cf.y = tag << 6;
                mi.conn[k] = cf;
                continue;
            }
            const int fn = neighbour_face[k] - 1;
            if (nb < 1 || nb > Ne || fn < 0 || fn > 5 || (nb - 1 == e && fn == f)) {
                std::snprintf(buf, sizeof(buf), "wt2_ctx_create_mesh: face %d of element %lld: bad neighbour %d / face %d",
                              f + 1, (long long)(e + 1), nb, fn + 1);
                set_error(buf);
                return WT2_ERR_ARG;
            }
            const int64_t en = nb - 1;
            int ori = -1;
            if (orientation) {
                ori = orientation[k];
                if (ori < 0 || ori > 7) {
                    std::snprintf(buf, sizeof(buf), "wt2_ctx_create_mesh: face %d of element %lld: orientation %d not in 0..7",
                                  f + 1, (long long)(e + 1), ori);
                    set_error(buf);
                    return WT2_ERR_ARG;
                }
            } else {
                // the D4 map under which the two faces' nodes coincide
                double best = 1e300;
                for (int o = 0; o < 8; ++o) {
                    double dmax = 0.0;
                    for (int q = 0; q < N && dmax < best; ++q)
                        for (int p = 0; p < N; ++p) {
                            int pn, qn;
                            orient_pq(o, N, p, q, &pn, &qn);
                            const int n0 = face_node(N, f, p, q), n1 = face_node(N, fn, pn, qn);
                            for (int a = 0; a < 3; ++a) dmax = std::max(dmax, std::fabs(X(e, n0, a) - X(en, n1, a)));
                        }
                    if (dmax < best) { best = dmax; ori = o; }
                }
                if (!(best <= 1e-8 * size)) {
                    std::snprintf(buf, sizeof(buf),
                                  "wt2_ctx_create_mesh: face %d of element %lld does not meet face %d of element %d "
                                  "(closest node distance %g)", f + 1, (long long)(e + 1), fn + 1, nb, best);
                    set_error(buf);
                    return WT2_ERR_ARG;
                }
            }
            cf.x = int(en);
            cf.y = fn | (ori << 3);
            mi.conn[k] = cf;
        }
    }
    // every shared face is shared both ways, with inverse node maps
    for (int64_t e = 0; e < Ne; ++e)
        for (int f = 0; f < 6; ++f) {
            const int2 cf = mi.conn[size_t(6 * e + f)];
            if (cf.x < 0) continue;
            const int fn = cf.y & 7, ori = (cf.y >> 3) & 7;
            const int2 back = mi.conn[size_t(6 * int64_t(cf.x) + fn)];
            bool ok = back.x == e && (back.y & 7) == f;
            const int orib = (back.y >> 3) & 7;
            for (int q = 0; q < N && ok; ++q)
                for (int p = 0; p < N && ok; ++p) {
                    int pn, qn, pb, qb;
                    orient_pq(ori, N, p, q, &pn, &qn);
                    orient_pq(orib, N, pn, qn, &pb, &qb);
                    ok = pb == p && qb == q;
                }
            if (!ok) {
                std::snprintf(buf, sizeof(buf),
                              "wt2_ctx_create_mesh: face %d of element %lld and face %d of element %d do not name each other "
                              "with inverse orientations", f + 1, (long long)(e + 1), fn + 1, cf.x + 1);
                set_error(buf);
                return WT2_ERR_ARG;
            }
        }
    wt2_mesh_desc d = *desc;
    for (int a = 0; a < 3; ++a) {
        d.M[a] = Mf[a];
        d.x0[a] = 0.0;
        d.x1[a] = 1.0;
        d.periodic[a] = 0;
    }
    for (int f = 0; f < 6; ++f) d.face_kind[f] = 0;
    d.warp_A = 0.0;
    return ctx_create_impl(&d, &mi, out);
}

static int ctx_create_impl(const wt2_mesh_desc* desc, const MeshIn* mi, wt2_ctx** out)
{
    if (!desc || !out) { set_error("wt2_ctx_create: null pointer"); return WT2_ERR_ARG; }
    *out = nullptr;
    if (desc->N < NMIN || desc->N > NMAX) {
        set_error("wt2_ctx_create: N must be in 3..10");
        return WT2_ERR_SHAPE;
    }
    for (int a = 0; a < 3; ++a) {
        if (desc->M[a] < 1) { set_error("wt2_ctx_create: M per axis must be >= 1"); return WT2_ERR_SHAPE; }
        if (desc->M[a] > MMAX_AXIS) {
            set_error("wt2_ctx_create: more than 64 elements along one axis is not supported");
            return WT2_ERR_UNSUPPORTED;
        }
        if (desc->N >= 9 && desc->M[a] > 32) {
            set_error("wt2_ctx_create: N >= 9 supports at most 32 elements along one axis");
            return WT2_ERR_UNSUPPORTED;
        }
        if (!(desc->x1[a] > desc->x0[a])) { set_error("wt2_ctx_create: x1 <= x0"); return WT2_ERR_ARG; }
        if (desc->periodic[a] && desc->M[a] < 2) {
            set_error("wt2_ctx_create: periodic axis needs M >= 2");
            return WT2_ERR_SHAPE;
        }
    }
    if (desc->precision != WT2_FP64 && desc->precision != WT2_MIXED) {
        set_error("wt2_ctx_create: precision must be WT2_FP64 or WT2_MIXED");
        return WT2_ERR_ARG;
    }
    if (!(desc->tau > 0) || !(desc->eps_ko >= 0)) {
        set_error("wt2_ctx_create: tau must be > 0 and eps_ko >= 0");
        return WT2_ERR_ARG;
    }
    for (int f = 0; f < 6; ++f) {
        if (desc->face_kind[f] != 0 && desc->face_kind[f] != SOMMERFELD_TAG) {
            set_error("wt2_ctx_create: face_kind must be 0 (Dirichlet) or 7 (Sommerfeld)");
            return WT2_ERR_ARG;
        }
        if (desc->face_kind[f] == SOMMERFELD_TAG && desc->periodic[f / 2]) {
            set_error("wt2_ctx_create: a Sommerfeld face on a periodic axis");
            return WT2_ERR_ARG;
        }
    }
    if (!(desc->sommerfeld_inv_R >= 0) || !std::isfinite(desc->sommerfeld_inv_R)) {
        set_error("wt2_ctx_create: sommerfeld_inv_R must be finite and >= 0");
        return WT2_ERR_ARG;
    }
    wt2_ctx* c = new wt2_ctx();
    c->d = *desc;
    c->N = desc->N;
    c->Ne = desc->M[0] * desc->M[1] * desc->M[2];
    c->npts = int64_t(c->N) * c->N * c->N * c->Ne;
    gll(c->N, c->xi, c->w, c->G);
    for (int a = 0; a < 3; ++a) c->h[a] = (desc->x1[a] - desc->x0[a]) / desc->M[a];
    c->curved = desc->warp_A != 0.0 || mi != nullptr;
    c->mesh_som = mi && mi->som;
    const int N = c->N;

    // Coercivity of the 1D SIPG operator: the flat operator is a Kronecker sum
    // of these, so its H-symmetric part is NSD iff each axis's is.
    c->max_sym_eig = -1e300;
    for (int a = 0; a < 3; ++a) {
        const int Mt = std::min(desc->M[a], 8);
        const double ev = sipg_max_sym_eig(N, Mt, desc->periodic[a] != 0, desc->tau, c->G, c->w);
        c->max_sym_eig = std::max(c->max_sym_eig, ev);
    }
    if (c->max_sym_eig > 1e-9 * desc->tau && !(desc->flags & WT2_ALLOW_NONCOERCIVE)) {
        char buf[320];
        std::snprintf(buf, sizeof(buf),
                      "wt2_ctx_create: tau=%g is below SIPG coercivity for N=%d (max eig of sym(HL) = %g > 0); "
                      "use tau >= 1.5*(N-1)^2 or pass WT2_ALLOW_NONCOERCIVE",
                      desc->tau, N, c->max_sym_eig);
        set_error(buf);
        delete c;
        return WT2_ERR_TAU;
    }

    // mu: 30 power iterations on D^2 along axis 1 from the alternating start
    // vector (src/wave3d_curved.jl:158-178).  D_1 acts line by line and every
    // x-line of that start vector is +-p, p(i,ex) = (-1)^(i+ex) (0-based:
    // +1 when i+ex is even), so after the first normalisation the iteration
    // is the 1D one on p.
    {
        const int M = desc->M[0], L = M * N;
        std::vector<double> v(L), wv(L);
        for (int ex = 0; ex < M; ++ex)
            for (int i = 0; i < N; ++i) v[ex * N + i] = ((i + ex) % 2 == 0) ? 1.0 : -1.0;
        double mu2 = 0.0;
        for (int it = 0; it < 30; ++it) {
            apply_D_1d(N, M, desc->periodic[0] != 0, c->h[0], c->G, c->w, v.data(), wv.data());
            apply_D_1d(N, M, desc->periodic[0] != 0, c->h[0], c->G, c->w, wv.data(), v.data());
            double s = 0.0;
            for (int q = 0; q < L; ++q) s += v[q] * v[q];
            mu2 = std::sqrt(s);
            if (!(mu2 > 0)) break;
            for (int q = 0; q < L; ++q) v[q] /= mu2;
        }
        c->mu = std::sqrt(mu2);
        c->inv_mu5 = c->mu > 0 ? std::pow(1.0 / c->mu, 5) : 0.0;
        const double koT = desc->eps_ko * c->inv_mu5;
        c->kos = koT > 0 ? std::pow(koT, 1.0 / 6.0) : 0.0;
    }
    {
        double gap = 1e300;
        for (int i = 1; i < N; ++i) gap = std::min(gap, c->xi[i] - c->xi[i - 1]);
        c->h_min = gap * std::min(c->h[0], std::min(c->h[1], c->h[2]));
    }
    // Pass kernels and tile sizes (measured on the 32^3 x 8^3 workload): the x
    // pass is always LDS-free; in FP64 the y and z passes are LDS-free too
    // (strided loads merged in L2, 8-wavefront tiles), in mixed precision they
    // stage through LDS (z tiles of 8 wavefronts, y of 4).
    const bool mixed_ctx = desc->precision == WT2_MIXED;
    const char* env_direct = std::getenv("WT2_DIRECT");
    for (int a = 0; a < 3; ++a) {
        c->direct[a] = a == 0 ? 1 : (mixed_ctx ? 0 : 1);
        if (env_direct && a > 0) c->direct[a] = std::atoi(env_direct) != 0;
        const int want = c->direct[a] ? (a == 0 ? 4 : 8) : (a == 2 ? 8 : 4);
        c->nb[a] = default_nb(N, desc->M[a], want);
        if (a > 0 && !env_direct) {
            // y and z: LDS-free with the exchange through LDS (coalesced lanes),
            // the largest tile of at most 256 work-items (measured better than
            // 512 at 32^3 x 8^3), else the smallest that fits 512
            for (int nb = N; nb >= 1; --nb)
                if (N % nb == 0 && N * nb * desc->M[a] <= 256) { c->direct[a] = 2; c->nb[a] = nb; break; }
            if (c->direct[a] != 2 && N * desc->M[a] <= 512) { c->direct[a] = 2; c->nb[a] = 1; }
        }
        const char* env = std::getenv("WT2_TILE_NB");
        if (env) {
            int t[3];
            if (std::sscanf(env, "%d,%d,%d", &t[0], &t[1], &t[2]) == 3 && t[a] >= 1 && N % t[a] == 0)
                c->nb[a] = t[a];
        }
    }

    {
        // the folded 8-pass schedule by default (32^3 x 8^3: 1.66 ms per FP64
        // step against 1.91 ms for the 12-pass one, 1.16 against 1.29 ms mixed,
        // with the XCD-ordered z passes); WT2_FOLD=0 selects the 12-pass one
        const char* ev = std::getenv("WT2_FOLD");
        c->fold = ev ? std::atoi(ev) != 0 : 1;
        fold_order(c);
    }
    {
        std::lock_guard<std::mutex> lk(g_tune_mu);
        auto it = g_tune.find(tune_key(c));
        if (it != g_tune.end())
            for (int a = 0; a < 3; ++a) { c->nb[a] = it->second.nb[a]; c->direct[a] = it->second.direct[a]; }
    }
    if (hipGetDevice(&c->device) != hipSuccess ||
        hipStreamCreateWithFlags(&c->own_stream, hipStreamNonBlocking) != hipSuccess) {
        set_error("wt2_ctx_create: no HIP device / stream");
        delete c;
        return WT2_ERR_HIP;
    }
    c->stream = c->own_stream;
    const size_t n = size_t(c->npts);
    const size_t blk = 2 * n * elem_size(c);
    int rc = alloc_dev(c, &c->Sb, blk);
    if (rc == WT2_OK) rc = alloc_dev(c, &c->Pb, blk);
    if (rc == WT2_OK && is_mixed(c)) rc = alloc_dev(c, reinterpret_cast<void**>(&c->Yf), 2 * n * sizeof(float));
    if (rc == WT2_OK) rc = alloc_dev(c, reinterpret_cast<void**>(&c->Lu), n * sizeof(double));
    if (rc == WT2_OK) rc = alloc_dev(c, reinterpret_cast<void**>(&c->red), (RED_BLOCKS + 8) * sizeof(double));
    if (rc == WT2_OK) {
        const int ts = op_table_size(N);
        std::vector<double> tab(ts);
        std::vector<float> tab32(ts), tab32c(ts);
        build_op_table(N, c->G, c->w, tab.data());
        for (int q = 0; q < ts; ++q) tab32[q] = float(tab[q]);
        build_op_table32(N, c->G, c->w, tab32c.data(), c->G32);
        rc = alloc_dev(c, &c->op64, ts * sizeof(double));
        if (rc == WT2_OK) rc = alloc_dev(c, &c->op32, ts * sizeof(float));
        if (rc == WT2_OK) rc = alloc_dev(c, &c->op32c, ts * sizeof(float));
        if (rc == WT2_OK && (hipMemcpy(c->op64, tab.data(), ts * sizeof(double), hipMemcpyHostToDevice) != hipSuccess ||
                             hipMemcpy(c->op32, tab32.data(), ts * sizeof(float), hipMemcpyHostToDevice) != hipSuccess ||
                             hipMemcpy(c->op32c, tab32c.data(), ts * sizeof(float), hipMemcpyHostToDevice) != hipSuccess)) {
            set_error("wt2_ctx_create: operator table upload failed");
            rc = WT2_ERR_HIP;
        }
    }
    if (rc == WT2_OK && mi)
        rc = curved_init_mesh(c->cd, *desc, c->xi, c->w, c->G, mi->xyz, mi->conn, c->stream, &c->bytes);
    else if (rc == WT2_OK && c->curved)
        rc = curved_init(c->cd, *desc, c->xi, c->w, c->G, nullptr, c->stream, &c->bytes);
    if (rc == WT2_OK && c->curved) {
        c->cd.op64 = c->op64; c->cd.op32 = c->op32c;
        std::memcpy(c->cd.G32, c->G32, sizeof(c->G32));
    }
    if (rc != WT2_OK) {
        wt2_ctx_destroy(c);
        return rc;
    }
    *out = c;
    return WT2_OK;
}

void wt2_ctx_destroy(wt2_ctx* c)
{
    if (!c) return;
    if (c->gvalid) hipGraphExecDestroy(c->gexec);
    hipFree(c->Sb); hipFree(c->Pb); hipFree(c->Yf);
    for (int f = 0; f < 2; ++f) { hipFree(c->Y[f]); hipFree(c->O[f]); }
    hipFree(c->Lu); hipFree(c->red); hipFree(c->Edev);
    hipFree(c->op64); hipFree(c->op32); hipFree(c->op32c);
    hipFree(c->fo_D);
    hipFree(c->fo_rk);
    hipFree(c->fc_JAp); hipFree(c->fc_Jp); hipFree(c->fc_ell); hipFree(c->fc_F); hipFree(c->fc_ko); hipFree(c->fc_ops);
    hipFree(c->fc_fb);
    curved_free(c->cd);
    if (c->own_stream) hipStreamDestroy(c->own_stream);
    delete c;
}

int wt2_ctx_info(const wt2_ctx* c, wt2_info* info)
{
    if (!c || !info) { set_error("wt2_ctx_info: null pointer"); return WT2_ERR_ARG; }
    {
        const int crc = ensure_coefs(c);
        if (crc) return crc;
    }
    info->npoints = c->npts;
    info->Ne = c->Ne;
    info->curved = c->curved;
    info->mu = c->mu;
    if (fo_general(c) && c->fc_JAp) {
        // general mesh with the first-order path set up: mu of its KO derivative
        const int mrc = fc_mu_ready(const_cast<wt2_ctx*>(c));
        if (mrc) return mrc;
        info->mu = c->fc_mu;
    }
    info->h_min = c->curved ? c->cd.h_min : c->h_min;
    info->tau = c->d.tau;
    info->max_sym_eig = c->max_sym_eig;
    info->device_bytes = c->bytes;
    {
        int n = 0;
        for (int st = 1; st <= 4; ++st) {
            StageItems it;
            stage_items(const_cast<wt2_ctx*>(c), st, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, 0.0, it);
            n += it.has_curved ? 2 : 0;    // face flux + element kernel
            for (int q = 0; q < 3; ++q) n += it.has_pass[q] ? 1 : 0;
        }
        info->launches_per_step = n;
    }
    info->reserved = 0;
    return WT2_OK;
}

int wt2_set_stream(wt2_ctx* c, void* stream)
{
    if (!c) { set_error("wt2_set_stream: null ctx"); return WT2_ERR_ARG; }
    c->stream = stream ? static_cast<hipStream_t>(stream) : c->own_stream;
    if (c->gvalid) { hipGraphExecDestroy(c->gexec); c->gvalid = false; }
    return WT2_OK;
}

int wt2_coords(const wt2_ctx* c, double* xyz)
{
    if (!c || !xyz) { set_error("wt2_coords: null pointer"); return WT2_ERR_ARG; }
    if (!c->cd.xyz.empty())
        std::memcpy(xyz, c->cd.xyz.data(), c->cd.xyz.size() * sizeof(double));
    else
        mesh_coords(c->d, c->xi, xyz);
    return WT2_OK;
}

int wt2_ctx_set_background(wt2_ctx* c, const double* bg10)
{
    if (!c || !bg10) { set_error("wt2_ctx_set_background: null pointer"); return WT2_ERR_ARG; }
    for (int a = 0; a < 4; ++a)
        if (a == 3 ? c->mesh_som : som_axis(c, a)) {
            set_error("wt2_ctx_set_background: Sommerfeld faces are built for meshes without a background");
            return WT2_ERR_UNSUPPORTED;
        }
    if (c->gvalid) { hipGraphExecDestroy(c->gexec); c->gvalid = false; }
    int rc;
    if (!c->curved) {
        rc = curved_init(c->cd, c->d, c->xi, c->w, c->G, bg10, c->stream, &c->bytes);
        if (rc == WT2_OK) {
            c->cd.op64 = c->op64; c->cd.op32 = c->op32c;
            std::memcpy(c->cd.G32, c->G32, sizeof(c->G32));
        }
        if (rc == WT2_OK) c->curved = 1;
    } else {
        rc = curved_set_background(c->cd, c->d, c->xi, c->w, c->G, bg10, c->stream);
    }
    return rc;
}

int wt2_rhs3d(wt2_ctx* c, double* udd, const double* u, const double* ud,
              const double bdry_values[6], int32_t flags)
{
    if (!c || !udd || !u || !bdry_values) { set_error("wt2_rhs3d: null pointer"); return WT2_ERR_ARG; }
    {
        const int crc = ensure_coefs(c);
        if (crc) return crc;
    }
    const bool host = flags & WT2_HOST_PTRS;
    const double* src[2] = {u, ud};
    double* din[2];
    double* dout[1];
    double* user[1] = {udd};
    int rc = stage_in(c, 2, src, din, host);
    if (!rc) rc = out_buf(c, 1, user, dout, host);
    if (!rc) rc = run_lap(c, din[0], din[1], dout[0], bdry_values, false, ud != nullptr);
    if (!rc && host) rc = copy_out(c, 1, user, dout);
    if (!rc) rc = sync(c);
    return rc;
}

int wt2_rhs_full(wt2_ctx* c, double* du, double* dud, const double* u, const double* ud,
                 int32_t flags)
{
    if (!c || !du || !dud || !u || !ud) { set_error("wt2_rhs_full: null pointer"); return WT2_ERR_ARG; }
    {
        const int crc = ensure_coefs(c);
        if (crc) return crc;
    }
    const bool host = flags & WT2_HOST_PTRS;
    const double* src[2] = {u, ud};
    double* din[2];
    double* dout[2];
    double* user[2] = {du, dud};
    int rc = stage_in(c, 2, src, din, host);
    if (!rc) rc = out_buf(c, 2, user, dout, host);
    if (!rc) rc = run_stage(c, 0, din[0], din[1], dout[0], dout[1], nullptr, nullptr, 0.0);
    if (!rc && host) rc = copy_out(c, 2, user, dout);
    if (!rc) rc = sync(c);
    return rc;
}

// Start of a run: in mixed contexts the FP32 mirror of Y; in a folded
// schedule P = the stage-1 contribution of axis e_4 on Y (or its mirror).
static int run_prologue(wt2_ctx* c, const double* Yu, const double* Yv)
{
    const bool mixed = is_mixed(c);
    if (mixed) {
        const int rc = pass_rc(launch_mirror32(Yu, Yv, c->Yf, c->Yf + c->npts, c->npts, c->stream), "state mirror");
        if (rc) return rc;
    }
    if (!fold_on(c)) return WT2_OK;
    const int ax = c->fe[3];
    PassDesc p = base_pass(c, ax, c->d.bdry_values);
    p.variant = mixed ? VAR_MIX_IN32 : VAR_F64;
    const bool ko = c->d.eps_ko > 0 && c->kos > 0;
    p.lap = 1; p.ko = ko; p.needv = ko || som_axis(c, ax);
    p.Su = mixed ? static_cast<const void*>(c->Yf) : Yu;
    p.Sv = mixed ? static_cast<const void*>(c->Yf + c->npts) : Yv;
    p.P = c->Pb;
    p.epi = EPI_FIRST;
    return pass_rc(launch_flat_pass(p, c->stream), "flat pass (prologue)");
}

static int rk4_step_launches(wt2_ctx* c, double* Yu, double* Yv, double dt)
{
    for (int s = 1; s <= 4; ++s) {
        int rc = run_stage(c, s, nullptr, nullptr, nullptr, nullptr, Yu, Yv, dt);
        if (rc) return rc;
    }
    return WT2_OK;
}

int wt2_rk4_steps(wt2_ctx* c, double* u, double* ud, double dt, int32_t nsteps,
                  int32_t energy_every, double* energy_out, int32_t flags)
{
    if (!c || !u || !ud) { set_error("wt2_rk4_steps: null pointer"); return WT2_ERR_ARG; }
    {
        const int crc = ensure_coefs(c);
        if (crc) return crc;
    }
    if (nsteps < 0 || !(dt > 0) || !std::isfinite(dt)) { set_error("wt2_rk4_steps: bad dt / nsteps"); return WT2_ERR_ARG; }
    const bool host = flags & WT2_HOST_PTRS;
    const bool want_e = energy_out && energy_every > 0;
    const double* src[2] = {u, ud};
    double* Y[2];
    int rc = stage_in(c, 2, src, Y, host);
    if (rc) return rc;
    const int ne = want_e ? 1 + nsteps / energy_every : 0;
    if (want_e && ne > c->Ecap) {
        hipFree(c->Edev);
        c->Edev = nullptr;
        rc = alloc_dev(c, reinterpret_cast<void**>(&c->Edev), size_t(ne) * sizeof(double));
        if (rc) return rc;
        c->Ecap = ne;
    }
    if (want_e) {
        rc = run_energy(c, Y[0], Y[1], c->Edev);
        if (rc) return rc;
    }
    const bool use_graph = !(flags & WT2_NO_GRAPH);
    rc = run_prologue(c, Y[0], Y[1]);
    if (rc) return rc;
    for (int st = 0; st < nsteps; ++st) {
        if (use_graph) {
            if (!(c->gvalid && c->g_dt == dt && c->g_u == Y[0] && c->g_v == Y[1])) {
                if (c->gvalid) { hipGraphExecDestroy(c->gexec); c->gvalid = false; }
                // first step of a new configuration runs eagerly (sets kernel
                // attributes), then the step is captured for the rest
                rc = rk4_step_launches(c, Y[0], Y[1], dt);
                if (rc) return rc;
                hipGraph_t g;
                HIPCHK(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
                rc = rk4_step_launches(c, Y[0], Y[1], dt);
                hipError_t ce = hipStreamEndCapture(c->stream, &g);
                if (rc) return rc;
                if (ce != hipSuccess) return pass_rc(ce, "graph capture");
                hipError_t ie = hipGraphInstantiate(&c->gexec, g, nullptr, nullptr, 0);
                hipGraphDestroy(g);
                if (ie != hipSuccess) return pass_rc(ie, "graph instantiate");
                c->gvalid = true;
                c->g_dt = dt; c->g_u = Y[0]; c->g_v = Y[1];
            } else {
                HIPCHK(hipGraphLaunch(c->gexec, c->stream));
            }
        } else {
            rc = rk4_step_launches(c, Y[0], Y[1], dt);
            if (rc) return rc;
        }
        if (want_e && (st + 1) % energy_every == 0) {
            rc = run_energy(c, Y[0], Y[1], c->Edev + (st + 1) / energy_every);
            if (rc) return rc;
        }
    }
    if (host) {
        double* user[2] = {u, ud};
        rc = copy_out(c, 2, user, Y);
        if (rc) return rc;
    }
    if (want_e) HIPCHK(hipMemcpyAsync(energy_out, c->Edev, size_t(ne) * sizeof(double),
                                      hipMemcpyDeviceToHost, c->stream));
    rc = sync(c);
    if (rc) return rc;
    if (want_e) {
        for (int q = 0; q < ne; ++q)
            if (!std::isfinite(energy_out[q])) {
                set_error("wt2_rk4_steps: non-finite energy (the evolution blew up)");
                return WT2_ERR_NONFINITE;
            }
    }
    return WT2_OK;
}

// ---- the entry point that allocates nothing (wt2_rk4_steps_noalloc) ----
//
// Its energy history and the reduction partials live in `scratch`, a fixed
// array of this code object that every GPU holds from the library's load until
// the process ends.  Host code never passes `scratch` itself to a kernel or a
// copy: the array's device address, per device, comes from
// hipGetSymbolAddress on the first call on that device.
constexpr int NOALLOC_SCRATCH = RED_BLOCKS + 8 + WT2_NOALLOC_MAX_ENERGIES;
static __device__ __align__(256) double scratch[NOALLOC_SCRATCH];

namespace {
constexpr int MAX_DEVICES = 64;
std::mutex g_scratch_mu;
double* g_scratch_ptr[MAX_DEVICES] = {};

// Device address of `scratch` on `dev`, looked up once per device.
int scratch_on(int dev, double** out)
{
    if (dev < 0 || dev >= MAX_DEVICES) { set_error("wt2_rk4_steps_noalloc: device index out of range"); return WT2_ERR_ARG; }
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        int cur = -1;
        HIPCHK(hipGetDevice(&cur));
        if (cur != dev) HIPCHK(hipSetDevice(dev));
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (cur != dev) (void)hipSetDevice(cur);
        if (e != hipSuccess || !p) {
            set_error(std::string("wt2_rk4_steps_noalloc: hipGetSymbolAddress(scratch) failed: ") + hipGetErrorString(e));
            return WT2_ERR_HIP;
        }
        g_scratch_ptr[dev] = static_cast<double*>(p);
    }
    *out = g_scratch_ptr[dev];
    return WT2_OK;
}
}  // namespace

int wt2_rk4_steps_noalloc(wt2_ctx* c, double* u, double* ud, double dt, int32_t nsteps,
                          int32_t energy_every, double* energy_out)
{
    if (!c || !u || !ud) { set_error("wt2_rk4_steps_noalloc: null pointer"); return WT2_ERR_ARG; }
    if (nsteps < 0 || !(dt > 0) || !std::isfinite(dt)) { set_error("wt2_rk4_steps_noalloc: bad dt / nsteps"); return WT2_ERR_ARG; }
    if (misaligned(u) || misaligned(ud)) { set_error("wt2_rk4_steps_noalloc: device pointers must be 16-byte aligned"); return WT2_ERR_ARG; }
    if (c->curved && c->cd.pending) {
        set_error("wt2_rk4_steps_noalloc: the context's curved coefficients are not computed yet "
                  "(call wt2_ctx_info once after creating it)");
        return WT2_ERR_STATE;
    }
    const bool want_e = energy_out && energy_every > 0;
    const int64_t ne = want_e ? 1 + int64_t(nsteps) / energy_every : 0;
    if (ne > WT2_NOALLOC_MAX_ENERGIES) {
        set_error("wt2_rk4_steps_noalloc: " + std::to_string(ne) + " energy samples exceed the maximum of " +
                  std::to_string(WT2_NOALLOC_MAX_ENERGIES) + " (WT2_NOALLOC_MAX_ENERGIES)");
        return WT2_ERR_ARG;
    }
    double* sp = nullptr;
    int rc = scratch_on(c->device, &sp);
    if (rc) return rc;
    double* partial = sp;
    double* Ehist = sp + RED_BLOCKS + 8;
    if (want_e) {
        rc = run_energy(c, u, ud, Ehist, partial);
        if (rc) return rc;
    }
    rc = run_prologue(c, u, ud);
    if (rc) return rc;
    for (int st = 0; st < nsteps; ++st) {
        rc = rk4_step_launches(c, u, ud, dt);
        if (rc) return rc;
        if (want_e && (st + 1) % energy_every == 0) {
            rc = run_energy(c, u, ud, Ehist + (st + 1) / energy_every, partial);
            if (rc) return rc;
        }
    }
    if (want_e) HIPCHK(hipMemcpyAsync(energy_out, Ehist, size_t(ne) * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    rc = sync(c);
    if (rc) return rc;
    for (int64_t q = 0; q < ne; ++q)
        if (!std::isfinite(energy_out[q])) {
            set_error("wt2_rk4_steps_noalloc: non-finite energy (the evolution blew up)");
            return WT2_ERR_NONFINITE;
        }
    return WT2_OK;
}

int wt2_launch_pass(wt2_ctx* c, int32_t axis, int32_t stage, double* u, double* ud, double dt)
{
    if (!c || !u || !ud) { set_error("wt2_launch_pass: null pointer"); return WT2_ERR_ARG; }
    if (axis < -1 || axis > 3 || stage < 1 || stage > 4) { set_error("wt2_launch_pass: bad axis / stage"); return WT2_ERR_ARG; }
    {
        const int crc = ensure_coefs(c);
        if (crc) return crc;
    }
    if (axis == -1 && !c->curved) { set_error("wt2_launch_pass: axis -1 needs a curved context"); return WT2_ERR_ARG; }
    if (axis == 3) return run_lap(c, u, ud, c->Lu, c->d.bdry_values, false, true);
    return run_stage(c, stage, nullptr, nullptr, nullptr, nullptr, u, ud, dt, axis);
}

}  // extern "C"


extern "C" {

// Autotuner (no counterpart in the reference): picks, per axis, the pass kernel
// (LDS-free or LDS-staged) and the element planes per tile by timing every
// valid candidate on the context's stream with the given state (device
// pointers; the state is only read), the other axes at their best so far
// (z, then y, then x; twice for a folded context).  A folded context times stages 1-3 of the step with the
// candidate (six passes: every pass kernel the step runs, stage 4 being stage
// 2's with Y written); the passes timed in their step order, not one pass
// launched back to back, which favoured tiles that lost in the step (32^3 x
// 8^3 FP64: 1.287 against 1.266 ms).  Otherwise that axis's stage-2 pass is
// timed alone.  Results are cached per (N, M, precision, curved, KO,
// periodicity, schedule) for later contexts.
int wt2_ctx_autotune(wt2_ctx* c, double* u, double* ud, double dt, int32_t flags)
{
    if (!c || !u || !ud || !(dt > 0)) { set_error("wt2_ctx_autotune: bad arguments"); return WT2_ERR_ARG; }
    if (misaligned(u) || misaligned(ud)) { set_error("wt2_ctx_autotune: device pointers must be 16-byte aligned"); return WT2_ERR_ARG; }
    {
        const int crc = ensure_coefs(c);
        if (crc) return crc;
    }
    const TuneKey key = tune_key(c);
    {
        std::lock_guard<std::mutex> lk(g_tune_mu);
        auto it = g_tune.find(key);
        if (it != g_tune.end() && !(flags & 1)) {
            for (int a = 0; a < 3; ++a) { c->nb[a] = it->second.nb[a]; c->direct[a] = it->second.direct[a]; }
            if (c->gvalid) { (void)hipGraphExecDestroy(c->gexec); c->gvalid = false; }
            return WT2_OK;
        }
    }
    if (c->gvalid) { (void)hipGraphExecDestroy(c->gexec); c->gvalid = false; }
    int rc = WT2_OK;
    if (is_mixed(c))    // stage 1 reads the FP32 mirror of the state
        rc = pass_rc(launch_mirror32(u, ud, c->Yf, c->Yf + c->npts, c->npts, c->stream), "state mirror");
    if (!rc) rc = run_stage(c, 1, nullptr, nullptr, nullptr, nullptr, u, ud, dt, -2, false);   // fill S, P, A
    if (rc) return rc;
    hipEvent_t e0, e1;
    HIPCHK(hipEventCreate(&e0));
    HIPCHK(hipEventCreate(&e1));
    TuneVal best;
    const int N = c->N;
    // folded contexts: two sweeps, since an axis's best tile depends on the
    // tiles of the passes around it in the step
    const int sweeps = fold_on(c) ? 2 : 1;
    for (int it = 0; it < 3 * sweeps; ++it) {
        const int a = 2 - it % 3;
        float tbest = 1e30f;
        int bnb = c->nb[a], bdir = c->direct[a];
        const int lpw = 64 / c->d.M[a];
        for (int dir = 2; dir >= 0; --dir) {
            // the folded schedule needs LDS-free passes along its e axes: an
            // LDS-staged candidate there would switch it off (8 -> 12 launches
            // per step)
            if (dir == 0 && fold_axis(c, a) && c->fold && !c->curved) continue;
            for (int nb = 1; nb <= N; ++nb) {
                if (!tiles_ok(c, a, nb, dir)) continue;
                c->direct[a] = dir;
                c->nb[a] = nb;
                const bool seq = fold_on(c);
                auto once = [&]() {
                    if (!seq) return run_stage(c, 2, nullptr, nullptr, nullptr, nullptr, u, ud, dt, a, false);
                    for (int st = 1; st <= 3; ++st) {
                        const int r = run_stage(c, st, nullptr, nullptr, nullptr, nullptr, u, ud, dt);
                        if (r) return r;
                    }
                    return int(WT2_OK);
                };
                rc = once();
                if (rc) { set_error(""); rc = 0; continue; }     // configuration not launchable: skip it
                // best of three trials (single trials vary by several percent
                // from launch to launch)
                const int reps = seq ? 2 : 6;
                float ms = 1e30f;
                for (int trial = 0; trial < 3 && !rc; ++trial) {
                    (void)hipEventRecord(e0, c->stream);
                    for (int r = 0; r < reps && !rc; ++r) rc = once();
                    (void)hipEventRecord(e1, c->stream);
                    (void)hipEventSynchronize(e1);
                    float t = 0.f;
                    (void)hipEventElapsedTime(&t, e0, e1);
                    ms = std::min(ms, t / reps);
                }
                if (rc) return rc;
                if (ms < tbest) { tbest = ms; bnb = nb; bdir = dir; }
            }
        }
        c->nb[a] = bnb;
        c->direct[a] = bdir;
        best.nb[a] = bnb; best.direct[a] = bdir; best.ms[a] = tbest;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    std::lock_guard<std::mutex> lk(g_tune_mu);
    g_tune[key] = best;
    return WT2_OK;
}

int wt2_ctx_set_tiles(wt2_ctx* c, const int32_t* nb3, const int32_t* direct3)
{
    if (!c || !nb3 || !direct3) { set_error("wt2_ctx_set_tiles: null pointer"); return WT2_ERR_ARG; }
    for (int a = 0; a < 3; ++a) {
        const int nb = nb3[a];
        const int lpw = 64 / c->d.M[a];
        if (!tiles_ok(c, a, nb, direct3[a])) {
            set_error("wt2_ctx_set_tiles: kernel kind must be 0, 1 or 2 and planes per tile must divide N "
                      "and fit one workgroup (8 wavefronts; 16 for the LDS-staged kernel in mixed precision)");
            return WT2_ERR_ARG;
        }
        (void)lpw;
    }
    for (int a = 0; a < 3; ++a) { c->nb[a] = nb3[a]; c->direct[a] = direct3[a]; }
    if (c->gvalid) { (void)hipGraphExecDestroy(c->gexec); c->gvalid = false; }
    return WT2_OK;
}

int wt2_ctx_tiles(const wt2_ctx* c, int32_t* nb3, int32_t* direct3)
{
    if (!c) { set_error("wt2_ctx_tiles: null ctx"); return WT2_ERR_ARG; }
    for (int a = 0; a < 3; ++a) {
        if (nb3) nb3[a] = c->nb[a];
        if (direct3) direct3[a] = c->direct[a];
    }
    return WT2_OK;
}

int wt2_energy(wt2_ctx* c, const double* u, const double* ud, double* E, int32_t flags)
{
    if (!c || !u || !ud || !E) { set_error("wt2_energy: null pointer"); return WT2_ERR_ARG; }
    {
        const int crc = ensure_coefs(c);
        if (crc) return crc;
    }
    const bool host = flags & WT2_HOST_PTRS;
    const double* src[2] = {u, ud};
    double* din[2];
    int rc = stage_in(c, 2, src, din, host);
    if (!rc) rc = run_energy(c, din[0], din[1], c->red + RED_BLOCKS);
    if (rc) return rc;
    HIPCHK(hipMemcpyAsync(E, c->red + RED_BLOCKS, sizeof(double), hipMemcpyDeviceToHost, c->stream));
    return sync(c);
}

}  // extern "C"

// ------------------------------------------------- first-order (Phi, Pi) system

namespace {

bool fo_general(const wt2_ctx* c) { return c->cd.conn != nullptr; }

void fc_desc(const wt2_ctx* c, FcDesc& d)
{
    std::memset(&d, 0, sizeof(d));
    d.N = c->N;
    d.npts = c->npts;
    d.G = c->fc_ops;
    d.w = c->fc_ops + 256;
    d.conn = c->cd.conn;
    d.JAp = c->fc_JAp; d.Jp = c->fc_Jp; d.ell = c->fc_ell;
    d.D = c->fo_D; d.F = c->fc_F; d.ko = c->fc_ko; d.fb = c->fc_fb;
    d.partial = c->red;
}

// The first-order path: affine uniform meshes (per-axis apply_D!), or general
// meshes given by coordinates and connectivity (metric terms set up here on
// first use).  Warped uniform meshes and context backgrounds are refused (the
// coefficients come in coef11; a warped mesh can be handed over as a general
// mesh).
int fo_ready(wt2_ctx* c, const char* who)
{
    if (!fo_general(c) && (c->curved || c->d.warp_A != 0.0)) {
        set_error(std::string(who) + ": the first-order path is built for affine meshes and for meshes given by "
                                     "coordinates and connectivity, without a context background (the "
                                     "coefficients come in coef11)");
        return WT2_ERR_UNSUPPORTED;
    }
    const size_t n = size_t(c->npts);
    if (!c->fo_D) {
        const int rc = alloc_dev(c, reinterpret_cast<void**>(&c->fo_D), 3 * n * sizeof(double));
        if (rc) return rc;
    }
    if (!fo_general(c) || c->fc_JAp) return WT2_OK;
    int rc = alloc_dev(c, reinterpret_cast<void**>(&c->fc_JAp), 9 * n * sizeof(double));
    if (!rc) rc = alloc_dev(c, reinterpret_cast<void**>(&c->fc_Jp), n * sizeof(double));
    if (!rc) rc = alloc_dev(c, reinterpret_cast<void**>(&c->fc_ell), 3 * size_t(c->Ne) * sizeof(double));
    if (!rc) rc = alloc_dev(c, reinterpret_cast<void**>(&c->fc_F), 3 * n * sizeof(double));
    if (!rc) rc = alloc_dev(c, reinterpret_cast<void**>(&c->fc_fb), 6 * size_t(c->Ne) * c->N * c->N * sizeof(double));
    if (!rc) rc = alloc_dev(c, reinterpret_cast<void**>(&c->fc_ko), 12 * n * sizeof(double));
    if (!rc) rc = alloc_dev(c, reinterpret_cast<void**>(&c->fc_ops), (256 + 16) * sizeof(double));
    double* xyz = nullptr;
    if (!rc && hipMalloc(&xyz, 3 * n * sizeof(double)) != hipSuccess) {
        set_error(std::string(who) + ": device staging allocation failed");
        rc = WT2_ERR_OOM;
    }
    if (!rc) {
        double ops[256 + 16] = {0};
        for (int q = 0; q < c->N * c->N; ++q) ops[q] = c->G[q];
        for (int q = 0; q < c->N; ++q) ops[256 + q] = c->w[q];
        rc = pass_rc(hipMemcpyAsync(c->fc_ops, ops, sizeof(ops), hipMemcpyHostToDevice, c->stream), "metric set-up copy");
        if (!rc) rc = pass_rc(hipMemcpyAsync(xyz, c->cd.xyz.data(), 3 * n * sizeof(double), hipMemcpyHostToDevice,
                                             c->stream), "metric set-up copy");
        FcDesc d;
        fc_desc(c, d);
        d.xyz = xyz;
        if (!rc) rc = pass_rc(launch_fc(d, 0, c->stream), "metric terms");
        if (!rc) rc = sync(c);
    }
    if (xyz) (void)hipFree(xyz);
    c->fc_mu = -1.0;
    if (rc) {
        hipFree(c->fc_JAp); c->fc_JAp = nullptr;
    }
    return rc;
}

// mu of the general-mesh KO derivative, computed once
int fc_mu_ready(wt2_ctx* c)
{
    if (c->fc_mu >= 0) return WT2_OK;
    FcDesc d;
    fc_desc(c, d);
    double mu = 0.0;
    const int rc = pass_rc(fc_mu(d, c->fc_ko, c->fc_ko + c->npts, c->stream, &mu), "KO spectral radius");
    if (!rc) c->fc_mu = mu;
    return rc;
}

void fo_desc(const wt2_ctx* c, FoDesc& fd)
{
    std::memset(&fd, 0, sizeof(fd));
    fd.N = c->N;
    for (int a = 0; a < 3; ++a) { fd.M[a] = c->d.M[a]; fd.periodic[a] = c->d.periodic[a]; fd.h[a] = c->h[a]; }
    for (int i = 0; i < c->N; ++i) fd.w[i] = c->w[i];
    fd.op64 = c->op64;
    fd.D = c->fo_D;
    // WT2_FO_FUSE=0: the pointwise step as a kernel of its own (same results)
    const char* ev = std::getenv("WT2_FO_FUSE");
    fd.fuse = ev ? (std::atoi(ev) != 0) : 1;
}

// Device copies of host arrays for a host-pointer call (freed by the caller).
struct FoTemp {
    std::vector<void*> bufs;
    ~FoTemp() { for (void* p : bufs) hipFree(p); }
    int in(wt2_ctx* c, const double* src, size_t count, const double** dst)
    {
        void* p = nullptr;
        if (hipMalloc(&p, count * sizeof(double)) != hipSuccess) { set_error("device staging allocation failed"); return WT2_ERR_OOM; }
        bufs.push_back(p);
        if (src && hipMemcpyAsync(p, src, count * sizeof(double), hipMemcpyHostToDevice, c->stream) != hipSuccess) {
            set_error("host-to-device copy failed");
            return WT2_ERR_HIP;
        }
        *dst = static_cast<const double*>(p);
        return WT2_OK;
    }
};

bool fo_aligned(std::initializer_list<const void*> ps)
{
    for (const void* p : ps)
        if (p && misaligned(p)) return false;
    return true;
}

}  // namespace

extern "C" {

int wt2_fo_coef(wt2_ctx* c, const double* bg10, double* coef11, int32_t flags)
{
    if (!c || !bg10 || !coef11) { set_error("wt2_fo_coef: null pointer"); return WT2_ERR_ARG; }
    const bool host = flags & WT2_HOST_PTRS;
    const size_t n = size_t(c->npts);
    FoTemp tmp;
    const double* bg = bg10;
    double* out = coef11;
    int rc = WT2_OK;
    if (host) {
        const double* o = nullptr;
        rc = tmp.in(c, bg10, 10 * n, &bg);
        if (!rc) rc = tmp.in(c, nullptr, 11 * n, &o);
        out = const_cast<double*>(o);
    } else if (!fo_aligned({bg10, coef11})) {
        set_error("device field pointers must be 16-byte aligned");
        return WT2_ERR_ARG;
    }
    if (!rc) rc = pass_rc(launch_fo_coef(bg, out, c->npts, c->stream), "first-order coefficients");
    if (!rc && host)
        rc = pass_rc(hipMemcpyAsync(coef11, out, 11 * n * sizeof(double), hipMemcpyDeviceToHost, c->stream), "copy out");
    if (!rc) rc = sync(c);
    return rc;
}

int wt2_fo_rhs(wt2_ctx* c, double* dPhi, double* dPi, const double* Phi, const double* Pi,
               const double* coef11, double eps_ko, const wt2_bc3d* bc3d, int32_t flags)
{
    if (!c || !dPhi || !dPi || !Phi || !Pi || !coef11) { set_error("wt2_fo_rhs: null pointer"); return WT2_ERR_ARG; }
    if (!(eps_ko >= 0) || !std::isfinite(eps_ko)) { set_error("wt2_fo_rhs: eps_ko must be finite and >= 0"); return WT2_ERR_ARG; }
    if (bc3d) {
        for (int f = 0; f < 6; ++f)
            if (bc3d->kinds[f] < WT2_BC_DIRICHLET || bc3d->kinds[f] > WT2_BC_FULL_DIRICHLET) {
                set_error("wt2_fo_rhs: bc3d kinds must be WT2_BC_DIRICHLET..WT2_BC_FULL_DIRICHLET");
                return WT2_ERR_ARG;
            }
        if (!(bc3d->sigma >= 0) || !std::isfinite(bc3d->sigma)) { set_error("wt2_fo_rhs: bc3d sigma must be finite and >= 0"); return WT2_ERR_ARG; }
    }
    int rc = fo_ready(c, "wt2_fo_rhs");
    if (rc) return rc;
    const bool host = flags & WT2_HOST_PTRS;
    const size_t n = size_t(c->npts);
    FoDesc fd;
    fo_desc(c, fd);
    fd.Phi = Phi; fd.Pi = Pi; fd.coef = coef11; fd.dPhi = dPhi; fd.dPi = dPi;
    fd.gPhi = bc3d ? bc3d->gPhi : nullptr;
    fd.gPi = bc3d ? bc3d->gPi : nullptr;
    FoTemp tmp;
    if (host) {
        const double* o1 = nullptr; const double* o2 = nullptr;
        rc = tmp.in(c, Phi, n, &fd.Phi);
        if (!rc) rc = tmp.in(c, Pi, n, &fd.Pi);
        if (!rc) rc = tmp.in(c, coef11, 11 * n, &fd.coef);
        if (!rc && fd.gPhi) rc = tmp.in(c, bc3d->gPhi, n, &fd.gPhi);
        if (!rc && fd.gPi) rc = tmp.in(c, bc3d->gPi, n, &fd.gPi);
        if (!rc) rc = tmp.in(c, nullptr, n, &o1);
        if (!rc) rc = tmp.in(c, nullptr, n, &o2);
        fd.dPhi = const_cast<double*>(o1); fd.dPi = const_cast<double*>(o2);
        if (rc) return rc;
    } else if (!fo_aligned({dPhi, dPi, Phi, Pi, coef11, fd.gPhi, fd.gPi})) {
        set_error("device field pointers must be 16-byte aligned");
        return WT2_ERR_ARG;
    }
    if (fo_general(c)) {
        // curvilinear path: one kind on every outer face but the excision tag
        const double* gD = bc3d ? bc3d->gD : nullptr;
        if (host && gD) rc = tmp.in(c, gD, 3 * n, &gD);
        else if (!host && gD && misaligned(gD)) { set_error("device field pointers must be 16-byte aligned"); return WT2_ERR_ARG; }
        if (!rc && eps_ko > 0) rc = fc_mu_ready(c);
        if (rc) return rc;
        FcDesc d;
        fc_desc(c, d);
        d.Phi = fd.Phi; d.Pi = fd.Pi; d.coef = fd.coef; d.dPhi = fd.dPhi; d.dPi = fd.dPi;
        d.koT = (eps_ko > 0 && c->fc_mu > 0) ? eps_ko * std::pow(1.0 / c->fc_mu, 5) : 0.0;
        d.has_bc = bc3d != nullptr;
        if (bc3d) {
            d.kind = bc3d->kinds[0];
            d.excision_tag = bc3d->excision_tag;
            d.sigma = bc3d->sigma;
            d.gPhi = fd.gPhi; d.gPi = fd.gPi; d.gD = gD;
        }
        rc = pass_rc(launch_fc(d, 1, c->stream), "first-order RHS (general mesh)");
    } else {
        fd.koT = eps_ko * c->inv_mu5;
        if (bc3d) {
            for (int f = 0; f < 6; ++f) fd.kinds[f] = bc3d->kinds[f];
            fd.sigma = bc3d->sigma;
        } else {
            for (int f = 0; f < 6; ++f) fd.kinds[f] = WT2_BC_EXCISION;   // no boundary terms
            fd.sigma = 0.0;
        }
        rc = pass_rc(launch_fo_rhs(fd, c->stream), "first-order RHS");
    }
    if (!rc && host) {
        rc = pass_rc(hipMemcpyAsync(dPhi, fd.dPhi, n * sizeof(double), hipMemcpyDeviceToHost, c->stream), "copy out");
        if (!rc) rc = pass_rc(hipMemcpyAsync(dPi, fd.dPi, n * sizeof(double), hipMemcpyDeviceToHost, c->stream), "copy out");
    }
    if (!rc) rc = sync(c);
    return rc;
}

int wt2_fo_rk4_stage(wt2_ctx* c, int32_t stage, double* Phi, double* Pi, const double* coef11, double dt,
                     double eps_ko, const wt2_bc3d* bc3d)
{
    if (!c || !Phi || !Pi || !coef11) { set_error("wt2_fo_rk4_stage: null pointer"); return WT2_ERR_ARG; }
    if (stage < 1 || stage > 4 || !(dt > 0) || !std::isfinite(dt)) { set_error("wt2_fo_rk4_stage: bad stage / dt"); return WT2_ERR_ARG; }
    if (misaligned(Phi) || misaligned(Pi)) { set_error("wt2_fo_rk4_stage: device pointers must be 16-byte aligned"); return WT2_ERR_ARG; }
    int rc = fo_ready(c, "wt2_fo_rk4_stage");
    if (rc) return rc;
    const size_t n = size_t(c->npts);
    const size_t st = (n + 1) & ~size_t(1);    // field stride of the registers, 16-byte aligned
    if (!c->fo_rk) {
        rc = alloc_dev(c, reinterpret_cast<void**>(&c->fo_rk), 6 * st * sizeof(double));
        if (rc) return rc;
    }
    double* R = c->fo_rk;                      // S (2 fields), A (2), k (2)
    const double* in0 = stage == 1 ? Phi : R;
    const double* in1 = stage == 1 ? Pi : R + st;
    rc = wt2_fo_rhs(c, R + 4 * st, R + 5 * st, in0, in1, coef11, eps_ko, bc3d, 0);
    if (rc) return rc;
    rc = pass_rc(launch_fo_rk_update(Phi, Pi, R, int64_t(st), int64_t(n), stage, dt, c->stream), "first-order RK4 update");
    if (!rc) rc = sync(c);
    return rc;
}

int wt2_fo_energy(wt2_ctx* c, const double* Phi, const double* Pi, const double* coef11, double* E, int32_t flags)
{
    if (!c || !Phi || !Pi || !coef11 || !E) { set_error("wt2_fo_energy: null pointer"); return WT2_ERR_ARG; }
    int rc = fo_ready(c, "wt2_fo_energy");
    if (rc) return rc;
    const bool host = flags & WT2_HOST_PTRS;
    const size_t n = size_t(c->npts);
    FoDesc fd;
    fo_desc(c, fd);
    fd.Phi = Phi; fd.Pi = Pi; fd.coef = coef11;
    FoTemp tmp;
    if (host) {
        rc = tmp.in(c, Phi, n, &fd.Phi);
        if (!rc) rc = tmp.in(c, Pi, n, &fd.Pi);
        if (!rc) rc = tmp.in(c, coef11, 11 * n, &fd.coef);
        if (rc) return rc;
    } else if (!fo_aligned({Phi, Pi, coef11})) {
        set_error("device field pointers must be 16-byte aligned");
        return WT2_ERR_ARG;
    }
    if (fo_general(c)) {
        FcDesc d;
        fc_desc(c, d);
        d.Phi = fd.Phi; d.Pi = fd.Pi; d.coef = fd.coef;
        d.grad_only = 1;
        rc = pass_rc(launch_fc(d, 2, c->stream), "first-order energy (general mesh)");
        std::vector<double> part(RED_BLOCKS);
        if (!rc) rc = pass_rc(hipMemcpyAsync(part.data(), c->red, RED_BLOCKS * sizeof(double), hipMemcpyDeviceToHost,
                                             c->stream), "copy out");
        if (!rc) rc = sync(c);
        if (rc) return rc;
        double s = 0.0;
        for (double p : part) s += p;
        *E = s;
        return WT2_OK;
    }
    fd.grad_only = 1;
    for (int f = 0; f < 6; ++f) fd.kinds[f] = WT2_BC_EXCISION;
    fd.partial = c->red;
    fd.Eout = c->red + RED_BLOCKS;
    rc = pass_rc(launch_fo_rhs(fd, c->stream), "first-order gradient");
    if (!rc) rc = pass_rc(launch_fo_energy(fd, c->stream), "first-order energy");
    if (!rc) rc = pass_rc(hipMemcpyAsync(E, c->red + RED_BLOCKS, sizeof(double), hipMemcpyDeviceToHost, c->stream), "copy out");
    if (!rc) rc = sync(c);
    return rc;
}

}  // extern "C"

// ------------------------------------------------------------------ batches
//
// A batch advances many independent instances (contexts with their own mesh,
// N, precision, dt and step count, and device-resident state) together: every
// RK4 stage is one launch per group of instances that share a launch shape
// (phase, N, precision, pass-kernel kind), over a packed work list of all
// their tiles or elements.  The launches of one step are captured as a
// hipGraph and replayed; the plans and the graph are rebuilt only when the set
// of unfinished instances changes.

namespace {

struct BatchSlot {
    wt2_ctx* ctx = nullptr;
    double* u = nullptr;
    double* v = nullptr;
    double dt = 0.0;
    int64_t nsteps = 0, done = 0;
    int32_t every = 0;
    double* Edev = nullptr;      // energy samples, written on the device in step order
    int32_t ecap = 0, ne = 0;
};

struct StagePlans {
    std::vector<CurvedBatchPlan> curved;     // phase 0
    std::vector<FlatBatchPlan> flat[3];      // phases 1..3 (z, y, x)
};

}  // namespace

struct wt2_batch {
    std::vector<BatchSlot> slots;
    hipStream_t stream = nullptr;
    bool dirty = true;
    StagePlans plans[4];
    hipGraphExec_t gexec = nullptr;
    bool gvalid = false;
    double* Ed = nullptr;
};

static void batch_free_plans(wt2_batch* b)
{
    // queued replays may still read the plans' work lists and the graph
    (void)hipStreamSynchronize(b->stream);
    for (auto& sp : b->plans) {
        for (auto& p : sp.curved) curved_batch_free(p);
        sp.curved.clear();
        for (auto& v : sp.flat) {
            for (auto& p : v) flat_batch_free(p);
            v.clear();
        }
    }
    if (b->gvalid) { (void)hipGraphExecDestroy(b->gexec); b->gvalid = false; }
}

static bool batch_active(const BatchSlot& s) { return s.ctx && s.done < s.nsteps; }

// Group key of a pass: everything a single launch must share.
static std::string pass_key(const PassDesc& p)
{
    char buf[96];
    std::snprintf(buf, sizeof(buf), "%d/%d/%d/%d/%d/%d/%d/%d/%d", p.N, p.variant, p.axis,
                  p.direct, p.epi, p.lap, p.ko, p.needv, p.fold);
    return buf;
}

static int batch_build(wt2_batch* b)
{
    batch_free_plans(b);
    for (int stage = 1; stage <= 4; ++stage) {
        StagePlans& sp = b->plans[stage - 1];
        std::map<std::string, std::vector<PassDesc>> groups[3];
        std::map<std::string, std::pair<std::vector<const CurvedData*>, std::vector<CurvedCall>>> cgroups;
        for (auto& s : b->slots) {
            if (!batch_active(s)) continue;
            StageItems it;
            stage_items(s.ctx, stage, nullptr, nullptr, nullptr, nullptr, s.u, s.v, s.dt, it);
            if (it.has_curved) {
                // uniform and general-connectivity contexts in separate groups,
                // so each runs the element-kernel instantiation of its single run
                char key[64];
                std::snprintf(key, sizeof(key), "%d/%d/%d/%d", s.ctx->N, it.cc.mixed, it.cc.in32,
                              s.ctx->cd.conn != nullptr ? 1 : 0);
                auto& g = cgroups[key];
                g.first.push_back(&s.ctx->cd);
                g.second.push_back(it.cc);
            }
            for (int q = 0; q < 3; ++q)
                if (it.has_pass[q]) groups[q][pass_key(it.pass[q])].push_back(it.pass[q]);
        }
        for (auto& kv : cgroups) {
            CurvedBatchPlan plan;
            hipError_t e = curved_batch_build(kv.second.first.data(), kv.second.second.data(),
                                              int(kv.second.first.size()), plan);
            if (e != hipSuccess) { curved_batch_free(plan); return pass_rc(e, "batch build (curved)"); }
            sp.curved.push_back(plan);
        }
        for (int q = 0; q < 3; ++q)
            for (auto& kv : groups[q]) {
                FlatBatchPlan plan;
                hipError_t e = flat_batch_build(kv.second.data(), int(kv.second.size()), plan);
                if (e != hipSuccess) { flat_batch_free(plan); return pass_rc(e, "batch build (pass)"); }
                sp.flat[q].push_back(plan);
            }
    }
    b->dirty = false;
    return WT2_OK;
}

static int batch_launch_step(wt2_batch* b)
{
    for (int stage = 0; stage < 4; ++stage) {
        const StagePlans& sp = b->plans[stage];
        for (const auto& p : sp.curved) {
            int rc = pass_rc(curved_batch_launch(p, b->stream), "batch curved operator");
            if (rc) return rc;
        }
        for (int q = 0; q < 3; ++q)
            for (const auto& p : sp.flat[q]) {
                int rc = pass_rc(flat_batch_launch(p, b->stream), "batch pass");
                if (rc) return rc;
            }
    }
    return WT2_OK;
}

extern "C" {

int wt2_batch_create(wt2_batch** out)
{
    if (!out) { set_error("wt2_batch_create: null pointer"); return WT2_ERR_ARG; }
    wt2_batch* b = new wt2_batch();
    if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess ||
        hipMalloc(&b->Ed, sizeof(double) * (RED_BLOCKS + 8)) != hipSuccess) {
        set_error("wt2_batch_create: HIP stream / allocation failed");
        if (b->stream) (void)hipStreamDestroy(b->stream);
        delete b;
        return WT2_ERR_HIP;
    }
    *out = b;
    return WT2_OK;
}

void wt2_batch_destroy(wt2_batch* b)
{
    if (!b) return;
    batch_free_plans(b);
    (void)hipStreamSynchronize(b->stream);
    for (auto& s : b->slots) {
        if (s.ctx) s.ctx->stream = s.ctx->own_stream;
        if (s.Edev) (void)hipFree(s.Edev);
    }
    (void)hipFree(b->Ed);
    (void)hipStreamDestroy(b->stream);
    delete b;
}

int wt2_batch_add(wt2_batch* b, wt2_ctx* ctx, double* u, double* ud, double dt, int64_t nsteps,
                  int32_t energy_every, int32_t* slot)
{
    if (!b || !ctx || !u || !ud || !slot) { set_error("wt2_batch_add: null pointer"); return WT2_ERR_ARG; }
    {
        const int crc = ensure_coefs(ctx);
        if (crc) return crc;
    }
    if (!(dt > 0) || nsteps < 0 || energy_every < 0) { set_error("wt2_batch_add: bad dt / nsteps / energy_every"); return WT2_ERR_ARG; }
    if (misaligned(u) || misaligned(ud)) { set_error("wt2_batch_add: state pointers must be 16-byte aligned"); return WT2_ERR_ARG; }
    for (auto& s : b->slots)
        if (s.ctx == ctx) { set_error("wt2_batch_add: context already in the batch"); return WT2_ERR_STATE; }
    if (ctx->gvalid) { (void)hipGraphExecDestroy(ctx->gexec); ctx->gvalid = false; }
    (void)hipStreamSynchronize(ctx->stream);
    ctx->stream = b->stream;
    BatchSlot s;
    s.ctx = ctx; s.u = u; s.v = ud; s.dt = dt; s.nsteps = nsteps; s.done = 0; s.every = energy_every;
    if (energy_every > 0) {
        const int64_t cap = nsteps / energy_every + 2;
        if (cap > INT32_MAX / 2) { ctx->stream = ctx->own_stream; set_error("wt2_batch_add: too many energy samples"); return WT2_ERR_ARG; }
        s.ecap = int32_t(cap);
        if (hipMalloc(&s.Edev, sizeof(double) * size_t(cap)) != hipSuccess) {
            ctx->stream = ctx->own_stream;
            set_error("wt2_batch_add: energy buffer allocation failed");
            return WT2_ERR_OOM;
        }
        int rc = run_energy(ctx, u, ud, s.Edev);
        if (rc) { (void)hipFree(s.Edev); ctx->stream = ctx->own_stream; return rc; }
        s.ne = 1;
    }
    {
        const int prc = run_prologue(ctx, u, ud);      // folded schedule: P for the first stage
        if (prc) { if (s.Edev) (void)hipFree(s.Edev); ctx->stream = ctx->own_stream; return prc; }
    }
    int id = -1;
    for (size_t q = 0; q < b->slots.size(); ++q)
        if (!b->slots[q].ctx) { id = int(q); break; }
    if (id < 0) { id = int(b->slots.size()); b->slots.push_back(s); }
    else b->slots[id] = s;
    *slot = id;
    if (nsteps > 0) b->dirty = true;
    return WT2_OK;
}

int wt2_batch_remove(wt2_batch* b, int32_t slot)
{
    if (!b || slot < 0 || slot >= int32_t(b->slots.size()) || !b->slots[slot].ctx) {
        set_error("wt2_batch_remove: no such slot");
        return WT2_ERR_ARG;
    }
    (void)hipStreamSynchronize(b->stream);
    BatchSlot& s = b->slots[slot];
    if (batch_active(s)) b->dirty = true;
    s.ctx->stream = s.ctx->own_stream;
    if (s.Edev) (void)hipFree(s.Edev);
    s = BatchSlot();
    return WT2_OK;
}

int wt2_batch_step(wt2_batch* b, int32_t max_steps, int32_t* finished)
{
    const int rc = wt2_batch_step_async(b, max_steps, finished);
    if (rc) return rc;
    HIPCHK(hipStreamSynchronize(b->stream));
    return WT2_OK;
}

int wt2_batch_sync(wt2_batch* b)
{
    if (!b) { set_error("wt2_batch_sync: null pointer"); return WT2_ERR_ARG; }
    HIPCHK(hipStreamSynchronize(b->stream));
    return WT2_OK;
}

int wt2_batch_step_async(wt2_batch* b, int32_t max_steps, int32_t* finished)
{
    if (!b || max_steps < 0) { set_error("wt2_batch_step: bad arguments"); return WT2_ERR_ARG; }
    int32_t nfin = 0;
    int32_t left = max_steps;
    while (left > 0) {
        int64_t k = left;
        bool any = false;
        for (auto& s : b->slots) {
            if (!batch_active(s)) continue;
            any = true;
            k = std::min<int64_t>(k, s.nsteps - s.done);
            if (s.every > 0) k = std::min<int64_t>(k, s.every - s.done % s.every);
        }
        if (!any) break;
        if (b->dirty) {
            int rc = batch_build(b);
            if (rc) return rc;
        }
        int64_t r = k;
        if (!b->gvalid) {
            // one eager step (sets kernel attributes), then capture the step
            int rc = batch_launch_step(b);
            if (rc) return rc;
            --r;
            hipGraph_t g;
            HIPCHK(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
            rc = batch_launch_step(b);
            hipError_t ce = hipStreamEndCapture(b->stream, &g);
            if (rc) return rc;
            if (ce != hipSuccess) return pass_rc(ce, "batch graph capture");
            hipError_t ie = hipGraphInstantiate(&b->gexec, g, nullptr, nullptr, 0);
            (void)hipGraphDestroy(g);
            if (ie != hipSuccess) return pass_rc(ie, "batch graph instantiate");
            b->gvalid = true;
        }
        for (int64_t q = 0; q < r; ++q) HIPCHK(hipGraphLaunch(b->gexec, b->stream));
        left -= int32_t(k);
        for (auto& s : b->slots) {
            if (!batch_active(s)) continue;
            s.done += k;
            if (s.every > 0 && s.done % s.every == 0 && s.ne < s.ecap) {
                // sampled on the stream in step order; read back by wt2_batch_slot
                int rc = run_energy(s.ctx, s.u, s.v, s.Edev + s.ne);
                if (rc) return rc;
                ++s.ne;
            }
            if (s.done >= s.nsteps) {
                ++nfin;
                b->dirty = true;     // the rebuild waits for the queued steps, then drops the graph
            }
        }
    }
    if (finished) *finished = nfin;
    return WT2_OK;
}

int wt2_batch_slot(const wt2_batch* b, int32_t slot, int64_t* steps_done, double* energies, int32_t cap,
                   int32_t* n_energies)
{
    if (!b || slot < 0 || slot >= int32_t(b->slots.size()) || !b->slots[slot].ctx) {
        set_error("wt2_batch_slot: no such slot");
        return WT2_ERR_ARG;
    }
    const BatchSlot& s = b->slots[slot];
    if (steps_done) *steps_done = s.done;
    const int32_t n = s.ne;
    if (n_energies) *n_energies = n;
    if (energies && cap > 0 && n > 0) {
        const int32_t m = std::min(n, cap);
        HIPCHK(hipMemcpyAsync(energies, s.Edev, sizeof(double) * size_t(m), hipMemcpyDeviceToHost, b->stream));
        HIPCHK(hipStreamSynchronize(b->stream));
        for (int32_t q = 0; q < m; ++q)
            if (!std::isfinite(energies[q])) {
                set_error("wt2_batch_slot: non-finite energy sample (the evolution blew up)");
                return WT2_ERR_NONFINITE;
            }
    }
    return WT2_OK;
}

int wt2_batch_info(const wt2_batch* b, int32_t* launches_per_step, int32_t* active)
{
    if (!b || !launches_per_step || !active) { set_error("wt2_batch_info: null pointer"); return WT2_ERR_ARG; }
    int32_t n = 0;
    if (!b->dirty)
        for (const auto& sp : b->plans) {
            n += int32_t(sp.curved.size());
            for (int q = 0; q < 3; ++q) n += int32_t(sp.flat[q].size());
        }
    *launches_per_step = n;
    int32_t a = 0;
    for (const auto& s : b->slots) a += batch_active(s) ? 1 : 0;
    *active = a;
    return WT2_OK;
}

}  // extern "C"
