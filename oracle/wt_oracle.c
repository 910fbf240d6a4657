/*
 * wt_oracle.c -- CPU FP64 oracle (TEST INFRASTRUCTURE ONLY; see wt_oracle.h).
 *
 * Straight element-by-element loops, OpenMP over elements, no attempt at
 * speed beyond that.  Each function cites the reference file:line whose
 * semantics it restates.  HexSBPSAT/HexMeshes (where the reference's
 * arithmetic lives) are not vendored in the reference tree, so the operator
 * follows the specification text instead: README.md:57-81 (SIPG),
 * METHODS.md:256-298 (apply_D!, KO), and the conventions of
 * src/wave_lap_strong.jl:27-62 (G sweeps, face rows, centred-flux SAT).
 */
#include "wt_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

struct wto_ctx {
    wto_desc d;
    int N, Ne, npe;
    long npts;
    double xi[WTO_NMAX], w[WTO_NMAX], G[WTO_NMAX * WTO_NMAX];
    double h[3];      /* unwarped element widths per axis */
    int curved;       /* warp or background present */
    double* coef;     /* curved: 13 per node (see wto_coefs) */
    double* coords;   /* (3,N,N,N,Ne) */
    double* hphys;    /* H3*|J| per node */
    double mu, inv_mu5;
    int* fconn;       /* general mesh: per (e, f) neighbour (-1 outer), its face, D4 code, tag */
    /* general mesh, first-order curvilinear path: metric terms |J| dxi^a/dx_n
     * (conservative-curl form, 9 per node, [a*3+n]), |J| per node, element
     * extent along each reference axis, mu of the curvilinear apply_D! */
    double* JAp;
    double* Jp;
    double* ell;
    double mu_c, inv_mu5_c;
};

/* ------------------------------------------------------------------ GLL */

/* Legendre P_n and P_{n-1} at x by the three-term recurrence. */
static void legendre(int n, double x, double* pn, double* pnm1)
{
    double p0 = 1.0, p1 = x;
    if (n == 0) { *pn = 1.0; *pnm1 = 0.0; return; }
    for (int m = 2; m <= n; ++m) {
        double p2 = ((2.0 * m - 1.0) * x * p1 - (m - 1.0) * p0) / m;
        p0 = p1; p1 = p2;
    }
    *pn = p1; *pnm1 = p0;
}

/* GLL nodes/weights/differentiation matrix on [0,1] (HexSBPSAT make_element /
 * make_operators, used at src/wave.jl:603; elem.xs in [0,1] per
 * src/wave_curved_rhs.jl:366-369; SBP identity HG + (HG)^T = B, README.md:13-15).
 * Nodes: endpoints and roots of P'_{N-1} by Newton on
 * (x P_n - P_{n-1}) from Chebyshev-Gauss-Lobatto starts; weights
 * 2/(n(n+1) P_n^2) on [-1,1]; G by the barycentric formula with the
 * negative-sum diagonal. */
int wto_gll(int N, double* xi, double* w, double* G)
{
    if (N < 2 || N > WTO_NMAX) return -1;
    int n = N - 1;
    double x[WTO_NMAX];
    for (int k = 0; k < N; ++k) {
        double xk = -cos(M_PI * k / n);
        for (int it = 0; it < 100; ++it) {
            double pn, pnm1;
            legendre(n, xk, &pn, &pnm1);
            double dx = (xk * pn - pnm1) / ((n + 1.0) * pn);
            xk -= dx;
            if (fabs(dx) < 1e-16) break;
        }
        x[k] = xk;
    }
    x[0] = -1.0; x[n] = 1.0;
    for (int k = 0; k < N; ++k) {
        /* symmetrise: nodes come in +/- pairs */
        double s = 0.5 * (x[k] - x[n - k]);
        x[k] = s; x[n - k] = -s;
    }
    if (N % 2 == 1) x[n / 2] = 0.0;
    for (int k = 0; k < N; ++k) {
        double pn, pnm1;
        legendre(n, x[k], &pn, &pnm1);
        w[k] = 0.5 * 2.0 / (n * (n + 1.0) * pn * pn);
        xi[k] = 0.5 * (x[k] + 1.0);
    }
    double lam[WTO_NMAX];
    for (int j = 0; j < N; ++j) {
        double p = 1.0;
        for (int k = 0; k < N; ++k) if (k != j) p *= (xi[j] - xi[k]);
        lam[j] = 1.0 / p;
    }
    for (int i = 0; i < N; ++i) {
        double diag = 0.0;
        for (int j = 0; j < N; ++j) {
            if (i == j) continue;
            double v = (lam[j] / lam[i]) / (xi[i] - xi[j]);
            G[i * N + j] = v;
            diag -= v;
        }
        G[i * N + i] = diag;
    }
    return 0;
}

/* ---------------------------------------------------------- mesh helpers */

static inline long nidx(int N, int i, int j, int k, long e)
{
    return i + (long)N * (j + (long)N * (k + (long)N * e));
}

/* node index from (axis a, row along a, tangential p, q); face-local (p,q)
 * are (j,k), (i,k), (i,j) for a = 0,1,2 (src/wave_lap_strong.jl:71-74). */
static inline int fidx(int N, int a, int r, int p, int q)
{
    if (a == 0) return r + N * (p + N * q);
    if (a == 1) return p + N * (r + N * q);
    return p + N * (q + N * r);
}

/* Neighbour element across face (axis a, side s: 0 low, 1 high); -1 if
 * boundary.  Elements are x-fastest (HexMeshes make_uniform_hex ordering is
 * not in the reference tree; x-fastest is this project's documented choice). */
static long neighbour(const wto_ctx* c, long e, int a, int s)
{
    const int* M = c->d.M;
    int ec[3];
    ec[0] = (int)(e % M[0]);
    ec[1] = (int)((e / M[0]) % M[1]);
    ec[2] = (int)(e / ((long)M[0] * M[1]));
    int t = ec[a] + (s ? 1 : -1);
    if (t < 0 || t >= M[a]) {
        if (!c->d.periodic[a]) return -1;
        t = (t + M[a]) % M[a];
    }
    ec[a] = t;
    return ec[0] + (long)M[0] * (ec[1] + (long)M[1] * ec[2]);
}

/* warp displacement s(xi) = sin(2 pi xi1) sin(2 pi xi2) sin(2 pi xi3) and its
 * gradient; physical x_i = x0_i + L_i (xi_i + A s(xi)).  The map is the one the
 * curved-background configuration states; HexMeshes' make_warped_uniform_hex
 * (test/dirichlet_warped_tests.jl:23,40) is not in the reference tree. */
static void warp(double A, const double xi[3], double* s, double ds[3])
{
    double tp = 2.0 * M_PI;
    double s1 = sin(tp * xi[0]), s2 = sin(tp * xi[1]), s3 = sin(tp * xi[2]);
    double c1 = cos(tp * xi[0]), c2 = cos(tp * xi[1]), c3 = cos(tp * xi[2]);
    *s = A * s1 * s2 * s3;
    ds[0] = A * tp * c1 * s2 * s3;
    ds[1] = A * tp * s1 * c2 * s3;
    ds[2] = A * tp * s1 * s2 * c3;
}

/* -------------------------------------------------------------- create */

static double det3(const double J[3][3])
{
    return J[0][0] * (J[1][1] * J[2][2] - J[1][2] * J[2][1])
         - J[0][1] * (J[1][0] * J[2][2] - J[1][2] * J[2][0])
         + J[0][2] * (J[1][0] * J[2][1] - J[1][1] * J[2][0]);
}

static void inv3(const double J[3][3], double det, double R[3][3])
{
    R[0][0] = (J[1][1] * J[2][2] - J[1][2] * J[2][1]) / det;
    R[0][1] = (J[0][2] * J[2][1] - J[0][1] * J[2][2]) / det;
    R[0][2] = (J[0][1] * J[1][2] - J[0][2] * J[1][1]) / det;
    R[1][0] = (J[1][2] * J[2][0] - J[1][0] * J[2][2]) / det;
    R[1][1] = (J[0][0] * J[2][2] - J[0][2] * J[2][0]) / det;
    R[1][2] = (J[0][2] * J[1][0] - J[0][0] * J[1][2]) / det;
    R[2][0] = (J[1][0] * J[2][1] - J[1][1] * J[2][0]) / det;
    R[2][1] = (J[0][1] * J[2][0] - J[0][0] * J[2][1]) / det;
    R[2][2] = (J[0][0] * J[1][1] - J[0][1] * J[1][0]) / det;
}

static void apply_D_line1d(const wto_ctx* c, int M, int periodic, double h,
                           const double* in, double* out);

static wto_ctx* create_impl(const wto_desc* d, const double* bg, const double* mesh_xyz);
static void curv_setup(wto_ctx* c);

/* Planted mistake in the ADM coefficients (0 = none), for the pin tests that
 * must show they catch it; read when a context is created. */
static int g_plant = 0;

void wto_plant_mistake(int code) { g_plant = code; }

wto_ctx* wto_create(const wto_desc* d, const double* bg)
{
    return create_impl(d, bg, NULL);
}

/* D4 map of face node (p, q) to the neighbour face's node: swap first, then
 * reverse p' (bit 1), then reverse q' (bit 2). */
static void orient_map(int ori, int N, int p, int q, int* pn, int* qn)
{
    int a = (ori & 1) ? q : p, b = (ori & 1) ? p : q;
    *pn = (ori & 2) ? N - 1 - a : a;
    *qn = (ori & 4) ? N - 1 - b : b;
}

/* A mesh given by node coordinates xyz (3,N,N,N,Ne) and 1-based (6,Ne)
 * neighbour / neighbour_face / bdry arrays (the reference's MeshGeometry and
 * MeshConnectivity fields, src/wave.jl:471-473, src/wave_strong_rhs.jl:41-73):
 * the D4 map of each shared face is the one under which the face nodes
 * coincide; J = dx/dxi is G applied to xyz.  NULL on inconsistent input. */
wto_ctx* wto_create_mesh(const wto_desc* d, long Ne, const double* xyz, const int* nbr,
                         const int* nface, const int* bdry)
{
    if (d->N < 2 || d->N > WTO_NMAX || Ne < 1 || d->eps_ko != 0.0) return NULL;
    wto_desc dm = *d;
    dm.M[0] = (int)Ne; dm.M[1] = dm.M[2] = 1;
    for (int a = 0; a < 3; ++a) { dm.x0[a] = 0.0; dm.x1[a] = 1.0; dm.periodic[a] = 0; }
    dm.warp_A = 0.0;
    for (int f = 0; f < 6; ++f) dm.face_kind[f] = 0;
    const int N = d->N, npe = N * N * N;
    int* fc = (int*)malloc(sizeof(int) * 24 * Ne);
    for (long e = 0; e < Ne; ++e)
        for (int f = 0; f < 6; ++f) {
            int* o = fc + 4 * (6 * e + f);
            int nb = nbr[6 * e + f];
            if (nb == 0) {
                int tag = bdry[6 * e + f];
                if (tag < 1 || tag > 8) { free(fc); return NULL; }
                o[0] = -1; o[1] = 0; o[2] = 0; o[3] = tag;
                continue;
            }
            int fn = nface[6 * e + f] - 1;
            if (nb < 1 || nb > Ne || fn < 0 || fn > 5) { free(fc); return NULL; }
            long en = nb - 1;
            double span = 0.0, best = 1e300;
            int bo = -1;
            for (int n = 0; n < npe; ++n)
                for (int a = 0; a < 3; ++a)
                    span = fmax(span, fabs(xyz[3 * (e * npe + n) + a] - xyz[3 * e * npe + a]));
            for (int ori = 0; ori < 8; ++ori) {
                double dmax = 0.0;
                for (int q = 0; q < N; ++q)
                    for (int p = 0; p < N; ++p) {
                        int pn, qn;
                        orient_map(ori, N, p, q, &pn, &qn);
                        long n0 = e * npe + fidx(N, f / 2, (f & 1) ? N - 1 : 0, p, q);
                        long n1 = en * npe + fidx(N, fn / 2, (fn & 1) ? N - 1 : 0, pn, qn);
                        for (int a = 0; a < 3; ++a) dmax = fmax(dmax, fabs(xyz[3 * n0 + a] - xyz[3 * n1 + a]));
                    }
                if (dmax < best) { best = dmax; bo = ori; }
            }
            if (!(best <= 1e-8 * span)) { free(fc); return NULL; }
            o[0] = (int)en; o[1] = fn; o[2] = bo; o[3] = 0;
        }
    wto_ctx* c = create_impl(&dm, NULL, xyz);
    if (!c) { free(fc); return NULL; }
    c->fconn = fc;
    curv_setup(c);
    return c;
}

static wto_ctx* create_impl(const wto_desc* d, const double* bg, const double* mesh_xyz)
{
    if (d->N < 2 || d->N > WTO_NMAX) return NULL;
    for (int a = 0; a < 3; ++a) if (d->M[a] < 1) return NULL;
    wto_ctx* c = (wto_ctx*)calloc(1, sizeof(wto_ctx));
    c->d = *d;
    int N = c->N = d->N;
    c->npe = N * N * N;
    c->Ne = d->M[0] * d->M[1] * d->M[2];
    c->npts = (long)c->npe * c->Ne;
    wto_gll(N, c->xi, c->w, c->G);
    for (int a = 0; a < 3; ++a) c->h[a] = (d->x1[a] - d->x0[a]) / d->M[a];
    c->curved = (d->warp_A != 0.0) || (bg != NULL) || (mesh_xyz != NULL);
    /* Sommerfeld faces: non-periodic axes, no background (the drag is the
     * flat-spacetime condition of rhs_wave3d!) */
    for (int f = 0; f < 6; ++f) {
        if (d->face_kind[f] != 0 && d->face_kind[f] != 7) { free(c); return NULL; }
        if (d->face_kind[f] == 7 && (bg != NULL || d->periodic[f / 2])) { free(c); return NULL; }
    }
    c->coords = (double*)malloc(sizeof(double) * 3 * c->npts);
    c->hphys = (double*)malloc(sizeof(double) * c->npts);
    if (c->curved) c->coef = (double*)malloc(sizeof(double) * 13 * c->npts);

    #pragma omp parallel for schedule(static)
    for (long e = 0; e < c->Ne; ++e) {
        int ec[3] = {(int)(e % d->M[0]), (int)((e / d->M[0]) % d->M[1]),
                     (int)(e / ((long)d->M[0] * d->M[1]))};
        for (int k = 0; k < N; ++k)
        for (int j = 0; j < N; ++j)
        for (int i = 0; i < N; ++i) {
            long n = nidx(N, i, j, k, e);
            int ii[3] = {i, j, k};
            double xi[3], L[3];
            for (int a = 0; a < 3; ++a) {
                xi[a] = (ec[a] + c->xi[ii[a]]) / d->M[a];
                L[a] = d->x1[a] - d->x0[a];
            }
            double s = 0.0, ds[3] = {0, 0, 0};
            if (d->warp_A != 0.0) warp(d->warp_A, xi, &s, ds);
            double J[3][3];
            if (mesh_xyz) {
                for (int a = 0; a < 3; ++a) {
                    c->coords[3 * n + a] = mesh_xyz[3 * n + a];
                    for (int b = 0; b < 3; ++b) {
                        double acc = 0.0;
                        for (int p = 0; p < N; ++p) {
                            int jj[3] = {i, j, k};
                            jj[b] = p;
                            acc += c->G[ii[b] * N + p] * mesh_xyz[3 * nidx(N, jj[0], jj[1], jj[2], e) + a];
                        }
                        J[a][b] = acc;
                    }
                }
            } else {
                for (int a = 0; a < 3; ++a) {
                    c->coords[3 * n + a] = d->x0[a] + L[a] * (xi[a] + s);
                    for (int b = 0; b < 3; ++b)
                        J[a][b] = L[a] * ((a == b ? 1.0 : 0.0) + ds[b]) / d->M[b];
                }
            }
            double detJ = det3(J);
            double H3 = c->w[i] * c->w[j] * c->w[k];
            if (!c->curved) { c->hphys[n] = H3 * fabs(detJ); continue; }
            double Ji[3][3];
            inv3(J, detJ, Ji);
            double aJ = fabs(detJ);
            /* ADM coefficients per node (src/wave3d_curved.jl:85-94): sqrt det
             * gamma and the cofactor inverse gamma^{ij}. */
            double al = 1.0, be[3] = {0, 0, 0};
            double g[6] = {1, 0, 0, 1, 0, 1};
            if (bg) {
                const double* b = bg + 10 * n;
                al = b[0]; be[0] = b[1]; be[1] = b[2]; be[2] = b[3];
                for (int q = 0; q < 6; ++q) g[q] = b[4 + q];
            }
            double g11 = g[0], g12 = g[1], g13 = g[2], g22 = g[3], g23 = g[4], g33 = g[5];
            double det = g11 * (g22 * g33 - g23 * g23) - g12 * (g12 * g33 - g23 * g13)
                       + g13 * (g12 * g23 - g22 * g13);
            double sg = sqrt(det);
            double gu[3][3];
            gu[0][0] = (g22 * g33 - g23 * g23) / det;
            gu[0][1] = gu[1][0] = (g13 * g23 - g12 * g33) / det;
            gu[0][2] = gu[2][0] = (g12 * g23 - g13 * g22) / det;
            gu[1][1] = (g11 * g33 - g13 * g13) / det;
            gu[1][2] = gu[2][1] = (g12 * g13 - g11 * g23) / det;
            gu[2][2] = (g11 * g22 - g12 * g12) / det;
            /* second-order form of the conservative ADM system
             * (src/wave3d_curved.jl:1-12): K^{ij} = sqrt(g)(alpha g^{ij} - b^i b^j/alpha),
             * c^i = sqrt(g) b^i / alpha; pulled back to the unit element. */
            double K[3][3], cv[3];
            for (int p = 0; p < 3; ++p) {
                cv[p] = (g_plant == WTO_PLANT_FLIP_C ? -1.0 : 1.0) * sg * be[p] / al;
                for (int q = 0; q < 3; ++q)
                    K[p][q] = sg * (al * gu[p][q] - (g_plant == WTO_PLANT_DROP_BB ? 0.0 : be[p] * be[q] / al));
            }
            double Kh[3][3], ch[3];
            for (int a = 0; a < 3; ++a) {
                ch[a] = 0.0;
                for (int p = 0; p < 3; ++p) ch[a] += aJ * Ji[a][p] * cv[p];
                for (int b = 0; b < 3; ++b) {
                    double sacc = 0.0;
                    for (int p = 0; p < 3; ++p)
                        for (int q = 0; q < 3; ++q) sacc += Ji[a][p] * K[p][q] * Ji[b][q];
                    Kh[a][b] = aJ * sacc;
                }
            }
            double* o = c->coef + 13 * n;
            o[0] = Kh[0][0]; o[1] = Kh[0][1]; o[2] = Kh[0][2];
            o[3] = Kh[1][1]; o[4] = Kh[1][2]; o[5] = Kh[2][2];
            o[6] = ch[0]; o[7] = ch[1]; o[8] = ch[2];
            o[9] = g_plant == WTO_PLANT_WGT_SQRTG ? al * sg / aJ : al / (sg * aJ);
            for (int a = 0; a < 3; ++a) {
                double kap = Kh[a][a];
                for (int b = 0; b < 3; ++b) if (b != a) kap += fabs(Kh[a][b]);
                o[10 + a] = kap;
            }
            c->hphys[n] = H3 * aJ;
        }
    }

    /* mu: spectral radius of apply_D! along axis 1 by 30 power iterations on
     * D^2 from the alternating +-1 start vector (src/wave3d_curved.jl:158-178).
     * On the x-fastest mesh every x-line of that start vector is +-p with
     * p(i,ex) = isodd(i+ex) ? 1 : -1 (1-based), and D_1 acts line by line, so
     * after the first normalisation the iteration equals the 1D one on p. */
    {
        int M = d->M[0], L = M * N;
        double* v = (double*)malloc(sizeof(double) * L);
        double* wv = (double*)malloc(sizeof(double) * L);
        for (int ex = 0; ex < M; ++ex)
            for (int i = 0; i < N; ++i)
                v[ex * N + i] = ((i + 1 + ex + 1) % 2 == 1) ? 1.0 : -1.0;
        double mu2 = 0.0;
        for (int it = 0; it < 30; ++it) {
            apply_D_line1d(c, M, d->periodic[0], c->h[0], v, wv);
            apply_D_line1d(c, M, d->periodic[0], c->h[0], wv, v);
            double s = 0.0;
            for (int q = 0; q < L; ++q) s += v[q] * v[q];
            mu2 = sqrt(s);
            if (!(mu2 > 0)) break;
            for (int q = 0; q < L; ++q) v[q] /= mu2;
        }
        c->mu = sqrt(mu2);
        c->inv_mu5 = c->mu > 0 ? pow(1.0 / c->mu, 5) : 0.0;
        free(v); free(wv);
    }
    return c;
}

void wto_destroy(wto_ctx* c)
{
    if (!c) return;
    free(c->coef); free(c->coords); free(c->hphys); free(c->fconn);
    free(c->JAp); free(c->Jp); free(c->ell); free(c);
}

long wto_npoints(const wto_ctx* c) { return c->npts; }
int wto_is_curved(const wto_ctx* c) { return c->curved; }
double wto_mu(const wto_ctx* c) { return c->mu; }

void wto_coords(const wto_ctx* c, double* xyz)
{
    memcpy(xyz, c->coords, sizeof(double) * 3 * c->npts);
}

void wto_hphys(const wto_ctx* c, double* hw)
{
    memcpy(hw, c->hphys, sizeof(double) * c->npts);
}

int wto_coefs(const wto_ctx* c, double* out13)
{
    if (!c->curved) return -1;
    memcpy(out13, c->coef, sizeof(double) * 13 * c->npts);
    return 0;
}

/* Smallest physical spacing between reference-axis-adjacent nodes over all
 * three reference directions (src/evolve.jl:60-69). */
double wto_hmin(const wto_ctx* c)
{
    int N = c->N;
    double h = INFINITY;
    for (long e = 0; e < c->Ne; ++e)
        for (int k = 0; k < N; ++k)
        for (int j = 0; j < N; ++j)
        for (int i = 0; i < N; ++i) {
            const double* x = c->coords + 3 * nidx(N, i, j, k, e);
            int ii[3] = {i, j, k};
            for (int a = 0; a < 3; ++a) {
                if (ii[a] == 0) continue;
                int jj[3] = {i, j, k};
                jj[a] -= 1;
                const double* y = c->coords + 3 * nidx(N, jj[0], jj[1], jj[2], e);
                double dd = sqrt((x[0] - y[0]) * (x[0] - y[0]) + (x[1] - y[1]) * (x[1] - y[1])
                                 + (x[2] - y[2]) * (x[2] - y[2]));
                if (dd < h) h = dd;
            }
        }
    return h;
}

/* ------------------------------------------------------ 1D line kernels */

/* One element segment of the SAT-coupled first derivative (apply_D!,
 * METHODS.md:256-264): (1/h) G x plus, at interior faces only, the centred
 * flux f_sign/(2 H_face h) (x_nbr - x) (form of src/wave_lap_strong.jl:48-62).
 * xl / xr: neighbour face values, or NULL at a boundary face (one-sided). */
static const double ghost_zero = 0.0;

static void D_segment(const wto_ctx* c, double h, const double* x,
                      const double* xl, const double* xr, double* y)
{
    int N = c->N;
    for (int i = 0; i < N; ++i) {
        double s = 0.0;
        for (int p = 0; p < N; ++p) s += c->G[i * N + p] * x[p];
        y[i] = s / h;
    }
    if (xl) y[0] += -1.0 / (2.0 * c->w[0] * h) * (*xl - x[0]);
    if (xr) y[N - 1] += 1.0 / (2.0 * c->w[N - 1] * h) * (*xr - x[N - 1]);
}

/* ghost0 = 0: outer faces one-sided (apply_D!, METHODS.md:262-264);
 * ghost0 = 1: outer faces take the centred-flux SAT against a zero ghost value,
 * which makes H D exactly skew on a bounded line too (the KO derivative). */
static void apply_D_line1d_g(const wto_ctx* c, int M, int periodic, double h,
                             const double* in, double* out, int ghost0)
{
    int N = c->N;
    const double* gz = ghost0 ? &ghost_zero : NULL;
    for (int e = 0; e < M; ++e) {
        const double* xl = gz; const double* xr = gz;
        if (e > 0) xl = in + (e - 1) * N + N - 1;
        else if (periodic) xl = in + (M - 1) * N + N - 1;
        if (e < M - 1) xr = in + (e + 1) * N;
        else if (periodic) xr = in;
        D_segment(c, h, in + e * N, xl, xr, out + e * N);
    }
}

static void apply_D_line1d(const wto_ctx* c, int M, int periodic, double h,
                           const double* in, double* out)
{
    apply_D_line1d_g(c, M, periodic, h, in, out, 0);
}

/* One element segment of the flat 1D SIPG Laplacian (README.md:57-81):
 *   L u = (1/h^2) { G G u + H^-1 [ alpha G^T B du - 1/2 B d(Gu) - tau |B| du ] }
 * du = u - u_nbr; alpha = 1/2 at interior faces; at an outer Dirichlet face
 * alpha = 1, du = u - g and d(Gu) = 0 (mirror convention).
 * ul / ur: neighbour element segment (N values) or NULL at a boundary face,
 * in which case gl / gr is the Dirichlet value.  somL / somR: that outer face
 * is a Sommerfeld (tag 7) face instead -- free for the SIPG terms
 * (src/wave.jl:446-451) -- and carries the Bayliss-Turkel drag of
 * src/wave.jl:453-518 on its face node:
 *   y -= wF / Hphys (u_t + d_n u + u / R),   wF / Hphys = 1 / (h w_r)
 * on an affine mesh (the tangential face weights cancel); v = NULL takes
 * u_t = 0, which leaves the H-symmetric operator of the energy
 * 1/2 v'Hv + 1/2 (Gu)'H(Gu) + 1/(2R) sum_faces u^2.  Result added into y. */
static void sipg_segment_s(const wto_ctx* c, double h, const double* u,
                           const double* ul, double gl, const double* ur, double gr,
                           int somL, int somR, const double* v, double* y);

static void sipg_segment(const wto_ctx* c, double h, const double* u,
                         const double* ul, double gl, const double* ur, double gr,
                         double* y)
{
    sipg_segment_s(c, h, u, ul, gl, ur, gr, 0, 0, NULL, y);
}

static void sipg_segment_s(const wto_ctx* c, double h, const double* u,
                           const double* ul, double gl, const double* ur, double gr,
                           int somL, int somR, const double* v, double* y)
{
    int N = c->N;
    const double* G = c->G;
    const double* w = c->w;
    double tau = c->d.tau;
    double gu[WTO_NMAX];
    for (int i = 0; i < N; ++i) {
        double s = 0.0;
        for (int p = 0; p < N; ++p) s += G[i * N + p] * u[p];
        gu[i] = s;
    }
    double duL, dgL, aL, duR, dgR, aR;
    if (ul) {
        double g_nb = 0.0;
        for (int p = 0; p < N; ++p) g_nb += G[(N - 1) * N + p] * ul[p];
        duL = u[0] - ul[N - 1]; dgL = gu[0] - g_nb; aL = 0.5;
    } else if (somL) { duL = 0.0; dgL = 0.0; aL = 0.0; }
    else { duL = u[0] - gl; dgL = 0.0; aL = 1.0; }
    if (ur) {
        double g_nb = 0.0;
        for (int p = 0; p < N; ++p) g_nb += G[0 * N + p] * ur[p];
        duR = u[N - 1] - ur[0]; dgR = gu[N - 1] - g_nb; aR = 0.5;
    } else if (somR) { duR = 0.0; dgR = 0.0; aR = 0.0; }
    else { duR = u[N - 1] - gr; dgR = 0.0; aR = 1.0; }
    double ih2 = 1.0 / (h * h);
    for (int i = 0; i < N; ++i) {
        double s = 0.0;
        for (int p = 0; p < N; ++p) s += G[i * N + p] * gu[p];
        /* alpha G^T B du: B = diag(-1,0..0,+1) */
        s += (aL * G[0 * N + i] * (-duL) + aR * G[(N - 1) * N + i] * duR) / w[i];
        if (i == 0) s += (0.5 * dgL - tau * duL) / w[0];
        if (i == N - 1) s += (-0.5 * dgR - tau * duR) / w[N - 1];
        y[i] += s * ih2;
    }
    {
        /* outward normal derivative: -(1/h) Gu at the low face, +(1/h) Gu at the
         * high one; without v the u_t part is zero (the energy's operator) */
        double iR = c->d.sommerfeld_inv_R;
        if (!ul && somL) y[0] -= ((v ? v[0] : 0.0) - gu[0] / h + iR * u[0]) / (h * w[0]);
        if (!ur && somR) y[N - 1] -= ((v ? v[N - 1] : 0.0) + gu[N - 1] / h + iR * u[N - 1]) / (h * w[N - 1]);
    }
}

/* ----------------------------------------------------- flat operators */

static void gather_line(const wto_ctx* c, const double* f, long e, int a, int p, int q, double* out)
{
    int N = c->N;
    for (int r = 0; r < N; ++r) out[r] = f[(long)e * c->npe + fidx(N, a, r, p, q)];
}

/* Flat (affine, Minkowski) Laplacian: Kronecker sum of 1D SIPG operators
 * along the three axes (the face terms of one axis involve only that axis,
 * tangential H cancels). */
/* v (u_t) of the stage, or NULL for u_t = 0 (the energy's operator); it only
 * enters through the Sommerfeld faces' drag. */
static void lap_flat(const wto_ctx* c, const double* u, const double* v, const double* bv, double* y)
{
    int N = c->N;
    #pragma omp parallel for schedule(static)
    for (long e = 0; e < c->Ne; ++e) {
        double* ye = y + e * c->npe;
        for (int n = 0; n < c->npe; ++n) ye[n] = 0.0;
        double us[WTO_NMAX], ul[WTO_NMAX], ur[WTO_NMAX], vs[WTO_NMAX], acc[WTO_NMAX];
        for (int a = 0; a < 3; ++a) {
            long el = neighbour(c, e, a, 0), er = neighbour(c, e, a, 1);
            int somL = c->d.face_kind[2 * a] == 7, somR = c->d.face_kind[2 * a + 1] == 7;
            for (int q = 0; q < N; ++q)
            for (int p = 0; p < N; ++p) {
                gather_line(c, u, e, a, p, q, us);
                if (el >= 0) gather_line(c, u, el, a, p, q, ul);
                if (er >= 0) gather_line(c, u, er, a, p, q, ur);
                if (v) gather_line(c, v, e, a, p, q, vs);
                for (int r = 0; r < N; ++r) acc[r] = 0.0;
                sipg_segment_s(c, c->h[a], us, el >= 0 ? ul : NULL, bv[2 * a],
                               er >= 0 ? ur : NULL, bv[2 * a + 1], somL, somR, v ? vs : NULL, acc);
                for (int r = 0; r < N; ++r) ye[fidx(N, a, r, p, q)] += acc[r];
            }
        }
    }
}

static void apply_D_g(const wto_ctx* c, int axis, const double* in, double* out, int ghost0);

void wto_apply_D(const wto_ctx* c, int axis, const double* in, double* out)
{
    apply_D_g(c, axis, in, out, 0);
}

void wto_apply_Dko(const wto_ctx* c, int axis, const double* in, double* out)
{
    apply_D_g(c, axis, in, out, 1);
}

static void apply_D_g(const wto_ctx* c, int axis, const double* in, double* out, int ghost0)
{
    int N = c->N, a = axis;
    #pragma omp parallel for schedule(static)
    for (long e = 0; e < c->Ne; ++e) {
        long el = neighbour(c, e, a, 0), er = neighbour(c, e, a, 1);
        double x[WTO_NMAX], y[WTO_NMAX];
        for (int q = 0; q < N; ++q)
        for (int p = 0; p < N; ++p) {
            gather_line(c, in, e, a, p, q, x);
            double xl = 0, xr = 0;
            if (el >= 0) xl = in[el * c->npe + fidx(N, a, N - 1, p, q)];
            if (er >= 0) xr = in[er * c->npe + fidx(N, a, 0, p, q)];
            D_segment(c, c->h[a], x, (el >= 0 || ghost0) ? &xl : NULL,
                      (er >= 0 || ghost0) ? &xr : NULL, y);
            for (int r = 0; r < N; ++r) out[e * c->npe + fidx(N, a, r, p, q)] = y[r];
        }
    }
}

/* ---------------------------------------------------- curved operator */

/* Reference-coordinate gradient of element field f at node (i,j,k). */
static void grad_node(const wto_ctx* c, const double* fe, int i, int j, int k, double g[3])
{
    int N = c->N;
    const double* G = c->G;
    double s0 = 0, s1 = 0, s2 = 0;
    for (int p = 0; p < N; ++p) {
        s0 += G[i * N + p] * fe[p + N * (j + N * k)];
        s1 += G[j * N + p] * fe[i + N * (p + N * k)];
        s2 += G[k * N + p] * fe[i + N * (j + N * p)];
    }
    g[0] = s0; g[1] = s1; g[2] = s2;
}

static inline double Kab(const double* o, int a, int b)
{
    static const int map[3][3] = {{0, 1, 2}, {1, 3, 4}, {2, 4, 5}};
    return o[map[a][b]];
}

/* Curved / background operator.  Writes Ltilde u + A v (unweighted) into y;
 * with_adv = 0 drops the u_t advection part.  The second-order part is SIPG
 * for d_a(K^ab d_b u) in unit-element coordinates:
 *   vol:  sum_a G_a F^a,   F^a = sum_b K^ab G_b u
 *   face (axis a, outward sign n, row r, weight w_r):
 *     y += (1/w_r) [ -1/2 n dF^a - tau kappa du ]            (average flux, penalty)
 *     y += alpha n sum_b (H_b^-1 G_b^T H_b)(K^ab du) on the face (symmetrising term;
 *          the b = a part spreads along the normal line with weight 1/w_m)
 *   kappa = K^aa + sum_{b!=a}|K^ab|, averaged over the two sides.
 * It reduces to the README.md:57-81 operator when K = h I.  Advection
 * A v = sum_a [c^a G_a v + G_a(c^a v)] - (1/w_r) n c^a (v - v_nbr) at faces
 * (split form, skew in the H norm), v_nbr = 0 at Dirichlet faces. */
static void op_curved(const wto_ctx* c, const double* u, const double* v,
                      const double* bv, int with_adv, double* y)
{
    int N = c->N, npe = c->npe;
    const double* G = c->G;
    const double* w = c->w;
    double tau = c->d.tau;
    #pragma omp parallel
    {
        double* F = (double*)malloc(sizeof(double) * 3 * npe);
        double* cvv = (double*)malloc(sizeof(double) * 3 * npe);
        #pragma omp for schedule(static)
        for (long e = 0; e < c->Ne; ++e) {
            const double* ue = u + e * npe;
            const double* ve = v ? v + e * npe : NULL;
            const double* co = c->coef + 13 * e * (long)npe;
            double* ye = y + e * npe;
            for (int k = 0; k < N; ++k)
            for (int j = 0; j < N; ++j)
            for (int i = 0; i < N; ++i) {
                int n = i + N * (j + N * k);
                double g[3];
                grad_node(c, ue, i, j, k, g);
                const double* o = co + 13 * n;
                for (int a = 0; a < 3; ++a)
                    F[a * npe + n] = Kab(o, a, 0) * g[0] + Kab(o, a, 1) * g[1] + Kab(o, a, 2) * g[2];
                if (with_adv)
                    for (int a = 0; a < 3; ++a) cvv[a * npe + n] = o[6 + a] * ve[n];
            }
            for (int k = 0; k < N; ++k)
            for (int j = 0; j < N; ++j)
            for (int i = 0; i < N; ++i) {
                int n = i + N * (j + N * k);
                double s = 0.0;
                for (int p = 0; p < N; ++p) {
                    s += G[i * N + p] * F[0 * npe + p + N * (j + N * k)];
                    s += G[j * N + p] * F[1 * npe + i + N * (p + N * k)];
                    s += G[k * N + p] * F[2 * npe + i + N * (j + N * p)];
                }
                if (with_adv) {
                    const double* o = co + 13 * n;
                    double gv[3];
                    grad_node(c, ve, i, j, k, gv);
                    s += o[6] * gv[0] + o[7] * gv[1] + o[8] * gv[2];
                    for (int p = 0; p < N; ++p) {
                        s += G[i * N + p] * cvv[0 * npe + p + N * (j + N * k)];
                        s += G[j * N + p] * cvv[1 * npe + i + N * (p + N * k)];
                        s += G[k * N + p] * cvv[2 * npe + i + N * (j + N * p)];
                    }
                }
                ye[n] = s;
            }
            /* faces */
            for (int a = 0; a < 3; ++a)
            for (int sd = 0; sd < 2; ++sd) {
                /* neighbour element, its face fn (axis an, row rn, outward
                 * sign nrm_n), the D4 code of the face node map, the tag */
                long en;
                int fn, ori = 0, tag = 0;
                if (c->fconn) {
                    const int* fc = c->fconn + 4 * (6 * e + 2 * a + sd);
                    en = fc[0]; fn = fc[1]; ori = fc[2]; tag = fc[3];
                } else {
                    en = neighbour(c, e, a, sd);
                    fn = 2 * a + (1 - sd);
                }
                int an = fn / 2;
                int r = sd ? N - 1 : 0, rn = (fn & 1) ? N - 1 : 0;
                double nrm = sd ? 1.0 : -1.0, nrm_n = (fn & 1) ? 1.0 : -1.0;
                double alpha = en >= 0 ? 0.5 : 1.0;
                double gval = c->fconn ? (tag >= 1 && tag <= 6 ? bv[tag - 1] : 0.0) : bv[2 * a + sd];
                const double* un = en >= 0 ? u + en * npe : NULL;
                const double* vn = (en >= 0 && with_adv) ? v + en * npe : NULL;
                const double* con = en >= 0 ? c->coef + 13 * en * (long)npe : NULL;
                int som = en < 0 && (c->fconn ? tag == 7 : c->d.face_kind[2 * a + sd] == 7);
                /* excision face (tag 8, src/wave_strong_rhs.jl:188-213): no SAT */
                if (en < 0 && c->fconn && tag == 8) continue;
                for (int q = 0; q < N; ++q)
                for (int p = 0; p < N; ++p) {
                    int n = fidx(N, a, r, p, q);
                    int pn, qn;
                    orient_map(ori, N, p, q, &pn, &qn);
                    int nn = fidx(N, an, rn, pn, qn);
                    const double* o = co + 13 * n;
                    if (som) {
                        /* Sommerfeld face (src/wave.jl:444-518): no SIPG terms,
                         * and the drag -(wF/Hphys)(u_t + d_n u + u/R).  Unweighted
                         * (times wgt = 1/|J| later): |n_u| d_n u = n K^ab G_b u
                         * (outward normal n), |n_u| = |J| |grad xi_a| =
                         * sqrt(K^aa / wgt), and wF/Hphys = |n_u| / (w_r |J|). */
                        double gs[3];
                        int i0 = n % N, j0 = (n / N) % N, k0 = n / (N * N);
                        grad_node(c, ue, i0, j0, k0, gs);
                        double Fs = Kab(o, a, 0) * gs[0] + Kab(o, a, 1) * gs[1] + Kab(o, a, 2) * gs[2];
                        double nu = sqrt(Kab(o, a, a) / o[9]);
                        double vv = v ? v[e * npe + n] : 0.0;
                        ye[n] -= (nrm * Fs + nu * (vv + c->d.sommerfeld_inv_R * ue[n])) / w[r];
                        continue;
                    }
                    double du, dF = 0.0, kap = o[10 + a];
                    if (en >= 0) {
                        du = ue[n] - un[nn];
                        int ii = nn % N, jj = (nn / N) % N, kk = nn / (N * N);
                        double gs[3], gn[3];
                        int i0 = n % N, j0 = (n / N) % N, k0 = n / (N * N);
                        grad_node(c, ue, i0, j0, k0, gs);
                        grad_node(c, un, ii, jj, kk, gn);
                        const double* on = con + 13 * nn;
                        double Fs = Kab(o, a, 0) * gs[0] + Kab(o, a, 1) * gs[1] + Kab(o, a, 2) * gs[2];
                        double Fn = Kab(on, an, 0) * gn[0] + Kab(on, an, 1) * gn[1] + Kab(on, an, 2) * gn[2];
                        /* average outward flux: the neighbour's outward normal
                         * is nrm_n along its own axis an */
                        dF = Fs + nrm * nrm_n * Fn;
                        kap = 0.5 * (o[10 + a] + on[10 + an]);
                    } else {
                        du = ue[n] - gval;
                    }
                    ye[n] += (-0.5 * nrm * dF - tau * kap * du) / w[r];
                    /* symmetrising term, b = a: along the normal line */
                    double sa = alpha * nrm * Kab(o, a, a) * du;
                    for (int m = 0; m < N; ++m)
                        ye[fidx(N, a, m, p, q)] += G[r * N + m] * sa / w[m];
                    /* b != a: along the tangential directions within the face */
                    int tb[2], tv[2];
                    if (a == 0) { tb[0] = 1; tb[1] = 2; }
                    else if (a == 1) { tb[0] = 0; tb[1] = 2; }
                    else { tb[0] = 0; tb[1] = 1; }
                    tv[0] = p; tv[1] = q;
                    for (int t = 0; t < 2; ++t) {
                        int b = tb[t];
                        double sb = alpha * nrm * Kab(o, a, b) * du * w[tv[t]];
                        for (int m = 0; m < N; ++m) {
                            int pm = t == 0 ? m : p, qm = t == 0 ? q : m;
                            ye[fidx(N, a, r, pm, qm)] += G[tv[t] * N + m] * sb / (w[m] * w[r]);
                        }
                    }
                    if (with_adv) {
                        double vnb = vn ? vn[nn] : 0.0;
                        ye[n] += -nrm * o[6 + a] * (ve[n] - vnb) / w[r];
                    }
                }
            }
        }
        free(F); free(cvv);
    }
}

/* ------------------------------------------------------------ public */

static const double zero6[6] = {0, 0, 0, 0, 0, 0};

/* Boundary value per (axis, side): the outer face f = 2a+side+1 carries tag f
 * (faces 1..6 tagged 1..6), so the Dirichlet value is bdry_values[f-1]
 * (src/wave.jl:567-571). */
void wto_lap(const wto_ctx* c, const double* u, const double* ud, double* udd)
{
    if (!c->curved) { lap_flat(c, u, ud, c->d.bdry_values, udd); return; }
    op_curved(c, u, ud, c->d.bdry_values, ud != NULL, udd);
    #pragma omp parallel for schedule(static)
    for (long n = 0; n < c->npts; ++n) udd[n] *= c->coef[13 * n + 9];
}

/* out += eps mu^-5 sum_d D_d^6 f with the six-fold ping-pong of
 * src/wave3d_curved.jl:182-192.  D_d is apply_D!'s SAT-coupled derivative
 * except that outer faces see a zero ghost value (H D skew, so H D^6 <= 0 on
 * the bounded box as well); mu stays the reference's apply_D! value. */
static void ko_add(const wto_ctx* c, const double* f, double* out, double* s1, double* s2)
{
    double koT = c->d.eps_ko * c->inv_mu5;
    for (int a = 0; a < 3; ++a) {
        apply_D_g(c, a, f, s1, 1);
        apply_D_g(c, a, s1, s2, 1);
        apply_D_g(c, a, s2, s1, 1);
        apply_D_g(c, a, s1, s2, 1);
        apply_D_g(c, a, s2, s1, 1);
        apply_D_g(c, a, s1, s2, 1);
        #pragma omp parallel for schedule(static)
        for (long n = 0; n < c->npts; ++n) out[n] += koT * s2[n];
    }
}

/* Second-order-in-space system u_t = v, v_t = lap(u, v) + KO(v).  KO acts on
 * u_t only: with H L symmetric NSD and H K NSD the energy
 * 1/2 v^T W v - 1/2 u^T H L u then obeys dE/dt = v^T H K v <= 0, whereas KO
 * on u as well (the reference's first-order system damps both Phi and Pi,
 * src/wave3d_curved.jl:241-248) breaks that bound here because D^6 does not
 * commute with the SIPG L. */
void wto_rhs(const wto_ctx* c, const double* u, const double* v, double* du, double* dv)
{
    wto_lap(c, u, v, dv);
    memcpy(du, v, sizeof(double) * c->npts);
    if (c->d.eps_ko != 0.0) {
        double* s1 = (double*)malloc(sizeof(double) * c->npts);
        double* s2 = (double*)malloc(sizeof(double) * c->npts);
        ko_add(c, v, dv, s1, s2);
        free(s1); free(s2);
    }
}

/* E = 1/2 <v, v>_W + 1/2 <u, -Ltilde u>_H3 with zero boundary data
 * (test/kernels3d_tests.jl:32-47); W = Hphys on flat meshes and H3/wgt =
 * H3 sqrt(g)|J|/alpha with a background. */
double wto_energy(const wto_ctx* c, const double* u, const double* v)
{
    double* Lu = (double*)malloc(sizeof(double) * c->npts);
    double E = 0.0;
    if (!c->curved) {
        lap_flat(c, u, NULL, zero6, Lu);
        #pragma omp parallel for reduction(+ : E) schedule(static)
        for (long n = 0; n < c->npts; ++n)
            E += 0.5 * c->hphys[n] * (v[n] * v[n] - u[n] * Lu[n]);
    } else {
        op_curved(c, u, NULL, zero6, 0, Lu);
        int N = c->N;
        #pragma omp parallel for reduction(+ : E) schedule(static)
        for (long n = 0; n < c->npts; ++n) {
            int i = (int)(n % N), j = (int)((n / N) % N), k = (int)((n / (N * N)) % N);
            double H3 = c->w[i] * c->w[j] * c->w[k];
            E += 0.5 * (H3 / c->coef[13 * n + 9] * v[n] * v[n] - H3 * u[n] * Lu[n]);
        }
    }
    free(Lu);
    return E;
}

/* Classical RK4 with the stage/combination order of test/wave3d_tests.jl:32-44. */
void wto_rk4(const wto_ctx* c, double* u, double* v, double dt, int nsteps, double* energy_out)
{
    long n = c->npts;
    double* k = (double*)malloc(sizeof(double) * 8 * n);
    double* us = (double*)malloc(sizeof(double) * n);
    double* vs = (double*)malloc(sizeof(double) * n);
    if (energy_out) energy_out[0] = wto_energy(c, u, v);
    for (int st = 0; st < nsteps; ++st) {
        wto_rhs(c, u, v, k + 0 * n, k + 1 * n);
        #pragma omp parallel for schedule(static)
        for (long q = 0; q < n; ++q) { us[q] = u[q] + dt / 2 * k[q]; vs[q] = v[q] + dt / 2 * k[n + q]; }
        wto_rhs(c, us, vs, k + 2 * n, k + 3 * n);
        #pragma omp parallel for schedule(static)
        for (long q = 0; q < n; ++q) { us[q] = u[q] + dt / 2 * k[2 * n + q]; vs[q] = v[q] + dt / 2 * k[3 * n + q]; }
        wto_rhs(c, us, vs, k + 4 * n, k + 5 * n);
        #pragma omp parallel for schedule(static)
        for (long q = 0; q < n; ++q) { us[q] = u[q] + dt * k[4 * n + q]; vs[q] = v[q] + dt * k[5 * n + q]; }
        wto_rhs(c, us, vs, k + 6 * n, k + 7 * n);
        #pragma omp parallel for schedule(static)
        for (long q = 0; q < n; ++q) {
            u[q] += dt / 6 * (k[q] + 2 * k[2 * n + q] + 2 * k[4 * n + q] + k[6 * n + q]);
            v[q] += dt / 6 * (k[n + q] + 2 * k[3 * n + q] + 2 * k[5 * n + q] + k[7 * n + q]);
        }
        if (energy_out) energy_out[st + 1] = wto_energy(c, u, v);
    }
    free(k); free(us); free(vs);
}

int wto_assemble_1d(int N, int M, int periodic, double tau, double h, int kind, double* A)
{
    wto_ctx c;
    memset(&c, 0, sizeof(c));
    c.N = N;
    c.d.tau = tau;
    if (wto_gll(N, c.xi, c.w, c.G)) return -1;
    int L = M * N;
    double* x = (double*)calloc(L, sizeof(double));
    double* y = (double*)calloc(L, sizeof(double));
    for (int col = 0; col < L; ++col) {
        memset(x, 0, sizeof(double) * L);
        x[col] = 1.0;
        if (kind == 1 || kind == 2) {
            apply_D_line1d_g(&c, M, periodic, h, x, y, kind == 2);
        } else {
            memset(y, 0, sizeof(double) * L);
            for (int e = 0; e < M; ++e) {
                const double* ul = NULL; const double* ur = NULL;
                if (e > 0) ul = x + (e - 1) * N; else if (periodic) ul = x + (M - 1) * N;
                if (e < M - 1) ur = x + (e + 1) * N; else if (periodic) ur = x;
                sipg_segment(&c, h, x + e * N, ul, 0.0, ur, 0.0, y + e * N);
            }
        }
        for (int row = 0; row < L; ++row) A[(long)row * L + col] = y[row];
    }
    free(x); free(y);
    return 0;
}

/* ------------------------------------------ first-order (Phi, Pi) system */

/* coef11 from the background (sample_background3d!'s per-node kernel,
 * src/wave3d_curved.jl:75-96): alpha, sqrt(det gamma), beta1..3 and the
 * cofactor inverse gamma^{ij}; field q of node n at coef11[q * n_pts + n]. */
void wto_fo_coef(long npts, const double* bg10, double* coef11)
{
    #pragma omp parallel for schedule(static)
    for (long n = 0; n < npts; ++n) {
        const double* b = bg10 + 10 * n;
        double g11 = b[4], g12 = b[5], g13 = b[6], g22 = b[7], g23 = b[8], g33 = b[9];
        double det = g11 * (g22 * g33 - g23 * g23) - g12 * (g12 * g33 - g23 * g13) + g13 * (g12 * g23 - g22 * g13);
        double* o = coef11;
        o[0 * npts + n] = b[0];
        o[1 * npts + n] = sqrt(det);
        o[2 * npts + n] = b[1];
        o[3 * npts + n] = b[2];
        o[4 * npts + n] = b[3];
        o[5 * npts + n] = (g22 * g33 - g23 * g23) / det;
        o[6 * npts + n] = (g13 * g23 - g12 * g33) / det;
        o[7 * npts + n] = (g12 * g23 - g13 * g22) / det;
        o[8 * npts + n] = (g11 * g33 - g13 * g13) / det;
        o[9 * npts + n] = (g12 * g13 - g11 * g23) / det;
        o[10 * npts + n] = (g11 * g22 - g12 * g12) / det;
    }
}

/* wave3d_curved_rhs! on the axis-aligned affine path (metric === nothing,
 * src/wave3d_curved.jl:205-258):
 *   DPhi_d = apply_D!(Phi, d),  F^i = b^i Pi + alpha sqrtg g^{ij} DPhi_j,
 *   dPi = sum_i apply_D!(F^i, i),  dPhi = b . DPhi + (alpha / sqrtg) Pi,
 * then eps mu^-5 D_d^6 on both fields with the one-sided apply_D!
 * (_ko_axis3!, :182-192), then the affine outer-boundary pass
 * (_bc3d_affine_kernel!, src/boundaries3d.jl:76-126) when bc is given. */
void wto_fo_rhs(const wto_ctx* c, const double* Phi, const double* Pi, const double* coef11,
                double eps_ko, const wto_bc3d* bc, double* dPhi, double* dPi)
{
    long n = c->npts;
    const double *al = coef11, *sg = coef11 + n, *b[3] = {coef11 + 2 * n, coef11 + 3 * n, coef11 + 4 * n};
    const double* gu = coef11 + 5 * n;    /* 11 12 13 22 23 33 */
    static const int gmap[3][3] = {{0, 1, 2}, {1, 3, 4}, {2, 4, 5}};
    double* D = (double*)malloc(sizeof(double) * 3 * n);
    double* F = (double*)malloc(sizeof(double) * n);
    double* s = (double*)malloc(sizeof(double) * n);
    for (int d = 0; d < 3; ++d) apply_D_g(c, d, Phi, D + d * n, 0);
    for (int i = 0; i < 3; ++i) {
        #pragma omp parallel for schedule(static)
        for (long q = 0; q < n; ++q)
            F[q] = b[i][q] * Pi[q] + al[q] * sg[q] * (gu[gmap[i][0] * n + q] * D[q] + gu[gmap[i][1] * n + q] * D[n + q] +
                                                     gu[gmap[i][2] * n + q] * D[2 * n + q]);
        apply_D_g(c, i, F, s, 0);
        #pragma omp parallel for schedule(static)
        for (long q = 0; q < n; ++q) dPi[q] = (i == 0 ? 0.0 : dPi[q]) + s[q];
    }
    #pragma omp parallel for schedule(static)
    for (long q = 0; q < n; ++q)
        dPhi[q] = b[0][q] * D[q] + b[1][q] * D[n + q] + b[2][q] * D[2 * n + q] + (al[q] / sg[q]) * Pi[q];
    if (eps_ko != 0.0) {
        double koT = eps_ko * c->inv_mu5;
        double* s2 = (double*)malloc(sizeof(double) * n);
        for (int fld = 0; fld < 2; ++fld) {
            const double* x = fld == 0 ? Phi : Pi;
            double* out = fld == 0 ? dPhi : dPi;
            for (int d = 0; d < 3; ++d) {
                apply_D_g(c, d, x, s, 0);
                apply_D_g(c, d, s, s2, 0);
                apply_D_g(c, d, s2, s, 0);
                apply_D_g(c, d, s, s2, 0);
                apply_D_g(c, d, s2, s, 0);
                apply_D_g(c, d, s, s2, 0);
                #pragma omp parallel for schedule(static)
                for (long q = 0; q < n; ++q) out[q] += koT * s2[q];
            }
        }
        free(s2);
    }
    if (bc) {
        int N = c->N;
        #pragma omp parallel for schedule(static)
        for (long q = 0; q < n; ++q) {
            long e = q / c->npe;
            int r = (int)(q % c->npe);
            int nd[3] = {r % N, (r / N) % N, r / (N * N)};
            long ec[3] = {e % c->d.M[0], (e / c->d.M[0]) % c->d.M[1], e / ((long)c->d.M[0] * c->d.M[1])};
            double dF = 0.0, dP = 0.0;
            for (int f = 0; f < 6; ++f) {
                int dn = f / 2, hi = f % 2;
                if (c->d.periodic[dn]) continue;                       /* no outer face */
                if (ec[dn] != (hi ? c->d.M[dn] - 1 : 0)) continue;     /* bdry[f, e] == 0 */
                int kind = bc->kinds[f];
                if (kind == WTO_BC_EXCISION) continue;
                int fn = hi ? N - 1 : 0;
                if (nd[dn] != fn) continue;
                double nh = hi ? 1.0 : -1.0;
                double a = al[q] / sg[q];
                double gunn = gu[gmap[dn][dn] * n + q];
                double a_n = al[q] * sqrt(gunn);
                double bax = b[dn][q];
                double bn = nh * bax;
                double wt = (1.0 / c->h[dn]) / c->w[fn];
                if (kind == WTO_BC_FULL_DIRICHLET) {
                    double tau = bc->sigma * (fabs(a_n - bax) + fabs(a_n + bax)) * wt;
                    double gF = bc->gPhi ? bc->gPhi[q] : 0.0, gP = bc->gPi ? bc->gPi[q] : 0.0;
                    dF += -tau * (Phi[q] - gF);
                    dP += -tau * (Pi[q] - gP);
                } else {
                    double qn = nh * D[dn * n + q];
                    double rr = Pi[q] + ((bn + a_n) / a) * qn;
                    double g = (kind == WTO_BC_SOMMERFELD || !bc->gPhi) ? 0.0 : bc->gPhi[q];
                    double s_in = a_n + bn;
                    dP += -bc->sigma * fabs(s_in) * wt * (rr - g);
                }
            }
            dPhi[q] += dF;
            dPi[q] += dP;
        }
    }
    free(D); free(F); free(s);
}

/* wave3d_energy (src/wave3d_curved.jl:268-287):
 * E = sum 1/2 [Pi^2 / sqrtg + sqrtg g^{ij} DPhi_i DPhi_j] Hphys. */
double wto_fo_energy(const wto_ctx* c, const double* Phi, const double* Pi, const double* coef11)
{
    long n = c->npts;
    const double *sg = coef11 + n, *gu = coef11 + 5 * n;
    double* D = (double*)malloc(sizeof(double) * 3 * n);
    for (int d = 0; d < 3; ++d) apply_D_g(c, d, Phi, D + d * n, 0);
    double E = 0.0;
    #pragma omp parallel for reduction(+ : E) schedule(static)
    for (long q = 0; q < n; ++q) {
        double d1 = D[q], d2 = D[n + q], d3 = D[2 * n + q];
        E += (Pi[q] * Pi[q] / sg[q] +
              sg[q] * (gu[q] * d1 * d1 + gu[3 * n + q] * d2 * d2 + gu[5 * n + q] * d3 * d3 + 2 * gu[n + q] * d1 * d2 +
                       2 * gu[2 * n + q] * d1 * d3 + 2 * gu[4 * n + q] * d2 * d3)) * c->hphys[q] / 2;
    }
    free(D);
    return E;
}

/* ------------------------------ first-order (Phi, Pi) system, curvilinear */

/* The curvilinear path of wave3d_curved_rhs! (metric given,
 * src/wave3d_curved.jl:218-238, METHODS.md "Curvilinear path"): metric terms
 * in the conservative-curl form computed from the node coordinates with the
 * element's own differentiation matrix,
 *   J a^i_n = (grad_xi x C^n)_i,  C^n = 1/2 (X_l grad_xi X_m - X_m grad_xi X_l),
 * (n, l, m) and (i, j, k) cyclic, so that sum_i D_i(J a^i_n) = 0 discretely
 * (free stream); the split form 1/2 (conservative + advective) of the
 * gradient and the divergence (a skew-adjoint pair in the mass
 * Hd = w_i w_j w_k |J|) with the centred-flux SAT across element faces
 * (neighbour values through the D4 face map), one-sided at outer faces.
 * The metric is stored with the element's orientation divided out
 * (JAp = sign(J) J a, |J|), so the formulas hold on left-handed elements. */

static void elem_dref(const wto_ctx* c, const double* f, int b, int i, int j, int k, double* out)
{
    int N = c->N;
    int ii[3] = {i, j, k};
    double s = 0.0;
    for (int p = 0; p < N; ++p) {
        int jj[3] = {i, j, k};
        jj[b] = p;
        s += c->G[ii[b] * N + p] * f[jj[0] + N * (jj[1] + N * jj[2])];
    }
    *out = s;
}

static void curv_setup(wto_ctx* c)
{
    int N = c->N, npe = c->npe;
    long Ne = c->Ne, n = c->npts;
    c->JAp = (double*)malloc(sizeof(double) * 9 * n);
    c->Jp = (double*)malloc(sizeof(double) * n);
    c->ell = (double*)malloc(sizeof(double) * 3 * Ne);
    #pragma omp parallel for schedule(static)
    for (long e = 0; e < Ne; ++e) {
        double* X = (double*)malloc(sizeof(double) * 3 * npe);
        double* C = (double*)malloc(sizeof(double) * 9 * npe);   /* C^n_k at [(n*3+k)*npe + node] */
        for (int q = 0; q < npe; ++q)
            for (int a = 0; a < 3; ++a) X[a * npe + q] = c->coords[3 * (e * npe + q) + a];
        for (int q = 0; q < npe; ++q) {
            int i = q % N, j = (q / N) % N, k = q / (N * N);
            double dX[3][3];   /* dX[m][kk] = D_kk X_m */
            for (int m = 0; m < 3; ++m)
                for (int kk = 0; kk < 3; ++kk) elem_dref(c, X + m * npe, kk, i, j, k, &dX[m][kk]);
            for (int nn = 0; nn < 3; ++nn) {
                int l = (nn + 1) % 3, m = (nn + 2) % 3;
                for (int kk = 0; kk < 3; ++kk)
                    C[(nn * 3 + kk) * npe + q] = 0.5 * (X[l * npe + q] * dX[m][kk] - X[m * npe + q] * dX[l][kk]);
            }
            double det = dX[0][0] * (dX[1][1] * dX[2][2] - dX[1][2] * dX[2][1]) -
                         dX[0][1] * (dX[1][0] * dX[2][2] - dX[1][2] * dX[2][0]) +
                         dX[0][2] * (dX[1][0] * dX[2][1] - dX[1][1] * dX[2][0]);
            c->Jp[e * npe + q] = det;   /* signed for now */
        }
        for (int q = 0; q < npe; ++q) {
            int i = q % N, j = (q / N) % N, k = q / (N * N);
            double sg = c->Jp[e * npe + q] < 0 ? -1.0 : 1.0;
            for (int ia = 0; ia < 3; ++ia) {
                int ja = (ia + 1) % 3, ka = (ia + 2) % 3;
                for (int nn = 0; nn < 3; ++nn) {
                    double d1, d2;
                    elem_dref(c, C + (nn * 3 + ka) * npe, ja, i, j, k, &d1);
                    elem_dref(c, C + (nn * 3 + ja) * npe, ka, i, j, k, &d2);
                    c->JAp[9 * (e * npe + q) + ia * 3 + nn] = sg * (d1 - d2);
                }
            }
        }
        for (int q = 0; q < npe; ++q) c->Jp[e * npe + q] = fabs(c->Jp[e * npe + q]);
        for (int a = 0; a < 3; ++a) {
            double lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};
            for (int qq = 0; qq < N; ++qq)
                for (int p = 0; p < N; ++p) {
                    int n0 = fidx(N, a, 0, p, qq), n1 = fidx(N, a, N - 1, p, qq);
                    for (int m = 0; m < 3; ++m) { lo[m] += X[m * npe + n0]; hi[m] += X[m * npe + n1]; }
                }
            double d2 = 0.0;
            for (int m = 0; m < 3; ++m) d2 += (hi[m] - lo[m]) * (hi[m] - lo[m]);
            c->ell[3 * e + a] = sqrt(d2) / (N * N);
        }
        free(X); free(C);
    }
    c->mu_c = -1.0;
}

int wto_curv_metric(const wto_ctx* c, double* JA9, double* Jp)
{
    if (!c->JAp) return -1;
    memcpy(JA9, c->JAp, sizeof(double) * 9 * c->npts);
    memcpy(Jp, c->Jp, sizeof(double) * c->npts);
    return 0;
}

/* the face (axis a, side sd) of element e: neighbour, its face, D4 code, tag */
static void curv_face(const wto_ctx* c, long e, int f, long* en, int* fn, int* ori, int* tag)
{
    const int* fc = c->fconn + 4 * (6 * e + f);
    *en = fc[0]; *fn = fc[1]; *ori = fc[2]; *tag = fc[3];
}

/* D = grad Phi (3 fields of npts): split form + centred SAT. */
void wto_curv_grad(const wto_ctx* c, const double* Phi, double* D3)
{
    int N = c->N, npe = c->npe;
    long n = c->npts;
    #pragma omp parallel for schedule(static)
    for (long e = 0; e < c->Ne; ++e) {
        const double* fe = Phi + e * npe;
        const double* JA = c->JAp + 9 * e * npe;
        double* W = (double*)malloc(sizeof(double) * 3 * npe);
        double out[3][WTO_NMAX * WTO_NMAX * WTO_NMAX];
        for (int q = 0; q < npe; ++q) for (int m = 0; m < 3; ++m) out[m][q] = 0.0;
        for (int nn = 0; nn < 3; ++nn) {
            for (int q = 0; q < npe; ++q)
                for (int a = 0; a < 3; ++a) W[a * npe + q] = JA[9 * q + a * 3 + nn] * fe[q];
            for (int q = 0; q < npe; ++q) {
                int i = q % N, j = (q / N) % N, k = q / (N * N);
                double cons = 0.0, adv = 0.0;
                for (int a = 0; a < 3; ++a) {
                    double d1, d2;
                    elem_dref(c, W + a * npe, a, i, j, k, &d1);
                    elem_dref(c, fe, a, i, j, k, &d2);
                    cons += d1;
                    adv += JA[9 * q + a * 3 + nn] * d2;
                }
                out[nn][q] = 0.5 * (cons + adv) / c->Jp[e * npe + q];
            }
        }
        for (int f = 0; f < 6; ++f) {
            long en; int fn, ori, tag;
            curv_face(c, e, f, &en, &fn, &ori, &tag);
            if (en < 0) continue;
            int a = f / 2, r = (f & 1) ? N - 1 : 0, an = fn / 2, rn = (fn & 1) ? N - 1 : 0;
            double s = (f & 1) ? 1.0 : -1.0;
            for (int qq = 0; qq < N; ++qq)
                for (int p = 0; p < N; ++p) {
                    int pn, qn;
                    orient_map(ori, N, p, qq, &pn, &qn);
                    int q = fidx(N, a, r, p, qq), qnb = fidx(N, an, rn, pn, qn);
                    double jump = Phi[en * npe + qnb] - fe[q];
                    double sc = 0.5 * s * jump / (c->w[r] * c->Jp[e * npe + q]);
                    for (int nn = 0; nn < 3; ++nn) out[nn][q] += sc * JA[9 * q + a * 3 + nn];
                }
        }
        for (int nn = 0; nn < 3; ++nn)
            for (int q = 0; q < npe; ++q) D3[nn * n + e * npe + q] = out[nn][q];
        free(W);
    }
}

/* out = div F (F: 3 fields of npts): split form + centred SAT. */
void wto_curv_div(const wto_ctx* c, const double* F3, double* out)
{
    int N = c->N, npe = c->npe;
    long n = c->npts;
    #pragma omp parallel for schedule(static)
    for (long e = 0; e < c->Ne; ++e) {
        const double* JA = c->JAp + 9 * e * npe;
        double* W = (double*)malloc(sizeof(double) * 3 * npe);
        double Fe[3][WTO_NMAX * WTO_NMAX * WTO_NMAX];
        for (int q = 0; q < npe; ++q)
            for (int m = 0; m < 3; ++m) Fe[m][q] = F3[m * n + e * npe + q];
        for (int q = 0; q < npe; ++q)
            for (int a = 0; a < 3; ++a)
                W[a * npe + q] = JA[9 * q + a * 3 + 0] * Fe[0][q] + JA[9 * q + a * 3 + 1] * Fe[1][q] +
                                 JA[9 * q + a * 3 + 2] * Fe[2][q];
        for (int q = 0; q < npe; ++q) {
            int i = q % N, j = (q / N) % N, k = q / (N * N);
            double cons = 0.0, adv = 0.0;
            for (int a = 0; a < 3; ++a) {
                double d1;
                elem_dref(c, W + a * npe, a, i, j, k, &d1);
                cons += d1;
                for (int m = 0; m < 3; ++m) {
                    double d2;
                    elem_dref(c, Fe[m], a, i, j, k, &d2);
                    adv += JA[9 * q + a * 3 + m] * d2;
                }
            }
            out[e * npe + q] = 0.5 * (cons + adv) / c->Jp[e * npe + q];
        }
        for (int f = 0; f < 6; ++f) {
            long en; int fn, ori, tag;
            curv_face(c, e, f, &en, &fn, &ori, &tag);
            if (en < 0) continue;
            int a = f / 2, r = (f & 1) ? N - 1 : 0, an = fn / 2, rn = (fn & 1) ? N - 1 : 0;
            double s = (f & 1) ? 1.0 : -1.0, sn = (fn & 1) ? 1.0 : -1.0;
            for (int qq = 0; qq < N; ++qq)
                for (int p = 0; p < N; ++p) {
                    int pn, qn;
                    orient_map(ori, N, p, qq, &pn, &qn);
                    int q = fidx(N, a, r, p, qq), qnb = fidx(N, an, rn, pn, qn);
                    long gn = en * npe + qnb;
                    const double* JAn = c->JAp + 9 * gn;
                    double Wn = JAn[an * 3 + 0] * F3[gn] + JAn[an * 3 + 1] * F3[n + gn] + JAn[an * 3 + 2] * F3[2 * n + gn];
                    double qo = s * W[a * npe + q], qb = sn * Wn;
                    out[e * npe + q] += -0.5 * (qo + qb) / (c->w[r] * c->Jp[e * npe + q]);
                }
        }
        free(W);
    }
}

/* The KO derivative on a general mesh, along each element's reference axis
 * `axis`: (1/ell) [G u + s/(2 w_r) (u_nbr - u)] at faces with a neighbour,
 * one-sided at outer faces (apply_D!'s form with the element's extent ell
 * along that axis in place of h). */
void wto_curv_D(const wto_ctx* c, int axis, const double* in, double* out)
{
    int N = c->N, npe = c->npe;
    #pragma omp parallel for schedule(static)
    for (long e = 0; e < c->Ne; ++e) {
        const double* fe = in + e * npe;
        double il = 1.0 / c->ell[3 * e + axis];
        for (int q = 0; q < npe; ++q) {
            int i = q % N, j = (q / N) % N, k = q / (N * N);
            double d;
            elem_dref(c, fe, axis, i, j, k, &d);
            out[e * npe + q] = il * d;
        }
        for (int sd = 0; sd < 2; ++sd) {
            int f = 2 * axis + sd;
            long en; int fn, ori, tag;
            curv_face(c, e, f, &en, &fn, &ori, &tag);
            if (en < 0) continue;
            int r = sd ? N - 1 : 0, an = fn / 2, rn = (fn & 1) ? N - 1 : 0;
            double s = sd ? 1.0 : -1.0;
            for (int qq = 0; qq < N; ++qq)
                for (int p = 0; p < N; ++p) {
                    int pn, qn;
                    orient_map(ori, N, p, qq, &pn, &qn);
                    int q = fidx(N, axis, r, p, qq), qnb = fidx(N, an, rn, pn, qn);
                    out[e * npe + q] += il * s * (in[en * npe + qnb] - fe[q]) / (2.0 * c->w[r]);
                }
        }
    }
}

/* mu of the curvilinear KO derivative: 30 power iterations of D_1^2 from the
 * alternating start vector isodd(i+j+k+m) (src/wave3d_curved.jl:158-178). */
double wto_curv_mu(const wto_ctx* c)
{
    wto_ctx* cc = (wto_ctx*)c;
    if (cc->mu_c >= 0) return cc->mu_c;
    long n = c->npts;
    int N = c->N, npe = c->npe;
    double* v = (double*)malloc(sizeof(double) * n);
    double* w = (double*)malloc(sizeof(double) * n);
    for (long q = 0; q < n; ++q) {
        long e = q / npe;
        int r = (int)(q % npe);
        int i = r % N, j = (r / N) % N, k = r / (N * N);
        v[q] = ((i + 1 + j + 1 + k + 1 + e + 1) % 2 == 1) ? 1.0 : -1.0;
    }
    double mu2 = 0.0;
    for (int it = 0; it < 30; ++it) {
        wto_curv_D(c, 0, v, w);
        wto_curv_D(c, 0, w, v);
        double s = 0.0;
        for (long q = 0; q < n; ++q) s += v[q] * v[q];
        mu2 = sqrt(s);
        if (!(mu2 > 0)) break;
        for (long q = 0; q < n; ++q) v[q] /= mu2;
    }
    cc->mu_c = sqrt(mu2);
    cc->inv_mu5_c = cc->mu_c > 0 ? pow(1.0 / cc->mu_c, 5) : 0.0;
    free(v); free(w);
    return cc->mu_c;
}

/* wave3d_curved_rhs! with metric (src/wave3d_curved.jl:218-256): the
 * gradient and divergence above, KO eps mu^-5 D_d^6 per reference axis on both
 * fields, then the curvilinear boundary pass _apply_bc3d_curv!
 * (src/boundaries3d.jl:131-236): one kind bc->kinds[0] on every outer face
 * whose tag is not bc->excision_tag, with the physical outward unit normal
 * n = s JAp^a / |JAp^a| and the lift wt = |JAp^a| / (w_r |J|). */
void wto_fo_rhs_curv(const wto_ctx* c, const double* Phi, const double* Pi, const double* coef11,
                     double eps_ko, const wto_bc3d* bc, double* dPhi, double* dPi)
{
    long n = c->npts;
    int N = c->N, npe = c->npe;
    const double *al = coef11, *sg = coef11 + n, *b[3] = {coef11 + 2 * n, coef11 + 3 * n, coef11 + 4 * n};
    const double* gu = coef11 + 5 * n;
    static const int gmap[3][3] = {{0, 1, 2}, {1, 3, 4}, {2, 4, 5}};
    double* D = (double*)malloc(sizeof(double) * 3 * n);
    double* F = (double*)malloc(sizeof(double) * 3 * n);
    wto_curv_grad(c, Phi, D);
    #pragma omp parallel for schedule(static)
    for (long q = 0; q < n; ++q) {
        for (int i = 0; i < 3; ++i)
            F[i * n + q] = b[i][q] * Pi[q] + al[q] * sg[q] * (gu[gmap[i][0] * n + q] * D[q] + gu[gmap[i][1] * n + q] * D[n + q] +
                                                            gu[gmap[i][2] * n + q] * D[2 * n + q]);
        dPhi[q] = b[0][q] * D[q] + b[1][q] * D[n + q] + b[2][q] * D[2 * n + q] + (al[q] / sg[q]) * Pi[q];
    }
    wto_curv_div(c, F, dPi);
    if (eps_ko != 0.0) {
        double koT = eps_ko * pow(1.0 / wto_curv_mu(c), 5);
        double* s1 = (double*)malloc(sizeof(double) * n);
        double* s2 = (double*)malloc(sizeof(double) * n);
        for (int fld = 0; fld < 2; ++fld) {
            const double* x = fld == 0 ? Phi : Pi;
            double* o = fld == 0 ? dPhi : dPi;
            for (int d = 0; d < 3; ++d) {
                wto_curv_D(c, d, x, s1);
                wto_curv_D(c, d, s1, s2);
                wto_curv_D(c, d, s2, s1);
                wto_curv_D(c, d, s1, s2);
                wto_curv_D(c, d, s2, s1);
                wto_curv_D(c, d, s1, s2);
                for (long q = 0; q < n; ++q) o[q] += koT * s2[q];
            }
        }
        free(s1); free(s2);
    }
    if (bc) {
        int kind = bc->kinds[0];
        for (long e = 0; e < c->Ne; ++e)
            for (int f = 0; f < 6; ++f) {
                long en; int fn, ori, tag;
                curv_face(c, e, f, &en, &fn, &ori, &tag);
                if (en >= 0 || tag == bc->excision_tag || kind == WTO_BC_EXCISION) continue;
                int a = f / 2, r = (f & 1) ? N - 1 : 0;
                double s = (f & 1) ? 1.0 : -1.0;
                for (int qq = 0; qq < N; ++qq)
                    for (int p = 0; p < N; ++p) {
                        long q = e * npe + fidx(N, a, r, p, qq);
                        const double* JA = c->JAp + 9 * q;
                        double nf[3] = {s * JA[a * 3 + 0], s * JA[a * 3 + 1], s * JA[a * 3 + 2]};
                        double JF = sqrt(nf[0] * nf[0] + nf[1] * nf[1] + nf[2] * nf[2]);
                        double nx = nf[0] / JF, ny = nf[1] / JF, nz = nf[2] / JF;
                        double av = al[q] / sg[q];
                        double a_n = al[q] * sqrt(gu[q] * nx * nx + gu[3 * n + q] * ny * ny + gu[5 * n + q] * nz * nz +
                                                  2 * gu[n + q] * nx * ny + 2 * gu[2 * n + q] * nx * nz +
                                                  2 * gu[4 * n + q] * ny * nz);
                        double bn = b[0][q] * nx + b[1][q] * ny + b[2][q] * nz;
                        double wt = JF / (c->w[r] * c->Jp[q]);
                        if (kind == WTO_BC_FULL_DIRICHLET) {
                            double tau = bc->sigma * (fabs(a_n - bn) + fabs(a_n + bn)) * wt;
                            double gF = bc->gPhi ? bc->gPhi[q] : 0.0, gP = bc->gPi ? bc->gPi[q] : 0.0;
                            dPhi[q] += -tau * (Phi[q] - gF);
                            dPi[q] += -tau * (Pi[q] - gP);
                        } else {
                            double qn = nx * D[q] + ny * D[n + q] + nz * D[2 * n + q];
                            double rr = Pi[q] + ((bn + a_n) / av) * qn;
                            double g = 0.0;
                            if (kind == WTO_BC_DIRICHLET) {
                                double gP = bc->gPi ? bc->gPi[q] : 0.0;
                                double gd = bc->gD ? nx * bc->gD[q] + ny * bc->gD[n + q] + nz * bc->gD[2 * n + q] : 0.0;
                                g = gP + ((bn + a_n) / av) * gd;
                            }
                            dPi[q] += -bc->sigma * fabs(a_n + bn) * wt * (rr - g);
                        }
                    }
            }
    }
    free(D); free(F);
}

/* wave3d_energy with metric: E = sum 1/2 [Pi^2/sqrtg + sqrtg g^{ij} D_i D_j] Hd,
 * Hd = w_i w_j w_k |J|. */
double wto_fo_energy_curv(const wto_ctx* c, const double* Phi, const double* Pi, const double* coef11)
{
    long n = c->npts;
    int N = c->N, npe = c->npe;
    const double *sg = coef11 + n, *gu = coef11 + 5 * n;
    double* D = (double*)malloc(sizeof(double) * 3 * n);
    wto_curv_grad(c, Phi, D);
    double E = 0.0;
    for (long q = 0; q < n; ++q) {
        int r = (int)(q % npe);
        double Hd = c->w[r % N] * c->w[(r / N) % N] * c->w[r / (N * N)] * c->Jp[q];
        double d1 = D[q], d2 = D[n + q], d3 = D[2 * n + q];
        E += (Pi[q] * Pi[q] / sg[q] +
              sg[q] * (gu[q] * d1 * d1 + gu[3 * n + q] * d2 * d2 + gu[5 * n + q] * d3 * d3 + 2 * gu[n + q] * d1 * d2 +
                       2 * gu[2 * n + q] * d1 * d3 + 2 * gu[4 * n + q] * d2 * d3)) * Hd / 2;
    }
    free(D);
    return E;
}
