/*
 * wt_oracle.h -- CPU FP64 oracle for the 3D SBP-SAT wave RHS and RK4 step.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load this library, and only to check results
 * against (or as the timed CPU baseline).  The product path (libwt2.so) never
 * calls it.
 *
 * It restates, element by element, the reference's hot path:
 *   - GLL element operators (HexSBPSAT make_element/make_operators, used at
 *     src/wave.jl:603; README.md:10-20),
 *   - the SIPG Laplacian of rhs_wave3d! / apply_laplacian! (src/wave.jl:573-611,
 *     formula README.md:57-81),
 *   - the ADM background coefficients of sample_background3d!
 *     (src/wave3d_curved.jl:75-124) folded into a second-order (u, u_t) form,
 *   - Kreiss-Oliger eps*mu^-5*D^6 per axis with the SAT-coupled apply_D!
 *     (src/wave3d_curved.jl:158-192; METHODS.md:256-298), applied to u_t
 *     (see wto_rhs for why), outer faces closed with a zero ghost value,
 *   - classical RK4 (test/wave3d_tests.jl:32-44),
 *   - the H-norm energy (test/kernels3d_tests.jl:32-47),
 *   - the first-order (Phi, Pi) system of wave3d_curved_rhs! with its KO,
 *     affine boundary pass and energy (src/wave3d_curved.jl:205-287,
 *     src/boundaries3d.jl:56-126),
 *   - general connectivity (multi-patch meshes whose faces meet with a D4
 *     orientation, src/wave_strong_rhs.jl:41-73; tags 1..6 Dirichlet, 7
 *     Sommerfeld, 8 excision) for the second-order operator, and the
 *     curvilinear first-order path on such meshes (metric given:
 *     src/wave3d_curved.jl:218-256, src/boundaries3d.jl:131-280).
 * Every function is checked against known answers, closed forms or
 * invariants in tests/test_oracle_pins.py.  Parity unpinned: on elements that
 * are not affine, the curvilinear KO derivative takes the element's extent
 * along each reference axis as its h (HexSBPSAT's apply_D!, which fixes that
 * choice in the reference, is not in the reference tree); on affine elements
 * it reduces to the affine path's KO, which is pinned.
 * Layout: fields are (N,N,N,Ne) column-major (i fastest), elements x-fastest
 * e = ex + Mx*(ey + My*ez); faces 1..6 = (-x,+x,-y,+y,-z,+z).
 */
#ifndef WT_ORACLE_H
#define WT_ORACLE_H
#ifdef __cplusplus
extern "C" {
#endif

#define WTO_NMAX 16

typedef struct wto_desc {
    int N;                 /* GLL nodes per direction                      */
    int M[3];              /* elements per axis                            */
    double x0[3], x1[3];   /* domain extent per axis                       */
    int periodic[3];       /* 1 = periodic along that axis                 */
    double warp_A;         /* warp amplitude (0 = affine)                  */
    double tau;            /* SIPG penalty, unit-element units             */
    double eps_ko;         /* Kreiss-Oliger epsilon (0 disables)           */
    double bdry_values[6]; /* Dirichlet value per boundary tag 1..6        */
    int face_kind[6];      /* outer face f = 1..6: 0 Dirichlet (tag f), 7 Sommerfeld */
    double sommerfeld_inv_R; /* 1/R of the BGT-1 condition (0: plane-wave BGT-0) */
} wto_desc;

typedef struct wto_ctx wto_ctx;

/* GLL nodes xi on [0,1], weights w (sum 1), differentiation matrix
 * G[i*N+p] = l_p'(xi_i).  Returns 0 on success. */
int wto_gll(int N, double* xi, double* w, double* G);

/* bg: NULL for Minkowski, else 10 doubles per node, node-major
 * (alpha, beta1, beta2, beta3, g11, g12, g13, g22, g23, g33). */
wto_ctx* wto_create(const wto_desc* d, const double* bg);
/* A mesh given by coordinates (3,N,N,N,Ne) and 1-based (6,Ne) neighbour,
 * neighbour_face and bdry arrays (d's N, tau, bdry_values, sommerfeld_inv_R
 * apply; eps_ko must be 0).  NULL when faces do not meet. */
wto_ctx* wto_create_mesh(const wto_desc* d, long Ne, const double* xyz, const int* nbr,
                         const int* nface, const int* bdry);
void wto_destroy(wto_ctx* c);

long wto_npoints(const wto_ctx* c);
int wto_is_curved(const wto_ctx* c);
double wto_mu(const wto_ctx* c);
double wto_hmin(const wto_ctx* c);
void wto_coords(const wto_ctx* c, double* xyz);    /* (3,N,N,N,Ne) */
void wto_hphys(const wto_ctx* c, double* hw);      /* H3*|J| per node */
/* curved coefficients per node: K11 K12 K13 K22 K23 K33 c1 c2 c3 wgt kap1 kap2 kap3 */
int wto_coefs(const wto_ctx* c, double* out13);

/* Drop-in RHS (rhs_wave3d! semantics): udd = L u (flat; ud unused) or
 * wgt*(Ltilde u + A ud) on curved/background meshes.  No KO. */
void wto_lap(const wto_ctx* c, const double* u, const double* ud, double* udd);
/* Full first-order RHS: du = v, dv = lap(u,v) + KO(v). */
void wto_rhs(const wto_ctx* c, const double* u, const double* v, double* du, double* dv);
/* out = D_axis in (SAT-coupled first derivative, apply_D! semantics:
 * one-sided at outer faces). */
void wto_apply_D(const wto_ctx* c, int axis, const double* in, double* out);
/* The KO derivative: as wto_apply_D but outer faces see a zero ghost value. */
void wto_apply_Dko(const wto_ctx* c, int axis, const double* in, double* out);
/* Classical RK4, nsteps steps of dt, in place on (u, v).  energy_out (nullable)
 * receives nsteps+1 energies (before the first step and after each). */
void wto_rk4(const wto_ctx* c, double* u, double* v, double dt, int nsteps, double* energy_out);
double wto_energy(const wto_ctx* c, const double* u, const double* v);

/* First-order (Phi, Pi) ADM system, axis-aligned affine meshes.
 * coef11: 11 fields of n_pts values, SoA, in the order of make_coef3d
 * (alpha, sqrtg, b1, b2, b3, gu11, gu12, gu13, gu22, gu23, gu33). */
#define WTO_BC_DIRICHLET      1   /* src/boundaries1d.jl:59-62 */
#define WTO_BC_SOMMERFELD     2
#define WTO_BC_EXCISION       3
#define WTO_BC_FULL_DIRICHLET 4
typedef struct wto_bc3d {
    int kinds[6];          /* per face -x,+x,-y,+y,-z,+z (curvilinear: kinds[0] on every outer face) */
    int excision_tag;      /* curvilinear: outer faces with this tag get no SAT (0: none) */
    double sigma;
    const double* gPhi;    /* nullable targets (per node)                    */
    const double* gPi;
    const double* gD;      /* curvilinear Dirichlet: grad Phi target, 3 fields (nullable) */
} wto_bc3d;
void wto_fo_coef(long npts, const double* bg10, double* coef11);
void wto_fo_rhs(const wto_ctx* c, const double* Phi, const double* Pi, const double* coef11,
                double eps_ko, const wto_bc3d* bc, double* dPhi, double* dPi);
double wto_fo_energy(const wto_ctx* c, const double* Phi, const double* Pi, const double* coef11);

/* The first-order system on a general mesh (wto_create_mesh): the
 * curvilinear path of wave3d_curved_rhs! (metric given).  Metric terms
 * JAp[9 n + 3 a + m] = sign(J) J dxi^a/dx_m in the conservative-curl form and
 * |J| per node; split-form gradient (D3: 3 fields) and divergence (F3: 3
 * fields) with the centred-flux SAT; the KO derivative along each element's
 * reference axis; mu of that derivative; the RHS and the energy
 * (mass w_i w_j w_k |J|). */
int wto_curv_metric(const wto_ctx* c, double* JA9, double* Jp);
void wto_curv_grad(const wto_ctx* c, const double* Phi, double* D3);
void wto_curv_div(const wto_ctx* c, const double* F3, double* out);
void wto_curv_D(const wto_ctx* c, int axis, const double* in, double* out);
double wto_curv_mu(const wto_ctx* c);
void wto_fo_rhs_curv(const wto_ctx* c, const double* Phi, const double* Pi, const double* coef11,
                     double eps_ko, const wto_bc3d* bc, double* dPhi, double* dPi);
double wto_fo_energy_curv(const wto_ctx* c, const double* Phi, const double* Pi, const double* coef11);

/* Plant a known mistake in the ADM coefficients of contexts created after the
 * call (0 restores the correct formulas).  Only for pin tests, which must show
 * that they fail on each of these:
 *   WTO_PLANT_DROP_BB    K^ij = sqrt(g) alpha g^ij (the -b^i b^j / alpha term dropped),
 *   WTO_PLANT_WGT_SQRTG  wgt = alpha sqrt(g) / |J| instead of alpha / (sqrt(g) |J|),
 *   WTO_PLANT_FLIP_C     c^i = -sqrt(g) b^i / alpha. */
#define WTO_PLANT_DROP_BB   1
#define WTO_PLANT_WGT_SQRTG 2
#define WTO_PLANT_FLIP_C    3
void wto_plant_mistake(int code);

/* Dense 1D operators on a line of M elements (for spectrum / symmetry pins).
 * kind 0: SIPG Laplacian with homogeneous Dirichlet (or periodic) ends,
 * kind 1: SAT-coupled first derivative D (one-sided ends), kind 2: the KO
 * derivative (zero-ghost ends).  A is (M*N)x(M*N) row-major. */
int wto_assemble_1d(int N, int M, int periodic, double tau, double h, int kind, double* A);

#ifdef __cplusplus
}
#endif
#endif
