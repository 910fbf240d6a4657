/*
 * wt2.h -- C ABI of libwt2.so: MI355X (gfx950) kernels for the 3D SBP-SAT
 * scalar-wave right-hand side and the classical RK4 step.
 *
 * Each entry point names the reference (eschnett/WaveToySecondOrder.jl) call it
 * replaces.  All functions are extern "C", take plain pointers and sizes,
 * return 0 on success or a negative WT2_ERR_* code (never abort), and leave a
 * message for wt2_last_error() (thread-local).
 *
 * Field layout (every double* / float* field argument): Julia column-major
 * (N,N,N,Ne), i.e. C index i + N*(j + N*(k + N*e)), elements x-fastest
 * e = ex + Mx*(ey + My*ez).  Faces 1..6 are (-x,+x,-y,+y,-z,+z); an outer face
 * f carries boundary tag f and Dirichlet value bdry_values[f-1], unless
 * face_kind[f-1] marks it a Sommerfeld face (tag 7, src/wave.jl:23).
 *
 * Pointers are device pointers (16-byte aligned) unless WT2_HOST_PTRS is passed
 * in flags, in which case they are host arrays copied in and out around the call.
 * Contexts are not thread-safe; separate contexts are independent.  Calls are
 * synchronous with respect to the host (they end with a stream sync).
 */
#ifndef WT2_H
#define WT2_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- return codes ---- */
#define WT2_OK                   0
#define WT2_ERR_ARG             -1   /* null pointer / bad size / bad flag      */
#define WT2_ERR_SHAPE           -2   /* N or M out of the supported range       */
#define WT2_ERR_TAU             -3   /* SIPG penalty below coercivity           */
#define WT2_ERR_UNSUPPORTED     -4   /* feature not built (e.g. M > 64 per axis) */
#define WT2_ERR_HIP             -5   /* HIP runtime error                       */
#define WT2_ERR_OOM             -6   /* device allocation failed                */
#define WT2_ERR_NONFINITE       -7   /* non-finite value detected in the state  */
#define WT2_ERR_STATE           -8   /* call not valid in this context state    */

/* ---- precision ---- */
#define WT2_FP64   0   /* everything in FP64                                      */
#define WT2_MIXED  1   /* FP32 RHS bulk; FP64 state, RK sums, geometry and energy */

/* ---- flags ---- */
#define WT2_HOST_PTRS         1   /* field pointers are host arrays               */
#define WT2_ALLOW_NONCOERCIVE 2   /* ctx_create: accept tau below coercivity      */
#define WT2_NO_GRAPH          4   /* rk4_steps: launch kernels one by one         */

/* Mesh / operator description (HexMeshes make_uniform_hex /
 * make_warped_uniform_hex, src/evolve.jl:915, test/dirichlet_warped_tests.jl:23;
 * Params3d tau/bdry_values, src/wave.jl:67-95; eps_KO,
 * src/wave3d_curved.jl:205-213). */
typedef struct wt2_mesh_desc {
    int32_t N;               /* GLL nodes per direction, 3..10               */
    int32_t M[3];            /* elements per axis, 1..64                     */
    double  x0[3], x1[3];    /* domain per axis                              */
    int32_t periodic[3];     /* 1 = periodic along that axis                 */
    double  warp_A;          /* x = x0 + L (xi + A sin sin sin); 0 = affine  */
    double  tau;             /* SIPG penalty in unit-element units           */
    double  eps_ko;          /* Kreiss-Oliger epsilon; 0 disables KO         */
    double  bdry_values[6];  /* Dirichlet value per tag 1..6 (RK steps)      */
    int32_t precision;       /* WT2_FP64 or WT2_MIXED                        */
    int32_t flags;           /* WT2_ALLOW_NONCOERCIVE                        */
    int32_t face_kind[6];    /* per face: 0 = Dirichlet (SIPG, tag f), 7 =
                                Sommerfeld (free for SIPG plus the
                                Bayliss-Turkel drag u_t + d_n u + u/R -> 0,
                                src/wave.jl:444-518); bounded axes, flat
                                or warped meshes without a background       */
    double  sommerfeld_inv_R;/* 1/R of that drag; 0 = plane-wave (BGT-0)     */
} wt2_mesh_desc;

typedef struct wt2_ctx wt2_ctx;

typedef struct wt2_info {
    int64_t npoints;         /* N^3 * Ne                                     */
    int32_t Ne;
    int32_t curved;          /* 1 if warp or background set                  */
    double  mu;              /* spectral radius of apply_D! along axis 1     */
    double  h_min;           /* smallest node spacing (src/evolve.jl:60-69)  */
    double  tau;             /* penalty in use                               */
    double  max_sym_eig;     /* max eigenvalue of sym(H L) of the 1D operator */
    int64_t device_bytes;    /* device memory held by the context            */
    int32_t launches_per_step; /* kernel launches of one RK4 step (8 folded, 12 flat, curved more) */
    int32_t reserved;
} wt2_info;

/* GLL nodes xi on [0,1], weights w (sum 1) and differentiation matrix
 * G[i*N+p] = l_p'(xi_i).  Replaces HexSBPSAT make_element/make_operators
 * (src/wave.jl:603; test/kernels3d_tests.jl:69-70).  Host only. */
int wt2_gll(int32_t N, double* xi, double* w, double* G);

/* Create a context: mesh, operators, geometry, workspace and a HIP stream on
 * the current device.  Replaces make_uniform_hex + make_element +
 * make_operators + make_geometry + make_workspace + to_device
 * (src/evolve.jl:910-936) and make_wave3d_workspace's mu
 * (src/wave3d_curved.jl:158-178).  Fails with WT2_ERR_TAU when the 1D SIPG
 * operator is indefinite at desc->tau unless WT2_ALLOW_NONCOERCIVE is set. */
int wt2_ctx_create(const wt2_mesh_desc* desc, wt2_ctx** out);
/* Create a context on a mesh given by node coordinates and face connectivity
 * (multi-patch meshes such as the cubed cube or inflated cube, whose patches
 * meet with orientation != 0).  Replaces make_geometry + make_workspace +
 * to_device of such a mesh (src/evolve.jl:916-936) and the connectivity reads
 * of rhs_wave3d! (src/wave.jl:471-473; src/wave_strong_rhs.jl:41-73,
 * 316-324).  Arrays are the reference's MeshGeometry/MeshConnectivity fields,
 * column-major: coords (3,N,N,N,Ne); neighbour, neighbour_face, bdry (6,Ne)
 * with 1-based element and face numbers (neighbour 0 = outer face).
 * orientation may be NULL: each shared face's D4 map is then found from the
 * face nodes' coordinates (they must coincide to 1e-8 of the element size).
 * Otherwise orientation[f,e] = s + 2 rp + 4 rq maps this face's node (p,q)
 * to the neighbour face's (p',q') = (q,p) if s else (p,q), then p' -> N-1-p'
 * if rp, q' -> N-1-q' if rq.  Outer faces carry bdry tag 1..6 (Dirichlet
 * value desc->bdry_values[tag-1]), 7 (Sommerfeld) or 8 (excision: no SAT,
 * src/wave_strong_rhs.jl:188-213).  desc supplies N, tau,
 * precision, flags, bdry_values and sommerfeld_inv_R; eps_ko must be 0.  The
 * geometry J = dx/dxi is G applied to coords.  Ne must factor as
 * M0*M1*M2 with each factor within wt2_ctx_create's per-axis limits. */
int wt2_ctx_create_mesh(const wt2_mesh_desc* desc, int64_t Ne, const double* coords,
                        const int32_t* neighbour, const int32_t* neighbour_face,
                        const int32_t* orientation, const int32_t* bdry, wt2_ctx** out);
/* Device bytes that wt2_ctx_create (and, with with_background != 0 on a flat
 * mesh, wt2_ctx_set_background) will allocate for desc, without allocating:
 * the figure wt2_info.device_bytes reports afterwards.  Caller-owned state and
 * the staging buffers of host-pointer calls are not included.  No counterpart
 * in the reference (it allocates on the host); used for admission by memory
 * footprint. */
int wt2_ctx_footprint(const wt2_mesh_desc* desc, int32_t with_background, int64_t* bytes);
void wt2_ctx_destroy(wt2_ctx* ctx);
int wt2_ctx_info(const wt2_ctx* ctx, wt2_info* info);

/* Static ADM background per node (host FP64, 10 values per node, node-major:
 * alpha, beta1..3, gamma11, gamma12, gamma13, gamma22, gamma23, gamma33).
 * Replaces sample_background3d! + make_coef3d (src/wave3d_curved.jl:106-124),
 * sampled once because the background is static.  Switches the context to the
 * curved operator. */
int wt2_ctx_set_background(wt2_ctx* ctx, const double* bg10);

/* Node coordinates (3,N,N,N,Ne), host array (geom.coords). */
int wt2_coords(const wt2_ctx* ctx, double* xyz);

/* rhs_wave3d!(udd, u, ud, bdry_values; geom, ops, work, tau)
 * (src/wave.jl:583-611): udd = L_SIPG u with Dirichlet values bdry_values[tag-1]
 * on outer faces, and the Sommerfeld faces' drag, which reads ud at the face
 * nodes (ud = NULL: taken as zero).  On a curved/background context udd also carries the u_t
 * advection terms of the ADM operator.  No KO (the reference RHS has none). */
int wt2_rhs3d(wt2_ctx* ctx, double* udd, const double* u, const double* ud,
              const double bdry_values[6], int32_t flags);

/* Full first-order right-hand side with KO, the (u, u_t) analogue of
 * wave3d_curved_rhs! (src/wave3d_curved.jl:205-258):
 *   du = ud,  dud = rhs_wave3d!(u, ud) + eps mu^-5 sum_d D_d^6 ud.
 * D_d is apply_D!'s SAT-coupled derivative with the outer faces closed against
 * a zero ghost value (H D skew, so the term only removes energy); KO acts on
 * u_t only because on u it would not commute with the SIPG operator. */
int wt2_rhs_full(wt2_ctx* ctx, double* du, double* dud, const double* u,
                 const double* ud, int32_t flags);

/* nsteps classical RK4 steps of size dt on (u, ud) in place (the integrator
 * step! loop of src/evolve.jl:1028-1031 with RK4(), stage order of
 * test/wave3d_tests.jl:32-44).  Each stage is the RHS fused with the stage
 * update.  The system is linear and autonomous (static background, constant
 * boundary data), so the RK4 polynomial is evaluated in Horner form --
 * S1 = y + dt/4 f(y), S2 = y + dt/3 f(S1), S3 = y + dt/2 f(S2),
 * y <- y + dt f(S3) -- equal to the classical stage sums up to rounding, with
 * no accumulator register.  WT2_MIXED: the stage inputs and partial sums are
 * FP32; stages 1-3 read an FP32 mirror of (u, ud) made at the start of the
 * call (they round their result to FP32, so Y is needed only to FP32
 * precision), stage 4 reads and writes (u, ud) in FP64 and renews the mirror.
 * If energy_out is non-null and energy_every > 0 it receives
 * 1 + nsteps/energy_every energies (host array): before the first step and
 * after every energy_every-th step. */
int wt2_rk4_steps(wt2_ctx* ctx, double* u, double* ud, double dt, int32_t nsteps,
                  int32_t energy_every, double* energy_out, int32_t flags);

/* No counterpart in the reference: wt2_rk4_steps for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  The same steps, kernels and results as
 * wt2_rk4_steps on device pointers, but it allocates nothing: the context's
 * workspace (allocated by wt2_ctx_create) holds the stage registers, and the
 * energy history and its reduction partials live in a fixed device array of
 * the library, sized for WT2_NOALLOC_MAX_ENERGIES samples (about 0.5 MiB,
 * held by every GPU the library runs on until the process ends); more samples
 * are refused with WT2_ERR_ARG.  Kernels are launched one by one (no graph is
 * instantiated).  u, ud are device pointers (16-byte aligned); energy_out is a
 * host array as in wt2_rk4_steps.  A curved context must have its
 * coefficients computed first (any wt2_ctx_info call does it), else
 * WT2_ERR_STATE.  All calls on one GPU share the one array: calls that overlap
 * in time on one GPU (from several host threads) race. */
#define WT2_NOALLOC_MAX_ENERGIES 65536
int wt2_rk4_steps_noalloc(wt2_ctx* ctx, double* u, double* ud, double dt, int32_t nsteps,
                          int32_t energy_every, double* energy_out);

/* Discrete energy E = 1/2<ud,ud>_W + 1/2<u,-L u>_H (zero boundary data)
 * (test/kernels3d_tests.jl:32-47; src/wave3d_curved.jl:268-287), FP64,
 * deterministic reduction order. */
int wt2_energy(wt2_ctx* ctx, const double* u, const double* ud, double* E, int32_t flags);

/* ---- the first-order (Phi, Pi) system on affine meshes ----
 * coef11: 11 fields of npoints values, one after the other, in the order of
 * make_coef3d / sample_background3d! (src/wave3d_curved.jl:75-124):
 * alpha, sqrt(det gamma), beta1, beta2, beta3, gamma^11, gamma^12, gamma^13,
 * gamma^22, gamma^23, gamma^33.  The context must be flat (no warp, no
 * background); its mu is the reference's (apply_D! power iteration). */
#define WT2_BC_DIRICHLET      1   /* src/boundaries1d.jl:59-62 */
#define WT2_BC_SOMMERFELD     2
#define WT2_BC_EXCISION       3
#define WT2_BC_FULL_DIRICHLET 4

/* make_bc3d (src/boundaries3d.jl:44-54).  Affine meshes: one kind per side.
 * General meshes (wt2_ctx_create_mesh): kinds[0] on every outer face except
 * those whose bdry tag equals excision_tag, which get no SAT
 * (_apply_bc3d_curv!, src/boundaries3d.jl:131-165). */
typedef struct wt2_bc3d {
    int32_t kinds[6];        /* WT2_BC_* per face -x,+x,-y,+y,-z,+z          */
    int32_t excision_tag;    /* general meshes: outer faces with this tag get no SAT (0: none) */
    double  sigma;           /* penalty scale (reference default 1)           */
    const double* gPhi;      /* field-radiation / full-state targets, per node */
    const double* gPi;       /* (nullable: zero)                               */
    const double* gD;        /* general meshes, WT2_BC_DIRICHLET: the grad Phi
                                target, 3 fields of npoints (nullable: zero)  */
} wt2_bc3d;

/* sample_background3d! (src/wave3d_curved.jl:75-115) on the device: coef11
 * from 10 background values per node (node-major, as wt2_ctx_set_background). */
int wt2_fo_coef(wt2_ctx* ctx, const double* bg10, double* coef11, int32_t flags);

/* wave3d_curved_rhs!(Phidot, Pidot, Phi, Pi, coef; geom, ops, ws, eps_KO, bc3d)
 * on the affine path (metric === nothing, src/wave3d_curved.jl:205-258):
 * DPhi_d = apply_D!(Phi, d); F^i = b^i Pi + alpha sqrtg g^{ij} DPhi_j;
 * Pidot = sum_i apply_D!(F^i, i); Phidot = b . DPhi + (alpha/sqrtg) Pi; then
 * eps mu^-5 D_d^6 on both fields; then the affine boundary pass
 * (src/boundaries3d.jl:76-126) when bc3d is non-null.  FP64.
 * On a context from wt2_ctx_create_mesh, the curvilinear path (metric given,
 * :218-256): the gradient and divergence in the split form with metric terms
 * in the conservative-curl form (free stream to rounding) and the centred
 * SAT across element faces, KO along each element's reference axes (the
 * element's extent in place of h; mu from 30 power iterations of that
 * derivative), and the physical-normal boundary pass _apply_bc3d_curv!
 * (src/boundaries3d.jl:131-280). */
int wt2_fo_rhs(wt2_ctx* ctx, double* dPhi, double* dPi, const double* Phi, const double* Pi,
               const double* coef11, double eps_ko, const wt2_bc3d* bc3d, int32_t flags);

/* One classical RK4 stage of the first-order system in place
 * (OrdinaryDiffEq RK4() stages of src/evolve.jl:1309-1313, stage order of
 * test/wave3d_tests.jl:32-44): the stage's right-hand side as wt2_fo_rhs, on
 * the state (stage 1) or the context's stage input, followed by one update
 * pass over both fields (stages 1-3 write the next stage input and the
 * accumulator, stage 4 the new state).  Phi and Pi are device pointers
 * (16-byte aligned).  The caller may change the boundary data of bc3d between
 * stages (time-dependent targets). */
int wt2_fo_rk4_stage(wt2_ctx* ctx, int32_t stage, double* Phi, double* Pi, const double* coef11, double dt,
                     double eps_ko, const wt2_bc3d* bc3d);

/* wave3d_energy (src/wave3d_curved.jl:268-287):
 * E = sum 1/2 [Pi^2/sqrtg + sqrtg g^{ij} DPhi_i DPhi_j] Hphys, FP64,
 * deterministic reduction order. */
int wt2_fo_energy(wt2_ctx* ctx, const double* Phi, const double* Pi, const double* coef11, double* E,
                  int32_t flags);

/* Use an external HIP stream (hipStream_t passed as void*); NULL restores the
 * context's own stream. */
int wt2_set_stream(wt2_ctx* ctx, void* stream);

/* ---- batches: many independent evolutions advanced together ----
 * No counterpart in the reference (it evolves one mesh at a time); this is the
 * batched form of wt2_rk4_steps.  Each instance is a context plus its
 * device-resident state (u, ud: device pointers owned by the caller) with its
 * own dt and number of steps; every RK4 stage runs as one launch per group of
 * instances sharing N, precision and pass kind, over a packed work list, and
 * the step is replayed from a hipGraph while the set of unfinished instances
 * is unchanged.  Adding a context moves its kernels onto the batch's stream
 * until it is removed; a context belongs to at most one batch. */
typedef struct wt2_batch wt2_batch;
int wt2_batch_create(wt2_batch** out);
void wt2_batch_destroy(wt2_batch* b);
/* energy_every > 0 records the FP64 energy at the start and every energy_every
 * steps of this instance. */
int wt2_batch_add(wt2_batch* b, wt2_ctx* ctx, double* u, double* ud, double dt, int64_t nsteps,
                  int32_t energy_every, int32_t* slot);
int wt2_batch_remove(wt2_batch* b, int32_t slot);
/* Advance every unfinished instance by up to max_steps steps (each stops at its
 * own nsteps); *finished (nullable) counts the instances that finished. */
int wt2_batch_step(wt2_batch* b, int32_t max_steps, int32_t* finished);
/* As wt2_batch_step, but returns once the steps are queued on the batch's
 * stream (it still waits where the set of unfinished instances changes and
 * the work lists are rebuilt).  Step counts advance at queue time; the state
 * of a finished instance is final after wt2_batch_sync, wt2_batch_remove or a
 * wt2_batch_slot call that copies energies. */
int wt2_batch_step_async(wt2_batch* b, int32_t max_steps, int32_t* finished);
int wt2_batch_sync(wt2_batch* b);
/* Progress of one slot: steps done, and the energy samples recorded so far
 * (kept on the device while stepping; copying them waits for the batch's
 * stream).  Returns WT2_ERR_NONFINITE when a copied sample is not finite. */
int wt2_batch_slot(const wt2_batch* b, int32_t slot, int64_t* steps_done, double* energies,
                   int32_t cap, int32_t* n_energies);

/* Kernel launches of one RK4 step of the batch as last planned by a step call
 * (0 before the first step and after the set of unfinished instances changed),
 * and the number of unfinished instances.  No counterpart in the reference. */
int wt2_batch_info(const wt2_batch* b, int32_t* launches_per_step, int32_t* active);

/* Autotuner (no counterpart in the reference): times every valid pass-kernel
 * kind and tile size per axis on this context with state (u, ud) (device
 * pointers, only read) and keeps the fastest (a flat context with the folded
 * RK4 schedule times stages 1-3 of the step with each candidate); the choice is
 * cached per (N, M, precision, curved, KO, periodicity, schedule) and applied to
 * later contexts of the same shape.  flags & 1 forces a re-measurement. */
int wt2_ctx_autotune(wt2_ctx* ctx, double* u, double* ud, double dt, int32_t flags);
/* Current tile choice: element planes per tile and LDS-free flag per axis. */
int wt2_ctx_tiles(const wt2_ctx* ctx, int32_t* nb3, int32_t* kind3);
/* Set the pass kernel kind and planes per tile of each axis, e.g. a table
 * the autotuner produced earlier.  kind3[a]: 0 = LDS-staged tile, 1 = LDS-free
 * with the line's elements in the lanes of one wavefront (shuffles), 2 =
 * LDS-free with adjacent lines in adjacent lanes and the neighbour exchange
 * through LDS.  nb3[a] must divide N and the tile must fit one workgroup.
 * wt2_ctx_tiles reports the same two arrays.  No counterpart in the
 * reference. */
int wt2_ctx_set_tiles(wt2_ctx* ctx, const int32_t* nb3, const int32_t* kind3);

/* Diagnostics (no counterpart in the reference; used by bench.py to time one
 * kernel in isolation with events on the context stream): enqueue, without
 * synchronising, the single pencil pass of RK4 stage `stage` (1..4) along
 * `axis` (0..2) on state (u, ud) (device pointers) -- or, for a curved
 * context and axis == -1, the curved element operator of that stage; with
 * axis == 3, the drop-in right-hand side of wt2_rhs3d (the context's own
 * bdry_values) into an internal buffer.  For axes -1..2 the context registers
 * must have been initialised by a wt2_rk4_steps call with the same state
 * pointers (in mixed contexts, together with the FP32 mirror of the state). */
int wt2_launch_pass(wt2_ctx* ctx, int32_t axis, int32_t stage, double* u, double* ud, double dt);

const char* wt2_last_error(void);
const char* wt2_version(void);

#ifdef __cplusplus
}
#endif
#endif /* WT2_H */
