"""Pins of the CPU oracle against what the reference's tests and the mathematics fix.

The reference ships no golden vectors (its tests are properties, closed forms
and tolerances), and its arithmetic (HexSBPSAT/HexMeshes) is not in the
reference tree, so every oracle function is tied down here by a check a
plausible mistake (a dropped term, a wrong sign or index) would fail.
"""
import math

import numpy as np
import pytest
from numpy.polynomial import legendre as npleg

import wt_oracle as wo


# ------------------------------------------------------------------ element

@pytest.mark.parametrize("N", [3, 4, 5, 6, 8, 10])
def test_gll_matches_numpy_and_sbp(N):
    """GLL nodes = endpoints + roots of P'_{N-1} (numpy.polynomial.legendre),
    weights 2/(n(n+1)P_n^2) sum to 1, G exact on degree <= N-1, HG+(HG)^T = B
    (README.md:13-15), G 1 = 0."""
    xi, w, G = wo.gll(N)
    c = np.zeros(N); c[-1] = 1.0
    r = np.sort(npleg.legroots(npleg.legder(c)))
    x = np.concatenate([[-1.0], r, [1.0]])
    assert np.max(np.abs((x + 1) / 2 - xi)) < 1e-14
    pn = npleg.legval(x, c)
    w_ref = 2.0 / ((N - 1) * N * pn ** 2) / 2.0
    assert np.max(np.abs(w - w_ref)) < 1e-14
    assert abs(w.sum() - 1.0) < 1e-14
    H = np.diag(w)
    B = np.zeros((N, N)); B[0, 0] = -1; B[-1, -1] = 1
    assert np.max(np.abs(H @ G + (H @ G).T - B)) < 1e-13
    for p in range(1, N):
        assert np.max(np.abs(G @ xi ** p - p * xi ** (p - 1))) < 1e-12
    assert np.max(np.abs(G @ np.ones(N))) < 1e-13
    # quadrature exact to degree 2N-3
    for p in range(0, 2 * N - 2):
        assert abs(w @ xi ** p - 1.0 / (p + 1)) < 1e-14


# --------------------------------------------------------------- 1D SIPG

@pytest.mark.parametrize("N", [4, 6, 8, 10])
def test_sipg_1d_symmetric_nsd_at_tau_eff(N):
    """H L symmetric; NSD at tau = 1.5 (N-1)^2 (src/evolve.jl:964-968); smallest
    Dirichlet eigenvalue -> -pi^2 on [0,1] (test/dirichlet_cube_tests.jl:110-130)."""
    M = 4
    h = 1.0 / M
    _, w, _ = wo.gll(N)
    A = wo.assemble_1d(N, M, 0, 1.5 * (N - 1) ** 2, h, 0)
    S = np.diag(np.tile(w, M) * h) @ A
    assert np.max(np.abs(S - S.T)) <= 1e-12 * np.max(np.abs(S))
    ev = np.linalg.eigvalsh(0.5 * (S + S.T))
    assert ev.max() < 0
    lam = np.linalg.eigvals(A)
    assert np.max(np.abs(lam.imag)) < 1e-8 * np.max(np.abs(lam))
    assert abs(np.max(lam.real) + math.pi ** 2) < 1e-2 * math.pi ** 2


def test_sipg_literal_tau20_indefinite_for_n8():
    """The configs' literal tau = 20 is below SIPG coercivity at N = 8 (a positive
    eigenvalue appears), while N = 4 is fine (README.md:77-81 'tau >~ 17')."""
    _, w8, _ = wo.gll(8)
    A8 = wo.assemble_1d(8, 4, 0, 20.0, 0.25, 0)
    assert np.max(np.linalg.eigvals(A8).real) > 1.0
    A4 = wo.assemble_1d(4, 4, 0, 20.0, 0.25, 0)
    assert np.max(np.linalg.eigvals(A4).real) < 0


def test_sipg_periodic_annihilates_constants_and_kinks_removed():
    """Periodic: L 1 = 0 (test/wave_lap_strong_mesh_tests.jl:45-53); the kink mode
    (continuous piecewise linear) is NOT in the kernel once the gradient jump is
    penalised (README.md:35-55)."""
    N, M = 5, 4
    A = wo.assemble_1d(N, M, 1, 1.5 * (N - 1) ** 2, 1.0 / M, 0)
    assert np.max(np.abs(A @ np.ones(M * N))) < 1e-10
    xi, _, _ = wo.gll(N)
    x = np.concatenate([(e + xi) / M for e in range(M)])
    kink = np.abs(x - 0.5)   # continuous, kinked at 0.5 (an element face for M=4)
    assert np.max(np.abs(wo.assemble_1d(N, M, 0, 1.5 * (N - 1) ** 2, 1.0 / M, 0) @ kink)) > 1.0


def test_sipg_second_derivative_accuracy():
    """On a smooth function vanishing at the ends, L u -> u'' at spectral rate."""
    errs = []
    for M in (2, 4, 8):
        N = 6
        xi, _, _ = wo.gll(N)
        x = np.concatenate([(e + xi) / M for e in range(M)])
        A = wo.assemble_1d(N, M, 0, 1.5 * (N - 1) ** 2, 1.0 / M, 0)
        u = np.sin(math.pi * x)
        errs.append(np.max(np.abs(A @ u + math.pi ** 2 * u)))
    assert errs[1] < errs[0] / 8 and errs[2] < errs[1] / 8


# ------------------------------------------------------------------ D / KO

@pytest.mark.parametrize("N", [4, 8])
def test_apply_D_skew_periodic_and_exact(N):
    """H D + (H D)^T = 0 on a periodic line (METHODS.md:259-262); D exact on
    global polynomials of degree <= N-1 away from SAT faces (continuous data:
    SAT jump terms vanish)."""
    M = 4
    h = 1.0 / M
    _, w, _ = wo.gll(N)
    D = wo.assemble_1d(N, M, 1, 0.0, h, 1)
    HD = np.diag(np.tile(w, M) * h) @ D
    assert np.max(np.abs(HD + HD.T)) < 1e-12 * np.max(np.abs(HD))
    Dn = wo.assemble_1d(N, M, 0, 0.0, h, 1)
    xi, _, _ = wo.gll(N)
    x = np.concatenate([(e + xi) / M for e in range(M)])
    for p in range(1, N):
        assert np.max(np.abs(Dn @ x ** p - p * x ** (p - 1))) < 1e-9


def test_mu_power_iteration_matches_dense_spectral_radius():
    """mu (30 power iterations on D^2 from the alternating vector,
    src/wave3d_curved.jl:158-178) is within a few percent of rho(D)
    (METHODS.md: 'accurate to a few %')."""
    for N, M in ((4, 4), (8, 4), (6, 3)):
        o = wo.Oracle(N, M, eps_ko=0.1)
        D = wo.assemble_1d(N, M, 0, 0.0, 1.0 / M, 1)
        rho = np.max(np.abs(np.linalg.eigvals(D)))
        assert abs(o.mu / rho - 1) < 0.05


def test_ko_derivative_skew_on_bounded_line():
    """The KO derivative (apply_D! with outer faces closed against a zero ghost)
    is exactly H-skew on a bounded line, so H D^6 <= 0 there; apply_D! itself
    (one-sided ends) satisfies H D + (H D)^T = B instead (SBP)."""
    for N in (4, 8):
        M = 3
        _, w, _ = wo.gll(N)
        H = np.diag(np.tile(w, M) / M)
        Dk = wo.assemble_1d(N, M, 0, 0.0, 1.0 / M, 2)
        HD = H @ Dk
        assert np.max(np.abs(HD + HD.T)) < 1e-12 * np.max(np.abs(HD))
        K = H @ np.linalg.matrix_power(Dk, 6)
        assert np.max(np.linalg.eigvalsh(0.5 * (K + K.T))) < 1e-9 * np.max(np.abs(K))
        D1 = wo.assemble_1d(N, M, 0, 0.0, 1.0 / M, 1)
        B = np.zeros((M * N, M * N)); B[0, 0] = -1; B[-1, -1] = 1
        assert np.max(np.abs(H @ D1 + (H @ D1).T - B)) < 1e-12 * np.max(np.abs(H @ D1))


def test_ko_is_dissipative_in_H():
    """D^6 is NSD in the H norm on a periodic line (D skew => D^6 = -(D^3)^T H D^3)."""
    N, M = 6, 4
    _, w, _ = wo.gll(N)
    D = wo.assemble_1d(N, M, 1, 0.0, 1.0 / M, 1)
    H = np.diag(np.tile(w, M) / M)
    K = H @ np.linalg.matrix_power(D, 6)
    assert np.max(np.linalg.eigvalsh(0.5 * (K + K.T))) < 1e-8 * np.max(np.abs(K))


# ------------------------------------------------------------------- 3D flat

def _dense(o, f, n):
    A = np.zeros((n, n))
    x = np.zeros(n)
    for j in range(n):
        x[:] = 0; x[j] = 1
        A[:, j] = f(x)
    return A


def test_flat_3d_is_kronecker_sum():
    """The flat 3D operator equals L1 (x) I (x) I + I (x) L2 (x) I + I (x) I (x) L3 once
    the element-major layout is mapped to global lexicographic node order."""
    N, M = 3, (2, 3, 2)
    o = wo.Oracle(N, M, x1=(1.0, 1.5, 0.8))
    A = _dense(o, o.lap, o.n)
    L = [wo.assemble_1d(N, M[a], 0, 1.5 * (N - 1) ** 2, [1.0, 1.5, 0.8][a] / M[a], 0) for a in range(3)]
    # global index g = gx + nx*(gy + ny*gz), gx = ex*N + i
    nx, ny, nz = M[0] * N, M[1] * N, M[2] * N
    perm = np.zeros(o.n, dtype=int)
    q = 0
    for ez in range(M[2]):
        for ey in range(M[1]):
            for ex in range(M[0]):
                for k in range(N):
                    for j in range(N):
                        for i in range(N):
                            gx, gy, gz = ex * N + i, ey * N + j, ez * N + k
                            perm[q] = gx + nx * (gy + ny * gz)
                            q += 1
    K = (np.kron(np.eye(nz), np.kron(np.eye(ny), L[0])) + np.kron(np.eye(nz), np.kron(L[1], np.eye(nx)))
         + np.kron(L[2], np.eye(nx * ny)))
    Kp = K[np.ix_(perm, perm)]
    assert np.max(np.abs(A - Kp)) < 1e-10 * np.max(np.abs(A))


def _sinmode(o):
    x = o.coords()
    return np.sin(math.pi * x[:, 0]) * np.sin(math.pi * x[:, 1]) * np.sin(math.pi * x[:, 2])


def test_dirichlet_rayleigh_quotient_converges_to_3pi2():
    """test/dirichlet_cube_tests.jl:58-108: Rayleigh quotient -> 3 pi^2, < 1e-3 at
    M = 4 and >= 4x per halving (N = 4)."""
    errs = []
    for M in (2, 4, 8):
        o = wo.Oracle(4, M)
        u = _sinmode(o)
        H = o.hphys()
        lam = -np.dot(u, H * o.lap(u)) / np.dot(u, H * u)
        errs.append(abs(lam - 3 * math.pi ** 2) / (3 * math.pi ** 2))
    assert errs[1] < 1e-3
    assert errs[1] < errs[0] / 4 and errs[2] < errs[1] / 4


def test_dirichlet_spectrum_real_nonpositive():
    """test/dirichlet_cube_tests.jl:110-130 (N = 3, M = 2)."""
    o = wo.Oracle(3, 2, tau=1.5 * 4)
    A = _dense(o, o.lap, o.n)
    lam = np.linalg.eigvals(A)
    assert np.max(lam.real) < 1e-10 * np.max(np.abs(lam))
    assert np.max(np.abs(lam.imag)) < 1e-8 * np.max(np.abs(lam))
    assert np.min(lam.real) < -1.0


def test_dirichlet_values_enter_by_tag():
    """bdry_values[tag-1] is the Dirichlet value of outer face tag f = 1..6 =
    (-x,+x,-y,+y,-z,+z) (src/wave.jl:567-571): with u = 0 and data 1 on tag f
    only, L u is non-zero exactly on the elements touching face f and the
    Nitsche penalty pulls u towards the data (positive at the face nodes); a
    constant u with matching data on every face gives L u = 0."""
    N, M = 4, 3
    for f in range(6):
        bv = [0.0] * 6
        bv[f] = 1.0
        o = wo.Oracle(N, M, bdry_values=tuple(bv))
        x = o.coords()
        y = o.lap(np.zeros(o.n))
        a, hi = f // 2, f % 2
        e = np.arange(o.n) // N ** 3
        ec = (e % M, (e // M) % M, e // (M * M))[a]
        near = ec == (M - 1 if hi else 0)
        on_face = np.abs(x[:, a] - (1.0 if hi else 0.0)) < 1e-12
        assert np.max(np.abs(y[~near])) == 0.0
        assert np.all(y[on_face] > 0)
    c = 0.7
    o = wo.Oracle(N, M, bdry_values=(c,) * 6)
    assert np.max(np.abs(o.lap(np.full(o.n, c)))) < 1e-9


def test_full_rhs_spectrum_stable_with_ko():
    """test/wave3d_tests.jl:72-89: max Re(lambda) <= 1e-5 |lambda|_max with eps_KO = 0.1."""
    o = wo.Oracle(3, 2, eps_ko=0.1)
    n = o.n

    def f(x):
        du, dv = o.rhs(x[:n].copy(), x[n:].copy())
        return np.concatenate([du, dv])
    A = _dense(o, f, 2 * n)
    lam = np.linalg.eigvals(A)
    assert np.max(lam.real) <= 1e-5 * np.max(np.abs(lam))


def test_rk4_step_is_fourth_order_taylor_polynomial():
    """For y' = A y one classical RK4 step is sum_{k<=4} (dt A)^k/k! y0 exactly."""
    o = wo.Oracle(3, 2, eps_ko=0.1)
    n = o.n

    def f(x):
        du, dv = o.rhs(x[:n].copy(), x[n:].copy())
        return np.concatenate([du, dv])
    A = _dense(o, f, 2 * n)
    rng = np.random.default_rng(7)
    y0 = rng.standard_normal(2 * n)
    dt = 0.2 * o.hmin
    P = np.eye(2 * n); T = np.eye(2 * n)
    for k in range(1, 5):
        T = T @ (dt * A) / k
        P = P + T
    y1 = P @ y0
    u1, v1 = o.rk4(y0[:n], y0[n:], dt, 1)
    assert np.max(np.abs(np.concatenate([u1, v1]) - y1)) < 1e-12 * np.max(np.abs(y1))


def test_eigenmode_evolution_converges_and_conserves_energy():
    """Exact standing wave cos(sqrt(3) pi t) sin sin sin (the plain-method CPU
    configuration, at reduced size): error ratio > 4 per M doubling
    (test/evolve3d_tests.jl:14-21) and energy drift < 1e-3 without KO
    (test/wave3d_tests.jl:50-69)."""
    errs = []
    t1 = 0.1
    for M in (2, 4):
        o = wo.Oracle(4, M)
        u0 = _sinmode(o)
        v0 = np.zeros_like(u0)
        dt = 0.2 * o.hmin
        ns = int(math.ceil(t1 / dt))
        dt = t1 / ns
        u, v, E = o.rk4(u0, v0, dt, ns, energy=True)
        ex = math.cos(math.sqrt(3) * math.pi * t1) * u0
        H = o.hphys()
        errs.append(math.sqrt(np.sum((u - ex) ** 2 * H)))
        assert abs(E[-1] / E[0] - 1) < 1e-3
    assert errs[0] / errs[1] > 4


def test_energy_nonincreasing_with_ko_dirichlet_noise():
    """test/evolve3d_tests.jl:31-36 pattern: noise IC, KO on, closed Dirichlet box:
    the discrete energy does not grow."""
    o = wo.Oracle(4, 2, eps_ko=0.1)
    rng = np.random.default_rng(20260604)
    amp = math.sqrt(np.finfo(float).eps)
    u0 = amp * rng.standard_normal(o.n)
    v0 = amp * rng.standard_normal(o.n)
    dt = 0.2 * o.hmin
    _, _, E = o.rk4(u0, v0, dt, 60, energy=True)
    assert np.all(np.diff(E) <= 1e-12 * E[0])
    assert E[-1] < E[0]


# ------------------------------------------------------------------ curved

def _flat_bg(n):
    bg = np.zeros((n, 10))
    bg[:, 0] = 1.0
    bg[:, 4] = bg[:, 7] = bg[:, 9] = 1.0
    return bg


def _bump_bg(x):
    r2 = np.sum((x - 0.5) ** 2, axis=1)
    e = np.exp(-r2 / 0.1)
    bg = np.zeros((len(x), 10))
    bg[:, 0] = 1 - 0.2 * e
    bg[:, 1:4] = 0.05 * e[:, None] * (x - 0.5)
    g = 1 + 0.2 * e
    bg[:, 4] = bg[:, 7] = bg[:, 9] = g
    return bg


def test_curved_reduces_to_flat():
    """Warp A = 0 with alpha = 1, beta = 0, gamma = delta: the curved operator
    equals the flat one (test/dirichlet_warped_tests.jl:19-36 reduction)."""
    N, M = 4, (2, 3, 2)
    bv = (0.1, -0.2, 0.3, 0.0, 0.5, -0.4)
    of = wo.Oracle(N, M, x1=(1.0, 1.2, 0.9), bdry_values=bv)
    oc = wo.Oracle(N, M, x1=(1.0, 1.2, 0.9), bdry_values=bv, bg=_flat_bg(of.n))
    rng = np.random.default_rng(2026_06_05)
    u = rng.standard_normal(of.n)
    v = rng.standard_normal(of.n)
    a, b = of.lap(u), oc.lap(u, v)
    assert np.max(np.abs(a - b)) < 1e-12 * np.max(np.abs(a))


def test_curved_symmetric_nsd_and_free_stream():
    """Warped mesh + ADM bump background: H3 Ltilde symmetric (SIPG), NSD at the
    effective tau, and a constant is annihilated on a periodic warped mesh
    (free-stream, test/wave3d_tests.jl:121-125)."""
    N, M = 3, 2
    o0 = wo.Oracle(N, M, warp_A=0.05)
    x = o0.coords()
    o = wo.Oracle(N, M, warp_A=0.05, bg=_bump_bg(x))
    n = o.n
    A = _dense(o, lambda z: o.lap(z, np.zeros(n)), n)
    co = o.coefs()
    wgt = co[:, 9]
    _, w, _ = wo.gll(N)
    H3 = np.tile(np.einsum("k,j,i->kji", w, w, w).reshape(-1), o.Ne)
    S = np.diag(H3 / wgt) @ A    # H3 Ltilde = (H3/wgt)(wgt Ltilde)
    assert np.max(np.abs(S - S.T)) < 1e-11 * np.max(np.abs(S))
    assert np.max(np.linalg.eigvalsh(0.5 * (S + S.T))) < 0
    op = wo.Oracle(N, M, warp_A=0.05, periodic=True, bg=_bump_bg(wo.Oracle(N, M, warp_A=0.05, periodic=True).coords()))
    c = np.full(op.n, 2.5)
    assert np.max(np.abs(op.lap(c, np.zeros(op.n)))) < 1e-9


def test_curved_advection_is_skew_in_energy_norm():
    """The u_t advection part A (split form + SAT) is skew in the H3/wgt norm, so
    with KO off the energy is conserved up to time-integration error."""
    N, M = 3, 2
    x = wo.Oracle(N, M, warp_A=0.05).coords()
    o = wo.Oracle(N, M, warp_A=0.05, bg=_bump_bg(x))
    n = o.n
    Av = _dense(o, lambda z: o.lap(np.zeros(n), z), n)
    co = o.coefs()
    _, w, _ = wo.gll(N)
    H3 = np.tile(np.einsum("k,j,i->kji", w, w, w).reshape(-1), o.Ne)
    S = np.diag(H3 / co[:, 9]) @ Av
    assert np.max(np.abs(S + S.T)) < 1e-11 * max(np.max(np.abs(S)), 1e-300)
    assert np.max(np.abs(S)) > 0


def test_warped_rayleigh_quotient():
    """test/dirichlet_warped_tests.jl:59-78: A = 0.05, Rayleigh error < 5% at
    M = 4; geometric-mean ratio > 4; < 1e-3 at M = 8 (N = 4)."""
    errs = []
    for M in (2, 4, 8):
        o = wo.Oracle(4, M, warp_A=0.05)
        x = o.coords()
        u = np.sin(math.pi * x[:, 0]) * np.sin(math.pi * x[:, 1]) * np.sin(math.pi * x[:, 2])
        H = o.hphys()
        lam = -np.dot(u, H * o.lap(u, np.zeros(o.n))) / np.dot(u, H * u)
        errs.append(abs(lam - 3 * math.pi ** 2) / (3 * math.pi ** 2))
    assert errs[1] < 0.05
    assert math.sqrt((errs[0] / errs[1]) * (errs[1] / errs[2])) > 4
    assert errs[2] < 1e-3


def test_curved_energy_conserved_without_ko():
    """Warped mesh + ADM background, KO off: the semi-discrete energy is
    conserved (H3 Ltilde symmetric, advection skew), so the drift over a fixed
    time is RK4's own error: non-increasing and shrinking at >= 4th order in dt."""
    N, M = 4, 2
    x = wo.Oracle(N, M, warp_A=0.05).coords()
    o = wo.Oracle(N, M, warp_A=0.05, bg=_bump_bg(x))
    xx = o.coords()
    u0 = np.exp(-np.sum((xx - 0.45) ** 2, axis=1) / 0.02)
    v0 = np.zeros(o.n)
    drift = []
    for f in (1, 2):
        dt = 0.2 * o.hmin / f
        _, _, E = o.rk4(u0, v0, dt, 20 * f, energy=True)
        assert np.all(np.diff(E) <= 1e-13 * E[0])
        drift.append(abs(E[-1] / E[0] - 1))
    assert drift[1] < drift[0] / 16


# ------------------------------------------------------------------ ADM coefficients against the PDE

_TP = 2 * math.pi


def _adm_bg(x):
    """A smooth periodic static background with a varying lapse, a shift of
    |beta| ~ 0.25 and an off-diagonal metric (subluminal: K stays SPD)."""
    X, Y, Z = x[..., 0], x[..., 1], x[..., 2]
    al = 1 + 0.2 * np.sin(_TP * X) * np.cos(_TP * Y)
    be = 0.25 * np.stack([np.cos(_TP * Y), 0.6 * np.sin(_TP * Z), -0.8 * np.cos(_TP * (X + Z))], -1)
    g = np.empty(X.shape + (3, 3))
    g[..., 0, 0] = 1.2 + np.sin(_TP * Z) / 5
    g[..., 0, 1] = g[..., 1, 0] = np.cos(_TP * X) / 10
    g[..., 0, 2] = g[..., 2, 0] = np.sin(_TP * (Y + Z)) / 20
    g[..., 1, 1] = 1 + 0.15 * np.cos(_TP * (X - Y))
    g[..., 1, 2] = g[..., 2, 1] = -0.07 * np.sin(_TP * X)
    g[..., 2, 2] = 0.9 + np.sin(_TP * Y) / 10
    return al, be, g


def _adm_bg10(x):
    al, be, g = _adm_bg(x)
    return np.concatenate([al[:, None], be, g[:, 0, 0:3], g[:, 1, 1:3], g[:, 2, 2:3]], axis=1).reshape(-1)


def _adm_fields(x):
    """Periodic closed-form u, u_t and their gradients."""
    X, Y, Z = x[..., 0], x[..., 1], x[..., 2]
    u = np.sin(_TP * X + 1) * np.cos(_TP * Y) + np.sin(_TP * Z)
    ut = np.cos(_TP * (X + Y + Z)) + 0.5 * np.sin(_TP * (X - Z))
    gu = np.stack([_TP * np.cos(_TP * X + 1) * np.cos(_TP * Y), -_TP * np.sin(_TP * X + 1) * np.sin(_TP * Y),
                   _TP * np.cos(_TP * Z)], -1)
    s = -_TP * np.sin(_TP * (X + Y + Z))
    c = 0.5 * _TP * np.cos(_TP * (X - Z))
    gut = np.stack([s + c, s, s - c], -1)
    return u, ut, gu, gut


def _adm_K_c(x):
    """The continuous coefficients of the covariant wave equation
    d_mu(sqrt(-g) g^{mu nu} d_nu u) = 0 with sqrt(-g) = alpha sqrt(gamma):
    K^{ij} = sqrt(-g) g^{ij} = sqrt(gamma)(alpha gamma^{ij} - b^i b^j / alpha),
    c^i = sqrt(-g) g^{ti} = sqrt(gamma) b^i / alpha
    (src/wave3d_curved.jl:1-12; METHODS.md:142-150)."""
    al, be, g = _adm_bg(x)
    gi = np.linalg.inv(g)
    sg = np.sqrt(np.linalg.det(g))
    K = sg[..., None, None] * (al[..., None, None] * gi - be[..., :, None] * be[..., None, :] / al[..., None, None])
    return al, be, sg, K, (sg / al)[..., None] * be


def _adm_rhs_exact(x):
    """u_tt = b^i d_i u_t + (alpha / sqrt(gamma)) d_i(K^{ij} d_j u + c^i u_t), the
    divergence by 4th-order central differences (step 1e-3) of the flux's closed form."""
    def flux(y):
        _, _, _, K, c = _adm_K_c(y)
        _, ut, gu, _ = _adm_fields(y)
        return np.einsum("...ij,...j->...i", K, gu) + c * ut[..., None]

    d = 1e-3
    div = np.zeros(x.shape[0])
    for i in range(3):
        e = np.zeros(3)
        e[i] = d
        div += (8 * (flux(x + e)[:, i] - flux(x - e)[:, i]) - (flux(x + 2 * e)[:, i] - flux(x - 2 * e)[:, i])) / (12 * d)
    al, be, sg, _, _ = _adm_K_c(x)
    _, _, _, gut = _adm_fields(x)
    return np.sum(be * gut, axis=1) + al / sg * div


def _adm_pde_errors(N, Ms, plant=0, warp_A=0.05):
    errs = []
    for M in Ms:
        x = wo.Oracle(N, M, periodic=True, warp_A=warp_A).coords()
        with wo.planted_mistake(plant):
            o = wo.Oracle(N, M, periodic=True, warp_A=warp_A, bg=_adm_bg10(x))
        u, ut, _, _ = _adm_fields(x)
        got = o.lap(np.ascontiguousarray(u), np.ascontiguousarray(ut))
        ex = _adm_rhs_exact(x)
        errs.append(np.max(np.abs(got - ex)) / np.max(np.abs(ex)))
    return errs


def test_adm_second_order_rhs_converges_to_the_pde():
    """The second-order ADM operator (wto_lap with a background on a periodic
    warped mesh) against the continuous right-hand side of the covariant wave
    equation on a static background,
    u_tt = b^i d_i u_t + (alpha/sqrt(gamma)) [d_i(K^{ij} d_j u) + d_i(c^i u_t)],
    with lapse, shift and an off-diagonal metric: spectral convergence in M at
    N = 8.  Each of three plausible coefficient mistakes planted in the oracle
    (the -b b / alpha term of K dropped, sqrt(gamma) inverted in the weight, the
    sign of c flipped) leaves an O(1e-2..1) error that does not converge."""
    ok = _adm_pde_errors(8, (2, 4))
    assert ok[1] < 5e-3 and ok[0] / ok[1] > 16, ok
    for plant in (wo.PLANT_DROP_BB, wo.PLANT_WGT_SQRTG, wo.PLANT_FLIP_C):
        bad = _adm_pde_errors(8, (2, 4), plant)
        assert bad[1] > 5 * ok[1] and bad[0] / bad[1] < 4, (plant, bad, ok)


def _adm_fo_minus_so(N, M, plant=0):
    o0 = wo.Oracle(N, M, periodic=True)
    x = o0.coords()
    bg = _adm_bg10(x)
    with wo.planted_mistake(plant):
        so = wo.Oracle(N, M, periodic=True, bg=bg)
    coef = wo.fo_coef(bg)
    al, be, g = _adm_bg(x)
    sg = np.sqrt(np.linalg.det(g))
    u, ut, _, _ = _adm_fields(x)
    u = np.ascontiguousarray(u)
    ut = np.ascontiguousarray(ut)
    Du = np.stack([o0.apply_D(a, u) for a in range(3)], 1)
    Dut = np.stack([o0.apply_D(a, ut) for a in range(3)], 1)
    Pi = np.ascontiguousarray(sg / al * (ut - np.sum(be * Du, 1)))
    _, dPi = o0.fo_rhs(u, Pi, coef, eps_ko=0.0)
    utt_fo = np.sum(be * Dut, 1) + al / sg * dPi
    utt_so = so.lap(u, ut)
    return np.max(np.abs(utt_fo - utt_so)) / np.max(np.abs(utt_so))


def test_adm_second_order_equals_first_order_form():
    """The second-order ADM operator against the reference's own first-order
    system (wave3d_curved_rhs!, src/wave3d_curved.jl:205-258, pinned to the PDE
    by test_fo_rhs_converges_to_the_pde): with Phi = u and
    Pi = (sqrt(gamma)/alpha)(u_t - b^i D_i u), u_tt = b^i D_i u_t +
    (alpha/sqrt(gamma)) dPi.  On a periodic affine mesh and data continuous across
    element faces (a smooth function at the GLL nodes) the SIPG jump terms
    vanish and the centred-flux SAT of apply_D! carries the same face term as
    SIPG's flux average, so the two forms agree to rounding; each planted
    coefficient mistake breaks that by 4e-2 or more."""
    for M in (2, 3):
        assert _adm_fo_minus_so(8, M) < 1e-11
    for plant in (wo.PLANT_DROP_BB, wo.PLANT_WGT_SQRTG, wo.PLANT_FLIP_C):
        assert _adm_fo_minus_so(8, 2, plant) > 1e-2, plant


# ------------------------------------------------------------------ Sommerfeld faces

def _face_measure(o, f, faces):
    """Surface integral over the given 1-based faces of the unit box of f(x,y,z),
    by a tensor Gauss-Legendre rule of its own (exact for the polynomials used)."""
    g, w = npleg.leggauss(8)
    g = 0.5 * (g + 1); w = 0.5 * w
    tot = 0.0
    for face in faces:
        a, side = (face - 1) // 2, float((face - 1) % 2)
        b, c = [d for d in range(3) if d != a]
        for p, wp in zip(g, w):
            for q, wq in zip(g, w):
                x = [0.0, 0.0, 0.0]
                x[a] = side; x[b] = p; x[c] = q
                tot += wp * wq * f(*x)
    return tot


def test_sommerfeld_surface_terms_closed_form():
    """Bayliss-Turkel faces (tag 7, src/wave.jl:453-518): <v, L(0, v)>_H is minus
    the surface integral of v^2 over the Sommerfeld faces, <1, L(1, 0)>_H is minus
    their area over R, and with R = inf a constant lies in the operator's kernel
    (the faces are free for the SIPG terms)."""
    faces = (1, 2, 4)
    R = 3.0
    o = wo.Oracle(5, 2, sommerfeld_faces=faces, sommerfeld_R=R)
    x = o.coords()
    H = o.hphys()
    fv = lambda X, Y, Z: 1.0 + 0.5 * X - Y * Y + 0.25 * X * Z
    v = fv(x[:, 0], x[:, 1], x[:, 2])
    Lv = o.lap(np.zeros(o.n), v)
    assert abs(np.dot(H * v, Lv) + _face_measure(o, lambda *X: fv(*X) ** 2, faces)) < 1e-12
    one = np.ones(o.n)
    allf = (1, 2, 3, 4, 5, 6)
    oR = wo.Oracle(5, 2, sommerfeld_faces=allf, sommerfeld_R=R)
    assert abs(np.dot(oR.hphys(), oR.lap(one)) + 6 / R) < 1e-12
    # Dirichlet faces still pull constants down, Sommerfeld ones at R = inf do not
    oi = wo.Oracle(5, 2, sommerfeld_faces=allf)
    assert np.max(np.abs(oi.lap(one))) < 1e-10
    assert np.max(np.abs(oi.lap(one, np.zeros(oi.n)))) < 1e-10
    od = wo.Oracle(5, 2)
    assert np.max(np.abs(od.lap(one))) > 1.0
    # Sommerfeld faces only on bounded axes, and not with a background
    with pytest.raises(ValueError):
        wo.Oracle(4, 2, periodic=(True, False, False), sommerfeld_faces=(2,))
    with pytest.raises(ValueError):
        wo.Oracle(4, 2, bg=_flat_bg(4 ** 3 * 8), sommerfeld_faces=(1,))


def test_sommerfeld_operator_symmetric_nsd():
    """With u_t = 0 the Sommerfeld faces leave L H-symmetric and negative
    semidefinite (the operator of the energy 1/2|v|^2 + 1/2|grad u|^2 +
    1/(2R) |u|^2 on those faces); a wrong sign in the normal derivative or a
    missing weight breaks the symmetry."""
    o = wo.Oracle(4, (2, 2, 1), sommerfeld_faces=(1, 4, 5), sommerfeld_R=2.0)
    A = _dense(o, o.lap, o.n)
    H = o.hphys()
    S = H[:, None] * A
    assert np.max(np.abs(S - S.T)) < 1e-10 * np.max(np.abs(S))
    ev = np.linalg.eigvalsh(0.5 * (S + S.T))
    assert ev.max() < 1e-10 * np.abs(ev).max()


def test_sommerfeld_energy_rate_and_decay():
    """dE/dt = <v, L(u, v) - L(u, 0)>_H = -(surface integral of v^2) for any u,
    and RK4 with Sommerfeld faces and noise data never gains energy."""
    faces = (2, 3, 6)
    o = wo.Oracle(4, 2, sommerfeld_faces=faces, sommerfeld_R=1.5, eps_ko=0.1)
    rng = np.random.default_rng(77)
    u = rng.standard_normal(o.n)
    x = o.coords()
    fv = lambda X, Y, Z: np.cos(X + 2 * Y) * (1 + Z)
    v = fv(x[:, 0], x[:, 1], x[:, 2])
    H = o.hphys()
    rate = np.dot(H * v, o.lap(u, v) - o.lap(u))
    ref = _face_measure(o, lambda *X: fv(*X) ** 2, faces)
    assert abs(rate + ref) < 2e-4 * ref      # GLL face quadrature of a non-polynomial
    amp = math.sqrt(np.finfo(float).eps)
    u0 = amp * rng.standard_normal(o.n)
    v0 = amp * rng.standard_normal(o.n)
    _, _, E = o.rk4(u0, v0, 0.2 * o.hmin, 60, energy=True)
    assert np.all(np.diff(E) <= 1e-12 * E[0])
    assert E[-1] < 0.9 * E[0]


def test_sommerfeld_absorbs_outgoing_plane_wave():
    """A right-moving pulse f(x - t) leaves through a Sommerfeld face at x = 1
    (exact for normal incidence at R = inf) and is reflected back by a Dirichlet
    face; y and z are periodic.  What stays behind is discretisation error and
    falls fast with resolution."""
    def run(faces, M=12):
        o = wo.Oracle(6, (M, 1, 1), periodic=(False, True, True), sommerfeld_faces=faces)
        x = o.coords()
        s = (x[:, 0] - 0.5) / 0.08
        u0 = np.exp(-s * s)
        v0 = 2 * s / 0.08 * u0      # u_t = -u_x for f(x - t)
        dt = 0.25 * o.hmin
        ns = int(math.ceil(0.9 / dt))
        _, _, E = o.rk4(u0, v0, 0.9 / ns, ns, energy=True)
        return E
    Es = run((1, 2))
    Ed = run(())
    assert abs(Ed[-1] / Ed[0] - 1) < 1e-3
    assert Es[-1] < 1e-6 * Es[0]
    Ec = run((1, 2), M=6)
    assert Es[-1] / Es[0] < 1e-3 * Ec[-1] / Ec[0]


# ------------------------------------------------------------------ first-order (Phi, Pi) system

def _flat_bg(n):
    return np.tile([1.0, 0, 0, 0, 1.0, 0, 0, 1.0, 0, 1.0], n)


def _dxmin(N, M):
    xi, _, _ = wo.gll(N)
    return np.min(np.diff(xi)) / M


def test_fo_periodic_plane_wave_energy_and_convergence():
    """test/wave3d_tests.jl:48-69: periodic plane wave sin(k(x+y+z) - wt) on the
    flat background, N = 4, M = 4 and 8, to t ~ 0.2 with dt = 0.1 dx_min: energy
    drift < 1e-3 and the error falls by more than 4x."""
    N = 4
    kap = 2 * math.pi; om = kap * math.sqrt(3.0)
    errs = []
    for M in (4, 8):
        o = wo.Oracle(N, M, periodic=True)
        coef = wo.fo_coef(_flat_bg(o.n))
        x = o.coords(); s = x.sum(axis=1)
        Phi = np.sin(kap * s); Pi = -om * np.cos(kap * s)
        E0 = o.fo_energy(Phi, Pi, coef)
        dt = 0.1 * _dxmin(N, M); nst = int(math.ceil(0.2 / dt))
        Phi, Pi = o.fo_rk4(Phi, Pi, coef, dt, nst, eps_ko=0.0)
        E1 = o.fo_energy(Phi, Pi, coef)
        assert abs(E1 / E0 - 1) < 1e-3
        ex = np.sin(kap * s - om * nst * dt)
        errs.append(math.sqrt(np.sum((Phi - ex) ** 2 * o.hphys())))
    assert errs[0] / errs[1] > 4


@pytest.mark.parametrize("periodic", [True, False])
def test_fo_spectrum_stable(periodic):
    """test/wave3d_tests.jl:71-90: N = 4, M = 2, KO 0.1, periodic or Sommerfeld on
    all six faces: max Re(lambda) <= 1e-5 max |lambda| of the dense RHS."""
    o = wo.Oracle(4, 2, periodic=periodic)
    coef = wo.fo_coef(_flat_bg(o.n))
    bc = None if periodic else wo.make_bc3d(["sommerfeld"] * 6)
    nn = o.n
    A = np.zeros((2 * nn, 2 * nn))
    z = np.zeros(nn)
    for j in range(2 * nn):
        P = z.copy(); Q = z.copy()
        (P if j < nn else Q)[j % nn] = 1.0
        a, b = o.fo_rhs(P, Q, coef, 0.1, bc)
        A[:nn, j] = a; A[nn:, j] = b
    lam = np.linalg.eigvals(A)
    assert np.max(lam.real) <= 1e-5 * np.max(np.abs(lam))


def test_fo_noise_bounded_with_sommerfeld():
    """test/wave3d_tests.jl:92-101: noise of size sqrt(eps), Sommerfeld faces, KO
    0.1, to t = 8: stays finite and below 1000x its size."""
    N, M = 4, 2
    o = wo.Oracle(N, M)
    coef = wo.fo_coef(_flat_bg(o.n))
    rng = np.random.default_rng(20260604)
    amp = math.sqrt(np.finfo(float).eps)
    Phi = amp * rng.standard_normal(o.n); Pi = amp * rng.standard_normal(o.n)
    dt = 0.1 * _dxmin(N, M); nst = int(math.ceil(8 / dt))
    Phi, Pi = o.fo_rk4(Phi, Pi, coef, dt, nst, eps_ko=0.1, bc=wo.make_bc3d(["sommerfeld"] * 6))
    assert np.all(np.isfinite(Phi)) and np.max(np.abs(Phi)) < 1000 * amp


def _metric_bg(x, alpha_on=False, shift_on=False):
    """A smooth periodic static background: gamma SPD with off-diagonal terms;
    alpha = 1 and beta = 0 unless switched on."""
    X, Y, Z = (2 * math.pi * x[:, d] for d in range(3))
    n = x.shape[0]
    bg = np.zeros((n, 10))
    bg[:, 0] = 1.0 + (0.2 * np.sin(X) * np.cos(Y) if alpha_on else 0.0)
    if shift_on:
        bg[:, 1] = 0.1 * np.cos(Y); bg[:, 2] = 0.05 * np.sin(Z); bg[:, 3] = -0.08 * np.cos(X + Z)
    bg[:, 4] = 1.2 + 0.2 * np.sin(Z)
    bg[:, 5] = 0.1 * np.cos(X)
    bg[:, 6] = 0.05 * np.sin(Y + Z)
    bg[:, 7] = 1.0 + 0.15 * np.cos(X - Y)
    bg[:, 8] = -0.07 * np.sin(X)
    bg[:, 9] = 0.9 + 0.1 * np.sin(Y)
    return bg.reshape(-1)


def test_fo_energy_exactly_conserved_on_static_metric():
    """With alpha = 1, beta = 0 and any static gamma the semi-discrete energy of
    wave3d_energy is conserved exactly on a periodic mesh (H D skew): its rate,
    taken as the central difference of the quadratic E along the RHS, vanishes
    to rounding; with a varying lapse it does not (the rate is O(1))."""
    o = wo.Oracle(4, (3, 2, 2), periodic=True)
    x = o.coords()
    rng = np.random.default_rng(5)
    Phi = rng.standard_normal(o.n); Pi = rng.standard_normal(o.n)
    rates = []
    for alpha_on in (False, True):
        coef = wo.fo_coef(_metric_bg(x, alpha_on=alpha_on))
        dPhi, dPi = o.fo_rhs(Phi, Pi, coef, eps_ko=0.0)
        d = 1e-3
        Ep = o.fo_energy(Phi + d * dPhi, Pi + d * dPi, coef)
        Em = o.fo_energy(Phi - d * dPhi, Pi - d * dPi, coef)
        scale = o.fo_energy(Phi, Pi, coef) * math.sqrt(np.sum(dPhi ** 2 + dPi ** 2) / np.sum(Phi ** 2 + Pi ** 2))
        rates.append(abs((Ep - Em) / (2 * d)) / scale)
    assert rates[0] < 1e-11
    assert rates[1] > 1e-3 * 1e-3 and rates[1] > 1e4 * rates[0]


def test_fo_rhs_converges_to_the_pde():
    """The discrete RHS against the continuous one,
    dPhi/dt = b.grad Phi + (alpha/sqrtg) Pi, dPi/dt = div(b Pi + alpha sqrtg g^-1 grad Phi),
    for smooth periodic data and a background with lapse, shift and an
    off-diagonal metric (the divergence of the continuous flux by 4th-order
    central differences of its closed form, step 1e-3): spectral convergence
    in M at N = 6."""
    tp = 2 * math.pi

    def bgv(x):
        X, Y, Z = x[..., 0], x[..., 1], x[..., 2]
        al = 1 + 0.2 * np.sin(tp * X) * np.cos(tp * Y)
        be = np.stack([0.1 * np.cos(tp * Y), 0.05 * np.sin(tp * Z), -0.08 * np.cos(tp * (X + Z))], -1)
        g = np.empty(X.shape + (3, 3))
        g[..., 0, 0] = 1.2 + np.sin(tp * Z) / 5
        g[..., 0, 1] = g[..., 1, 0] = np.cos(tp * X) / 10
        g[..., 0, 2] = g[..., 2, 0] = np.sin(tp * (Y + Z)) / 20
        g[..., 1, 1] = 1 + 0.15 * np.cos(tp * (X - Y))
        g[..., 1, 2] = g[..., 2, 1] = -0.07 * np.sin(tp * X)
        g[..., 2, 2] = 0.9 + np.sin(tp * Y) / 10
        return al, be, g

    def fields(x):
        X, Y, Z = x[..., 0], x[..., 1], x[..., 2]
        Phi = np.sin(tp * X + 1) * np.cos(tp * Y) + np.sin(tp * Z)
        Pi = np.cos(tp * (X + Y + Z))
        grad = np.stack([tp * np.cos(tp * X + 1) * np.cos(tp * Y), -tp * np.sin(tp * X + 1) * np.sin(tp * Y),
                         tp * np.cos(tp * Z)], -1)
        return Phi, Pi, grad

    def flux(x):
        al, be, g = bgv(x)
        _, Pi, grad = fields(x)
        sg = np.sqrt(np.linalg.det(g))
        return be * Pi[..., None] + (al * sg)[..., None] * np.einsum("...ij,...j->...i", np.linalg.inv(g), grad)

    def div_flux(x):
        d = 1e-3
        out = np.zeros(x.shape[0])
        for i in range(3):
            e = np.zeros(3); e[i] = d
            out += (8 * (flux(x + e)[:, i] - flux(x - e)[:, i]) - (flux(x + 2 * e)[:, i] - flux(x - 2 * e)[:, i])) / (12 * d)
        return out

    errs = []
    for M in (2, 4):
        o = wo.Oracle(6, M, periodic=True)
        x = o.coords()
        al, be, g = bgv(x)
        bg = np.concatenate([al[:, None], be, g[:, 0, 0:3], g[:, 1, 1:3], g[:, 2, 2:3]], axis=1).reshape(-1)
        coef = wo.fo_coef(bg)
        Phi, Pi, grad = fields(x)
        sg = np.sqrt(np.linalg.det(g))
        dPhi_ex = np.sum(be * grad, axis=1) + al / sg * Pi
        dPi_ex = div_flux(x)
        dP, dQ = o.fo_rhs(np.ascontiguousarray(Phi), np.ascontiguousarray(Pi), coef, eps_ko=0.0)
        errs.append((np.max(np.abs(dP - dPhi_ex)) / np.max(np.abs(dPhi_ex)),
                     np.max(np.abs(dQ - dPi_ex)) / np.max(np.abs(dPi_ex))))
    assert errs[1][0] < 1e-3 and errs[1][1] < 1e-2, errs
    assert errs[0][0] / errs[1][0] > 16 and errs[0][1] / errs[1][1] > 8, errs


def test_fo_boundary_pass_terms():
    """The affine boundary pass (src/boundaries3d.jl:76-126) on a flat background:
    full-state Dirichlet pulls (Phi, Pi) to the targets at the face nodes only, with
    rate sigma (|a_n - b| + |a_n + b|) / (h w); Sommerfeld drives r = Pi + d_n Phi;
    excision adds nothing."""
    N, M = 4, 2
    o = wo.Oracle(N, M)
    coef = wo.fo_coef(_flat_bg(o.n))
    rng = np.random.default_rng(9)
    Phi = rng.standard_normal(o.n); Pi = rng.standard_normal(o.n)
    base = o.fo_rhs(Phi, Pi, coef, eps_ko=0.0)
    ex = o.fo_rhs(Phi, Pi, coef, eps_ko=0.0, bc=wo.make_bc3d(["excision"] * 6))
    assert np.array_equal(base[0], ex[0]) and np.array_equal(base[1], ex[1])
    x = o.coords()
    xi, w, _ = wo.gll(N)
    h = 1.0 / M
    on1 = np.isclose(x[:, 0], 0.0)              # face 1 (-x)
    gP = np.full(o.n, 0.3); gQ = np.full(o.n, -0.2)
    fd = o.fo_rhs(Phi, Pi, coef, eps_ko=0.0,
                  bc=wo.make_bc3d(["full_dirichlet", "excision", "excision", "excision", "excision", "excision"],
                                  sigma=0.5, gPhi=gP, gPi=gQ))
    rate = 0.5 * 2.0 / (h * w[0])              # a_n = 1, beta = 0
    assert np.allclose(fd[0][on1] - base[0][on1], -rate * (Phi[on1] - 0.3), rtol=1e-12, atol=1e-12)
    assert np.allclose(fd[1][on1] - base[1][on1], -rate * (Pi[on1] + 0.2), rtol=1e-12, atol=1e-12)
    assert np.array_equal(fd[0][~on1], base[0][~on1]) and np.array_equal(fd[1][~on1], base[1][~on1])
    # Sommerfeld on +x: r = Pi + d_x Phi (n = +1), rate 1 / (h w_N)
    onx = np.isclose(x[:, 0], 1.0)
    so = o.fo_rhs(Phi, Pi, coef, eps_ko=0.0,
                  bc=wo.make_bc3d(["excision", "sommerfeld", "excision", "excision", "excision", "excision"]))
    Dx = o.apply_D(0, Phi)
    assert np.allclose(so[1][onx] - base[1][onx], -(Pi[onx] + Dx[onx]) / (h * w[-1]), rtol=1e-12, atol=1e-10)
    assert np.array_equal(so[0], base[0])


def _warped_face_weights(o, N, M, A, faces):
    """Per node: sum over the Sommerfeld faces through it of w_p w_q |n_u|, with
    n_u = J_p x J_q from the warp map x = x0 + L (xi + A s(xi)),
    s = sin(2 pi xi_1) sin(2 pi xi_2) sin(2 pi xi_3), xi the global unit
    coordinates (element e_a covers [e_a, e_a + 1] / M_a)."""
    xi, w, _ = wo.gll(N)
    out = np.zeros(o.n)
    tp = 2 * math.pi
    for e in range(M[0] * M[1] * M[2]):
        ec = (e % M[0], (e // M[0]) % M[1], e // (M[0] * M[1]))
        for f in faces:
            a, side = (f - 1) // 2, (f - 1) % 2
            if ec[a] != (M[a] - 1 if side else 0):
                continue
            r = N - 1 if side else 0
            pax = [b for b in range(3) if b != a]
            for q in range(N):
                for pp in range(N):
                    idx = [0, 0, 0]
                    idx[a] = r; idx[pax[0]] = pp; idx[pax[1]] = q
                    g = [(ec[b] + xi[idx[b]]) / M[b] for b in range(3)]
                    sn = [math.sin(tp * g[b]) for b in range(3)]
                    cs = [math.cos(tp * g[b]) for b in range(3)]
                    ds = [tp * cs[0] * sn[1] * sn[2], tp * sn[0] * cs[1] * sn[2], tp * sn[0] * sn[1] * cs[2]]
                    J = np.array([[((1.0 if i == b else 0.0) + A * ds[b]) / M[b] for b in range(3)] for i in range(3)])
                    nu = np.linalg.norm(np.cross(J[:, pax[0]], J[:, pax[1]]))
                    n = e * N ** 3 + idx[0] + N * (idx[1] + N * idx[2])
                    out[n] += w[pp] * w[q] * nu
    return out


def test_sommerfeld_on_warped_meshes():
    """Sommerfeld faces on a warped mesh (no background): the drag
    -(wF/Hphys)(u_t + d_n u + u/R) of src/wave.jl:444-518 with the physical
    normal.  (1) dE/dt = <v, L(u, v) - L(u, 0)>_W equals minus the sum of
    w_p w_q |J_p x J_q| v^2 over the Sommerfeld face nodes, |J_p x J_q| taken from
    the warp map itself; (2) with u_t = 0 the operator stays W-symmetric and
    negative semidefinite; (3) as A -> 0 it tends to the flat Sommerfeld
    operator linearly in A."""
    N, M, A = 4, (2, 3, 2), 0.05
    faces = (1, 2, 3, 4, 5, 6)
    o = wo.Oracle(N, M, warp_A=A, sommerfeld_faces=faces, sommerfeld_R=1.5)
    rng = np.random.default_rng(5)
    u = rng.standard_normal(o.n); v = rng.standard_normal(o.n)
    W = o.hphys()
    rate = np.dot(W * v, o.lap(u, v) - o.lap(u))
    ref = np.dot(_warped_face_weights(o, N, M, A, faces), v * v)
    assert abs(rate + ref) < 1e-11 * ref
    # a subset of faces and R = inf
    sub = (2, 3, 5)
    o2 = wo.Oracle(N, M, warp_A=A, sommerfeld_faces=sub)
    rate2 = np.dot(o2.hphys() * v, o2.lap(u, v) - o2.lap(u))
    assert abs(rate2 + np.dot(_warped_face_weights(o2, N, M, A, sub), v * v)) < 1e-11 * abs(rate2)
    # (2)
    o3 = wo.Oracle(3, (2, 2, 1), warp_A=A, sommerfeld_faces=(1, 4, 5), sommerfeld_R=2.0)
    S = o3.hphys()[:, None] * _dense(o3, o3.lap, o3.n)
    assert np.max(np.abs(S - S.T)) < 1e-10 * np.max(np.abs(S))
    ev = np.linalg.eigvalsh(0.5 * (S + S.T))
    assert ev.max() < 1e-10 * np.abs(ev).max()
    # (3)
    of = wo.Oracle(4, (3, 2, 2), sommerfeld_faces=(2, 3, 6), sommerfeld_R=2.0)
    d = []
    for eps in (1e-7, 1e-8):
        oc = wo.Oracle(4, (3, 2, 2), warp_A=eps, sommerfeld_faces=(2, 3, 6), sommerfeld_R=2.0)
        a = of.lap(u[:of.n], v[:of.n]); b = oc.lap(u[:of.n], v[:of.n])
        d.append(np.max(np.abs(a - b)) / np.max(np.abs(a)))
    assert d[0] < 1e-5 and 8 < d[0] / d[1] < 12


def _mesh_oracle(mesh, **kw):
    return wo.Oracle.from_mesh(mesh.N, mesh.coords, mesh.neighbour, mesh.neighbour_face, mesh.bdry, **kw)


def test_oracle_general_connectivity_reduces_to_uniform_operator():
    # an affine mesh given by coordinates, with its elements renumbered and each
    # element's axes turned by a symmetry of the cube, is the uniform operator
    import wavetoy_hip as wh
    N, M = 4, (3, 2, 2)
    rng = np.random.default_rng(3)
    bv = (0.3, -0.2, 0.1, 0.4, -0.5, 0.25)
    ref = wo.Oracle(N, M, bdry_values=bv)
    X = ref.coords().reshape(-1, N, N, N, 3)
    Ne = X.shape[0]
    perm = rng.permutation(Ne)
    idx = np.arange(Ne * N ** 3).reshape(Ne, N, N, N)
    Xn, src = np.empty_like(X), np.empty_like(idx)
    for en, eo in enumerate(perm):
        axes = tuple(rng.permutation(3))
        A, I = np.transpose(X[eo], axes + (3,)), np.transpose(idx[eo], axes)
        for d in range(3):
            if rng.integers(0, 2):
                A, I = np.flip(A, axis=d), np.flip(I, axis=d)
        Xn[en], src[en] = A, I
    src = src.reshape(-1)
    nbr, nface, bdry = wh.connect_by_coords(N, Xn)
    for e in range(Ne):
        for f in range(6):
            if nbr[e, f] == 0:
                c = wh.core._face_nodes(Xn[e:e + 1], f).reshape(-1, 3).mean(axis=0)
                a = int(np.argmin(np.minimum(np.abs(c), np.abs(1 - c))))
                bdry[e, f] = 2 * a + (2 if c[a] > 0.5 else 1)
    om = wo.Oracle.from_mesh(N, Xn, nbr, nface, bdry, bdry_values=bv)
    u = rng.standard_normal(ref.n)
    want, got = ref.lap(u), om.lap(u[src])
    # the turned elements meet with every D4 code and with like and unlike
    # sides, so an unmapped neighbour node or a wrong outward sign of the
    # neighbour's flux breaks this equality
    assert np.max(np.abs(got - want[src])) < 1e-11 * np.max(np.abs(want))


def test_oracle_cubed_cube_rayleigh_and_symmetry():
    import wavetoy_hip as wh
    # test/dirichlet_cubed_cube_tests.jl:54-81 (N = 4, M = 2, 4): lambda = 3 pi^2 / 4
    lam = 3 * math.pi ** 2 / 4
    errs = []
    for M in (2, 4):
        mesh = wh.make_cubed_cube_mesh(4, M, 0.3)
        o = _mesh_oracle(mesh)
        x = o.coords()
        u = np.prod(np.cos(0.5 * math.pi * x), axis=1)
        H = o.hphys()
        errs.append(abs(-np.dot(u * H, o.lap(u)) / np.dot(u * H, u) - lam))
    assert errs[1] < errs[0] / 4 and errs[1] < 1e-2, errs
    mesh = wh.make_inflated_cube_mesh(3, 0.1, 0.3, 1.0, 1)
    o = _mesh_oracle(mesh)
    H = o.hphys()
    rng = np.random.default_rng(0)
    u, v = rng.standard_normal(o.n), rng.standard_normal(o.n)
    a, b = np.dot(v * H, o.lap(u)), np.dot(u * H, o.lap(v))
    assert abs(a - b) < 1e-11 * math.sqrt(abs(np.dot(u * H, o.lap(u)) * np.dot(v * H, o.lap(v))))
    assert np.dot(u * H, o.lap(u)) < 0


# ------------------------------------------------------------------ first-order system, curvilinear meshes

def _flat_bg10(n, shift=(0.0, 0.0, 0.0)):
    bg = np.zeros((n, 10))
    bg[:, 0] = 1.0
    bg[:, 1:4] = shift
    bg[:, 4] = bg[:, 7] = bg[:, 9] = 1.0
    return bg


def _outer_nodes(mesh):
    N = mesh.N
    out = np.zeros((mesh.Ne, N, N, N), bool)
    for e, f in zip(*np.nonzero(mesh.neighbour == 0)):
        sl = [e, slice(None), slice(None), slice(None)]
        sl[3 - f // 2] = N - 1 if f % 2 else 0
        out[tuple(sl)] = True
    return out.reshape(-1)


def test_curv_metric_free_stream_and_skew_pair():
    """Curvilinear first-order path (src/wave3d_curved.jl:218-238, METHODS.md
    "Curvilinear path"): with the conservative-curl metric terms a constant
    has a zero gradient and a constant vector a zero divergence to rounding
    (the discrete metric identities) on the cubed cube, the inflated cube and
    the radial shell; the gradient of a linear function is exact where the
    element maps are polynomials of degree < N (cubed cube); and the split-form
    gradient and divergence with the centred SAT are a skew-adjoint pair in
    Hd = w_i w_j w_k |J|: <v, div F> + <grad v, F> = 0 for fields that vanish at
    the outer face nodes (the interface terms cancel, across unlike and
    turned faces and left-handed elements)."""
    import wavetoy_hip as wh
    meshes = [wh.make_cubed_cube_mesh(4, 2, 0.3), wh.make_inflated_cube_mesh(4, 0.1, 0.3, 1.0, 1),
              wh.make_radial_shell_mesh(5, 0.5, 1.0, 1)]
    rng = np.random.default_rng(11)
    for q, mesh in enumerate(meshes):
        o = _mesh_oracle(mesh)
        JA, J = o.curv_metric()
        scale = np.max(np.abs(o.curv_grad(rng.standard_normal(o.n))))
        assert np.max(np.abs(o.curv_grad(np.full(o.n, 2.5)))) < 1e-13 * scale
        assert np.max(np.abs(o.curv_div(np.stack([np.ones(o.n), -0.5 * np.ones(o.n), 0.3 * np.ones(o.n)])))) < 1e-13 * scale
        if q == 0:
            x = o.coords()
            g = o.curv_grad(x[:, 0] + 2 * x[:, 1] - x[:, 2])
            assert np.max(np.abs(g - np.array([1.0, 2.0, -1.0])[:, None])) < 1e-12
        out = _outer_nodes(mesh)
        v = rng.standard_normal(o.n)
        F = rng.standard_normal((3, o.n))
        v[out] = 0.0
        F[:, out] = 0.0
        _, w, _ = wo.gll(mesh.N)
        Hd = np.tile(np.einsum("k,j,i->kji", w, w, w).reshape(-1), mesh.Ne) * J
        a = np.dot(v * Hd, o.curv_div(F))
        b = np.sum(o.curv_grad(v) * F * Hd)
        assert abs(a + b) < 1e-12 * abs(a), (q, a, b)


def test_curv_grad_div_converge():
    """Gradient and divergence of smooth closed-form fields on the inflated
    cube (13 patches, spherical shells): max-norm error falls with M, by more
    than 4x per halving at N = 4 and 12x at N = 6."""
    import wavetoy_hip as wh
    for N, Ms, rate in ((4, (1, 2, 4), 4.0), (6, (1, 2), 12.0)):
        errs = []
        for M in Ms:
            o = _mesh_oracle(wh.make_inflated_cube_mesh(N, 0.1, 0.3, 1.0, M))
            x = o.coords()
            phi = np.sin(2 * x[:, 0] + 1) * np.cos(1.5 * x[:, 1]) + x[:, 2] ** 3
            gex = np.stack([2 * np.cos(2 * x[:, 0] + 1) * np.cos(1.5 * x[:, 1]),
                            -1.5 * np.sin(2 * x[:, 0] + 1) * np.sin(1.5 * x[:, 1]), 3 * x[:, 2] ** 2])
            F = np.stack([np.sin(x[:, 1]) + x[:, 0] ** 2, x[:, 0] * x[:, 2], np.cos(x[:, 0] + x[:, 2])])
            dex = 2 * x[:, 0] - np.sin(x[:, 0] + x[:, 2])
            eg = np.max(np.abs(o.curv_grad(phi) - gex)) / np.max(np.abs(gex))
            ed = np.max(np.abs(o.curv_div(F) - dex)) / np.max(np.abs(dex))
            errs.append(max(eg, ed))
        for a, b in zip(errs, errs[1:]):
            assert a / b > rate, (N, errs)


def test_curv_fo_reduces_to_the_affine_path():
    """An axis-aligned uniform mesh handed over as a general mesh: the
    curvilinear right-hand side (metric terms, split form, centred SAT, KO
    along each element axis with the element extent as h, the physical-normal
    boundary pass with Sommerfeld or full-Dirichlet data) equals the affine
    path's wave3d_curved_rhs!, itself pinned to the PDE and the reference's
    tests, to 1e-12; mu and the energy agree too."""
    import wavetoy_hip as wh
    N, M = 5, (3, 2, 2)
    ref = wo.Oracle(N, M, x1=(1.0, 0.7, 0.8))
    X = ref.coords().reshape(-1, N, N, N, 3)
    nbr, nface, bdry = wh.connect_by_coords(N, X, 7)
    om = wo.Oracle.from_mesh(N, X, nbr, nface, bdry)
    x = ref.coords()
    rng = np.random.default_rng(3)
    bg = _flat_bg10(ref.n, 0.05 * rng.standard_normal(3))
    bg[:, 0] = 1 + 0.1 * np.sin(3 * x[:, 0])
    bg[:, 4:10] = (1.1, 0.05, 0.02, 0.9, -0.03, 1.2)
    coef = wo.fo_coef(bg.reshape(-1))
    Phi, Pi = rng.standard_normal(ref.n), rng.standard_normal(ref.n)
    assert abs(om.curv_mu() / ref.mu - 1) < 1e-12
    for eps in (0.0, 0.1):
        for kind, kw in (("sommerfeld", {}), ("full_dirichlet", dict(gPhi=0.3 * Phi, gPi=-0.2 * Pi))):
            a = ref.fo_rhs(Phi, Pi, coef, eps, wo.make_bc3d([kind] * 6, **kw))
            b = om.fo_rhs_curv(Phi, Pi, coef, eps, wo.make_bc3d(kind, **kw))
            for i in range(2):
                assert np.max(np.abs(a[i] - b[i])) < 1e-12 * np.max(np.abs(a[i])), (eps, kind, i)
    assert abs(om.fo_energy_curv(Phi, Pi, coef) / ref.fo_energy(Phi, Pi, coef) - 1) < 1e-13


def test_curv_fo_sommerfeld_spectrum_and_excision_drain():
    """On the 7-patch ball (N = 3) with the Sommerfeld outer sphere and no KO
    the first-order operator's spectrum has max Re lambda <= rounding (the
    reference's curved-Sommerfeld stability check); on the radial shell with
    the outward radial shift (superluminal at the excised inner sphere, tag 8,
    no SAT there) a noisy evolution drains out with non-increasing energy
    (METHODS.md "3D BH-excision")."""
    import wavetoy_hip as wh
    mesh = wh.make_cubed_ball_mesh(3, 1, 1.0, outer_tag=7)
    o = _mesh_oracle(mesh)
    n = o.n
    coef = wo.fo_coef(_flat_bg10(n).reshape(-1))
    bc = wo.make_bc3d("sommerfeld")
    A = np.zeros((2 * n, 2 * n))
    for j in range(2 * n):
        e = np.zeros(2 * n)
        e[j] = 1.0
        a, b = o.fo_rhs_curv(e[:n], e[n:], coef, 0.0, bc)
        A[:n, j], A[n:, j] = a, b
    lam = np.linalg.eigvals(A)
    assert lam.real.max() < 1e-12 * np.abs(lam).max()
    R1, R2, V = 0.5, 1.0, 1.5
    mesh = wh.make_radial_shell_mesh(4, R1, R2, 2)
    o = _mesh_oracle(mesh)
    x = o.coords()
    r = np.linalg.norm(x, axis=1)
    br = V + (0.05 - V) * (r - R1) / (R2 - R1)
    bg = _flat_bg10(o.n)
    bg[:, 1:4] = (br / r)[:, None] * x
    coef = wo.fo_coef(bg.reshape(-1))
    bc = wo.make_bc3d("sommerfeld", excision_tag=8)
    rng = np.random.default_rng(0)
    Phi, Pi = 1e-3 * rng.standard_normal(o.n), 1e-3 * rng.standard_normal(o.n)
    dt = 0.1 * o.hmin / (V + 1)
    E = [o.fo_energy_curv(Phi, Pi, coef)]
    for _ in range(6):
        Phi, Pi = o.fo_rk4_curv(Phi, Pi, coef, dt, 20, eps_ko=0.1, bc=bc)
        E.append(o.fo_energy_curv(Phi, Pi, coef))
    assert np.all(np.diff(E) < 0) and E[-1] < 0.2 * E[0], E


def test_curv_fo_plane_wave_converges_with_dirichlet_data():
    """The curvilinear field-radiation Dirichlet pass with exact boundary data
    (Pi and grad Phi of the plane wave, src/boundaries3d.jl:131-236): RK4 of the
    plane wave sin(k (x + y + z) - k sqrt(3) t) on the cubed cube to t = 0.05;
    the Hd-norm error falls with M."""
    import wavetoy_hip as wh
    k = 2 * math.pi / 2
    w = k * math.sqrt(3)
    errs = []
    for M in (1, 2):
        mesh = wh.make_cubed_cube_mesh(5, M, 0.3)
        o = _mesh_oracle(mesh)
        x = o.coords()
        ph = lambda t: np.sin(k * x.sum(axis=1) - w * t)
        pi = lambda t: -w * np.cos(k * x.sum(axis=1) - w * t)
        coef = wo.fo_coef(_flat_bg10(o.n).reshape(-1))
        JA, J = o.curv_metric()
        _, wq, _ = wo.gll(5)
        Hd = np.tile(np.einsum("k,j,i->kji", wq, wq, wq).reshape(-1), mesh.Ne) * J
        dt, T = 0.1 * o.hmin / math.sqrt(3), 0.05
        ns = int(math.ceil(T / dt))
        dt = T / ns
        Phi, Pi = ph(0.0), pi(0.0)
        t = 0.0
        for _ in range(ns):
            def rhs(P, Q, tt):
                g = k * np.cos(k * x.sum(axis=1) - w * tt)
                bc = wo.make_bc3d("dirichlet", gPi=pi(tt), gD=np.stack([g, g, g]))
                return o.fo_rhs_curv(P, Q, coef, 0.0, bc)
            k1 = rhs(Phi, Pi, t)
            k2 = rhs(Phi + dt / 2 * k1[0], Pi + dt / 2 * k1[1], t + dt / 2)
            k3 = rhs(Phi + dt / 2 * k2[0], Pi + dt / 2 * k2[1], t + dt / 2)
            k4 = rhs(Phi + dt * k3[0], Pi + dt * k3[1], t + dt)
            Phi = Phi + dt / 6 * (k1[0] + 2 * k2[0] + 2 * k3[0] + k4[0])
            Pi = Pi + dt / 6 * (k1[1] + 2 * k2[1] + 2 * k3[1] + k4[1])
            t += dt
        errs.append(math.sqrt(np.sum(Hd * (Phi - ph(t)) ** 2) / np.sum(Hd * ph(t) ** 2)))
    assert errs[1] < errs[0] / 4 and errs[1] < 5e-3, errs
