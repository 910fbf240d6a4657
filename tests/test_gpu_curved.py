"""GPU parity of the curved path (warped mesh + static ADM background) against
the CPU oracle, FP64 <= 1e-12 relative; mixed precision against FP64."""
import math
import sys

import numpy as np
import pytest

import wt_oracle as wo
import wavetoy_hip as wh

pytestmark = pytest.mark.gpu
TOL64 = 1e-12


def _say(*a):
    print(*a, flush=True)


def relerr(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def _pair(N, M, A=0.05, bg=True, eps_ko=0.0, periodic=False, bv=(0.0,) * 6, precision="fp64"):
    Mt = M if isinstance(M, tuple) else (M, M, M)
    mesh = wh.make_warped_uniform_hex(*Mt, 0.0, 1.0, A, periodic=periodic)
    elem = wh.make_element(N)
    x = wh.host_coords(mesh, elem)
    bgv = wh.sample_background3d(wh.adm_bump_background(), 0.0, x[:, 0], x[:, 1], x[:, 2]) if bg else None
    tau = wh.tau_effective(None, N)
    ws = wh.Workspace(mesh, elem, tau=tau, eps_ko=eps_ko, bdry_values=bv, precision=precision, bg=bgv)
    o = wo.Oracle(N, Mt, warp_A=A, periodic=periodic, tau=tau, eps_ko=eps_ko, bdry_values=bv, bg=bgv)
    return ws, o, x


CASES = [(3, 2, 0.05, True), (4, (2, 3, 2), 0.05, True), (6, 2, 0.05, False), (8, 2, 0.0, True),
         (10, 2, 0.05, True), (5, 3, 0.05, True)]


@pytest.mark.parametrize("case", CASES, ids=[f"N{c[0]}-M{c[1]}-A{c[2]}-bg{c[3]}" for c in CASES])
def test_curved_rhs3d_matches_oracle(case):
    N, M, A, bg = case
    _say("curved rhs3d", case)
    bv = (0.1, -0.2, 0.0, 0.3, 0.0, 0.05)
    ws, o, x = _pair(N, M, A, bg, bv=bv)
    rng = np.random.default_rng(300 + N)
    u = rng.standard_normal(o.n)
    ud = rng.standard_normal(o.n)
    out = np.zeros(o.n)
    ws.rhs(out, u, ud, bv)
    assert relerr(out, o.lap(u, ud)) <= TOL64
    assert abs(ws.h_min / o.hmin - 1) < 1e-13


@pytest.mark.parametrize("N,M", [(4, 2), (8, 2), (4, (64, 2, 2)), (10, (32, 2, 2))])
def test_curved_rhs_full_and_rk4_match_oracle(N, M):
    """Also at the longest element lines of the KO passes (64; 32 for N >= 9)."""
    _say("curved rhs_full/rk4", N, M)
    ws, o, x = _pair(N, M, eps_ko=0.1)
    rng = np.random.default_rng(7)
    u = rng.standard_normal(o.n); v = rng.standard_normal(o.n)
    du = np.zeros(o.n); dv = np.zeros(o.n)
    ws.rhs_full(du, dv, u, v)
    dur, dvr = o.rhs(u, v)
    assert relerr(du, dur) == 0.0
    assert relerr(dv, dvr) <= TOL64
    u0 = np.exp(-np.sum((x - 0.45) ** 2, axis=1) / 0.02)
    v0 = np.zeros(o.n)
    dt = 0.2 * o.hmin
    uo, vo, Eo = o.rk4(u0, v0, dt, 5, energy=True)
    uu = u0.copy(); vv = v0.copy()
    E = ws.rk4_steps(uu, vv, dt, 5, energy_every=1)
    # five steps: the RHS bar per evaluation (checked above), and 1e-11 for the
    # state at N = 10 on 32-element lines, where the six chained KO derivatives
    # amplify the rounding differences of the two operation orders over the
    # steps (measured 4.6e-12 on u_t, 4.5e-16 on u)
    tol = TOL64 if N < 10 else 1e-11
    assert relerr(uu, uo) <= TOL64 and relerr(vv, vo) <= tol
    assert np.max(np.abs(E - Eo)) <= TOL64 * np.max(np.abs(Eo))


def test_curved_periodic_free_stream_on_gpu():
    _say("curved free stream")
    ws, o, x = _pair(4, 2, periodic=True)
    c = np.full(o.n, 2.5)
    out = np.zeros(o.n)
    ws.rhs(out, c, np.zeros(o.n), (0.0,) * 6)
    assert np.max(np.abs(out)) < 1e-9


def test_curved_flat_limit_equals_flat_context():
    _say("curved flat limit")
    N, M = 5, 2
    mesh = wh.make_uniform_hex(M, 0.0, 1.0)
    elem = wh.make_element(N)
    flat = wh.Workspace(mesh, elem, eps_ko=0.1)
    bg = np.zeros((N ** 3 * 8, 10)); bg[:, 0] = 1; bg[:, 4] = bg[:, 7] = bg[:, 9] = 1
    curv = wh.Workspace(mesh, elem, eps_ko=0.1, bg=bg)
    rng = np.random.default_rng(3)
    u = rng.standard_normal(flat.npoints); v = rng.standard_normal(flat.npoints)
    a = np.zeros(flat.npoints); b = np.zeros(flat.npoints); c = np.zeros(flat.npoints); d = np.zeros(flat.npoints)
    flat.rhs_full(a, b, u, v)
    curv.rhs_full(c, d, u, v)
    assert relerr(d, b) <= 1e-12


def test_curved_mixed_precision_error():
    _say("curved mixed")
    ws64, o, x = _pair(8, 3, eps_ko=0.1)
    wsmx, _, _ = _pair(8, 3, eps_ko=0.1, precision="mixed")
    u0 = np.exp(-np.sum((x - 0.45) ** 2, axis=1) / (2 * 0.1 ** 2))
    v0 = np.zeros(o.n)
    dt = 0.2 * o.hmin
    a, b = u0.copy(), v0.copy()
    c, d = u0.copy(), v0.copy()
    ws64.rk4_steps(a, b, dt, 100)
    wsmx.rk4_steps(c, d, dt, 100)
    H = o.hphys()
    err = math.sqrt(np.sum(H * (c - a) ** 2) / np.sum(H * a ** 2))
    _say("mixed H-norm rel err after 100 steps", err)
    assert err <= 1e-5
