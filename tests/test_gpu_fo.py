"""The first-order (Phi, Pi) ADM path (wave3d_curved_rhs!, affine meshes)
through the C ABI against the CPU oracle.

FP64 bar: max |gpu - oracle| <= 1e-12 * max |oracle| for the RHS, the
coefficient fields and the energy.
"""
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
    sys.stdout.flush()


def relerr(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def _bg(x, seed):
    """A smooth static background with lapse, shift and off-diagonal metric."""
    rng = np.random.default_rng(seed)
    ph = rng.uniform(0, 2 * math.pi, 8)
    X, Y, Z = (2 * math.pi * x[:, d] for d in range(3))
    n = x.shape[0]
    bg = np.zeros((n, 10))
    bg[:, 0] = 1.0 + 0.2 * np.sin(X + ph[0]) * np.cos(Y + ph[1])
    bg[:, 1] = 0.1 * np.cos(Y + ph[2]); bg[:, 2] = 0.05 * np.sin(Z + ph[3]); bg[:, 3] = -0.08 * np.cos(X + Z + ph[4])
    bg[:, 4] = 1.2 + 0.2 * np.sin(Z + ph[5])
    bg[:, 5] = 0.1 * np.cos(X)
    bg[:, 6] = 0.05 * np.sin(Y + Z)
    bg[:, 7] = 1.0 + 0.15 * np.cos(X - Y + ph[6])
    bg[:, 8] = -0.07 * np.sin(X)
    bg[:, 9] = 0.9 + 0.1 * np.sin(Y + ph[7])
    return bg


def _pair(N, M, periodic=False):
    Mt = M if isinstance(M, tuple) else (M, M, M)
    pt = periodic if isinstance(periodic, tuple) else (periodic,) * 3
    mesh = wh.make_uniform_hex(Mt, 0.0, 1.0, pt)
    ws = wh.Workspace(mesh, wh.make_element(N))
    o = wo.Oracle(N, Mt, periodic=pt)
    return ws, o


CASES = [
    # N, M, periodic, eps, bc kinds (None: no boundary pass)
    (3, 2, True, 0.0, None),
    (4, (3, 2, 4), False, 0.1, ("sommerfeld",) * 6),
    (5, 3, (True, False, False), 0.1, (1, 1, "dirichlet", "sommerfeld", "full_dirichlet", "excision")),
    (6, (2, 4, 3), False, 0.0, ("full_dirichlet", "dirichlet", "sommerfeld", "excision", "dirichlet", "full_dirichlet")),
    (8, 2, False, 0.1, ("sommerfeld",) * 6),
    (10, (2, 2, 3), (False, True, False), 0.1, ("dirichlet", "sommerfeld", 1, 1, "full_dirichlet", "sommerfeld")),
    (4, (40, 2, 2), False, 0.1, ("sommerfeld",) * 6),     # 40 elements on one line
    (3, (2, 64, 2), (False, True, False), 0.1, None),      # 64: one line per 8-lane group
]
IDS = [f"N{c[0]}-M{c[1]}" for c in CASES]


@pytest.mark.parametrize("case", CASES, ids=IDS)
def test_fo_rhs_matches_oracle(case):
    import torch
    N, M, per, eps, kinds = case
    _say("fo rhs case", case)
    ws, o = _pair(N, M, per)
    x = o.coords()
    bg = _bg(x, N)
    coef_ref = wo.fo_coef(bg.reshape(-1))
    coef = ws.fo_coef(torch.from_numpy(bg.reshape(-1).copy()).cuda())
    assert relerr(coef.cpu().numpy(), coef_ref.reshape(-1)) <= TOL64
    rng = np.random.default_rng(100 + N)
    Phi = rng.standard_normal(o.n); Pi = rng.standard_normal(o.n)
    gP = rng.standard_normal(o.n); gQ = rng.standard_normal(o.n)
    bc_ref = None if kinds is None else wo.make_bc3d(kinds, sigma=0.8, gPhi=gP, gPi=gQ)
    dP_ref, dQ_ref = o.fo_rhs(Phi, Pi, coef_ref, eps, bc_ref)
    t = lambda a: torch.from_numpy(a).cuda()
    dP = torch.empty(o.n, dtype=torch.float64, device="cuda"); dQ = torch.empty_like(dP)
    bc = None if kinds is None else wh.make_bc3d(kinds, sigma=0.8, gPhi=t(gP), gPi=t(gQ))
    ws.fo_rhs(dP, dQ, t(Phi), t(Pi), coef, eps, bc)
    assert relerr(dP.cpu().numpy(), dP_ref) <= TOL64
    assert relerr(dQ.cpu().numpy(), dQ_ref) <= TOL64
    # host arrays give the same result bit for bit
    hP = np.zeros(o.n); hQ = np.zeros(o.n)
    bch = None if kinds is None else wh.make_bc3d(kinds, sigma=0.8, gPhi=gP, gPi=gQ)
    ws.fo_rhs(hP, hQ, Phi, Pi, coef.cpu().numpy(), eps, bch)
    assert np.array_equal(hP, dP.cpu().numpy()) and np.array_equal(hQ, dQ.cpu().numpy())


@pytest.mark.parametrize("N,M,per", [(4, 3, True), (6, (2, 3, 2), False), (8, 2, (True, False, True))])
def test_fo_energy_matches_oracle(N, M, per):
    _say("fo energy", N, M, per)
    ws, o = _pair(N, M, per)
    x = o.coords()
    coef = wo.fo_coef(_bg(x, 7).reshape(-1)).reshape(-1)
    rng = np.random.default_rng(3)
    Phi = rng.standard_normal(o.n); Pi = rng.standard_normal(o.n)
    assert abs(ws.fo_energy(Phi, Pi, coef) / o.fo_energy(Phi, Pi, coef) - 1) <= TOL64


def test_fo_rk4_matches_oracle():
    """Five RK4 steps of the first-order system (KO, Sommerfeld faces) driven
    from the host in the stage order of test/wave3d_tests.jl:32-44."""
    import torch
    _say("fo rk4")
    N, M = 6, (3, 2, 4)
    ws, o = _pair(N, M)
    x = o.coords()
    coef_h = wo.fo_coef(_bg(x, 11).reshape(-1))
    coef = torch.from_numpy(coef_h.reshape(-1).copy()).cuda()
    Phi0 = np.exp(-np.sum((x - 0.5) ** 2, axis=1) / 0.05); Pi0 = np.zeros(o.n)
    dt = 0.1 * o.hmin; ns = 5
    bc_ref = wo.make_bc3d(["sommerfeld"] * 6)
    Pr, Qr = o.fo_rk4(Phi0, Pi0, coef_h, dt, ns, eps_ko=0.1, bc=bc_ref)
    bc = wh.make_bc3d(["sommerfeld"] * 6)
    P = torch.from_numpy(Phi0).cuda(); Q = torch.from_numpy(Pi0).cuda()
    k = [torch.empty_like(P) for _ in range(8)]

    def rhs(a, b, p, q):
        ws.fo_rhs(a, b, p, q, coef, 0.1, bc)

    for _ in range(ns):
        rhs(k[0], k[1], P, Q)
        rhs(k[2], k[3], P + dt / 2 * k[0], Q + dt / 2 * k[1])
        rhs(k[4], k[5], P + dt / 2 * k[2], Q + dt / 2 * k[3])
        rhs(k[6], k[7], P + dt * k[4], Q + dt * k[5])
        P = P + dt / 6 * (k[0] + 2 * k[2] + 2 * k[4] + k[6])
        Q = Q + dt / 6 * (k[1] + 2 * k[3] + 2 * k[5] + k[7])
    assert relerr(P.cpu().numpy(), Pr) <= TOL64 and relerr(Q.cpu().numpy(), Qr) <= TOL64


def test_fo_energy_conserved_at_size():
    """Size-independent property at 16^3 elements of N = 8 (2.1M points): with
    alpha = 1, beta = 0 and a static metric on a periodic mesh the energy rate
    <dE, RHS> vanishes to rounding (central difference of the quadratic E
    along the RHS), with a varying lapse it does not."""
    import torch
    _say("fo energy rate at size")
    N, M = 8, 16
    ws = wh.Workspace(wh.make_uniform_hex(M, 0.0, 1.0, True), wh.make_element(N))
    x = torch.from_numpy(ws.coords())
    rates = []
    for lapse in (False, True):
        bg = _bg(x.numpy(), 4)
        if not lapse:
            bg[:, 0] = 1.0
        bg[:, 1:4] = 0.0
        coef = ws.fo_coef(torch.from_numpy(bg.reshape(-1).copy()).cuda())
        X = x.cuda()
        P = (torch.sin(2 * math.pi * X[:, 0]) * torch.cos(2 * math.pi * X[:, 1]) + torch.sin(4 * math.pi * X[:, 2])).contiguous()
        Q = torch.cos(2 * math.pi * (X[:, 0] + X[:, 2])).contiguous()
        dP = torch.empty_like(P); dQ = torch.empty_like(Q)
        ws.fo_rhs(dP, dQ, P, Q, coef, 0.0)
        d = 1e-3
        Ep = ws.fo_energy(P + d * dP, Q + d * dQ, coef)
        Em = ws.fo_energy(P - d * dP, Q - d * dQ, coef)
        scale = ws.fo_energy(P, Q, coef) * math.sqrt(float(((dP ** 2 + dQ ** 2).sum() / (P ** 2 + Q ** 2).sum()).item()))
        rates.append(abs((Ep - Em) / (2 * d)) / scale)
    assert rates[0] < 1e-10
    assert rates[1] > 1e4 * rates[0]


def test_fo_rejects_curved_contexts_and_bad_boundary_kinds():
    _say("fo errors")
    mesh = wh.make_warped_uniform_hex(2, 2, 2, 0.0, 1.0, 0.05)
    ws = wh.Workspace(mesh, wh.make_element(4))
    n = ws.npoints
    z = np.zeros(n); c = np.ones(11 * n)
    with pytest.raises(wh.WT2Error):
        ws.fo_rhs(np.zeros(n), np.zeros(n), z, z, c, 0.1)
    ws2, _ = _pair(4, 2)
    with pytest.raises(wh.WT2Error):
        ws2.fo_rhs(np.zeros(n), np.zeros(n), z, z, c, 0.1, wh.make_bc3d([0, 1, 1, 1, 1, 1]))
    with pytest.raises(wh.WT2Error):
        ws2.fo_rhs(np.zeros(n), np.zeros(n), z, z, c, -1.0)


def test_fused_and_separate_pointwise_step_agree_bitwise(monkeypatch):
    """The z pass doing the pointwise flux step (default) and the same step as a
    kernel of its own (WT2_FO_FUSE=0) give the same bits, with every boundary
    kind and KO."""
    import torch
    _say("fo fused vs separate")
    N, M, per, eps, kinds = CASES[5]
    ws, o = _pair(N, M, per)
    x = o.coords()
    coef = ws.fo_coef(torch.from_numpy(_bg(x, 3).reshape(-1).copy()).cuda())
    rng = np.random.default_rng(9)
    t = lambda a: torch.from_numpy(a).cuda()
    Phi = t(rng.standard_normal(o.n)); Pi = t(rng.standard_normal(o.n))
    bc = wh.make_bc3d(kinds, sigma=0.8, gPhi=t(rng.standard_normal(o.n)), gPi=t(rng.standard_normal(o.n)))
    out = []
    for fuse in ("1", "0"):
        monkeypatch.setenv("WT2_FO_FUSE", fuse)
        dP = torch.empty_like(Phi); dQ = torch.empty_like(Phi)
        ws.fo_rhs(dP, dQ, Phi, Pi, coef, eps, bc)
        out.append((dP.cpu().numpy(), dQ.cpu().numpy()))
    assert np.array_equal(out[0][0], out[1][0]) and np.array_equal(out[0][1], out[1][1])
