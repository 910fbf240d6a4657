"""Sommerfeld (tag 7) faces on flat meshes through the C ABI against the CPU
oracle: the faces are free for the SIPG terms and carry the Bayliss-Turkel
drag -(u_t + d_n u + u/R) / (h w) of src/wave.jl:444-518.

FP64 bar: max |gpu - oracle| <= 1e-12 * max |oracle|; mixed precision: RK4
H-norm error <= 1e-5.
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


def _pair(N, M, faces, R=math.inf, periodic=False, eps_ko=0.0, bv=(0.0,) * 6, precision="fp64"):
    Mt = M if isinstance(M, tuple) else (M, M, M)
    pt = periodic if isinstance(periodic, tuple) else (periodic,) * 3
    mesh = wh.make_uniform_hex(Mt, 0.0, 1.0, pt)
    elem = wh.make_element(N)
    tau = wh.tau_effective(None, N)
    ws = wh.Workspace(mesh, elem, tau=tau, eps_ko=eps_ko, bdry_values=bv, precision=precision,
                      sommerfeld_faces=faces, sommerfeld_R=R)
    o = wo.Oracle(N, Mt, periodic=pt, tau=tau, eps_ko=eps_ko, bdry_values=bv,
                  sommerfeld_faces=faces, sommerfeld_R=R)
    return ws, o


CASES = [
    # N, M, faces, R, periodic, bv
    (3, 2, (1, 2, 3, 4, 5, 6), math.inf, False, (0.0,) * 6),
    (4, (3, 2, 5), (2, 3, 6), 2.5, False, (0.1, -0.2, 0.3, 0.4, -0.5, 0.6)),
    (5, 3, (3, 4), 1.0, (True, False, True), (0.0,) * 6),
    (6, 4, (1,), math.inf, False, (0.2,) * 6),
    (8, (2, 3, 2), (1, 2, 3, 4, 5, 6), 4.0, False, (0.0,) * 6),
    (10, 2, (5, 6), math.inf, (True, True, False), (0.0,) * 6),
    (4, (16, 2, 3), (1, 2), 3.0, False, (0.0,) * 6),
]
IDS = [f"N{c[0]}-M{c[1]}-f{''.join(map(str, c[2]))}" for c in CASES]


@pytest.mark.parametrize("case", CASES, ids=IDS)
def test_rhs3d_matches_oracle(case):
    """The drop-in RHS (single-pass kernel) with u_t given, and with u_t = NULL
    (the drag then takes u_t = 0)."""
    N, M, faces, R, per, bv = case
    _say("sommerfeld rhs3d", case)
    ws, o = _pair(N, M, faces, R, per, bv=bv)
    rng = np.random.default_rng(300 + N)
    u = rng.standard_normal(o.n)
    ud = rng.standard_normal(o.n)
    out = np.zeros(o.n)
    ws.rhs(out, u, ud, bv)
    assert relerr(out, o.lap(u, ud)) <= TOL64
    ws.rhs(out, u, None, bv)
    assert relerr(out, o.lap(u)) <= TOL64


@pytest.mark.parametrize("case", CASES, ids=IDS)
def test_rhs_full_with_ko_matches_oracle(case):
    """The pencil passes (u_t read on the Sommerfeld axes) with KO."""
    N, M, faces, R, per, bv = case
    _say("sommerfeld rhs_full", case)
    ws, o = _pair(N, M, faces, R, per, eps_ko=0.1, bv=bv)
    rng = np.random.default_rng(400 + N)
    u = rng.standard_normal(o.n)
    v = rng.standard_normal(o.n)
    du_ref, dv_ref = o.rhs(u, v)
    du = np.zeros(o.n); dv = np.zeros(o.n)
    ws.rhs_full(du, dv, u, v)
    assert relerr(du, du_ref) == 0.0
    assert relerr(dv, dv_ref) <= TOL64


@pytest.mark.parametrize("N,M,faces,R", [(4, 3, (1, 2, 3, 4, 5, 6), math.inf), (6, (3, 2, 4), (2, 3, 5), 2.0),
                                         (8, 2, (4, 6), 1.5)])
def test_rk4_steps_and_energy_match_oracle(N, M, faces, R):
    _say("sommerfeld rk4", N, M, faces, R)
    ws, o = _pair(N, M, faces, R, eps_ko=0.1)
    x = o.coords()
    u0 = np.exp(-np.sum((x - 0.6) ** 2, axis=1) / 0.05)
    v0 = 0.3 * np.sin(2 * math.pi * x[:, 1])
    dt = 0.2 * o.hmin
    ns = 7
    uo, vo, Eo = o.rk4(u0, v0, dt, ns, energy=True)
    u = u0.copy(); v = v0.copy()
    E = ws.rk4_steps(u, v, dt, ns, energy_every=1)
    assert relerr(u, uo) <= TOL64
    assert relerr(v, vo) <= TOL64
    assert np.max(np.abs(E - Eo)) <= TOL64 * np.max(np.abs(Eo))
    assert abs(ws.energy(u0, v0) / o.energy(u0, v0) - 1) <= TOL64


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_every_tile_choice_with_sommerfeld_faces(prec):
    """Every pass kernel kind on every axis (the LDS-staged kernel keeps u_t in
    a second LDS slot on Sommerfeld axes): 3 RK4 steps against the oracle."""
    import itertools
    import torch
    _say(f"sommerfeld tile choices {prec}")
    N, M = 6, (4, 3, 5)
    ws, o = _pair(N, M, (1, 2, 4, 5), 3.0, eps_ko=0.1, precision=prec)
    x = o.coords()
    u0 = np.exp(-np.sum((x - 0.45) ** 2, axis=1) / 0.03)
    v0 = np.sin(math.pi * x[:, 1]) * 0.1 + 0.05
    dt = 0.2 * o.hmin
    uo, vo = o.rk4(u0, v0, dt, 3)
    H = o.hphys()
    for kinds in itertools.product(["lds", "shfl", "xchg"], repeat=3):
        for nb in ((1, 1, 1), (3, 3, 6)):
            ws.set_tiles(nb, kinds)
            u = torch.from_numpy(u0).cuda(); v = torch.from_numpy(v0).cuda()
            ws.rk4_steps(u, v, dt, 3)
            uh = u.cpu().numpy()
            if prec == "fp64":
                assert relerr(uh, uo) <= TOL64, (kinds, nb)
                assert relerr(v.cpu().numpy(), vo) <= TOL64, (kinds, nb)
            else:
                d = uh - uo
                assert math.sqrt(np.sum(H * d * d) / np.sum(H * uo * uo)) <= 1e-5, (kinds, nb)


def test_folded_schedule_with_sommerfeld_faces(monkeypatch):
    """The folded RK4 schedule computes the next stage's drag from the new
    stage input in registers."""
    import torch
    _say("sommerfeld folded schedule")
    monkeypatch.setenv("WT2_FOLD", "1")
    ws, o = _pair(6, (5, 3, 4), (1, 2, 3, 4, 5, 6), 2.0, eps_ko=0.1)
    assert ws.info()["launches_per_step"] == 8
    x = o.coords()
    u0 = np.exp(-np.sum((x - 0.5) ** 2, axis=1) / 0.02)
    v0 = 0.1 * np.sin(math.pi * x[:, 0]) + 0.02
    dt = 0.2 * o.hmin
    uo, vo = o.rk4(u0, v0, dt, 6)
    u = torch.from_numpy(u0).cuda(); v = torch.from_numpy(v0).cuda()
    ws.rk4_steps(u, v, dt, 6)
    assert relerr(u.cpu().numpy(), uo) <= TOL64 and relerr(v.cpu().numpy(), vo) <= TOL64


def test_batch_with_sommerfeld_instances_equals_single_runs():
    import torch
    _say("sommerfeld batch")
    specs = [(4, 3, (1, 2, 3, 4, 5, 6), math.inf, "fp64"), (6, 2, (2, 5), 2.0, "mixed"),
             (5, (3, 2, 4), (), math.inf, "fp64")]
    singles, keep = [], []
    b = wh.Batch()
    slots = []
    for q, (N, M, faces, R, prec) in enumerate(specs):
        for rep in range(2):
            ws, o = _pair(N, M, faces, R, eps_ko=0.1, precision=prec)
            x = o.coords()
            u = torch.from_numpy(np.exp(-np.sum((x - 0.4) ** 2, axis=1) / 0.03)).cuda()
            v = torch.from_numpy(0.2 * np.cos(math.pi * x[:, 2])).cuda()
            dt = 0.2 * o.hmin
            if rep == 0:
                E = ws.rk4_steps(u, v, dt, 5 + q, energy_every=1)
                singles.append((u.cpu().numpy(), v.cpu().numpy(), E))
            else:
                keep.append((ws, u, v))
                slots.append(b.add(ws, u, v, dt, 5 + q, energy_every=1))
    fin = 0
    while fin < len(specs):
        fin += b.step(3)
    torch.cuda.synchronize()
    for q, s in enumerate(slots):
        _, E = b.slot_info(s)
        _, u, v = keep[q]
        assert np.array_equal(u.cpu().numpy(), singles[q][0])
        assert np.array_equal(v.cpu().numpy(), singles[q][1])
        assert np.array_equal(E, singles[q][2])
    b.close()


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_outgoing_plane_wave_leaves_at_size(prec):
    """Size-independent property at 32 x 4 x 4 elements of N = 8: a pulse
    f(x - t) leaves through the Sommerfeld face at x = 1 (energy left < 1e-8 of
    the start, and not growing between samples beyond the rounding of the
    arithmetic: 1e-12 of the start in FP64, 1e-8 with the FP32 RHS), where a
    Dirichlet box keeps it."""
    import torch
    _say(f"sommerfeld plane wave {prec}")
    N, M = 8, (32, 4, 4)
    res = {}
    for faces in ((1, 2), ()):
        mesh = wh.make_uniform_hex(M, 0.0, (1.0, 0.125, 0.125), (False, True, True))
        ws = wh.Workspace(mesh, wh.make_element(N), tau=wh.tau_effective(None, N), precision=prec,
                          sommerfeld_faces=faces)
        x = wh.host_coords(mesh, wh.make_element(N))
        s = (x[:, 0] - 0.4) / 0.06
        u0 = np.exp(-s * s)
        v0 = 2 * s / 0.06 * u0
        dt = 0.25 * ws.h_min
        ns = int(math.ceil(0.9 / dt))
        u = torch.from_numpy(u0).cuda(); v = torch.from_numpy(v0).cuda()
        res[faces] = ws.rk4_steps(u, v, 0.9 / ns, ns, energy_every=max(1, ns // 20))
    Es, Ed = res[(1, 2)], res[()]
    assert Es[-1] < 1e-8 * Es[0]
    assert np.all(np.diff(Es) <= (1e-12 if prec == "fp64" else 1e-8) * Es[0])
    assert abs(Ed[-1] / Ed[0] - 1) < 1e-4


# ------------------------------------------------------- warped meshes

CURVED = [
    # N, M, warp A, faces, R
    (3, 2, 0.05, (1, 2, 3, 4, 5, 6), math.inf),
    (4, (2, 3, 2), 0.05, (2, 3, 6), 2.5),
    (6, 2, 0.03, (1, 4), 1.0),
    (8, (2, 2, 3), 0.05, (5, 6), math.inf),
    (10, 2, 0.05, (1, 2, 3, 4, 5, 6), 3.0),
]
CIDS = [f"N{c[0]}-M{c[1]}-f{''.join(map(str, c[3]))}" for c in CURVED]


def _cpair(N, M, A, faces, R, eps_ko=0.0, precision="fp64"):
    Mt = M if isinstance(M, tuple) else (M, M, M)
    mesh = wh.make_warped_uniform_hex(*Mt, 0.0, 1.0, A)
    tau = wh.tau_effective(None, N)
    ws = wh.Workspace(mesh, wh.make_element(N), tau=tau, eps_ko=eps_ko, precision=precision,
                      sommerfeld_faces=faces, sommerfeld_R=R)
    o = wo.Oracle(N, Mt, warp_A=A, tau=tau, eps_ko=eps_ko, sommerfeld_faces=faces, sommerfeld_R=R)
    return ws, o


@pytest.mark.parametrize("case", CURVED, ids=CIDS)
def test_warped_rhs3d_matches_oracle(case):
    """The curved operator's Sommerfeld faces (drag with the physical normal)
    in the drop-in RHS, with u_t given and with u_t = NULL."""
    N, M, A, faces, R = case
    _say("warped sommerfeld rhs3d", case)
    ws, o = _cpair(N, M, A, faces, R)
    rng = np.random.default_rng(500 + N)
    u = rng.standard_normal(o.n)
    ud = rng.standard_normal(o.n)
    out = np.zeros(o.n)
    ws.rhs(out, u, ud, (0.0,) * 6)
    assert relerr(out, o.lap(u, ud)) <= TOL64
    ws.rhs(out, u, None, (0.0,) * 6)
    assert relerr(out, o.lap(u)) <= TOL64


@pytest.mark.parametrize("case", CURVED[1:4], ids=CIDS[1:4])
def test_warped_rk4_steps_and_energy_match_oracle(case):
    """5 RK4 steps with KO 0.1 and the energy at every step (FP64, 1e-12), and
    the mixed path within the H-norm bar."""
    N, M, A, faces, R = case
    _say("warped sommerfeld rk4", case)
    ws, o = _cpair(N, M, A, faces, R, eps_ko=0.1)
    x = o.coords()
    u0 = np.exp(-np.sum((x - 0.55) ** 2, axis=1) / 0.04)
    v0 = 0.2 * np.sin(2 * math.pi * x[:, 0]) * np.cos(math.pi * x[:, 2])
    dt = 0.2 * o.hmin
    uo, vo, Eo = o.rk4(u0, v0, dt, 5, energy=True)
    u = u0.copy(); v = v0.copy()
    E = ws.rk4_steps(u, v, dt, 5, energy_every=1)
    assert relerr(u, uo) <= TOL64 and relerr(v, vo) <= TOL64
    assert np.max(np.abs(E - Eo)) <= TOL64 * np.max(np.abs(Eo))
    assert np.all(np.diff(Eo) <= 1e-12 * Eo[0])
    wm, _ = _cpair(N, M, A, faces, R, eps_ko=0.1, precision="mixed")
    u = u0.copy(); v = v0.copy()
    wm.rk4_steps(u, v, dt, 5)
    H = o.hphys()
    d = u - uo
    assert math.sqrt(np.sum(H * d * d) / np.sum(H * uo * uo)) <= 1e-5


def test_warped_sommerfeld_batch_equals_single_runs_and_background_refused():
    import torch
    _say("warped sommerfeld batch")
    N, M, A, faces, R = CURVED[1]
    runs = []
    for rep in range(2):
        ws, o = _cpair(N, M, A, faces, R, eps_ko=0.1)
        x = o.coords()
        u = torch.from_numpy(np.exp(-np.sum((x - 0.5) ** 2, axis=1) / 0.03)).cuda()
        v = torch.from_numpy(0.1 * np.cos(math.pi * x[:, 1])).cuda()
        runs.append((ws, u, v, 0.2 * o.hmin))
    ws, u, v, dt = runs[0]
    E = ws.rk4_steps(u, v, dt, 6, energy_every=2)
    b = wh.Batch()
    ws2, u2, v2, _ = runs[1]
    s = b.add(ws2, u2, v2, dt, 6, energy_every=2)
    fin = 0
    while fin < 1:
        fin += b.step(4)
    torch.cuda.synchronize()
    _, E2 = b.slot_info(s)
    assert np.array_equal(u2.cpu().numpy(), u.cpu().numpy()) and np.array_equal(v2.cpu().numpy(), v.cpu().numpy())
    assert np.array_equal(E2, E)
    b.close()
    mesh = wh.make_warped_uniform_hex(2, 2, 2, 0.0, 1.0, 0.05)
    x = wh.host_coords(mesh, wh.make_element(4))
    bg = wh.sample_background3d(wh.adm_bump_background(), 0.0, x[:, 0], x[:, 1], x[:, 2])
    with pytest.raises(wh.WT2Error):
        wh.Workspace(mesh, wh.make_element(4), bg=bg, sommerfeld_faces=(1,))
