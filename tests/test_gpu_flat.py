"""GPU parity of the flat (affine) path through the C ABI against the CPU oracle.

FP64 bar: max |gpu - oracle| <= 1e-12 * max |oracle| (the stated RHS
tolerance); mixed precision (FP32 RHS bulk, FP64 state): relative
RHS error <= 2e-5 and RK4 H-norm error <= 1e-5.
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


def _pair(N, M, x1=1.0, periodic=False, eps_ko=0.0, bv=(0.0,) * 6, precision="fp64", tau=None):
    Mt = M if isinstance(M, tuple) else (M, M, M)
    x1t = x1 if isinstance(x1, tuple) else (x1, x1, x1)
    pt = periodic if isinstance(periodic, tuple) else (periodic,) * 3
    mesh = wh.make_uniform_hex(Mt, 0.0, x1t, pt)
    elem = wh.make_element(N)
    tau = wh.tau_effective(None, N) if tau is None else tau
    ws = wh.Workspace(mesh, elem, tau=tau, eps_ko=eps_ko, bdry_values=bv, precision=precision)
    o = wo.Oracle(N, Mt, x1=x1t, periodic=pt, tau=tau, eps_ko=eps_ko, bdry_values=bv)
    return ws, o


CASES = [
    # N, M, x1, periodic, bv
    (3, 2, 1.0, False, (0.0,) * 6),
    (4, (3, 2, 5), (1.0, 0.7, 1.3), False, (0.1, -0.2, 0.3, 0.4, -0.5, 0.6)),
    (5, 3, 1.0, (True, False, True), (0.0, 0.0, 0.2, -0.1, 0.0, 0.0)),
    (6, 4, 1.0, False, (0.0,) * 6),
    (7, (2, 3, 3), 1.0, True, (0.0,) * 6),
    (8, 4, 1.0, False, (0.3,) * 6),
    (9, 2, 1.0, False, (0.0,) * 6),
    (10, 3, 1.0, False, (0.0,) * 6),
    (4, (16, 2, 3), 1.0, False, (0.0,) * 6),   # 16 elements on one axis: 4 lines per wave
    (3, (33, 2, 2), 1.0, False, (0.0,) * 6),   # 33: one line per wave, partial lanes
    # 512 element columns: the drop-in RHS marches up the columns along z
    (8, (32, 16, 3), 1.0, (True, False, False), (0.1, -0.2, 0.3, 0.4, -0.5, 0.6)),
]


@pytest.mark.parametrize("case", CASES, ids=[f"N{c[0]}-M{c[1]}" for c in CASES])
def test_rhs3d_matches_oracle(case):
    N, M, x1, per, bv = case
    _say("rhs3d case", case)
    ws, o = _pair(N, M, x1, per, bv=bv)
    rng = np.random.default_rng(1000 + N)
    u = rng.standard_normal(o.n)
    ud = rng.standard_normal(o.n)
    ref = o.lap(u)
    out = np.zeros(o.n)
    ws.rhs(out, u, ud, bv)
    assert relerr(out, ref) <= TOL64


@pytest.mark.parametrize("case", CASES[:8], ids=[f"N{c[0]}-M{c[1]}" for c in CASES[:8]])
def test_rhs_full_with_ko_matches_oracle(case):
    N, M, x1, per, bv = case
    _say("rhs_full case", case)
    ws, o = _pair(N, M, x1, per, eps_ko=0.1, bv=bv)
    assert abs(ws.mu / o.mu - 1) < 1e-12
    rng = np.random.default_rng(2000 + N)
    u = rng.standard_normal(o.n)
    v = rng.standard_normal(o.n)
    du_ref, dv_ref = o.rhs(u, v)
    du = np.zeros(o.n); dv = np.zeros(o.n)
    ws.rhs_full(du, dv, u, v)
    assert relerr(du, du_ref) == 0.0          # u' = u_t exactly
    assert relerr(dv, dv_ref) <= TOL64


@pytest.mark.parametrize("N,M", [(4, 3), (8, 2), (6, (2, 3, 4)), (10, 2)])
def test_rk4_steps_match_oracle(N, M):
    _say("rk4 case", N, M)
    ws, o = _pair(N, M, eps_ko=0.1, bv=(0.0,) * 6)
    x = o.coords()
    u0 = np.exp(-np.sum((x - 0.45) ** 2, axis=1) / 0.05)
    v0 = 0.3 * np.sin(2 * math.pi * x[:, 1])
    dt = 0.2 * o.hmin
    ns = 7
    uo, vo, Eo = o.rk4(u0, v0, dt, ns, energy=True)
    u = u0.copy(); v = v0.copy()
    E = ws.rk4_steps(u, v, dt, ns, energy_every=1)
    assert relerr(u, uo) <= TOL64
    assert relerr(v, vo) <= TOL64
    assert np.max(np.abs(E - Eo)) <= TOL64 * np.max(np.abs(Eo))


def test_device_pointers_graph_and_determinism():
    import torch
    _say("device ptr / graph / determinism")
    ws, o = _pair(8, 4, eps_ko=0.1)
    x = o.coords()
    u0 = np.exp(-np.sum((x - 0.5) ** 2, axis=1) / 0.02)
    v0 = np.zeros(o.n)
    dt = 0.2 * o.hmin
    runs = []
    for graph in (True, True, False):
        u = torch.from_numpy(u0).cuda()
        v = torch.from_numpy(v0).cuda()
        ws.rk4_steps(u, v, dt, 9, graph=graph)
        runs.append((u.cpu().numpy(), v.cpu().numpy()))
    uh = u0.copy(); vh = v0.copy()
    ws.rk4_steps(uh, vh, dt, 9)
    for r in runs[1:] + [(uh, vh)]:
        assert np.array_equal(r[0], runs[0][0]) and np.array_equal(r[1], runs[0][1])
    uo, vo = o.rk4(u0, v0, dt, 9)
    assert relerr(runs[0][0], uo) <= TOL64


def test_energy_matches_oracle():
    _say("energy")
    for N, M in ((4, 3), (8, 2)):
        ws, o = _pair(N, M)
        rng = np.random.default_rng(5)
        u = rng.standard_normal(o.n); v = rng.standard_normal(o.n)
        assert abs(ws.energy(u, v) / o.energy(u, v) - 1) <= TOL64


def test_rhs_wave3d_signature_and_tau_mismatch():
    _say("rhs_wave3d signature")
    N, M = 4, 2
    mesh = wh.make_uniform_hex(M, 0.0, 1.0)
    elem = wh.make_element(N)
    geom = wh.make_geometry(mesh, elem)
    ops = wh.make_operators(elem)
    work = wh.make_workspace(geom, tau=13.5)
    p = wh.Params3d(A=1.0, k=(math.pi,) * 3, omega=math.sqrt(3) * math.pi, tau=13.5)
    x = wh.host_coords(mesh, elem)
    u, ud = wh.eigenmode_cartesian(x, 0.0)
    udd = np.zeros_like(u)
    wh.rhs_wave3d(udd, u, ud, p, geom=geom, ops=ops, work=work)
    assert relerr(udd, wo.Oracle(N, M).lap(u)) <= TOL64
    with pytest.raises(ValueError):
        wh.rhs_wave3d(udd, u, ud, wh.Params3d(tau=20.0), geom=geom, ops=ops, work=work)


# ------------------------------------------------------------ mixed precision

def test_mixed_rhs_close_to_fp64():
    _say("mixed rhs")
    ws, o = _pair(8, 3, eps_ko=0.1, precision="mixed")
    x = o.coords()
    u = np.exp(-np.sum((x - 0.5) ** 2, axis=1) / 0.05)
    v = np.sin(math.pi * x[:, 0]) * np.sin(math.pi * x[:, 1])
    du_ref, dv_ref = o.rhs(u, v)
    du = np.zeros(o.n); dv = np.zeros(o.n)
    ws.rhs_full(du, dv, u, v)
    assert relerr(dv, dv_ref) <= 2e-5
    assert relerr(du, du_ref) == 0.0


@pytest.mark.parametrize("M", [(32, 16, 3), (4, 3, 5)])
def test_mixed_rhs3d_close_to_oracle(M):
    """The drop-in RHS in FP32 (at N = 8 on 512 element columns the
    column-marching kernel, else three wavefronts per element) against the
    FP64 oracle: relative error <= 2e-5."""
    _say(f"mixed rhs3d M={M}")
    bv = (0.1, -0.2, 0.3, 0.4, -0.5, 0.6)
    ws, o = _pair(8, M, bv=bv, precision="mixed")
    rng = np.random.default_rng(77)
    u = rng.standard_normal(o.n)
    ud = rng.standard_normal(o.n)
    out = np.zeros(o.n)
    ws.rhs(out, u, ud, bv)
    assert relerr(out, o.lap(u)) <= 2e-5


def test_mixed_rk4_hnorm_error():
    _say("mixed rk4")
    ws, o = _pair(8, 3, eps_ko=0.1, precision="mixed")
    x = o.coords()
    u0 = np.sin(math.pi * x[:, 0]) * np.sin(math.pi * x[:, 1]) * np.sin(math.pi * x[:, 2])
    v0 = np.zeros(o.n)
    dt = 0.2 * o.hmin
    ns = 60
    uo, vo = o.rk4(u0, v0, dt, ns)
    u = u0.copy(); v = v0.copy()
    ws.rk4_steps(u, v, dt, ns)
    H = o.hphys()
    err = math.sqrt(np.sum((u - uo) ** 2 * H)) / math.sqrt(np.sum(uo ** 2 * H))
    assert err <= 1e-5


@pytest.mark.parametrize("warp", [0.0, 0.05])
def test_mixed_state_mirror_follows_the_state(warp):
    """Mixed stages 1-3 read an FP32 mirror of the state that each call makes
    afresh: overwriting the state in place between two calls (same device
    pointers, so the captured step graph is replayed) must give the oracle's
    evolution of the new state; energy sampling between steps leaves the
    mirror alone (bitwise the same state as without sampling)."""
    import torch
    _say(f"mixed state mirror warp={warp}")
    N, M = 6, 3
    mesh = wh.make_warped_uniform_hex(M, M, M, 0.0, 1.0, warp) if warp else wh.make_uniform_hex(M, 0.0, 1.0)
    elem = wh.make_element(N)
    tau = wh.tau_effective(None, N)
    ws = wh.Workspace(mesh, elem, tau=tau, eps_ko=0.1, precision="mixed")
    o = wo.Oracle(N, M, tau=tau, eps_ko=0.1, warp_A=warp)
    x = o.coords()
    uA = np.exp(-np.sum((x - 0.4) ** 2, axis=1) / 0.05)
    uB = np.sin(math.pi * x[:, 0]) * np.sin(math.pi * x[:, 1]) * np.sin(math.pi * x[:, 2])
    vB = 0.3 * np.cos(math.pi * x[:, 2])
    dt = 0.2 * ws.h_min
    ns = 12
    u = torch.from_numpy(uA.copy()).cuda()
    v = torch.zeros_like(u)
    ws.rk4_steps(u, v, dt, ns)
    u.copy_(torch.from_numpy(uB))
    v.copy_(torch.from_numpy(vB))
    ws.rk4_steps(u, v, dt, ns)
    uo, vo = o.rk4(uB, vB, dt, ns)
    H = o.hphys()
    err = math.sqrt(np.sum((u.cpu().numpy() - uo) ** 2 * H)) / math.sqrt(np.sum(uo ** 2 * H))
    assert err <= 1e-5, err
    # the same evolution of B with energies sampled every 3 steps
    u2 = torch.from_numpy(uB.copy()).cuda()
    v2 = torch.from_numpy(vB.copy()).cuda()
    ws.rk4_steps(u2, v2, dt, ns, energy_every=3)
    u3 = torch.from_numpy(uB.copy()).cuda()
    v3 = torch.from_numpy(vB.copy()).cuda()
    ws.rk4_steps(u3, v3, dt, ns)
    assert torch.equal(u2, u3) and torch.equal(v2, v3)


# -------------------------------------------- full-size properties (16.8M points)

def test_full_size_linearity_symmetry_energy():
    """At the plain FP64 GPU configuration's size (32^3 elements x 8^3 nodes) the
    oracle is too slow to run whole, so check what is size independent: the
    RHS is linear, H L is symmetric, and the energy with KO on does not grow."""
    import torch
    _say("full size 32^3 x 8^3")
    N, M = 8, 32
    mesh = wh.make_uniform_hex(M, 0.0, 1.0)
    elem = wh.make_element(N)
    ws = wh.Workspace(mesh, elem, tau=73.5, eps_ko=0.1)
    n = ws.npoints
    g = torch.Generator(device="cuda").manual_seed(11)
    a = torch.randn(n, dtype=torch.float64, device="cuda", generator=g)
    b = torch.randn(n, dtype=torch.float64, device="cuda", generator=g)
    zero = torch.zeros(n, dtype=torch.float64, device="cuda")
    La = torch.empty_like(a); Lb = torch.empty_like(a); Lab = torch.empty_like(a)
    ws.rhs(La, a, zero, (0.0,) * 6)
    ws.rhs(Lb, b, zero, (0.0,) * 6)
    ws.rhs(Lab, 2.0 * a - 3.0 * b, zero, (0.0,) * 6)
    lin = (Lab - (2.0 * La - 3.0 * Lb)).abs().max().item() / Lab.abs().max().item()
    assert lin < 1e-13
    H = torch.from_numpy(wh.Geometry(mesh, elem).Hphys).cuda()
    s1 = torch.sum(H * a * Lb).item()
    s2 = torch.sum(H * b * La).item()
    assert abs(s1 - s2) <= 1e-10 * math.sqrt(torch.sum(H * a * La).abs().item() * torch.sum(H * b * Lb).abs().item())
    # energy non-increasing over a few steps of a resolved pulse
    x = torch.from_numpy(wh.host_coords(mesh, elem)).cuda()
    c = torch.tensor([0.42, 0.55, 0.48], dtype=torch.float64, device="cuda")
    u = torch.exp(-((x - c) ** 2).sum(dim=1) / 0.01).contiguous()
    v = torch.zeros_like(u)
    dt = 0.2 * ws.h_min
    E = ws.rk4_steps(u, v, dt, 6, energy_every=1)
    assert np.all(np.diff(E) <= 1e-13 * E[0])


@pytest.mark.parametrize("N,M", [(10, 9), (10, 12), (8, 16)])
def test_mixed_large_lds_tiles(N, M):
    """Mixed-precision z/y passes whose LDS tiles exceed 64 KB (the dynamic-LDS
    limit is raised per kernel): 4 RK4 steps against the FP64 path."""
    import torch
    _say(f"mixed large-LDS N={N} M={M}")
    mesh = wh.make_uniform_hex(M, 0.0, 1.0)
    elem = wh.make_element(N)
    x = wh.host_coords(mesh, elem)
    u0, v0 = wh.gaussian_pulse(x, (0.45, 0.5, 0.55), width=0.1)
    res = {}
    for prec in ("fp64", "mixed"):
        ws = wh.Workspace(mesh, elem, tau=wh.tau_effective(20.0, N), eps_ko=0.1, precision=prec)
        u = torch.from_numpy(u0).cuda(); v = torch.from_numpy(v0).cuda()
        ws.rk4_steps(u, v, 0.2 * ws.h_min, 4)
        res[prec] = u.cpu().numpy()
        ws.close()
    H = wh.Geometry(mesh, elem).Hphys
    d = res["mixed"] - res["fp64"]
    assert math.sqrt(np.sum(H * d * d) / np.sum(H * res["fp64"] ** 2)) <= 1e-5


def test_autotuned_tiles_give_the_same_evolution():
    import torch
    _say("autotune")
    N, M = 6, (5, 3, 4)
    ws, o = _pair(N, M, eps_ko=0.1)
    x = o.coords()
    u0 = np.exp(-np.sum((x - 0.5) ** 2, axis=1) / 0.02)
    v0 = np.zeros(o.n)
    dt = 0.2 * o.hmin
    u = torch.from_numpy(u0).cuda(); v = torch.from_numpy(v0).cuda()
    ws.rk4_steps(u, v, dt, 5)
    u_def = u.cpu().numpy()
    u = torch.from_numpy(u0).cuda(); v = torch.from_numpy(v0).cuda()
    ws.autotune(u, v, dt, force=True)
    t = ws.tiles()
    assert len(t["planes_per_tile"]) == 3 and all(N % nb == 0 for nb in t["planes_per_tile"])
    ws.rk4_steps(u, v, dt, 5)
    uo, _ = o.rk4(u0, v0, dt, 5)
    assert relerr(u.cpu().numpy(), uo) <= TOL64
    assert relerr(u_def, uo) <= TOL64


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_every_tile_choice_gives_the_same_evolution(prec):
    """Every pass kernel kind (LDS-staged, LDS-free with shuffles, LDS-free with
    the exchange through LDS) on every axis, including the x pass that carries
    the RK4 update, with two tile sizes: 3 RK4 steps with KO against the oracle
    (FP64 <= 1e-12; mixed H-norm <= 1e-5)."""
    import itertools
    import torch
    _say(f"tile choices {prec}")
    N, M = 6, (4, 3, 5)
    ws, o = _pair(N, M, eps_ko=0.1, precision=prec)
    x = o.coords()
    u0 = np.exp(-np.sum((x - 0.45) ** 2, axis=1) / 0.03)
    v0 = np.sin(math.pi * x[:, 1]) * 0.1
    dt = 0.2 * o.hmin
    uo, vo = o.rk4(u0, v0, dt, 3)
    H = o.hphys()
    for kinds in itertools.product(["lds", "shfl", "xchg"], repeat=3):
        for nb in ((1, 1, 1), (3, 2, 6)):
            lds = kinds
            ws.set_tiles(nb, kinds)
            u = torch.from_numpy(u0).cuda(); v = torch.from_numpy(v0).cuda()
            ws.rk4_steps(u, v, dt, 3)
            uh = u.cpu().numpy()
            if prec == "fp64":
                assert relerr(uh, uo) <= TOL64, (lds, nb)
            else:
                d = uh - uo
                assert math.sqrt(np.sum(H * d * d) / np.sum(H * uo * uo)) <= 1e-5, (lds, nb)


@pytest.mark.parametrize("kind", ["shfl", "xchg"])
def test_exchange_kernels_drop_in_and_rhs_outputs(kind):
    """The drop-in RHS output and the (u', u_t') output epilogues through the
    exchange kernels (x pass kind set explicitly), against the oracle."""
    _say(f"rhs outputs {kind}")
    N, M = 5, (6, 2, 3)
    ws, o = _pair(N, M, eps_ko=0.1, periodic=(True, False, False))
    ws.set_tiles((1, 1, 1), (kind, "xchg", "xchg"))
    x = o.coords()
    u = np.cos(2 * math.pi * x[:, 0]) * np.sin(math.pi * x[:, 1]) + x[:, 2] ** 2
    v = np.sin(2 * math.pi * x[:, 0]) * x[:, 1]
    du_ref, dv_ref = o.rhs(u, v)
    du = np.zeros(o.n); dv = np.zeros(o.n)
    ws.rhs_full(du, dv, u, v)
    assert relerr(du, du_ref) <= TOL64 and relerr(dv, dv_ref) <= TOL64


@pytest.mark.parametrize("prec,fold", [("fp64", 1), ("mixed", 1), ("fp64", 0), ("mixed", 0)])
def test_folded_and_unfolded_schedules(prec, fold, monkeypatch):
    """The folded RK4 schedule (the default, WT2_FOLD=1: 8 passes per step, the
    first pass of each stage computed by the previous stage's RK pass) and the
    12-pass one (WT2_FOLD=0): 6 steps with KO against the oracle, and a batch of
    it equal to the single run."""
    import torch
    _say(f"schedule {prec} fold={fold}")
    monkeypatch.setenv("WT2_FOLD", str(fold))
    N, M = 6, (5, 3, 4)
    ws, o = _pair(N, M, eps_ko=0.1, precision=prec)
    assert ws.info()["launches_per_step"] == (8 if fold else 12)
    x = o.coords()
    u0 = np.exp(-np.sum((x - 0.5) ** 2, axis=1) / 0.02)
    v0 = 0.1 * np.sin(math.pi * x[:, 0])
    dt = 0.2 * o.hmin
    uo, vo = o.rk4(u0, v0, dt, 6)
    u = torch.from_numpy(u0).cuda(); v = torch.from_numpy(v0).cuda()
    E = ws.rk4_steps(u, v, dt, 6, energy_every=2)
    uh = u.cpu().numpy()
    if prec == "fp64":
        assert relerr(uh, uo) <= TOL64 and relerr(v.cpu().numpy(), vo) <= TOL64
    else:
        H = o.hphys()
        d = uh - uo
        assert math.sqrt(np.sum(H * d * d) / np.sum(H * uo * uo)) <= 1e-5
    # the same context in a batch: bitwise the single run
    u2 = torch.from_numpy(u0).cuda(); v2 = torch.from_numpy(v0).cuda()
    b = wh.Batch()
    slot = b.add(ws, u2, v2, dt, 6, energy_every=2)
    for _ in range(4):
        if b.step(4):
            break
    done, E2 = b.slot_info(slot)
    b.remove(slot)
    b.close()
    assert done == 6 and np.array_equal(u2.cpu().numpy(), uh) and np.array_equal(E2, E)


EDGE_CASES = [
    # N, M, periodic: the longest element lines the pass kernels take (64 per axis,
    # 32 for N >= 9), and a single element along y or z (not x: mu comes from
    # the x lines, and on a single element G is nilpotent, so mu -> 0 and the
    # KO scale mu^-5 is unbounded, in the oracle as in the reference's formula)
    (3, (64, 2, 2), False),
    (4, (2, 64, 2), (False, True, False)),
    (8, (2, 2, 64), False),
    (9, (32, 2, 2), False),
    (10, (2, 32, 2), False),
    (5, (3, 1, 2), False),
    (6, (3, 1, 1), (True, False, False)),
]


@pytest.mark.parametrize("N,M,per", EDGE_CASES, ids=[f"N{c[0]}-M{'x'.join(map(str, c[1]))}" for c in EDGE_CASES])
def test_longest_and_shortest_element_lines(N, M, per):
    """Drop-in RHS and 3 RK4 steps with KO at the per-axis line-length limits and
    with one element along an axis, FP64 against the oracle (1e-12) and the
    mixed path within its H-norm bar (1e-5).  The box's sides are proportional
    to M (cubic elements): mu, and with it the KO scale and the dt cap, comes
    from the x axis alone (src/wave3d_curved.jl:158-178), so strongly
    anisotropic elements would make KO along a finer axis unstable at that dt,
    in the oracle as on the device."""
    _say("edge case", N, M, per)
    x1 = tuple(m / max(M) for m in M)
    ws, o = _pair(N, M, x1=x1, periodic=per, eps_ko=0.1)
    x = o.coords()
    c = 0.4 * np.asarray(x1)
    u0 = np.exp(-np.sum(((x - c) / np.asarray(x1)) ** 2, axis=1) / 0.05)
    v0 = 0.2 * np.cos(2 * math.pi * x[:, 2] / x1[2]) * np.sin(math.pi * x[:, 0] / x1[0])
    out = np.zeros(o.n)
    ws.rhs(out, u0, None, (0.0,) * 6)
    assert relerr(out, o.lap(u0)) <= TOL64
    dt = 0.2 * o.hmin
    uo, vo = o.rk4(u0, v0, dt, 3)
    u = u0.copy(); v = v0.copy()
    ws.rk4_steps(u, v, dt, 3)
    assert relerr(u, uo) <= TOL64 and relerr(v, vo) <= TOL64
    wm, _ = _pair(N, M, x1=x1, periodic=per, eps_ko=0.1, precision="mixed")
    u = u0.copy(); v = v0.copy()
    wm.rk4_steps(u, v, dt, 3)
    H = o.hphys()
    d = u - uo
    assert math.sqrt(np.sum(H * d * d) / np.sum(H * uo * uo)) <= 1e-5
