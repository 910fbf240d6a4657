"""Seeded random configurations through the C ABI against the CPU oracle: N,
per-axis M (up to the line-length limits), periodicity, warp, Sommerfeld face
subsets and R, Dirichlet values, KO, FP64 and mixed -- drop-in RHS, full RHS
and RK4 steps with energies.  FP64 bar 1e-12 (RHS) and 1e-11 (states after
steps); mixed: H-norm 1e-5."""
import math
import sys

import numpy as np
import pytest

import wt_oracle as wo
import wavetoy_hip as wh

pytestmark = pytest.mark.gpu


def _case(seed):
    rng = np.random.default_rng(seed)
    N = int(rng.integers(3, 11))
    warp = bool(rng.integers(0, 2))
    lim = 32 if N >= 9 else 64
    M = [int(rng.integers(1, 5)) for _ in range(3)]
    M[int(rng.integers(0, 3))] = int(rng.integers(2, min(lim, 12) + 1))
    M[0] = max(M[0], 2)                       # mu comes from the x lines (needs two elements)
    per = tuple(bool(rng.integers(0, 4) == 0) and M[a] >= 2 for a in range(3))
    if warp:
        per = (False, False, False)
        # the warped cube is the unit cube: keep x the finest axis (mu, the KO
        # scale and the dt cap come from the x lines)
        k = int(np.argmax(M))
        M[0], M[k] = M[k], M[0]
    faces = tuple(f for f in range(1, 7) if not per[(f - 1) // 2] and rng.integers(0, 3) == 0)
    R = math.inf if rng.integers(0, 2) else float(rng.uniform(0.5, 4.0))
    bv = tuple(float(x) for x in rng.uniform(-0.5, 0.5, 6)) if not warp else (0.0,) * 6
    eps = float(rng.choice([0.0, 0.1]))
    return dict(N=N, M=tuple(M), per=per, warp=0.04 if warp else 0.0, faces=faces, R=R, bv=bv, eps=eps, seed=seed)


CASES = [_case(s) for s in range(12)]


@pytest.mark.parametrize("c", CASES, ids=[f"seed{c['seed']}-N{c['N']}-M{'x'.join(map(str, c['M']))}" for c in CASES])
def test_random_configuration_matches_oracle(c):
    print("random case", c, flush=True)
    sys.stdout.flush()
    N, M = c["N"], c["M"]
    x1 = tuple(m / max(M) for m in M)              # cubic elements
    tau = wh.tau_effective(None, N)
    if c["warp"]:
        mesh = wh.make_warped_uniform_hex(*M, 0.0, 1.0, c["warp"])
        okw = dict(warp_A=c["warp"])
    else:
        mesh = wh.make_uniform_hex(M, 0.0, x1, c["per"])
        okw = dict(x1=x1, periodic=c["per"])
    mk = lambda prec: wh.Workspace(mesh, wh.make_element(N), tau=tau, eps_ko=c["eps"], bdry_values=c["bv"],
                                   precision=prec, sommerfeld_faces=c["faces"], sommerfeld_R=c["R"])
    ws = mk("fp64")
    o = wo.Oracle(N, M, tau=tau, eps_ko=c["eps"], bdry_values=c["bv"], sommerfeld_faces=c["faces"],
                  sommerfeld_R=c["R"], **okw)
    rng = np.random.default_rng(1000 + c["seed"])
    u = rng.standard_normal(o.n); v = rng.standard_normal(o.n)
    out = np.zeros(o.n)
    ws.rhs(out, u, v, c["bv"])
    ref = o.lap(u, v)
    assert np.max(np.abs(out - ref)) <= 1e-12 * np.max(np.abs(ref))
    du = np.zeros(o.n); dv = np.zeros(o.n)
    ws.rhs_full(du, dv, u, v)
    dur, dvr = o.rhs(u, v)
    assert np.array_equal(du, dur)
    assert np.max(np.abs(dv - dvr)) <= 1e-12 * np.max(np.abs(dvr))
    x = o.coords()
    u0 = np.exp(-np.sum((x - 0.45 * np.max(x, axis=0)) ** 2, axis=1) / 0.05)
    v0 = 0.1 * np.sin(3.0 * x[:, 1]) + 0.05
    dt = 0.2 * o.hmin
    uo, vo, Eo = o.rk4(u0, v0, dt, 3, energy=True)
    uu = u0.copy(); vv = v0.copy()
    E = ws.rk4_steps(uu, vv, dt, 3, energy_every=1)
    assert np.max(np.abs(uu - uo)) <= 1e-11 * np.max(np.abs(uo))
    assert np.max(np.abs(vv - vo)) <= 1e-11 * np.max(np.abs(vo))
    assert np.max(np.abs(E - Eo)) <= 1e-11 * np.max(np.abs(Eo))
    wm = mk("mixed")
    uu = u0.copy(); vv = v0.copy()
    wm.rk4_steps(uu, vv, dt, 3)
    H = o.hphys()
    d = uu - uo
    assert math.sqrt(np.sum(H * d * d) / np.sum(H * uo * uo)) <= 1e-5
