"""The batched and the served paths against the CPU oracle, instance by
instance (not only against the single-instance HIP path).

The instances are the batched workload's draws (per instance N ~ U{4,6,8,10},
half flat, half warped with the ADM bump background, Gaussian pulse with a
seeded centre and width), with M capped at 5 elements per side and a few
uneven RK4 step counts so that the oracle finishes in seconds.  FP64: the
state after the steps equals the oracle's to 1e-11 relative (max norm).
Mixed: the H-norm error against the oracle's FP64 state is at most 1e-5.
"""
import math
import os
import sys

import numpy as np
import pytest

import wt_oracle as wo
import wavetoy_hip as wh
from wavetoy_hip import serving as sv

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
import bench  # noqa: E402

pytestmark = pytest.mark.gpu

TOL64 = 1e-11
TOL_MIXED_H = 1e-5


def _specs(n=32, seed=3):
    out = []
    for q, (N, M, curved, centre, width, _) in enumerate(bench.batch_specs(n, seed)):
        out.append((N, 2 + (M - 2) % 4, curved, centre, width, 6 + (q * 7) % 17))
    # every N, flat and curved, appears
    assert {(s[0], s[2]) for s in out} == {(N, c) for N in (4, 6, 8, 10) for c in (False, True)}
    return out


def _oracle_for(N, M, curved, x):
    bg = sv.sample_background3d(sv.adm_bump_background(), 0.0, x[:, 0], x[:, 1], x[:, 2]) if curved else None
    return wo.Oracle(N, M, warp_A=sv.WARP_A if curved else 0.0, tau=wh.tau_effective(sv.TAU_LITERAL, N),
                     eps_ko=sv.EPS_KO, bg=bg)


def _hnorm_rel(a, b, H):
    return math.sqrt(np.sum(H * (a - b) ** 2) / np.sum(H * b ** 2))


def _check(label, prec, N, M, curved, o, u, v, uo, vo):
    if prec == "fp64":
        for got, ref in ((u, uo), (v, vo)):
            err = np.max(np.abs(got - ref)) / np.max(np.abs(ref))
            assert err <= TOL64, f"{label}: N={N} M={M} curved={curved} fp64 error {err:.3e}"
    else:
        H = o.hphys()
        err = _hnorm_rel(u, uo, H)
        assert err <= TOL_MIXED_H, f"{label}: N={N} M={M} curved={curved} mixed H-norm error {err:.3e}"


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_batch_matches_oracle_per_instance(prec):
    import torch
    specs = _specs()
    print(f"batch vs oracle: {len(specs)} instances, {prec}", flush=True)
    b = wh.Batch()
    insts = []
    for N, M, curved, centre, width, ns in specs:
        ws, u, v, dt, _ = bench.make_instance(wh, torch, (N, M, curved, centre, width, 1.0), prec)
        insts.append((ws, u, v, dt, ns, u.cpu().numpy(), v.cpu().numpy()))
        b.add(ws, u, v, dt, ns)
    fin = 0
    while fin < len(insts):
        fin += b.step(5)
    torch.cuda.synchronize()
    b.close()
    for (N, M, curved, _, _, ns), (ws, u, v, dt, _, u0, v0) in zip(specs, insts):
        x = wh.host_coords(ws.mesh, ws.elem)
        o = _oracle_for(N, M, curved, x)
        uo, vo = o.rk4(u0, v0, dt, ns)
        _check("batch", prec, N, M, curved, o, u.cpu().numpy(), v.cpu().numpy(), uo, vo)
        ws.close()


@pytest.mark.parametrize("prec", ["mixed", "fp64"])
def test_served_requests_match_oracle(prec):
    """Requests arriving 1 ms apart under a budget of about three live requests,
    so that they queue, join running batches and leave them; every response
    against the oracle's RK4 of the same request."""
    rng = np.random.default_rng(17)
    reqs = []
    for q, (N, M, curved, centre, width, _) in enumerate(_specs(16, seed=1)):
        reqs.append(sv.Request(rid=q, N=N, M=M, curved=curved, centre=tuple(centre), width=width,
                               T=float(rng.uniform(0.02, 0.05)), arrival=0.001 * q))
    fps = [sv.request_footprint(r, prec, 4) for r in reqs]
    budget = int(3.2 * max(fps))
    print(f"serving vs oracle: {len(reqs)} requests, {prec}, budget {budget} B", flush=True)
    srv = sv.Server(budget_bytes=budget, precision=prec, energy_every=4, chunk_ms=1.0)
    stats = srv.run(reqs)
    assert len(stats.responses) == len(reqs) and not any(r.error for r in stats.responses)
    assert stats.max_live >= 2
    for resp in sorted(stats.responses, key=lambda r: r.rid):
        r = reqs[resp.rid]
        mesh = r.mesh()
        x = wh.host_coords(mesh, wh.make_element(r.N))
        u0, v0 = wh.gaussian_pulse(x, r.centre, width=r.width, amplitude=r.amplitude)
        o = _oracle_for(r.N, r.M, r.curved, x)
        uo, vo, E = o.rk4(u0, v0, resp.dt, resp.nsteps, energy=True)
        if prec == "fp64":
            err = np.max(np.abs(resp.u - uo)) / np.max(np.abs(uo))
            assert err <= TOL64, f"request {r.rid} N={r.N} M={r.M}: fp64 error {err:.3e}"
            assert np.allclose(resp.energy, E[::4][:len(resp.energy)], rtol=1e-11, atol=0)
        else:
            err = _hnorm_rel(resp.u, uo, o.hphys())
            assert err <= TOL_MIXED_H, f"request {r.rid} N={r.N} M={r.M}: mixed H-norm error {err:.3e}"
