"""Batched evolutions (many uneven instances per launch) give bitwise the
results of running each instance alone, energies included."""
import math

import numpy as np
import pytest

import wavetoy_hip as wh

pytestmark = pytest.mark.gpu


def _instances():
    # (N, M, curved, precision)
    return [(4, 2, False, "fp64"), (6, 3, True, "fp64"), (8, 2, False, "mixed"), (4, 5, True, "mixed"),
            (10, 2, False, "fp64"), (6, 4, False, "fp64"), (4, 3, False, "fp64"), (8, 3, True, "fp64")]


def _make(N, M, curved, prec, seed):
    import torch
    if curved:
        mesh = wh.make_warped_uniform_hex(M, M, M, 0.0, 1.0, 0.05)
    else:
        mesh = wh.make_uniform_hex(M, 0.0, 1.0)
    elem = wh.make_element(N)
    x = wh.host_coords(mesh, elem)
    bg = wh.sample_background3d(wh.adm_bump_background(), 0.0, x[:, 0], x[:, 1], x[:, 2]) if curved else None
    ws = wh.Workspace(mesh, elem, tau=wh.tau_effective(20.0, N), eps_ko=0.1, precision=prec, bg=bg)
    rng = np.random.default_rng(seed)
    c = rng.uniform(0.3, 0.7, 3)
    w = rng.uniform(0.05, 0.2)
    u0, v0 = wh.gaussian_pulse(x, c, width=w)
    dt = 0.2 * ws.h_min
    return ws, torch.from_numpy(u0).cuda(), torch.from_numpy(v0).cuda(), dt


def test_batch_equals_single_runs():
    import torch
    print("batch vs single", flush=True)
    insts = _instances()
    nsteps = [7, 12, 5, 9, 3, 12, 10, 6]
    singles = []
    for q, (N, M, cv, pr) in enumerate(insts):
        ws, u, v, dt = _make(N, M, cv, pr, q)
        E = ws.rk4_steps(u, v, dt, nsteps[q], energy_every=1 if q % 2 == 0 else 3)
        singles.append((u.cpu().numpy(), v.cpu().numpy(), E))
    b = wh.Batch()
    keep = []
    slots = []
    for q, (N, M, cv, pr) in enumerate(insts):
        ws, u, v, dt = _make(N, M, cv, pr, q)
        keep.append((ws, u, v))
        slots.append(b.add(ws, u, v, dt, nsteps[q], energy_every=1 if q % 2 == 0 else 3))
    fin = 0
    while fin < len(insts):
        fin += b.step(4)
    torch.cuda.synchronize()
    for q, s in enumerate(slots):
        done, E = b.slot_info(s)
        assert done == nsteps[q]
        ws, u, v = keep[q]
        assert np.array_equal(u.cpu().numpy(), singles[q][0]), f"instance {q} u differs"
        assert np.array_equal(v.cpu().numpy(), singles[q][1]), f"instance {q} u_t differs"
        assert np.array_equal(E, singles[q][2]), f"instance {q} energies differ"
    b.close()
    # contexts are usable alone again after the batch is gone
    ws, u, v = keep[0]
    ws.rk4_steps(u, v, 1e-3, 1)


def test_batch_add_remove_midway():
    import torch
    print("batch add/remove", flush=True)
    b = wh.Batch()
    wsa, ua, va, dta = _make(6, 3, False, "fp64", 11)
    wsb, ub, vb, dtb = _make(4, 2, True, "fp64", 12)
    ref_u = ua.clone(); ref_v = va.clone()
    wsa2, _, _, _ = _make(6, 3, False, "fp64", 11)
    wsa2.rk4_steps(ref_u, ref_v, dta, 10)
    sa = b.add(wsa, ua, va, dta, 10)
    b.step(4)
    sb = b.add(wsb, ub, vb, dtb, 8)
    b.step(3)
    b.remove(sb)
    b.step(100)
    assert b.slot_info(sa)[0] == 10
    assert torch.equal(ua, ref_u) and torch.equal(va, ref_v)
