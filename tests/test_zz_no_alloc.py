"""wt2_rk4_steps_noalloc (the entry point that allocates nothing: its energy
history and reduction partials live in the library's fixed device array)
against wt2_rk4_steps on the same context and state: bitwise equal u, u_t
and energy history, on two calls in a row, the second with a smaller mesh,
then on a curved context; more energy samples than the array holds are
refused before anything runs."""
import ctypes

import numpy as np
import pytest

import wavetoy_hip as wh
from wavetoy_hip import _lib

pytestmark = pytest.mark.gpu


def _dp(t):
    return ctypes.cast(ctypes.c_void_p(t.data_ptr()), ctypes.POINTER(ctypes.c_double))


def _noalloc(ws, u, v, dt, nsteps, energy_every):
    ne = 1 + nsteps // energy_every if energy_every > 0 else 0
    E = np.zeros(max(ne, 1))
    rc = _lib.lib().wt2_rk4_steps_noalloc(ws.ctx, _dp(u), _dp(v), float(dt), int(nsteps), int(energy_every),
                                           E.ctypes.data_as(ctypes.POINTER(ctypes.c_double)) if ne else None)
    return rc, E[:ne]


def _workspace(N, M, prec, curved):
    if curved:
        mesh = wh.make_warped_uniform_hex(M, M, M, 0.0, 1.0, 0.05)
    else:
        mesh = wh.make_uniform_hex(M, 0.0, 1.0)
    elem = wh.make_element(N)
    x = wh.host_coords(mesh, elem)
    bg = wh.sample_background3d(wh.adm_bump_background(), 0.0, x[:, 0], x[:, 1], x[:, 2]) if curved else None
    ws = wh.Workspace(mesh, elem, tau=wh.tau_effective(20.0, N), eps_ko=0.1, precision=prec, bg=bg)
    ws.info()      # computes a curved context's coefficients (the entry point refuses pending ones)
    u0, v0 = wh.gaussian_pulse(x, (0.45, 0.55, 0.5), width=0.12)
    return ws, u0, v0 + 0.1 * u0


def _compare(N, M, prec, curved, nsteps, every):
    import torch
    print(f"noalloc vs rk4_steps: N={N} M={M} {prec} curved={curved} steps={nsteps} every={every}", flush=True)
    ws, u0, v0 = _workspace(N, M, prec, curved)
    dt = 0.2 * ws.h_min
    ua, va = torch.from_numpy(u0).cuda(), torch.from_numpy(v0).cuda()
    Ea = ws.rk4_steps(ua, va, dt, nsteps, energy_every=every)
    ub, vb = torch.from_numpy(u0).cuda(), torch.from_numpy(v0).cuda()
    torch.cuda.synchronize()
    rc, Eb = _noalloc(ws, ub, vb, dt, nsteps, every)
    assert rc == 0, _lib.lib().wt2_last_error()
    torch.cuda.synchronize()
    assert torch.equal(ua, ub) and torch.equal(va, vb)
    assert np.array_equal(Ea, Eb)
    assert len(Eb) == 1 + nsteps // every and np.all(np.isfinite(Eb))
    ws.close()


def test_noalloc_equals_rk4_steps_two_calls_second_smaller():
    _compare(6, 4, "fp64", False, 12, 3)
    _compare(4, 2, "mixed", False, 5, 1)


def test_noalloc_equals_rk4_steps_curved():
    _compare(8, 2, "fp64", True, 6, 2)
    _compare(4, 3, "mixed", True, 4, 4)


def test_noalloc_refuses_more_energy_samples_than_its_array():
    import torch
    ws, u0, v0 = _workspace(4, 2, "fp64", False)
    u, v = torch.from_numpy(u0).cuda(), torch.from_numpy(v0).cuda()
    before = u.clone()
    nsteps = 70000
    E = np.zeros(nsteps + 1)
    rc = _lib.lib().wt2_rk4_steps_noalloc(ws.ctx, _dp(u), _dp(v), 1e-4, nsteps, 1,
                                           E.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
    assert rc == _lib.WT2_ERR_ARG
    assert b"WT2_NOALLOC_MAX_ENERGIES" in _lib.lib().wt2_last_error()
    torch.cuda.synchronize()
    assert torch.equal(u, before)          # nothing ran
    ws.close()


def test_host_mirror_equals_rk4_steps():
    """Workspace.rk4_steps_noalloc (the Python mirror of the entry point)
    gives rk4_steps's bits."""
    import torch
    ws, u0, v0 = _workspace(5, 3, "fp64", False)
    dt = 0.2 * ws.h_min
    ua, va = torch.from_numpy(u0).cuda(), torch.from_numpy(v0).cuda()
    ub, vb = ua.clone(), va.clone()
    Ea = ws.rk4_steps(ua, va, dt, 6, energy_every=2)
    Eb = ws.rk4_steps_noalloc(ub, vb, dt, 6, energy_every=2)
    torch.cuda.synchronize()
    assert torch.equal(ua, ub) and torch.equal(va, vb) and np.array_equal(Ea, Eb)
    ws.close()
