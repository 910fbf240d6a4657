"""GPU parity of the first-order (Phi, Pi) path on general meshes (the
curvilinear path of wave3d_curved_rhs!, metric given) against the CPU oracle:
the right-hand side with lapse, shift and metric, KO, and each boundary kind
(Sommerfeld, field-radiation Dirichlet with data, full Dirichlet, excision),
the energy, on the cubed cube, the inflated cube and the radial shell -- FP64
to 1e-12; and the second-order operator with excision faces (tag 8)."""
import numpy as np
import pytest

import wt_oracle as wo
import wavetoy_hip as wh

pytestmark = pytest.mark.gpu


def _oracle(mesh, **kw):
    return wo.Oracle.from_mesh(mesh.N, mesh.coords, mesh.neighbour, mesh.neighbour_face, mesh.bdry, **kw)


def _bg(x, rng):
    n = x.shape[0]
    r = np.linalg.norm(x, axis=1)
    bg = np.zeros((n, 10))
    bg[:, 0] = 1.0 + 0.1 * np.sin(2 * x[:, 0]) * np.cos(x[:, 1])
    bg[:, 1:4] = 0.1 * np.stack([np.cos(x[:, 1]), np.sin(x[:, 2]), -np.cos(x[:, 0] + x[:, 2])], 1)
    bg[:, 4] = 1.1 + 0.1 * np.sin(x[:, 2])
    bg[:, 5] = 0.05 * np.cos(x[:, 0])
    bg[:, 6] = 0.02 * np.sin(r)
    bg[:, 7] = 1.0 + 0.05 * np.cos(x[:, 0] - x[:, 1])
    bg[:, 8] = -0.03 * np.sin(x[:, 0])
    bg[:, 9] = 0.95 + 0.05 * np.sin(x[:, 1])
    return bg


def _relerr(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


MESHES = [("cubed_cube", lambda: wh.make_cubed_cube_mesh(4, 2, 0.3)),
          ("inflated_cube", lambda: wh.make_inflated_cube_mesh(5, 0.1, 0.3, 1.0, 1, outer_tag=7)),
          ("radial_shell", lambda: wh.make_radial_shell_mesh(4, 0.5, 1.0, 2))]


@pytest.mark.parametrize("name,make", MESHES, ids=[m[0] for m in MESHES])
def test_fo_rhs_and_energy_on_general_meshes_match_oracle(name, make):
    print("first-order general mesh", name, flush=True)
    mesh = make()
    N = mesh.N
    ws = wh.Workspace(mesh, wh.make_element(N))
    o = _oracle(mesh)
    x = o.coords()
    rng = np.random.default_rng(len(name))
    coef = wo.fo_coef(_bg(x, rng).reshape(-1)).reshape(-1)
    assert _relerr(ws.fo_coef(_bg(x, rng).reshape(-1).copy()).reshape(-1), coef) < 1e-15
    Phi, Pi = rng.standard_normal(o.n), rng.standard_normal(o.n)
    gPhi, gPi, gD = rng.standard_normal(o.n), rng.standard_normal(o.n), rng.standard_normal(3 * o.n)
    exc = wh.EXCISION_BDRY_TAG if name == "radial_shell" else 0
    cases = [("sommerfeld", {}), ("dirichlet", dict(gPi=gPi, gD=gD)), ("full_dirichlet", dict(gPhi=gPhi, gPi=gPi)),
             ("excision", {})]
    for eps in (0.0, 0.1):
        for kind, kw in cases:
            dP, dQ = np.zeros(o.n), np.zeros(o.n)
            ws.fo_rhs(dP, dQ, Phi, Pi, coef, eps, wh.make_bc3d(kind, excision_tag=exc, **kw))
            rP, rQ = o.fo_rhs_curv(Phi, Pi, coef, eps, wo.make_bc3d(kind, excision_tag=exc, **kw))
            assert _relerr(dP, rP) < 1e-12 and _relerr(dQ, rQ) < 1e-12, (name, eps, kind, _relerr(dP, rP),
                                                                          _relerr(dQ, rQ))
    assert abs(ws.fo_energy(Phi, Pi, coef) / o.fo_energy_curv(Phi, Pi, coef) - 1) < 1e-13
    ws.close()


def test_fo_general_mesh_device_pointers_and_rk4():
    """Device-resident fields: 10 RK4 steps of the first-order system with
    Sommerfeld outer faces on the inflated cube equal the oracle's to 1e-11."""
    import torch
    mesh = wh.make_inflated_cube_mesh(4, 0.1, 0.3, 1.0, 1, outer_tag=7)
    ws = wh.Workspace(mesh, wh.make_element(4))
    o = _oracle(mesh)
    x = o.coords()
    bg = np.zeros((o.n, 10))
    bg[:, 0] = 1.0
    bg[:, 4] = bg[:, 7] = bg[:, 9] = 1.0
    coef = wo.fo_coef(bg.reshape(-1))
    Phi0 = np.exp(-np.sum(x ** 2, axis=1) / 0.05)
    Pi0 = np.zeros(o.n)
    dt = 0.1 * o.hmin
    P, Q = torch.from_numpy(Phi0).cuda(), torch.from_numpy(Pi0).cuda()
    cd = torch.from_numpy(coef).cuda()
    bc = wh.make_bc3d("sommerfeld")
    k = [torch.empty_like(P) for _ in range(8)]
    for _ in range(10):
        ws.fo_rhs(k[0], k[1], P, Q, cd, 0.1, bc)
        ws.fo_rhs(k[2], k[3], P + dt / 2 * k[0], Q + dt / 2 * k[1], cd, 0.1, bc)
        ws.fo_rhs(k[4], k[5], P + dt / 2 * k[2], Q + dt / 2 * k[3], cd, 0.1, bc)
        ws.fo_rhs(k[6], k[7], P + dt * k[4], Q + dt * k[5], cd, 0.1, bc)
        P = P + dt / 6 * (k[0] + 2 * k[2] + 2 * k[4] + k[6])
        Q = Q + dt / 6 * (k[1] + 2 * k[3] + 2 * k[5] + k[7])
    rP, rQ = o.fo_rk4_curv(Phi0, Pi0, coef, dt, 10, eps_ko=0.1, bc=wo.make_bc3d("sommerfeld"))
    assert _relerr(P.cpu().numpy(), rP) < 1e-11 and _relerr(Q.cpu().numpy(), rQ) < 1e-11
    ws.close()


def test_second_order_operator_with_excision_faces_matches_oracle():
    """The second-order drop-in RHS and RK4 steps on the radial shell whose
    inner sphere carries the excision tag (no SAT there) and the outer one
    Dirichlet data: equal to the oracle to 1e-12 / 1e-11."""
    mesh = wh.make_radial_shell_mesh(4, 0.5, 1.0, 2, outer_tag=1)
    bv = (0.3, 0.0, 0.0, 0.0, 0.0, 0.0)
    tau = wh.tau_effective(None, 4)
    ws = wh.Workspace(mesh, wh.make_element(4), tau=tau, bdry_values=bv)
    o = _oracle(mesh, tau=tau, bdry_values=bv)
    rng = np.random.default_rng(8)
    u, v = rng.standard_normal(o.n), rng.standard_normal(o.n)
    out = np.zeros(o.n)
    ws.rhs(out, u, v, bv)
    assert _relerr(out, o.lap(u, v)) < 1e-12
    x = o.coords()
    u0 = np.exp(-np.sum((x - 0.7) ** 2, axis=1) / 0.05)
    v0 = np.zeros(o.n)
    dt = 0.2 * ws.h_min
    uu, vv = u0.copy(), v0.copy()
    ws.rk4_steps(uu, vv, dt, 5)
    ru, rv = o.rk4(u0, v0, dt, 5)
    assert _relerr(uu, ru) < 1e-11 and _relerr(vv, rv) < 1e-11
    ws.close()
