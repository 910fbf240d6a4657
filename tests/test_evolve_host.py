"""Host logic of the evolve3d drivers (no GPU): the initial-data closed forms,
face classification and boundary-kind checks of the first-order driver, and
its argument errors (test/evolve3d_tests.jl:39-42, src/boundaries1d.jl:97-138).

The closed forms are pinned by the PDE they solve: second differences in t
and in x, y, z (step 1e-3, O(h^2)) agree to 1e-4 relative, and the stated
u_t is the time derivative of u (central difference, step 1e-4, to 1e-5)."""
import math

import numpy as np
import pytest

import wavetoy_hip as wh
from wavetoy_hip import evolve as ev

PTS = np.array([[0.7, 0.6, 0.55], [0.52, 0.5, 0.49], [0.8, 0.3, 0.45], [0.1, 0.9, 0.2]])
CTR = (0.5, 0.5, 0.5)


def _pde_residuals(f, t, h=1e-3):
    u = f(PTS, t)[0]
    utt = (f(PTS, t + h)[0] - 2 * u + f(PTS, t - h)[0]) / h ** 2
    lap = sum((f(PTS + h * np.eye(3)[a], t)[0] - 2 * u + f(PTS - h * np.eye(3)[a], t)[0]) / h ** 2
              for a in range(3))
    k = h / 10
    ut = (f(PTS, t + k)[0] - f(PTS, t - k)[0]) / (2 * k)
    scale_tt = np.max(np.abs(utt)) + 1e-300
    scale_t = np.max(np.abs(ut)) + 1e-300
    return np.max(np.abs(utt - lap)) / scale_tt, np.max(np.abs(ut - f(PTS, t)[1])) / scale_t


@pytest.mark.parametrize("t", [0.0, 0.07, 0.2])
def test_outgoing_pulse_solves_the_wave_equation(t):
    f = lambda X, tt: wh.outgoing_pulse(X, tt, A=1.3, s0=0.25, sigma=0.07, centre=CTR)
    r_wave, r_t = _pde_residuals(f, t)
    assert r_wave < 1e-4 and r_t < 1e-5


def test_outgoing_pulse_is_regular_at_the_centre_and_outgoing():
    """At r = 0 the closed form's limit: u = 2 A b / s^2 exp(-b^2 / 2 s^2); the
    pulse peak moves out at unit speed."""
    A, s0, sg = 1.0, 0.25, 0.05
    u, ud = wh.outgoing_pulse(np.array([list(CTR)]), 0.1, A=A, s0=s0, sigma=sg, centre=CTR)
    b = 0.1 + s0
    assert np.isfinite(ud[0]) and abs(u[0] - 2 * A * b / sg ** 2 * math.exp(-b * b / (2 * sg * sg))) < 1e-12
    r = np.linspace(0.0, 0.49, 4901)
    X = np.stack([0.5 + r, 0.5 + 0 * r, 0.5 + 0 * r], axis=1)
    for t in (0.0, 0.1):
        u, _ = wh.outgoing_pulse(X, t, A=A, s0=s0, sigma=sg, centre=CTR)
        rp = r[np.argmax(np.abs(u * r))]
        assert abs(rp - (s0 + t)) < 2e-3


@pytest.mark.parametrize("n", [1, 2])
def test_radial_eigenmode_solves_the_wave_equation(n):
    f = lambda X, tt: wh.eigenmode_radial(X, tt, A=0.8, R=0.5, n=n, centre=CTR)
    r_wave, r_t = _pde_residuals(f, 0.13)
    assert r_wave < 1e-4 and r_t < 1e-5
    u, _ = wh.eigenmode_radial(np.array([list(CTR), [1.0, 0.5, 0.5]]), 0.0, A=0.8, R=0.5, n=n, centre=CTR)
    assert u[0] == 0.8 and abs(u[1]) < 1e-16          # value A at r = 0, node of sin at r = R


def test_face_classification():
    """classify_face1d: s_R = a - beta, s_L = -a - beta against the outward sign."""
    assert wh.classify_face1d(1.0, 0.0, 1) == wh.FACE_SUBLUMINAL
    assert wh.classify_face1d(1.0, 0.5, -1) == wh.FACE_SUBLUMINAL
    assert wh.classify_face1d(1.0, 2.0, 1) == wh.FACE_INFLOW       # both characteristics move in -x
    assert wh.classify_face1d(1.0, 2.0, -1) == wh.FACE_OUTFLOW
    assert wh.classify_face1d(1.0, -2.0, 1) == wh.FACE_OUTFLOW
    assert wh.classify_face1d(1.0, 1.0, 1) == wh.FACE_SONIC
    # a = alpha sqrt(gamma^nn): alpha = 2, gamma^yy = 4 makes a = 4 > beta = 3
    assert wh.classify_face3d(2.0, 0.0, 3.0, 0.0, 1.0, 4.0, 1.0, 2, 1) == wh.FACE_SUBLUMINAL
    assert wh.classify_face3d(1.0, 0.0, 3.0, 0.0, 1.0, 1.0, 1.0, 2, 1) == wh.FACE_INFLOW


def test_boundary_kinds_are_checked_against_the_face_class():
    BC = wh.BC3d._NAMES
    wh.validate_bc1d(wh.FACE_SUBLUMINAL, BC["sommerfeld"], "-x side")
    wh.validate_bc1d(wh.FACE_SUBLUMINAL, BC["dirichlet"], "-x side")
    wh.validate_bc1d(wh.FACE_OUTFLOW, BC["excision"], "-x side")
    wh.validate_bc1d(wh.FACE_INFLOW, BC["full_dirichlet"], "-x side")
    for cls, kind in ((wh.FACE_SUBLUMINAL, "excision"), (wh.FACE_SUBLUMINAL, "full_dirichlet"),
                      (wh.FACE_OUTFLOW, "sommerfeld"), (wh.FACE_INFLOW, "dirichlet"), (wh.FACE_SONIC, "sommerfeld")):
        with pytest.raises(ValueError):
            wh.validate_bc1d(cls, BC[kind], "+y side")


def test_face_kinds_of_the_cube():
    """bc = "auto": Minkowski gives Sommerfeld on every side; a shift of 2 along
    x makes -x an outflow (excision) side and +x an inflow (full Dirichlet)
    side; a 6-tuple is checked against the classes."""
    BC = wh.BC3d._NAMES
    N, M = 4, (3, 2, 2)
    n = N ** 3 * 12
    x = np.zeros((n, 3))
    bgf, _, _, _, speed = ev._background3d("minkowski")
    bg = bgf(x)
    assert ev._face_kinds("auto", bg, M, N) == (BC["sommerfeld"],) * 6
    assert speed == math.sqrt(3.0)
    bgf, _, _, _, speed = ev._background3d("constant_shift", (2.0, 0.1, 0.0))
    bg = bgf(x)
    k = ev._face_kinds("auto", bg, M, N)
    assert k == (BC["excision"], BC["full_dirichlet"]) + (BC["sommerfeld"],) * 4
    assert abs(speed - (2.1 + math.sqrt(3.0))) < 1e-15
    ok = ("excision", "full_dirichlet", "dirichlet", "sommerfeld", "sommerfeld", "dirichlet")
    assert ev._face_kinds(ok, bg, M, N)[2] == BC["dirichlet"]
    with pytest.raises(ValueError):
        ev._face_kinds(("sommerfeld",) * 6, bg, M, N)


def test_radial_shift_background():
    """The excision shell's background (src/evolve.jl:1129-1145): flat alpha and
    gamma, an outward shift ramping linearly from V at R1 to 1/20 at R2."""
    bgf, phi, pi, d, speed = ev._background3d("radial_shift", (1.5, 0.0, 0.0), 0.5, 1.0)
    x = np.array([[0.5, 0.0, 0.0], [0.0, 0.0, -1.0], [0.0, 0.75, 0.0]])
    bg = bgf(x)
    assert np.allclose(bg[:, 0], 1.0) and np.allclose(bg[:, [4, 7, 9]], 1.0) and np.allclose(bg[:, [5, 6, 8]], 0.0)
    assert np.allclose(bg[0, 1:4], [1.5, 0, 0]) and np.allclose(bg[1, 1:4], [0, 0, -0.05])
    assert np.allclose(bg[2, 1:4], [0, 0.5 * (1.5 + 0.05), 0])
    assert speed == 2.5


def test_side_nodes():
    N, M = 4, (3, 2, 5)
    mid = (N + 1) // 2 - 1
    assert ev._side_node(M, N, 0) == mid * N + mid * N * N                      # element 0, i = 0
    e = 2                                                                       # ex = 2
    assert ev._side_node(M, N, 1) == e * N ** 3 + (N - 1) + mid * N + mid * N * N
    e = 3 * 2 * 4                                                               # ez = 4
    assert ev._side_node(M, N, 5) == e * N ** 3 + mid + mid * N + (N - 1) * N * N


def test_inverse_metric_diagonal():
    rng = np.random.default_rng(1)
    A = rng.standard_normal((3, 3))
    g = A @ A.T + 3 * np.eye(3)
    gu = np.linalg.inv(g)
    d = ev._inverse_metric_diag((g[0, 0], g[0, 1], g[0, 2], g[1, 1], g[1, 2], g[2, 2]))
    assert np.allclose(d, np.diag(gu), rtol=1e-13)


def test_driver_argument_errors():
    with pytest.raises(ValueError):
        wh.evolve3d(formulation="bogus", N=4, M=2)
    with pytest.raises(ValueError):
        wh.evolve3d(formulation="conservative", N=4, M=2, mesh_kind="bogus")
    with pytest.raises(ValueError, match="exact"):
        wh.evolve3d(formulation="conservative", N=4, M=2, mesh_kind="cubed_cube", bc="dirichlet", ic="gaussian")
    with pytest.raises(ValueError, match="sommerfeld"):
        wh.evolve3d(formulation="conservative", N=4, M=2, mesh_kind="cubed_cube", bc="auto")
    with pytest.raises(ValueError):
        wh.evolve3d(formulation="conservative", N=4, M=2, bc="sommerfeld")
    with pytest.raises(ValueError):
        wh.evolve3d(formulation="conservative", N=4, M=2, ic="radial")
    with pytest.raises(ValueError):
        wh.evolve3d(formulation="conservative", N=4, M=2, background="radial_shift")
    with pytest.raises(ValueError):
        wh.evolve3d(formulation="conservative", N=4, M=2, bc=("excision",) * 6)
    with pytest.raises(ValueError):
        wh.evolve3d(N=4, M=2, outer_bc="sommerfeld", nsteps=1)


def test_evolve3d_mesh_kind_arguments():
    with pytest.raises(ValueError, match="mesh_kind"):
        wh.evolve3d(mesh_kind="bogus_shell", device=False)
    with pytest.raises(ValueError, match="spherical"):
        wh.evolve3d(mesh_kind="cubed_cube", outer_bc="sommerfeld", device=False)
    with pytest.raises(ValueError, match="outer_bc"):
        wh.evolve3d(mesh_kind="inflated_cube", outer_bc="absorbing", device=False)


def test_spectral_radius_estimate_against_dense_eigenvalues():
    """The patch-mesh time step's spectral-radius estimate (power iteration in
    the H |J| inner product with a Rayleigh-quotient stop) against a dense
    eigenvalue solve of the operator on a small cubed ball (the oracle's
    operator, the same SIPG form as the device's)."""
    import wt_oracle as wo
    mesh = wh.make_cubed_ball_mesh(3, 1, 1.0)
    o = wo.Oracle.from_mesh(3, mesh.coords, mesh.neighbour, mesh.neighbour_face, mesh.bdry, tau=8.0 * 4)
    H = ev._hphys(3, mesh.coords)
    n = o.n
    z = np.zeros(n)
    A = np.zeros((n, n))
    for j in range(n):
        e = np.zeros(n)
        e[j] = 1.0
        A[:, j] = o.lap(e, z)
    rho = np.max(np.abs(np.linalg.eigvals(A)))
    est = ev._spectral_radius(lambda v: o.lap(np.ascontiguousarray(v), z), H, n, tol=1e-4)
    assert est <= rho * (1 + 1e-9) and est > 0.99 * rho, (est, rho)
    with pytest.raises(RuntimeError):
        ev._spectral_radius(lambda v: o.lap(np.ascontiguousarray(v), z), H, n, tol=1e-15, max_iters=3)
