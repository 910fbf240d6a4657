"""Meshes given by coordinates and connectivity (wt2_ctx_create_mesh): the
multi-patch cubed ball, whose patches meet with orientation != 0, and
relabelled copies of uniform meshes.

The oracle restates the general-connectivity path too (wto_create_mesh: the
same D4 encoding, the same matching of face nodes by coordinates), so the
1e-12 parity test below shows the two agree; the mathematics pins both
independently:
  * on affine elements the operator is the uniform-mesh one (oracle-pinned)
    whatever the element numbering and the orientation of each element's
    local axes;
  * relabelling the elements of a curved mesh and turning each element's
    axes by a symmetry of the cube changes nothing but rounding;
  * the SIPG operator is symmetric in the H3 |J| inner product and negative
    semi-definite at zero Dirichlet data on the cubed ball;
  * the standing radial mode on the cubed ball converges to the exact
    solution as N grows;
  * Sommerfeld outer faces only dissipate energy.
"""
import math

import numpy as np
import pytest

import wavetoy_hip as wh

pytestmark = pytest.mark.gpu


def _orient_pq(ori, N, p, q):
    a, b = (q, p) if ori & 1 else (p, q)
    if ori & 2:
        a = N - 1 - a
    if ori & 4:
        b = N - 1 - b
    return a, b


def _face_node_index(N, f, p, q):
    ax, r = f // 2, (N - 1 if f % 2 else 0)
    if ax == 0:
        return r + N * (p + N * q)
    if ax == 1:
        return p + N * (r + N * q)
    return p + N * (q + N * r)


def _orientations(mesh):
    """The D4 code of every shared face, found by brute force over the 8
    maps of the face nodes (encoding of wt2_ctx_create_mesh)."""
    N, X = mesh.N, mesh.coords.reshape(mesh.Ne, -1, 3)
    ori = np.zeros((mesh.Ne, 6), dtype=np.int32)
    for e in range(mesh.Ne):
        for f in range(6):
            if mesh.neighbour[e, f] == 0:
                continue
            en, fn = mesh.neighbour[e, f] - 1, mesh.neighbour_face[e, f] - 1
            best = None
            for o in range(8):
                d = 0.0
                for q in range(N):
                    for p in range(N):
                        pn, qn = _orient_pq(o, N, p, q)
                        d = max(d, np.max(np.abs(X[e, _face_node_index(N, f, p, q)]
                                                 - X[en, _face_node_index(N, fn, pn, qn)])))
                if best is None or d < best[0]:
                    best = (d, o)
            ori[e, f] = best[1]
    return ori


def _relabel(X, rng):
    """Permute the elements and apply a random symmetry of the cube (axis
    permutation and reversals) to each element's node array.  Returns the new
    coordinates and, per new node, the flat index of the old node it is."""
    Ne, N = X.shape[0], X.shape[1]
    perm = rng.permutation(Ne)
    idx = np.arange(Ne * N ** 3).reshape(Ne, N, N, N)
    Xn = np.empty_like(X)
    src = np.empty_like(idx)
    for en, eo in enumerate(perm):
        axes = tuple(rng.permutation(3))
        flips = rng.integers(0, 2, size=3)
        A, I = np.transpose(X[eo], axes + (3,)), np.transpose(idx[eo], axes)
        for d in range(3):
            if flips[d]:
                A, I = np.flip(A, axis=d), np.flip(I, axis=d)
        Xn[en], src[en] = A, I
    return Xn, src.reshape(-1)


def _rhs(work, u, ud, bv):
    out = np.zeros_like(u)
    work.rhs(out, u, ud, bv)
    return out


def _abs_detJ(mesh):
    """|det dx/dxi| per node, G applied to the coordinates."""
    N = mesh.N
    G = wh.make_element(N).G
    X = mesh.coords                                  # [e, k, j, i, c]
    J = np.stack([np.einsum("ip,ekjpc->ekjic", G, X),
                  np.einsum("jp,ekpic->ekjic", G, X),
                  np.einsum("kp,epjic->ekjic", G, X)], axis=-1)
    return np.abs(np.linalg.det(J)).reshape(-1)


def _h3(N, Ne):
    w = wh.make_element(N).w
    return np.tile(np.einsum("k,j,i->kji", w, w, w).reshape(-1), Ne)


@pytest.mark.parametrize("N,M,precision", [(4, (3, 2, 2), "fp64"), (6, (2, 2, 3), "fp64"), (5, (2, 3, 2), "mixed")])
def test_affine_mesh_any_labelling_matches_uniform_operator(N, M, precision):
    rng = np.random.default_rng(N)
    mesh = wh.make_uniform_hex(M, 0.0, 1.0)
    elem = wh.make_element(N)
    bv = (0.3, -0.2, 0.1, 0.4, -0.5, 0.25)
    ref = wh.Workspace(mesh, elem, bdry_values=bv, precision=precision)
    X = ref.coords().reshape(mesh.Ne, N, N, N, 3)
    Xn, src = _relabel(X, rng)
    # outer face tags: the face's side of the unit cube, so Dirichlet values follow
    nbr, nface, bdry = wh.connect_by_coords(N, Xn)
    for e in range(mesh.Ne):
        for f in range(6):
            if nbr[e, f] == 0:
                c = wh.core._face_nodes(Xn[e:e + 1], f).reshape(-1, 3).mean(axis=0)
                a = int(np.argmin(np.minimum(np.abs(c), np.abs(1 - c))))
                bdry[e, f] = 2 * a + (1 if c[a] > 0.5 else 0) + 1
    hm = wh.HexMesh(N, Xn, nbr, nface, bdry)
    work = wh.Workspace(hm, elem, bdry_values=bv, precision=precision)
    assert work.info()["curved"] == 1
    u = rng.standard_normal(ref.npoints)
    ud = rng.standard_normal(ref.npoints)
    want = _rhs(ref, u, ud, bv)
    got = _rhs(work, u[src], ud[src], bv)
    tol = (1e-11 if precision == "fp64" else 2e-5) * np.max(np.abs(want))
    assert np.max(np.abs(got - want[src])) < tol
    assert np.array_equal(work.coords().reshape(-1, 3), Xn.reshape(-1, 3))


@pytest.mark.parametrize("explicit", [False, True])
def test_cubed_ball_relabelling_invariance(explicit):
    N, M = 5, 2
    rng = np.random.default_rng(11)
    ball = wh.make_cubed_ball_mesh(N, M, R=1.0, outer_tag=7)
    Xn, src = _relabel(ball.coords, rng)
    nbr, nface, bdry = wh.connect_by_coords(N, Xn, outer_tag=7)
    hm = wh.HexMesh(N, Xn, nbr, nface, bdry)
    ori = _orientations(ball)
    assert np.any(ori[ball.neighbour > 0] != 0)       # the patches meet turned
    if explicit:
        ball.orientation = ori
        hm.orientation = _orientations(hm)
    elem = wh.make_element(N)
    a = wh.Workspace(ball, elem, sommerfeld_R=1.0)
    b = wh.Workspace(hm, elem, sommerfeld_R=1.0)
    u = rng.standard_normal(a.npoints)
    ud = rng.standard_normal(a.npoints)
    want = _rhs(a, u, ud, (0.0,) * 6)
    got = _rhs(b, u[src], ud[src], (0.0,) * 6)
    assert np.max(np.abs(got - want[src])) < 1e-11 * np.max(np.abs(want))


@pytest.mark.parametrize("N", [4, 6])
def test_cubed_ball_operator_symmetric_and_nonpositive(N):
    ball = wh.make_cubed_ball_mesh(N, 2, R=1.0)
    work = wh.Workspace(ball, wh.make_element(N))
    W = _h3(N, ball.Ne) * _abs_detJ(ball)            # H3 / wgt with wgt = 1 / |J|
    rng = np.random.default_rng(N)
    u, v = rng.standard_normal(work.npoints), rng.standard_normal(work.npoints)
    z = np.zeros_like(u)
    Lu, Lv = _rhs(work, u, z, (0.0,) * 6), _rhs(work, v, z, (0.0,) * 6)
    a, b = np.dot(v * W, Lu), np.dot(u * W, Lv)
    scale = math.sqrt(abs(np.dot(u * W, Lu) * np.dot(v * W, Lv)))
    assert abs(a - b) < 1e-11 * scale
    assert np.dot(u * W, Lu) < 0 and np.dot(v * W, Lv) < 0


def _radial_mode_error(N, M, t1):
    ball = wh.make_cubed_ball_mesh(N, M, R=1.0)
    work = wh.Workspace(ball, wh.make_element(N))
    X = ball.coords.reshape(-1, 3)
    u, ud = wh.eigenmode_radial(X, 0.0, R=1.0)
    dt0 = wh.recommended_dt_rk4(work, cfl=0.2)
    n = max(1, math.ceil(t1 / dt0))
    work.rk4_steps(u, ud, t1 / n, n)
    ue, _ = wh.eigenmode_radial(X, t1, R=1.0)
    return np.max(np.abs(u - ue)), u


def test_cubed_ball_standing_wave_converges():
    e4, _ = _radial_mode_error(4, 2, 0.25)
    e7, u7 = _radial_mode_error(7, 2, 0.25)
    assert np.all(np.isfinite(u7))
    assert e7 < 2e-4 and e7 < e4 / 10, (e4, e7)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_cubed_ball_sommerfeld_dissipates(precision):
    N = 6
    ball = wh.make_cubed_ball_mesh(N, 2, R=1.0, outer_tag=7)
    work = wh.Workspace(ball, wh.make_element(N), sommerfeld_R=1.0, precision=precision)
    X = ball.coords.reshape(-1, 3)
    u, ud = wh.outgoing_pulse(X, 0.0, s0=0.3, sigma=0.1)
    dt = wh.recommended_dt_rk4(work, cfl=0.2)
    n = 20 * math.ceil(1.2 / dt / 20)                # to t ~ 1.2: the pulse has crossed r = 1
    E = work.rk4_steps(u, ud, dt, n, energy_every=20)
    assert np.all(np.isfinite(u))
    assert np.all(np.diff(E) <= 1e-12 * E[0]), E
    assert E[-1] < 0.5 * E[0], E                     # it left through the sphere


def test_cubed_ball_mixed_matches_fp64():
    N = 6
    ball = wh.make_cubed_ball_mesh(N, 2, R=1.0)
    X = ball.coords.reshape(-1, 3)
    res = {}
    for prec in ("fp64", "mixed"):
        work = wh.Workspace(ball, wh.make_element(N), precision=prec)
        u, ud = wh.eigenmode_radial(X, 0.0, R=1.0)
        n = math.ceil(0.25 / wh.recommended_dt_rk4(work, cfl=0.2))
        work.rk4_steps(u, ud, 0.25 / n, n)
        res[prec] = u
    assert np.all(np.isfinite(res["fp64"]))
    assert np.max(np.abs(res["mixed"] - res["fp64"])) < 1e-5 * np.max(np.abs(res["fp64"]))


def _rayleigh(mesh, u):
    """-<u, L u>_Hphys / <u, u>_Hphys with Hphys = w_i w_j w_k |J|, zero
    Dirichlet data (test/dirichlet_cubed_cube_tests.jl:58-73)."""
    work = wh.Workspace(mesh, wh.make_element(mesh.N))
    W = _h3(mesh.N, mesh.Ne) * _abs_detJ(mesh)
    Lu = _rhs(work, u, np.zeros_like(u), (0.0,) * 6)
    return -np.dot(u * W, Lu) / np.dot(u * W, u)


def test_cubed_cube_rayleigh_quotient_converges():
    # test/dirichlet_cubed_cube_tests.jl:54-81: u = cos(pi x/2) cos(pi y/2)
    # cos(pi z/2) on [-1, 1]^3, lambda = 3 pi^2 / 4, N = 4, M = 2, 4, 8
    lam = 3 * math.pi ** 2 / 4
    errs = []
    for M in (2, 4, 8):
        mesh = wh.make_cubed_cube_mesh(4, M, 0.3)
        X = mesh.coords.reshape(-1, 3)
        u = np.prod(np.cos(0.5 * math.pi * X), axis=1)
        errs.append(abs(_rayleigh(mesh, u) - lam))
    assert all(e > 0 for e in errs) and errs[-1] < errs[0], errs
    assert (errs[0] / errs[-1]) ** 0.5 > 4.0, errs
    assert errs[-1] < 1e-3, errs


def test_inflated_cube_rayleigh_quotient():
    # test/dirichlet_cubed_sphere_tests.jl:84-125: u = j0(pi r), lambda = pi^2
    # on the ball R2 = 1 (L = 0.1, R1 = 0.3), N = 4
    errs = []
    for M in (2, 4):
        mesh = wh.make_inflated_cube_mesh(4, 0.1, 0.3, 1.0, M)
        u, _ = wh.eigenmode_radial(mesh.coords.reshape(-1, 3), 0.0, R=1.0)
        errs.append(abs(_rayleigh(mesh, u) - math.pi ** 2))
    assert errs[1] / math.pi ** 2 < 1e-3, errs
    assert errs[1] < errs[0] / 10, errs


def test_evolve3d_on_patch_meshes():
    # src/evolve.jl:913-928, 961-968: the inflated cube with the j0 standing
    # mode (exact solution for l2_err) and the cubed cube with the standing
    # cartesian wave; tau = 8 (N-1)^2 on both
    r = wh.evolve3d(mesh_kind="inflated_cube", ic="radial", N=6, M=2, t1=0.2, Nt=3)
    assert r["tau"] == 8 * 25 and r["x0"] == -1.0 and r["sommerfeld_R"] == math.inf
    assert np.all(np.isfinite(r["l2_err"])) and r["l2_err"][0] < 1e-12 and r["l2_err"][-1] < 2e-3
    assert r["xs_line"].size > 0 and abs(r["y_target"]) < 1e-12 and abs(r["z_target"]) < 1e-12
    r = wh.evolve3d(mesh_kind="cubed_cube", ic="cartesian", ic_wavenumber=math.pi, N=6, M=2, t1=0.1, Nt=2)
    assert np.all(np.isfinite(r["u_final"])) and r["l2_err"][-1] < 1e-3, r["l2_err"]


def test_evolve3d_inflated_cube_sommerfeld_lets_the_pulse_out():
    kw = dict(mesh_kind="inflated_cube", ic="outgoing", N=6, M=2, t1=1.2, Nt=4)
    som = wh.evolve3d(outer_bc="sommerfeld", **kw)
    dirichlet = wh.evolve3d(outer_bc="dirichlet", **kw)
    assert som["sommerfeld_R"] == 1.0
    assert np.all(np.diff(som["energy"]) <= 1e-12 * som["energy"][0])
    assert som["energy"][-1] < 0.2 * dirichlet["energy"][-1], (som["energy"], dirichlet["energy"])


def test_batch_of_patch_meshes_matches_single_runs():
    # batches (one launch per kernel group per stage) give the bits of rk4_steps,
    # also with a uniform warped context of the same N and precision in the mix
    import torch
    specs = [(wh.make_inflated_cube_mesh(5, 0.1, 0.3, 1.0, 2, outer_tag=7), "fp64"),
             (wh.make_cubed_cube_mesh(5, 2, 0.3), "fp64"),
             (wh.make_warped_uniform_hex(3, 3, 3, -1.0, 1.0, 0.05), "fp64"),
             (wh.make_cubed_ball_mesh(5, 2, outer_tag=7), "mixed")]
    singles, runs = [], []
    for mesh, prec in specs:
        x = mesh.coords.reshape(-1, 3) if hasattr(mesh, "coords") else wh.host_coords(mesh, wh.make_element(5))
        u0 = np.exp(-np.sum((x - 0.1) ** 2, axis=1) / 0.05)
        v0 = 0.1 * np.cos(math.pi * x[:, 2])
        for rep in range(2):
            ws = wh.Workspace(mesh, wh.make_element(5), precision=prec, sommerfeld_R=1.0)
            u, v = torch.from_numpy(u0.copy()).cuda(), torch.from_numpy(v0.copy()).cuda()
            dt = wh.recommended_dt_rk4(ws, 0.2)
            if rep == 0:
                E = ws.rk4_steps(u, v, dt, 6, energy_every=2)
                singles.append((u.cpu().numpy(), v.cpu().numpy(), E))
            else:
                runs.append((ws, u, v, dt))
    b = wh.Batch()
    slots = [b.add(ws, u, v, dt, 6, energy_every=2) for ws, u, v, dt in runs]
    fin = 0
    while fin < len(slots):
        fin += b.step(4)
    torch.cuda.synchronize()
    for s, (ws, u, v, _), (us, vs, Es) in zip(slots, runs, singles):
        _, E = b.slot_info(s)
        assert np.array_equal(u.cpu().numpy(), us) and np.array_equal(v.cpu().numpy(), vs)
        assert np.array_equal(E, Es)
    b.close()


@pytest.mark.parametrize("kind,prec", [("inflated_som", "fp64"), ("cubed_cube", "fp64"), ("ball", "fp64"),
                                       ("inflated_som", "mixed")])
def test_patch_mesh_rhs_matches_oracle(kind, prec):
    import wt_oracle as wo
    N = 5
    mesh = {"inflated_som": lambda: wh.make_inflated_cube_mesh(N, 0.1, 0.3, 1.0, 2, outer_tag=7),
            "cubed_cube": lambda: wh.make_cubed_cube_mesh(N, 2, 0.3),
            "ball": lambda: wh.make_cubed_ball_mesh(N, 2)}[kind]()
    bv = (0.3, -0.2, 0.1, 0.4, -0.5, 0.25)
    o = wo.Oracle.from_mesh(N, mesh.coords, mesh.neighbour, mesh.neighbour_face, mesh.bdry,
                            bdry_values=bv, sommerfeld_R=2.0)
    work = wh.Workspace(mesh, wh.make_element(N), bdry_values=bv, sommerfeld_R=2.0, precision=prec)
    rng = np.random.default_rng(7)
    u, ud = rng.standard_normal(o.n), rng.standard_normal(o.n)
    want = o.lap(u, ud)
    got = _rhs(work, u, ud, bv)
    tol = (1e-12 if prec == "fp64" else 2e-5) * np.max(np.abs(want))
    assert np.max(np.abs(got - want)) < tol
