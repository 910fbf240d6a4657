"""C-ABI checks that need no GPU: the library loads, exports every symbol
include/wt2.h declares, and rejects bad arguments with the documented codes
before touching a device."""
import ctypes
import os
import re

import numpy as np
import pytest

import wt_oracle as wo
import wavetoy_hip as wh
from wavetoy_hip import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    text = open(os.path.join(ROOT, "include", "wt2.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(wt2_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    L = _lib.lib()
    declared = _declared()
    assert declared, "no declarations parsed"
    for name in declared:
        assert hasattr(L, name), f"{name} declared in include/wt2.h but not exported"
    assert sorted(_lib.EXPORTS) == declared


def _c_param_counts():
    text = open(os.path.join(ROOT, "include", "wt2.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    out = {}
    for name, params in re.findall(r"\b(wt2_[a-z0-9_]+)\s*\(([^)]*)\)\s*;", text):
        params = params.strip()
        out[name] = 0 if params in ("", "void") else params.count(",") + 1
    return out


def test_julia_binding_matches_the_header():
    """Every ccall of the Julia binding names a declared entry with the
    declared number of arguments, and its BC3d mirrors wt2_bc3d field by
    field (no Julia toolchain here, so the binding is checked as text)."""
    jl = open(os.path.join(ROOT, "wavetoysecondorder.jl_amd", "julia", "WaveToyHIP.jl")).read()
    counts = _c_param_counts()
    calls = re.findall(r"ccall\(\(:(wt2_[a-z0-9_]+), libwt2\), \w+,\s*\(([^()]*(?:\{[^}]*\}[^()]*)*)\)", jl)
    assert len(calls) == jl.count("ccall(") and len(calls) >= 15
    for name, types in calls:
        assert name in counts, name
        n = len([t for t in re.split(r",(?![^{]*\})", types) if t.strip()])
        assert n == counts[name], (name, n, counts[name])
    body = re.search(r"struct BC3d\n(.*?)\nend", jl, flags=re.S).group(1)
    jl_types = re.findall(r"::\s*([A-Za-z{}0-9,]+)", body)
    c_types = [t for _, t in _lib.BC3d._fields_]
    want = {"NTuple{6,Int32}": ctypes.c_int32 * 6, "Int32": ctypes.c_int32, "Float64": ctypes.c_double,
            "Ptr{Float64}": ctypes.c_void_p}
    assert len(jl_types) == len(c_types)
    for jt, ct in zip(jl_types, c_types):
        w = want[jt]
        assert ctypes.sizeof(w) == ctypes.sizeof(ct) and ctypes.alignment(w) == ctypes.alignment(ct), (jt, ct)


def test_version_string():
    assert b"gfx950" in _lib.lib().wt2_version()


@pytest.mark.parametrize("N", [3, 4, 6, 8, 10])
def test_gll_matches_oracle(N):
    e = wh.make_element(N)
    xi, w, G = wo.gll(N)
    assert np.max(np.abs(e.xs - xi)) < 1e-15
    assert np.max(np.abs(e.w - w)) < 1e-15
    assert np.max(np.abs(e.G - G)) < 1e-12 * np.max(np.abs(G))


def _desc(**kw):
    d = _lib.MeshDesc()
    d.N = kw.get("N", 4)
    for a in range(3):
        d.M[a] = kw.get("M", 2)
        d.x0[a] = 0.0
        d.x1[a] = 1.0
        d.periodic[a] = 0
    d.tau = kw.get("tau", 1.5 * (d.N - 1) ** 2)
    d.eps_ko = kw.get("eps", 0.0)
    d.precision = kw.get("precision", 0)
    d.flags = kw.get("flags", 0)
    for f in range(6):
        d.face_kind[f] = kw.get("face_kind", (0,) * 6)[f]
    d.sommerfeld_inv_R = kw.get("inv_R", 0.0)
    if kw.get("periodic_x"):
        d.periodic[0] = 1
    d.warp_A = kw.get("warp_A", 0.0)
    return d


def _create(d):
    ctx = ctypes.c_void_p()
    rc = _lib.lib().wt2_ctx_create(ctypes.byref(d), ctypes.byref(ctx))
    return rc, ctx


@pytest.mark.parametrize("kw,code", [
    (dict(N=2), _lib.WT2_ERR_SHAPE),
    (dict(N=11), _lib.WT2_ERR_SHAPE),
    (dict(M=0), _lib.WT2_ERR_SHAPE),
    (dict(M=65), _lib.WT2_ERR_UNSUPPORTED),
    (dict(precision=7), _lib.WT2_ERR_ARG),
    (dict(tau=-1.0), _lib.WT2_ERR_ARG),
    (dict(eps=-0.1), _lib.WT2_ERR_ARG),
    (dict(N=8, tau=20.0), _lib.WT2_ERR_TAU),   # literal tau = 20 is not coercive at N = 8
    (dict(N=10, tau=20.0), _lib.WT2_ERR_TAU),
    (dict(face_kind=(0, 3, 0, 0, 0, 0)), _lib.WT2_ERR_ARG),          # only 0 and 7 are face kinds
    (dict(face_kind=(7, 0, 0, 0, 0, 0), periodic_x=True), _lib.WT2_ERR_ARG),
    (dict(face_kind=(7,) * 6, inv_R=-1.0), _lib.WT2_ERR_ARG),
])
def test_ctx_create_rejects(kw, code):
    rc, ctx = _create(_desc(**kw))
    assert rc == code
    assert not ctx.value
    assert len(_lib.lib().wt2_last_error()) > 0


def test_null_pointers_rejected():
    L = _lib.lib()
    assert L.wt2_ctx_create(None, None) == _lib.WT2_ERR_ARG
    assert L.wt2_gll(4, None, None, None) == _lib.WT2_ERR_ARG
    assert L.wt2_rhs3d(None, None, None, None, None, 0) == _lib.WT2_ERR_ARG
    assert L.wt2_rk4_steps(None, None, None, 0.1, 1, 0, None, 0) == _lib.WT2_ERR_ARG
    assert L.wt2_energy(None, None, None, None, 0) == _lib.WT2_ERR_ARG
    assert L.wt2_ctx_info(None, None) == _lib.WT2_ERR_ARG
    assert L.wt2_fo_rhs(None, None, None, None, None, None, 0.1, None, 0) == _lib.WT2_ERR_ARG
    assert L.wt2_fo_energy(None, None, None, None, None, 0) == _lib.WT2_ERR_ARG
    assert L.wt2_fo_coef(None, None, None, 0) == _lib.WT2_ERR_ARG


def test_host_descriptors_validate_their_arguments():
    import wavetoy_hip as wh
    mesh = wh.make_uniform_hex(2, 0.0, 1.0)
    elem = wh.make_element(4)
    d = wh.mesh_desc(mesh, elem, sommerfeld_faces=(2, 5), sommerfeld_R=4.0)
    assert list(d.face_kind) == [0, 7, 0, 0, 7, 0] and d.sommerfeld_inv_R == 0.25
    assert wh.mesh_desc(mesh, elem).sommerfeld_inv_R == 0.0
    with pytest.raises(ValueError):
        wh.mesh_desc(mesh, elem, sommerfeld_faces=(7,))
    with pytest.raises(ValueError):
        wh.mesh_desc(mesh, elem, sommerfeld_faces=(1,), sommerfeld_R=0.0)
    bc = wh.make_bc3d(["sommerfeld", "dirichlet", "excision", "full_dirichlet", 2, 1], sigma=0.5)
    assert bc.kinds == (2, 1, 3, 4, 2, 1) and bc.sigma == 0.5
    with pytest.raises(KeyError):
        wh.make_bc3d(["absorbing"] * 6)
    with pytest.raises(ValueError):
        wh.make_bc3d(["sommerfeld"] * 5)


def test_host_coords_match_oracle():
    for A in (0.0, 0.05):
        mesh = (wh.make_warped_uniform_hex(3, 2, 2, 0.0, 1.0, A) if A
                else wh.make_uniform_hex((3, 2, 2), 0.0, 1.0))
        e = wh.make_element(5)
        x = wh.host_coords(mesh, e)
        o = wo.Oracle(5, (3, 2, 2), warp_A=A)
        assert np.max(np.abs(x - o.coords())) < 1e-14


def test_tau_effective_matches_reference_driver_scaling():
    # src/evolve.jl:964-968: tau = 1.5 (N-1)^2 on cubical meshes
    assert wh.tau_effective(20.0, 8) == 73.5
    assert wh.tau_effective(20.0, 4) == 20.0
    assert wh.tau_effective(None, 10) == 121.5


def _ball(N=3, M=1, tag=1):
    return wh.make_cubed_ball_mesh(N, M, R=1.0, outer_tag=tag)


def _create_mesh(m, eps=0.0, orientation=None, nbr=None, nface=None, bdry=None):
    d = _desc(N=m.N, eps=eps)
    ctx = ctypes.c_void_p()
    arr = lambda a: np.ascontiguousarray(a, dtype=np.int32)  # noqa: E731
    nbr, nface, bdry = arr(m.neighbour if nbr is None else nbr), arr(m.neighbour_face if nface is None else nface), \
        arr(m.bdry if bdry is None else bdry)
    ori = None if orientation is None else arr(orientation)
    rc = _lib.lib().wt2_ctx_create_mesh(ctypes.byref(d), m.Ne, m.coords.ctypes.data, nbr.ctypes.data,
                                        nface.ctypes.data, None if ori is None else ori.ctypes.data,
                                        bdry.ctypes.data, ctypes.byref(ctx))
    if rc == 0:
        _lib.lib().wt2_ctx_destroy(ctx)
    return rc


def test_cubed_ball_mesh_connectivity():
    m = wh.make_cubed_ball_mesh(4, 2, R=2.0, outer_tag=7)
    assert m.Ne == 7 * 8
    r = np.linalg.norm(m.coords.reshape(-1, 3), axis=1)
    assert abs(r.max() - 2.0) < 1e-14
    outer = m.neighbour == 0
    assert outer.sum() == 6 * 4 and np.all(m.bdry[outer] == 7) and np.all(m.bdry[~outer] == 0)
    e, f = np.nonzero(~outer)
    en, fn = m.neighbour[e, f] - 1, m.neighbour_face[e, f] - 1
    assert np.all(m.neighbour[en, fn] == e + 1) and np.all(m.neighbour_face[en, fn] == f + 1)
    # the outer faces lie on the sphere
    for e, f in zip(*np.nonzero(outer)):
        x = wh.core._face_nodes(m.coords[e:e + 1], f).reshape(-1, 3)
        assert np.max(np.abs(np.linalg.norm(x, axis=1) - 2.0)) < 1e-14


def test_ctx_create_mesh_validates_before_touching_a_device():
    m = _ball()
    ok = (0, _lib.WT2_ERR_HIP)        # valid input: created, or no device here
    assert _create_mesh(m) in ok
    assert _create_mesh(m, eps=0.1) == _lib.WT2_ERR_UNSUPPORTED
    bad = m.bdry.copy()
    bad[m.neighbour == 0] = 9
    assert _create_mesh(m, bdry=bad) == _lib.WT2_ERR_ARG
    nb = m.neighbour.copy()
    e, f = map(int, np.argwhere(nb > 0)[0])
    nb[e, f] = m.Ne + 1
    assert _create_mesh(m, nbr=nb) == _lib.WT2_ERR_ARG
    # a face named as the neighbour of a face it does not touch
    nf = m.neighbour_face.copy()
    nf[e, f] = nf[e, f] % 6 + 1
    assert _create_mesh(m, nface=nf) == _lib.WT2_ERR_ARG
    assert "does not meet" in _lib.lib().wt2_last_error().decode() or \
        "name each other" in _lib.lib().wt2_last_error().decode()
    # orientations that are not each other's inverse
    ori = np.zeros((m.Ne, 6), dtype=np.int32)
    ori[e, f] = 2
    assert _create_mesh(m, orientation=ori) == _lib.WT2_ERR_ARG
    ori[e, f] = 8
    assert _create_mesh(m, orientation=ori) == _lib.WT2_ERR_ARG
    # 67 elements (prime > 64) cannot be tiled as M0 x M1 x M2
    m67 = wh.HexMesh(3, np.tile(m.coords[:1], (67, 1, 1, 1, 1)), np.zeros((67, 6)), np.zeros((67, 6)),
                     np.ones((67, 6)))
    assert _create_mesh(m67) == _lib.WT2_ERR_UNSUPPORTED
    with pytest.raises(ValueError):
        wh.Workspace(m, wh.make_element(3), sommerfeld_faces=(1,))
    with pytest.raises(ValueError):
        wh.Workspace(m, wh.make_element(4))


@pytest.mark.parametrize("make,patches,tags", [
    (lambda: wh.make_cubed_cube_mesh(3, 2, 0.3), 7, list(range(7))),
    (lambda: wh.make_inflated_cube_mesh(3, 0.1, 0.3, 1.0, 2), 13, [0, 1]),
])
def test_patch_mesh_families(make, patches, tags):
    # test/dirichlet_cubed_cube_tests.jl:21-32, test/dirichlet_cubed_sphere_tests.jl:44-82
    m = make()
    assert m.Ne == patches * 8
    assert sorted(set(m.bdry.reshape(-1).tolist())) == tags
    if patches == 13:
        for e, f in zip(*np.nonzero(m.neighbour == 0)):
            x = wh.core._face_nodes(m.coords[e:e + 1], f).reshape(-1, 3)
            assert np.max(np.abs(np.linalg.norm(x, axis=1) - 1.0)) < 1e-12
    assert _create_mesh(m) in (0, _lib.WT2_ERR_HIP)
