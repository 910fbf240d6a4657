"""Host-side mirror of the reference's hot-path API over libwt2.so.

Names follow eschnett/WaveToySecondOrder.jl (Julia's `f!` becomes `f`; the
first argument is still the one that is overwritten):

    Params3d                    src/wave.jl:67-95
    make_element/make_operators HexSBPSAT (used at src/wave.jl:603)
    make_uniform_hex            HexMeshes (src/evolve.jl:915)
    make_warped_uniform_hex     HexMeshes (test/dirichlet_warped_tests.jl:23,40)
    make_geometry/make_workspace/to_device   src/evolve.jl:935-936
    rhs_wave3d                  rhs_wave3d! (src/wave.jl:573-611)
    wave3d_rhs_full             (u, u_t) analogue of wave3d_curved_rhs! (src/wave3d_curved.jl:205-258)
    AnalyticBackground3D, sample_background3d, make_coef3d   src/wave3d_curved.jl:36-124
    discrete_energy             test/kernels3d_tests.jl:32-47
    recommended_dt_rk4          dt = cfl * h_min (src/evolve.jl:1207-1213)

Field arrays are either torch CUDA tensors (used in place on the device) or
numpy float64 arrays (copied in and out); both hold the Julia column-major
(N,N,N,Ne) layout, i.e. a C-contiguous (Ne,N,N,N) block, flattened or not.
"""
import ctypes
import math

import numpy as np

from . import _lib
from ._lib import WT2_FP64, WT2_MIXED, WT2_HOST_PTRS, WT2_ALLOW_NONCOERCIVE, WT2Error, check, lib

SOMMERFELD_BDRY_TAG = 7     # src/wave.jl:23

try:  # torch is plumbing for device memory; numpy host arrays work without it
    import torch
except ImportError:  # pragma: no cover
    torch = None


# ----------------------------------------------------------------- params

class Params3d:
    """Params3d(A, k, omega, tau, bdry_values, sommerfeld_R=inf) (src/wave.jl:67-95)."""

    def __init__(self, A=1.0, k=(0.0, 0.0, 0.0), omega=0.0, tau=None, bdry_values=(0.0,) * 6,
                 sommerfeld_R=math.inf):
        self.A = float(A)
        self.k = tuple(float(x) for x in k)
        self.omega = float(omega)
        self.tau = None if tau is None else float(tau)
        self.bdry_values = tuple(float(x) for x in bdry_values)
        if len(self.bdry_values) != 6:
            raise ValueError("bdry_values must have 6 entries (one per boundary tag 1..6)")
        self.sommerfeld_R = float(sommerfeld_R)


def tau_effective(tau, N):
    """Penalty actually used: max(tau, 1.5 (N-1)^2), the reference driver's
    scaling for cubical meshes (src/evolve.jl:964-968), which keeps SIPG coercive."""
    t = 1.5 * (N - 1) ** 2
    return t if tau is None else max(float(tau), t)


# ---------------------------------------------------------------- element

class Element:
    def __init__(self, N):
        xi = np.zeros(N); w = np.zeros(N); G = np.zeros((N, N))
        check(lib().wt2_gll(N, xi.ctypes.data, w.ctypes.data, G.ctypes.data))
        self.N = N
        self.xs = xi
        self.w = w
        self.G = G
        self.h = float(np.min(np.diff(xi)))   # smallest node gap on the unit element


class SBPOps:
    def __init__(self, elem):
        self.N = elem.N
        self.G = elem.G
        self.H = np.diag(elem.w)


def make_element(N):
    return Element(N)


def make_operators(elem):
    return SBPOps(elem)


# ------------------------------------------------------------------- mesh

class Mesh:
    def __init__(self, M, x0, x1, periodic, warp_A=0.0):
        self.M = tuple(int(m) for m in M)
        self.x0 = tuple(float(v) for v in x0)
        self.x1 = tuple(float(v) for v in x1)
        self.periodic = tuple(bool(p) for p in periodic)
        self.warp_A = float(warp_A)
        self.Ne = self.M[0] * self.M[1] * self.M[2]


def _triple(v):
    if isinstance(v, (tuple, list)):
        return tuple(v)
    return (v, v, v)


def make_uniform_hex(M, x0=0.0, x1=1.0, periodic=False):
    """M^3 (or Mx,My,Mz) affine hex elements on [x0,x1]^3, x-fastest element order,
    outer face f tagged f (1..6 = -x,+x,-y,+y,-z,+z)."""
    return Mesh(_triple(M), _triple(x0), _triple(x1), _triple(periodic))


def make_warped_uniform_hex(Mx, My, Mz, x0, x1, A, periodic=False):
    """Warped cube: x = x0 + L (xi + A sin(2 pi xi1) sin(2 pi xi2) sin(2 pi xi3))."""
    return Mesh((Mx, My, Mz), _triple(x0), _triple(x1), _triple(periodic), warp_A=A)


def _face_nodes(X, f):
    """Nodes of face f (0..5 = -x,+x,-y,+y,-z,+z) of element arrays X
    (Ne, N, N, N, 3) indexed [e, k, j, i], as (Ne, q, p, 3): the face's (p, q)
    run over the two tangential axes in increasing order, as in the
    reference's face extraction (src/wave_strong_rhs.jl:41-73)."""
    ax, r = f // 2, (0 if f % 2 == 0 else -1)
    if ax == 0:
        return X[:, :, :, r, :]
    if ax == 1:
        return X[:, :, r, :, :]
    return X[:, r, :, :, :]


class HexMesh:
    """A hex mesh given by node coordinates and face connectivity: the
    reference's MeshGeometry.coords and MeshConnectivity fields (neighbour,
    neighbour_face, orientation, bdry; 1-based, column-major (6, Ne), here as
    (Ne, 6) arrays), for meshes that are not one uniform block, such as the
    multi-patch cubed ball.  coords: (Ne, N, N, N, 3) or (N^3 Ne, 3), node
    index i + N (j + N (k + N e)).  orientation None: found from the face
    nodes' coordinates by the library."""

    M = (1, 1, 1)
    x0 = (0.0, 0.0, 0.0)
    x1 = (1.0, 1.0, 1.0)
    periodic = (False, False, False)
    warp_A = 0.0

    def __init__(self, N, coords, neighbour, neighbour_face, bdry, orientation=None):
        self.N = int(N)
        X = np.ascontiguousarray(coords, dtype=np.float64)
        if X.size % (3 * self.N ** 3):
            raise ValueError("coords must hold 3 N^3 values per element")
        self.Ne = X.size // (3 * self.N ** 3)
        self.coords = X.reshape(self.Ne, self.N, self.N, self.N, 3)

        def tab(a):
            t = np.ascontiguousarray(a, dtype=np.int32)
            if t.shape != (self.Ne, 6):
                raise ValueError("connectivity arrays are (Ne, 6)")
            return t
        self.neighbour = tab(neighbour)
        self.neighbour_face = tab(neighbour_face)
        self.bdry = tab(bdry)
        self.orientation = None if orientation is None else tab(orientation)


def connect_by_coords(N, coords, outer_tag=1):
    """neighbour, neighbour_face and bdry of elements that meet face to face:
    two faces are shared when their node centroids coincide (to 1e-9 of the
    smallest element); every other face is an outer face with outer_tag."""
    from scipy.spatial import cKDTree
    X = np.asarray(coords, dtype=np.float64).reshape(-1, N, N, N, 3)
    Ne = X.shape[0]
    cen = np.stack([_face_nodes(X, f).mean(axis=(1, 2)) for f in range(6)], axis=1).reshape(-1, 3)
    span = (X.reshape(Ne, -1, 3).max(axis=1) - X.reshape(Ne, -1, 3).min(axis=1)).max(axis=1)
    tol = 1e-9 * float(span.min())
    nbr = np.zeros((Ne, 6), dtype=np.int32)
    nface = np.zeros((Ne, 6), dtype=np.int32)
    bdry = np.full((Ne, 6), int(outer_tag), dtype=np.int32)
    for a, b in cKDTree(cen).query_pairs(tol):
        ea, fa, eb, fb = a // 6, a % 6, b // 6, b % 6
        if nbr[ea, fa] or nbr[eb, fb]:
            raise ValueError("a face meets more than one other face")
        nbr[ea, fa], nface[ea, fa], nbr[eb, fb], nface[eb, fb] = eb + 1, fb + 1, ea + 1, fa + 1
    bdry[nbr > 0] = 0
    return nbr, nface, bdry


def _cube_surface(a):
    """Point (s, t) of the face (axis d, sign sg) of the cube [-a, a]^3, in the
    patch frame (components along axes d+1, d+2, d)."""
    return lambda S, T, sg: (a * S, a * T, sg * a * np.ones_like(S))


def _sphere_surface(r):
    """Point (s, t) of the sphere of radius r over the cube face (axis d, sign
    sg), equiangular: direction (tan(pi s / 4), tan(pi t / 4), sg)."""
    def f(S, T, sg):
        tu, tv = np.tan(0.25 * np.pi * S), np.tan(0.25 * np.pi * T)
        k = r / np.sqrt(1.0 + tu * tu + tv * tv)
        return k * tu, k * tv, sg * k
    return f


def _patch_mesh(N, M, a, layers, core=True):
    """Coordinates (Ne, N, N, N, 3) of a patch mesh: the cube [-a, a]^3 of M^3
    affine elements (x fastest; left out with core=False), then for each layer
    (inner, outer surface) six patches of M^3 elements, one per cube face
    (axis d, sign -, +), with local axes (s, t, rho): s, t in [-1, 1] along
    axes d+1, d+2 (cyclic), rho in [0, 1] outward,
    x = (1 - rho) inner(s, t) + rho outer(s, t)."""
    xs = make_element(N).xs
    g = (np.arange(M)[:, None] + xs[None, :]) / M          # (element, node) -> [0, 1]
    els = []
    c = -a + 2.0 * a * g
    for ez in range(M if core else 0):
        for ey in range(M):
            for ex in range(M):
                Z, Y, Xc = np.meshgrid(c[ez], c[ey], c[ex], indexing="ij")
                els.append(np.stack([Xc, Y, Z], axis=-1))
    st = -1.0 + 2.0 * g
    for inner, outer in layers:
        for d in range(3):
            p1, p2 = (d + 1) % 3, (d + 2) % 3
            for sg in (-1.0, 1.0):
                for er in range(M):
                    for et in range(M):
                        for es in range(M):
                            RHO, T, S = np.meshgrid(g[er], st[et], st[es], indexing="ij")
                            pi_, po = inner(S, T, sg), outer(S, T, sg)
                            x = np.empty(S.shape + (3,))
                            for comp, ax in enumerate((p1, p2, d)):
                                x[..., ax] = (1 - RHO) * pi_[comp] + RHO * po[comp]
                            els.append(x)
    return np.stack(els)


def _tag_cube_sides(mesh, half):
    """Outer faces on the cube [-half, half]^3 get the reference's tags 1..6
    (-x, +x, -y, +y, -z, +z)."""
    for e, f in zip(*np.nonzero(mesh.neighbour == 0)):
        cen = _face_nodes(mesh.coords[e:e + 1], f).reshape(-1, 3).mean(axis=0)
        ax = int(np.argmax(np.abs(cen)))
        mesh.bdry[e, f] = 2 * ax + (2 if cen[ax] > 0 else 1)
    return mesh


def make_cubed_cube_mesh(N, M, R=0.3):
    """The cube [-1, 1]^3 as 7 patches of M^3 elements (make_cubed_cube_mesh,
    src/evolve.jl:916-918; test/dirichlet_cubed_cube_tests.jl): the inner cube
    [-R, R]^3 and six wedges between its faces and the outer cube's.  Outer
    faces carry tags 1..6 by side (-x, +x, -y, +y, -z, +z)."""
    X = _patch_mesh(N, M, R, [(_cube_surface(R), _cube_surface(1.0))])
    return _tag_cube_sides(HexMesh(N, X, *connect_by_coords(N, X)), 1.0)


def make_inflated_cube_mesh(N, L=0.1, R1=0.3, R2=1.0, M=2, outer_tag=1):
    """The ball |x| <= R2 as 13 patches of M^3 elements
    (make_inflated_cube_mesh, src/evolve.jl:919-921;
    test/dirichlet_cubed_sphere_tests.jl): the cube [-L, L]^3, six inflation
    patches from its faces to the sphere R1, six shells from R1 to R2.  The
    outer faces lie on the sphere R2 and carry outer_tag (1 Dirichlet, 7
    Sommerfeld)."""
    X = _patch_mesh(N, M, L, [(_cube_surface(L), _sphere_surface(R1)),
                              (_sphere_surface(R1), _sphere_surface(R2))])
    return HexMesh(N, X, *connect_by_coords(N, X, outer_tag))


def make_cubed_ball_mesh(N, M, R=1.0, a=None, outer_tag=1):
    """The ball |x| <= R as 7 patches of M^3 elements: the cube [-a, a]^3
    (default a = R/2) and six patches from its faces to the sphere R.  The
    patches meet with orientation != 0; outer faces carry outer_tag."""
    a = 0.5 * R if a is None else float(a)
    X = _patch_mesh(N, M, a, [(_cube_surface(a), _sphere_surface(R))])
    return HexMesh(N, X, *connect_by_coords(N, X, outer_tag))


EXCISION_BDRY_TAG = 8    # excision / pure-outflow outer faces: no SAT (src/wave_strong_rhs.jl:188-213)


def make_radial_shell_mesh(N, R1=0.5, R2=1.0, M=2, inner_tag=EXCISION_BDRY_TAG, outer_tag=7):
    """The spherical shell R1 <= |x| <= R2 as 6 patches of M^3 elements
    (make_radial_shell_mesh, src/evolve.jl:922-930): one per cube face,
    equiangular in the angles, linear in the radius.  The inner sphere carries
    inner_tag (8: excision, no SAT) and the outer one outer_tag (7:
    Sommerfeld, 1: Dirichlet)."""
    X = _patch_mesh(N, M, R1, [(_sphere_surface(R1), _sphere_surface(R2))], core=False)
    nbr, nface, bdry = connect_by_coords(N, X, outer_tag)
    mesh = HexMesh(N, X, nbr, nface, bdry)
    for e, f in zip(*np.nonzero(mesh.neighbour == 0)):
        r = np.linalg.norm(_face_nodes(mesh.coords[e:e + 1], f).reshape(-1, 3), axis=1).mean()
        mesh.bdry[e, f] = inner_tag if abs(r - R1) < abs(r - R2) else outer_tag
    return mesh


# -------------------------------------------------------- device workspace

def mesh_desc(mesh, elem, tau=None, eps_ko=0.0, bdry_values=(0.0,) * 6, precision="fp64",
              allow_noncoercive=False, sommerfeld_faces=(), sommerfeld_R=math.inf):
    """wt2_mesh_desc of a mesh: sommerfeld_faces lists the outer faces (1..6 =
    -x,+x,-y,+y,-z,+z) that carry the Sommerfeld tag (SOMMERFELD_BDRY_TAG = 7,
    src/wave.jl:23) instead of a Dirichlet value; sommerfeld_R is the radius of
    their u/R term (inf: plane-wave condition)."""
    d = _lib.MeshDesc()
    d.N = elem.N
    for a in range(3):
        d.M[a] = mesh.M[a]
        d.x0[a] = mesh.x0[a]
        d.x1[a] = mesh.x1[a]
        d.periodic[a] = int(mesh.periodic[a])
    d.warp_A = mesh.warp_A
    d.tau = tau_effective(None, elem.N) if tau is None else float(tau)
    d.eps_ko = float(eps_ko)
    for f in range(6):
        d.bdry_values[f] = float(bdry_values[f])
    d.precision = {"fp64": WT2_FP64, "mixed": WT2_MIXED}[precision]
    d.flags = WT2_ALLOW_NONCOERCIVE if allow_noncoercive else 0
    som = tuple(int(f) for f in sommerfeld_faces)
    if any(f < 1 or f > 6 for f in som):
        raise ValueError("sommerfeld_faces are face numbers 1..6")
    for f in range(6):
        d.face_kind[f] = SOMMERFELD_BDRY_TAG if (f + 1) in som else 0
    R = float(sommerfeld_R)
    if not R > 0:
        raise ValueError("sommerfeld_R must be > 0")
    d.sommerfeld_inv_R = 0.0 if math.isinf(R) else 1.0 / R
    return d


def ctx_footprint(mesh, elem, precision="fp64", with_background=False):
    """Device bytes a Workspace on this mesh allocates (wt2_ctx_footprint),
    computed without touching the device."""
    d = mesh_desc(mesh, elem, precision=precision)
    b = ctypes.c_int64()
    check(lib().wt2_ctx_footprint(ctypes.byref(d), int(bool(with_background)), ctypes.byref(b)))
    return b.value


# pass kernel kinds: LDS-staged tile; LDS-free with shuffles; LDS-free with the
# neighbour exchange through LDS
KERNEL_KINDS = ("lds", "shfl", "xchg")


class Workspace:
    """A libwt2 context: mesh + operators + geometry + device workspace + stream
    (the reference's geom/ops/work bundle after to_device)."""

    def __init__(self, mesh, elem, tau=None, eps_ko=0.0, bdry_values=(0.0,) * 6,
                 precision="fp64", allow_noncoercive=False, bg=None, sommerfeld_faces=(),
                 sommerfeld_R=math.inf):
        d = mesh_desc(mesh, elem, tau, eps_ko, bdry_values, precision, allow_noncoercive,
                      sommerfeld_faces, sommerfeld_R)
        self.desc = d
        self.mesh = mesh
        self.elem = elem
        self.N = elem.N
        self.precision = precision
        self.ctx = ctypes.c_void_p()
        if isinstance(mesh, HexMesh):
            if mesh.N != elem.N:
                raise ValueError("mesh coordinates are for another N")
            if sommerfeld_faces:
                raise ValueError("on a HexMesh, Sommerfeld faces are the outer faces with bdry tag 7")
            ori = None if mesh.orientation is None else mesh.orientation.ctypes.data
            check(lib().wt2_ctx_create_mesh(ctypes.byref(d), mesh.Ne, mesh.coords.ctypes.data,
                                            mesh.neighbour.ctypes.data, mesh.neighbour_face.ctypes.data,
                                            ori, mesh.bdry.ctypes.data, ctypes.byref(self.ctx)))
        else:
            check(lib().wt2_ctx_create(ctypes.byref(d), ctypes.byref(self.ctx)))
        self.npoints = elem.N ** 3 * mesh.Ne
        if bg is not None:
            self.set_background(bg)

    def close(self):
        if getattr(self, "ctx", None) and self.ctx.value:
            lib().wt2_ctx_destroy(self.ctx)
            self.ctx = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def info(self):
        inf = _lib.Info()
        check(lib().wt2_ctx_info(self.ctx, ctypes.byref(inf)))
        return {k: getattr(inf, k) for k, _ in inf._fields_}

    @property
    def mu(self):
        return self.info()["mu"]

    @property
    def h_min(self):
        return self.info()["h_min"]

    @property
    def tau(self):
        return self.desc.tau

    def set_background(self, bg10):
        bg = np.ascontiguousarray(bg10, dtype=np.float64).reshape(-1)
        if bg.size != 10 * self.npoints:
            raise ValueError("background needs 10 values per node")
        check(lib().wt2_ctx_set_background(self.ctx, bg.ctypes.data))

    def coords(self):
        x = np.zeros((self.npoints, 3))
        check(lib().wt2_coords(self.ctx, x.ctypes.data))
        return x

    # -- pointer plumbing
    def _ptr(self, a, writable=False, mult=1):
        if a is None:
            return None, False
        if torch is not None and isinstance(a, torch.Tensor):
            if a.dtype != torch.float64 or not a.is_contiguous() or a.numel() != mult * self.npoints:
                raise ValueError(f"field tensors must be contiguous float64 with {mult} x N^3*Ne entries")
            if not a.is_cuda:
                raise ValueError("torch field tensors must live on the GPU")
            return a.data_ptr(), False
        if isinstance(a, np.ndarray):
            if a.dtype != np.float64 or not a.flags["C_CONTIGUOUS"] or a.size != mult * self.npoints:
                raise ValueError(f"field arrays must be C-contiguous float64 with {mult} x N^3*Ne entries")
            if writable and not a.flags["WRITEABLE"]:
                raise ValueError("output array is read-only")
            return a.ctypes.data, True
        raise TypeError("fields must be torch CUDA tensors or numpy arrays")

    def _ptrs(self, *arrs, writable=(), mult=None):
        ptrs, hosts = [], set()
        for i, a in enumerate(arrs):
            p, h = self._ptr(a, i in writable, (mult or {}).get(i, 1))
            ptrs.append(p)
            if a is not None:
                hosts.add(h)
        if len(hosts) > 1:
            raise ValueError("mix of host and device arrays in one call")
        return ptrs, (WT2_HOST_PTRS if hosts == {True} else 0)

    def _sync_torch(self):
        if torch is not None and torch.cuda.is_available() and torch.cuda.is_initialized():
            torch.cuda.current_stream().synchronize()

    def rhs(self, udd, u, ud, bdry_values):
        (pa, pu, pud), fl = self._ptrs(udd, u, ud, writable=(0,))
        bv = (ctypes.c_double * 6)(*bdry_values)
        self._sync_torch()
        check(lib().wt2_rhs3d(self.ctx, pa, pu, pud, ctypes.cast(bv, ctypes.c_void_p), fl))
        return udd

    def rhs_full(self, du, dud, u, ud):
        (p1, p2, p3, p4), fl = self._ptrs(du, dud, u, ud, writable=(0, 1))
        self._sync_torch()
        check(lib().wt2_rhs_full(self.ctx, p1, p2, p3, p4, fl))
        return du, dud

    def rk4_steps(self, u, ud, dt, nsteps, energy_every=0, graph=True):
        (pu, pud), fl = self._ptrs(u, ud, writable=(0, 1))
        ne = 1 + nsteps // energy_every if energy_every > 0 else 0
        E = np.zeros(max(ne, 1))
        flags = fl | (0 if graph else _lib.WT2_NO_GRAPH)
        self._sync_torch()
        check(lib().wt2_rk4_steps(self.ctx, pu, pud, float(dt), int(nsteps), int(energy_every),
                                  E.ctypes.data if ne else None, flags))
        return E[:ne] if ne else None

    def rk4_steps_noalloc(self, u, ud, dt, nsteps, energy_every=0):
        """rk4_steps for callers that cannot allocate device memory
        (wt2_rk4_steps_noalloc): device tensors only, kernels launched one by
        one, the energy history in the library's fixed device array (at most
        WT2_NOALLOC_MAX_ENERGIES samples).  Calls that overlap in time on one
        GPU race."""
        (pu, pud), fl = self._ptrs(u, ud, writable=(0, 1))
        if fl:
            raise ValueError("rk4_steps_noalloc needs device tensors")
        ne = 1 + nsteps // energy_every if energy_every > 0 else 0
        E = np.zeros(max(ne, 1))
        self._sync_torch()
        check(lib().wt2_rk4_steps_noalloc(self.ctx, pu, pud, float(dt), int(nsteps), int(energy_every),
                                          E.ctypes.data if ne else None))
        return E[:ne] if ne else None

    def set_stream(self, stream_handle):
        """Run this context's kernels on an external HIP stream (int handle, e.g.
        torch.cuda.current_stream().cuda_stream); None restores its own."""
        check(lib().wt2_set_stream(self.ctx, stream_handle))

    def launch_pass(self, axis, stage, u, ud, dt):
        """Diagnostics: enqueue one pencil pass (axis 0..2) or the curved
        operator (axis -1) of RK4 stage `stage` without synchronising."""
        (pu, pud), fl = self._ptrs(u, ud)
        if fl:
            raise ValueError("launch_pass needs device tensors")
        check(lib().wt2_launch_pass(self.ctx, int(axis), int(stage), pu, pud, float(dt)))

    def autotune(self, u, ud, dt, force=False):
        """Pick the fastest pass kernel / tile size per axis on this state (device
        tensors); cached per shape for later workspaces."""
        (pu, pud), fl = self._ptrs(u, ud)
        if fl:
            raise ValueError("autotune needs device tensors")
        self._sync_torch()
        check(lib().wt2_ctx_autotune(self.ctx, pu, pud, float(dt), 1 if force else 0))
        return self.tiles()

    def tiles(self):
        nb = (ctypes.c_int32 * 3)()
        d = (ctypes.c_int32 * 3)()
        check(lib().wt2_ctx_tiles(self.ctx, ctypes.cast(nb, ctypes.c_void_p), ctypes.cast(d, ctypes.c_void_p)))
        return {"planes_per_tile": tuple(nb), "kernel": tuple(KERNEL_KINDS[x] for x in d)}

    def set_tiles(self, planes_per_tile, kernel):
        """Choose the pass kernel ("lds", "shfl" or "xchg", or 0/1/2) and the
        element planes per tile of each axis."""
        nb = (ctypes.c_int32 * 3)(*[int(x) for x in planes_per_tile])
        d = (ctypes.c_int32 * 3)(*[KERNEL_KINDS.index(x) if isinstance(x, str) else int(x) for x in kernel])
        check(lib().wt2_ctx_set_tiles(self.ctx, ctypes.cast(nb, ctypes.c_void_p), ctypes.cast(d, ctypes.c_void_p)))

    def energy(self, u, ud):
        (pu, pud), fl = self._ptrs(u, ud)
        E = ctypes.c_double()
        self._sync_torch()
        check(lib().wt2_energy(self.ctx, pu, pud, ctypes.byref(E), fl))
        return E.value

    # -- the first-order (Phi, Pi) system (affine meshes)
    def fo_coef(self, bg10, out=None):
        """sample_background3d! (src/wave3d_curved.jl:75-115) on the device:
        the 11 coefficient fields (alpha, sqrtg, b1..3, gu11, gu12, gu13, gu22,
        gu23, gu33), each npoints long, one after the other, from 10 background
        values per node.  bg10: (npoints, 10) host array or device tensor."""
        if out is None:
            if torch is not None and isinstance(bg10, torch.Tensor):
                out = torch.empty(11 * self.npoints, dtype=torch.float64, device=bg10.device)
            else:
                out = np.empty(11 * self.npoints)
        if isinstance(bg10, np.ndarray):
            bg10 = np.ascontiguousarray(bg10, dtype=np.float64).reshape(-1)
        (pb, po), fl = self._ptrs(bg10, out, writable=(1,), mult={0: 10, 1: 11})
        self._sync_torch()
        check(lib().wt2_fo_coef(self.ctx, pb, po, fl))
        return out

    def fo_rhs(self, dPhi, dPi, Phi, Pi, coef, eps_ko=0.1, bc3d=None):
        """wave3d_curved_rhs!(Phidot, Pidot, Phi, Pi, coef; eps_KO, bc3d) on the
        affine path (src/wave3d_curved.jl:205-258)."""
        g = (None, None, None) if bc3d is None else (bc3d.gPhi_arr, bc3d.gPi_arr, bc3d.gD_arr)
        (p1, p2, p3, p4, pc, pg1, pg2, pg3), fl = self._ptrs(dPhi, dPi, Phi, Pi, coef, g[0], g[1], g[2],
                                                             writable=(0, 1), mult={4: 11, 7: 3})
        bc = None
        if bc3d is not None:
            bc = _lib.BC3d()
            for f in range(6):
                bc.kinds[f] = bc3d.kinds[f]
            bc.excision_tag = bc3d.excision_tag
            bc.sigma = bc3d.sigma
            bc.gPhi = pg1
            bc.gPi = pg2
            bc.gD = pg3
        self._sync_torch()
        check(lib().wt2_fo_rhs(self.ctx, p1, p2, p3, p4, pc, float(eps_ko),
                               None if bc is None else ctypes.byref(bc), fl))
        return dPhi, dPi

    def fo_rk4_stage(self, stage, Phi, Pi, coef, dt, eps_ko=0.1, bc3d=None):
        """One fused RK4 stage of the first-order system in place on (Phi, Pi),
        CUDA float64 tensors (wt2_fo_rk4_stage)."""
        g = (None, None, None) if bc3d is None else (bc3d.gPhi_arr, bc3d.gPi_arr, bc3d.gD_arr)
        (p0, p1, pc, pg1, pg2, pg3), fl = self._ptrs(Phi, Pi, coef, g[0], g[1], g[2], writable=(0, 1),
                                                     mult={2: 11, 5: 3})
        if fl:
            raise ValueError("fo_rk4_stage needs device tensors")
        bc = None
        if bc3d is not None:
            bc = _lib.BC3d()
            for f in range(6):
                bc.kinds[f] = bc3d.kinds[f]
            bc.excision_tag = bc3d.excision_tag
            bc.sigma = bc3d.sigma
            bc.gPhi, bc.gPi, bc.gD = pg1, pg2, pg3
        self._sync_torch()
        check(lib().wt2_fo_rk4_stage(self.ctx, int(stage), p0, p1, pc,
                                     float(dt), float(eps_ko), None if bc is None else ctypes.byref(bc)))

    def fo_energy(self, Phi, Pi, coef):
        (p1, p2, pc), fl = self._ptrs(Phi, Pi, coef, mult={2: 11})
        E = ctypes.c_double()
        self._sync_torch()
        check(lib().wt2_fo_energy(self.ctx, p1, p2, pc, ctypes.byref(E), fl))
        return E.value


class Geometry:
    def __init__(self, mesh, elem):
        self.mesh = mesh
        self.elem = elem
        self.N = elem.N
        self.Ne = mesh.Ne
        ws = None
        self._coords = None

    @property
    def coords(self):
        """(Ne, N, N, N, 3)-shaped view of the node coordinates (host)."""
        if self._coords is None:
            self._coords = host_coords(self.mesh, self.elem)
        return self._coords

    @property
    def Hphys(self):
        """H3 * |J| per node (host); only the affine case is closed form here."""
        if self.mesh.warp_A != 0.0:
            raise NotImplementedError("Hphys on warped meshes comes from the device context")
        w = self.elem.w
        hvol = np.prod([(b - a) / m for a, b, m in zip(self.mesh.x0, self.mesh.x1, self.mesh.M)])
        H3 = np.einsum("k,j,i->kji", w, w, w)
        return np.broadcast_to(H3 * hvol, (self.Ne,) + H3.shape).reshape(-1).copy()


def host_coords(mesh, elem):
    """Node coordinates, flat (npoints, 3), computed on the host like wt2_coords."""
    N = elem.N
    M = mesh.M
    ez, ey, ex = np.meshgrid(np.arange(M[2]), np.arange(M[1]), np.arange(M[0]), indexing="ij")
    ex, ey, ez = ex.reshape(-1), ey.reshape(-1), ez.reshape(-1)
    xs = elem.xs
    k, j, i = np.meshgrid(xs, xs, xs, indexing="ij")
    xi0 = (ex[:, None, None, None] + i[None]) / M[0]
    xi1 = (ey[:, None, None, None] + j[None]) / M[1]
    xi2 = (ez[:, None, None, None] + k[None]) / M[2]
    s = 0.0
    if mesh.warp_A != 0.0:
        tp = 2.0 * np.pi
        s = mesh.warp_A * np.sin(tp * xi0) * np.sin(tp * xi1) * np.sin(tp * xi2)
    out = np.empty(xi0.shape + (3,))
    for a, xi in enumerate((xi0, xi1, xi2)):
        out[..., a] = mesh.x0[a] + (mesh.x1[a] - mesh.x0[a]) * (xi + s)
    return out.reshape(-1, 3)


def make_geometry(mesh, elem):
    return Geometry(mesh, elem)


def make_workspace(geom, tau=None, eps_ko=0.0, bdry_values=(0.0,) * 6, precision="fp64", bg=None,
                   allow_noncoercive=False):
    return Workspace(geom.mesh, geom.elem, tau=tau, eps_ko=eps_ko, bdry_values=bdry_values,
                     precision=precision, bg=bg, allow_noncoercive=allow_noncoercive)


to_device = make_workspace


# ------------------------------------------------------------- RHS / energy

def rhs_wave3d(udd, u, ud, params, geom=None, ops=None, work=None):
    """rhs_wave3d!(udd, u, ud, params; geom, ops, work) (src/wave.jl:573-611):
    udd = L_SIPG u with Dirichlet values params.bdry_values[tag-1].  work is a
    Workspace (its tau is the context's; params.tau must agree).  Tag-7
    Sommerfeld faces are the workspace's (Workspace(sommerfeld_faces=...,
    sommerfeld_R=...)); params.sommerfeld_R must agree with its R."""
    if work is None:
        raise ValueError("rhs_wave3d needs work= (a Workspace from make_workspace)")
    if params.tau is not None and abs(params.tau - work.tau) > 1e-12 * work.tau:
        raise ValueError(f"params.tau={params.tau} differs from the workspace tau={work.tau}")
    inv_R = work.desc.sommerfeld_inv_R
    pR = 0.0 if math.isinf(params.sommerfeld_R) else 1.0 / params.sommerfeld_R
    if any(work.desc.face_kind[f] == SOMMERFELD_BDRY_TAG for f in range(6)) and abs(pR - inv_R) > 1e-12 * max(inv_R, 1.0):
        raise ValueError(f"params.sommerfeld_R={params.sommerfeld_R} differs from the workspace's Sommerfeld R")
    return work.rhs(udd, u, ud, params.bdry_values)


def wave3d_rhs_full(du, dud, u, ud, work):
    """(du, dud) = (ud + KO(u), L u + KO(ud)) with KO = eps mu^-5 sum_d D_d^6."""
    return work.rhs_full(du, dud, u, ud)


class BC3d:
    """make_bc3d(kinds; sigma, gPhi, gPi) (src/boundaries3d.jl:44-54): a boundary
    kind per face (-x,+x,-y,+y,-z,+z) as a name (:dirichlet, :sommerfeld,
    :excision, :full_dirichlet) or a WT2_BC_* code, the penalty scale and the
    optional per-node targets (arrays on the same side as the fields)."""

    _NAMES = {"dirichlet": _lib.WT2_BC_DIRICHLET, "sommerfeld": _lib.WT2_BC_SOMMERFELD,
              "excision": _lib.WT2_BC_EXCISION, "full_dirichlet": _lib.WT2_BC_FULL_DIRICHLET}

    def __init__(self, kinds, sigma=1.0, gPhi=None, gPi=None, gD=None, excision_tag=0):
        if isinstance(kinds, (str, int)):
            kinds = (kinds,) * 6
        if len(kinds) != 6:
            raise ValueError("make_bc3d needs six face kinds")
        self.kinds = tuple(self._NAMES[k] if isinstance(k, str) else int(k) for k in kinds)
        self.sigma = float(sigma)
        self.gPhi_arr = gPhi
        self.gPi_arr = gPi
        self.gD_arr = gD                 # general meshes, Dirichlet kind: grad Phi target (3 fields)
        self.excision_tag = int(excision_tag)


def make_bc3d(kinds, sigma=1.0, gPhi=None, gPi=None, gD=None, excision_tag=0):
    """make_bc3d; on general meshes one kind (a single name) for every outer
    face but those tagged excision_tag (make_bc3d(kinds; excision_tag),
    src/boundaries3d.jl:44-54, 131-165)."""
    return BC3d(kinds, sigma, gPhi, gPi, gD, excision_tag)


# characteristic classes of a boundary face (src/boundaries1d.jl:79-83)
FACE_SUBLUMINAL, FACE_OUTFLOW, FACE_INFLOW, FACE_SONIC = 1, 2, 3, 4
_FACE_CLASS_NAMES = {FACE_SUBLUMINAL: "subluminal", FACE_OUTFLOW: "superluminal outflow",
                     FACE_INFLOW: "superluminal inflow", FACE_SONIC: "sonic"}


def classify_face1d(a, beta, nsign, sonic_tol=np.finfo(np.float64).eps ** 0.25):
    """classify_face1d (src/boundaries1d.jl:97-108): the class of a face with
    normal characteristic speed a = alpha sqrt(gamma^nn), shift component beta
    along the normal axis and outward sign nsign = +-1."""
    s_R = a - beta
    s_L = -a - beta
    tol = sonic_tol * a
    if abs(s_R) <= tol or abs(s_L) <= tol:
        return FACE_SONIC
    out_R = s_R * nsign > 0
    out_L = s_L * nsign > 0
    if out_R and out_L:
        return FACE_OUTFLOW
    if not out_R and not out_L:
        return FACE_INFLOW
    return FACE_SUBLUMINAL


def classify_face3d(alpha, b1, b2, b3, gu11, gu22, gu33, axis, nsign,
                    sonic_tol=np.finfo(np.float64).eps ** 0.25):
    """classify_face3d (src/boundaries3d.jl:25-32); axis 1..3."""
    gunn = (gu11, gu22, gu33)[axis - 1]
    bax = (b1, b2, b3)[axis - 1]
    return classify_face1d(alpha * math.sqrt(gunn), bax, nsign, sonic_tol)


def validate_bc1d(cls, kind, face):
    """validate_bc1d (src/boundaries1d.jl:120-138): subluminal faces take
    Dirichlet or Sommerfeld, outflow faces excision, inflow faces the full
    state; sonic faces none.  Raises ValueError."""
    if cls == FACE_SONIC:
        raise ValueError(f"{face}: the normal shift equals the light speed there (|beta_n| = alpha sqrt(gamma^nn)), "
                         "so one characteristic speed is zero and none of the boundary kinds applies")
    ok = ((cls == FACE_SUBLUMINAL and kind in (_lib.WT2_BC_DIRICHLET, _lib.WT2_BC_SOMMERFELD))
          or (cls == FACE_OUTFLOW and kind == _lib.WT2_BC_EXCISION)
          or (cls == FACE_INFLOW and kind == _lib.WT2_BC_FULL_DIRICHLET))
    if not ok:
        adm = {FACE_SUBLUMINAL: "dirichlet or sommerfeld", FACE_OUTFLOW: "excision",
               FACE_INFLOW: "full_dirichlet"}[cls]
        raise ValueError(f"{face} boundary face is {_FACE_CLASS_NAMES[cls]}; admissible boundary conditions: {adm}")


def wave3d_curved_rhs(dPhi, dPi, Phi, Pi, coef, work, eps_ko=0.1, bc3d=None):
    """wave3d_curved_rhs!(Phidot, Pidot, Phi, Pi, coef; geom, ops, ws, eps_KO, bc3d)
    (src/wave3d_curved.jl:205-258), the affine first-order path; coef from
    work.fo_coef."""
    return work.fo_rhs(dPhi, dPi, Phi, Pi, coef, eps_ko, bc3d)


def wave3d_energy(Phi, Pi, coef, work):
    """wave3d_energy (src/wave3d_curved.jl:268-287)."""
    return work.fo_energy(Phi, Pi, coef)


def discrete_energy(u, ud, work):
    """E = 1/2 <ud,ud>_H + 1/2 <u, -L u>_H with zero boundary data
    (test/kernels3d_tests.jl:32-47), FP64 on the device."""
    return work.energy(u, ud)


# ------------------------------------------------------------- background

class AnalyticBackground3D:
    """Closures alpha_fn(t,x,y,z), beta_fn -> (b1,b2,b3), gamma_fn -> (g11,g12,g13,g22,g23,g33)
    (src/wave3d_curved.jl:45-57); vectorised over numpy arrays."""

    def __init__(self, alpha_fn, beta_fn, gamma_fn):
        self.alpha_fn = alpha_fn
        self.beta_fn = beta_fn
        self.gamma_fn = gamma_fn


def sample_background3d(bg, t, xg, yg, zg):
    """Per-node (alpha, beta1..3, gamma11..33) as an (n, 10) array, the input of
    Workspace.set_background (sample_background3d!, src/wave3d_curved.jl:106-115)."""
    n = np.size(xg)
    out = np.empty((n, 10))
    out[:, 0] = np.broadcast_to(bg.alpha_fn(t, xg, yg, zg), (n,))
    b = bg.beta_fn(t, xg, yg, zg)
    for q in range(3):
        out[:, 1 + q] = np.broadcast_to(b[q], (n,))
    g = bg.gamma_fn(t, xg, yg, zg)
    for q in range(6):
        out[:, 4 + q] = np.broadcast_to(g[q], (n,))
    return out


def make_coef3d(geom):
    return np.zeros((geom.N ** 3 * geom.Ne, 10))


def adm_bump_background(amp_lapse=0.2, amp_shift=0.05, amp_metric=0.2, width2=0.1, centre=0.5):
    """The curved-background configuration: with r = |x - centre|,
    alpha = 1 - 0.2 exp(-r^2/0.1), beta^i = 0.05 exp(-r^2/0.1)(x_i - centre),
    gamma_ij = (1 + 0.2 exp(-r^2/0.1)) delta_ij."""
    def env(x, y, z):
        return np.exp(-((x - centre) ** 2 + (y - centre) ** 2 + (z - centre) ** 2) / width2)

    def alpha(t, x, y, z):
        return 1.0 - amp_lapse * env(x, y, z)

    def beta(t, x, y, z):
        e = amp_shift * env(x, y, z)
        return (e * (x - centre), e * (y - centre), e * (z - centre))

    def gamma(t, x, y, z):
        g = 1.0 + amp_metric * env(x, y, z)
        z0 = np.zeros_like(g)
        return (g, z0, z0, g, z0, g)

    return AnalyticBackground3D(alpha, beta, gamma)


# -------------------------------------------------------------------- ICs

def eigenmode_cartesian(coords, t, A=1.0, kx=math.pi, ky=math.pi, kz=math.pi, omega=None,
                        x0=0.0, x1=1.0):
    """eigenmode_cartesian! (src/wave.jl:163-185): u, u_t of
    A sin(kx X) sin(ky Y) sin(kz Z) cos(omega t), X = (x - x0)/(x1 - x0)."""
    L = x1 - x0
    if omega is None:
        omega = math.sqrt(kx * kx + ky * ky + kz * kz) / L
    X = (coords[:, 0] - x0) / L
    Y = (coords[:, 1] - x0) / L
    Z = (coords[:, 2] - x0) / L
    s = A * np.sin(kx * X) * np.sin(ky * Y) * np.sin(kz * Z)
    return s * math.cos(omega * t), -omega * s * math.sin(omega * t)


def gaussian_pulse(coords, centre, width=None, denom=None, amplitude=1.0):
    """Gaussian pulse u = A exp(-|x-c|^2 / denom), u_t = 0.  denom defaults to
    2 width^2 (src/evolve.jl:1260-1263)."""
    if denom is None:
        denom = 2.0 * width * width
    c = np.asarray(centre, dtype=np.float64)
    r2 = np.sum((coords - c[None, :]) ** 2, axis=1)
    return amplitude * np.exp(-r2 / denom), np.zeros(len(coords))


def eigenmode_radial(coords, t, A=1.0, R=0.5, n=1, centre=(0.0, 0.0, 0.0)):
    """eigenmode_radial! (src/wave.jl:207-226): u = A sinc(n r / R) cos(omega t),
    u_t = -A omega sinc(n r / R) sin(omega t), omega = n pi / R, with
    sinc(y) = sin(pi y) / (pi y) (1 at y = 0)."""
    c = np.asarray(centre, dtype=np.float64)
    r = np.sqrt(np.sum((coords - c[None, :]) ** 2, axis=1))
    omega = n * math.pi / R
    s = A * np.sinc(n * r / R)
    return s * math.cos(omega * t), -omega * s * math.sin(omega * t)


def _sinhc(x):
    """sinh(x)/x, 1 at x = 0 (series below x^2 = 1e-3)."""
    x = np.asarray(x, dtype=np.float64)
    x2 = x * x
    small = x2 < 1e-3
    safe = np.where(small, 1.0, x)
    return np.where(small, 1.0 + x2 * (1.0 / 6 + x2 * (1.0 / 120 + x2 / 5040)), np.sinh(safe) / safe)


def outgoing_pulse(coords, t, A=1.0, s0=0.25, sigma=0.05, centre=(0.0, 0.0, 0.0)):
    """outgoing_pulse! (src/wave.jl:292-320): the outgoing spherical pulse
    u = [F(t - r) - F(t + r)] / r, F(s) = A exp(-(s + s0)^2 / (2 sigma^2)), in
    the closed form with b = t + s0:
    u = 2 A b / sigma^2 sinhc(b r / sigma^2) exp(-(b^2 + r^2) / (2 sigma^2)),
    u_t = -2 A / sigma^2 exp(...) (b^2 / sigma^2 sinhc(b r / sigma^2) - cosh(b r / sigma^2))."""
    c = np.asarray(centre, dtype=np.float64)
    r = np.sqrt(np.sum((coords - c[None, :]) ** 2, axis=1))
    is2 = 1.0 / (sigma * sigma)
    b = t + s0
    e = np.exp(-0.5 * (b * b + r * r) * is2)
    sh = _sinhc(b * r * is2)
    u = 2.0 * A * b * is2 * sh * e
    ud = -2.0 * A * is2 * e * (b * b * is2 * sh - np.cosh(b * r * is2))
    return u, ud


def recommended_dt_rk4(work, cfl=0.2):
    """dt = cfl * h_min (h_min per src/evolve.jl:60-69), capped by the KO limit
    1.4/(eps mu) (src/evolve.jl:1211-1213)."""
    dt = cfl * work.h_min
    eps = work.desc.eps_ko
    if eps > 0:
        dt = min(dt, 1.4 / (eps * work.mu))
    return dt


# ------------------------------------------------------------------ batches

class Batch:
    """Many independent evolutions advanced together (wt2_batch_*): each slot is
    a Workspace with device-resident state and its own dt and step count."""

    def __init__(self):
        self.ptr = ctypes.c_void_p()
        check(lib().wt2_batch_create(ctypes.byref(self.ptr)))
        self.slots = {}

    def close(self):
        if getattr(self, "ptr", None) and self.ptr.value:
            lib().wt2_batch_destroy(self.ptr)
            self.ptr = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def add(self, work, u, ud, dt, nsteps, energy_every=0):
        if torch is None or not (isinstance(u, torch.Tensor) and u.is_cuda):
            raise TypeError("batch state must be torch CUDA tensors (device-resident)")
        (pu, pud), fl = work._ptrs(u, ud, writable=(0, 1))
        slot = ctypes.c_int32()
        work._sync_torch()
        check(lib().wt2_batch_add(self.ptr, work.ctx, pu, pud, float(dt), int(nsteps), int(energy_every),
                                  ctypes.byref(slot)))
        self.slots[slot.value] = (work, u, ud, int(nsteps), int(energy_every))
        return slot.value

    def remove(self, slot):
        check(lib().wt2_batch_remove(self.ptr, int(slot)))
        self.slots.pop(slot, None)

    def step(self, max_steps):
        fin = ctypes.c_int32()
        check(lib().wt2_batch_step(self.ptr, int(max_steps), ctypes.byref(fin)))
        return fin.value

    def step_async(self, max_steps):
        """Queue up to max_steps steps of every unfinished instance; returns the
        number of instances whose last step is now queued."""
        fin = ctypes.c_int32()
        check(lib().wt2_batch_step_async(self.ptr, int(max_steps), ctypes.byref(fin)))
        return fin.value

    def sync(self):
        check(lib().wt2_batch_sync(self.ptr))

    def launches_per_step(self):
        """Kernel launches of one RK4 step as last planned (0 before the first
        step or after the set of unfinished instances changed)."""
        n = ctypes.c_int32()
        a = ctypes.c_int32()
        check(lib().wt2_batch_info(self.ptr, ctypes.byref(n), ctypes.byref(a)))
        return n.value

    def steps_done(self, slot):
        """Steps queued so far for a slot (no device synchronisation)."""
        done = ctypes.c_int64()
        check(lib().wt2_batch_slot(self.ptr, int(slot), ctypes.byref(done), None, 0, None))
        return done.value

    def slot_info(self, slot):
        done = ctypes.c_int64()
        n = ctypes.c_int32()
        check(lib().wt2_batch_slot(self.ptr, int(slot), ctypes.byref(done), None, 0, ctypes.byref(n)))
        E = np.zeros(max(n.value, 1))
        check(lib().wt2_batch_slot(self.ptr, int(slot), None, E.ctypes.data, n.value, None))
        return done.value, E[:n.value]
