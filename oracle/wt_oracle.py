"""ctypes binding of the CPU FP64 oracle (libwt_oracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, to check results against.  The product library never loads
it.  The C source (wt_oracle.c) cites the reference file:line each function
restates.
"""
import ctypes
import math
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "libwt_oracle.so")
_lib = None

NMAX = 16


class Desc(ctypes.Structure):
    _fields_ = [
        ("N", ctypes.c_int),
        ("M", ctypes.c_int * 3),
        ("x0", ctypes.c_double * 3),
        ("x1", ctypes.c_double * 3),
        ("periodic", ctypes.c_int * 3),
        ("warp_A", ctypes.c_double),
        ("tau", ctypes.c_double),
        ("eps_ko", ctypes.c_double),
        ("bdry_values", ctypes.c_double * 6),
        ("face_kind", ctypes.c_int * 6),
        ("sommerfeld_inv_R", ctypes.c_double),
    ]


BC_DIRICHLET, BC_SOMMERFELD, BC_EXCISION, BC_FULL_DIRICHLET = 1, 2, 3, 4   # src/boundaries1d.jl:59-62
_BC_NAMES = {"dirichlet": 1, "sommerfeld": 2, "excision": 3, "full_dirichlet": 4}


class BC3d(ctypes.Structure):
    _fields_ = [
        ("kinds", ctypes.c_int * 6),
        ("excision_tag", ctypes.c_int),
        ("sigma", ctypes.c_double),
        ("gPhi", ctypes.POINTER(ctypes.c_double)),
        ("gPi", ctypes.POINTER(ctypes.c_double)),
        ("gD", ctypes.POINTER(ctypes.c_double)),
    ]


def make_bc3d(kinds, sigma=1.0, gPhi=None, gPi=None, gD=None, excision_tag=0):
    """make_bc3d (src/boundaries3d.jl:44-54): kinds per face (-x,+x,-y,+y,-z,+z),
    names or BC_* codes (a single name: every face); gPhi / gPi: optional
    per-node targets; gD: the grad Phi target (3 fields) of the curvilinear
    Dirichlet kind; excision_tag: curvilinear outer faces with this mesh tag
    get no SAT."""
    if isinstance(kinds, (str, int)):
        kinds = [kinds] * 6
    b = BC3d()
    for f in range(6):
        k = kinds[f]
        b.kinds[f] = _BC_NAMES[k] if isinstance(k, str) else int(k)
    b.excision_tag = int(excision_tag)
    b.sigma = float(sigma)
    b._keep = [None if g is None else np.ascontiguousarray(g, dtype=np.float64).reshape(-1) for g in (gPhi, gPi, gD)]
    b.gPhi = _p(b._keep[0])
    b.gPi = _p(b._keep[1])
    b.gD = _p(b._keep[2])
    return b


def fo_coef(bg10):
    """sample_background3d!'s coefficient bundle (src/wave3d_curved.jl:75-96)
    from 10 background values per node: (11, npts) alpha, sqrtg, b1..3,
    gu11, gu12, gu13, gu22, gu23, gu33."""
    bg = np.ascontiguousarray(bg10, dtype=np.float64).reshape(-1)
    n = bg.size // 10
    out = np.zeros(11 * n)
    lib().wto_fo_coef(n, _p(bg), _p(out))
    return out.reshape(11, n)


def build():
    """Compile libwt_oracle.so with the Makefile beside this file."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        dp = ctypes.POINTER(ctypes.c_double)
        vp = ctypes.c_void_p
        L.wto_gll.argtypes = [ctypes.c_int, dp, dp, dp]
        L.wto_create.argtypes = [ctypes.POINTER(Desc), dp]
        L.wto_create.restype = vp
        L.wto_create_mesh.argtypes = [ctypes.POINTER(Desc), ctypes.c_long, dp, vp, vp, vp]
        L.wto_create_mesh.restype = vp
        L.wto_destroy.argtypes = [vp]
        L.wto_npoints.argtypes = [vp]
        L.wto_npoints.restype = ctypes.c_long
        L.wto_is_curved.argtypes = [vp]
        L.wto_mu.argtypes = [vp]
        L.wto_mu.restype = ctypes.c_double
        L.wto_hmin.argtypes = [vp]
        L.wto_hmin.restype = ctypes.c_double
        L.wto_coords.argtypes = [vp, dp]
        L.wto_hphys.argtypes = [vp, dp]
        L.wto_coefs.argtypes = [vp, dp]
        L.wto_lap.argtypes = [vp, dp, dp, dp]
        L.wto_rhs.argtypes = [vp, dp, dp, dp, dp]
        L.wto_apply_D.argtypes = [vp, ctypes.c_int, dp, dp]
        L.wto_apply_Dko.argtypes = [vp, ctypes.c_int, dp, dp]
        L.wto_rk4.argtypes = [vp, dp, dp, ctypes.c_double, ctypes.c_int, dp]
        L.wto_energy.argtypes = [vp, dp, dp]
        L.wto_energy.restype = ctypes.c_double
        L.wto_fo_coef.argtypes = [ctypes.c_long, dp, dp]
        L.wto_fo_rhs.argtypes = [vp, dp, dp, dp, ctypes.c_double, ctypes.POINTER(BC3d), dp, dp]
        L.wto_fo_energy.argtypes = [vp, dp, dp, dp]
        L.wto_fo_energy.restype = ctypes.c_double
        L.wto_plant_mistake.argtypes = [ctypes.c_int]
        L.wto_curv_metric.argtypes = [vp, dp, dp]
        L.wto_curv_grad.argtypes = [vp, dp, dp]
        L.wto_curv_div.argtypes = [vp, dp, dp]
        L.wto_curv_D.argtypes = [vp, ctypes.c_int, dp, dp]
        L.wto_curv_mu.argtypes = [vp]
        L.wto_curv_mu.restype = ctypes.c_double
        L.wto_fo_rhs_curv.argtypes = [vp, dp, dp, dp, ctypes.c_double, ctypes.POINTER(BC3d), dp, dp]
        L.wto_fo_energy_curv.argtypes = [vp, dp, dp, dp]
        L.wto_fo_energy_curv.restype = ctypes.c_double
        L.wto_assemble_1d.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                      ctypes.c_double, ctypes.c_double, ctypes.c_int, dp]
        _lib = L
    return _lib


def _p(a):
    if a is None:
        return None
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def gll(N):
    xi = np.zeros(N); w = np.zeros(N); G = np.zeros((N, N))
    if lib().wto_gll(N, _p(xi), _p(w), _p(G)) != 0:
        raise ValueError("N out of range")
    return xi, w, G


PLANT_DROP_BB, PLANT_WGT_SQRTG, PLANT_FLIP_C = 1, 2, 3


class planted_mistake:
    """Context manager: oracle contexts created inside it carry a known mistake
    in their ADM coefficients (PLANT_*), so a pin test can show it catches it."""

    def __init__(self, code):
        self.code = int(code)

    def __enter__(self):
        lib().wto_plant_mistake(self.code)
        return self

    def __exit__(self, *exc):
        lib().wto_plant_mistake(0)
        return False


def assemble_1d(N, M, periodic, tau, h, kind):
    A = np.zeros((M * N, M * N))
    lib().wto_assemble_1d(N, M, int(periodic), tau, h, kind, _p(A))
    return A


class Oracle:
    """One mesh + operator set; arrays are flat float64 (N,N,N,Ne) Fortran-order
    fields stored as 1-D numpy arrays of length N^3*Ne."""

    def __init__(self, N, M, x0=0.0, x1=1.0, periodic=False, warp_A=0.0, tau=None,
                 eps_ko=0.0, bdry_values=(0.0,) * 6, bg=None, sommerfeld_faces=(), sommerfeld_R=math.inf):
        if isinstance(M, int):
            M = (M, M, M)
        if not isinstance(x0, (tuple, list)):
            x0 = (x0,) * 3
        if not isinstance(x1, (tuple, list)):
            x1 = (x1,) * 3
        if not isinstance(periodic, (tuple, list)):
            periodic = (periodic,) * 3
        if tau is None:
            tau = 1.5 * (N - 1) ** 2
        d = Desc()
        d.N = N
        for a in range(3):
            d.M[a] = M[a]; d.x0[a] = x0[a]; d.x1[a] = x1[a]; d.periodic[a] = int(periodic[a])
        d.warp_A = warp_A; d.tau = tau; d.eps_ko = eps_ko
        for f in range(6):
            d.bdry_values[f] = bdry_values[f]
            d.face_kind[f] = 7 if (f + 1) in tuple(sommerfeld_faces) else 0
        d.sommerfeld_inv_R = 0.0 if math.isinf(sommerfeld_R) else 1.0 / sommerfeld_R
        self.desc = d
        self.N, self.M = N, tuple(M)
        self.Ne = M[0] * M[1] * M[2]
        self.n = N ** 3 * self.Ne
        self._bg = None if bg is None else np.ascontiguousarray(bg, dtype=np.float64)
        self.h = self._bg  # keep alive
        self.ctx = lib().wto_create(ctypes.byref(d), _p(self._bg))
        if not self.ctx:
            raise ValueError("wto_create failed")

    @classmethod
    def from_mesh(cls, N, coords, neighbour, neighbour_face, bdry, tau=None, bdry_values=(0.0,) * 6,
                  sommerfeld_R=math.inf):
        """A mesh given by node coordinates (N^3 Ne, 3) and (Ne, 6) 1-based
        neighbour / neighbour_face / bdry arrays (tag 1..6 Dirichlet, 7
        Sommerfeld on outer faces)."""
        self = cls.__new__(cls)
        d = Desc()
        d.N = N
        d.tau = 1.5 * (N - 1) ** 2 if tau is None else tau
        d.eps_ko = 0.0
        for f in range(6):
            d.bdry_values[f] = bdry_values[f]
        d.sommerfeld_inv_R = 0.0 if math.isinf(sommerfeld_R) else 1.0 / sommerfeld_R
        X = np.ascontiguousarray(coords, dtype=np.float64).reshape(-1)
        self.Ne = X.size // (3 * N ** 3)
        tabs = [np.ascontiguousarray(t, dtype=np.int32).reshape(-1) for t in (neighbour, neighbour_face, bdry)]
        self.desc, self.N, self.M = d, N, (self.Ne, 1, 1)
        self.n = N ** 3 * self.Ne
        self._bg = None
        self.h = (X, tabs)  # keep alive
        self.ctx = lib().wto_create_mesh(ctypes.byref(d), self.Ne, _p(X), tabs[0].ctypes.data, tabs[1].ctypes.data,
                                           tabs[2].ctypes.data)
        if not self.ctx:
            raise ValueError("wto_create_mesh failed")
        return self

    def __del__(self):
        # at interpreter exit the module globals may already be gone
        if getattr(self, "ctx", None) and lib is not None:
            lib().wto_destroy(self.ctx)
            self.ctx = None

    def _new(self):
        return np.zeros(self.n)

    @property
    def mu(self):
        return lib().wto_mu(self.ctx)

    @property
    def hmin(self):
        return lib().wto_hmin(self.ctx)

    def coords(self):
        x = np.zeros(3 * self.n)
        lib().wto_coords(self.ctx, _p(x))
        return x.reshape(self.n, 3)

    def hphys(self):
        h = self._new()
        lib().wto_hphys(self.ctx, _p(h))
        return h

    def coefs(self):
        c = np.zeros(13 * self.n)
        if lib().wto_coefs(self.ctx, _p(c)) != 0:
            return None
        return c.reshape(self.n, 13)

    def lap(self, u, ud=None):
        y = self._new()
        lib().wto_lap(self.ctx, _p(u), _p(ud), _p(y))
        return y

    def rhs(self, u, v):
        du = self._new(); dv = self._new()
        lib().wto_rhs(self.ctx, _p(u), _p(v), _p(du), _p(dv))
        return du, dv

    def apply_D(self, axis, x):
        y = self._new()
        lib().wto_apply_D(self.ctx, axis, _p(x), _p(y))
        return y

    def apply_Dko(self, axis, x):
        y = self._new()
        lib().wto_apply_Dko(self.ctx, axis, _p(x), _p(y))
        return y

    def rk4(self, u, v, dt, nsteps, energy=False):
        u = np.array(u, dtype=np.float64, copy=True)
        v = np.array(v, dtype=np.float64, copy=True)
        E = np.zeros(nsteps + 1) if energy else None
        lib().wto_rk4(self.ctx, _p(u), _p(v), dt, nsteps, _p(E))
        return (u, v, E) if energy else (u, v)

    def energy(self, u, v):
        return lib().wto_energy(self.ctx, _p(u), _p(v))

    # first-order (Phi, Pi) system, affine meshes
    def fo_rhs(self, Phi, Pi, coef, eps_ko=0.1, bc=None):
        """wave3d_curved_rhs!(Phidot, Pidot, Phi, Pi, coef; eps_KO, bc3d)
        (src/wave3d_curved.jl:205-258); coef: (11, npts) from fo_coef."""
        coef = np.ascontiguousarray(coef, dtype=np.float64).reshape(-1)
        assert coef.size == 11 * self.n
        dPhi = self._new(); dPi = self._new()
        lib().wto_fo_rhs(self.ctx, _p(Phi), _p(Pi), _p(coef), float(eps_ko),
                         None if bc is None else ctypes.byref(bc), _p(dPhi), _p(dPi))
        return dPhi, dPi

    def fo_energy(self, Phi, Pi, coef):
        coef = np.ascontiguousarray(coef, dtype=np.float64).reshape(-1)
        return lib().wto_fo_energy(self.ctx, _p(Phi), _p(Pi), _p(coef))

    def fo_rk4(self, Phi, Pi, coef, dt, nsteps, eps_ko=0.1, bc=None):
        """Classical RK4 of the first-order system in the stage order of
        test/wave3d_tests.jl:32-44."""
        Phi = np.array(Phi, dtype=np.float64, copy=True)
        Pi = np.array(Pi, dtype=np.float64, copy=True)
        for _ in range(nsteps):
            k1 = self.fo_rhs(Phi, Pi, coef, eps_ko, bc)
            k2 = self.fo_rhs(Phi + dt / 2 * k1[0], Pi + dt / 2 * k1[1], coef, eps_ko, bc)
            k3 = self.fo_rhs(Phi + dt / 2 * k2[0], Pi + dt / 2 * k2[1], coef, eps_ko, bc)
            k4 = self.fo_rhs(Phi + dt * k3[0], Pi + dt * k3[1], coef, eps_ko, bc)
            Phi += dt / 6 * (k1[0] + 2 * k2[0] + 2 * k3[0] + k4[0])
            Pi += dt / 6 * (k1[1] + 2 * k2[1] + 2 * k3[1] + k4[1])
        return Phi, Pi


def _c(a):
    return np.ascontiguousarray(a, dtype=np.float64).reshape(-1)


def _curv_methods():
    """The curvilinear first-order path on general meshes (Oracle.from_mesh)."""

    def curv_metric(self):
        """(JAp (npts, 3, 3) [a, m] = sign(J) J dxi^a/dx_m, |J| (npts,))."""
        JA = np.zeros(9 * self.n); J = np.zeros(self.n)
        if lib().wto_curv_metric(self.ctx, _p(JA), _p(J)) != 0:
            raise ValueError("not a general mesh")
        return JA.reshape(self.n, 3, 3), J

    def curv_grad(self, Phi):
        D = np.zeros(3 * self.n)
        lib().wto_curv_grad(self.ctx, _p(_c(Phi)), _p(D))
        return D.reshape(3, self.n)

    def curv_div(self, F3):
        out = np.zeros(self.n)
        lib().wto_curv_div(self.ctx, _p(_c(F3)), _p(out))
        return out

    def curv_D(self, axis, x):
        out = np.zeros(self.n)
        lib().wto_curv_D(self.ctx, int(axis), _p(_c(x)), _p(out))
        return out

    def curv_mu(self):
        return lib().wto_curv_mu(self.ctx)

    def fo_rhs_curv(self, Phi, Pi, coef, eps_ko=0.1, bc=None):
        coef = _c(coef)
        assert coef.size == 11 * self.n
        dPhi = self._new(); dPi = self._new()
        lib().wto_fo_rhs_curv(self.ctx, _p(_c(Phi)), _p(_c(Pi)), _p(coef), float(eps_ko),
                              None if bc is None else ctypes.byref(bc), _p(dPhi), _p(dPi))
        return dPhi, dPi

    def fo_energy_curv(self, Phi, Pi, coef):
        return lib().wto_fo_energy_curv(self.ctx, _p(_c(Phi)), _p(_c(Pi)), _p(_c(coef)))

    def fo_rk4_curv(self, Phi, Pi, coef, dt, nsteps, eps_ko=0.1, bc=None):
        """Classical RK4 of the curvilinear first-order system (stage order of
        test/wave3d_tests.jl:32-44)."""
        Phi = np.array(Phi, dtype=np.float64, copy=True)
        Pi = np.array(Pi, dtype=np.float64, copy=True)
        for _ in range(nsteps):
            k1 = self.fo_rhs_curv(Phi, Pi, coef, eps_ko, bc)
            k2 = self.fo_rhs_curv(Phi + dt / 2 * k1[0], Pi + dt / 2 * k1[1], coef, eps_ko, bc)
            k3 = self.fo_rhs_curv(Phi + dt / 2 * k2[0], Pi + dt / 2 * k2[1], coef, eps_ko, bc)
            k4 = self.fo_rhs_curv(Phi + dt * k3[0], Pi + dt * k3[1], coef, eps_ko, bc)
            Phi += dt / 6 * (k1[0] + 2 * k2[0] + 2 * k3[0] + k4[0])
            Pi += dt / 6 * (k1[1] + 2 * k2[1] + 2 * k3[1] + k4[1])
        return Phi, Pi

    for f in (curv_metric, curv_grad, curv_div, curv_D, curv_mu, fo_rhs_curv, fo_energy_curv, fo_rk4_curv):
        setattr(Oracle, f.__name__, f)


_curv_methods()
