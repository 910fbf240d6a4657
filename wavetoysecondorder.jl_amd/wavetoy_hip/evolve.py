"""evolve3d: the driver of the fused RK4 path (the (u, u_t) counterpart of
_evolve3d_strong / _evolve3d_conservative, src/evolve.jl:877-1073,1149-1366).

It builds the mesh, context, initial data and dt, runs classical RK4 on the
device in chunks between samples, and returns a dict with the reference
driver's outputs (src/evolve.jl:1067-1072): sample times, the line of nodes
y = slice_y, z = slice_z (xs_line) with u and u_t on it at every sample
(us, uds: n_line x Nt), the H-norm L2 error against an exact solution (when
there is one), the energy history, the final state, dt, dx, tau, mu.
"""
import math

import numpy as np

from . import core


def build_slice(coords, y_target, z_target, atol=math.sqrt(np.finfo(np.float64).eps)):
    """Nodes on the line y = y_target, z = z_target: indices into the flat
    field (one per distinct x, element-major scan order) and their x, sorted by
    x (_build_slice_3d, src/evolve.jl:104-123)."""
    hit = np.nonzero((np.abs(coords[:, 1] - y_target) < atol) & (np.abs(coords[:, 2] - z_target) < atol))[0]
    idx, xs = [], []
    for n in hit:
        x = coords[n, 0]
        if not any(abs(x0 - x) < atol for x0 in xs):
            idx.append(n)
            xs.append(x)
    order = np.argsort(np.asarray(xs), kind="stable")
    return np.asarray(idx, dtype=np.int64)[order], np.asarray(xs)[order]


def evolve3d(formulation="strong", **kwargs):
    """evolve3d(; formulation, kwargs...) (src/evolve.jl:1083-1089): the
    second-order (u, u_t) driver (formulation "strong") or the first-order
    (Phi, Pi) ADM driver ("conservative")."""
    if formulation == "conservative":
        return _evolve3d_conservative(**kwargs)
    if formulation != "strong":
        raise ValueError(f"evolve3d: formulation must be 'strong' or 'conservative'; got {formulation!r}")
    return _evolve3d_strong(**kwargs)


def _hphys(N, X):
    """Hphys = w_i w_j w_k |det dx/dxi| per node of coordinates X (Ne, N, N, N, 3),
    dx/dxi by G (the geometry wt2_ctx_create_mesh builds)."""
    e = core.make_element(N)
    J = np.stack([np.einsum("ip,ekjpc->ekjic", e.G, X),
                  np.einsum("jp,ekpic->ekjic", e.G, X),
                  np.einsum("kp,epjic->ekjic", e.G, X)], axis=-1)
    w3 = np.einsum("k,j,i->kji", e.w, e.w, e.w)
    return (np.abs(np.linalg.det(J)) * w3[None]).reshape(-1)


def _spectral_radius(apply, H, n, tol=1e-3, max_iters=500, seed=0):
    """Largest |eigenvalue| of the u -> u_tt operator `apply` at u_t = 0 and
    zero boundary data (the role of spectral_radius_estimate in the reference's
    recommended_dt, src/evolve.jl:27-36).  The operator is symmetric and
    negative semi-definite in the H inner product (H = w_i w_j w_k |J|), so the
    power iteration runs in that inner product with the Rayleigh quotient
    <v, Lv>_H / <v, v>_H, which approaches rho from below; it stops when two
    successive quotients agree to `tol` (relative) and raises RuntimeError if
    that does not happen within max_iters."""
    v = np.random.default_rng(seed).standard_normal(n)
    lam_old = None
    for _ in range(max_iters):
        v /= math.sqrt(float(np.dot(v * H, v)))
        w = apply(v)
        lam = abs(float(np.dot(v * H, w)))
        if lam_old is not None and abs(lam - lam_old) <= tol * lam:
            return lam
        lam_old = lam
        v = w
    raise RuntimeError(f"evolve3d: the spectral-radius estimate did not converge in {max_iters} iterations")


def _work_operator(work):
    """u -> u_tt at u_t = 0 and zero boundary data, through the device."""
    z = np.zeros(work.npoints)
    out = np.empty(work.npoints)

    def apply(v):
        work.rhs(out, np.ascontiguousarray(v), z, (0.0,) * 6)
        return out.copy()
    return apply


def _evolve3d_strong(N=5, M=8, x0=0.0, x1=1.0, warp_A=0.0, background=None, ic="cartesian",
                     ic_wavenumber=3 * math.pi, ic_centre=(0.5, 0.5, 0.5), ic_width=None, ic_denom=None,
                     ic_radial_mode=1, ic_radius=None, ic_pulse_offset=None, ic_pulse_width=None,
                     outer_bc="dirichlet", tau=None, eps_ko=0.0, t0=0.0, t1=1.0, nsteps=None, dt=None,
                     cfl=0.2, Nt=200, precision="fp64", energy_every=None, device=True, slice_y=None,
                     slice_z=None, mesh_kind="cubical", R=0.1, L=0.1, R1=0.3, R2=1.0):
    """The (u, u_t) driver (_evolve3d_strong, src/evolve.jl:877-1073) on the
    cube mesh (warp_A = 0) or the warped cube.  ic: "cartesian" (standing
    wave), "radial" (eigenmode_radial, radius ic_radius = L/2 and mode
    ic_radial_mode about the cube centre), "outgoing" (outgoing_pulse,
    offset ic_pulse_offset = L/4 and width ic_pulse_width = offset/5 about the
    cube centre) -- these three carry the exact solution for l2_err -- or
    "gaussian".  Outer faces are Dirichlet (the reference takes
    outer_bc = "sommerfeld" only on its spherical meshes).  The defaults are
    the reference's (N = 5, M = 8, wavenumber 3 pi, t1 = 1, Nt = 200 samples,
    no KO: rhs_wave3d! has none); the time stepper is classical RK4 with dt =
    cfl h_min (the reference picks a symplectic method by N, src/evolve.jl:27-36,
    and its dt from a spectral-radius estimate); nsteps, when given, overrides
    t1.

    mesh_kind (src/evolve.jl:913-928): "cubical" (above), "cubed_cube" (the
    cube [-1, 1]^3 as 7 patches around the inner cube [-R, R]^3, outer faces
    Dirichlet), "inflated_cube" (the ball of radius R2 as 13 patches: cube
    [-L, L]^3, inflation to the sphere R1, shells to R2; outer_bc "dirichlet"
    or "sommerfeld" with sommerfeld_R = R2) or "radial_shell" (the shell
    R1 <= r <= R2 as 6 patches, the inner sphere an excision surface, tag 8,
    with no SAT; outer_bc as for the inflated cube).  On those the initial data are
    centred at the origin, the sampled line is y = z = 0 and tau defaults to
    8 (N-1)^2 (the reference's tau_mult, src/evolve.jl:962-968)."""
    if outer_bc not in ("dirichlet", "sommerfeld"):
        raise ValueError(f"evolve3d: unknown outer_bc {outer_bc!r}")
    if outer_bc == "sommerfeld" and mesh_kind not in ("inflated_cube", "radial_shell"):
        raise ValueError(f"evolve3d: outer_bc = {outer_bc!r} is only supported on the spherical meshes "
                         "(inflated cube, radial shell); the cube takes 'dirichlet'")
    elem = core.make_element(N)
    som_R = math.inf
    if mesh_kind == "cubical":
        mesh = (core.make_warped_uniform_hex(M, M, M, x0, x1, warp_A) if warp_A
                else core.make_uniform_hex(M, x0, x1))
        tau_used = core.tau_effective(tau, N)
        coords = core.host_coords(mesh, elem)
    elif mesh_kind in ("cubed_cube", "inflated_cube", "radial_shell"):
        if mesh_kind == "cubed_cube":
            x0, x1 = -1.0, 1.0
            mesh = core.make_cubed_cube_mesh(N, M, R)
        else:
            x0, x1 = -float(R2), float(R2)
            tag = 7 if outer_bc == "sommerfeld" else 1
            if mesh_kind == "inflated_cube":
                mesh = core.make_inflated_cube_mesh(N, L, R1, R2, M, outer_tag=tag)
            else:
                # the inner sphere R1 is an excision surface (tag 8, no SAT)
                mesh = core.make_radial_shell_mesh(N, R1, R2, M, outer_tag=tag)
            som_R = float(R2) if outer_bc == "sommerfeld" else math.inf
        tau_used = 8.0 * (N - 1) ** 2 if tau is None else float(tau)
        coords = mesh.coords.reshape(-1, 3)
    else:
        raise ValueError(f"evolve3d: unknown mesh_kind {mesh_kind!r} (cubical, cubed_cube, inflated_cube, "
                         "radial_shell)")
    patch = mesh_kind != "cubical"
    bg = None
    if background == "adm_bump":
        bgobj = core.adm_bump_background()
        bg = core.sample_background3d(bgobj, 0.0, coords[:, 0], coords[:, 1], coords[:, 2])
    elif background is not None:
        bg = np.asarray(background, dtype=np.float64)
    work = core.Workspace(mesh, elem, tau=tau_used, eps_ko=eps_ko, precision=precision, bg=bg,
                          sommerfeld_R=som_R)
    if dt is None:
        dt = core.recommended_dt_rk4(work, cfl)
        if patch:
            # the penalty 8 (N-1)^2 raises the spectral radius above what
            # cfl h_min allows for: RK4's stability interval on the imaginary
            # axis is 2 sqrt(2), kept at 2.5 / sqrt(rho)
            rho = _spectral_radius(_work_operator(work), _hphys(N, mesh.coords), work.npoints)
            # (the Rayleigh quotient converges to within 1e-3 from below; the
            # 2.5 against the stability limit 2.83 keeps a 12 % margin)
            dt = min(dt, 2.5 / math.sqrt(rho * (1 + 1e-2)))
    if nsteps is None:
        if t1 is None:
            raise ValueError("give nsteps or t1")
        nsteps = max(1, int(math.ceil((t1 - t0) / dt - 1e-12)))
        dt = (t1 - t0) / nsteps
    exact = None
    if ic == "cartesian":
        k = ic_wavenumber
        u0, v0 = core.eigenmode_cartesian(coords, t0, kx=k, ky=k, kz=k, x0=x0, x1=x1)
        omega = math.sqrt(3.0) * k / (x1 - x0)
        exact = lambda t: core.eigenmode_cartesian(coords, t, kx=k, ky=k, kz=k, omega=omega,
                                                   x0=x0, x1=x1)[0]
    elif ic == "radial":
        L = x1 - x0
        R = L / 2 if ic_radius is None else float(ic_radius)
        ctr = ((x0 + x1) / 2,) * 3
        u0, v0 = core.eigenmode_radial(coords, t0, R=R, n=ic_radial_mode, centre=ctr)
        exact = lambda t: core.eigenmode_radial(coords, t, R=R, n=ic_radial_mode, centre=ctr)[0]
    elif ic == "outgoing":
        L = x1 - x0
        s0 = L / 4 if ic_pulse_offset is None else float(ic_pulse_offset)
        sig = s0 / 5 if ic_pulse_width is None else float(ic_pulse_width)
        ctr = ((x0 + x1) / 2,) * 3
        u0, v0 = core.outgoing_pulse(coords, t0, s0=s0, sigma=sig, centre=ctr)
        exact = lambda t: core.outgoing_pulse(coords, t, s0=s0, sigma=sig, centre=ctr)[0]
    elif ic == "gaussian":
        u0, v0 = core.gaussian_pulse(coords, ic_centre, width=ic_width, denom=ic_denom)
    else:
        raise ValueError(f"evolve3d: unknown ic {ic!r} (cartesian, radial, outgoing, gaussian)")

    if device:
        import torch
        u = torch.from_numpy(u0).cuda()
        v = torch.from_numpy(v0).cuda()
        fetch = lambda a: a.cpu().numpy()
    else:
        u, v = u0.copy(), v0.copy()
        fetch = lambda a: a
    if Nt < 2:
        Nt = 2
    # the sampled line along x: the reference's y = z = 1/4 of the cube, or
    # (when that is not a node line of this mesh) the node line nearest to it
    def nearest(col):
        q = 0.0 if patch else x0 + 0.25 * (x1 - x0)
        vals = coords[:, col]
        return float(vals[np.argmin(np.abs(vals - q))])
    y_target = nearest(1) if slice_y is None else float(slice_y)
    z_target = nearest(2) if slice_z is None else float(slice_z)
    slice_idx, xs_line = build_slice(coords, y_target, z_target)
    if slice_idx.size == 0:
        raise ValueError(f"evolve3d: slice at y={y_target}, z={z_target} hit no GLL nodes")
    us, uds = [], []
    bounds = np.unique(np.round(np.linspace(0, nsteps, Nt)).astype(int))
    ts, l2, energies = [], [], []
    H = None
    if patch:
        H = _hphys(N, mesh.coords)
    elif mesh.warp_A == 0.0:
        H = core.Geometry(mesh, elem).Hphys

    def sample(step):
        t = t0 + step * dt
        ts.append(t)
        uh, vh = fetch(u), fetch(v)
        # every sample, as the reference's driver asserts (src/evolve.jl:1039)
        if not (np.all(np.isfinite(uh)) and np.all(np.isfinite(vh))):
            raise FloatingPointError(f"evolve3d: non-finite state at t = {t} (step {step})")
        energies.append(work.energy(u, v))
        us.append(uh[slice_idx])
        uds.append(vh[slice_idx])
        if exact is not None and H is not None:
            err = uh - exact(t)
            l2.append(float(np.sqrt(np.sum(err * err * H))))

    sample(0)
    done = 0
    for b in bounds[1:]:
        work.rk4_steps(u, v, dt, int(b - done))
        done = int(b)
        sample(done)
    uf, vf = fetch(u), fetch(v)
    if not (np.all(np.isfinite(uf)) and np.all(np.isfinite(vf))):
        raise FloatingPointError("evolve3d: non-finite state")
    return dict(ts=np.array(ts), xs_line=xs_line, us=np.stack(us, axis=1), uds=np.stack(uds, axis=1),
                l2_err=np.array(l2), energy=np.array(energies), u_final=uf,
                ud_final=vf, dt=dt, nsteps=nsteps, dx=work.h_min, tau=tau_used, mu=work.mu,
                y_target=y_target, z_target=z_target, mesh_kind=mesh_kind, outer_bc=outer_bc,
                sommerfeld_R=som_R, x0=x0, x1=x1,
                integrator_name="RK4", work=work, coords=coords, mesh=mesh)


# ------------------------------------------------ first-order (Phi, Pi) driver

def _background3d(kind, shift=(0.0, 0.0, 0.0), R1=0.5, R2=1.0):
    """The built-in static backgrounds (_background3d, src/evolve.jl:1109-1145):
    Minkowski or a constant shift (alpha = 1, gamma = I, with the exact plane
    wave Phi = sin(k (x + y + z) - omega t), k = 2 pi,
    omega = -k (b1 + b2 + b3) + k sqrt(3)), or the radial shift of the
    excision shell (flat alpha, gamma; beta = b_r x / r with b_r ramping
    linearly from V = shift[0] (> 1) at R1 to 1/20 at R2; no exact solution).
    Returns a function of the node coordinates (n, 3) giving the 10 background
    values per node, Phi_e, Pi_e, grad Phi_e (3 functions) and the largest
    characteristic speed."""
    if kind in ("minkowski", "constant_shift"):
        b = (0.0, 0.0, 0.0) if kind == "minkowski" else tuple(float(v) for v in shift)
        k = 2.0 * math.pi
        s = b[0] + b[1] + b[2]
        omega = -k * s + k * math.sqrt(3.0)
        row = np.array([1.0, b[0], b[1], b[2], 1.0, 0.0, 0.0, 1.0, 0.0, 1.0])
        bgf = lambda x: np.broadcast_to(row, (x.shape[0], 10)).copy()
        phi_e = lambda t, x, y, z, lib=np: lib.sin(k * (x + y + z) - omega * t)
        pi_e = lambda t, x, y, z, lib=np: (-omega - k * s) * lib.cos(k * (x + y + z) - omega * t)
        d_e = lambda t, x, y, z, lib=np: k * lib.cos(k * (x + y + z) - omega * t)
        return bgf, phi_e, pi_e, (d_e, d_e, d_e), abs(b[0]) + abs(b[1]) + abs(b[2]) + math.sqrt(3.0)
    if kind == "radial_shift":
        V, Vout = float(shift[0]), 1.0 / 20.0

        def bgf(x):
            r = np.linalg.norm(x, axis=1)
            br = V + (Vout - V) * (r - R1) / (R2 - R1)
            out = np.zeros((x.shape[0], 10))
            out[:, 0] = 1.0
            out[:, 1:4] = (br / r)[:, None] * x
            out[:, 4] = out[:, 7] = out[:, 9] = 1.0
            return out
        z = lambda t, x, y, zz, lib=np: 0.0 * x
        return bgf, z, z, (z, z, z), V + 1.0
    raise ValueError(f"evolve3d (conservative): unknown background {kind!r} (minkowski, constant_shift, radial_shift)")


def _side_node(M, N, f):
    """Flat node index of the representative node of boundary side f (0..5 =
    -x, +x, -y, +y, -z, +z): the first element on that side, the side's row,
    the middle tangential node (_side_node_3d, src/evolve.jl:1094-1106)."""
    ax, hi = f // 2, f % 2
    ec = [0, 0, 0]
    if hi:
        ec[ax] = M[ax] - 1
    e = ec[0] + M[0] * (ec[1] + M[1] * ec[2])
    mid = (N + 1) // 2 - 1
    ijk = [mid, mid, mid]
    ijk[ax] = N - 1 if hi else 0
    return e * N ** 3 + ijk[0] + N * (ijk[1] + N * ijk[2])


def _inverse_metric_diag(g):
    """gamma^11, gamma^22, gamma^33 from (g11, g12, g13, g22, g23, g33)."""
    g11, g12, g13, g22, g23, g33 = g
    c11 = g22 * g33 - g23 * g23
    c22 = g11 * g33 - g13 * g13
    c33 = g11 * g22 - g12 * g12
    det = g11 * c11 - g12 * (g12 * g33 - g23 * g13) + g13 * (g12 * g23 - g22 * g13)
    return c11 / det, c22 / det, c33 / det


def _face_kinds(bc, bg10, M, N):
    """The six boundary kinds of a non-periodic cube (src/evolve.jl:1218-1243):
    each face is classified from the background at its representative node;
    bc = "auto" picks Sommerfeld / excision / full Dirichlet by class, a
    6-tuple of names or WT2_BC_* codes is checked against the class."""
    BC = core.BC3d._NAMES
    kinds = []
    for f in range(6):
        q = bg10[_side_node(M, N, f)]
        gu = _inverse_metric_diag(q[4:10])
        cls = core.classify_face3d(q[0], q[1], q[2], q[3], gu[0], gu[1], gu[2], f // 2 + 1, 1 if f % 2 else -1)
        if bc == "auto":
            kind = {core.FACE_SUBLUMINAL: BC["sommerfeld"], core.FACE_OUTFLOW: BC["excision"]}.get(
                cls, BC["full_dirichlet"])
        else:
            kind = BC[bc[f]] if isinstance(bc[f], str) else int(bc[f])
        core.validate_bc1d(cls, kind, ("-x", "+x", "-y", "+y", "-z", "+z")[f] + " side")
        kinds.append(kind)
    return tuple(kinds)


def _evolve3d_conservative(N=4, M=8, x0=0.0, x1=1.0, mesh_kind="cubical", R=0.3, L=0.2, R1=0.5, R2=1.0,
                           background="minkowski", shift=(0.0, 0.0, 0.0), ic="exact", bc="periodic", ic_width=0.15,
                           noise_amp=math.sqrt(np.finfo(np.float64).eps), eps_ko=0.0, t0=0.0, t1=1.0,
                           Nt=200, cfl=0.1, slice_y=None, slice_z=None, seed=0):
    """The first-order (Phi, Pi) ADM driver (_evolve3d_conservative,
    src/evolve.jl:1149-1366): wave3d_curved_rhs! on the device under classical
    RK4 with fixed dt = cfl h_min / max_speed (capped by 1.4 / (eps_ko mu)),
    the last step cut to end at t1; each stage is one wt2_fo_rk4_stage call
    (the right-hand side and one fused update pass).  The reference picks RK4
    for N <= 4 and higher-order Runge-Kutta methods above; this driver runs
    RK4 for every N.

    mesh_kind: "cubical" (the affine path; bc "periodic", "auto" or a 6-tuple
    of kinds, classified per face) or "cubed_cube", "inflated_cube",
    "radial_shell" (the curvilinear path; bc "sommerfeld" or "dirichlet" on
    every outer face, "periodic" meaning Sommerfeld; the radial shell's inner
    sphere is an excision surface, tag 8).  ic: "exact" (plane wave; l2_err
    against it, Dirichlet data from it), "gaussian" (width ic_width about the
    domain centre) or "noise".  Returns the reference's outputs: ts,
    ts_actual, xs_line, perm, Phis, Pis (line samples), l2_err, energy,
    Phi_final, Pi_final, dt, dx, kinds."""
    import torch
    curv = mesh_kind != "cubical"
    if mesh_kind not in ("cubical", "cubed_cube", "inflated_cube", "radial_shell"):
        raise ValueError(f"evolve3d conservative: unknown mesh_kind {mesh_kind!r} "
                         "(cubical, cubed_cube, inflated_cube, radial_shell)")
    if background == "radial_shift" and mesh_kind != "radial_shell":
        raise ValueError("evolve3d (conservative): background 'radial_shift' belongs to the radial shell "
                         "(mesh_kind='radial_shell')")
    if not curv and not (bc in ("periodic", "auto") or (isinstance(bc, (tuple, list)) and len(bc) == 6)):
        raise ValueError(f"evolve3d: bc must be 'periodic', 'auto' or a 6-tuple; got {bc!r}")
    if ic not in ("exact", "gaussian", "noise"):
        raise ValueError(f"evolve3d: unknown ic {ic!r} (exact, gaussian, noise)")
    periodic = bc == "periodic" and not curv
    ck = None
    if curv:
        ck = "sommerfeld" if bc == "periodic" else bc
        if ck not in ("sommerfeld", "dirichlet"):
            raise ValueError(f"evolve3d: curvilinear ({mesh_kind}) bc must be 'sommerfeld' or 'dirichlet'; got {bc!r}")
        if ck == "dirichlet" and ic != "exact":
            raise ValueError("evolve3d: curvilinear bc='dirichlet' needs ic='exact' (it injects the exact solution)")
    outer_tag = 1 if ck == "dirichlet" else core.SOMMERFELD_BDRY_TAG
    if mesh_kind == "cubical":
        mesh = core.make_uniform_hex(M, x0, x1, periodic)
    elif mesh_kind == "cubed_cube":
        mesh = core.make_cubed_cube_mesh(N, M, R)
        x0, x1 = -1.0, 1.0
    elif mesh_kind == "inflated_cube":
        mesh = core.make_inflated_cube_mesh(N, L, R1, R2, M, outer_tag=outer_tag)
        x0, x1 = -R2, R2
    else:
        mesh = core.make_radial_shell_mesh(N, R1, R2, M, outer_tag=outer_tag)
        x0, x1 = -R2, R2
    elem = core.make_element(N)
    coords = mesh.coords.reshape(-1, 3) if curv else core.host_coords(mesh, elem)
    n = coords.shape[0]
    bgf, phi_e, pi_e, d_e, max_speed = _background3d(background, shift, R1, R2)
    bg10 = bgf(coords)
    kinds = None
    if not curv and not periodic:
        kinds = _face_kinds(bc, bg10, mesh.M, N)
    work = core.Workspace(mesh, elem)
    coef = work.fo_coef(torch.from_numpy(bg10.reshape(-1).copy()).cuda())
    dx = work.h_min
    dt = cfl * dx / max_speed

    xg, yg, zg = (torch.from_numpy(coords[:, a].copy()).cuda() for a in range(3))
    if ic == "exact":
        Phi0 = phi_e(t0, coords[:, 0], coords[:, 1], coords[:, 2])
        Pi0 = pi_e(t0, coords[:, 0], coords[:, 1], coords[:, 2])
    elif ic == "gaussian":
        c0 = (x0 + x1) / 2
        r2 = np.sum((coords - c0) ** 2, axis=1)
        Phi0 = np.exp(-r2 / (2 * ic_width * ic_width))
        Pi0 = np.zeros(n)
    else:
        rng = np.random.default_rng(seed)
        Phi0 = noise_amp * rng.standard_normal(n)
        Pi0 = noise_amp * rng.standard_normal(n)
    Phi = torch.from_numpy(np.array(np.broadcast_to(Phi0, (n,)), dtype=np.float64)).cuda()
    Pi = torch.from_numpy(np.array(np.broadcast_to(Pi0, (n,)), dtype=np.float64)).cuda()

    withdata = ic == "exact"
    curv_dir = curv and ck == "dirichlet"
    needdata = (kinds is not None and core.BC3d._NAMES["full_dirichlet"] in kinds) or curv_dir
    gPhi = torch.zeros_like(Phi) if needdata and not curv else None
    gPi = torch.zeros_like(Phi) if needdata else None
    gD = torch.zeros(3 * n, dtype=torch.float64, device="cuda") if curv_dir else None
    if curv:
        bc3d = core.make_bc3d(ck, gPi=gPi, gD=gD,
                              excision_tag=core.EXCISION_BDRY_TAG if mesh_kind == "radial_shell" else 0)
    else:
        bc3d = None if kinds is None else core.make_bc3d(kinds, gPhi=gPhi, gPi=gPi)
    energy0 = work.fo_energy(Phi, Pi, coef)        # sets up a general mesh's metric terms
    if eps_ko != 0:
        dt = min(dt, 1.4 / (eps_ko * work.info()["mu"]))

    def set_data(t):
        if not (needdata and withdata):
            return
        gPi.copy_(pi_e(t, xg, yg, zg, lib=torch))
        if curv_dir:
            for a in range(3):
                gD[a * n:(a + 1) * n].copy_(d_e[a](t, xg, yg, zg, lib=torch))
        else:
            gPhi.copy_(phi_e(t, xg, yg, zg, lib=torch))

    def rk4_step(t, h):
        for stage, ts in ((1, t), (2, t + h / 2), (3, t + h / 2), (4, t + h)):
            set_data(ts)
            work.fo_rk4_stage(stage, Phi, Pi, coef, h, eps_ko, bc3d)

    c_mid = (x0 + x1) / 2
    y_target = c_mid if slice_y is None else float(slice_y)
    z_target = c_mid if slice_z is None else float(slice_z)
    sidx, xs_line = build_slice(coords, y_target, z_target)
    if sidx.size == 0 and curv and slice_y is None and slice_z is None:
        # no node on the centre line (even N on the inner cube): the nearest node line
        q = int(np.argmin((coords[:, 1] - y_target) ** 2 + (coords[:, 2] - z_target) ** 2))
        y_target, z_target = float(coords[q, 1]), float(coords[q, 2])
        sidx, xs_line = build_slice(coords, y_target, z_target)
    if sidx.size == 0:
        raise ValueError(f"evolve3d: slice y={y_target} z={z_target} hit no nodes")
    H = _hphys(N, mesh.coords) if curv else core.Geometry(mesh, elem).Hphys
    ts = np.linspace(t0, t1, Nt)
    Phis = np.empty((len(sidx), Nt)); Pis = np.empty_like(Phis)
    ts_actual = np.empty(Nt); l2_err = np.empty(Nt); energy = np.empty(Nt)
    t = float(t0)
    nsteps = 0
    for q, tq in enumerate(ts):
        while t < tq:
            last = t1 - t <= dt * (1 + 1e-12)   # the step that ends the interval lands on t1
            h = t1 - t if last else dt
            rk4_step(t, h)
            t = t1 if last else t + h
            nsteps += 1
        ts_actual[q] = t
        Ph, Qh = Phi.cpu().numpy(), Pi.cpu().numpy()
        if not (np.all(np.isfinite(Ph)) and np.all(np.isfinite(Qh))):
            raise FloatingPointError("evolve3d: non-finite state")
        Phis[:, q] = Ph[sidx]; Pis[:, q] = Qh[sidx]
        ref = phi_e(t, coords[:, 0], coords[:, 1], coords[:, 2]) if withdata else 0.0
        l2_err[q] = math.sqrt(float(np.sum((Ph - ref) ** 2 * H)))
        energy[q] = energy0 if (q == 0 and t == t0) else work.fo_energy(Phi, Pi, coef)
    return dict(ts=ts, ts_actual=ts_actual, xs_line=xs_line, perm=np.arange(len(sidx)), Phis=Phis, Pis=Pis,
                l2_err=l2_err, energy=energy, Phi_final=Phi.cpu().numpy(), Pi_final=Pi.cpu().numpy(),
                dt=dt, dx=dx, nsteps=nsteps, kinds=kinds if not curv else (ck,), background=background, ic=ic, bc=bc,
                mesh_kind=mesh_kind, x0=x0, x1=x1, y_target=y_target, z_target=z_target,
                integrator_name="RK4", work=work, coords=coords, coef=coef, mesh=mesh)
