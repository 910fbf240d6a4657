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


def evolve3d(N=8, M=4, x0=0.0, x1=1.0, warp_A=0.0, background=None, ic="cartesian",
             ic_wavenumber=math.pi, ic_centre=(0.5, 0.5, 0.5), ic_width=None, ic_denom=None,
             tau=None, eps_ko=0.1, t0=0.0, t1=None, nsteps=None, dt=None, cfl=0.2, Nt=2,
             precision="fp64", energy_every=None, device=True, slice_y=None, slice_z=None):
    mesh = (core.make_warped_uniform_hex(M, M, M, x0, x1, warp_A) if warp_A
            else core.make_uniform_hex(M, x0, x1))
    elem = core.make_element(N)
    tau_used = core.tau_effective(tau, N)
    bg = None
    coords = core.host_coords(mesh, elem)
    if background == "adm_bump":
        bgobj = core.adm_bump_background()
        bg = core.sample_background3d(bgobj, 0.0, coords[:, 0], coords[:, 1], coords[:, 2])
    elif background is not None:
        bg = np.asarray(background, dtype=np.float64)
    work = core.Workspace(mesh, elem, tau=tau_used, eps_ko=eps_ko, precision=precision, bg=bg)
    if dt is None:
        dt = core.recommended_dt_rk4(work, cfl)
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
    elif ic == "gaussian":
        u0, v0 = core.gaussian_pulse(coords, ic_centre, width=ic_width, denom=ic_denom)
    else:
        raise ValueError(f"unknown ic {ic}")

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
        q = x0 + 0.25 * (x1 - x0)
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
    if mesh.warp_A == 0.0:
        H = core.Geometry(mesh, elem).Hphys

    def sample(step):
        t = t0 + step * dt
        ts.append(t)
        energies.append(work.energy(u, v))
        uh, vh = fetch(u), fetch(v)
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
                y_target=y_target, z_target=z_target,
                integrator_name="RK4", work=work, coords=coords)
