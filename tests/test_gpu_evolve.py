"""End-to-end GPU evolutions: the plain-method CPU configuration run in full
on the device, compared with the exact standing wave, the CPU oracle, and the
energy bound (test/evolve3d_tests.jl:11-43 pattern)."""
import math

import numpy as np
import pytest

import wt_oracle as wo
import wavetoy_hip as wh

pytestmark = pytest.mark.gpu


def test_plain_cpu_configuration_full_run():
    """4^3 elements x 8^3 nodes (32,768 points), Dirichlet, tau_eff, KO eps 0.1,
    u0 = sin(pi x) sin(pi y) sin(pi z), u_t = 0, RK4 dt = 0.2 h_min, 200 steps."""
    print("plain CPU configuration, 200 steps", flush=True)
    r = wh.evolve3d(N=8, M=4, ic="cartesian", tau=20.0, eps_ko=0.1, cfl=0.2, nsteps=200, Nt=5,
                    device=True)
    assert r["tau"] == 73.5
    assert abs(r["dt"] - 0.2 * r["dx"]) < 1e-15
    t1 = r["ts"][-1]
    x = r["coords"]
    u0 = np.sin(math.pi * x[:, 0]) * np.sin(math.pi * x[:, 1]) * np.sin(math.pi * x[:, 2])
    exact = math.cos(math.sqrt(3) * math.pi * t1) * u0
    assert np.max(np.abs(r["u_final"] - exact)) < 1e-4   # measured 1.75e-5 at t = 0.641
    assert np.all(np.diff(r["energy"]) <= 1e-14 * r["energy"][0])
    assert r["energy"][-1] / r["energy"][0] > 1 - 1e-4
    o = wo.Oracle(8, 4, tau=73.5, eps_ko=0.1)
    uo, vo = o.rk4(u0, np.zeros_like(u0), r["dt"], 200)
    assert np.max(np.abs(r["u_final"] - uo)) <= 1e-12 * np.max(np.abs(uo))
    assert np.max(np.abs(r["ud_final"] - vo)) <= 1e-12 * np.max(np.abs(vo))
    # the sampled line y = z = 1/4 (src/evolve.jl:1008-1012): distinct x, sorted
    xs = r["xs_line"]
    assert len(xs) == 4 * 7 + 1 and np.all(np.diff(xs) > 0)
    assert r["y_target"] == 0.25 and r["z_target"] == 0.25
    on = (np.abs(x[:, 1] - 0.25) < 1e-12) & (np.abs(x[:, 2] - 0.25) < 1e-12)
    assert r["us"].shape == (len(xs), 5) and r["uds"].shape == (len(xs), 5)
    line0 = np.sin(math.pi * xs) * math.sin(math.pi * 0.25) ** 2
    assert np.allclose(r["us"][:, 0], line0, atol=1e-14)
    # last sample = the final state on the line
    xf = x[on]
    order = np.argsort(xf[:, 0], kind="stable")   # shared face nodes: the lower element's first
    uf_line = r["u_final"][on][order]
    _, first = np.unique(np.round(xf[order, 0], 12), return_index=True)
    assert np.allclose(r["us"][:, -1], uf_line[first], atol=1e-15)


def test_convergence_order_against_exact_solution():
    """Error at fixed t falls by > 4x per element doubling (N = 4), the
    reference's convergence criterion (test/evolve3d_tests.jl:14-21)."""
    errs = []
    for M in (2, 4, 8):
        r = wh.evolve3d(N=4, M=M, ic="cartesian", eps_ko=0.0, t1=0.25, cfl=0.1, Nt=2)
        errs.append(r["l2_err"][-1])
    print("l2 errors", errs, flush=True)
    assert errs[0] / errs[1] > 4 and errs[1] / errs[2] > 4


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_curved_gaussian_evolution_energy(precision):
    """Warped mesh + ADM bump background, Gaussian pulse (width 0.1, denominator
    2 w^2 as src/evolve.jl:1262): energy non-increasing with KO on."""
    print("curved evolution", precision, flush=True)
    r = wh.evolve3d(N=6, M=4, warp_A=0.05, background="adm_bump", ic="gaussian",
                    ic_centre=(0.45, 0.52, 0.5), ic_width=0.1, eps_ko=0.1, nsteps=60, Nt=7,
                    precision=precision)
    E = r["energy"]
    assert np.all(np.isfinite(E))
    assert np.all(np.diff(E) <= 1e-10 * E[0])
