"""Serving on the device: the footprint figure matches what a context
allocates, and every request served in the mix of a continuous batch returns
the same final u and energy history as the request run alone through the
single-instance path -- bitwise (tolerance 0), since each instance's tiles,
kernels and energy reduction are the same wherever it is packed."""
import numpy as np
import pytest

import wavetoy_hip as wh
from wavetoy_hip import serving as sv

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
@pytest.mark.parametrize("curved", [False, True])
def test_footprint_equals_allocated_bytes(prec, curved):
    r = sv.Request(rid=0, N=6, M=3, curved=curved, centre=(0.5, 0.5, 0.5), width=0.1, T=0.1)
    ws, u, v, dt, ns = sv.build_instance(r, prec)
    try:
        fp = wh.ctx_footprint(r.mesh(), wh.make_element(r.N), prec, with_background=curved)
        assert ws.info()["device_bytes"] == fp
    finally:
        ws.close()


def _requests():
    # small, uneven requests; half arrive together (and queue behind the
    # budget), the rest arrive 2 ms apart and join a running batch
    shapes = [(4, 2, False, 0.60), (6, 3, True, 0.30), (4, 3, False, 0.40), (8, 2, True, 0.20),
              (6, 2, False, 0.50), (4, 4, True, 0.20), (10, 2, False, 0.10), (6, 3, False, 0.24)]
    rng = np.random.default_rng(11)
    out = []
    for q, (N, M, curved, T) in enumerate(shapes):
        out.append(sv.Request(rid=q, N=N, M=M, curved=curved, centre=tuple(rng.uniform(0.3, 0.7, 3)),
                              width=float(rng.uniform(0.05, 0.2)), T=T, arrival=0.0 if q < 4 else 0.002 * q))
    return out


@pytest.mark.parametrize("prec", ["mixed", "fp64"])
def test_served_results_equal_alone(prec):
    reqs = _requests()
    print(f"serving {len(reqs)} requests, {prec}", flush=True)
    fps = [sv.request_footprint(r, prec, 5) for r in reqs]
    budget = int(2.2 * max(fps))            # at most two or three live at a time: requests queue
    srv = sv.Server(budget_bytes=budget, precision=prec, energy_every=5, chunk_ms=2.0)
    stats = srv.run(reqs)
    assert len(stats.responses) == len(reqs) and not any(r.error for r in stats.responses)
    assert stats.peak_bytes <= budget
    assert stats.max_live >= 2               # requests really overlapped on the device
    for resp in sorted(stats.responses, key=lambda r: r.rid):
        r = reqs[resp.rid]
        ws, u, v, dt, ns = sv.build_instance(r, prec)
        E = ws.rk4_steps(u, v, dt, ns, energy_every=5)
        ua = u.cpu().numpy()
        ws.close()
        assert resp.nsteps == ns and resp.dt == dt
        assert np.array_equal(resp.u, ua), f"request {r.rid}: u differs from the alone run"
        assert np.array_equal(resp.energy, E), f"request {r.rid}: energy history differs"
        assert resp.latency >= 0 and resp.finished >= resp.admitted >= r.arrival - 1e-9
        assert np.all(np.isfinite(resp.energy))
        if not r.curved:    # flat SIPG + KO + Dirichlet: the energy does not grow
            assert resp.energy[-1] <= resp.energy[0] * (1 + 1e-9)


def test_server_rejects_oversized_request():
    r = sv.Request(rid=0, N=4, M=2, curved=False, centre=(0.5, 0.5, 0.5), width=0.1, T=0.05)
    srv = sv.Server(budget_bytes=1000, precision="mixed")
    stats = srv.run([r])
    assert len(stats.responses) == 1 and stats.responses[0].error
