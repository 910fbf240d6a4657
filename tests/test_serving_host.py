"""Host-side serving logic: arrival-order admission by memory footprint, the
FIFO latency model of the one-at-a-time baseline, request draws, and the
footprint figure (wt2_ctx_footprint is computed on the host, no device)."""
from collections import deque

import numpy as np
import pytest

import wavetoy_hip as wh
from wavetoy_hip import serving as sv


def _req(rid, N, M, curved=False, T=1.0, arrival=0.0):
    return sv.Request(rid=rid, N=N, M=M, curved=curved, centre=(0.5, 0.5, 0.5), width=0.1, T=T, arrival=arrival)


def test_admission_is_fifo_with_head_of_line_blocking():
    fp = {0: 40, 1: 50, 2: 5, 3: 10}
    q = deque(_req(i, 4, 2) for i in range(4))
    adm, rej, used = sv.admit_fifo(q, 0, 100, lambda r: fp[r.rid])
    # 40 + 50 fit; request 2 (5) would fit too, but must wait behind none -> also admitted
    assert [r.rid for r, _ in adm] == [0, 1, 2] and used == 95 and not rej
    assert [r.rid for r in q] == [3]        # 95 + 10 > 100: blocks the head
    adm, rej, used = sv.admit_fifo(q, used - 40, 100, lambda r: fp[r.rid])
    assert [r.rid for r, _ in adm] == [3] and used == 65


def test_admission_blocks_later_small_requests_behind_a_large_head():
    fp = {0: 90, 1: 1}
    q = deque([_req(0, 4, 2), _req(1, 4, 2)])
    adm, rej, used = sv.admit_fifo(q, 20, 100, lambda r: fp[r.rid])
    assert adm == [] and used == 20 and len(q) == 2


def test_admission_rejects_requests_larger_than_the_budget():
    fp = {0: 500, 1: 10}
    q = deque([_req(0, 4, 2), _req(1, 4, 2)])
    adm, rej, used = sv.admit_fifo(q, 0, 100, lambda r: fp[r.rid])
    assert [r.rid for r in rej] == [0] and [r.rid for r, _ in adm] == [1] and used == 10


def test_admission_live_point_cap():
    q = deque([_req(0, 4, 4), _req(1, 4, 4), _req(2, 4, 4)])   # 4096 points each
    adm, _, _ = sv.admit_fifo(q, 0, 10 ** 12, lambda r: 1, live_points=0, max_live_points=9000)
    assert len(adm) == 2 and len(q) == 1
    # an empty server always admits its head, even above the cap
    q = deque([_req(0, 10, 12)])
    adm, _, _ = sv.admit_fifo(q, 0, 10 ** 12, lambda r: 1, live_points=0, max_live_points=10)
    assert len(adm) == 1


def test_fifo_latencies_closed_form():
    lat = sv.fifo_latencies([0.0, 0.5, 0.6, 10.0], [1.0, 1.0, 1.0, 2.0])
    # finishes at 1, 2, 3, 12
    assert np.allclose(lat, [1.0, 1.5, 2.4, 2.0])


def test_draw_requests_distribution():
    reqs = sv.draw_requests(2000, 100.0, seed=3)
    arr = np.array([r.arrival for r in reqs])
    assert np.all(np.diff(arr) >= 0)
    gaps = np.diff(np.concatenate([[0.0], arr]))
    assert abs(gaps.mean() - 0.01) < 0.001                     # Poisson rate 100/s
    assert set(r.N for r in reqs) == {4, 6, 8, 10}
    assert min(r.M for r in reqs) == 2 and max(r.M for r in reqs) == 12
    assert sum(r.curved for r in reqs) == 1000
    assert all(0.5 <= r.T <= 2.0 and 0.05 <= r.width <= 0.2 for r in reqs)
    assert all(0.3 <= c <= 0.7 for r in reqs for c in r.centre)
    again = sv.with_arrivals(reqs[:5], 50.0, seed=1)
    assert [(r.N, r.M, r.T) for r in again] == [(r.N, r.M, r.T) for r in reqs[:5]]


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_footprint_scales_with_points_and_precision(prec):
    m = wh.make_uniform_hex(4, 0.0, 1.0)
    e = wh.make_element(6)
    flat = wh.ctx_footprint(m, e, prec)
    curved = wh.ctx_footprint(m, e, prec, with_background=True)
    n = 6 ** 3 * 4 ** 3
    es = 8 if prec == "fp64" else 4
    # two two-field register blocks (stage input S, partial sums P) + the FP64
    # Lu field (+ mixed: the two-field FP32 mirror of the state), plus small tables
    mirror = 0 if prec == "fp64" else 2 * n * 4
    assert 4 * n * es + 8 * n + mirror <= flat <= 4 * n * es + 8 * n + mirror + 16384
    # curved adds 13 FP64 + 13 FP32 coefficients, the energy weight, the face
    # flux buffer and the 5 FP64 + 5 FP32 face-node coefficients
    assert curved - flat == 13 * n * 12 + 8 * n + 64 * 6 * 36 * 8 + 5 * 64 * 6 * 36 * 12
    r = _req(0, 6, 4, curved=True, T=1.0)
    assert sv.request_footprint(r, prec, 20) >= curved + 16 * n
    assert sv.request_footprint(r, prec, 0) == curved + 16 * n
