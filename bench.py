"""Benchmark: grid-point updates/s per RK4 step on the plain FP64 GPU workload
(flat cube [0,1]^3, 32^3 elements x 8^3 GLL nodes = 16,777,216 points,
Dirichlet outer faces, SIPG tau_eff = 73.5, KO eps = 0.1, Gaussian pulse
exp(-|x-x0|^2/0.01) with x0 ~ U[0.3,0.7]^3 from the seed, u_t = 0,
dt = 0.2 h_min), with the mixed FP32-RHS / FP64-state path measured beside it
on the same initial data, and the CPU oracle timed on a bounded sample.

    python bench.py [--gpus 1] [--steps 500] [--warmup 20] [--seed 0] [--no-cpu] [--repeats 5]

Prints one JSON line.  Only one GPU is used; --gpus > 1 exits with a message.
"""
import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "wavetoysecondorder.jl_amd"))
sys.path.insert(0, os.path.join(ROOT, "oracle"))

import numpy as np  # noqa: E402

N_NODES, M_ELEMS = 8, 32
EPS_KO = 0.1
TAU_LITERAL = 20.0
PEAK_HBM_GBS = 8000.0

# Compulsory HBM bytes per grid point of each pencil pass of one RK4 step
# (reads + writes of the fields the pass must touch), unfolded schedule, RK4
# in Horner form (stage s writes S = Y + c_s dt k, stage 4 writes Y):
#   fp64: the first pass reads S(u,u_t) 16 + writes P 8; the middle one adds a
#   P read 8; the final (RK) pass at stage 1 reads Y 16 + P 8, writes S 16;
#   stages 2,3 read S 16 + P 8 + Y 16 and write S 16; stage 4 reads S 16 +
#   P 8 + Y 16, writes Y 16.
PASS_BYTES = {
    "fp64": {"first": [24, 24, 24, 24], "mid": [32, 32, 32, 32], "final": [40, 56, 56, 56]},
    #   mixed: S and P in FP32; Y in FP64, and its FP32 mirror Yf (made at the
    #   start of a run) in place of Y wherever the result is an FP32 stage
    #   input: stage 1 reads Yf as S, stages 2, 3 read Yf; stage 4 reads
    #   S 8 + P 4 + Y 16 and writes Y 16 + Yf 8.
    "mixed": {"first": [12, 12, 12, 12], "mid": [16, 16, 16, 16], "final": [20, 28, 28, 52]},
}
# Whole-step minimum of the classical 3-register stage form (Y, S, A: 16
# field-pair accesses per point): the algorithmic bytes the roofline uses.
FUSED_MIN_BYTES = {"fp64": 256, "mixed": 208}
# Whole-step minimum of the Horner form the kernels run (Y, S: 11 field-pair
# accesses): stage 1 Y -> S, stages 2, 3 S, Y -> S, stage 4 S, Y -> Y (mixed:
# Yf in place of Y at stages 1-3, stage 4 also writes Yf).
HORNER_MIN_BYTES = {"fp64": 176, "mixed": 112}


def step_bytes(prec, launches=12):
    """Bytes per point the passes of one step move: 12 launches as tabled, or
    8 (folded: the 4 first passes gone, +8 B (FP64) / 4 B (mixed) P write in
    each RK pass)."""
    t = PASS_BYTES[prec]
    full = sum(sum(t[k]) for k in ("first", "mid", "final"))
    if launches == 12:
        return full
    return full - sum(t["first"]) + 4 * (8 if prec == "fp64" else 4)


def initial_data(wh, mesh, elem, seed):
    rng = np.random.default_rng(seed)
    x0 = rng.uniform(0.3, 0.7, size=3)
    x = wh.host_coords(mesh, elem)
    u0, v0 = wh.gaussian_pulse(x, x0, denom=0.01)
    return u0, v0, x0


def run_gpu(wh, torch, prec, u0, v0, steps, warmup, stream, tune=True, repeats=0):
    mesh = wh.make_uniform_hex(M_ELEMS, 0.0, 1.0)
    elem = wh.make_element(N_NODES)
    tau = wh.tau_effective(TAU_LITERAL, N_NODES)
    ws = wh.Workspace(mesh, elem, tau=tau, eps_ko=EPS_KO, precision=prec)
    ws.set_stream(stream.cuda_stream)
    dt = wh.recommended_dt_rk4(ws, 0.2)
    with torch.cuda.stream(stream):
        u0d = torch.from_numpy(u0).to("cuda")
        v0d = torch.from_numpy(v0).to("cuda")
        u, v = u0d.clone(), v0d.clone()
    stream.synchronize()
    if tune:
        ws.autotune(u, v, dt)          # tile shapes per axis (cached per shape)
    # warm-up (captures the step graph for these state pointers), then restart
    # from the initial data in the same buffers
    ws.rk4_steps(u, v, dt, warmup)
    with torch.cuda.stream(stream):
        u.copy_(u0d)
        v.copy_(v0d)
    del u0d, v0d
    ev0 = torch.cuda.Event(enable_timing=True)
    ev1 = torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ev0.record(stream)
    ws.rk4_steps(u, v, dt, steps)       # returns after a stream sync
    ev1.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ev_ms = ev0.elapsed_time(ev1)
    # per-pass timing (diagnostics): each pencil pass of stage 2 alone, repeated
    pass_ms = {}
    reps = 10
    # (with the folded schedule stage 2 runs z and x, stage 3 z and y)
    for name, axis, stage in (("X", 0, 2), ("Y", 1, 3), ("Z", 2, 2)):
        a = torch.cuda.Event(enable_timing=True)
        b = torch.cuda.Event(enable_timing=True)
        ws.launch_pass(axis, stage, u, v, dt)   # warm
        a.record(stream)
        for _ in range(reps):
            ws.launch_pass(axis, stage, u, v, dt)
        b.record(stream)
        torch.cuda.synchronize()
        pass_ms[name] = a.elapsed_time(b) / reps
    # the drop-in right-hand side (wt2_rhs3d's kernel) alone, same way (more
    # launches: one is about 0.1 ms)
    a = torch.cuda.Event(enable_timing=True)
    b = torch.cuda.Event(enable_timing=True)
    for _ in range(5):
        ws.launch_pass(3, 1, u, v, dt)
    rreps = 100
    a.record(stream)
    for _ in range(rreps):
        ws.launch_pass(3, 1, u, v, dt)
    b.record(stream)
    torch.cuda.synchronize()
    rhs_ms = a.elapsed_time(b) / rreps
    # repeats (diagnostics, after the timed region): `repeats` more runs of
    # `steps` steps on a copy of the state, for the median and minimum
    rep_ms = []
    if repeats > 0:
        with torch.cuda.stream(stream):
            ur, vr = u.clone(), v.clone()
        ws.rk4_steps(ur, vr, dt, 1)          # graph capture for these pointers
        for _ in range(repeats):
            a.record(stream)
            ws.rk4_steps(ur, vr, dt, steps)
            b.record(stream)
            torch.cuda.synchronize()
            rep_ms.append(a.elapsed_time(b) / steps)
        del ur, vr
    ws.set_stream(None)
    E = ws.energy(u, v)
    lps = ws.info()["launches_per_step"]
    return dict(ws=ws, u=u.cpu().numpy(), v=v.cpu().numpy(), dt=dt, wall=wall, ev_ms=ev_ms, lps=lps,
                pass_ms=pass_ms, rhs_ms=rhs_ms, energy=E, npts=ws.npoints, h_min=ws.h_min, mu=ws.mu, tau=tau,
                rep_ms=rep_ms)


def repeat_summary(rep_ms):
    """Median and minimum ms per step over the repeat runs (None without repeats)."""
    if not rep_ms:
        return None
    return {"runs": len(rep_ms), "median": float(np.median(rep_ms)), "min": float(min(rep_ms))}


def cpu_baseline(u0, v0, dt, budget_s=12.0):
    import wt_oracle as wo
    o = wo.Oracle(N_NODES, M_ELEMS, tau=73.5, eps_ko=EPS_KO)
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    t0 = time.perf_counter()
    o.rk4(u0, v0, dt, 1)
    t1 = time.perf_counter() - t0
    n = max(1, min(5, int(budget_s / max(t1, 1e-3))))
    if n > 1:
        t0 = time.perf_counter()
        o.rk4(u0, v0, dt, n)
        t1 = (time.perf_counter() - t0) / n
    return dict(value=o.n / t1, unit="grid-point updates/s", cores=threads, kind="port",
                sample=f"CPU oracle (C, OpenMP, FP64), {n} RK4 step(s) of the full 32^3 x 8^3 "
                       f"mesh after one untimed step; {t1:.2f} s per step on {threads} threads")


def curved_run(wh, torch, args):
    """Warped cube x_i = xi_i + 0.05 sin(2 pi xi1) sin(2 pi xi2) sin(2 pi xi3), 24^3
    elements x 10^3 nodes (13,824,000 points), ADM bump background, Gaussian
    pulse width 0.1 (denominator 2 w^2) with seeded centre in [0.3,0.7]^3,
    Dirichlet, tau_eff, KO 0.1, RK4 dt = 0.2 h_min; FP64 and mixed timed on the
    same data, mixed error against FP64 at the end."""
    N, M = 10, 24
    mesh = wh.make_warped_uniform_hex(M, M, M, 0.0, 1.0, 0.05)
    elem = wh.make_element(N)
    x = wh.host_coords(mesh, elem)
    bg = wh.sample_background3d(wh.adm_bump_background(), 0.0, x[:, 0], x[:, 1], x[:, 2])
    rng = np.random.default_rng(args.seed)
    c = rng.uniform(0.3, 0.7, size=3)
    u0, v0 = wh.gaussian_pulse(x, c, width=0.1)
    del x
    tau = wh.tau_effective(TAU_LITERAL, N)
    res = {}
    stream = torch.cuda.Stream()
    for prec in ("fp64", "mixed"):
        print(f"curved: {prec}", file=sys.stderr, flush=True)
        ws = wh.Workspace(mesh, elem, tau=tau, eps_ko=EPS_KO, precision=prec, bg=bg)
        ws.set_stream(stream.cuda_stream)
        dt = wh.recommended_dt_rk4(ws, 0.2)
        steps = args.steps if args.t_final is None else int(math.ceil(args.t_final / dt - 1e-9))
        if args.t_final is not None:
            dt = args.t_final / steps
        with torch.cuda.stream(stream):
            u = torch.from_numpy(u0).to("cuda")
            v = torch.from_numpy(v0).to("cuda")
        stream.synchronize()
        if not args.no_tune:
            ws.autotune(u, v, dt)                             # tile shapes per axis (cached per shape)
        ws.rk4_steps(u.clone(), v.clone(), dt, 2)          # warm-up + graph capture
        E0 = ws.energy(u, v)
        t0 = time.perf_counter()
        chunk = 500
        done = 0
        while done < steps:
            k = min(chunk, steps - done)
            ws.rk4_steps(u, v, dt, k)
            done += k
        wall = time.perf_counter() - t0
        E1 = ws.energy(u, v)
        # the drop-in RHS alone (wt2_rhs3d's two kernels: face fluxes, volume)
        ea = torch.cuda.Event(enable_timing=True)
        eb = torch.cuda.Event(enable_timing=True)
        for _ in range(3):
            ws.launch_pass(3, 1, u, v, dt)
        ea.record(stream)
        for _ in range(20):
            ws.launch_pass(3, 1, u, v, dt)
        eb.record(stream)
        torch.cuda.synchronize()
        rhs_ms = ea.elapsed_time(eb) / 20
        res[prec] = dict(ms_per_step=wall * 1e3 / steps, steps=steps, dt=dt,
                         value=ws.npoints * steps / wall, energy0=E0, energy1=E1,
                         u=u.cpu().numpy(), h_min=ws.h_min, rhs_ms=rhs_ms)
        ws.set_stream(None)
        ws.close()
    o = wh.Workspace(mesh, elem, tau=tau, bg=bg)
    # H-norm from the energy weights: |d|_H^2 via the v^2 term of the energy with u = 0
    d = res["mixed"]["u"] - res["fp64"]["u"]
    z = np.zeros_like(d)
    hn_d = o.energy(z, d)
    hn_u = o.energy(z, res["fp64"]["u"])
    out = {
        "config": {"workload": "warped cube A=0.05, 24^3 x 10^3 (13,824,000 points), ADM bump "
                               "background, Gaussian width 0.1, Dirichlet, tau_eff=121.5, KO 0.1",
                   "points": o.npoints, "seed": args.seed, "pulse_centre": [float(q) for q in c]},
        "fp64": {k: v for k, v in res["fp64"].items() if k != "u"},
        "mixed": {k: v for k, v in res["mixed"].items() if k != "u"},
        "mixed_speedup": res["mixed"]["value"] / res["fp64"]["value"],
        "mixed_rel_err_weighted_l2_vs_fp64": math.sqrt(hn_d / hn_u),
        "mixed_max_rel_err_vs_fp64": float(np.max(np.abs(d)) / np.max(np.abs(res["fp64"]["u"]))),
        "t_final": res["fp64"]["steps"] * res["fp64"]["dt"],
        # drop-in RHS: u, u_t in, u_tt out (24 B) + 13 static coefficients per
        # node (104 B FP64, 52 B FP32)
        "rhs_roofline": {prec: {"ms": res[prec]["rhs_ms"], "algorithmic_bytes_per_point": bpp,
                                "achieved": bpp * o.npoints / (res[prec]["rhs_ms"] * 1e-3) / 1e9,
                                "peak": PEAK_HBM_GBS, "unit": "GB/s",
                                "frac": bpp * o.npoints / (res[prec]["rhs_ms"] * 1e-3) / 1e9 / PEAK_HBM_GBS}
                         for prec, bpp in (("fp64", 128), ("mixed", 76))},
    }
    return out


def make_instance(wh, torch, spec, precision):
    """One evolution of the batched workload: M^3 elements x N^3 nodes, flat or
    warped + ADM bump background, Gaussian pulse; returns (workspace, u, u_t, dt, nsteps)."""
    N, M, curved, centre, width, T = spec
    if curved:
        mesh = wh.make_warped_uniform_hex(M, M, M, 0.0, 1.0, 0.05)
    else:
        mesh = wh.make_uniform_hex(M, 0.0, 1.0)
    elem = wh.make_element(N)
    x = wh.host_coords(mesh, elem)
    bg = wh.sample_background3d(wh.adm_bump_background(), 0.0, x[:, 0], x[:, 1], x[:, 2]) if curved else None
    ws = wh.Workspace(mesh, elem, tau=wh.tau_effective(TAU_LITERAL, N), eps_ko=EPS_KO,
                      precision=precision, bg=bg)
    u0, v0 = wh.gaussian_pulse(x, centre, width=width)
    dt = 0.2 * ws.h_min
    nsteps = max(1, int(math.ceil(T / dt - 1e-9)))
    dt = T / nsteps
    return ws, torch.from_numpy(u0).cuda(), torch.from_numpy(v0).cuda(), dt, nsteps


def batch_specs(n, seed, t_final=1.0):
    rng = np.random.default_rng(seed)
    specs = []
    for q in range(n):
        M = int(rng.integers(2, 13))
        N = int(rng.choice([4, 6, 8, 10]))
        centre = rng.uniform(0.3, 0.7, size=3)
        width = float(rng.uniform(0.05, 0.2))
        specs.append((N, M, q % 2 == 1, centre, width, t_final))
    return specs


def run_batch(wh, torch, specs, precision, chunk=64):
    """All instances in one wt2_batch, stepped to their ends; returns the live
    instances (workspaces still open), the point-updates, the wall time of the
    stepping, the tuned shapes and the launches of one RK4 step with every
    instance live."""
    insts = [make_instance(wh, torch, sp, precision) for sp in specs]
    shapes = {}
    for (ws, u, v, dt, ns), sp in zip(insts, specs):
        key = (sp[0], sp[1], sp[2])
        if key not in shapes:   # autotune once per shape; later workspaces reuse the cached choice
            shapes[key] = ws.autotune(u.clone(), v.clone(), dt)
    # workspaces created before their shape was tuned pick the choice up now
    for (ws, u, v, dt, ns), sp in zip(insts, specs):
        ws.autotune(u, v, dt)
    b = wh.Batch()
    for ws, u, v, dt, ns in insts:
        b.add(ws, u, v, dt, ns)
    work = sum(ws.npoints * ns for ws, u, v, dt, ns in insts)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    fin = b.step(1)
    launches = b.launches_per_step()
    while fin < len(insts):
        fin += b.step(chunk)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    b.close()
    return insts, work, wall, shapes, launches


def batch_run(wh, torch, args):
    """The batched workload: args.instances independent evolutions, per instance
    M ~ U{2..12} elements per side and N ~ U{4,6,8,10}, half flat, half warped
    with the ADM bump background, Gaussian pulse (centre U[0.3,0.7]^3, width
    U[0.05,0.2], amplitude 1), RK4 with dt = 0.2 h_min to t = 1 (or --t-final).
    FP64 first, then mixed; every instance's mixed result against its FP64 one
    (relative error in the energy norm of the u_t term, which is the H norm,
    and in the max norm)."""
    T = 1.0 if args.t_final is None else args.t_final
    specs = batch_specs(args.instances, args.seed, T)
    out = {"config": {"workload": f"{len(specs)} instances, M~U{{2..12}}, N~U{{4,6,8,10}}, half warped+ADM, "
                                  f"Gaussian pulses, RK4 dt=0.2 h_min to t={T}, one wt2_batch per precision "
                                  f"(hipGraph per step)", "seed": args.seed}}
    u64 = []
    for prec in ("fp64", "mixed"):
        print(f"batch: {len(specs)} instances, {prec}", file=sys.stderr, flush=True)
        insts, work, wall, shapes, launches = run_batch(wh, torch, specs, prec)
        out[prec] = {"point_updates": work, "wall_s": wall, "value": work / wall,
                     "total_points": sum(i[0].npoints for i in insts),
                     "distinct_shapes_tuned": len(shapes), "launches_per_step_all_live": launches}
        print(f"batch: {prec} {work / wall:.3e} point-updates/s, {wall:.1f} s, "
              f"{launches} launches per step", file=sys.stderr, flush=True)
        if prec == "fp64":
            u64 = [i[1].cpu().numpy() for i in insts]
        else:
            out["tiles_example"] = {f"N{k[0]}M{k[1]}{'c' if k[2] else 'f'}": v for k, v in list(shapes.items())[:6]}
            hn, mx = [], []
            for (ws, u, v, dt, ns), ref in zip(insts, u64):
                r = torch.from_numpy(ref).cuda()
                d = (u - r).contiguous()
                z = torch.zeros_like(d)
                hn.append(math.sqrt(ws.energy(z, d) / ws.energy(z, r)))
                mx.append(float((d.abs().max() / r.abs().max()).item()))
            w = int(np.argmax(hn))
            out["mixed_vs_fp64_per_instance"] = {
                "hnorm_rel_err_max": max(hn), "hnorm_rel_err_median": float(np.median(hn)),
                "max_rel_err_max": max(mx), "max_rel_err_median": float(np.median(mx)),
                "worst_instance": {"N": specs[w][0], "M": specs[w][1], "warped": bool(specs[w][2]),
                                   "width": specs[w][4]}}
        for i in insts:
            i[0].close()
        del insts
        torch.cuda.empty_cache()
    out["mixed_speedup"] = out["mixed"]["value"] / out["fp64"]["value"]
    # the plain FP64 single-instance path (wt2_rk4_steps, one instance at a
    # time, tuned tiles, state resident) on a quarter of the same list (the
    # flat and the warped instance of every fourth pair, so both kinds keep
    # their share): its point-updates/s against the mixed batch's
    sample = [sp for q, sp in enumerate(specs) if q % 8 < 2]
    print(f"batch: {len(sample)} of the instances one at a time, fp64", file=sys.stderr, flush=True)
    work1, wall1 = 0, 0.0
    for sp in sample:
        ws, u, v, dt, ns = make_instance(wh, torch, sp, "fp64")
        ws.autotune(u.clone(), v.clone(), dt)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        ws.rk4_steps(u, v, dt, ns)
        torch.cuda.synchronize()
        wall1 += time.perf_counter() - t0
        work1 += ws.npoints * ns
        ws.close()
    out["one_at_a_time_fp64"] = {"instances": len(sample), "sample": "instances q with q mod 8 in {0, 1} (a flat and a warped one of every fourth pair)",
                                 "point_updates": work1, "wall_s": wall1, "value": work1 / wall1}
    out["mixed_batch_vs_one_at_a_time_fp64"] = out["mixed"]["value"] / (work1 / wall1)
    return out


def serving_run(wh, torch, args, rates=None):
    """The serving workload: requests drawn like the batched workload with
    T ~ U[0.5,2], Poisson arrivals at each rate, served by the continuous-
    batching server (mixed precision, 200 GB budget); beside it the same
    requests one at a time through the plain FP64 single-instance path, whose
    latencies follow from a FIFO queue on the same arrival times."""
    from wavetoy_hip import serving as sv
    reqs = sv.draw_requests(args.requests, None, args.seed)
    work = sum(r.npoints for r in reqs)
    mean_work = float(np.mean([sv.request_work_estimate(r) for r in reqs]))
    srv = sv.Server(budget_bytes=200e9, precision="mixed", energy_every=20)
    t0 = time.perf_counter()
    srv.prewarm(reqs, verbose=True)
    prewarm_s = time.perf_counter() - t0
    print(f"serving: {len(reqs)} requests, {work} points, tuned in {prewarm_s:.1f} s", file=sys.stderr, flush=True)
    out = {"config": {"workload": f"{len(reqs)} requests drawn like the batched workload, T~U[0.5,2], "
                                  "Poisson arrivals; continuous batching, mixed precision, 200 GB budget",
                      "seed": args.seed, "requests": len(reqs), "tile_tuning_s_before_start": prewarm_s,
                      "mean_request_point_updates": mean_work}}
    fp64_resp, service, work64 = sv.run_one_at_a_time(reqs, "fp64", energy_every=20)
    out["one_at_a_time_fp64"] = {"point_updates": work64, "busy_s": float(sum(service)),
                                 "throughput_point_updates_per_s": work64 / sum(service),
                                 "saturation_rate_per_s": len(reqs) / sum(service)}
    print(f"serving: one-at-a-time FP64 {sum(service):.1f} s", file=sys.stderr, flush=True)
    if rates is None and args.rates == "auto":
        # saturation of the continuous batch: one closed-loop run with every
        # request present at t = 0 gives the batch's own service rate
        st = srv.run(sv.with_arrivals(reqs, 1e9, args.seed))
        sat = len(reqs) / st.wall
        out["saturation"] = {"rate_per_s": sat, "closed_run_wall_s": st.wall,
                             "throughput_point_updates_per_s": st.summary()["throughput_point_updates_per_s"]}
        print(f"serving: saturation {sat:.2f} requests/s", file=sys.stderr, flush=True)
        rates = [round(f * sat, 2) for f in (0.25, 0.5, 0.8)] + [2.0 * round(sat, 2)]
    rates = rates or [float(r) for r in args.rates.split(",")]
    by_rid = {r.rid: r for r in fp64_resp}
    out["rates"] = {}
    for lam in rates:
        rq = sv.with_arrivals(reqs, lam, args.seed + int(lam))
        st = srv.run(rq)
        summ = st.summary()
        lat_seq = sv.fifo_latencies([r.arrival for r in rq], service)
        # mixed vs FP64 result of each request: relative H-free L2 error of the final u
        errs = []
        for resp in st.responses:
            ref = by_rid[resp.rid].u
            errs.append(float(np.linalg.norm(resp.u - ref) / max(np.linalg.norm(ref), 1e-300)))
        summ["one_at_a_time_fp64_p50_s"] = float(np.percentile(lat_seq, 50))
        summ["one_at_a_time_fp64_p99_s"] = float(np.percentile(lat_seq, 99))
        summ["throughput_vs_one_at_a_time_fp64"] = summ["throughput_point_updates_per_s"] / \
            out["one_at_a_time_fp64"]["throughput_point_updates_per_s"]
        errs = np.asarray(errs)
        w = int(np.argmax(errs))
        wr = reqs[st.responses[w].rid]
        summ["mixed_vs_fp64_worst_request"] = {"N": wr.N, "M": wr.M, "warped": wr.curved, "T": wr.T,
                                               "width": wr.width, "final_u_l2": float(np.linalg.norm(by_rid[wr.rid].u))}
        summ["mixed_vs_fp64_rel_l2_err_max"] = float(np.max(errs))
        summ["mixed_vs_fp64_rel_l2_err_median"] = float(np.median(errs))
        summ["mixed_vs_fp64_nonfinite"] = int(np.sum(~np.isfinite(errs)))
        summ["rate_per_s"] = lam
        out["rates"][f"{lam:g}"] = summ
        print(f"serving: rate {lam}: p50 {summ['p50_latency_s']:.3f} s p99 {summ['p99_latency_s']:.3f} s "
              f"{summ['throughput_point_updates_per_s']:.3e} pt-upd/s", file=sys.stderr, flush=True)
    return out


KIND_NAMES = ("lds", "shfl", "xchg")


def passes_run(wh, torch, args):
    """Each pencil-pass kernel kind and tile size of each axis, stage 2 of the
    RK4 step, timed alone (10 launches, HIP events on the pass stream): the
    plain workload (32^3 x 8^3, FP64 and mixed) and the KO-only passes of the
    curved workload (24^3 x 10^3).  With --no-sweep: the default tiles only,
    each pass launched 3 times."""
    out = {}
    cases = [(8, 32, "fp64", False), (8, 32, "mixed", False), (10, 24, "fp64", True), (10, 24, "mixed", True)]
    for N, M, prec, curved in cases:
        mesh = (wh.make_warped_uniform_hex(M, M, M, 0.0, 1.0, 0.05) if curved
                else wh.make_uniform_hex(M, 0.0, 1.0))
        elem = wh.make_element(N)
        ws = wh.Workspace(mesh, elem, tau=wh.tau_effective(TAU_LITERAL, N), eps_ko=EPS_KO, precision=prec)
        st = torch.cuda.Stream()
        ws.set_stream(st.cuda_stream)
        gen = torch.Generator(device="cuda").manual_seed(args.seed)
        u = torch.randn(ws.npoints, dtype=torch.float64, device="cuda", generator=gen) * 1e-3
        v = torch.randn(ws.npoints, dtype=torch.float64, device="cuda", generator=gen) * 1e-3
        dt = 0.2 * ws.h_min
        ws.rk4_steps(u, v, dt, 1)           # initialise the stage registers
        torch.cuda.synchronize()
        base = ws.tiles()
        key = f"N{N}_M{M}_{prec}{'_curved_ko' if curved else ''}"
        res = {"default": base}
        for axis in (2, 1, 0):
            if args.no_sweep:
                for _ in range(3):
                    ws.launch_pass(axis, 2, u, v, dt)
                continue
            times = {}
            for kind in (2, 1, 0):
                for nb in range(1, N + 1):
                    nbs = list(base["planes_per_tile"])
                    ks = [KIND_NAMES.index(k) for k in base["kernel"]]
                    nbs[axis] = nb
                    ks[axis] = kind
                    try:
                        ws.set_tiles(nbs, ks)
                    except wh.WT2Error:
                        continue                # not a valid tile for this shape
                    ws.launch_pass(axis, 2, u, v, dt)
                    a = torch.cuda.Event(enable_timing=True)
                    b = torch.cuda.Event(enable_timing=True)
                    a.record(st)
                    for _ in range(10):
                        ws.launch_pass(axis, 2, u, v, dt)
                    b.record(st)
                    torch.cuda.synchronize()
                    times[f"{KIND_NAMES[kind]}{nb}"] = round(a.elapsed_time(b) / 10 * 1e3, 1)
            res["xyz"[axis]] = times
        torch.cuda.synchronize()
        ws.close()
        out[key] = res
    return {"config": {"workload": "pencil passes of RK4 stage 2, each kernel kind and tile (us per launch)"},
            "passes": out}


FO_ALG_BYTES = 120   # first-order RHS: read Phi, Pi and 11 coefficients, write Phidot, Pidot (FP64)


def fo_run(wh, torch, args):
    """The first-order (Phi, Pi) ADM right-hand side (wt2_fo_rhs) on the flat
    32^3 x 8^3 mesh with the ADM bump background, KO 0.1 and Sommerfeld faces:
    RHS evaluations per second, with the CPU oracle timed on 16^3 x 8^3."""
    mesh = wh.make_uniform_hex(M_ELEMS, 0.0, 1.0)
    elem = wh.make_element(N_NODES)
    ws = wh.Workspace(mesh, elem)
    x = wh.host_coords(mesh, elem)
    bg = wh.sample_background3d(wh.adm_bump_background(), 0.0, x[:, 0], x[:, 1], x[:, 2])
    coef = ws.fo_coef(torch.from_numpy(bg.reshape(-1).copy()).cuda())
    Phi0, _ = wh.gaussian_pulse(x, (0.5, 0.45, 0.55), width=0.1)
    Phi = torch.from_numpy(Phi0).cuda()
    Pi = torch.zeros_like(Phi)
    dP = torch.empty_like(Phi); dQ = torch.empty_like(Phi)
    bc = wh.make_bc3d(["sommerfeld"] * 6)
    stream = torch.cuda.Stream()
    ws.set_stream(stream.cuda_stream)
    K = max(1, min(args.steps, 200))
    with torch.cuda.stream(stream):
        for _ in range(max(args.warmup, 2)):
            ws.fo_rhs(dP, dQ, Phi, Pi, coef, 0.1, bc)
        torch.cuda.synchronize()
        a = torch.cuda.Event(enable_timing=True); b = torch.cuda.Event(enable_timing=True)
        t0 = time.perf_counter()
        a.record(stream)
        for _ in range(K):
            ws.fo_rhs(dP, dQ, Phi, Pi, coef, 0.1, bc)
        b.record(stream)
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
    ev_ms = a.elapsed_time(b) / K
    ws.set_stream(None)
    npts = ws.npoints
    out = {"metric": "first-order (Phi, Pi) RHS evaluations: grid points/s", "value": npts * K / wall,
           "unit": "grid points/s", "n_gpus": 1, "steps": K, "warmup": args.warmup,
           "ms_per_step": wall * 1e3 / K, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
           "dtype": "f64", "data": "synthetic (Gaussian Phi, Pi = 0, ADM bump background)",
           "config": {"workload": "wave3d_curved_rhs! affine path, flat cube 32^3 elements x 8^3 GLL nodes "
                                  "(16,777,216 points), ADM bump background, KO eps=0.1, Sommerfeld on six faces",
                      "points": npts},
           "roofline": {"bound": "hbm", "achieved": FO_ALG_BYTES * npts / (ev_ms * 1e-3) / 1e9, "peak": PEAK_HBM_GBS,
                        "unit": "GB/s", "frac": FO_ALG_BYTES * npts / (ev_ms * 1e-3) / 1e9 / PEAK_HBM_GBS,
                        "traffic": None, "kernel": "fo_pass_kernel<8> (six passes per RHS)",
                        "algorithmic_bytes_per_point": FO_ALG_BYTES, "event_ms_per_rhs": ev_ms}}
    if not args.no_cpu:
        import wt_oracle as wo
        Mc = 16
        o = wo.Oracle(N_NODES, Mc)
        xc = o.coords()
        bgc = wh.sample_background3d(wh.adm_bump_background(), 0.0, xc[:, 0], xc[:, 1], xc[:, 2])
        cc = wo.fo_coef(bgc.reshape(-1))
        Pc = np.exp(-np.sum((xc - 0.5) ** 2, axis=1) / 0.02)
        Qc = np.zeros_like(Pc)
        bcc = wo.make_bc3d(["sommerfeld"] * 6)
        o.fo_rhs(Pc, Qc, cc, 0.1, bcc)
        t0 = time.perf_counter()
        reps = 3
        for _ in range(reps):
            o.fo_rhs(Pc, Qc, cc, 0.1, bcc)
        tc = (time.perf_counter() - t0) / reps
        out["cpu_baseline"] = {"value": o.n / tc, "unit": "grid points/s",
                               "cores": int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1)),
                               "kind": "port", "sample": f"CPU oracle (C, OpenMP, FP64), {reps} first-order RHS "
                                                         f"evaluations on 16^3 x 8^3 elements, {tc:.2f} s each"}
    return out


FO_CURV_ALG_BYTES = 200   # read Phi, Pi, 11 coefficients, 9 metric terms J a^i_n and |J|; write Phidot, Pidot


def fo_curv_run(wh, torch, args):
    """The first-order (Phi, Pi) system on a curvilinear multi-patch mesh (the
    metric path of wave3d_curved_rhs!): the inflated cube, 13 patches of 8^3
    elements, N = 8 (3,407,872 points), ADM bump background, Sommerfeld outer
    sphere, no KO (the reference's default): RHS evaluations (wt2_fo_rhs) and
    RK4 steps of fused stages (wt2_fo_rk4_stage) per second, the CPU oracle
    timed on the same mesh family with 4^3 elements per patch."""
    mesh = wh.make_inflated_cube_mesh(N_NODES, 0.1, 0.3, 1.0, 8, outer_tag=7)
    ws = wh.Workspace(mesh, wh.make_element(N_NODES))
    x = mesh.coords.reshape(-1, 3)
    bg = wh.sample_background3d(wh.adm_bump_background(), 0.0, x[:, 0], x[:, 1], x[:, 2])
    coef = ws.fo_coef(torch.from_numpy(bg.reshape(-1).copy()).cuda())
    Phi0, _ = wh.gaussian_pulse(x, (0.1, -0.05, 0.05), width=0.1)
    Phi = torch.from_numpy(Phi0).cuda()
    Pi = torch.zeros_like(Phi)
    dP = torch.empty_like(Phi); dQ = torch.empty_like(Phi)
    bc = wh.make_bc3d("sommerfeld")
    stream = torch.cuda.Stream()
    ws.set_stream(stream.cuda_stream)
    K = max(1, min(args.steps, 200))
    dt = 0.2 * ws.h_min
    npts = ws.npoints
    with torch.cuda.stream(stream):
        for _ in range(max(args.warmup, 2)):
            ws.fo_rhs(dP, dQ, Phi, Pi, coef, 0.0, bc)
        torch.cuda.synchronize()
        a = torch.cuda.Event(enable_timing=True); b = torch.cuda.Event(enable_timing=True)
        t0 = time.perf_counter()
        a.record(stream)
        for _ in range(K):
            ws.fo_rhs(dP, dQ, Phi, Pi, coef, 0.0, bc)
        b.record(stream)
        torch.cuda.synchronize()
        wall_rhs = time.perf_counter() - t0
        ev_rhs = a.elapsed_time(b) / K
        for _ in range(max(args.warmup, 1)):
            for st in range(1, 5):
                ws.fo_rk4_stage(st, Phi, Pi, coef, dt, 0.0, bc)
        torch.cuda.synchronize()
        c = torch.cuda.Event(enable_timing=True); d = torch.cuda.Event(enable_timing=True)
        t0 = time.perf_counter()
        c.record(stream)
        for _ in range(K):
            for st in range(1, 5):
                ws.fo_rk4_stage(st, Phi, Pi, coef, dt, 0.0, bc)
        d.record(stream)
        torch.cuda.synchronize()
        wall_step = time.perf_counter() - t0
        ev_step = c.elapsed_time(d) / K
    assert bool(torch.isfinite(Phi).all()) and bool(torch.isfinite(Pi).all())
    ws.set_stream(None)
    gbs = FO_CURV_ALG_BYTES * npts / (ev_rhs * 1e-3) / 1e9
    out = {"metric": "first-order (Phi, Pi) RK4 steps on a curvilinear mesh: grid-point updates/s",
           "value": npts * K / wall_step, "unit": "grid-point updates/s", "n_gpus": 1, "steps": K,
           "warmup": args.warmup, "ms_per_step": wall_step * 1e3 / K, "event_ms_per_step": ev_step,
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
           "data": "synthetic (Gaussian Phi, Pi = 0, ADM bump background)",
           "config": {"workload": "wave3d_curved_rhs! metric path, inflated cube 13 patches of 8^3 elements, "
                                  "N = 8 (3,407,872 points), ADM bump background, Sommerfeld outer sphere, "
                                  "KO off, dt = 0.2 h_min", "points": npts, "dt": dt},
           "rhs": {"ms": wall_rhs * 1e3 / K, "event_ms": ev_rhs, "points_per_s": npts * K / wall_rhs},
           "roofline": {"bound": "hbm", "achieved": gbs, "peak": PEAK_HBM_GBS, "unit": "GB/s",
                        "frac": gbs / PEAK_HBM_GBS, "traffic": None,
                        "kernel": "fc_grad_kernel<8> + fc_div_kernel<8> (one RHS)",
                        "algorithmic_bytes_per_point": FO_CURV_ALG_BYTES}}
    ws.close()
    if not args.no_cpu:
        import wt_oracle as wo
        small = wh.make_inflated_cube_mesh(N_NODES, 0.1, 0.3, 1.0, 4, outer_tag=7)
        o = wo.Oracle.from_mesh(small.N, small.coords, small.neighbour, small.neighbour_face, small.bdry)
        xc = o.coords()
        bgc = wh.sample_background3d(wh.adm_bump_background(), 0.0, xc[:, 0], xc[:, 1], xc[:, 2])
        cc = wo.fo_coef(bgc.reshape(-1))
        Pc = np.exp(-np.sum(xc ** 2, axis=1) / 0.02)
        Qc = np.zeros_like(Pc)
        bcc = wo.make_bc3d("sommerfeld")
        o.fo_rhs_curv(Pc, Qc, cc, 0.0, bcc)
        reps, t0 = 0, time.perf_counter()
        while reps < 3 or time.perf_counter() - t0 < 5.0:
            o.fo_rhs_curv(Pc, Qc, cc, 0.0, bcc)
            reps += 1
        tc = (time.perf_counter() - t0) / reps
        out["cpu_baseline"] = {"value": o.n / tc, "unit": "grid points/s (RHS evaluations)",
                               "cores": int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1)),
                               "kind": "port", "sample": f"CPU oracle (C, FP64), {reps} curvilinear first-order "
                                                         f"RHS evaluations on the inflated cube with 4^3 elements "
                                                         f"per patch ({o.n} points), {tc:.3f} s each"}
    return out


def mfma_ab_run(args):
    """The per-axis contractions of the drop-in RHS at N = 8, FP64, on VALU
    (the even/odd split the kernels use) and on MFMA (v_mfma_f64_16x16x4f64
    with [G G; G] as one operand), 32^3 elements: the standalone program
    wt2_mfma_ab built next to libwt2.so, run as a child process."""
    import subprocess
    exe = os.path.join(ROOT, "wavetoysecondorder.jl_amd", "wavetoy_hip", "wt2_mfma_ab")
    if not os.path.exists(exe):
        raise SystemExit("bench.py: wt2_mfma_ab is not built (run __graft_entry__.build())")
    r = subprocess.run([exe, str(M_ELEMS)], capture_output=True, text=True, timeout=300)
    if r.returncode != 0:
        raise SystemExit(f"bench.py: wt2_mfma_ab failed ({r.returncode}): {r.stderr.strip()} {r.stdout.strip()}")
    res = json.loads(r.stdout.strip().splitlines()[-1])
    return {"metric": "per-axis G u and G G u of every element line, three axes (us per launch)",
            "config": {"workload": f"{M_ELEMS}^3 elements x 8^3 nodes, FP64, contractions of the drop-in RHS "
                                   "without its face terms, element staged in LDS, running sum in LDS"},
            **res}


def patch_run(wh, torch, args):
    """The multi-patch inflated cube (13 patches of 8^3 elements, N = 8:
    3,407,872 points; patches meet with orientation != 0) with a Sommerfeld
    outer sphere: RK4 steps of an outgoing pulse, FP64 and mixed, and the
    drop-in RHS (wt2_rhs3d) on its own.  Roofline of the curved operator on
    128 algorithmic bytes per point (FP64: u, u_t, 13 coefficients read, the
    output written) and 76 (mixed: the coefficients in FP32)."""
    ball = wh.make_inflated_cube_mesh(N_NODES, 0.1, 0.3, 1.0, 8, outer_tag=7)
    elem = wh.make_element(N_NODES)
    x = ball.coords.reshape(-1, 3)
    u0, v0 = wh.outgoing_pulse(x, 0.0, s0=0.5, sigma=0.1)
    K = max(1, min(args.steps, 200))
    res = {}
    for prec in ("fp64", "mixed"):
        ws = wh.Workspace(ball, elem, precision=prec, sommerfeld_R=1.0)
        stream = torch.cuda.Stream()
        ws.set_stream(stream.cuda_stream)
        dt = wh.recommended_dt_rk4(ws, 0.2)
        u = torch.from_numpy(u0).cuda(); v = torch.from_numpy(v0).cuda(); o = torch.empty_like(u)
        with torch.cuda.stream(stream):
            ws.rk4_steps(u, v, dt, max(args.warmup, 2))
            torch.cuda.synchronize()
            a = torch.cuda.Event(enable_timing=True); b = torch.cuda.Event(enable_timing=True)
            t0 = time.perf_counter()
            a.record(stream)
            ws.rk4_steps(u, v, dt, K)
            b.record(stream)
            torch.cuda.synchronize()
            wall = time.perf_counter() - t0
            step_ms = a.elapsed_time(b) / K
            for _ in range(3):
                ws.rhs(o, u, v, (0.0,) * 6)
            torch.cuda.synchronize()
            c = torch.cuda.Event(enable_timing=True); d = torch.cuda.Event(enable_timing=True)
            c.record(stream)
            for _ in range(K):
                ws.rhs(o, u, v, (0.0,) * 6)
            d.record(stream)
            torch.cuda.synchronize()
            rhs_ms = c.elapsed_time(d) / K
        assert bool(torch.isfinite(u).all())
        ws.set_stream(None)
        alg = 128 if prec == "fp64" else 76
        gbs = alg * ws.npoints / (rhs_ms * 1e-3) / 1e9
        res[prec] = {"ms_per_step": wall * 1e3 / K, "event_ms_per_step": step_ms,
                     "point_updates_per_s": ws.npoints * K / wall, "rhs_ms": rhs_ms,
                     "rhs_roofline": {"bound": "hbm", "achieved": gbs, "peak": PEAK_HBM_GBS, "unit": "GB/s",
                                      "frac": gbs / PEAK_HBM_GBS, "algorithmic_bytes_per_point": alg}}
        npts = ws.npoints
        ws.close()
    return {"metric": "grid-point updates per second (RK4 step, multi-patch mesh)",
            "value": res["fp64"]["point_updates_per_s"], "unit": "grid-point updates/s", "n_gpus": 1,
            "steps": K, "warmup": args.warmup, "ms_per_step": res["fp64"]["ms_per_step"],
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
            "data": "synthetic (outgoing pulse, s0 = 0.5, sigma = 0.1)",
            "config": {"workload": "inflated cube, 13 patches of 8^3 elements (6656), N = 8 (3,407,872 points), "
                                   "Sommerfeld sphere R = 1, tau = 1.5 (N-1)^2", "points": npts},
            "fp64": res["fp64"], "mixed": res["mixed"],
            "mixed_speedup": res["fp64"]["ms_per_step"] / res["mixed"]["ms_per_step"],
            "cpu_baseline": None,
            "cpu_baseline_note": "timed on the headline (plain) configuration only"}


def mixed_error_run(wh, torch, args):
    """Growth in time of the mixed-vs-FP64 relative H-norm error of u and u_t on
    N = 10 meshes of --elems^3 elements, the warped ADM-bump configuration and
    its flat counterpart (KO 0.1, Dirichlet, dt = 0.2 h_min, Gaussian width 0.1),
    sampled at t = 0.125 ... 2."""
    N, M = 10, args.elems
    out = {"metric": "mixed-vs-FP64 relative H-norm error against time", "config": {
        "workload": f"N = 10, {M}^3 elements, warped A=0.05 + ADM bump background and flat, KO 0.1, "
                    "Gaussian width 0.1 centred at (0.55, 0.41, 0.32), dt = 0.2 h_min", "elements": M}}
    for kind in ("curved", "flat"):
        mesh = (wh.make_warped_uniform_hex(M, M, M, 0.0, 1.0, 0.05) if kind == "curved"
                else wh.make_uniform_hex(M, 0.0, 1.0))
        elem = wh.make_element(N)
        x = wh.host_coords(mesh, elem)
        bg = (wh.sample_background3d(wh.adm_bump_background(), 0.0, x[:, 0], x[:, 1], x[:, 2])
              if kind == "curved" else None)
        u0, v0 = wh.gaussian_pulse(x, (0.55, 0.41, 0.32), width=0.1)
        tau = wh.tau_effective(TAU_LITERAL, N)
        ws = {p: wh.Workspace(mesh, elem, tau=tau, eps_ko=EPS_KO, precision=p, bg=bg) for p in ("fp64", "mixed")}
        dt = wh.recommended_dt_rk4(ws["fp64"], 0.2)
        st = {p: (torch.from_numpy(u0).cuda(), torch.from_numpy(v0).cuda()) for p in ws}
        o = wh.Workspace(mesh, elem, tau=tau, bg=bg)
        z = torch.zeros_like(st["fp64"][0])
        rows, t = [], 0.0
        for T in (0.125, 0.25, 0.5, 1.0, 2.0):
            k = int(round((T - t) / dt))
            for p in ws:
                ws[p].rk4_steps(st[p][0], st[p][1], dt, k)
            t += k * dt
            du = (st["mixed"][0] - st["fp64"][0]).contiguous()
            dv = (st["mixed"][1] - st["fp64"][1]).contiguous()
            rows.append({"t": t, "steps": int(round(t / dt)),
                         "u": math.sqrt(o.energy(z, du) / o.energy(z, st["fp64"][0])),
                         "u_t": math.sqrt(o.energy(z, dv) / o.energy(z, st["fp64"][1]))})
            print(f"{kind} t={t:.3f} u {rows[-1]['u']:.3e} u_t {rows[-1]['u_t']:.3e}", file=sys.stderr, flush=True)
        out[kind] = {"dt": dt, "samples": rows}
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=500)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-mixed", action="store_true")
    ap.add_argument("--no-extra", action="store_true",
                    help="plain config: skip the curved, batch and serving legs")
    ap.add_argument("--extra-requests", type=int, default=48, help="plain config: requests of the serving leg")
    ap.add_argument("--repeats", type=int, default=5,
                    help="plain config: further runs of --steps steps after the timed one (median, minimum)")
    ap.add_argument("--no-tune", action="store_true", help="default tile shapes instead of the autotuner's")
    ap.add_argument("--config", choices=["plain", "curved", "batch", "serving", "passes", "fo", "fo_curv",
                                         "mixed_error", "patch", "mfma_ab"],
                    default="plain",
                    help="plain: the FP64 flat workload (headline); curved: the warped-mesh "
                         "ADM workload, FP64 vs mixed, error at --t-final")
    ap.add_argument("--instances", type=int, default=512, help="batch config: number of instances")
    ap.add_argument("--elems", type=int, default=12, help="mixed_error config: elements per axis")
    ap.add_argument("--requests", type=int, default=64, help="serving config: requests per rate")
    ap.add_argument("--no-sweep", action="store_true",
                    help="passes config: only the default tiles, each pass launched 3 times (for counter runs)")
    ap.add_argument("--rates", default="auto",
                    help="serving config: arrival rates (1/s), comma-separated, or 'auto' (fractions of the "
                         "measured saturation rate)")
    ap.add_argument("--t-final", type=float, default=None,
                    help="curved config: evolve to this time (default: --steps steps)")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="run only a plain device copy of 1 GiB (x5) for HBM-counter calibration")
    args = ap.parse_args()
    if args.calibrate_copy:
        import torch
        src = torch.empty(1 << 27, dtype=torch.float64, device="cuda").fill_(1.0)
        dst = torch.empty_like(src)
        for _ in range(5):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"calibrate_copy_bytes_read": src.numel() * 8, "bytes_written": dst.numel() * 8}))
        return
    if args.config == "mfma_ab":
        print(json.dumps(mfma_ab_run(args)), flush=True)
        return
    if args.gpus != 1:
        print(f"bench.py: measures one GPU only (asked for {args.gpus}); instances or element slabs "
              f"would be sharded one per GPU, which this benchmark does not run.", file=sys.stderr)
        sys.exit(2)

    import torch
    import wavetoy_hip as wh

    if args.config == "curved":
        print(json.dumps(curved_run(wh, torch, args)), flush=True)
        return
    if args.config == "batch":
        print(json.dumps(batch_run(wh, torch, args)), flush=True)
        return
    if args.config == "serving":
        print(json.dumps(serving_run(wh, torch, args)), flush=True)
        return
    if args.config == "passes":
        print(json.dumps(passes_run(wh, torch, args)), flush=True)
        return
    if args.config == "fo":
        print(json.dumps(fo_run(wh, torch, args)), flush=True)
        return
    if args.config == "patch":
        print(json.dumps(patch_run(wh, torch, args)), flush=True)
        return
    if args.config == "fo_curv":
        print(json.dumps(fo_curv_run(wh, torch, args)), flush=True)
        return
    if args.config == "mixed_error":
        print(json.dumps(mixed_error_run(wh, torch, args)), flush=True)
        return

    mesh = wh.make_uniform_hex(M_ELEMS, 0.0, 1.0)
    elem = wh.make_element(N_NODES)
    u0, v0, x0 = initial_data(wh, mesh, elem, args.seed)
    stream = torch.cuda.Stream()

    r64 = run_gpu(wh, torch, "fp64", u0, v0, args.steps, args.warmup, stream, not args.no_tune, args.repeats)
    npts, K = r64["npts"], args.steps
    value = npts * K / r64["wall"]
    # dominant kernel: the pencil-pass kernel (every launch of a step: 8 with
    # the folded schedule); average launch duration over the timed region from
    # the stream events.  Algorithmic bytes: the fused 3-register RK4 minimum,
    # 256 B per point and step (u, u_t of Y, S, A: 16 field-pair accesses),
    # i.e. 256 / launches B per point per launch; the passes themselves move
    # step_bytes("fp64", launches) (448 B folded, 512 B unfolded).
    lps = r64["lps"]
    nlaunch = lps * K
    avg_launch_s = r64["ev_ms"] * 1e-3 / nlaunch
    bytes_per_launch = FUSED_MIN_BYTES["fp64"] * npts / lps
    achieved = bytes_per_launch / avg_launch_s / 1e9
    design_GBs = step_bytes("fp64", lps) * npts / lps / avg_launch_s / 1e9
    # the drop-in RHS kernel: algorithmic 24 B per point (read u, u_t, write
    # u_tt); the Dirichlet RHS reads no u_t, so 16 B are compulsory
    rhs_s = r64["rhs_ms"] * 1e-3
    traffic = None
    tfile = os.path.join(ROOT, "profiles", "traffic_flat_pass_fp64_32x8.json")
    if os.path.exists(tfile):
        tj = json.load(open(tfile))
        traffic = tj.get("hbm_bytes_per_launch_by_launches_per_step", {}).get(str(lps), tj.get("hbm_bytes_per_launch"))

    out = {
        "metric": "grid-point updates/s per RK4 step, mixed vs FP64 at <=1e-5 err; p99 latency",
        "value": value,
        "unit": "grid-point updates/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": args.warmup,
        "ms_per_step": r64["wall"] * 1e3 / K,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded Gaussian pulse)",
        "config": {"workload": "flat cube [0,1]^3, 32^3 elements x 8^3 GLL nodes (16,777,216 points), "
                               "Dirichlet, SIPG tau_eff=73.5 (literal 20 is not coercive at N=8), "
                               "KO eps=0.1 on u_t, classical RK4 dt=0.2*h_min",
                   "points": npts, "dt": r64["dt"], "h_min": r64["h_min"], "mu": r64["mu"],
                   "seed": args.seed, "pulse_centre": [float(c) for c in x0]},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": PEAK_HBM_GBS, "unit": "GB/s",
                     "frac": achieved / PEAK_HBM_GBS, "traffic": traffic,
                     "kernel": f"flat_lean_kernel + flat_xpass_kernel <8,double,...> (all {lps} pencil passes of a step)",
                     "launches_per_step": lps,
                     "algorithmic_bytes_per_point_step": FUSED_MIN_BYTES["fp64"],
                     "pass_design_bytes_per_point_step": step_bytes("fp64", lps),
                     "pass_design_GBs": design_GBs},
        "rhs_roofline": {"bound": "hbm", "kernel": "flat_lap3_kernel<8,double> (wt2_rhs3d)",
                         "ms": r64["rhs_ms"], "achieved": 24 * npts / rhs_s / 1e9, "peak": PEAK_HBM_GBS,
                         "unit": "GB/s", "frac": 24 * npts / rhs_s / 1e9 / PEAK_HBM_GBS,
                         "algorithmic_bytes_per_point": 24, "compulsory_bytes_per_point": 16,
                         "compulsory_GBs": 16 * npts / rhs_s / 1e9},
        "ms_per_step_repeats": repeat_summary(r64["rep_ms"]),
        "tiles_fp64": r64["ws"].tiles(),
        "pass_ms_fp64": r64["pass_ms"],
        "energy_final": r64["energy"],
    }
    if not args.no_mixed:
        rmx = run_gpu(wh, torch, "mixed", u0, v0, K, args.warmup, stream, not args.no_tune, args.repeats)
        vmx = npts * K / rmx["wall"]
        H = wh.Geometry(mesh, elem).Hphys
        d = rmx["u"] - r64["u"]
        herr = math.sqrt(float(np.sum(H * d * d)) / float(np.sum(H * r64["u"] ** 2)))
        maxerr = float(np.max(np.abs(d)) / np.max(np.abs(r64["u"])))
        out["mixed"] = {"value": vmx, "ms_per_step": rmx["wall"] * 1e3 / K,
                        "speedup_vs_fp64": vmx / value, "hnorm_rel_err_vs_fp64": herr,
                        "max_rel_err_vs_fp64": maxerr, "dtype": "f32 RHS / f64 state",
                        "pass_ms": rmx["pass_ms"], "ms_per_step_repeats": repeat_summary(rmx["rep_ms"]),
                        "tiles": rmx["ws"].tiles(),
                        "roofline_achieved_GBs": FUSED_MIN_BYTES["mixed"] * npts * K / (rmx["ev_ms"] * 1e-3) / 1e9,
                        # the drop-in RHS of the mixed context (FP32 arithmetic, FP64 in and out)
                        "rhs_ms": rmx["rhs_ms"],
                        "rhs_frac_24B": 24 * npts / (rmx["rhs_ms"] * 1e-3) / 1e9 / PEAK_HBM_GBS}
        del rmx
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(u0, v0, r64["dt"])
    del r64
    torch.cuda.empty_cache()
    if not args.no_extra:
        # the other workloads of the metric, each in its own key: the curved
        # ADM step to t = 2 (FP64 and mixed, error at t = 2), the 512-instance
        # batch (FP64 and mixed, per-instance error), the serving loop
        # (saturation and p50 / p99 at rates below and above it)
        ns = argparse.Namespace(**vars(args))
        ns.t_final = 2.0
        t0 = time.perf_counter()
        print("extra: curved 24^3 x 10^3 to t = 2", file=sys.stderr, flush=True)
        out["curved_t2"] = curved_run(wh, torch, ns)
        out["curved_t2"]["leg_wall_s"] = time.perf_counter() - t0
        ns.t_final = 1.0
        ns.instances = 512
        t0 = time.perf_counter()
        out["batch_512"] = batch_run(wh, torch, ns)
        out["batch_512"]["leg_wall_s"] = time.perf_counter() - t0
        ns.requests = args.extra_requests
        ns.rates = "auto"
        t0 = time.perf_counter()
        out["serving"] = serving_run(wh, torch, ns)
        out["serving"]["leg_wall_s"] = time.perf_counter() - t0
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
