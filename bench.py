"""Benchmark: OSQP ADMM solves on one MI355X.

Default workload: the large random QP (n=50,000, m=500,000, A ~5e7 nonzeros,
P = MM' + 1e-2 I with M of density 1e-4), eps_abs = eps_rel = 1e-4, solved by
the FP32-PCG engine (FP32 CG with FP64 defect correction; FP64 residuals and
termination).  A "step" is one full cold-started osqp_solve of that QP with its
data already resident in HBM (setup excluded).  The plain FP64 HIP backend is
run on the same QP as the GPU comparison, and the CPU oracle's FP64 PCG path is
timed on a bounded prefix of the same solve as the CPU baseline.

  python bench.py [--steps K] [--warmup W] [--gpus 1] [--small]

Prints one JSON line on stdout; progress goes to stderr.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "QPs solved/s + p50/p99 latency at eps 1e-4; ADMM-iter HBM GB/s vs peak"
HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def percentile(xs, q):
    return float(np.percentile(np.asarray(xs), q)) if len(xs) else None


def measured_traffic(kernel):
    """HBM bytes per launch of the dominant kernel from the committed
    rocprofv3 FETCH_SIZE / WRITE_SIZE summary (calibrated), or None."""
    path = os.path.join(ROOT, "profiles", "panel_spmv_large_qp_traffic.md")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        for line in f:
            if line.startswith("traffic_bytes_per_launch:"):
                return float(line.split(":")[1])
    return None


def _mpc_stream(eng, reqs, arrivals, label):
    """Feeds requests to the engine at their arrival times (seconds from the
    start) and steps it until every result is seen; returns the elapsed time,
    per-request latencies (result seen by the host - arrival), statuses,
    objective values and ADMM iteration counts."""
    n_req = len(reqs)
    t_start = time.perf_counter()
    arr_of, lat, status, iters, objs = {}, [], {}, [], {}
    nxt = 0
    pending = set()
    t_log = time.perf_counter()
    while nxt < n_req or pending:
        if time.perf_counter() - t_log > 5.0:
            log(f"[serve] mpc {label}: {nxt} admitted, {len(lat)} done")
            t_log = time.perf_counter()
        now = time.perf_counter() - t_start
        # take at most 256 arrivals between two steps, so the engine works on
        # admitted requests while the host takes in the next ones
        k1 = nxt
        while k1 < n_req and arrivals[k1] <= now and len(pending) + (k1 - nxt) < 4096 and k1 - nxt < 256:
            k1 += 1
        if k1 > nxt:
            # the arrivals since the last step, queued in one call
            for k, pid in zip(range(nxt, k1), eng.add_many(reqs[nxt:k1])):
                arr_of[pid] = (k, arrivals[k])
                pending.add(pid)
            nxt = k1
        done = eng.step()
        t_seen = time.perf_counter() - t_start
        for pid in done:
            r = eng.poll(pid)
            k, a = arr_of[pid]
            lat.append(t_seen - a)
            status[k] = r.info.status_val
            objs[k] = r.info.obj_val
            iters.append(r.info.iter)
            pending.discard(pid)
    return time.perf_counter() - t_start, lat, status, objs, iters


def serve_mpc(n_req, seed):
    """MPC stream (nx=12, nu=4, N ~ U{10..50}) through the batched FP32-PCG
    engine.  Throughput: Poisson arrivals at a rate far above saturation (up to
    4096 in flight), so the engine is the bottleneck.  Latency under load: the
    same requests again, Poisson arrivals at 70 % of that throughput.  Latency
    = result seen by the host - arrival.  The plain FP64 engine (osqp_setup +
    osqp_solve per request, one at a time) on a sample of the same requests is
    the GPU comparison; statuses are compared on it."""
    import torch
    from osqp_amd import OSQP
    from osqp_amd import problems as pr
    from osqp_amd.serve import BatchEngine
    st = dict(eps_abs=1e-4, eps_rel=1e-4)
    rng = np.random.default_rng(seed + 17)
    Ns = rng.integers(10, 51, n_req)
    t0 = time.time()
    reqs = [pr.mpc(int(N), seed=seed * 100000 + k) for k, N in enumerate(Ns)]
    log(f"[serve] generated {n_req} MPC requests in {time.time() - t0:.1f}s")
    eng = BatchEngine(precision=32, max_active=4096, **st)
    eng.solve_batch(reqs[:512])  # warm-up: code paths and the team-size tuner
    rate = 1e6  # requests/s: far above saturation
    arrivals = np.cumsum(rng.exponential(1.0 / rate, n_req))
    torch.cuda.synchronize()
    c0 = eng.counters()
    elapsed, lat, status, objs, iters = _mpc_stream(eng, reqs, arrivals, "saturated")
    c1 = eng.counters()
    qps = n_req / elapsed
    rate70 = 0.7 * qps
    arrivals70 = np.cumsum(rng.exponential(1.0 / rate70, n_req))
    el70, lat70, status70, _, _ = _mpc_stream(eng, reqs, arrivals70, "70% load")
    tuning = [eng.tuning(b) for b in range(6)]
    eng.close()
    same_runs = sum(int(status70[k] == status[k]) for k in range(n_req))
    # plain FP64 engine, one request at a time (setup + solve), on a sample
    k_plain = 24
    same = 0
    plain = []
    ts = time.perf_counter()
    for k in range(k_plain):
        p = reqs[k]
        s = OSQP(precision=64)
        assert s.setup(p["P"], p["q"], p["A"], p["l"], p["u"], **st) == 0
        r = s.solve()
        same += int(r.info.status_val == status[k])
        plain.append(r)
        s.close()
    plain_qps = k_plain / (time.perf_counter() - ts)
    obj_err = max(abs(objs[k] - plain[k].info.obj_val) / max(1.0, abs(plain[k].info.obj_val)) for k in range(k_plain))
    pct = lambda v, q: float(np.percentile(v, q) * 1e3)
    log(f"[serve] batched FP32: {qps:.0f} QPs/s, p50 {pct(lat, 50):.1f} ms, p99 {pct(lat, 99):.1f} ms; "
        f"at {rate70:.0f}/s: p50 {pct(lat70, 50):.1f} ms, p99 {pct(lat70, 99):.1f} ms; "
        f"plain FP64: {plain_qps:.1f} QPs/s; statuses identical on {same}/{k_plain}")
    return {"workload": f"MPC stream nx=12 nu=4 N~U{{10..50}}, {n_req} requests, Poisson arrivals at "
                        f"{rate:.0e}/s, <=4096 in flight, eps 1e-4",
            "qps": qps, "latency_ms": {"p50": pct(lat, 50), "p99": pct(lat, 99)},
            "at_70pct_load": {"arrival_rate_per_s": rate70, "qps": n_req / el70,
                              "latency_ms": {"p50": pct(lat70, 50), "p99": pct(lat70, 99)},
                              "statuses_same_as_saturated_run": same_runs},
            "mean_admm_iters": float(np.mean(iters)),
            "statuses_solved": int(sum(1 for v in status.values() if v == 1)),
            "plain_fp64_qps": plain_qps, "ratio_vs_plain_fp64": qps / plain_qps,
            "plain_sample": k_plain, "statuses_identical_on_sample": same,
            "max_obj_rel_err_vs_fp64_on_sample": obj_err,
            "engine": {"epochs": c1["epochs"] - c0["epochs"], "gpu_s": c1["gpu_s"] - c0["gpu_s"],
                       "tuned_team_per_bucket": tuning}}


def serve_portfolio(n_problems, n_req, seed):
    """Portfolio stream: n ~ U{500..5000} assets, problems resident on the GPU;
    each request draws a new mu, replaces q (osqp_update_data_vec semantics)
    and re-solves warm-started.  Closed loop: an idle problem takes its next
    request as soon as its previous result is seen.  The plain FP64 engine,
    one resident solver per problem doing update(q) + solve, on a sample of
    the same requests, is the GPU comparison."""
    from osqp_amd import OSQP
    from osqp_amd import problems as pr
    from osqp_amd.serve import BatchEngine
    st = dict(eps_abs=1e-4, eps_rel=1e-4)
    rng = np.random.default_rng(seed + 29)
    sizes = rng.integers(500, 5001, n_problems)
    t0 = time.time()
    probs = [pr.portfolio(int(nn), seed=seed * 1000 + k) for k, nn in enumerate(sizes)]
    reqs = [(int(rng.integers(n_problems)), None) for _ in range(n_req)]
    reqs = [(k, pr.portfolio_q(rng.standard_normal(probs[k]["P"].shape[0] - probs[k]["k"]), probs[k]["k"]))
            for k, _ in reqs]
    log(f"[serve] generated {n_problems} portfolios, {n_req} requests in {time.time() - t0:.1f}s")
    eng = BatchEngine(precision=32, max_active=4096, **st)
    pid = [eng.add(p["P"], p["q"], p["A"], p["l"], p["u"], resident=True) for p in probs]
    t0 = time.perf_counter()
    ep = 0
    while eng.pending():
        eng.step()
        ep += 1
        if ep % 20 == 0:
            log(f"[serve] portfolio initial solves: {n_problems - eng.pending()}/{n_problems} done, "
                f"{time.perf_counter() - t0:.1f}s")
    for i in pid:
        eng.poll(i)
    log(f"[serve] portfolio initial solves done in {time.perf_counter() - t0:.1f}s")
    queue = {k: [] for k in range(n_problems)}
    for j, (k, q) in enumerate(reqs):
        queue[k].append(j)
    t_start = time.perf_counter()
    issued, lat, status, busy, objs = {}, [], {}, {}, {}
    for k in range(n_problems):
        if queue[k]:
            j = queue[k].pop(0)
            eng.resolve(pid[k], reqs[j][1])
            issued[pid[k]] = (j, time.perf_counter())
    t_log = time.perf_counter()
    while issued:
        if time.perf_counter() - t_log > 5.0:
            log(f"[serve] portfolio: {len(lat)}/{n_req} requests done")
            t_log = time.perf_counter()
        if time.perf_counter() - t_start > 120.0:
            # bound the section's run time: stop issuing, drain what is in flight
            for k in queue:
                queue[k] = []
        for i in eng.step():
            r = eng.poll(i)
            j, t_iss = issued.pop(i)
            lat.append(time.perf_counter() - t_iss)
            status[j] = r.info.status_val
            objs[j] = r.info.obj_val
            k = pid.index(i)
            if queue[k]:
                jn = queue[k].pop(0)
                eng.resolve(i, reqs[jn][1])
                issued[i] = (jn, time.perf_counter())
    elapsed = time.perf_counter() - t_start
    eng.close()
    n_done = len(lat)
    qps = n_done / elapsed
    # plain FP64: one resident solver per problem (setup outside the timing)
    sample = [j for j in range(len(reqs)) if j in status][:24]
    k_plain = len(sample)
    solvers = {}
    for j in sample:
        k = reqs[j][0]
        if k not in solvers:
            s = OSQP(precision=64)
            p = probs[k]
            assert s.setup(p["P"], p["q"], p["A"], p["l"], p["u"], **st) == 0
            s.solve()
            solvers[k] = s
    same = 0
    obj_err = 0.0
    ts = time.perf_counter()
    for j in sample:
        k, q = reqs[j]
        solvers[k].update(q=q)
        r = solvers[k].solve()
        same += int(r.info.status_val == status[j])
        obj_err = max(obj_err, abs(objs[j] - r.info.obj_val) / max(1.0, abs(r.info.obj_val)))
    plain_qps = k_plain / (time.perf_counter() - ts)
    for s in solvers.values():
        s.close()
    log(f"[serve] portfolio batched FP32: {qps:.0f} QPs/s, p50 {np.percentile(lat, 50) * 1e3:.1f} ms, "
        f"p99 {np.percentile(lat, 99) * 1e3:.1f} ms; plain FP64: {plain_qps:.1f} QPs/s")
    return {"workload": f"portfolio stream, {n_problems} resident problems n~U{{500..5000}}, {n_done} mu updates "
                        f"with warm-started re-solves, eps 1e-4, closed loop",
            "qps": qps, "latency_ms": {"p50": float(np.percentile(lat, 50) * 1e3),
                                       "p99": float(np.percentile(lat, 99) * 1e3)},
            "statuses_solved": int(sum(1 for v in status.values() if v == 1)),
            "plain_fp64_qps": plain_qps, "ratio_vs_plain_fp64": qps / plain_qps,
            "plain_sample": k_plain, "statuses_identical_on_sample": same,
            "max_obj_rel_err_vs_fp64_on_sample": obj_err}


def random_qp_small_section(seed):
    """The random-QP configuration that runs on the CPU today (n = 500,
    m = 5000, density 0.15, P = M M' + 1e-2 I, default settings, eps 1e-4):
    GPU solves with both engines, and the CPU oracle's direct (LDL') path --
    the reference's default -- timed on one core; errors against the oracle."""
    from oracle import OracleSolver
    from osqp_amd import OSQP
    from osqp_amd import problems as pr
    p = pr.random_qp(500, 5000, 0.15, 0.15, seed=seed)
    st = dict(eps_abs=1e-4, eps_rel=1e-4, verbose=0)
    o = OracleSolver()
    ts = time.perf_counter()
    assert o.setup(p["P"], p["q"], p["A"], p["l"], p["u"], linsys_solver=1, **st) == 0
    o.solve()
    t_cpu = time.perf_counter() - ts
    xo, yo, fo = np.array(o.x), np.array(o.y), o.info.obj_val
    out = {"workload": f"random QP n=500 m=5000 density 0.15 (nnz(A)={p['A'].nnz}), eps 1e-4, setup + cold solve",
           "cpu_oracle_direct": {"qps": 1.0 / t_cpu, "cores": 1, "status": o.info.status.decode()
                                 if isinstance(o.info.status, bytes) else str(o.info.status), "iter": o.info.iter}}
    for prec in (32, 64):
        s = OSQP(precision=prec)
        ts = time.perf_counter()
        assert s.setup(p["P"], p["q"], p["A"], p["l"], p["u"], **st) == 0
        r = s.solve()
        el = time.perf_counter() - ts
        s.close()
        out["gpu_fp32_pcg" if prec == 32 else "gpu_fp64_plain"] = {
            "qps": 1.0 / el, "status": r.info.status, "iter": r.info.iter,
            "err_vs_oracle": {"x_inf": float(np.max(np.abs(r.x - xo))), "y_inf": float(np.max(np.abs(r.y - yo))),
                              "obj_rel": abs(r.info.obj_val - fo) / max(1.0, abs(fo))}}
    log(f"[random-qp-small] CPU oracle {t_cpu * 1e3:.0f} ms; GPU FP32 {1e3 / out['gpu_fp32_pcg']['qps']:.0f} ms, "
        f"obj err {out['gpu_fp32_pcg']['err_vs_oracle']['obj_rel']:.1e}")
    return out


def lasso_section(seed):
    """The lasso configuration at full size (n = 20000, m_d = 200000, D of
    density 1e-3): one cold solve with the FP32-PCG engine and one with the
    plain FP64 engine, the error of the first against the second, and the FP64
    KKT certificate of the FP32 point."""
    from osqp_amd import OSQP
    from osqp_amd import problems as pr
    t0 = time.time()
    p = pr.lasso(20000, 200000, 1e-3, seed=seed)
    log(f"[lasso] generated n={p['P'].shape[0]} m={p['A'].shape[0]} nnz(A)={p['A'].nnz} in {time.time() - t0:.1f}s")
    st = dict(eps_abs=1e-4, eps_rel=1e-4, verbose=0)
    out = {}
    for prec in (32, 64):
        s = OSQP(precision=prec)
        assert s.setup(p["P"], p["q"], p["A"], p["l"], p["u"], **st) == 0
        ts = time.perf_counter()
        r = s.solve()
        out[prec] = (r, time.perf_counter() - ts)
        s.close()
    (r32, t32), (r64, t64) = out[32], out[64]
    c = pr.kkt_certificate(p, r32.x, r32.y)
    log(f"[lasso] FP32-PCG {t32 * 1e3:.0f} ms ({r32.info.status}, {r32.info.iter} it); FP64 {t64 * 1e3:.0f} ms")
    return {"workload": f"lasso n=20000 m_d=200000 density 1e-3 (QP n={p['P'].shape[0]} m={p['A'].shape[0]} "
                        f"nnz(A)={p['A'].nnz}), eps 1e-4, cold solve",
            "fp32_pcg": {"qps": 1.0 / t32, "status": r32.info.status, "iter": r32.info.iter},
            "fp64_plain": {"qps": 1.0 / t64, "status": r64.info.status, "iter": r64.info.iter},
            "err_vs_fp64": {"x_inf": float(np.max(np.abs(r32.x - r64.x))), "y_inf": float(np.max(np.abs(r32.y - r64.y))),
                            "obj_rel": abs(r32.info.obj_val - r64.info.obj_val) / max(1.0, abs(r64.info.obj_val))},
            "kkt_certificate": {"prim": c["prim"], "dual": c["dual"]}}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--small", action="store_true", help="n=5000, m=50000 instance (quick check)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--no-serving", action="store_true")
    ap.add_argument("--mpc-requests", type=int, default=4096)
    ap.add_argument("--profile-kernels", action="store_true",
                    help="set up the QP, launch only the timed SpMV kernels (reps=20 each, A then K) and a "
                         "256 MiB device copy (counter calibration); for rocprofv3 runs")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py runs on one GPU only: this build has no multi-GPU path (one MI355X per job); "
              "use --gpus 1", file=sys.stderr)
        sys.exit(2)

    import torch
    from osqp_amd import OSQP, device_name
    from osqp_amd import problems as pr

    n, m = (5000, 50000) if args.small else (50000, 500000)
    dens_M, dens_A = (1e-3, 2e-2) if args.small else (1e-4, 2e-3)
    t0 = time.time()
    prob = pr.random_qp(n, m, dens_M, dens_A, seed=args.seed)
    log(f"[bench] generated n={n} m={m} nnz(A)={prob['A'].nnz} nnz(P triu)={prob['P'].nnz} in {time.time() - t0:.1f}s")
    settings = dict(eps_abs=1e-4, eps_rel=1e-4, verbose=0)

    # ---- FP32-PCG engine (the measured configuration) -------------------------
    t0 = time.time()
    eng = OSQP(precision=32)
    rc = eng.setup(prob["P"], prob["q"], prob["A"], prob["l"], prob["u"], **settings)
    assert rc == 0, f"setup failed: {rc}"
    torch.cuda.synchronize()
    log(f"[bench] FP32-PCG setup {time.time() - t0:.2f}s on {device_name()}")
    if args.profile_kernels:
        kt = eng.time_kernels(reps=20)
        kt64 = eng.time_fp64_kernels(reps=20)
        a = torch.ones(64 << 20, dtype=torch.float32, device="cuda")
        b = torch.empty_like(a)
        for _ in range(5):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"profile_kernels": kt, "fp64_kernels": kt64, "copy_bytes_each_way": a.numel() * 4}),
              flush=True)
        eng.close()
        return
    for w in range(args.warmup):
        eng.cold_start()
        r = eng.solve()
        log(f"[bench] warmup {w}: status={r.info.status} iter={r.info.iter} t={r.info.solve_time:.3f}s")
    lat, iters, statuses = [], [], []
    st0 = eng.stats()
    torch.cuda.synchronize()
    t_start = time.perf_counter()
    for k in range(args.steps):
        ts = time.perf_counter()
        eng.cold_start()
        r = eng.solve()
        lat.append(time.perf_counter() - ts)
        iters.append(r.info.iter)
        statuses.append(r.info.status)
        log(f"[bench] step {k}: status={r.info.status} iter={r.info.iter} obj={r.info.obj_val:.8e} "
            f"lat={lat[-1] * 1e3:.1f}ms")
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t_start
    st1 = eng.stats()
    x32, y32, obj32 = r.x, r.y, r.info.obj_val
    cert32 = pr.kkt_certificate(prob, x32, y32)
    qps = args.steps / elapsed
    admm_iters = int(np.sum(iters))
    cg_iters = st1.pcg_iters - st0.pcg_iters
    ms_per_admm = elapsed * 1e3 / max(admm_iters, 1)

    # ---- whole ADMM iteration: algorithmic bytes / measured time ----------------
    # PCG bytes from the library (same formulas as the kernels' accounting);
    # host-core vector passes per ADMM iteration: compute_rhs 3n + 5m,
    # update_x 6n, update_z 11m, update_y 10m element touches of 8 bytes
    pb = eng.pcg_bytes()
    cg_per_admm = cg_iters / max(admm_iters, 1)
    admm_bytes = pb["per_solve"] + cg_per_admm * pb["per_cg_iter"] + 8.0 * (9 * n + 26 * m)
    admm_gbs = admm_bytes / (ms_per_admm * 1e-3) / 1e9

    # ---- dominant kernels: CG operator, timed with HIP events ------------------
    kt = eng.time_kernels(reps=20)
    dom = "k_Kv" if kt["ms_kv"] >= kt["ms_av"] else "k_Av_rho"
    dom_ms = max(kt["ms_kv"], kt["ms_av"])
    dom_bytes = kt["bytes_kv"] if dom == "k_Kv" else kt["bytes_av"]
    achieved = dom_bytes / (dom_ms * 1e-3) / 1e9
    log(f"[bench] kernels: Av {kt['ms_av']:.4f}ms {kt['bytes_av'] / kt['ms_av'] / 1e6:.0f} GB/s; "
        f"Kv {kt['ms_kv']:.4f}ms {kt['bytes_kv'] / kt['ms_kv'] / 1e6:.0f} GB/s")
    # the defect correction's FP64 panel SpMVs (once per ADMM iteration each)
    kt64 = eng.time_fp64_kernels(reps=20)
    fp64_kernels = None
    if kt64:
        fp64_kernels = {"A_x64": {"ms_per_launch": kt64["ms_a"], "bytes_per_launch": kt64["bytes_a"],
                                  "GBs": kt64["bytes_a"] / kt64["ms_a"] / 1e6},
                        "K_x64": {"ms_per_launch": kt64["ms_k"], "bytes_per_launch": kt64["bytes_k"],
                                  "GBs": kt64["bytes_k"] / kt64["ms_k"] / 1e6}}
        log(f"[bench] FP64 residual kernels: A {kt64['ms_a']:.4f}ms {fp64_kernels['A_x64']['GBs']:.0f} GB/s; "
            f"K {kt64['ms_k']:.4f}ms {fp64_kernels['K_x64']['GBs']:.0f} GB/s")
    eng.close()

    # ---- plain FP64 backend on the same QP (GPU comparison) -------------------
    fp64 = None
    if not args.no_fp64:
        e64 = OSQP(precision=64)
        assert e64.setup(prob["P"], prob["q"], prob["A"], prob["l"], prob["u"], **settings) == 0
        ts = time.perf_counter()
        r64 = e64.solve()
        t64 = time.perf_counter() - ts
        fp64 = dict(qps=1.0 / t64, status=r64.info.status, iter=r64.info.iter, obj=r64.info.obj_val,
                    err_x_inf=float(np.max(np.abs(x32 - r64.x))),
                    err_y_inf=float(np.max(np.abs(y32 - r64.y))),
                    err_obj_rel=float(abs(obj32 - r64.info.obj_val) / max(1.0, abs(r64.info.obj_val))))
        log(f"[bench] FP64 plain: status={r64.info.status} iter={r64.info.iter} t={t64:.3f}s "
            f"|dx|={fp64['err_x_inf']:.2e} |dobj|={fp64['err_obj_rel']:.2e}")
        e64.close()

    # ---- CPU baseline: oracle FP64 PCG, bounded prefix of the same solve -------
    cpu = None
    if not args.no_cpu:
        from oracle import OracleSolver
        k_cpu = 3
        o = OracleSolver()
        t0 = time.time()
        assert o.setup(prob["P"], prob["q"], prob["A"], prob["l"], prob["u"], linsys_solver=2,
                       max_iter=k_cpu, check_termination=0, **settings) == 0
        log(f"[bench] oracle setup {time.time() - t0:.1f}s")
        ts = time.perf_counter()
        o.solve()
        tc = time.perf_counter() - ts
        it_per_s = k_cpu / tc
        mean_iters = admm_iters / args.steps
        cpu = dict(value=it_per_s / mean_iters, unit="QPs/s", cores=1, kind="port",
                   sample=f"oracle FP64 PCG (oracle/osqp_oracle.c), first {k_cpu} ADMM iterations of the same "
                          f"n={n} m={m} QP ({tc:.1f}s, single thread); QPs/s = ADMM iter/s / "
                          f"{mean_iters:.0f} iterations per GPU solve")
        log(f"[bench] cpu oracle: {k_cpu} ADMM iterations in {tc:.1f}s")

    # ---- served MPC stream: batched FP32-PCG engine vs plain FP64 engine -------
    small = None if args.no_serving else random_qp_small_section(args.seed)
    lasso = None if args.no_serving else lasso_section(args.seed)
    serving = None if args.no_serving else serve_mpc(args.mpc_requests, args.seed)
    serving_pf = None if args.no_serving else serve_portfolio(64, 512, args.seed)

    out = {
        "metric": METRIC,
        "value": qps,
        "unit": "QPs/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": elapsed * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32 CG values / fp64 residuals",
        "data": "synthetic (seeded numpy generator of the configuration)",
        "config": {"workload": f"large random QP n={n} m={m} nnz(A)={prob['A'].nnz} P=MM'+1e-2I, eps 1e-4, FP32-PCG",
                   "n": n, "m": m, "nnz_A": int(prob["A"].nnz), "nnz_P_triu": int(prob["P"].nnz),
                   "eps_abs": 1e-4, "eps_rel": 1e-4, "seed": args.seed},
        "latency_ms": {"p50": percentile(lat, 50) * 1e3, "p99": percentile(lat, 99) * 1e3},
        "status": statuses,
        "admm_iters_per_solve": admm_iters / args.steps,
        "cg_iters_per_admm_iter": cg_iters / max(admm_iters, 1),
        "ms_per_admm_iter": ms_per_admm,
        "kkt_certificate": {"prim": cert32["prim"], "dual": cert32["dual"], "obj": cert32["obj"]},
        "admm_iter": {"bytes": admm_bytes, "ms": ms_per_admm, "achieved_GBs": admm_gbs,
                      "frac_of_peak": admm_gbs / HBM_PEAK_GBS, "bytes_per_cg_iter": pb["per_cg_iter"],
                      "bytes_per_linsys_outer": pb["per_solve"]},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": measured_traffic(dom), "kernel": dom,
                     "ms_per_launch": dom_ms, "bytes_per_launch": dom_bytes,
                     "other_kernel": {"name": "k_Av_rho" if dom == "k_Kv" else "k_Kv",
                                      "ms_per_launch": min(kt["ms_kv"], kt["ms_av"]),
                                      "bytes_per_launch": kt["bytes_av"] if dom == "k_Kv" else kt["bytes_kv"]},
                     "fp64_residual_kernels": fp64_kernels},
        "gpu_fp64_plain": fp64,
        "cpu_baseline": cpu,
        "random_qp_small": small,
        "lasso": lasso,
        "serving_mpc": serving,
        "serving_portfolio": serving_pf,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
