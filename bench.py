"""Benchmark: OSQP ADMM solves on one MI355X.

Default workload: the large random QP (n=50,000, m=500,000, A ~5e7 nonzeros,
P = MM' + 1e-2 I with M of density 1e-4), eps_abs = eps_rel = 1e-4, solved by
the FP32-PCG engine (FP32 CG with FP64 defect correction; FP64 residuals and
termination).  A "step" is one full cold-started osqp_solve of that QP with its
data already resident in HBM (setup excluded).  The plain FP64 HIP backend is
run on the same QP as the GPU comparison, and the CPU oracle's FP64 PCG path is
timed over one termination-check interval (25 ADMM iterations) of the same
solve as the CPU baseline.  Beside it: the random-QP, lasso, MPC-stream and
portfolio-stream configurations, each with its error against the FP64 oracle
and, for the streams, the oracle's direct path on a pool of the host's cores
(CPU model and core count reported).

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


def host_cpu():
    """CPU model and core counts of the host this runs on."""
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    aff = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
    return {"model": model, "nproc": os.cpu_count(), "affinity": aff}


def cpu_workers():
    """Worker processes for the CPU pools: the cores this process may use,
    at most 16 (a one-GPU job's share of the host)."""
    return max(1, min(16, host_cpu()["affinity"] or 1))


CHECK_INTERVAL = 25  # check_termination of the default settings (include/osqp_amd.h)


# ---- algorithmic bytes: the compulsory traffic of the stated algorithm -------
# (the counting convention of the measurement plan: each array element touched
# once per pass, int32 indices; value and PCG-vector bytes s_v = s_x = 4 in the
# FP32-PCG variant, ADMM state S = 8)
def b_spmv(nnz, rows, cols, s_v=4, s_i=4, s_x=4):
    """B_spmv(M) = nnz (s_v + s_i) + (rows + 1) s_i + cols s_x + rows s_x."""
    return nnz * (s_v + s_i) + (rows + 1) * s_i + cols * s_x + rows * s_x


def convention_bytes(n, m, nnz_a, nnz_pf, k_cg, checks_per_iter, s_v=4, s_x=4, S=8):
    """Per-unit figures of the convention for the large QP: one application
    of w = rho o (A p) and of K p = P p + sigma p + A' w, one CG iteration,
    one ADMM iteration (with k_cg CG iterations and checks_per_iter
    termination checks, each adding the FP64 products A x, A' y, P x and
    12 (n + m) state touches)."""
    av = b_spmv(nnz_a, m, n, s_v, 4, s_x) + m * s_x
    kv = b_spmv(nnz_pf, n, n, s_v, 4, s_x) + b_spmv(nnz_a, n, m, s_v, 4, s_x) + n * s_x
    cg = av + kv + 8 * n * s_x
    check = (b_spmv(nnz_a, m, n, 8, 4, 8) + b_spmv(nnz_a, n, m, 8, 4, 8) + b_spmv(nnz_pf, n, n, 8, 4, 8) +
             12 * (n + m) * 8)
    admm = (b_spmv(nnz_a, n, m, s_v, 4, s_x) + k_cg * cg + b_spmv(nnz_a, m, n, s_v, 4, s_x) + (5 * n + 9 * m) * S +
            (n + m) * (S + s_x) + checks_per_iter * check)
    return dict(av=av, kv=kv, cg=cg, check=check, admm=admm)


def lib_sync():
    """Waits for the library's queued GPU work (its own stream)."""
    from osqp_amd import load_library
    load_library().osqp_amd_algebra_sync()


def mpc_requests(seed, count):
    """Horizons and problem seeds of the MPC stream: request k has horizon
    Ns[k] ~ U{10..50} and problem seed seed * 100000 + k."""
    rng = np.random.default_rng(seed + 17)
    return rng.integers(10, 51, count), [seed * 100000 + k for k in range(count)]


def portfolio_stream(seed, n_problems, n_req):
    """The portfolio stream's problems (n ~ U{500..5000}) and requests
    (problem index, new q) in order."""
    from osqp_amd import problems as pr
    rng = np.random.default_rng(seed + 29)
    sizes = rng.integers(500, 5001, n_problems)
    probs = [pr.portfolio(int(nn), seed=seed * 1000 + k) for k, nn in enumerate(sizes)]
    ks = [int(rng.integers(n_problems)) for _ in range(n_req)]
    reqs = [(k, pr.portfolio_q(rng.standard_normal(probs[k]["P"].shape[0] - probs[k]["k"]), probs[k]["k"]))
            for k in ks]
    return probs, reqs


# ---- CPU baseline legs: the oracle (oracle/osqp_oracle.c, FP64) ----------------
# These run in a child process started before this process touches the GPU;
# the pools fork from that child, which never initialises the GPU.

def _cpu_mpc_chunk(job):
    """Oracle direct (LDL') path, setup + solve per request, one core."""
    from oracle import OracleSolver
    from osqp_amd import problems as pr
    Ns, seeds = job
    out = []
    for N, sd in zip(Ns, seeds):
        p = pr.mpc(int(N), seed=int(sd))
        t0 = time.perf_counter()
        o = OracleSolver()
        o.setup(p["P"], p["q"], p["A"], p["l"], p["u"], linsys_solver=1, eps_abs=1e-4, eps_rel=1e-4, verbose=0)
        o.solve()
        out.append((int(o.info.status_val), float(o.info.obj_val), time.perf_counter() - t0))
    return out


# the FP64 optimum of a request (objective error reference for both the GPU
# engine and the oracle's eps-1e-4 answers): the oracle's direct path at
# eps_abs = eps_rel = 1e-8; None where it does not reach that accuracy
OPT_SETTINGS = dict(linsys_solver=1, eps_abs=1e-8, eps_rel=1e-8, max_iter=200000, verbose=0)


def _cpu_mpc_optimum(job):
    from oracle import OracleSolver
    from osqp_amd import problems as pr
    Ns, seeds = job
    out = []
    for N, sd in zip(Ns, seeds):
        p = pr.mpc(int(N), seed=int(sd))
        o = OracleSolver()
        o.setup(p["P"], p["q"], p["A"], p["l"], p["u"], **OPT_SETTINGS)
        o.solve()
        out.append(float(o.info.obj_val) if o.info.status_val == 1 else None)
    return out


def _cpu_portfolio_optimum(job):
    """A resident portfolio's requests in stream order at the optimum settings
    (each warm-started from the previous optimum)."""
    from oracle import OracleSolver
    p, reqs = job
    o = OracleSolver()
    o.setup(p["P"], p["q"], p["A"], p["l"], p["u"], **OPT_SETTINGS)
    o.solve()
    out = []
    for j, q in reqs:
        o.update_data_vec(q=q)
        o.solve()
        out.append((j, float(o.info.obj_val) if o.info.status_val == 1 else None))
    return out


def obj_errors(objs, ref, keys):
    """Relative objective errors |f - f*| / max(1, |f*|) over the keys where
    both are present: max, p99, p50."""
    e = [abs(objs[k] - ref[k]) / max(1.0, abs(ref[k])) for k in keys
         if objs.get(k) is not None and ref[k] is not None]
    return {"requests": len(e), "max": float(np.max(e)) if e else None, "p99": percentile(e, 99),
            "p50": percentile(e, 50)}


def _cpu_portfolio_problem(job):
    """One resident portfolio on the oracle's direct path: setup + first solve
    (untimed), then its requests in stream order, each update_data_vec(q) +
    warm-started solve (timed)."""
    from oracle import OracleSolver
    p, reqs = job
    o = OracleSolver()
    o.setup(p["P"], p["q"], p["A"], p["l"], p["u"], linsys_solver=1, eps_abs=1e-4, eps_rel=1e-4, verbose=0)
    o.solve()
    out = []
    for j, q in reqs:
        t0 = time.perf_counter()
        o.update_data_vec(q=q)
        o.solve()
        out.append((j, int(o.info.status_val), float(o.info.obj_val), time.perf_counter() - t0))
    return out


def cpu_legs(seed, mpc_warm, mpc_req, pf_problems, pf_req, large_qp):
    import multiprocessing as mp
    from oracle import OracleSolver
    from osqp_amd import problems as pr
    W = cpu_workers()
    out = {"host": host_cpu(), "workers": W}
    ctx = mp.get_context("fork")
    # MPC stream: every timed request, W worker processes
    Ns, seeds = mpc_requests(seed, mpc_warm + mpc_req)
    Ns, seeds = Ns[mpc_warm:], seeds[mpc_warm:]
    chunks = [(Ns[i::W], seeds[i::W]) for i in range(W)]
    t0 = time.perf_counter()
    with ctx.Pool(W) as pool:
        res = pool.map(_cpu_mpc_chunk, chunks)
    wall = time.perf_counter() - t0
    status, obj, svc = [None] * mpc_req, [None] * mpc_req, [None] * mpc_req
    for i, rr in enumerate(res):
        for j, (st, ob, dt) in enumerate(rr):
            k = i + j * W
            status[k], obj[k], svc[k] = st, ob, dt
    out["mpc"] = {"qps": mpc_req / wall, "wall_s": wall, "status": status, "obj": obj,
                  "solve_ms": {"p50": percentile(svc, 50) * 1e3, "p99": percentile(svc, 99) * 1e3},
                  "single_core_qps": mpc_req / float(np.sum(svc))}
    log(f"[cpu] MPC: {mpc_req} requests on {W} processes in {wall:.1f}s ({mpc_req / wall:.0f} QPs/s)")
    t0 = time.perf_counter()
    with ctx.Pool(W) as pool:
        res = pool.map(_cpu_mpc_optimum, chunks)
    opt = [None] * mpc_req
    for i, rr in enumerate(res):
        for j, ob in enumerate(rr):
            opt[i + j * W] = ob
    out["mpc"]["opt"] = opt
    log(f"[cpu] MPC optima (eps 1e-8) in {time.perf_counter() - t0:.1f}s")
    # portfolio stream: one problem's requests in order per job
    probs, reqs = portfolio_stream(seed, pf_problems, pf_req)
    jobs = [(probs[k], [(j, q) for j, (kk, q) in enumerate(reqs) if kk == k]) for k in range(pf_problems)]
    t0 = time.perf_counter()
    with ctx.Pool(W) as pool:
        res = pool.map(_cpu_portfolio_problem, jobs)
    wall = time.perf_counter() - t0
    pst, pob, psv = {}, {}, []
    for rr in res:
        for j, st, ob, dt in rr:
            pst[j], pob[j] = st, ob
            psv.append(dt)
    busy = float(np.sum(psv))
    out["portfolio"] = {"status": [pst[j] for j in range(pf_req)], "obj": [pob[j] for j in range(pf_req)],
                        "qps_packed": pf_req / (busy / W), "single_core_qps": pf_req / busy,
                        "solve_ms": {"p50": percentile(psv, 50) * 1e3, "p99": percentile(psv, 99) * 1e3},
                        "wall_s_with_setups": wall}
    log(f"[cpu] portfolio: {pf_req} re-solves, {busy:.1f}s of solves on {W} processes")
    t0 = time.perf_counter()
    with ctx.Pool(W) as pool:
        res = pool.map(_cpu_portfolio_optimum, jobs)
    popt = {}
    for rr in res:
        for j, ob in rr:
            popt[j] = ob
    out["portfolio"]["opt"] = [popt[j] for j in range(pf_req)]
    log(f"[cpu] portfolio optima (eps 1e-8) in {time.perf_counter() - t0:.1f}s")
    # large QP: the oracle's FP64 PCG path over one termination-check interval
    if large_qp:
        n, m = large_qp
        p = pr.random_qp(n, m, 1e-4, 2e-3, seed=seed) if n == 50000 else pr.random_qp(n, m, 1e-3, 2e-2, seed=seed)
        o = OracleSolver()
        t0 = time.perf_counter()
        assert o.setup(p["P"], p["q"], p["A"], p["l"], p["u"], linsys_solver=2, max_iter=25, check_termination=25,
                       eps_abs=1e-4, eps_rel=1e-4, verbose=0) == 0
        t_setup = time.perf_counter() - t0
        t0 = time.perf_counter()
        o.solve()
        t25 = time.perf_counter() - t0
        out["large_qp"] = {"admm_iters": int(o.info.iter), "cg_iters": int(o.total_pcg_iters()), "solve_s": t25,
                           "setup_s": t_setup}
        log(f"[cpu] large QP: {o.info.iter} ADMM iterations ({o.total_pcg_iters()} CG) in {t25:.1f}s")
    return out


def run_cpu_legs(args, large_qp):
    """Runs the CPU legs in a child process (this process has not touched the
    GPU yet) and returns their results."""
    import subprocess
    cmd = [sys.executable, os.path.abspath(__file__), "--cpu-legs", "--seed", str(args.seed),
           "--mpc-requests", str(args.mpc_requests), "--mpc-warmup", str(args.mpc_warmup),
           "--portfolio-problems", str(args.portfolio_problems), "--portfolio-requests",
           str(args.portfolio_requests)]
    if large_qp:
        cmd += ["--cpu-large-qp", f"{large_qp[0]}x{large_qp[1]}"]
    r = subprocess.run(cmd, stdout=subprocess.PIPE, text=True, timeout=900)
    if r.returncode != 0:
        raise RuntimeError(f"CPU legs failed ({r.returncode})")
    return json.loads(r.stdout.strip().splitlines()[-1])


def oracle_fixture_errors(name, prob, x, y, obj):
    """Errors of a GPU point against the oracle's FP64 PCG answer held in
    tests/golden/<name>_oracle.npz (written by tests/golden/make_full_size_fixtures.py),
    when the fixture is present and was made from the same data."""
    path = os.path.join(ROOT, "tests", "golden", f"{name}_oracle.npz")
    if not os.path.exists(path):
        return None
    z = np.load(path)
    chk = np.array([prob["P"].nnz, prob["A"].nnz, np.sum(prob["P"].data), np.sum(prob["A"].data),
                    np.sum(prob["q"]), np.sum(np.clip(prob["l"], -1e30, 1e30)),
                    np.sum(np.clip(prob["u"], -1e30, 1e30))])
    if not np.allclose(chk, z["checksums"], rtol=1e-12, atol=0):
        return {"fixture": "different data"}
    info = z["info"]
    fo = float(info[3])
    return {"oracle_status": int(info[0]), "oracle_iter": int(info[1]), "oracle_obj": fo,
            "x_inf": float(np.max(np.abs(x - z["x"]))), "y_inf": float(np.max(np.abs(y - z["y"]))),
            "obj_rel": abs(obj - fo) / max(1.0, abs(fo)), "oracle_solve_s_on_build_host": float(info[8])}


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


MPC_INTAKE = 256  # arrivals taken in between two engine steps


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
        # take at most MPC_INTAKE arrivals between two steps, so the engine
        # works on admitted requests while the host takes in the next ones
        k1 = nxt
        while k1 < n_req and arrivals[k1] <= now and len(pending) + (k1 - nxt) < 4096 and k1 - nxt < MPC_INTAKE:
            k1 += 1
        if k1 > nxt:
            # the arrivals since the last step, queued in one call
            for k, pid in zip(range(nxt, k1), eng.add_many(reqs[nxt:k1])):
                arr_of[pid] = (k, arrivals[k])
                pending.add(pid)
            nxt = k1
        done = eng.step()
        t_seen = time.perf_counter() - t_start
        # the results' status, iterations and objective in one call (x and y
        # are not needed here)
        for pid, (st, it, ob) in eng.poll_info_many(done).items():
            k, a = arr_of[pid]
            lat.append(t_seen - a)
            status[k] = st
            objs[k] = ob
            iters.append(it)
            pending.discard(pid)
    return time.perf_counter() - t_start, lat, status, objs, iters


def serve_mpc(n_req, n_warm, seed, cpu):
    """MPC stream (nx=12, nu=4, N ~ U{10..50}) through the batched FP32-PCG
    engine: n_warm distinct warm-up requests (code paths and the team-size
    tuner), then n_req timed ones.  Throughput: Poisson arrivals at a rate far
    above saturation (up to 4096 in flight), so the engine is the bottleneck.
    Latency under load: the same requests again, Poisson arrivals at 70 % of
    that throughput.  Latency = result seen by the host - arrival.  Every
    timed request's status is compared with the CPU oracle's direct path
    (the reference's default), and its objective likewise.  The plain FP64
    engine (osqp_setup + osqp_solve per request, one at a time) on a sample
    of the same requests is the GPU comparison."""
    from osqp_amd import OSQP
    from osqp_amd import problems as pr
    from osqp_amd.serve import BatchEngine
    st = dict(eps_abs=1e-4, eps_rel=1e-4)
    Ns, seeds = mpc_requests(seed, n_warm + n_req)
    t0 = time.time()
    allreqs = [pr.mpc(int(N), seed=sd) for N, sd in zip(Ns, seeds)]
    warm, reqs = allreqs[:n_warm], allreqs[n_warm:]
    log(f"[serve] generated {n_warm} + {n_req} MPC requests in {time.time() - t0:.1f}s")
    rng = np.random.default_rng(seed + 18)
    eng = BatchEngine(precision=32, max_active=4096, **st)
    eng.solve_batch(warm)
    rate = 1e6  # requests/s: far above saturation
    arrivals = np.cumsum(rng.exponential(1.0 / rate, n_req))
    lib_sync()
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
    same_plain = 0
    ts = time.perf_counter()
    for k in range(k_plain):
        p = reqs[k]
        s = OSQP(precision=64)
        assert s.setup(p["P"], p["q"], p["A"], p["l"], p["u"], **st) == 0
        r = s.solve()
        same_plain += int(r.info.status_val == status[k])
        s.close()
    plain_qps = k_plain / (time.perf_counter() - ts)
    pct = lambda v, q: float(np.percentile(v, q) * 1e3)
    vs_oracle = None
    if cpu is not None:
        cst, cob = cpu["mpc"]["status"], cpu["mpc"]["obj"]
        errs = [abs(objs[k] - cob[k]) / max(1.0, abs(cob[k])) for k in range(n_req) if cst[k] == 1 and status[k] == 1]
        vs_oracle = {"requests": n_req, "statuses_identical": int(sum(int(status[k] == cst[k]) for k in range(n_req))),
                     "obj_rel_err_vs_direct": {"max": float(np.max(errs)) if errs else None,
                                               "p99": percentile(errs, 99), "p50": percentile(errs, 50)},
                     "obj_rel_err_vs_fp64_optimum": {
                         "gpu_batched_fp32": obj_errors(objs, cpu["mpc"]["opt"], range(n_req)),
                         "oracle_direct_eps_1e-4": obj_errors(dict(enumerate(cob)), cpu["mpc"]["opt"], range(n_req))}}
    log(f"[serve] batched FP32: {qps:.0f} QPs/s, p50 {pct(lat, 50):.1f} ms, p99 {pct(lat, 99):.1f} ms; "
        f"at {rate70:.0f}/s: p50 {pct(lat70, 50):.1f} ms, p99 {pct(lat70, 99):.1f} ms; "
        f"plain FP64: {plain_qps:.1f} QPs/s; vs oracle: {vs_oracle and vs_oracle['statuses_identical']}/{n_req}")
    out = {"workload": f"MPC stream nx=12 nu=4 N~U{{10..50}}, {n_req} requests after {n_warm} distinct warm-up "
                       f"requests, Poisson arrivals at {rate:.0e}/s, <=4096 in flight, eps 1e-4",
           "qps": qps, "latency_ms": {"p50": pct(lat, 50), "p99": pct(lat, 99)},
           "at_70pct_load": {"arrival_rate_per_s": rate70, "qps": n_req / el70,
                             "latency_ms": {"p50": pct(lat70, 50), "p99": pct(lat70, 99)},
                             "statuses_same_as_saturated_run": same_runs},
           "mean_admm_iters": float(np.mean(iters)),
           "statuses_solved": int(sum(1 for v in status.values() if v == 1)),
           "vs_cpu_oracle_direct": vs_oracle,
           "plain_fp64_qps": plain_qps, "ratio_vs_plain_fp64": qps / plain_qps,
           "plain_sample": k_plain, "statuses_identical_to_plain_on_sample": same_plain,
           "engine": {"epochs": c1["epochs"] - c0["epochs"], "gpu_s": c1["gpu_s"] - c0["gpu_s"],
                      "tuned_team_per_bucket": tuning}}
    if cpu is not None:
        c = cpu["mpc"]
        out["cpu_oracle"] = {"qps": c["qps"], "processes": cpu["workers"], "single_core_qps": c["single_core_qps"],
                             "solve_ms": c["solve_ms"], "kind": "port",
                             "sample": f"the same {n_req} timed requests, setup + solve each on the oracle's direct "
                                       f"(LDL') path, {cpu['workers']} processes"}
        out["ratio_vs_cpu_oracle"] = qps / c["qps"]
    return out


def serve_portfolio(n_problems, n_req, seed, cpu):
    """Portfolio stream: n ~ U{500..5000} assets, problems resident on the GPU;
    each request draws a new mu, replaces q (osqp_update_data_vec semantics)
    and re-solves warm-started.  Closed loop: an idle problem takes its next
    request as soon as its previous result is seen.  Statuses and objectives
    are compared with the CPU oracle's direct path on the same request
    sequence; the plain FP64 engine, one resident solver per problem doing
    update(q) + solve, on a sample of the same requests, is the GPU comparison."""
    from osqp_amd import OSQP
    from osqp_amd.serve import BatchEngine
    st = dict(eps_abs=1e-4, eps_rel=1e-4)
    t0 = time.time()
    probs, reqs = portfolio_stream(seed, n_problems, n_req)
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
    issued, lat, status, objs, its_of = {}, [], {}, {}, {}
    for k in range(n_problems):
        if queue[k]:
            j = queue[k].pop(0)
            eng.resolve(pid[k], reqs[j][1])
            issued[pid[k]] = (j, time.perf_counter())
    t_log = time.perf_counter()
    capped = False
    # completion times: the stream's end waits on the largest problems' last
    # solves (each problem takes its requests one after another), so the
    # rate up to half and nine tenths of the requests is stated beside the
    # whole stream's
    t_done = []
    while issued:
        if time.perf_counter() - t_log > 5.0:
            log(f"[serve] portfolio: {len(lat)}/{n_req} requests done")
            t_log = time.perf_counter()
        if not capped and time.perf_counter() - t_start > 120.0:
            # bound the section's run time: stop issuing, drain what is in flight
            capped = True
            for k in queue:
                queue[k] = []
        for i in eng.step():
            r = eng.poll(i)
            j, t_iss = issued.pop(i)
            lat.append(time.perf_counter() - t_iss)
            t_done.append(time.perf_counter() - t_start)
            status[j] = r.info.status_val
            objs[j] = r.info.obj_val
            its_of[j] = r.info.iter
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
    ts = time.perf_counter()
    for j in sample:
        k, q = reqs[j]
        solvers[k].update(q=q)
        solvers[k].solve()
    plain_qps = k_plain / (time.perf_counter() - ts)
    for s in solvers.values():
        s.close()
    vs_oracle = None
    if cpu is not None:
        cst, cob = cpu["portfolio"]["status"], cpu["portfolio"]["obj"]
        done = sorted(status)
        errs = [abs(objs[j] - cob[j]) / max(1.0, abs(cob[j])) for j in done if cst[j] == 1 and status[j] == 1]
        # the requests whose status differs from the oracle's (at most 10)
        diff = [{"request": j, "assets": int(probs[reqs[j][0]]["P"].shape[0] - probs[reqs[j][0]]["k"]),
                 "gpu_status": int(status[j]), "gpu_iter": int(its_of[j]), "oracle_status": int(cst[j])}
                for j in done if status[j] != cst[j]][:10]
        vs_oracle = {"requests": len(done), "statuses_identical": int(sum(int(status[j] == cst[j]) for j in done)),
                     "status_differences": diff,
                     "obj_rel_err_vs_direct": {"max": float(np.max(errs)) if errs else None,
                                               "p50": percentile(errs, 50)},
                     "obj_rel_err_vs_fp64_optimum": {
                         "gpu_batched_fp32": obj_errors(objs, cpu["portfolio"]["opt"], done),
                         "oracle_direct_eps_1e-4": obj_errors(dict(enumerate(cob)), cpu["portfolio"]["opt"], done)}}
    log(f"[serve] portfolio batched FP32: {qps:.1f} QPs/s, p50 {np.percentile(lat, 50) * 1e3:.1f} ms, "
        f"p99 {np.percentile(lat, 99) * 1e3:.1f} ms; plain FP64: {plain_qps:.1f} QPs/s")
    out = {"workload": f"portfolio stream, {n_problems} resident problems n~U{{500..5000}}, {n_done} mu updates "
                       f"with warm-started re-solves, eps 1e-4, closed loop",
           "qps": qps, "latency_ms": {"p50": float(np.percentile(lat, 50) * 1e3),
                                      "p99": float(np.percentile(lat, 99) * 1e3)},
           "rate_to_fraction_done": {f"{int(f * 100)}%": {"requests": int(f * n_done),
                                                          "seconds": t_done[int(f * n_done) - 1],
                                                          "qps": int(f * n_done) / t_done[int(f * n_done) - 1]}
                                     for f in (0.5, 0.9) if int(f * n_done) >= 1},
           "statuses_solved": int(sum(1 for v in status.values() if v == 1)),
           "vs_cpu_oracle_direct": vs_oracle,
           "plain_fp64_qps": plain_qps, "ratio_vs_plain_fp64": qps / plain_qps, "plain_sample": k_plain}
    if cpu is not None:
        c = cpu["portfolio"]
        out["cpu_oracle"] = {"qps": c["qps_packed"], "processes": cpu["workers"],
                             "single_core_qps": c["single_core_qps"], "solve_ms": c["solve_ms"], "kind": "port",
                             "sample": f"all {n_req} requests in stream order per problem, update_data_vec + "
                                       f"warm-started solve on the oracle's direct path; QPs/s = requests / "
                                       f"(solve seconds / {cpu['workers']} processes)"}
    return out


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
            "kkt_certificate": {"prim": c["prim"], "dual": c["dual"]},
            "err_vs_oracle_fp64_pcg": {"fp32_pcg": oracle_fixture_errors("lasso", p, r32.x, r32.y, r32.info.obj_val),
                                       "fp64_plain": oracle_fixture_errors("lasso", p, r64.x, r64.y, r64.info.obj_val)}}


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
    ap.add_argument("--mpc-requests", type=int, default=20000)
    ap.add_argument("--mpc-warmup", type=int, default=2000)
    ap.add_argument("--portfolio-problems", type=int, default=64)
    ap.add_argument("--portfolio-requests", type=int, default=512)
    ap.add_argument("--cpu-legs", action="store_true", help=argparse.SUPPRESS)
    ap.add_argument("--cpu-large-qp", default=None, help=argparse.SUPPRESS)
    ap.add_argument("--profile-kernels", action="store_true",
                    help="set up the QP, launch only the timed SpMV kernels (reps=20 each, A then K) and a "
                         "256 MiB device copy (counter calibration); for rocprofv3 runs")
    ap.add_argument("--profile-workload", choices=["large_qp", "lasso"], default="large_qp",
                    help="the QP --profile-kernels sets up: the large random QP or the full-size lasso")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py runs on one GPU only: this build has no multi-GPU path (one MI355X per job); "
              "use --gpus 1", file=sys.stderr)
        sys.exit(2)

    if args.cpu_legs:
        lq = tuple(int(v) for v in args.cpu_large_qp.split("x")) if args.cpu_large_qp else None
        print(json.dumps(cpu_legs(args.seed, args.mpc_warmup, args.mpc_requests, args.portfolio_problems,
                                  args.portfolio_requests, lq)), flush=True)
        return

    n, m = (5000, 50000) if args.small else (50000, 500000)
    # CPU legs first, in a child process, before this process touches the GPU
    cpu = None
    if not args.no_cpu:
        t0 = time.time()
        cpu = run_cpu_legs(args, (n, m))
        log(f"[bench] CPU legs done in {time.time() - t0:.0f}s on {cpu['host']}")

    from osqp_amd import OSQP, device_name
    from osqp_amd import problems as pr

    dens_M, dens_A = (1e-3, 2e-2) if args.small else (1e-4, 2e-3)
    t0 = time.time()
    if args.profile_kernels and args.profile_workload == "lasso":
        prob = pr.lasso(20000, 200000, 1e-3, seed=args.seed)
        n, m = prob["P"].shape[0], prob["A"].shape[0]
    else:
        prob = pr.random_qp(n, m, dens_M, dens_A, seed=args.seed)
    log(f"[bench] generated n={n} m={m} nnz(A)={prob['A'].nnz} nnz(P triu)={prob['P'].nnz} in {time.time() - t0:.1f}s")
    settings = dict(eps_abs=1e-4, eps_rel=1e-4, verbose=0)

    # ---- FP32-PCG engine (the measured configuration) -------------------------
    t0 = time.time()
    eng = OSQP(precision=32)
    rc = eng.setup(prob["P"], prob["q"], prob["A"], prob["l"], prob["u"], **settings)
    assert rc == 0, f"setup failed: {rc}"
    lib_sync()
    log(f"[bench] FP32-PCG setup {time.time() - t0:.2f}s on {device_name()}")
    if args.profile_kernels:
        import ctypes as C
        kt = eng.time_kernels(reps=20)
        kt64 = eng.time_fp64_kernels(reps=20)
        # a 256 MiB device-to-device copy, the counters' calibration
        hip = C.CDLL("libamdhip64.so.7")
        nb = 256 << 20
        a, b = C.c_void_p(), C.c_void_p()
        assert hip.hipMalloc(C.byref(a), C.c_size_t(nb)) == 0 and hip.hipMalloc(C.byref(b), C.c_size_t(nb)) == 0
        assert hip.hipMemset(a, 1, C.c_size_t(nb)) == 0
        for _ in range(5):
            assert hip.hipMemcpy(b, a, C.c_size_t(nb), 3) == 0  # device to device
        assert hip.hipDeviceSynchronize() == 0
        hip.hipFree(a)
        hip.hipFree(b)
        print(json.dumps({"profile_kernels": kt, "fp64_kernels": kt64, "copy_bytes_each_way": nb}), flush=True)
        eng.close()
        return
    for w in range(args.warmup):
        eng.cold_start()
        r = eng.solve()
        log(f"[bench] warmup {w}: status={r.info.status} iter={r.info.iter} t={r.info.solve_time:.3f}s")
    lat, iters, statuses = [], [], []
    st0 = eng.stats()
    lib_sync()
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
    lib_sync()
    elapsed = time.perf_counter() - t_start
    st1 = eng.stats()
    x32, y32, obj32 = r.x, r.y, r.info.obj_val
    cert32 = pr.kkt_certificate(prob, x32, y32)
    qps = args.steps / elapsed
    admm_iters = int(np.sum(iters))
    cg_iters = st1.pcg_iters - st0.pcg_iters
    ms_per_admm = elapsed * 1e3 / max(admm_iters, 1)

    # ---- whole ADMM iteration: algorithmic bytes / measured time ----------------
    # The convention's bytes (the stated algorithm's compulsory traffic, int32
    # indices) at the solve's own CG iterations per ADMM iteration and
    # termination checks per ADMM iteration.  Beside them, the bytes this
    # implementation's own format and passes need at minimum (from the
    # library: 16-bit panel indices, the incremental z on seven solves of
    # eight, the fused ADMM update at 6n + 12m state touches including the
    # next right-hand side and the PCG's w): fewer bytes for the same work.
    import scipy.sparse as sp
    cg_per_admm = cg_iters / max(admm_iters, 1)
    P_full = prob["P"] + sp.triu(prob["P"], 1).T
    checks_per_iter = np.ceil(admm_iters / args.steps / CHECK_INTERVAL) / (admm_iters / args.steps)
    conv = convention_bytes(n, m, int(prob["A"].nnz), int(P_full.nnz), cg_per_admm, checks_per_iter)
    admm_bytes = conv["admm"]
    admm_gbs = admm_bytes / (ms_per_admm * 1e-3) / 1e9
    pb = eng.pcg_bytes()
    check_b = eng.check_bytes()
    own_bytes = (pb["per_solve"] - 4.0 * m * 8 + cg_per_admm * pb["per_cg_iter"] + 8.0 * (6 * n + 12 * m) +
                 check_b * checks_per_iter)

    # ---- dominant kernels: CG operator, timed with HIP events ------------------
    kt = eng.time_kernels(reps=20)
    dom = "k_Kv" if kt["ms_kv"] >= kt["ms_av"] else "k_Av_rho"
    dom_ms = max(kt["ms_kv"], kt["ms_av"])
    # the convention's bytes of one application (K p: P_full, A', sigma p;
    # w = rho o (A p): A, rho); the panel format's own minimum beside them
    dom_bytes = conv["kv"] if dom == "k_Kv" else conv["av"]
    dom_own = kt["bytes_kv"] if dom == "k_Kv" else kt["bytes_av"]
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
                    err_vs_oracle_fp64_pcg=oracle_fixture_errors("large_qp", prob, r64.x, r64.y, r64.info.obj_val),
                    err_x_inf=float(np.max(np.abs(x32 - r64.x))),
                    err_y_inf=float(np.max(np.abs(y32 - r64.y))),
                    err_obj_rel=float(abs(obj32 - r64.info.obj_val) / max(1.0, abs(r64.info.obj_val))))
        log(f"[bench] FP64 plain: status={r64.info.status} iter={r64.info.iter} t={t64:.3f}s "
            f"|dx|={fp64['err_x_inf']:.2e} |dobj|={fp64['err_obj_rel']:.2e}")
        e64.close()

    # ---- CPU baseline: oracle FP64 PCG over one termination-check interval -----
    # Cost model: an ADMM iteration of the PCG path streams A and A' about
    # 2 + 2 k_cg times (right-hand side A'(rho b2), z = A x, and A, A' per CG
    # iteration), so the sample's seconds per ADMM iteration are scaled by that
    # count at the GPU solve's own k_cg, times the GPU solve's ADMM iterations.
    cpu_baseline = None
    if cpu is not None and "large_qp" in cpu:
        c = cpu["large_qp"]
        k_cpu = c["cg_iters"] / max(c["admm_iters"], 1)
        mean_iters = admm_iters / args.steps
        s_per_admm = c["solve_s"] / max(c["admm_iters"], 1) * (2 + 2 * cg_per_admm) / (2 + 2 * k_cpu)
        cpu_baseline = dict(value=1.0 / (s_per_admm * mean_iters), unit="QPs/s", cores=1, kind="port",
                            sample=f"oracle FP64 PCG (oracle/osqp_oracle.c, linsys_solver=2) on the same n={n} m={m} "
                                   f"QP: {c['admm_iters']} ADMM iterations ({c['cg_iters']} CG) in {c['solve_s']:.1f}s, "
                                   f"one thread; scaled to {mean_iters:.0f} ADMM iterations at the GPU's "
                                   f"{cg_per_admm:.2f} CG per iteration (A/A' passes 2 + 2 k_cg)",
                            host=cpu["host"])
        log(f"[bench] cpu oracle: {c['admm_iters']} ADMM iterations in {c['solve_s']:.1f}s -> "
            f"{cpu_baseline['value']:.2e} QPs/s")

    # ---- served MPC stream: batched FP32-PCG engine vs plain FP64 engine -------
    small = None if args.no_serving else random_qp_small_section(args.seed)
    lasso = None if args.no_serving else lasso_section(args.seed)
    serving = None if args.no_serving else serve_mpc(args.mpc_requests, args.mpc_warmup, args.seed, cpu)
    serving_pf = None if args.no_serving else serve_portfolio(args.portfolio_problems, args.portfolio_requests,
                                                              args.seed, cpu)

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
                      "frac_of_peak": admm_gbs / HBM_PEAK_GBS, "convention": "compulsory traffic of the stated "
                      "algorithm, int32 indices, FP32 values and PCG vectors, FP64 state",
                      "bytes_per_cg_iter": conv["cg"], "bytes_per_termination_check": conv["check"],
                      "termination_checks_per_admm_iter": checks_per_iter,
                      "own_format_minimum": {"bytes": own_bytes, "achieved_GBs": own_bytes / (ms_per_admm * 1e-3) / 1e9,
                                             "frac_of_peak": own_bytes / (ms_per_admm * 1e-3) / 1e9 / HBM_PEAK_GBS,
                                             "bytes_per_cg_iter": pb["per_cg_iter"],
                                             "bytes_per_linsys_outer": pb["per_solve"] - 4.0 * m * 8,
                                             "bytes_per_termination_check": check_b}},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": measured_traffic(dom), "kernel": dom,
                     "ms_per_launch": dom_ms, "bytes_per_launch": dom_bytes,
                     "own_format_bytes_per_launch": dom_own,
                     "own_format_GBs": dom_own / (dom_ms * 1e-3) / 1e9,
                     "other_kernel": {"name": "k_Av_rho" if dom == "k_Kv" else "k_Kv",
                                      "ms_per_launch": min(kt["ms_kv"], kt["ms_av"]),
                                      "bytes_per_launch": conv["av"] if dom == "k_Kv" else conv["kv"],
                                      "own_format_bytes_per_launch": kt["bytes_av"] if dom == "k_Kv" else kt["bytes_kv"]},
                     "fp64_residual_kernels": fp64_kernels},
        "err_vs_oracle_fp64_pcg": oracle_fixture_errors("large_qp", prob, x32, y32, obj32),
        "gpu_fp64_plain": fp64,
        "cpu_baseline": cpu_baseline,
        "random_qp_small": small,
        "lasso": lasso,
        "serving_mpc": serving,
        "serving_portfolio": serving_pf,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
