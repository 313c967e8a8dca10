"""Benchmark of the fortfem sparse-assembly + Jacobi-PCG path on one MI355X.

Workload (default ``c4``): P1 tetrahedral Poisson on the unit cube split into
128^3 hexes x 6 tets (12,582,912 tets, 2,146,689 dofs, 2,048,383 free), per-cell
kappa = exp(N(0,1)) from a seed, right-hand side U[-1, 1] per free dof, u = 0
on the boundary, Jacobi-PCG to a relative residual of 1e-10 (``--workload c2``:
P2 on 32^3 x 6 with the sin load).

One step = numeric assembly of the stiffness into the resident CSC pattern +
right-hand side + Dirichlet reduction + the solve, with every input already in
HBM.
The headline solve uses the mixed path (FP32 Jacobi-PCG iterations with FP64
x, reductions and residual refreshes); the FP64 path (the reference's
recurrence) is timed beside it and the mixed solution's relative error
against it is reported.

Prints one JSON line.  The dominant kernel's roofline figure comes from HIP
events recorded around every 8th FP32 SpMV launch of the iterations inside the
timed solves: ``achieved`` divides the algorithmic bytes of an FP32 SpMV over
the int32 CSR (8 nnz + 12 n + 4) by the launch's average duration, and
``achieved_format`` the bytes the tuned variant moves in its own storage
format.  The same line carries the served-stream keys (``served``: the first
requests of the seeded request stream, batched mixed against one-at-a-time
FP64, saturated and under Poisson arrivals) and the CPU baseline (the oracle
on all host cores on the full configuration).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


# the roofline kernel: the FP32 SpMV launch of the mixed solve's iterations
ROOFLINE_KERNEL = {"c4": "k_pcg_apdot", "c2": "k_pcg_apdot"}
ROOFLINE_TYPES = "float, float>"


def _pmc_pass(counter, cmd, tmp, kernel_regex=None, timeout=240):
    """One rocprofv3 --pmc pass (one counter) over a child run of this
    script: {kernel name: [(counter value, duration us), ...]} from its
    counter CSV, or raises."""
    import csv
    import glob
    import shutil
    import subprocess

    out = os.path.join(tmp, counter)
    shutil.rmtree(out, ignore_errors=True)
    prof = [shutil.which("rocprofv3") or "rocprofv3", "--pmc", counter, "--output-format", "csv",
            "-d", out, "-o", "p"]
    if kernel_regex:
        prof += ["--kernel-include-regex", kernel_regex]
    import signal

    env = dict(os.environ, TMPDIR=tmp)
    # its own process group, so that a time-out ends the profiled child too
    proc = subprocess.Popen(prof + ["--", sys.executable, os.path.abspath(__file__)] + cmd, cwd=tmp,
                            env=env, stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL,
                            start_new_session=True)
    try:
        rc = proc.wait(timeout=timeout)
    except subprocess.TimeoutExpired:
        os.killpg(proc.pid, signal.SIGKILL)
        proc.wait()
        raise RuntimeError(f"rocprofv3 --pmc {counter} did not finish in {timeout} s")
    files = glob.glob(os.path.join(out, "**", "*counter_collection.csv"), recursive=True)
    if rc != 0 or not files:
        raise RuntimeError(f"rocprofv3 --pmc {counter} exited {rc}")
    res = {}
    for row in csv.DictReader(open(files[0])):
        if row["Counter_Name"] != counter:
            continue
        d = (int(row["End_Timestamp"]) - int(row["Start_Timestamp"])) / 1e3
        res.setdefault(row["Kernel_Name"], []).append((float(row["Counter_Value"]), d))
    return res


def measure_traffic(workload: str):
    """HBM bytes per live launch of the roofline kernel, measured now: two
    rocprofv3 --pmc passes (FETCH_SIZE, then WRITE_SIZE, in KiB per dispatch)
    over a one-step child run of this workload, each scaled by what the same
    counter reads for a device copy of 1 GiB (--workload copy).  Returns
    (bytes or None, description)."""
    import shutil
    import tempfile

    if shutil.which("rocprofv3") is None:
        return None, "not measured: rocprofv3 is not on PATH"
    if any("rocprof" in os.environ.get(k, "") for k in ("LD_PRELOAD", "HSA_TOOLS_LIB")):
        return None, "not measured: this run is itself under a profiler"
    kname = ROOFLINE_KERNEL.get(workload)
    tmp = tempfile.mkdtemp(prefix="ffh_pmc_")
    try:
        child = ["--workload", workload, "--steps", "1", "--warmup", "0", "--no-cpu", "--no-served",
                 "--no-traffic"]
        kf = _pmc_pass("FETCH_SIZE", child, tmp, kname)
        kw = _pmc_pass("WRITE_SIZE", child, tmp, kname)
        cf = _pmc_pass("FETCH_SIZE", ["--workload", "copy", "--steps", "3"], tmp)
        cw = _pmc_pass("WRITE_SIZE", ["--workload", "copy", "--steps", "3"], tmp)

        def live(res):  # the roofline kernel's launches that did work (> 5 us)
            vals = [v for k, vs in res.items() if kname in k and ROOFLINE_TYPES in k
                    for v, d in vs if d > 5.0]
            if not vals:
                raise RuntimeError(f"no live {kname} launch in the counter output")
            return sum(vals) / len(vals)

        def copy_value(res):  # the copy kernel's dispatches (torch's multiply by 1.0)
            vals = [v for k, vs in res.items() if "MulFunctor" in k for v, _ in vs]
            if not vals:
                raise RuntimeError("no copy dispatch in the calibration counters")
            return float(np.median(vals))

        gib_kib = float(1 << 20)
        f_scale, w_scale = gib_kib / copy_value(cf), gib_kib / copy_value(cw)
        fk, wk = live(kf), live(kw)
        traffic = (f_scale * fk + w_scale * wk) * 1024.0
        return traffic, (f"measured in this run: {kname}<.., {ROOFLINE_TYPES} {f_scale:.2f} x FETCH_SIZE "
                         f"{fk:,.1f} KiB + "
                         f"{w_scale:.2f} x WRITE_SIZE {wk:,.1f} KiB per live launch (rocprofv3 --pmc "
                         f"over bench.py --workload {workload} --steps 1 --warmup 0; scale factors from "
                         f"a 1 GiB device copy, bench.py --workload copy)")
    except Exception as e:  # noqa: BLE001 -- the line reports why it has no figure
        return None, f"not measured: {e}"
    finally:
        shutil.rmtree(tmp, ignore_errors=True)


def spmv_bytes(nnz: int, n: int, sv: int) -> float:
    """Algorithmic bytes of one CSR-gather SpMV: values + int32 column indices
    + int32 row pointers + x once + y once."""
    return nnz * (sv + 4) + 4.0 * (n + 1) + n * 2.0 * sv


def dirspmv_bytes(nnz: int, n: int, sv: int) -> float:
    """Algorithmic bytes of the direction + SpMV launch of a PCG iteration:
    the SpMV's values, column indices and row pointers, the three vectors the
    new direction is formed from (r, D^-1, p_old) read once, and the new
    direction and A p written once."""
    return nnz * (sv + 4) + 4.0 * (n + 1) + n * 5.0 * sv


def build_problem(ff, workload: str):
    from fortfem_amd import workloads

    if workload == "c2":
        N, degree, cfg = 32, 2, 2
    elif workload == "c4":
        N, degree, cfg = 128, 1, 4
    else:
        raise SystemExit(f"unknown workload {workload}")
    v, t = ff.unit_box_mesh(N)
    kappa = workloads.lognormal_kappa(t.shape[0], cfg)
    p = ff.Problem(v, t)
    p.pattern_build(ff.capi.FAMILY_H1, degree)
    q = 2 * degree + 4
    p.assemble(q, cell_kappa=kappa)
    if workload == "c2":
        p.load(q, ff.capi.SOURCE_SIN_PRODUCT, [1.0, np.pi, np.pi, np.pi])
        p.dirichlet(0)
        rhs = None
    else:
        p.dirichlet(0)
        n_free = p.sizes()["n_free"]
        rhs = workloads.uniform_rhs(n_free, cfg)
        p.set_rhs(rhs)
    meta = dict(N=N, degree=degree, q=q, ncell=int(t.shape[0]), config=cfg)
    return p, v, t, kappa, rhs, meta


def c3_tensors(T: int, seed: int):
    """K_e = Q diag(exp(0.5 z)) Q^T, Q from a normalised 4-normal quaternion
    (uniform on SO(3)), z ~ N(0, 1): the per-cell SPD tensors of the
    tensor-weighted Nedelec workload."""
    from fortfem_amd import workloads

    g = workloads.normal(seed, 7 * T).reshape(T, 7)
    q = g[:, :4] / np.linalg.norm(g[:, :4], axis=1, keepdims=True)
    w, x, y, z = q.T
    Q = np.stack([
        np.stack([1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)], -1),
        np.stack([2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)], -1),
        np.stack([2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)], -1)], 1)
    d = np.exp(0.5 * g[:, 4:])
    K = np.einsum("tij,tj,tkj->tik", Q, d, Q)
    return 0.5 * (K + K.transpose(0, 2, 1))


def run_c3(args):
    """Tensor-weighted Nedelec curl-mass assembly on the 16^3x6 cube (24,576
    tets), orders 1..6, FP64 and FP32 contraction; assembly only (pattern,
    colouring and class tables are set-up, built once per order)."""
    import fortfem_amd as ff
    from fortfem_amd import workloads

    v, t = ff.unit_box_mesh(16)
    T = t.shape[0]
    K = c3_tensors(T, workloads.BASE_SEED + 3)
    rows = []
    for k in args.orders:
        t0 = time.perf_counter()
        p = ff.Problem(v, t)
        n, nnz = p.pattern_build(ff.capi.FAMILY_NEDELEC, k)
        t_pat = time.perf_counter() - t0
        t0 = time.perf_counter()
        p.assemble(-1, mass=1.0, cell_tensor=K)  # class tables + warm-up
        t_tab = time.perf_counter() - t0
        nl = k * (k + 2) * (k + 3) // 2
        row = {"order": k, "n": n, "nnz": nnz, "n_local": nl, "colors": p.sizes()["n_colors"],
               "pattern_s": round(t_pat, 2), "tables_s": round(t_tab, 2)}
        vals = {}
        for prec in ("fp64", "fp32"):
            ts = []
            for _ in range(max(1, args.steps)):
                t1 = time.perf_counter()
                p.assemble(-1, mass=1.0, reuse_tensor=True, precision=prec)
                ts.append(time.perf_counter() - t1)
            ms = float(np.median(ts)) * 1e3
            sv = 8 if prec == "fp64" else 4
            # per element entry: int32 slot + value store (first touch of a slot,
            # nnz of them) or read-modify-write (the rest); 12 FP64 terms per cell
            ent = float(T) * nl * nl
            by = 4.0 * ent + sv * nnz + 2.0 * sv * (ent - nnz) + 96.0 * T
            row[prec + "_ms"] = round(ms, 3)
            row[prec + "_gbs"] = round(by / (ms * 1e-3) / 1e9, 1)
            row[prec + "_gflops"] = round(24.0 * nl * nl * T / (ms * 1e-3) / 1e9, 1)
            vals[prec] = p.matrix(ff.capi.FULL).download().val
        d = vals["fp32"] - vals["fp64"]
        row["fp32_rel_fro"] = float(np.linalg.norm(d) / np.linalg.norm(vals["fp64"]))
        row["fp32_rel_max"] = float(np.max(np.abs(d)) / np.max(np.abs(vals["fp64"])))
        if args.point_tensors:
            row.update(c3_point_tensors(ff, p, v, t, K, k, nl, args.steps))
        rows.append(row)
        print(json.dumps({"progress": row}), file=sys.stderr, flush=True)
        del p, vals
    out = {"metric": "tensor-weighted Nedelec assembly time per order (FP64 and FP32 contraction)",
           "value": rows[-1]["fp64_ms"], "unit": "ms", "n_gpus": 1, "steps": args.steps,
           "higher_is_better": False, "dtype": "f64+f32", "data": "synthetic (seeded SPD tensors)",
           "config": {"workload": "c3: Nedelec k=%s on 16^3x6 (24576 tets), K_e = Q diag(exp(0.5 z)) Q^T, "
                                  "beta = 1" % ",".join(str(k) for k in args.orders)},
           "orders": rows}
    print(json.dumps(out))


MFMA_PEAK_TFLOPS = {"fp64": 78.6, "fp32": 157.3}


def c3_point_tensors(ff, p, v, t, K, k, nl, steps):
    """The same assembly with the coefficient evaluated at every point of the
    rule, K(x) = K_e (1 + 0.25 sin 2 pi x sin 2 pi y sin 2 pi z): the curl
    part is the dense per-cell C^T blockdiag(M_q) C product on MFMA
    (k_ned_point_curl), the mass part the class-table kernel.  Times the
    assembly with the point tensors resident; MFMA rate on the executed
    count (the kernel's 32 x 32 regions of the upper triangle, 16 x 16 x 4
    MFMAs over the rule padded to whole chunks) and the useful upper-triangle
    count n (n + 1) / 2 * 2 * 3 nq per cell."""
    T = t.shape[0]
    nq = ff.nedelec_load_points(k).shape[0]
    X = ff.nedelec_tensor_points(v, t, k)
    s = 1.0 + 0.25 * np.sin(2 * np.pi * X[..., 0]) * np.sin(2 * np.pi * X[..., 1]) * np.sin(2 * np.pi * X[..., 2])
    Kp = K[:, None] * s[..., None, None]
    p.assemble(-1, mass=1.0, point_tensor=Kp)
    del Kp, X, s
    ldn = 32 * ((nl + 31) // 32)
    ns = ldn // 32
    regions = ns * (ns + 1) // 2
    out = {"point_nq": nq}
    vals = {}
    for prec in ("fp64", "fp32"):
        pc = 4 if prec == "fp64" else 8
        steps_k = -(-nq // pc) * 3 * pc // 4
        executed = float(T) * regions * 4 * steps_k * 2.0 * 16 * 16 * 4
        useful = float(T) * nl * (nl + 1) / 2 * 2.0 * 3 * nq
        ts = []
        for _ in range(max(1, steps)):
            t1 = time.perf_counter()
            p.assemble(-1, mass=1.0, reuse_point_tensor=True, precision=prec)
            ts.append(time.perf_counter() - t1)
        ms = float(np.median(ts)) * 1e3
        out["point_" + prec + "_ms"] = round(ms, 3)
        out["point_" + prec + "_mfma_tflops_executed"] = round(executed / (ms * 1e-3) / 1e12, 2)
        out["point_" + prec + "_mfma_frac"] = round(executed / (ms * 1e-3) / 1e12 / MFMA_PEAK_TFLOPS[prec], 4)
        out["point_" + prec + "_tflops_useful"] = round(useful / (ms * 1e-3) / 1e12, 2)
        vals[prec] = p.matrix(ff.capi.FULL).download().val
    d = vals["fp32"] - vals["fp64"]
    out["point_fp32_rel_fro"] = float(np.linalg.norm(d) / np.linalg.norm(vals["fp64"]))
    out["point_fp32_rel_max"] = float(np.max(np.abs(d)) / np.max(np.abs(vals["fp64"])))
    return out


def served_stream(args):
    """Served stream: the first `requests` requests of the seeded stream of
    4096 (P1 35 %, P2 35 %, Nedelec k=1 15 %, k=2 15 %; cube N^3 x 6 with N in
    8..48; lognormal per-cell coefficient; 1..8 right-hand sides U[-1, 1]),
    every one of them, solved to a relative residual of 1e-10 by the batched
    server (FP32 iterations, FP64 refresh).  Three timed passes over the same
    requests: saturated (all queued at t = 0: served solves/s), Poisson
    arrivals at 0.8 x the saturated request rate (latency p50 / p99 from
    arrival to result), and the baseline, one right-hand side at a time by the
    FP64 Jacobi-PCG of the drop-in path on the resident problem (assembly per
    request included).  Patterns are built once per (N, family, order)
    before timing."""
    import fortfem_amd as ff
    from fortfem_amd import workloads

    stream = workloads.served_stream(args.stream)[:args.requests]
    meshes = {}
    for r in stream:
        if r["N"] not in meshes:
            meshes[r["N"]] = ff.unit_box_mesh(r["N"])
    srv = ff.Server(budget_bytes=int(args.budget_gb * (1 << 30)), max_systems=args.slots, chunk=args.chunk)
    srv.set_precision(args.precision)
    srv.use_graph(args.served_graph)
    probs = {}
    t0 = time.perf_counter()
    for r in stream:
        key = (r["N"], r["family"], r["order"])
        if key not in probs:
            v, t = meshes[r["N"]]
            n_free, nnz = srv.prepare(r["N"], r["family"], r["order"], v, t)
            p = ff.Problem(v, t)
            p.pattern_build(r["family"], r["order"])
            q = 2 * r["order"] + 4 if r["family"] == 1 else -1
            p.assemble(q, cell_kappa=np.ones(t.shape[0]), mass=0.0 if r["family"] == 1 else 1.0)
            p.dirichlet(0)
            probs[key] = (p, q, n_free, nnz, t.shape[0])
        r["n_free"], r["nnz"] = probs[key][2], probs[key][3]
        r["coeff"], r["rhs"] = workloads.request_data(r, probs[key][4], r["n_free"])
    t_prep = time.perf_counter() - t0
    mass = lambda r: 0.0 if r["family"] == 1 else 1.0
    solves = sum(r["nrhs"] for r in stream)
    dofs = sum(r["nrhs"] * r["n_free"] for r in stream)

    def serve(arrivals):
        """Submits request i at arrivals[i] seconds (steps the server while
        waiting); returns {index: (result, latency_s)} and the wall time."""
        t0 = time.perf_counter()
        pending, served, nxt = {}, {}, 0
        t_said = 0.0
        while nxt < len(stream) or pending:
            now = time.perf_counter() - t0
            if now - t_said >= 20.0:  # a line every 20 s on long streams
                t_said = now
                print(json.dumps({"progress": "served stream: %d of %d requests done at %.0f s"
                                  % (len(served), len(stream), now)}), file=sys.stderr, flush=True)
            while nxt < len(stream) and arrivals[nxt] <= now:
                r = stream[nxt]
                tk = srv.submit(r["N"], r["family"], r["order"], r["coeff"], r["rhs"], mass=mass(r),
                                tol=1e-10, max_iterations=args.max_iter)
                pending[tk[0]] = (tk, r, arrivals[nxt])
                nxt += 1
            if not pending:
                time.sleep(max(0.0, min(arrivals[nxt] - now, 0.05)))
                continue
            srv.step()
            for tid in list(pending):
                tk, r, ta = pending[tid]
                res = srv.poll(*tk)
                if res is not None:
                    served[r["index"]] = (res, time.perf_counter() - t0 - ta)
                    del pending[tid]
        return served, time.perf_counter() - t0

    def progress(msg):
        print(json.dumps({"progress": "served stream: " + msg}), file=sys.stderr, flush=True)

    served, t_sat = serve([0.0] * len(stream))
    progress("saturated pass %.2f s" % t_sat)
    lat_sat = np.array([served[r["index"]][1] for r in stream]) * 1e3
    rate = args.arrival_factor * len(stream) / t_sat  # requests per second
    gaps = -np.log(1.0 - workloads.uniform(workloads.BASE_SEED + 5 + 77, len(stream))) / rate
    arrivals = np.concatenate([[0.0], np.cumsum(gaps)[:-1]])
    served_p, t_poi = serve(list(arrivals))
    progress("Poisson pass %.2f s" % t_poi)
    lat = np.array([served_p[r["index"]][1] for r in stream]) * 1e3

    # FP64 baseline: one right-hand side at a time on the resident problem
    o64 = ff.solver_options(method="pcg", preconditioner="jacobi", tolerance=1e-10, max_iterations=args.max_iter,
                            breakdown="relative", check_interval=32)
    errs, t_base, iters64 = [], 0.0, 0
    free = {key: ~val[0].constrained() for key, val in probs.items()}
    for r in stream:
        p, q, n_free, nnz, T = probs[(r["N"], r["family"], r["order"])]
        t1 = time.perf_counter()
        p.assemble(q, cell_kappa=r["coeff"], mass=mass(r))
        t_base += time.perf_counter() - t1
        x_srv = served[r["index"]][0][0]
        for j in range(r["nrhs"]):
            p.set_rhs(r["rhs"][j])
            t1 = time.perf_counter()
            _, st = p.solve(o64, download=False)
            t_base += time.perf_counter() - t1
            iters64 += st.iterations
            x64 = p.solution()[free[(r["N"], r["family"], r["order"])]]
            errs.append(float(np.linalg.norm(x_srv[j] - x64) / max(np.linalg.norm(x64), 1e-300)))
        progress("FP64 one at a time: request %d done, %.2f s so far" % (r["index"], t_base))
    iters_srv = int(sum(int(served[r["index"]][0][1].sum()) for r in stream))
    conv = all(bool(served[r["index"]][0][2].all()) and bool(served_p[r["index"]][0][2].all()) for r in stream)
    same = all(np.array_equal(served[r["index"]][0][0], served_p[r["index"]][0][0]) for r in stream)
    kinds = {}
    for r in stream:
        kk = ("P%d" % r["order"]) if r["family"] == 1 else ("N%d" % r["order"])
        kinds[kk] = kinds.get(kk, 0) + 1
    return {
        "workload": "c5 served stream: requests 0..%d of the seeded 4096-request stream, all of them "
                    "(%d solves: %s), N in %d..%d" % (len(stream) - 1, solves, kinds,
                                                      min(r["N"] for r in stream), max(r["N"] for r in stream)),
        "solves_per_s": round(solves / t_sat, 2),
        "fp64_one_at_a_time_solves_per_s": round(solves / t_base, 2),
        "speedup_vs_fp64": round(t_base / t_sat, 2),
        "dof_per_s": round(dofs / t_sat, 1), "fp64_dof_per_s": round(dofs / t_base, 1),
        "p50_ms": round(float(np.percentile(lat, 50)), 1), "p99_ms": round(float(np.percentile(lat, 99)), 1),
        "arrival": "Poisson, rate %.3f requests/s (%.2f x the saturated rate), seeded" % (rate, args.arrival_factor),
        "poisson_wall_s": round(t_poi, 3),
        "saturated_p50_ms": round(float(np.percentile(lat_sat, 50)), 1),
        "saturated_p99_ms": round(float(np.percentile(lat_sat, 99)), 1),
        "max_rel_err_vs_fp64": max(errs), "all_converged": conv, "poisson_equals_saturated": same,
        "iterations_served": iters_srv, "iterations_fp64": iters64, "prepare_s": round(t_prep, 2),
        "chunk": args.chunk, "slots": args.slots, "precision_policy": args.precision,
        "hip_graph": bool(args.served_graph),
        "switched_to_fp64": srv.set_precision(-1),
        "batched_spmv_autotune": dict(zip(("system_types", "tiled", "split"), srv.product_forms())),
        "served_s": round(t_sat, 3), "fp64_s": round(t_base, 3),
    }


def run_c5(args):
    out = served_stream(args)
    line = {"metric": "served solves/s at 1e-8 rel err vs FP64 (batched mixed vs one-at-a-time FP64)",
            "value": out["solves_per_s"], "unit": "solves/s", "n_gpus": 1, "higher_is_better": True,
            "dtype": "fp32-iter+fp64-refresh", "data": "synthetic (seeded request stream)",
            "config": {"workload": out["workload"]}}
    line.update({k: v for k, v in out.items() if k != "workload"})
    print(json.dumps(line))


def one_step(ff, p, meta, rhs, opts):
    p.assemble(meta["q"], reuse_coefficients=True)
    if rhs is None:
        p.load(meta["q"], ff.capi.SOURCE_SIN_PRODUCT, [1.0, np.pi, np.pi, np.pi])
    else:
        p.set_rhs(rhs)
    return p.solve(opts, download=False)


def host_description():
    """nproc and the CPU model of the host the baseline runs on."""
    model = ""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        usable = len(os.sched_getaffinity(0))
    except AttributeError:
        usable = os.cpu_count() or 1
    return {"nproc": usable, "cpu_count": os.cpu_count(), "model": model}


def cpu_baseline(ff, p, meta, rhs, x_gpu, threads: int, reps: int, one_core: bool = True):
    """The oracle's restatement of the reference algorithm (C, OpenMP) timed
    on the full configuration on `threads` host threads: one step = the
    triplet assembly (the reference's quadrature-point element loop) +
    csc_from_triplet + the load (P2 workload) + the constrained reduction +
    Jacobi-PCG to a relative residual of 1e-10, the same work as one GPU
    step.  The dof map and the boundary flags are set-up on both sides (taken
    from the GPU problem, whose dof map and boundary equal the oracle's bit
    for bit).  One warm-up, then `reps` timed steps: median and p99.  The
    FP64 solution of the last step gives the GPU mixed solution's error
    against the oracle.  With one_core, one more step on a single thread
    (the reference's own loops are serial), timed once after the others."""
    import oracle as O
    from fortfem_amd import workloads

    N, d, q = meta["N"], meta["degree"], meta["q"]
    v, t = ff.unit_box_mesh(N)
    kappa = workloads.lognormal_kappa(t.shape[0], meta["config"])
    g = p.global_dofs(base=1)
    cons = p.constrained().astype(np.int32)
    G = int(g.max())
    sin_f = lambda x, y, z: np.sin(np.pi * x) * np.sin(np.pi * y) * np.sin(np.pi * z)

    def step(threads=threads):
        a = O.par_lagrange_stiffness_csc(v, t, d, q, g, G, threads, cell_kappa=kappa)
        b = O.lagrange_load(v, t, d, q, sin_f, g, G) if rhs is None else np.zeros(G)
        red, free, rb = O.constrained_reduction(a, cons, np.zeros(G), b)
        del a
        if rhs is not None:
            rb = rhs
        x, st = O.par_solve_sparse(red, rb, threads, tolerance=1e-10, max_iterations=100000,
                                   breakdown="relative")
        return x, st, free

    times = []
    for k in range(reps + 1):
        t0 = time.perf_counter()
        x, st, free = step()
        if k > 0:
            times.append(time.perf_counter() - t0)
        print(json.dumps({"progress": "cpu baseline step %d of %d: %.2f s" % (k, reps, time.perf_counter() - t0)}),
              file=sys.stderr, flush=True)
    med = float(np.median(times))
    err = float(np.linalg.norm(x_gpu[free - 1] - x) / np.linalg.norm(x))
    host = host_description()
    single = None
    if one_core and threads > 1:
        t0 = time.perf_counter()
        step(1)
        t1 = time.perf_counter() - t0
        print(json.dumps({"progress": "cpu baseline, one thread: %.2f s" % t1}), file=sys.stderr, flush=True)
        single = {"value": round(1.0 / t1, 5), "unit": "solves/s", "cores": 1, "s_per_step": round(t1, 3),
                  "sample": "one timed step of the same full configuration on one thread, after the "
                            f"{threads}-thread steps"}
    return {
        "value": round(1.0 / med, 4),
        "unit": "solves/s",
        "cores": threads,
        "kind": "port",
        "sample": (f"the full configuration, no scaling: oracle (C restatement of the reference's triplet "
                   f"assembly, csc_from_triplet, constrained reduction and Jacobi-PCG) on {threads} OpenMP "
                   f"threads, 1 warm-up + {reps} timed steps; dof map and boundary flags are set-up"),
        "median_s": round(med, 3),
        "p99_s": round(float(np.percentile(times, 99)), 3),
        "times_s": [round(x_, 3) for x_ in times],
        "iterations": st.iterations,
        "host": host,
        "compiler": "gcc -O2 -ffp-contract=off -fopenmp",
        "gpu_mixed_rel_err_vs_oracle_fp64": err,
        "one_core": single,
    }


def run_served_one(args):
    """One request of the seeded served stream (--request-index) alone on the
    batched server, with the precision policy --precision, for --steps server
    steps of --chunk iterations: a fixed kernel sequence for counter runs."""
    import fortfem_amd as ff
    from fortfem_amd import workloads

    r = workloads.served_stream(args.request_index + 1)[args.request_index]
    v, t = ff.unit_box_mesh(r["N"])
    srv = ff.Server(budget_bytes=int(args.budget_gb * (1 << 30)), max_systems=args.slots, chunk=args.chunk)
    srv.set_precision(args.precision)
    n, nnz = srv.prepare(0, r["family"], r["order"], v, t)
    coeff, rhs = workloads.request_data(r, t.shape[0], n)
    srv.submit(0, r["family"], r["order"], coeff, rhs, mass=0.0 if r["family"] == 1 else 1.0,
               tol=1e-10, max_iterations=args.max_iter)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        srv.step()
    dt = time.perf_counter() - t0
    print(json.dumps({"workload": "served1", "request": args.request_index, "family": r["family"],
                      "order": r["order"], "N": r["N"], "nrhs": r["nrhs"], "n": n, "nnz": nnz,
                      "precision_policy": args.precision, "steps": args.steps, "chunk": args.chunk,
                      "s": round(dt, 4)}))


def run_copy(args):
    """Calibration for the HBM-traffic counters: a device copy of known size
    (1 GiB of FP32 read, 1 GiB written, through torch's elementwise kernel),
    `steps` times.  FETCH_SIZE and WRITE_SIZE of these dispatches against
    2^30 bytes give the scale factors applied to the counters of the
    roofline kernel."""
    import torch

    n = 1 << 28
    x = torch.rand(n, device="cuda:0")
    y = torch.empty_like(x)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        torch.mul(x, 1.0, out=y)
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / args.steps
    assert torch.equal(x, y)
    print(json.dumps({"workload": "copy", "bytes_read": 4 * n, "bytes_written": 4 * n,
                      "launches": args.steps, "ms_per_copy": round(dt * 1e3, 3),
                      "gbs": round(8.0 * n / dt / 1e9, 1)}))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="c4",
                    help="c4 (default), c2, c3, c5, copy (traffic-counter calibration) or "
                         "served1 (one served request alone, for counter runs)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--no-served", action="store_true", help="skip the served-stream keys")
    ap.add_argument("--no-traffic", action="store_true",
                    help="skip the rocprofv3 --pmc passes that measure roofline.traffic")
    ap.add_argument("--cpu-threads", type=int, default=0, help="CPU baseline threads (0: min(16, nproc))")
    ap.add_argument("--cpu-reps", type=int, default=5, help="CPU baseline timed steps (after 1 warm-up)")
    ap.add_argument("--cpu-one-core", type=int, default=1,
                    help="1 (default): also time one CPU baseline step on a single thread")
    ap.add_argument("--graph", action="store_true", help="replay iteration chunks from HIP graphs")
    ap.add_argument("--point-tensors", action="store_true",
                    help="c3: also time the assembly with the tensor evaluated at every rule point (MFMA)")
    ap.add_argument("--orders", type=int, nargs="+", default=[1, 2, 3, 4, 5, 6],
                    help="Nedelec orders of the c3 workload")
    ap.add_argument("--requests", type=int, default=16, help="served: the first REQUESTS of the stream")
    ap.add_argument("--served-graph", type=int, default=1,
                    help="served: 1 (default) replays each step's chunk from a HIP graph, 0 launches one by one")
    ap.add_argument("--request-index", type=int, default=1,
                    help="served1: the request of the stream served alone")
    ap.add_argument("--stream", type=int, default=4096, help="served: length of the seeded stream")
    ap.add_argument("--arrival-factor", type=float, default=0.8,
                    help="served: Poisson arrival rate as a fraction of the saturated request rate")
    ap.add_argument("--slots", type=int, default=64, help="served: batch slots")
    ap.add_argument("--chunk", type=int, default=48, help="served: iterations per server step")
    ap.add_argument("--check-interval", type=int, default=48,
                    help="c2/c4: iterations per chunk (the mixed solve's FP64 refresh interval)")
    ap.add_argument("--budget-gb", type=float, default=64.0, help="served: device memory for admitted requests")
    ap.add_argument("--max-iter", type=int, default=200000, help="served: iteration cap per right-hand side")
    ap.add_argument("--precision", type=int, default=0,
                    help="served: 0 FP32 iterations (FP64 on stagnation), 1 FP64, 2 FP64 for Nedelec")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1 "
                                   "(multi-GPU sharding is a design note, not implemented)"}))
        sys.exit(2)

    if args.workload == "c3":
        run_c3(args)
        return
    if args.workload == "c5":
        run_c5(args)
        return
    if args.workload == "copy":
        run_copy(args)
        return
    if args.workload == "served1":
        run_served_one(args)
        return

    import fortfem_amd as ff

    t_setup = time.perf_counter()
    p, v, t, kappa, rhs, meta = build_problem(ff, args.workload)
    t_setup = time.perf_counter() - t_setup
    sizes = p.sizes()
    base = dict(method="pcg", preconditioner="jacobi", tolerance=1e-10, max_iterations=100000,
                breakdown="relative", check_interval=args.check_interval, use_graph=args.graph)
    o64 = ff.solver_options(**base)
    omx = ff.solver_options(**base, precision="mixed")

    # FP64 reference solution on the GPU and its timing
    x64, st64 = p.solve(o64)
    for _ in range(args.warmup):
        one_step(ff, p, meta, rhs, o64)
    lat64 = []
    for _ in range(args.steps):
        t0 = time.perf_counter()
        one_step(ff, p, meta, rhs, o64)
        lat64.append(time.perf_counter() - t0)

    # mixed path (headline)
    xm, stm = p.solve(omx)
    err = float(np.linalg.norm(xm - x64) / np.linalg.norm(x64))
    for _ in range(args.warmup):
        one_step(ff, p, meta, rhs, omx)
    p.ctx.set_profiling(True)
    lat = []
    prof = []
    t_all = time.perf_counter()
    for _ in range(args.steps):
        t0 = time.perf_counter()
        _, st = one_step(ff, p, meta, rhs, omx)
        lat.append(time.perf_counter() - t0)
        prof.append(p.ctx.last_solve_profile(1))
    t_all = time.perf_counter() - t_all
    p.ctx.set_profiling(False)

    n_free, nnz_free = sizes["n_free"], sizes["nnz_free"]
    cnt = sum(c for c, _ in prof)
    avg_ms = sum(c * m for c, m in prof) / max(cnt, 1)
    red = p.matrix(ff.capi.REDUCED)
    if ff.pcg_launches_per_iteration(n_free, "fp32") == 3:
        # the SpMV launch of the three-launch iteration: algorithmic bytes of
        # an FP32 SpMV over the int32 CSR (8 nnz + 12 n + 4), and beside them
        # the bytes of the tuned variant's own format (p gathered, A p written)
        kname = "k_pcg_apdot<float> (FP32 SpMV A p + p.Ap partials)"
        alg32 = spmv_bytes(nnz_free, n_free, 4)
        fmt32 = red.spmv_bytes("fp32")
    else:
        kname = "k_pcg_dirspmv<float> (FP32 direction update + SpMV + p.Ap)"
        alg32 = fmt32 = dirspmv_bytes(nnz_free, n_free, 4)
    rate = lambda b: b / (avg_ms * 1e-3) / 1e9 if avg_ms > 0 else 0.0
    achieved, achieved_fmt = rate(alg32), rate(fmt32)
    ms64, by64 = red.bench_spmv("fp64", 50)
    ms32, by32 = red.bench_spmv("fp32", 50)
    csr64, csr32 = spmv_bytes(nnz_free, n_free, 8), spmv_bytes(nnz_free, n_free, 4)

    value = args.steps / t_all
    lat_ms = np.array(lat) * 1e3
    out = {
        "metric": "solves/s & p99 latency at 1e-8 rel err (mixed vs FP64); SpMV HBM GB/s",
        "value": round(value, 3),
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(t_all / args.steps * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32-iter+fp64-refresh",
        "data": "synthetic (seeded lognormal kappa, closed-form load)",
        "config": {"workload": f"{args.workload}: P{meta['degree']} tet Poisson {meta['N']}^3x6 "
                               f"({meta['ncell']} tets), n_free {n_free}, nnz_free {nnz_free}, "
                               "Jacobi-PCG rel 1e-10, assembly+load+reduction+solve per step"},
        "p50_ms": round(float(np.percentile(lat_ms, 50)), 3),
        "p99_ms": round(float(np.percentile(lat_ms, 99)), 3),
        "rel_err_vs_fp64": err,
        "mixed_iterations": st.iterations,
        "mixed_fp64_refreshes": st.refinements,
        "fp64_solves_per_s": round(len(lat64) / sum(lat64), 3),
        "fp64_iterations": st64.iterations,
        "fp64_p99_ms": round(float(np.percentile(np.array(lat64) * 1e3, 99)), 3),
        "spmv_fp64_gbs": round(by64 / (ms64 * 1e-3) / 1e9, 1),
        "spmv_fp32_gbs": round(by32 / (ms32 * 1e-3) / 1e9, 1),
        "spmv_bytes_basis": "bytes of the tuned variant's own storage format (SELL padding, 16-bit "
                            "tile positions and tile column lists included); the csr_equivalent keys "
                            "divide the int32 CSR count nnz (s + 4) + 4 (n + 1) + 2 n s instead",
        "spmv_fp64_csr_equivalent_gbs": round(csr64 / (ms64 * 1e-3) / 1e9, 1),
        "spmv_fp32_csr_equivalent_gbs": round(csr32 / (ms32 * 1e-3) / 1e9, 1),
        "spmv_variants": {"fp64": red.spmv_variant("fp64"), "fp32": red.spmv_variant("fp32")},
        "pcg_launches_per_iteration": {"fp64": ff.pcg_launches_per_iteration(n_free, "fp64"),
                                       "fp32": ff.pcg_launches_per_iteration(n_free, "fp32")},
        "setup_s": round(t_setup, 2),
        "roofline": {
            "kernel": kname,
            "bound": "hbm",
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": None,
            "traffic_source": "not measured (--no-traffic)",
            "avg_launch_ms": avg_ms,
            "launches": cnt,
            "bytes_per_launch": alg32,
            "bytes_basis": "algorithmic: FP32 SpMV over the int32 CSR, 8 nnz + 12 n + 4 bytes "
                           "(values, column indices, row pointers, p read once, A p written once)",
            "format_bytes_per_launch": fmt32,
            "achieved_format": round(achieved_fmt, 1),
            "frac_format": round(achieved_fmt / HBM_PEAK_GBS, 4),
            "format_basis": "the bytes the tuned variant reads and writes in its own storage "
                            "format (padded SELL-64 entries with 16-bit tile positions, slice "
                            "offsets, row map, tile offsets and column lists, p, A p)",
        },
    }
    if not args.no_traffic:
        print(json.dumps({"progress": "traffic counters"}), file=sys.stderr, flush=True)
        tr, src = measure_traffic(args.workload)
        out["roofline"]["traffic"] = tr
        out["roofline"]["traffic_source"] = src
    if not args.no_served:
        print(json.dumps({"progress": "served stream"}), file=sys.stderr, flush=True)
        out["served"] = served_stream(args)
    if not args.no_cpu:
        threads = args.cpu_threads or min(16, host_description()["nproc"])
        out["cpu_baseline"] = cpu_baseline(ff, p, meta, rhs, xm, threads, args.cpu_reps,
                                           one_core=bool(args.cpu_one_core))
    print(json.dumps(out))


if __name__ == "__main__":
    main()
