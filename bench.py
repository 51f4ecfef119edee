"""Benchmark: time to 1e-8 relative residual of the FP32-preconditioned solve
against the FP64-only solve on one MI355X, plus the Laplacian SpMV's
achieved HBM bandwidth.

Workload (default): R-MAT scale 23, edge factor 16, (a,b,c,d) =
(.57,.19,.19,.05), symmetrised, self loops and duplicates removed, largest
connected component, pendant paths of length uniform{1,2,3} at 10% of the
vertices, edge weights lognormal(0, 1.5); one mean-zero N(0,1) right-hand
side (seed 0).  A step is one full solve to 1e-8 relative residual from a
device-resident right-hand side.

Prints one JSON line (rank 0).  Usage:
    python bench.py [--gpus 1] [--steps K] [--warmup W] [--scale 23]
"""
import argparse
import ctypes as C
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "time to 1e-8 rel. residual: FP32-precond vs FP64 on GPU; SpMV HBM GB/s"
HBM_PEAK_GBS = 8000.0


def traffic_of(kernel):
    """HBM bytes per launch of a kernel, from the committed counter summary
    (FETCH_SIZE x read calibration + WRITE_SIZE x write calibration, both
    calibrated on a copy kernel of known size), and the source commit that
    summary names; (None, None) without it."""
    path = os.path.join(ROOT, "profiles", "hbm_traffic_rmat23.txt")
    value, commit = None, None
    try:
        with open(path) as f:
            for line in f:
                if line.startswith("#") and "source commit" in line and commit is None:
                    commit = line.split("source commit", 1)[1].split()[0].strip(",")
                parts = line.split()
                if len(parts) >= 2 and parts[0] == kernel + ":":
                    value = float(parts[1])
    except OSError:
        pass
    return value, commit


def cpu_full_workload():
    """The CPU oracle on the full workload, measured once on the build
    machine (the run takes minutes): profiles/cpu_baseline_rmat23.txt."""
    path = os.path.join(ROOT, "profiles", "cpu_baseline_rmat23.txt")
    out = {}
    try:
        with open(path) as f:
            for line in f:
                parts = line.split()
                if len(parts) >= 2 and parts[0].endswith(":") and not line.startswith("#"):
                    try:
                        out[parts[0][:-1]] = float(parts[1])
                    except ValueError:
                        out[parts[0][:-1]] = " ".join(parts[1:])
    except OSError:
        return None
    return out or None


def cpu_baseline(scale, threads=None):
    """The CPU oracle (the project's C++ restatement, not the reference
    binary, which cannot be built without MPI/PETSc/CombBLAS) timed here on a
    bounded sample of the same construction: setup and solve, on one thread
    and on this process's thread share (OMP_NUM_THREADS, else all cores).
    Nothing is extrapolated; the full-size figures measured once on the
    build machine are reported beside it."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_ffi
    from ligmg_amd import graphs
    g = graphs.rmat_config(scale, "lognormal", pendants=True)
    b = graphs.rhs(g.n, 1, 0)
    nth = int(os.environ.get("OMP_NUM_THREADS") or os.cpu_count() or 1)
    legs = {}
    for th in (threads or (1, nth)):
        th = oracle_ffi.lib().oracle_set_threads(th)
        t = time.perf_counter()
        s = oracle_ffi.OracleSolver(g)
        ts = time.perf_counter() - t
        t = time.perf_counter()
        _, info, _ = s.solve(b)
        tv = time.perf_counter() - t
        legs[th] = {"setup_s": ts, "solve_s": tv, "iters": info.iters}
        del s
        print(f"# cpu oracle R-MAT {scale} threads {th}: setup {ts:.2f}s solve {tv:.2f}s iters {info.iters}",
              file=sys.stderr, flush=True)
    top = legs[max(legs)]
    return {
        "value": top["solve_s"],
        "unit": "s",
        "cores": max(legs),
        "kind": "port",
        "sample": (f"CPU oracle FP64 V-cycle/CG to 1e-8 on R-MAT {scale} of the same construction (n={g.n}, "
                   f"nnz={g.nnz}), measured in this run: "
                   + ", ".join(f"solve {v['solve_s']:.2f} s and setup {v['setup_s']:.2f} s on {k} thread(s)"
                               for k, v in sorted(legs.items())) + f", {top['iters']} iterations"),
        "legs": {str(k): v for k, v in legs.items()},
        "full_workload": cpu_full_workload(),
    }


def _line(metric, value, unit, hib, ms, args, dtype, config, **extra):
    out = {"metric": metric, "value": value, "unit": unit, "n_gpus": 1, "steps": args.steps,
           "warmup": args.warmup, "ms_per_step": ms, "higher_is_better": hib, "scaling": "weak",
           "vs_baseline": None, "dtype": dtype, "data": "synthetic", "config": config}
    out.update(extra)
    print(json.dumps(out), flush=True)


def _demean(X, comp_ptr):
    """X (n x k, host) with its mean removed on every connected component
    (comp_ptr: component offsets; None: one component)."""
    import numpy as np
    X = X.copy()
    bounds = [0, X.shape[0]] if comp_ptr is None else list(comp_ptr)
    for a, b in zip(bounds[:-1], bounds[1:]):
        X[a:b] -= X[a:b].mean(axis=0)
    return X


def _fp64_proxy_err(L, B, X, solve, comp_ptr=None):
    """Error of the solutions X (n x k) against an FP64 proxy for the
    sparse-direct answer: X plus one iterative-refinement step,
    X_p = X + L^+ (B - L X), the residual formed in FP64 on the host and
    means removed per component (a direct factorisation of these graphs does
    not fit).  Returns the largest relative error over the columns and the
    proxy's largest full relative residual."""
    import numpy as np
    Bh, Xh = np.asarray(B, dtype=np.float64), np.asarray(X, dtype=np.float64)
    R = _demean(Bh - L @ Xh, comp_ptr)
    Xp = _demean(Xh + solve(R), comp_ptr)
    Xd = _demean(Xh, comp_ptr)
    err = max(float(np.linalg.norm(Xd[:, c] - Xp[:, c]) / np.linalg.norm(Xp[:, c])) for c in range(Xh.shape[1]))
    res = max(float(np.linalg.norm(Bh[:, c] - L @ Xp[:, c]) / np.linalg.norm(Bh[:, c])) for c in range(Xh.shape[1]))
    return err, res


def bench_batched(args):
    """Per-right-hand-side time of batched solves: 48 independent
    Barabasi-Albert graphs (m=5, n uniform in [20k, 400k], seeds 0..47) with
    16 mean-zero N(0,1) right-hand sides each, and a 3-D 7-point grid 160^3
    (weights uniform[0.1, 10]) with 64; FP32 preconditioner inside FP64
    flexible CG, graph-captured cycles.  A step solves both batches."""
    import numpy as np
    import torch
    import ligmg_amd
    from ligmg_amd import _ffi, graphs
    ns = np.random.default_rng(0).integers(20000, 400001, 48)
    u = graphs.union([graphs.barabasi_albert(int(n), 5, seed=i) for i, n in enumerate(ns)])
    g3 = graphs.grid3d(160, 160, 160, wkind=_ffi.W_UNIFORM, p0=0.1, p1=10.0, seed=0)
    opts = ligmg_amd.Options(precision=_ffi.PREC_MIXED, outer_ksp_type=_ffi.KSP_FCG,
                             spmm_tune=0 if args.no_tune else 1)
    dev = torch.device("cuda:0")
    cases = []
    for name, g, k, cp in (("ba48", u, 16, u.comp_ptr), ("grid160", g3, 64, None)):
        B = torch.from_numpy(np.ascontiguousarray(
            graphs.rhs(g.n, k, 0, comp_ptr=cp).reshape(g.n, k, order="F"))).to(dev)
        t = time.perf_counter()
        s = ligmg_amd.Solver(g, opts, comp_ptr=cp)
        setup = time.perf_counter() - t
        cases.append((name, g, k, s, B, setup))
        print(f"# {name}: n={g.n} nnz={g.nnz} k={k} setup {setup:.2f}s", file=sys.stderr, flush=True)
    for _ in range(args.warmup):
        for name, g, k, s, B, _s in cases:
            s.solve_batch(B)
    torch.cuda.synchronize()
    per = {name: [] for name, *_ in cases}
    its, sols = {}, {}
    t0 = time.perf_counter()
    for _ in range(args.steps):
        for name, g, k, s, B, _s in cases:
            X, info, it = s.solve_batch(B)
            assert info.status == _ffi.CONVERGED and info.full_rel_residual <= 1e-7
            per[name].append(info.time)
            its[name] = (int(it.min()), int(it.max()))
            sols[name] = (X, info.full_rel_residual)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    nrhs = sum(k for _, _, k, *_ in cases)
    ms_step = 1e3 * wall / args.steps
    detail = {name: {"n": g.n, "nnz": g.nnz, "rhs": k, "setup_s": st, "batch_solve_s": min(per[name]),
                     "per_rhs_ms": 1e3 * min(per[name]) / k, "iters_min_max": its[name],
                     "full_rel_residual_max": sols[name][1]}
              for name, g, k, s, B, st in cases}
    # accuracy, outside the timed region: against the FP64 proxy
    for name, g, k, s, B, st in cases:
        cp = g.comp_ptr if name == "ba48" else None
        err, res = _fp64_proxy_err(g.to_scipy(), B.cpu().numpy(), sols[name][0].cpu().numpy(),
                                   lambda R: s.solve_batch(torch.from_numpy(np.ascontiguousarray(R)).to(dev))[0]
                                   .cpu().numpy(), cp)
        detail[name]["err_vs_fp64_proxy"] = {"max_over_columns": err, "proxy_full_rel_residual_max": res}
        print(f"# {name}: error against the FP64 proxy {err:.2e} (proxy residual {res:.2e})", file=sys.stderr,
              flush=True)
    # base-level SpMM (the batched solves' dominant kernel shape) and the
    # coarse dense apply on the matrix cores, timed with HIP events inside
    # the library on the solver's stream
    lib = _ffi.solver_lib()
    ms, nb, fl = C.c_double(), C.c_int64(), C.c_double()
    for name, g, k, s, B, st in cases:
        base = 1 if s.level_info(0)[0] == _ffi.LEVEL_ELIM else 0
        d = detail[name]
        for prec, lab in ((_ffi.PREC_MIXED, "fp32"), (_ffi.PREC_FP64, "fp64")):
            assert lib.ligmg_bench_spmm(s._h, base, prec, k, 10, C.byref(ms), C.byref(nb)) == 0
            d[f"spmm_{lab}"] = {"ms": ms.value, "bytes": nb.value, "gbs": nb.value / ms.value / 1e6,
                                "frac_of_8tbs": nb.value / ms.value / 1e6 / HBM_PEAK_GBS}
        assert lib.ligmg_bench_coarse(s._h, k, 20, C.byref(ms), C.byref(fl)) == 0
        nc = s.level_info(s.num_levels() - 1)[1]
        d["coarse_dense"] = {"n_c": nc, "kp": 1 << (k - 1).bit_length(), "ms": ms.value, "flop": fl.value,
                             "tflops": fl.value / ms.value / 1e9, "frac_of_fp64_mfma_78.6tf":
                             fl.value / ms.value / 1e9 / 78.6}
    for name, d in detail.items():
        print(f"# {name}: batch {d['batch_solve_s']:.3f}s, {d['per_rhs_ms']:.2f} ms per right-hand side, "
              f"iterations {d['iters_min_max']}; base SpMM fp32 {d['spmm_fp32']['ms']:.3f} ms "
              f"{d['spmm_fp32']['gbs']:.0f} GB/s, fp64 {d['spmm_fp64']['ms']:.3f} ms {d['spmm_fp64']['gbs']:.0f} GB/s; "
              f"coarse n_c={d['coarse_dense']['n_c']} {d['coarse_dense']['ms'] * 1e3:.1f} us "
              f"{d['coarse_dense']['tflops']:.3f} TFLOP/s", file=sys.stderr, flush=True)
    _line("batched solve time per right-hand side to 1e-8 (FP32 preconditioner, FP64 FCG)",
          ms_step / nrhs, "ms", False, ms_step, args, "f64 outer / f32 preconditioner",
          {"workload": "ba48x16rhs_plus_grid160^3x64rhs", "rhs": nrhs, "spmm_tune": 0 if args.no_tune else 1},
          batches=detail)


def bench_kcycle(args):
    """The plain method on an irregular graph: R-MAT scale 22 (edge factor
    16, a/b/c/d = 0.57/0.19/0.19/0.05, symmetrised, largest component),
    weights uniform[0.5, 1.5], one mean-zero N(0,1) right-hand side, FP64
    K-cycle (FCG(2) on every non-exact coarse level) inside FP64 FCG to 1e-8.
    A step is one solve; the fused coarse tail is on (its default)."""
    import numpy as np
    import torch
    import ligmg_amd
    from ligmg_amd import _ffi, graphs
    g = graphs.rmat_config(22, "uniform")
    dev = torch.device("cuda:0")
    b = torch.from_numpy(graphs.rhs(g.n, 1, 0)).to(dev)
    A = (torch.from_numpy(g.rowptr).to(dev), torch.from_numpy(g.colind).to(dev), torch.from_numpy(g.vals).to(dev))
    s = ligmg_amd.Solver(A, ligmg_amd.Options(kcycle=1, outer_ksp_type=_ffi.KSP_FCG))
    print(f"# kcycle: n={g.n} nnz={g.nnz} levels {s.num_levels()} fused tail from level {s.tail_start()} "
          f"setup {s.setup_time:.2f}s", file=sys.stderr, flush=True)
    for _ in range(args.warmup):
        s.solve(b)
    torch.cuda.synchronize()
    times = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        info = s.solve(b)
        assert info.status == _ffi.CONVERGED and info.rel_residual() <= 1e-8
        times.append(info.time())
    torch.cuda.synchronize()
    ms_step = 1e3 * (time.perf_counter() - t0) / args.steps
    # accuracy, outside the timed region: against the FP64 proxy
    err, res = _fp64_proxy_err(g.to_scipy(), b.cpu().numpy()[:, None], info.solution.cpu().numpy()[:, None],
                               lambda R: s.solve(torch.from_numpy(np.ascontiguousarray(R[:, 0])).to(dev))
                               .solution.cpu().numpy()[:, None])
    v = ligmg_amd.Solver(A, ligmg_amd.Options())
    vt = min(v.solve(b).time() for _ in range(3))
    print(f"# kcycle solve {min(times):.4f}s, {info.iters()} iterations; V-cycle/CG on the same graph {vt:.4f}s; "
          f"error against the FP64 proxy {err:.2e} (proxy residual {res:.2e})", file=sys.stderr, flush=True)
    _line("time to 1e-8 rel. residual, FP64 K-cycle inside FCG", ms_step / 1e3, "s", False, ms_step, args,
          "f64", {"workload": "rmat22_lcc_uniform0.5-1.5_kcycle_fcg", "n": g.n, "nnz": g.nnz},
          iters=info.iters(), setup_s=s.setup_time, vcycle_cg_s=vt, full_rel_residual=info.full_rel_residual,
          err_vs_fp64_proxy={"error": err, "proxy_full_rel_residual": res,
                             "proxy": "K-cycle solution + one iterative-refinement step (residual b - L x in "
                                      "FP64 on the host), mean removed"})


def bench_coarse(args):
    """The dense pseudo-inverse of a large coarsest level: a 3-D 7-point grid
    (weights uniform[0.5, 1.5]) whose hierarchy is capped at two levels
    (max_levels = 2), so the coarse level holds several thousand vertices.
    A step re-factors it as setup does (blocked Gauss-Jordan on the matrix
    cores) and applies it to a block of 64 right-hand sides."""
    import ligmg_amd
    from ligmg_amd import _ffi, graphs
    side = 40
    g = graphs.grid3d(side, side, side, wkind=_ffi.W_UNIFORM, p0=0.5, p1=1.5, seed=0)
    s = ligmg_amd.Solver(g, ligmg_amd.Options(max_levels=2))
    nc = s.level_info(s.num_levels() - 1)[1]
    lib = _ffi.solver_lib()
    tot, gj, fl = C.c_double(), C.c_double(), C.c_double()
    for _ in range(args.warmup):
        assert lib.ligmg_bench_pinv(s._h, C.byref(tot), C.byref(gj), C.byref(fl)) == 0
    tots, gjs = [], []
    for _ in range(args.steps):
        assert lib.ligmg_bench_pinv(s._h, C.byref(tot), C.byref(gj), C.byref(fl)) == 0
        tots.append(tot.value)
        gjs.append(gj.value)
    ms, fl64 = C.c_double(), C.c_double()
    assert lib.ligmg_bench_coarse(s._h, 64, 20, C.byref(ms), C.byref(fl64)) == 0
    tf = fl.value / (min(gjs) * 1e-3) / 1e12
    print(f"# coarse n_c={nc}: inversion {min(gjs):.2f} ms on the device ({tf:.2f} TFLOP/s FP64), "
          f"{min(tots):.1f} ms with host assembly and upload; apply to 64 columns {ms.value * 1e3:.1f} us "
          f"({fl64.value / (ms.value * 1e-3) / 1e12:.2f} TFLOP/s)", file=sys.stderr, flush=True)
    _line("dense coarse pseudo-inverse: device inversion time", min(gjs) / 1e3, "s", False, min(tots), args, "f64",
          {"workload": f"grid3d_{side}^3_uniform0.5-1.5_max_levels2_coarse_pinv", "n_coarse": nc},
          gj_tflops=tf, total_ms=min(tots),
          roofline={"bound": "mfma", "achieved": tf, "peak": 78.6, "unit": "TFLOP/s", "frac": tf / 78.6,
                    "traffic": None},
          apply64_us=ms.value * 1e3, apply64_tflops=fl64.value / (ms.value * 1e-3) / 1e12)


def bench_grad(args):
    """Forward + backward through x = L(w)^+ b on a random geometric graph:
    1,000,000 vertices in the unit square, radius for mean degree 12, LCC,
    w = exp(-d^2/r^2) in FP32, b = +1/-1 at 256 source/sink pairs, loss =
    sum over 10,000 random vertex pairs of (x_i - x_j)^2.  A step is one
    forward (assembly, setup, solve) and one backward (adjoint solve,
    edge-gradient kernel)."""
    import math
    import numpy as np
    import torch
    from ligmg_amd import graphs
    from ligmg_amd.autograd import EdgePattern, laplacian_solve
    n0 = 1000000
    r = math.sqrt(12 / (math.pi * n0))
    g = graphs.rgg(n0, r, seed=0)
    ei, ej, w = g.upper_edges()
    rng = np.random.default_rng(0)
    npairs = 256
    perm = rng.permutation(g.n)
    b = np.zeros(g.n)
    b[perm[:npairs]] = 1.0
    b[perm[npairs:2 * npairs]] = -1.0
    pi, pj = rng.integers(0, g.n, 10000), rng.integers(0, g.n, 10000)
    dev = torch.device("cuda:0")
    pat = EdgePattern(g.n, ei, ej, dev)
    tpi, tpj = torch.tensor(pi, device=dev), torch.tensor(pj, device=dev)
    w32 = torch.tensor(w.astype(np.float32), device=dev)
    bt = torch.tensor(b, device=dev)

    def step():
        wt = w32.clone().requires_grad_(True)
        bb = bt.clone().requires_grad_(True)
        torch.cuda.synchronize()
        t = time.perf_counter()
        x = laplacian_solve(wt, bb, pat)
        loss = ((x[tpi] - x[tpj]) ** 2).sum()
        torch.cuda.synchronize()
        tf = time.perf_counter() - t
        loss.backward()
        loss = loss.detach()
        torch.cuda.synchronize()
        return tf, time.perf_counter() - t - tf, loss.item()

    for _ in range(args.warmup):
        step()
    fw, bw = [], []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        a, c, loss = step()
        fw.append(a)
        bw.append(c)
    wall = time.perf_counter() - t0
    ms_step = 1e3 * wall / args.steps
    # accuracy, outside the timed region: the forward solution against the
    # FP64 proxy of L(w)^+ b (the gradients are checked against central
    # differences and a dense pseudo-inverse in the tests)
    import scipy.sparse as sp
    with torch.no_grad():
        x = laplacian_solve(w32, bt, pat)
        W = sp.coo_matrix((w32.double().cpu().numpy(), (ei, ej)), shape=(g.n, g.n)).tocsr()
        W = (W + W.T).tocsr()
        L = (sp.diags(np.asarray(W.sum(axis=1)).ravel()) - W).tocsr()
        err, res = _fp64_proxy_err(L, (b - b.mean())[:, None], x.cpu().numpy()[:, None],
                                   lambda R: laplacian_solve(w32, torch.tensor(R[:, 0], device=dev), pat)
                                   .cpu().numpy()[:, None])
    print(f"# rgg n={g.n} edges={len(w)}: forward {min(fw):.3f}s backward {min(bw):.3f}s loss {loss:.6g}; "
          f"forward error against the FP64 proxy {err:.2e} (proxy residual {res:.2e})", file=sys.stderr, flush=True)
    _line("forward+backward time of x = L(w)^+ b (gradients in w and b)", ms_step / 1e3, "s", False, ms_step,
          args, "f64", {"workload": "rgg1M_deg12_gauss_weights_256pairs_10kpair_loss", "n": g.n,
                        "edges": int(len(w))},
          forward_s=min(fw), backward_s=min(bw),
          err_vs_fp64_proxy={"forward_error": err, "proxy_full_rel_residual": res,
                             "proxy": "forward solution + one iterative-refinement step (residual b - L x in "
                                      "FP64 on the host), mean removed"})


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--scale", type=int, default=23)
    ap.add_argument("--fp64-steps", type=int, default=2)
    ap.add_argument("--cpu-scale", type=int, default=19)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-tune", action="store_true", help="batched workload: the fixed SpMM tile")
    ap.add_argument("--cpu-only", action="store_true",
                    help="only the CPU oracle leg (no GPU), at --cpu-scale with --cpu-threads")
    ap.add_argument("--cpu-threads", default="",
                    help="comma-separated thread counts of the CPU leg (default: 1 and the thread share)")
    ap.add_argument("--ordering", choices=["degree", "natural"], default="natural",
                    help="vertex order of the hierarchy (default: the input's)")
    ap.add_argument("--spmv-only", action="store_true",
                    help="setup + SpMV timing + copy calibration only (for counter runs)")
    ap.add_argument("--workload", choices=["rmat", "batched", "grad", "kcycle", "coarse"], default="rmat",
                    help="rmat: the headline solve (default); batched: 48 BA graphs x 16 right-hand "
                         "sides and a 160^3 grid x 64; grad: gradients through a 1M-vertex geometric graph; "
                         "kcycle: the FP64 K-cycle on R-MAT 22; coarse: the dense pseudo-inverse of a "
                         "large coarse level")
    args = ap.parse_args()
    if args.cpu_only:
        th = [int(t) for t in args.cpu_threads.split(",") if t] or None
        print(json.dumps({"cpu_baseline": cpu_baseline(args.cpu_scale, th)}), flush=True)
        return
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print("bench.py: this solver runs on one GPU only (no sharding or collectives); "
              "use --gpus 1", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch
    import ligmg_amd
    from ligmg_amd import _ffi, graphs

    assert torch.cuda.is_available(), "bench.py needs a GPU"
    if args.workload == "batched":
        return bench_batched(args)
    if args.workload == "grad":
        return bench_grad(args)
    if args.workload == "kcycle":
        return bench_kcycle(args)
    if args.workload == "coarse":
        return bench_coarse(args)
    dev = torch.device("cuda:0")
    t = time.perf_counter()
    g = graphs.rmat_config(args.scale, "lognormal", pendants=True)
    t_gen = time.perf_counter() - t
    print(f"# graph n={g.n} nnz={g.nnz} generated in {t_gen:.1f}s", file=sys.stderr, flush=True)
    b = torch.from_numpy(graphs.rhs(g.n, 1, 0)).to(dev)
    A = (torch.from_numpy(g.rowptr).to(dev), torch.from_numpy(g.colind).to(dev),
         torch.from_numpy(g.vals).to(dev))
    torch.cuda.synchronize()

    # kernels of known traffic calibrate the HBM counters: a 2 GiB fill
    # (WRITE_SIZE) and a 2 GiB sum (FETCH_SIZE)
    cal = torch.ones(1 << 28, dtype=torch.float64, device=dev)
    float(cal.sum())
    torch.cuda.synchronize()
    # achievable HBM bandwidth beside the vendor peak: a 2 GiB device copy
    # (2 GiB read + 2 GiB written per copy), HIP events around 5 copies
    dst = torch.empty_like(cal)
    dst.copy_(cal)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(5):
        dst.copy_(cal)
    e1.record()
    torch.cuda.synchronize()
    copy_gbs = 5 * 2 * cal.numel() * 8 / (e0.elapsed_time(e1) / 1e3) / 1e9
    print(f"# 2 GiB device copy {copy_gbs:.0f} GB/s (read + write)", file=sys.stderr, flush=True)
    del cal, dst

    # FP64-only solve (the plain method, the drop-in): V-cycle + CG
    verbose = int(os.environ.get("LIGMG_VERBOSE", "0"))
    order = _ffi.ORDER_DEGREE if args.ordering == "degree" else _ffi.ORDER_NATURAL
    if args.spmv_only:
        # base-level kernels only (counter runs): plain SpMV and fused
        # Chebyshev step, FP64 and FP32, 20 launches each
        sm = ligmg_amd.Solver(A, ligmg_amd.Options(precision=_ffi.PREC_MIXED, verbose=verbose,
                                                    ordering=order))
        lib = _ffi.solver_lib()
        ms, nbytes = C.c_double(), C.c_int64()
        base = 1 if sm.level_info(0)[0] == _ffi.LEVEL_ELIM else 0
        for fn in ("ligmg_bench_spmv", "ligmg_bench_smoother"):
            for prec in (_ffi.PREC_FP64, _ffi.PREC_MIXED):
                rc = getattr(lib, fn)(sm._h, base, prec, 20, C.byref(ms), C.byref(nbytes))
                assert rc == 0
                print(f"# {fn} prec={prec} {ms.value:.4f} ms {nbytes.value} B "
                      f"{nbytes.value / ms.value / 1e6:.1f} GB/s", file=sys.stderr, flush=True)
        return
    s64 = ligmg_amd.Solver(A, ligmg_amd.Options(verbose=verbose, ordering=order))
    print(f"# fp64 setup {s64.setup_time:.2f}s levels {s64.num_levels()}", file=sys.stderr,
          flush=True)
    info64 = None
    t64 = []
    for _ in range(args.fp64_steps):
        info64 = s64.solve(b)
        t64.append(info64.time())
    print(f"# fp64 solve {min(t64):.4f}s iters {info64.iters()} rel {info64.rel_residual():.2e} "
          f"full {info64.full_rel_residual:.2e}", file=sys.stderr, flush=True)
    base = info64.base_level
    # SpMV roofline on the base-level FP64 operator (the outer-loop matrix)
    lib = _ffi.solver_lib()
    ms, nbytes = C.c_double(), C.c_int64()
    rc = lib.ligmg_bench_spmv(s64._h, base, _ffi.PREC_FP64, 20, C.byref(ms), C.byref(nbytes))
    assert rc == 0
    spmv64 = {"ms": ms.value, "bytes": nbytes.value, "gbs": nbytes.value / ms.value / 1e6}
    rc = lib.ligmg_bench_spmv(s64._h, base, _ffi.PREC_MIXED, 20, C.byref(ms), C.byref(nbytes))
    assert rc == 0
    spmv32 = {"ms": ms.value, "bytes": nbytes.value, "gbs": nbytes.value / ms.value / 1e6}
    rc = lib.ligmg_bench_smoother(s64._h, base, _ffi.PREC_FP64, 20, C.byref(ms), C.byref(nbytes))
    assert rc == 0
    cheb64 = {"ms": ms.value, "bytes": nbytes.value, "gbs": nbytes.value / ms.value / 1e6}
    lvl = s64.level_info(base)
    setup64 = s64.setup_time
    x64 = info64.solution
    # FP64 proxy for the sparse-direct answer (a direct factorisation of
    # this graph does not fit): the FP64 solution plus one step of iterative
    # refinement, x_p = x + L^+ (b - L x), with the residual formed in FP64
    # on the host; each solve cuts the residual by 1e-8
    L = g.to_scipy()
    bh = b.cpu().numpy()
    xh = x64.cpu().numpy()
    r = bh - L @ xh
    corr = s64.solve(torch.from_numpy(r - r.mean()).to(dev)).solution.cpu().numpy()
    xp = xh + corr
    xp -= xp.mean()
    proxy_res = float(np.linalg.norm(bh - L @ xp) / np.linalg.norm(bh))
    print(f"# fp64 proxy: full rel residual {proxy_res:.2e}", file=sys.stderr, flush=True)
    s64.close()
    del s64
    torch.cuda.empty_cache()

    # FP32 hierarchy + cycle inside the FP64 flexible-CG outer loop
    s32 = ligmg_amd.Solver(A, ligmg_amd.Options(precision=_ffi.PREC_MIXED, outer_ksp_type=_ffi.KSP_FCG,
                                                 verbose=verbose, ordering=order))
    print(f"# mixed setup {s32.setup_time:.2f}s", file=sys.stderr, flush=True)
    for _ in range(args.warmup):
        s32.solve(b)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    infos = []
    for _ in range(args.steps):
        infos.append(s32.solve(b))
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    t32 = [i.time() for i in infos]
    info32 = infos[-1]
    err = float(torch.linalg.norm((info32.solution - info32.solution.mean()) -
                                  (x64 - x64.mean())) / torch.linalg.norm(x64 - x64.mean()))
    x32h = info32.solution.cpu().numpy()
    pn = np.linalg.norm(xp)
    err_proxy = {"fp32_precond": float(np.linalg.norm((x32h - x32h.mean()) - xp) / pn),
                 "fp64": float(np.linalg.norm((xh - xh.mean()) - xp) / pn),
                 "proxy": "FP64 solve + one iterative-refinement step (residual b - L x in FP64 on the host), "
                          f"mean removed; its full relative residual {proxy_res:.2e}"}
    print(f"# mixed solve {min(t32):.4f}s iters {info32.iters()} rel {info32.rel_residual():.2e} "
          f"full {info32.full_rel_residual:.2e} diff_vs_fp64 {err:.2e}", file=sys.stderr,
          flush=True)
    assert info32.rel_residual() <= 1e-8 and info64.rel_residual() <= 1e-8
    # roofline of the dominant kernel of the timed solve: the FP32 fused
    # Chebyshev-step SpMV on the base level, HIP events on the solver's stream
    rc = lib.ligmg_bench_smoother(s32._h, base, _ffi.PREC_MIXED, 20, C.byref(ms), C.byref(nbytes))
    assert rc == 0
    cheb32 = {"ms": ms.value, "bytes": nbytes.value, "gbs": nbytes.value / ms.value / 1e6}
    print(f"# base-level smoother step fp32 {cheb32['ms']:.4f} ms {cheb32['gbs']:.0f} GB/s; "
          f"fp64 {cheb64['ms']:.4f} ms {cheb64['gbs']:.0f} GB/s; spmv fp64 {spmv64['ms']:.4f} ms "
          f"{spmv64['gbs']:.0f} GB/s, fp32 {spmv32['ms']:.4f} ms {spmv32['gbs']:.0f} GB/s",
          file=sys.stderr, flush=True)

    ms_step = 1e3 * wall / args.steps
    traffic = traffic_of("smoother_fp32")
    out = {
        "metric": METRIC,
        "value": ms_step / 1e3,
        "unit": "s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64 outer / f32 preconditioner",
        "data": "synthetic",
        "config": {
            "workload": f"rmat{args.scale}_lcc_pendants_lognormal1.5_fp32precond_fcg",
            "ordering": args.ordering,
            "n": g.n, "nnz": g.nnz, "rhs": 1, "rtol": 1e-8,
            "base_level_n": lvl[1], "base_level_nnz": lvl[2],
        },
        "fp64_solve_s": min(t64),
        "fp32_precond_solve_s": min(t32),
        "speedup_fp32_vs_fp64": min(t64) / min(t32),
        "iters_fp32": info32.iters(), "iters_fp64": info64.iters(),
        "rel_residual_fp32": info32.rel_residual(), "full_rel_residual_fp32": info32.full_rel_residual,
        "rel_residual_fp64": info64.rel_residual(), "full_rel_residual_fp64": info64.full_rel_residual,
        "solution_diff_fp32_vs_fp64": err,
        "err_vs_fp64_proxy": err_proxy,
        "setup_s_fp64": setup64, "setup_s_mixed": s32.setup_time,
        "graph_gen_s": t_gen,
        "spmv_fp64": spmv64, "spmv_fp32": spmv32, "smoother_fp64": cheb64, "smoother_fp32": cheb32,
        "roofline": {
            "bound": "hbm",
            "achieved": cheb32["gbs"],
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": cheb32["gbs"] / HBM_PEAK_GBS,
            "traffic": traffic[0],
            "traffic_kind": ("L2 fill bytes per launch (FETCH_SIZE x 2.0 + WRITE_SIZE x 1.0, calibrated on "
                             "kernels of known size; FETCH_SIZE also counts Infinity Cache hits), read from "
                             "profiles/hbm_traffic_rmat23.txt, not measured in this run"),
            "traffic_source_commit": traffic[1],
            "kernel": ("k_spmv_units<float, ValArr<float>, EpiDiag<float, EpiCheb<float>>, true> "
                       "(base-level FP32 Chebyshev step)"),
        },
        "hbm_copy_gbs": copy_gbs,
    }
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(args.cpu_scale)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
