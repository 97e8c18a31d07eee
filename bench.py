#!/usr/bin/env python3
"""CPR-GMRES solve phase benchmark on one MI355X.

Workload (one "step" = one full solve): the 60x220x85 reservoir system
(1,122,000 cells, 3x3 FP64 blocks, 7,780,000 stored blocks; top 35 layers
smooth lognormal, bottom 50 layers sinuous channels with k ratio 1e3, 5 well
columns), ABF-decoupled, CPR-GMRES(30) to a relative residual of 1e-6 from
x0 = 0, with the pressure-AMG hierarchy built on the host.  b and x live in
HBM when the timed region starts; host setup (generation, decoupling, AMG
setup, upload) is reported separately.

Prints one JSON line (rank 0).  Also measured live on the solver's stream:
per-kernel-class device time and algorithmic bytes (roofline of the
dominant kernel), the BSR SpMV bandwidth (3x3 here, 4x4 on the 100^3
system), the reference-exact configuration (one-thread relax 6, one-rank
stage 2), BILU(0) stage 2, and the decoupling kernels.  The CPU baseline is
the reference's own whole solve of the same system (its library built from
its sources, one thread and all usable cores), or the oracle's
restatement where that build is absent.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_HBM_GBS = 8000.0
METRIC = "CPR-GMRES time-to-rtol 1e-6 (s) & BSR SpMV GB/s; mixed+learned vs plain FP64"


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", default="C2")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-kind", choices=("reference", "port"), default="reference",
                    help="CPU baseline: the reference built from its sources (oracle/_ref/ref_driver, when built) "
                         "or the oracle's restatement")
    ap.add_argument("--spmv-reps", type=int, default=50)
    ap.add_argument("--no-learned", action="store_true", help="skip the mixed-precision + learned-smoother leg")
    ap.add_argument("--train-seeds", type=int, default=20,
                    help="small systems labelled for training the learned smoother (0: no training, no learned leg)")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="only run three plain device copies of MIB MiB (counter calibration) and exit")
    ap.add_argument("--stage2-chunk", type=int, default=None,
                    help="block rows per stage-2 partition (0: one-rank sweeps); default: the package's value")
    ap.add_argument("--relax-chunk", type=int, default=None,
                    help="rows per relax-6 chunk (0: one-thread sweeps); default: the package's large-workload value")
    ap.add_argument("--arnoldi-kernel", type=int, default=2,
                    help="0: one cooperative launch per MGS step, 1: one launch per projection, "
                         "2: the one-launch step by an ordinary launch")
    ap.add_argument("--no-extra", action="store_true",
                    help="skip the subsystem legs (reference-exact, BILU(0), 4x4 SpMV, decoupling)")
    return ap.parse_args()


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


class _Roctx:
    """ROCTx ranges around the legs (rocprofv3 --marker-trace records them, so
    a kernel trace can be split by leg); no-ops where the library is absent."""

    def __init__(self):
        import ctypes
        self.lib = None
        for name in ("librocprofiler-sdk-roctx.so.1", "librocprofiler-sdk-roctx.so"):
            try:
                self.lib = ctypes.CDLL(name)
                break
            except OSError:
                continue
        if self.lib is not None:
            self.lib.roctxRangePushA.argtypes = [ctypes.c_char_p]

    def push(self, name):
        if self.lib is not None:
            self.lib.roctxRangePushA(name.encode())

    def pop(self):
        if self.lib is not None:
            self.lib.roctxRangePop()


def read_traffic(key):
    """HBM bytes per launch (kernel class or kernel) from the committed rocprof
    counter summary, if any."""
    path = os.path.join(ROOT, "profiles", "traffic.json")
    try:
        with open(path) as f:
            return json.load(f).get(key)
    except (OSError, ValueError):
        return None


def host_cpu():
    """(CPU model, logical CPUs, CPUs this process may run on)"""
    model = "unknown"
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
    except (AttributeError, OSError):
        usable = os.cpu_count() or 1
    return model, os.cpu_count() or 1, usable


def reference_cpu_solve(A, rhs, threads):
    """The reference's own whole solve phase (oracle/_ref/ref_driver: its jxf_
    library compiled from the reference sources, FP64) on the same raw
    system, ABF-decoupled by the reference itself, CPR-GMRES(30) to rtol 1e-6
    with OMP_NUM_THREADS=threads; None when the build is absent or fails."""
    import re
    import subprocess
    import tempfile

    import jxpamg_all_amd as jx
    exe = os.path.join(ROOT, "oracle", "_ref", "ref_driver")
    if not os.access(exe, os.X_OK):
        return None
    with tempfile.TemporaryDirectory() as tmp:
        fa, fb, fo = (os.path.join(tmp, n) for n in ("A.bin", "b.rhs", "out.dump"))
        jx.write_bsr_binary(fa, A)
        jx.write_rhs_text(fb, rhs)
        env = dict(os.environ, OMP_NUM_THREADS=str(threads))
        try:
            r = subprocess.run([exe, fa, fb, fo, "1e-6", "100", "3", "1", "0", "30"], capture_output=True,
                               text=True, timeout=900, env=env)
        except (OSError, subprocess.TimeoutExpired):
            return None
    m = re.search(r"ORACLE iters=(\d+) rel=\S+ true_rel=(\S+) solve=([0-9.]+)s setup=([0-9.]+)s", r.stderr)
    if r.returncode != 0 or not m:
        return None
    out = {"s": float(m.group(3)), "iterations": int(m.group(1)), "true_rel_res": float(m.group(2)),
           "threads": threads, "host_setup_s": float(m.group(4))}
    log(f"reference CPU solve, {threads} thread(s): {out}")
    return out


def cpu_baseline(A, rhs, Ad, bd, H, kind):
    """Whole C2 solves on the host: the reference build with one thread and
    with every usable core (capped at 16); the oracle restatement (one
    thread) when the reference build is absent."""
    model, ncpu, usable = host_cpu()
    threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or min(usable, 16)
    sample = "the whole solve phase (JXF_GMRESSolve from x0 = 0 to rtol 1e-6; host setup not counted)"
    if kind == "reference":
        one = reference_cpu_solve(A, rhs, 1)
        if one is not None:
            allc = reference_cpu_solve(A, rhs, threads) if threads > 1 else None
            best = allc if allc is not None and allc["s"] < one["s"] else one
            return {"value": round(best["s"], 2), "unit": "s", "cores": best["threads"], "kind": "reference",
                    "sample": f"{sample} of the same 60x220x85 system with the reference's FP64 jxf_ library built "
                              f"from its sources (oracle/_ref/ref_driver), OMP_NUM_THREADS={best['threads']}",
                    "cpu_model": model, "nproc": ncpu, "usable_cpus": usable,
                    "one_thread": one, "all_threads": allc}
    import numpy as np
    from oracle.cpu_ref import Oracle
    o = Oracle(Ad, H)
    t0 = time.time()
    xo, sto, histo, rc = o.gmres(bd, o.params(tol=1e-6, max_iter=100))
    t_cpu = time.time() - t0
    true = float(np.linalg.norm(bd - Ad.to_scipy() @ xo) / np.linalg.norm(bd))
    log(f"oracle CPU solve: {t_cpu:.1f} s, {sto.iterations} iterations")
    return {"value": round(t_cpu, 2), "unit": "s", "cores": 1, "kind": "port",
            "sample": f"{sample} of the same system with the oracle (single-thread restatement of the reference, "
                      f"pinned to it bit for bit): {sto.iterations} iterations, true rel. residual {true:.2e}",
            "cpu_model": model, "nproc": ncpu, "usable_cpus": usable}


def decouple_legs(A, rhs):
    """ABF and True-IMPES decoupling of the raw system on the GPU (in place,
    with the pressure matrix values in the same pass), wall time of the
    synchronous entry point over the algorithmic bytes: matrix values read
    and written once, row pointers, rhs read and written, pressure values
    written; True-IMPES adds its column-sum pass (values, column indices,
    row pointers read once, nb*b coefficients written and read)."""
    import numpy as np
    import torch

    import jxpamg_all_amd as jx
    ia = torch.tensor(A.ia, dtype=torch.int32, device="cuda")
    ja = torch.tensor(A.ja, dtype=torch.int32, device="cuda")
    a0 = torch.tensor(A.a, dtype=torch.float64, device="cuda")
    r0 = torch.tensor(rhs, dtype=torch.float64, device="cuda")
    a, r = a0.clone(), r0.clone()
    pa = torch.empty(A.nnzb, dtype=torch.float64, device="cuda")
    nnzb, nb, N, bb = A.nnzb, A.nb, A.n, A.b * A.b
    apply_bytes = 2 * 8 * nnzb * bb + 4 * (nb + 1) + 16 * N + 8 * nnzb
    coef_bytes = 8 * nnzb * bb + 4 * nnzb + 4 * (nb + 1) + 16 * N
    out = {}
    for name, dtype, by in (("abf", jx.DECOUP_ABF, apply_bytes), ("true_impes", jx.DECOUP_TIMPES,
                                                                  apply_bytes + coef_bytes)):
        ts = []
        for rep in range(4):
            a.copy_(a0)
            r.copy_(r0)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            rc = jx.decouple_device(nb, A.b, ia.data_ptr(), ja.data_ptr(), a.data_ptr(), r.data_ptr(), dtype, 0, 0,
                                    pa.data_ptr())
            torch.cuda.synchronize()
            if rep:
                ts.append(time.perf_counter() - t0)
            assert rc == 0, f"device decoupling {name} failed: {rc}"
        t = float(np.median(ts))
        out[name] = {"ms": round(t * 1e3, 3), "algorithmic_bytes": int(by), "GBs": round(by / t / 1e9, 1),
                     "frac": round(by / t / 1e9 / PEAK_HBM_GBS, 4)}
        log(f"decoupling on the GPU, {name}: {out[name]}")
    return out


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (the solve phase runs on a single MI355X); got --gpus {args.gpus}",
              file=sys.stderr)
        sys.exit(2)
    import numpy as np
    import torch

    if args.copy_probe:
        n = args.copy_probe * 1024 * 1024 // 8
        x = torch.ones(n, dtype=torch.float64, device="cuda")
        y = torch.empty_like(x)
        for _ in range(3):
            y.copy_(x)
        torch.cuda.synchronize()
        log(f"copy probe: 3 x {n * 8} bytes read and written")
        return

    import jxpamg_all_amd as jx

    assert torch.cuda.is_available(), "bench.py needs a GPU"
    torch.cuda.set_device(0)
    cfg = jx.CONFIGS[args.config]

    t0 = time.time()
    A, rhs = jx.generate(**cfg)
    Ad, bd = jx.decouple_abf(A, rhs)
    t_gen = time.time() - t0
    t0 = time.time()
    ia, ja, a = jx.extract_pressure(Ad)
    H = jx.amg_setup(ia, ja, a)
    t_amg = time.time() - t0
    log(f"{args.config}: nb={Ad.nb} nnzb={Ad.nnzb} levels={H.num_levels} sizes={[len(m[0]) - 1 for m in H.A]} "
        f"gen+decouple {t_gen:.1f}s amg {t_amg:.1f}s")
    chunk = jx.BENCH_RELAX_CHUNK if args.relax_chunk is None else args.relax_chunk
    s2chunk = jx.BENCH_STAGE2_CHUNK if args.stage2_chunk is None else args.stage2_chunk
    params = jx.default_params(tol=1e-6, max_iter=100, k_dim=30, relax_chunk=chunk, stage2_chunk=s2chunk,
                               arnoldi_kernel=args.arnoldi_kernel)
    t0 = time.time()
    s = jx.Solver(Ad, H, params)
    t_create = time.time() - t0
    log(f"solver create (layouts + upload) {t_create:.1f}s")

    b_d = torch.tensor(bd, dtype=torch.float64, device="cuda")
    x_d = torch.zeros_like(b_d)
    hist = np.zeros(params.max_iter + 1)

    for w in range(args.warmup):
        x_d.zero_()
        st = s.solve_device(b_d.data_ptr(), x_d.data_ptr(), hist)
        log(f"warmup {w}: iters {st.iterations} true_rel {st.true_rel_res:.3e} device {st.t_solve_ms:.1f} ms")
    rtx = _Roctx()
    times = []
    rtx.push("plain_fp64")
    for k in range(args.steps):
        x_d.zero_()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        st = s.solve_device(b_d.data_ptr(), x_d.data_ptr(), hist)
        torch.cuda.synchronize()
        times.append(time.perf_counter() - t0)
        log(f"step {k}: {times[-1] * 1e3:.1f} ms iters {st.iterations} true_rel {st.true_rel_res:.3e}")
    rtx.pop()
    sec = float(np.mean(times))

    def timed_solves(solver, n, name="", rng=None):
        ts, stx = [], None
        x_d.zero_()
        solver.solve_device(b_d.data_ptr(), x_d.data_ptr(), hist)
        if rng:
            rtx.push(rng)
        for k in range(n):
            x_d.zero_()
            torch.cuda.synchronize()
            t1 = time.perf_counter()
            stx = solver.solve_device(b_d.data_ptr(), x_d.data_ptr(), hist)
            torch.cuda.synchronize()
            ts.append(time.perf_counter() - t1)
            if name:
                log(f"{name} solve {k}: {ts[-1] * 1e3:.1f} ms")
        if rng:
            rtx.pop()
        return float(np.mean(ts)), stx

    mixed = None
    if not args.no_learned:
        # mixed precision + learned polynomial smoother.  Reference-style mixed
        # apply first: FP32 stage 1 (relax-6 chunks), FP64 GMRES, stage 2 as in
        # the reference's mixed driver (two block sweeps, w 1.0, omega 0.8:
        # src/test_bsr_cprgmres_hybrid.c:373-375, 648) with its blocks stored in
        # FP32.  Then the mixed + learned configuration: FP16 stage 1 (row
        # scaled, FP32 vectors) with a degree-2 polynomial smoother whose roots
        # the network predicts (trained here, on the GPU, on data the plain
        # solver labels itself), judged level by level against relax 6 (the
        # plain smoother, also the fallback), one stage-2 sweep with FP16
        # blocks (row-scaled); the same with the default Chebyshev roots is
        # the method without the network
        from jxpamg_all_amd import learned as lrn
        net, t_train = None, 0.0
        if args.train_seeds > 0:
            t0 = time.time()
            Xtr, Ttr, _ = lrn.make_dataset(range(1001, 1001 + args.train_seeds), precision=2, m=2)
            net = lrn.MLP(seed=0)
            losses = net.fit(Xtr, Ttr, steps=3000, lr=3e-3)
            t_train = time.time() - t0
            log(f"learned smoother: {len(Xtr)} samples, loss {losses[0]:.4f} -> {losses[-1]:.5f}, {t_train:.1f}s")
        legs = {}
        stage2_ref = dict(stage2_maxit=2, stage2_weight=1.0, stage2_omega=0.8, stage2_chunk=s2chunk, stage2_precision=1)
        stage2_one = dict(stage2_maxit=1, stage2_weight=1.0, stage2_omega=0.8, stage2_chunk=s2chunk, stage2_precision=2)
        mixed_kw = dict(precond_precision=2, poly_degree=2, relax_chunk=chunk, coarse_rows=4096, **stage2_one)
        runs = [("mixed_fp32_relax6", dict(precond_precision=1, relax_chunk=chunk, **stage2_ref)),
                ("mixed_default_poly", mixed_kw)]
        if net is not None:
            runs.append(("mixed_learned", mixed_kw))
        for name, kw in runs:
            pm = jx.default_params(tol=1e-6, max_iter=100, k_dim=30, arnoldi_kernel=args.arnoldi_kernel, **kw)
            sm = jx.Solver(Ad, H, pm)
            used = None
            if name in ("mixed_default_poly", "mixed_learned"):
                # judged level by level against relax 6, the fallback
                _, f_j, f_p = sm.apply_mlp(net.P.data_ptr()) if name == "mixed_learned" else sm.judge_poly()
                kinds = sm.level_smoothers()[:-1]
                used = {"relax6": kinds.count("R"), "default_roots": kinds.count("D"), "learned_roots": kinds.count("L")}
                log(f"{name}: polynomial on {used} level(s), V-cycle factor {f_j:.4f} (relax 6 everywhere "
                    f"{f_p:.4f}), level smoothers {sm.level_smoothers()[:-1]} (R relax 6, D default roots, L learned)")
            t_m, st_m = timed_solves(sm, args.steps, rng=name)
            assert st_m.converged == 1 and st_m.true_rel_res <= 1e-6, f"{name} did not reach rtol 1e-6"
            legs[name] = {"s": round(t_m, 4), "iterations": st_m.iterations, "true_rel_res": st_m.true_rel_res,
                          "speedup_vs_plain_fp64": round(sec / t_m, 3)}
            if used is not None:
                legs[name]["level_smoothers"] = used
                legs[name]["vcycle_factor"] = round(f_j, 4)
                legs[name]["vcycle_factor_relax6"] = round(f_p, 4)
            log(f"{name}: {t_m * 1e3:.1f} ms iters {st_m.iterations} true_rel {st_m.true_rel_res:.3e}")
            sm.close()
        legs["training_s"] = round(t_train, 2) if net is not None else None
        mixed = legs
    def profiled(solver):
        solver.profile(True)
        x_d.zero_()
        solver.solve_device(b_d.data_ptr(), x_d.data_ptr(), hist)
        pr, lv = solver.profile_read(), solver.profile_read_levels()
        solver.profile(False)
        return pr, lv

    def gbs(v):
        return round(v[2] / (v[1] * 1e-3) / 1e9, 1) if v[1] > 0 else None

    extra = None
    if not args.no_extra:
        extra = {}
        # the reference's own one-thread semantics: sequential relax-6 sweeps
        # (k_gs_sweep) and the one-rank stage 2 (k_bhgs_multi)
        pe = jx.default_params(tol=1e-6, max_iter=100, k_dim=30, relax_chunk=0, stage2_chunk=0,
                               arnoldi_kernel=args.arnoldi_kernel)
        se = jx.Solver(Ad, H, pe)
        log("reference-exact solver created")
        t_e, st_e = timed_solves(se, 1, "reference-exact", rng="reference_exact")
        pr, lv = profiled(se)
        log("reference-exact profiled solve done")
        se.close()
        extra["reference_exact"] = {
            "s": round(t_e, 4), "iterations": st_e.iterations, "true_rel_res": st_e.true_rel_res,
            "k_gs_sweep_per_level": {f"level {l}": {"launches": v[0], "ms": round(v[1], 3), "GBs": gbs(v)}
                                     for l, v in lv.items()},
            "k_bhgs_multi": {"launches": pr["bhgs_sweep"][0], "ms": round(pr["bhgs_sweep"][1], 3),
                             "GBs": gbs(pr["bhgs_sweep"])}}
        log(f"reference-exact (one-thread relax 6, one-rank stage 2): {t_e * 1e3:.1f} ms iters {st_e.iterations}")
        # stage 2 as BILU(0) corrections with level-scheduled block triangular solves (k_btri)
        pb = jx.default_params(tol=1e-6, max_iter=100, k_dim=30, relax_chunk=chunk, stage2_solver_type=5,
                               arnoldi_kernel=args.arnoldi_kernel)
        sb = jx.Solver(Ad, H, pb)
        t_b, st_b = timed_solves(sb, args.steps, "BILU(0) stage 2", rng="bilu0_stage2")
        pr, _ = profiled(sb)
        sb.close()
        extra["bilu0_stage2"] = {"s": round(t_b, 4), "iterations": st_b.iterations, "true_rel_res": st_b.true_rel_res,
                                 "k_btri": {"launches": pr["bilu_solve"][0], "ms": round(pr["bilu_solve"][1], 3),
                                            "GBs": gbs(pr["bilu_solve"]),
                                            "frac": round((gbs(pr["bilu_solve"]) or 0) / PEAK_HBM_GBS, 4)}}
        log(f"BILU(0) stage 2: {t_b * 1e3:.1f} ms iters {st_b.iterations} k_btri {extra['bilu0_stage2']['k_btri']}")
        # the same with FP16 factors (a power of two per block row) and FP32 inverted diagonals
        pb16 = jx.default_params(tol=1e-6, max_iter=100, k_dim=30, relax_chunk=chunk, stage2_solver_type=5,
                                 stage2_precision=2, arnoldi_kernel=args.arnoldi_kernel)
        sb16 = jx.Solver(Ad, H, pb16)
        t_b16, st_b16 = timed_solves(sb16, args.steps, "BILU(0) FP16 factors", rng="bilu0_stage2_f16")
        pr, _ = profiled(sb16)
        sb16.close()
        extra["bilu0_stage2_f16_factors"] = {
            "s": round(t_b16, 4), "iterations": st_b16.iterations, "true_rel_res": st_b16.true_rel_res,
            "k_btri": {"launches": pr["bilu_solve"][0], "ms": round(pr["bilu_solve"][1], 3),
                       "GBs": gbs(pr["bilu_solve"]), "frac": round((gbs(pr["bilu_solve"]) or 0) / PEAK_HBM_GBS, 4)}}
        log(f"BILU(0) FP16 factors: {t_b16 * 1e3:.1f} ms iters {st_b16.iterations} "
            f"k_btri {extra['bilu0_stage2_f16_factors']['k_btri']}")
        extra["decoupling"] = decouple_legs(A, rhs)
    iters = st.iterations
    assert st.converged == 1 and st.true_rel_res <= 1e-6, "solve did not reach rtol 1e-6"

    # per-kernel device time of one more solve, with events around every
    # launch: classes that are one kernel, and the smoother per level (one
    # kernel per level); the dominant kernel is the one with the most time
    s.profile(True)
    x_d.zero_()
    s.solve_device(b_d.data_ptr(), x_d.data_ptr(), hist)
    prof = s.profile_read()
    levels = s.profile_read_levels()
    s.profile(False)
    single = {"bhgs_sweep": ("k_bhgs_part<3, 4, double, double, 6, double>" if s2chunk else "k_bhgs_multi<3>"),
              "bsr_spmv": "k_bsr_spmv<3, 8, true>", "mgs": "k_mgs_coop<16>", "coarse_solve": "k_dense_gemv<double>",
              "bilu_solve": "k_btri"}
    cands = [(name, prof[c]) for c, name in single.items() if prof[c][0] > 0]

    def smoother_kernel(l):
        ia = H.A[l][0]
        m = (ia[-1] / (len(ia) - 1)) if len(ia) > 1 else 0.0
        if not chunk:
            return "k_gs_sweep<double, double>"
        if m < 12:
            return "k_gs_chunk_lds<4, double, double, true>"
        g = 4
        while g < 32 and g * 8 < m:
            g *= 2
        return f"k_gs_chunk_grp<4, {g}, double, double>"
    cands += [(smoother_kernel(l) + f" (level {l})", v) for l, v in levels.items()]
    dom_name, (n_l, ms_l, by_l) = max(cands, key=lambda c: c[1][1])
    per_launch_ms = ms_l / n_l
    achieved = (by_l / n_l) / (per_launch_ms * 1e-3) / 1e9
    traffic = read_traffic(dom_name.split(" (level")[0])
    roofline = {"bound": "hbm", "achieved": round(achieved, 2), "peak": PEAK_HBM_GBS, "unit": "GB/s",
                "frac": round(achieved / PEAK_HBM_GBS, 4), "traffic": traffic,
                "kernel": dom_name, "launches_per_solve": n_l,
                "mean_launch_us": round(per_launch_ms * 1e3, 2),
                "algorithmic_bytes_per_launch": round(by_l / n_l)}
    smoother_levels = {f"level {l} {smoother_kernel(l)}": {
        "launches": v[0], "ms": round(v[1], 3), "GBs": round(v[2] / (v[1] * 1e-3) / 1e9, 1),
        "traffic": read_traffic(smoother_kernel(l))} for l, v in levels.items()}
    breakdown = {k: {"launches": v[0], "ms": round(v[1], 3),
                     "GBs": round(v[2] / (v[1] * 1e-3) / 1e9, 1) if v[1] > 0 else None} for k, v in prof.items()}
    log(f"kernel classes: {json.dumps(breakdown)}")

    # BSR SpMV bandwidth (second half of the metric).  The ABF-decoupled
    # system's identity diagonal blocks are not stored (the kernel starts each
    # row from x_i), so the bytes the kernel moves are the off-diagonal
    # blocks with their column indices, row pointers, x and y: 564.4 MB here
    # against 649.6 MB in the all-blocks formula.  Timed over the SpMV
    # launches of two profiled solves (HIP events per launch) and back to back.
    nb_, bb_ = Ad.nb, Ad.b * Ad.b
    spmv_read = (Ad.nnzb - nb_) * (8 * bb_ + 4) + 4 * (nb_ + 1) + 16 * Ad.n
    spmv_all = Ad.nnzb * (8 * bb_ + 4) + 4 * (nb_ + 1) + 16 * Ad.n
    n_sp = ms_sp = 0.0
    for _ in range(2):
        s.profile(True)
        x_d.zero_()
        s.solve_device(b_d.data_ptr(), x_d.data_ptr(), hist)
        n_, ms_, _ = s.profile_read()["bsr_spmv"]
        s.profile(False)
        n_sp += n_
        ms_sp += ms_
    us_in = ms_sp / n_sp * 1e3
    y_d = torch.empty_like(b_d)
    s.bsr_matvec_timed(b_d.data_ptr(), y_d.data_ptr(), 5)
    rtx.push("spmv_back_to_back")
    ms = s.bsr_matvec_timed(b_d.data_ptr(), y_d.data_ptr(), args.spmv_reps)
    rtx.pop()
    spmv = {"kernel": "k_bsr_spmv<3, 8, true>", "bytes_read": spmv_read, "bytes_all_blocks_formula": spmv_all,
            "in_solve_launches": int(n_sp), "in_solve_mean_launch_us": round(us_in, 2),
            "achieved_GBs": round(spmv_read / (us_in * 1e-6) / 1e9, 1),
            "frac": round(spmv_read / (us_in * 1e-6) / 1e9 / PEAK_HBM_GBS, 4),
            "back_to_back_mean_launch_us": round(ms * 1e3, 2),
            "back_to_back_GBs": round(spmv_read / (ms * 1e-3) / 1e9, 1),
            "traffic": read_traffic("k_bsr_spmv<3, 8, true>")}
    # what plain streaming attains on this GPU, for scale: a read-only
    # reduction and a device copy over 2 GiB (torch's kernels)
    big = torch.ones(1 << 28, dtype=torch.float64, device="cuda")
    big2 = torch.empty_like(big)
    for key, fn, nbytes in (("device_read_GBs", lambda: big.sum(), 8 << 28),
                            ("device_copy_GBs", lambda: big2.copy_(big), 16 << 28)):
        fn()
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(10):
            fn()
        e1.record()
        torch.cuda.synchronize()
        spmv[key] = round(nbytes / (e0.elapsed_time(e1) / 10 * 1e-3) / 1e9, 1)
    del big, big2
    spmv["back_to_back_over_device_read"] = round(spmv["back_to_back_GBs"] / spmv["device_read_GBs"], 3)
    log(f"BSR SpMV {spmv}")
    assert s.error_flag() == 0
    if extra is not None:
        # 4x4 BSR SpMV on the 100^3 system (exact ABF decoupling, its own hierarchy)
        A4, r4 = jx.generate(**jx.CONFIGS["C4"])
        A4d, _ = jx.decouple_abf(A4, r4, exact=True)
        del A4, r4
        H4 = jx.amg_setup(*jx.extract_pressure(A4d))
        s4 = jx.Solver(A4d, H4, jx.default_params(tol=1e-8, max_iter=100, k_dim=30, relax_chunk=chunk,
                                                  stage2_chunk=s2chunk, arnoldi_kernel=args.arnoldi_kernel))
        x4 = torch.ones(A4d.n, dtype=torch.float64, device="cuda")
        y4 = torch.empty_like(x4)
        s4.bsr_matvec_timed(x4.data_ptr(), y4.data_ptr(), 5)
        ms4 = s4.bsr_matvec_timed(x4.data_ptr(), y4.data_ptr(), args.spmv_reps)
        # identity diagonal blocks after decoupling: not stored, not read
        by4 = (A4d.nnzb - A4d.nb) * (8 * 16 + 4) + 4 * (A4d.nb + 1) + 16 * A4d.n
        extra["spmv_4x4_100cubed"] = {"kernel": "k_bsr_spmv<4, 6, true>", "mean_launch_us": round(ms4 * 1e3, 2),
                                      "bytes_read": by4, "achieved_GBs": round(by4 / (ms4 * 1e-3) / 1e9, 1),
                                      "frac": round(by4 / (ms4 * 1e-3) / 1e9 / PEAK_HBM_GBS, 4)}
        log(f"BSR SpMV 4x4 {extra['spmv_4x4_100cubed']}")
        s4.close()
        del A4d, H4, x4, y4

    cpu = None if args.no_cpu else cpu_baseline(A, rhs, Ad, bd, H, args.cpu_kind)

    out = {"metric": METRIC, "value": round(sec, 4), "unit": "s", "n_gpus": 1, "steps": args.steps,
           "warmup": args.warmup, "ms_per_step": round(sec * 1e3, 2), "higher_is_better": False,
           "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
           "config": {"workload": f"{args.config} {cfg['nx']}x{cfg['ny']}x{cfg['nz']} 3x3 BSR FP64, ABF, "
                                  f"CPR(AMG V-cycle hSGS + 3 block HGS)-GMRES(30) to rtol 1e-6",
                      "relax6_rows_per_chunk": chunk, "stage2_rows_per_partition": s2chunk,
                      "cells": Ad.nb, "nnzb": Ad.nnzb, "iterations": iters,
                      "true_rel_res": st.true_rel_res, "amg_levels": [len(m[0]) - 1 for m in H.A],
                      "host_setup_s": {"generate_decouple": round(t_gen, 2), "amg": round(t_amg, 2),
                                       "upload": round(t_create, 2)}},
           "roofline": roofline, "spmv": spmv, "kernel_classes": breakdown, "smoother_levels": smoother_levels,
           "cpu_baseline": cpu,
           "mixed_precision_learned": mixed, "subsystems": extra}
    s.close()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
