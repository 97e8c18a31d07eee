"""Evaluation of the mixed-precision and learned configurations on one GPU.

    python -m jxpamg_all_amd.evaluate C4      # 100^3, 4x4 blocks (exact decoupling), FP16 stage 1, rtol 1e-8
    python -m jxpamg_all_amd.evaluate C5      # 120x440x85, FP32 preconditioner + learned smoother, 5 unseen seeds
    python -m jxpamg_all_amd.evaluate C3      # the 16x16x8 system with an FP32 hierarchy

Each configuration is solved with the plain FP64 GPU path and with the
mixed / learned variant; printed (one JSON object per configuration): time
to rtol, iterations, the FP64 true residual, and the error of the solution
against the FP64 solve.  Host setup (generation, decoupling, AMG) is timed
separately.
"""
import json
import sys
import time

import numpy as np


def log(msg):
    print(f"[evaluate] {msg}", file=sys.stderr, flush=True)


def _setup(cfg, exact=False):
    from . import amg_setup, decouple_abf, extract_pressure, generate
    t0 = time.time()
    A, rhs = generate(**cfg)
    Ad, bd = decouple_abf(A, rhs, exact=exact)
    H = amg_setup(*extract_pressure(Ad))
    return Ad, bd, H, time.time() - t0


def _solve(Ad, bd, H, params, reps=2, mlp=None, judge=None):
    """Time `reps` solves (after one untimed).  judge: None (the smoothers as
    configured), "plain" (relax 6 on every level), "default" (default
    Chebyshev roots judged level by level against relax 6) or "learned"
    (the network's roots, mlp, judged the same way)."""
    import torch

    from . import Solver
    s = Solver(Ad, H, params)
    info = None
    if judge == "plain":
        for l in range(s.num_levels):
            s.set_level_smoother(l, 1)
        info = {"polynomial_levels": 0, "factor": round(s.vcycle_factor(), 4)}
    elif judge in ("default", "learned"):
        _, f, fp = s.apply_mlp(mlp.P.data_ptr()) if judge == "learned" else s.judge_poly()
        kinds = s.level_smoothers()[:-1]
        info = {"polynomial_levels": sum(c != "R" for c in kinds), "learned_levels": kinds.count("L"),
                "factor": round(f, 4), "factor_relax6": round(fp, 4), "level_smoothers": kinds}
    b = torch.tensor(bd, dtype=torch.float64, device="cuda")
    x = torch.zeros_like(b)
    hist = np.zeros(params.max_iter + 1)
    s.solve_device(b.data_ptr(), x.data_ptr(), hist)
    ts = []
    for _ in range(reps):
        x.zero_()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        st = s.solve_device(b.data_ptr(), x.data_ptr(), hist)
        torch.cuda.synchronize()
        ts.append(time.perf_counter() - t0)
    s.close()
    out = {"s": round(float(np.mean(ts)), 4), "iterations": st.iterations, "converged": st.converged,
           "true_rel_res": st.true_rel_res}
    if info is not None:
        out["smoother"] = info
    return out, x.cpu().numpy()


def _rel(x, ref):
    return float(np.linalg.norm(x - ref) / max(np.linalg.norm(ref), 1e-300))


def evaluate_c3():
    from . import CONFIGS, default_params
    Ad, bd, H, t = _setup(CONFIGS["C1"])
    p64 = default_params(tol=1e-6, relax_chunk=4)
    r64, x64 = _solve(Ad, bd, H, p64)
    r32, x32 = _solve(Ad, bd, H, default_params(tol=1e-6, relax_chunk=4, precond_precision=1))
    r32["error_vs_fp64"] = _rel(x32, x64)
    return {"config": "C3 16x16x8 3x3, FP32 hierarchy and stage-1 vectors", "setup_s": round(t, 2), "fp64": r64,
            "fp32_stage1": r32, "iteration_delta": r32["iterations"] - r64["iterations"]}


def evaluate_c4():
    from . import CONFIGS, default_params
    log("C4: generating 100^3 4x4 system (exact decoupling: the reference's 4x4 block inverse is wrong)")
    Ad, bd, H, t = _setup(CONFIGS["C4"], exact=True)
    log(f"C4: setup {t:.1f}s, levels {[len(m[0]) - 1 for m in H.A]}")
    common = dict(tol=1e-8, max_iter=200, relax_chunk=4, stage2_chunk=4)
    r64, x64 = _solve(Ad, bd, H, default_params(**common))
    log(f"C4 fp64: {r64}")
    r16, x16 = _solve(Ad, bd, H, default_params(precond_precision=2, **common))
    r16["error_vs_fp64"] = _rel(x16, x64)
    log(f"C4 fp16: {r16}")
    rmx, xmx = _solve(Ad, bd, H, default_params(tol=1e-8, max_iter=200, **MIXED))
    rmx["error_vs_fp64"] = _rel(xmx, x64)
    log(f"C4 mixed: {rmx}")
    return {"config": "C4 100^3 4x4, exact ABF decoupling, rtol 1e-8: FP64; FP16 stage 1 (power-of-two row "
                      "scales, FP32 accumulation) with relax-6 chunks; the benchmark's mixed configuration",
            "setup_s": round(t, 2), "fp64": r64, "fp16_stage1": r16, "mixed": rmx,
            "iteration_delta": r16["iterations"] - r64["iterations"], "speedup": round(r64["s"] / r16["s"], 3),
            "speedup_mixed": round(r64["s"] / rmx["s"], 3)}


# the mixed + learned configuration of the benchmark: FP16 stage 1 (row-scaled
# level matrices, FP32 vectors), degree-2 polynomial smoother, one stage-2
# sweep (w 1.0, omega 0.8) with FP16 blocks; GMRES FP64
MIXED = dict(precond_precision=2, poly_degree=2, stage2_maxit=1, stage2_weight=1.0, stage2_omega=0.8,
             stage2_chunk=4, stage2_precision=2, relax_chunk=4, coarse_rows=4096)
# the 120x440x85 learned-smoother configuration: FP32 preconditioner (stage-1
# level matrices, vectors and accumulation; FP32 stage-2 blocks), degree-2
# polynomial smoother judged against relax 6 (chunks of 4 rows) level by
# level, the V-cycle ending at the first level of at most 4096 rows with an
# explicit inverse
C5_MIXED = dict(precond_precision=1, poly_degree=2, stage2_maxit=1, stage2_weight=1.0, stage2_omega=0.8,
                stage2_chunk=4, stage2_precision=1, relax_chunk=4, coarse_rows=4096)


def train_c5_network(train_seeds=20, precision=1, m=2):
    """The learned smoother's network, trained on the GPU on levels of
    train_seeds small systems labelled by the plain solver."""
    from . import learned as lrn
    t0 = time.time()
    Xtr, Ttr, info = lrn.make_dataset(range(1001, 1001 + train_seeds), precision=precision, m=m)
    net = lrn.MLP(seed=0)
    losses = net.fit(Xtr, Ttr, steps=3000, lr=3e-3)
    log(f"trained on {len(Xtr)} levels of {train_seeds} systems, loss {losses[0]:.4f} -> {losses[-1]:.5f}")
    return net, {"samples": len(Xtr), "loss_first": float(losses[0]), "loss_last": float(losses[-1]),
                 "seconds": round(time.time() - t0, 2)}


def evaluate_c5_seed(seed, net, reps=2):
    """One unseen 120x440x85 system: plain FP64 against the FP32
    preconditioner with relax 6 everywhere, with the default roots and with
    the learned roots (both judged against relax 6)."""
    from . import CONFIGS, default_params
    log(f"C5 seed {seed}: setup")
    Ad, bd, H, t = _setup(dict(CONFIGS["C5"], seed=seed))
    r64, x64 = _solve(Ad, bd, H, default_params(tol=1e-6, relax_chunk=4, stage2_chunk=4), reps=reps)
    log(f"C5 seed {seed}: fp64 {r64}")
    pm = default_params(tol=1e-6, **C5_MIXED)
    rec = {"seed": seed, "setup_s": round(t, 2), "fp64": r64}
    for name in ("plain", "default", "learned"):
        r, x = _solve(Ad, bd, H, pm, reps=reps, mlp=net, judge=name)
        r["error_vs_fp64"] = _rel(x, x64)
        r["speedup_vs_fp64"] = round(r64["s"] / r["s"], 3)
        rec[f"fp32_{name}"] = r
        log(f"C5 seed {seed}: fp32 {name} {r}")
    return rec


def evaluate_c5(seeds=(2001, 2002, 2003, 2004, 2005), train_seeds=20):
    net, training = train_c5_network(train_seeds)
    out = {"config": "C5 120x440x85 3x3: plain FP64 against the FP32 preconditioner with relax 6 (plain), and "
                     "with the degree-2 polynomial smoother on default and on learned roots, each judged level by "
                     "level against relax 6; one FP32-block stage-2 sweep; 5 unseen seeds",
           "training": training, "seeds": [evaluate_c5_seed(sd, net) for sd in seeds]}
    return out


def main(argv):
    import torch  # initialise the device through torch before the library touches it
    assert torch.cuda.is_available(), "evaluate needs a GPU"
    torch.cuda.set_device(0)
    which = argv[1:] or ["C3"]
    for w in which:
        fn = {"C3": evaluate_c3, "C4": evaluate_c4, "C5": evaluate_c5}[w]
        print(json.dumps(fn()), flush=True)


if __name__ == "__main__":
    main(sys.argv)
