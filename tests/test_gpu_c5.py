"""GPU: the learned-smoother workload at full size.  One unseen 120x440x85
system (4,488,000 cells, 3x3 blocks; seed outside the training seeds) solved
with the plain FP64 path and with the FP32 preconditioner whose AMG levels
smooth with relax 6 (the plain smoother), with the degree-2 polynomial on
default Chebyshev roots, and with the roots the network predicts; both
polynomial variants are judged level by level against relax 6, which is also
their fallback.  The network is trained here, on the GPU, on small systems
the plain solver labels itself."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
ev = pytest.importorskip("jxpamg_all_amd.evaluate")


def test_c5_unseen_seed_learned_default_and_plain_smoothers():
    torch.cuda.set_device(0)
    print("[gpu] C5: training the network (FP32 preconditioner, degree 2)", flush=True)
    net, training = ev.train_c5_network(train_seeds=20, precision=1, m=2)
    assert training["loss_last"] < 0.2 * training["loss_first"]
    rec = ev.evaluate_c5_seed(2004, net, reps=1)
    r64 = rec["fp64"]
    print(f"[gpu] C5 seed 2004 fp64: {r64}", flush=True)
    assert r64["converged"] == 1 and r64["true_rel_res"] <= 1e-6
    plain = rec["fp32_plain"]
    for name in ("plain", "default", "learned"):
        r = rec[f"fp32_{name}"]
        print(f"[gpu] C5 seed 2004 fp32 {name}: {r['s'] * 1e3:.1f} ms, {r['iterations']} iterations, "
              f"true rel. residual {r['true_rel_res']:.2e}, error vs FP64 {r['error_vs_fp64']:.2e}, "
              f"{r['speedup_vs_fp64']}x, smoother {r['smoother']}", flush=True)
        assert r["converged"] == 1 and r["true_rel_res"] <= 1e-6   # FP64-verified
        assert r["error_vs_fp64"] < 5e-3
        if name != "plain":
            sm = r["smoother"]
            # the judge's reference is relax 6 everywhere; the judged smoother never contracts worse
            assert abs(sm["factor_relax6"] - plain["smoother"]["factor"]) < 1e-3
            assert sm["factor"] <= sm["factor_relax6"] + 1e-12
            assert r["iterations"] <= plain["iterations"] + 1
    lrn, dft = rec["fp32_learned"], rec["fp32_default"]
    print(f"[gpu] C5 learned vs default: factor {lrn['smoother']['factor']} vs {dft['smoother']['factor']}, "
          f"iterations {lrn['iterations']} vs {dft['iterations']}, {lrn['s'] * 1e3:.1f} vs {dft['s'] * 1e3:.1f} ms",
          flush=True)
