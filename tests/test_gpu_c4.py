"""GPU: the 100^3 4x4 workload at full size (1,000,000 cells, 4-unknown
blocks, exact ABF decoupling: the reference's 4x4 block inverse is not an
inverse, jxf_bsr_decoup.c:112-183), solved to rtol 1e-8 with the plain FP64
path, with the FP16 stage 1 (level matrices scaled per row by a power of
two, FP32 accumulation) on relax-6 chunks, and with the benchmark's mixed
configuration (FP16 stage 1, degree-2 polynomial smoother, one FP16-block
stage-2 sweep).  Each solve is verified by its FP64 true residual; the
lower-precision preconditioners are compared with the FP64 solve in
iterations and in the solution."""
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
ev = pytest.importorskip("jxpamg_all_amd.evaluate")


def test_c4_fp64_fp16_and_mixed_at_full_size():
    torch.cuda.set_device(0)
    print("[gpu] C4: 100^3 4x4, rtol 1e-8", flush=True)
    rec = ev.evaluate_c4()
    print(f"[gpu] C4 setup {rec['setup_s']} s", flush=True)
    for key in ("fp64", "fp16_stage1", "mixed"):
        r = rec[key]
        print(f"[gpu] C4 {key}: {r['s'] * 1e3:.1f} ms, {r['iterations']} iterations, true rel. residual "
              f"{r['true_rel_res']:.2e}" + (f", error vs FP64 {r['error_vs_fp64']:.2e}" if key != "fp64" else ""),
              flush=True)
        assert r["converged"] == 1 and r["true_rel_res"] <= 1e-8   # FP64-verified
    r64 = rec["fp64"]
    assert 1 <= r64["iterations"] <= 30
    # the FP16 stage 1 keeps the FP64 iteration count within one
    assert abs(rec["iteration_delta"]) <= 1
    # a different preconditioner stops at a different x with the same
    # residual bound: the two solutions agree far beyond FP16 precision
    assert rec["fp16_stage1"]["error_vs_fp64"] < 1e-5
    assert rec["mixed"]["error_vs_fp64"] < 1e-5
    assert rec["mixed"]["iterations"] <= r64["iterations"] + 3
