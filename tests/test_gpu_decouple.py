"""GPU: decoupling and pressure extraction on device arrays through the
C-ABI, bit for bit against the reference build's outputs (decoup.npz: every
type and block size; the 60x220x85 system by SHA-256) and against the host
restatement."""
import hashlib

import numpy as np
import pytest

import jxpamg_all_amd as jx
from tests._common import bits, decoup_cases, decoup_system
from tests.conftest import fixture_gen_params, load_fixture

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _dev(a, dt=torch.float64):
    return torch.tensor(np.ascontiguousarray(a), dtype=dt, device="cuda")


def _device_decouple(A, rhs, dtype, thermal=0, pidx=0, pressure=True):
    ia, ja, a = _dev(A.ia, torch.int32), _dev(A.ja, torch.int32), _dev(A.a)
    r = _dev(rhs)
    pa = torch.full((A.nnzb,), np.nan, dtype=torch.float64, device="cuda") if pressure else None
    rc = jx.decouple_device(A.nb, A.b, ia.data_ptr(), ja.data_ptr(), a.data_ptr(), r.data_ptr(), dtype, thermal,
                            pidx, pa.data_ptr() if pressure else None)
    torch.cuda.synchronize()
    return rc, a.cpu().numpy(), r.cpu().numpy(), (pa.cpu().numpy() if pressure else None)


@pytest.mark.parametrize("key,b,dtype,thermal", decoup_cases())
def test_device_decoupling_bit_exact_vs_reference(key, b, dtype, thermal):
    d = load_fixture("decoup")
    A, rhs = decoup_system(d, b)
    print(f"[gpu] decouple {key}", flush=True)
    rc, a, r, pa = _device_decouple(A, rhs, dtype, thermal)
    assert rc == 0
    assert np.array_equal(bits(a), bits(d[key + "_a"]))
    assert np.array_equal(bits(r), bits(d[key + "_b"]))
    # the fused pressure values are the (0, 0) entries of the decoupled blocks
    assert np.array_equal(bits(pa), bits(d[key + "_a"].reshape(-1, b, b)[:, 0, 0]))


@pytest.mark.parametrize("dtype", [jx.DECOUP_TIMPES, jx.DECOUP_TIMPES2])
def test_device_true_impes_rejects_unsymmetric_pattern_untouched(dtype):
    d = load_fixture("decoup")
    _, rhs = decoup_system(d, 3)
    As = jx.BSRMatrix(len(d["asym_ia"]) - 1, 3, d["asym_ia"], d["asym_ja"], d["asym_a"])
    rc, a, r, _ = _device_decouple(As, rhs, dtype)
    assert rc == 4  # JXAMD_ERROR_ARG, nothing modified
    assert np.array_equal(bits(a), bits(As.a)) and np.array_equal(bits(r), bits(rhs))


def test_device_pressure_extraction_and_bad_arguments():
    A, rhs = jx.generate(**dict(jx.CONFIGS["C1"], nx=9, ny=7, nz=5, block_size=4, seed=3))
    ia, ja, a = _dev(A.ia, torch.int32), _dev(A.ja, torch.int32), _dev(A.a)
    for pidx in range(4):
        pia = torch.zeros(A.nb + 1, dtype=torch.int32, device="cuda")
        pja = torch.zeros(A.nnzb, dtype=torch.int32, device="cuda")
        pa = torch.zeros(A.nnzb, dtype=torch.float64, device="cuda")
        assert jx.lib().jxamd_bsr_extract_pressure_device(A.nb, 4, A.nnzb, ia.data_ptr(), ja.data_ptr(), a.data_ptr(),
                                                          pidx, pia.data_ptr(), pja.data_ptr(), pa.data_ptr(),
                                                          None) == 0
        torch.cuda.synchronize()
        hia, hja, ha = jx.extract_pressure(A, pidx)
        assert np.array_equal(pia.cpu().numpy(), hia) and np.array_equal(pja.cpu().numpy(), hja)
        assert np.array_equal(bits(pa.cpu().numpy()), bits(ha))
    L = jx.lib()
    p = a.data_ptr()
    assert L.jxamd_bsr_extract_pressure_device(A.nb, 4, A.nnzb, ia.data_ptr(), ja.data_ptr(), p, 4, None, None, p,
                                               None) == 4
    assert L.jxamd_bsr_decouple_device(A.nb, 5, ia.data_ptr(), ja.data_ptr(), p, None, 1, 0, 0, None, None) == 4
    assert L.jxamd_bsr_decouple_device(A.nb, 4, ia.data_ptr(), ja.data_ptr(), p, None, 7, 0, 0, None, None) == 4
    # nothing to do for an empty system
    assert L.jxamd_bsr_decouple_device(0, 3, None, None, None, None, 5, 0, 0, None, None) == 0


@pytest.mark.parametrize("name,dtype", [("c2_dec_abf", jx.DECOUP_ABF), ("c2_dec_timpes", jx.DECOUP_TIMPES)])
def test_device_decoupling_c2_matches_reference_hash(name, dtype):
    """the full 60x220x85 system decoupled on the GPU = the reference's (SHA-256
    of the decoupled values and rhs), with the pressure matrix of the same pass
    equal to the host extraction"""
    d = load_fixture(name)
    A, rhs = jx.generate(**fixture_gen_params(d))
    print(f"[gpu] decouple {name} nb={A.nb}", flush=True)
    rc, a, r, pa = _device_decouple(A, rhs, dtype)
    assert rc == 0
    assert hashlib.sha256(a.tobytes()).hexdigest() == bytes(d["dec_a_sha256"]).decode()
    assert hashlib.sha256(r.tobytes()).hexdigest() == bytes(d["dec_b_sha256"]).decode()
    assert np.array_equal(bits(pa), bits(a.reshape(-1, 3, 3)[:, 0, 0]))
