"""Host-side product code (no GPU): generator, reference formats, ABF
decoupling, pressure extraction, and the C-ABI surface of libjxamd.so."""
import hashlib
import os
import re
import tempfile

import numpy as np
import pytest

import jxpamg_all_amd as jx
from tests._common import bits, decoup_cases, decoup_system, system
from tests.conftest import ROOT, fixture_gen_params, load_fixture

ALL = ["c1", "tiny", "b4", "b2", "lvl3", "nodec", "restart", "layered", "timpes"]


def _sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


@pytest.mark.parametrize("name", ALL)
def test_generator_reproduces_fixture_inputs(name):
    d = load_fixture(name)
    A, rhs = jx.generate(**fixture_gen_params(d))
    assert _sha(A.a) == bytes(d["input_a_sha256"]).decode()
    assert _sha(rhs) == bytes(d["input_rhs_sha256"]).decode()


@pytest.mark.parametrize("b", [2, 3, 4])
def test_generator_structure(b):
    A, rhs = jx.generate(nx=7, ny=5, nz=4, block_size=b, seed=9)
    nx, ny, nz = 7, 5, 4
    nb = nx * ny * nz
    faces = (nx - 1) * ny * nz + nx * (ny - 1) * nz + nx * ny * (nz - 1)
    assert A.nb == nb and A.nnzb == nb + 2 * faces
    for i in range(nb):
        cols = A.ja[A.ia[i]:A.ia[i + 1]]
        assert cols[0] == i
        assert np.all(np.diff(cols[1:]) > 0)
    S = A.to_scipy()
    P = (S != 0).astype(int)
    assert (abs(P - P.T)).sum() == 0 or True  # values differ, block pattern symmetric below
    pat = {(i, j) for i in range(nb) for j in A.ja[A.ia[i]:A.ia[i + 1]]}
    assert all((j, i) in pat for (i, j) in pat)
    A2, rhs2 = jx.generate(nx=7, ny=5, nz=4, block_size=b, seed=9)
    assert np.array_equal(A.a, A2.a) and np.array_equal(rhs, rhs2)
    assert np.all(np.abs(rhs) < 1.0)


@pytest.mark.parametrize("name", ["tiny", "b4", "b2", "timpes"])
def test_abf_decoupling_bit_exact(name):
    d, A, rhs, Ad, bd, H, meta = system(name)
    assert np.array_equal(Ad.a, d["dec_a"])
    assert np.array_equal(bd, d["dec_b"])
    if name == "timpes":  # True-IMPES keeps the diagonal blocks' other rows
        return
    # diagonal blocks become the identity
    for i in range(Ad.nb):
        assert np.array_equal(Ad.a[Ad.ia[i] * Ad.b ** 2:(Ad.ia[i] + 1) * Ad.b ** 2], np.eye(Ad.b).ravel())


def test_abf_rhs_keeps_reference_in_place_product():
    """The reference multiplies the rhs in place; check that quirk directly on a
    1x1-block-row system against the aliased formula, and that it differs from D^-1 b."""
    b = 3
    D = np.array([[4.0, 1.0, 0.5], [0.3, 5.0, 0.2], [0.1, 0.4, 6.0]])
    A = jx.BSRMatrix(1, b, np.array([0, 1], np.int32), np.array([0], np.int32), D.ravel())
    r = np.array([1.0, 2.0, 3.0])
    Ad, rd = jx.decouple_abf(A, r)
    Bi = np.linalg.inv(D)
    v = r.copy()
    for k in range(b):
        v[k] = 0.0
        for q in range(b):
            v[k] += Bi[k, q] * v[q]
    assert np.allclose(rd, v, rtol=1e-12)
    assert not np.allclose(rd, Bi @ r)


def _inv4_reference_form(D):
    """The closed form of jxf_smat_inv_4x4 (jxf_bsr_decoup.c:112-183) written
    as what it computes: diagonal cofactors, off-diagonal (i, j) entry
    (-1)^(i+j) adj(D)_ji, over its own determinant expansion."""
    adj = np.linalg.inv(D) * np.linalg.det(D)
    sgn = np.array([[(-1) ** (i + j) for j in range(4)] for i in range(4)], float)
    out = sgn * adj.T
    np.fill_diagonal(out, np.diag(adj))
    # line 127: a11*M11 - a12*M12 + a13*M13 - a14*M14 with M1k as the code forms them
    m = [np.linalg.det(np.delete(np.delete(D, k, 0), 0, 1)) for k in range(4)]
    det = D[0, 0] * m[0] + D[0, 1] * m[1] + D[0, 2] * m[2] + D[0, 3] * m[3]
    return out / det


def test_abf_4x4_keeps_reference_inverse_and_exact_variant_is_exact():
    """b = 4: the bit-exact decoupling reproduces the reference's closed-form
    4x4 "inverse" (which is not the inverse; the b4 fixture pins it), and the
    exact variant gives D^-1 A_ij and D^-1 b, so its solution is A's."""
    rng = np.random.default_rng(4)
    D = rng.standard_normal((4, 4)) + 5 * np.eye(4)
    C = rng.standard_normal((4, 4))
    ia, ja = np.array([0, 2, 4], np.int32), np.array([0, 1, 1, 0], np.int32)
    D2 = rng.standard_normal((4, 4)) + 6 * np.eye(4)
    C2 = rng.standard_normal((4, 4))
    A = jx.BSRMatrix(2, 4, ia, ja, np.concatenate([D.ravel(), C.ravel(), D2.ravel(), C2.ravel()]))
    r = rng.standard_normal(8)
    Ad, _ = jx.decouple_abf(A, r)
    Wrong = _inv4_reference_form(D)
    assert np.allclose(Ad.a[16:32].reshape(4, 4), Wrong @ C, rtol=1e-10, atol=1e-12)
    assert not np.allclose(Wrong, np.linalg.inv(D), atol=1e-3)
    Ae, re_ = jx.decouple_abf(A, r, exact=True)
    assert np.allclose(Ae.a[16:32].reshape(4, 4), np.linalg.solve(D, C), rtol=1e-12, atol=1e-14)
    assert np.allclose(Ae.a[48:64].reshape(4, 4), np.linalg.solve(D2, C2), rtol=1e-12, atol=1e-14)
    assert np.allclose(re_, np.concatenate([np.linalg.solve(D, r[:4]), np.linalg.solve(D2, r[4:])]), rtol=1e-12)
    x = np.linalg.solve(A.to_scipy().toarray(), r)
    assert np.allclose(Ae.to_scipy() @ x, re_, rtol=1e-10, atol=1e-12)


def test_abf_5x5_reference_gauss_jordan_is_a_permutation():
    """Above 4x4 the reference's Gauss-Jordan (jxf_bsr_decoup.c:200-266)
    row-reduces in place with no identity beside it: the "inverse" it returns
    is a permutation matrix, so the off-diagonal blocks are only permuted.
    The exact variant gives D^-1 C."""
    rng = np.random.default_rng(5)
    D = rng.standard_normal((5, 5)) + 2 * np.eye(5)
    C = rng.standard_normal((5, 5))
    A = jx.BSRMatrix(1, 5, np.array([0, 2], np.int32), np.array([0, 0], np.int32),
                     np.concatenate([D.ravel(), C.ravel()]))
    Ad, _ = jx.decouple_abf(A)
    B = Ad.a[25:].reshape(5, 5)
    P = np.linalg.solve(C.T, B.T).T  # B = P C
    assert np.allclose(P, np.round(P), atol=1e-9)
    Pr = np.round(P)
    assert np.array_equal(np.sort(Pr.sum(0)), np.ones(5)) and np.array_equal(np.sort(Pr.sum(1)), np.ones(5))
    Ae, _ = jx.decouple_abf(A, exact=True)
    assert np.allclose(Ae.a[25:].reshape(5, 5), np.linalg.solve(D, C), rtol=1e-12, atol=1e-13)


def test_abf_exact_matches_reference_matrix_for_3x3():
    """b = 3: the reference's closed-form inverse is exact, so the matrix part
    of both variants is identical bit for bit; only the rhs differs."""
    d, A, rhs, Ad, bd, H, meta = system("c1")
    Ae, be = jx.decouple_abf(A, rhs, exact=True)
    assert np.array_equal(Ae.a, Ad.a)
    x = np.random.default_rng(0).standard_normal(A.n)
    b = A.to_scipy() @ x
    _, bx = jx.decouple_abf(A, b, exact=True)
    assert np.allclose(Ae.to_scipy() @ x, bx, rtol=1e-10, atol=1e-10 * np.abs(bx).max())


@pytest.mark.parametrize("key,b,dtype,thermal", decoup_cases())
def test_every_decoupling_type_bit_exact(key, b, dtype, thermal):
    """jxf_BSRMatrixDecouple of every type (ABF, ANL, SEM, QI, True-IMPES and
    its second form; thermal variants) against the reference build, bit for
    bit including signed zeros"""
    d = load_fixture("decoup")
    A, rhs = decoup_system(d, b)
    Ad, bd = jx.decouple(A, rhs, dtype, thermal)
    assert np.array_equal(bits(Ad.a), bits(d[key + "_a"]))
    assert np.array_equal(bits(bd), bits(d[key + "_b"]))
    if dtype == jx.DECOUP_ABF:
        Aa, ba = jx.decouple_abf(A, rhs)
        assert np.array_equal(bits(Aa.a), bits(Ad.a)) and np.array_equal(bits(ba), bits(bd))


@pytest.mark.parametrize("dtype", [jx.DECOUP_TIMPES, jx.DECOUP_TIMPES2])
def test_true_impes_on_unsymmetric_pattern_bit_exact(dtype):
    d = load_fixture("decoup")
    _, rhs = decoup_system(d, 3)
    As = jx.BSRMatrix(len(d["asym_ia"]) - 1, 3, d["asym_ia"], d["asym_ja"], d["asym_a"])
    Ad, bd = jx.decouple(As, rhs, dtype)
    assert np.array_equal(bits(Ad.a), bits(d[f"asym_t{dtype}_a"]))
    assert np.array_equal(bits(bd), bits(d[f"asym_t{dtype}_b"]))


def test_true_impes_column_sums_closed_form():
    """True-IMPES 3x3 recombines row 0 of every block of row i with the first
    two entries of the first row of (sum over rows of column block i)^{-1};
    restated here with numpy on a tiny system (a dropped column term or a
    transposed sum fails it)"""
    A, rhs = jx.generate(nx=3, ny=2, nz=2, block_size=3, seed=12)
    Ad, bd = jx.decouple(A, rhs, jx.DECOUP_TIMPES)
    blocks = A.a.reshape(-1, 3, 3)
    Ac = np.zeros((A.nb, 3, 3))
    for i in range(A.nb):
        for k in range(A.ia[i], A.ia[i + 1]):
            Ac[A.ja[k]] += blocks[k]
    for i in range(A.nb):
        w = np.linalg.inv(Ac[i])[0]
        for k in range(A.ia[i], A.ia[i + 1]):
            got = Ad.a.reshape(-1, 3, 3)[k]
            assert np.allclose(got[0], w[0] * blocks[k][0] + w[1] * blocks[k][1], rtol=1e-12, atol=1e-14)
            assert np.array_equal(got[1:], blocks[k][1:])
        assert np.isclose(bd[3 * i], w @ rhs[3 * i:3 * i + 3], rtol=1e-12)
        assert np.array_equal(bd[3 * i + 1:3 * i + 3], rhs[3 * i + 1:3 * i + 3])


@pytest.mark.parametrize("name,dtype", [("c2_dec_abf", 1), ("c2_dec_timpes", 5)])
def test_decoupled_c2_hash(name, dtype):
    """the full 60x220x85 system decoupled on the host = the reference's (SHA-256)"""
    d = load_fixture(name)
    A, rhs = jx.generate(**fixture_gen_params(d))
    Ad, bd = jx.decouple(A, rhs, dtype)
    assert _sha(Ad.a) == bytes(d["dec_a_sha256"]).decode()
    assert _sha(bd) == bytes(d["dec_b_sha256"]).decode()


def test_pressure_extraction_matches_reference():
    d, A, rhs, Ad, bd, H, meta = system("c1")
    ia, ja, a = jx.extract_pressure(Ad, 0)
    assert np.array_equal(ia, d["A0_ia"]) and np.array_equal(ja, d["A0_ja"])
    assert np.array_equal(a, d["A0_a"])


def test_decoupled_c1_hash():
    d, A, rhs, Ad, bd, H, meta = system("c1")
    assert _sha(Ad.a) == bytes(d["dec_a_sha256"]).decode()


def test_reference_file_formats_roundtrip():
    A, rhs = jx.generate(nx=5, ny=4, nz=3, block_size=3, seed=2)
    with tempfile.TemporaryDirectory() as t:
        jx.write_bsr_binary(os.path.join(t, "A.bin"), A)
        jx.write_rhs_text(os.path.join(t, "b.rhs"), rhs)
        B = jx.read_bsr_binary(os.path.join(t, "A.bin"))
        r = jx.read_rhs_text(os.path.join(t, "b.rhs"), A.n)
        assert np.array_equal(A.ia, B.ia) and np.array_equal(A.ja, B.ja) and np.array_equal(A.a, B.a)
        assert np.array_equal(r, rhs)
        # header layout of jxf_BSRMatrixPrint_Binary: ROW COL NNZ b manner, ROW+1, IA...
        raw = np.fromfile(os.path.join(t, "A.bin"), dtype=np.int32, count=6)
        assert list(raw) == [A.nb, A.nb, A.nnzb, 3, 0, A.nb + 1]
        # 1-based files are accepted
        hdr = np.fromfile(os.path.join(t, "A.bin"), dtype=np.uint8)
        ints = hdr.view(np.int32)
        off = 6
        ints[off:off + A.nb + 1] += 1
        off += A.nb + 1 + 1
        ints[off:off + A.nnzb] += 1
        hdr.tofile(os.path.join(t, "A1.bin"))
        C = jx.read_bsr_binary(os.path.join(t, "A1.bin"))
        assert np.array_equal(C.ia, A.ia) and np.array_equal(C.ja, A.ja)


def test_library_exports_every_declared_symbol():
    L = jx.lib()
    hdr = open(os.path.join(ROOT, "include", "jxamd.h")).read()
    names = set(re.findall(r"^\s*(?:int|void)\s+\**\s*(jxamd_\w+|JXFPAMG_\w+)\s*\(", hdr, re.M))
    assert len(names) > 20
    missing = [n for n in sorted(names) if not hasattr(L, n)]
    assert not missing, missing


def test_library_shares_torchs_hip_runtime_whatever_the_import_order():
    """libjxamd.so loaded in a fresh process before torch is imported ends up
    with one HIP runtime in the process (torch's): device memory from torch
    and the library's launches and synchronisations then refer to one device
    state (a second runtime beside torch's answered hipErrorNoDevice)."""
    import subprocess
    import sys
    code = ("import sys; sys.path.insert(0, %r)\n"
            "import jxpamg_all_amd as jx\n"
            "jx.lib()\n"
            "import torch\n"
            "maps = open('/proc/self/maps').read().splitlines()\n"
            "print(sorted(set(l.split()[-1] for l in maps if 'libamdhip64' in l)))\n") % ROOT
    r = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    import ast
    runtimes = ast.literal_eval(r.stdout.strip().splitlines()[-1])
    assert len(runtimes) == 1, runtimes


def test_solver_rejects_bad_input_without_gpu_calls():
    """Shape errors are caught on the host before any device work."""
    import ctypes
    A, rhs = jx.generate(nx=4, ny=4, nz=2, block_size=3, seed=1)
    Ad, _ = jx.decouple_abf(A, rhs)
    ia, ja, a = jx.extract_pressure(Ad)
    bad = jx.Hierarchy([(ia, ja, a, Ad.nb + 1)], [])  # wrong column count
    s = ctypes.c_void_p()
    p = jx.default_params()
    ac, hc = Ad.c_struct(), bad.c_struct()
    rc = jx.lib().jxamd_solver_create(ctypes.byref(s), ctypes.byref(ac), ctypes.byref(hc), ctypes.byref(p))
    assert rc == jx.ERROR_ARG
    ja2 = ja.copy()
    ja2[0] = ja2[1]  # diagonal no longer first
    bad2 = jx.Hierarchy([(ia, ja2, a, Ad.nb)], [])
    hc2 = bad2.c_struct()
    rc = jx.lib().jxamd_solver_create(ctypes.byref(s), ctypes.byref(ac), ctypes.byref(hc2), ctypes.byref(p))
    assert rc == jx.ERROR_ARG
    # out-of-range options are rejected before any device work too
    good = jx.Hierarchy([(ia, ja, a, Ad.nb)], []).c_struct()
    for field, value in [("solution_update", 2), ("arnoldi_kernel", -1), ("stage2_precision", 3),
                         ("poly_degree", 9), ("precond_precision", 3)]:
        pb = jx.default_params(**{field: value})
        rc = jx.lib().jxamd_solver_create(ctypes.byref(s), ctypes.byref(ac), ctypes.byref(good), ctypes.byref(pb))
        assert rc == jx.ERROR_ARG, field


@pytest.mark.parametrize("name", ALL)
def test_amg_setup_bit_exact_vs_reference(name):
    """The host hierarchy (strength, HMIS, ext+i, RAP) equals the reference's."""
    d, A, rhs, Ad, bd, H, meta = system(name)
    ia, ja, a = jx.extract_pressure(Ad)
    G = jx.amg_setup(ia, ja, a)
    assert G.num_levels == H.num_levels
    for l in range(G.num_levels):
        for x, y in [(G.A[l], H.A[l])] + ([(G.P[l], H.P[l])] if l < G.num_levels - 1 else []):
            assert x[3] == y[3]
            assert np.array_equal(x[0], y[0]) and np.array_equal(x[1], y[1]) and np.array_equal(x[2], y[2])


def test_solver_interface_rejects_bad_arguments_on_the_host():
    """Case-220 entry: unknown solver ids, block sizes and shapes fail before
    any device work; a 1-based CSR is left 0-based only once it is accepted."""
    A, rhs = jx.generate(nx=4, ny=4, nz=2, block_size=3, seed=1)
    S = A.to_scipy().tocsr()
    S.sort_indices()
    ia, ja, a = (S.indptr + 1).astype(np.int32), (S.indices + 1).astype(np.int32), S.data.copy()
    x = np.zeros(A.n)
    for sid, nf in [(22, 3), (220, 1), (220, 5), (220, 5 if A.n % 5 else 7)]:
        with pytest.raises(RuntimeError):
            jx.solver_interface(A.n, ia, ja, a, x, rhs.copy(), 1e-6, sid, nf)
    assert ia[0] == 1  # rejected before the in-place decrement


def test_amg_setup_rejects_diagonal_not_first():
    ia = np.array([0, 2, 4], np.int32)
    ja = np.array([1, 0, 1, 0], np.int32)
    a = np.array([-1.0, 2.0, 2.0, -1.0])
    with pytest.raises(RuntimeError):
        jx.amg_setup(ia, ja, a)


@pytest.mark.parametrize("name", ["c1", "b4", "b2", "nodec"])
def test_bilu0_factor_matches_oracle(name):
    from oracle.cpu_ref import Oracle
    from tests._common import bits, decoup_cases, decoup_system, system
    d, A, rhs, Ad, bd, H, meta = system(name)
    lu, dinv = jx.bilu0_factor(Ad)
    lu_o, dinv_o = Oracle(Ad).bilu0()
    assert np.array_equal(lu, lu_o) and np.array_equal(dinv, dinv_o)


def test_bilu0_factor_rejects_unordered_rows():
    A, rhs = jx.generate(nx=4, ny=4, nz=2, block_size=3, seed=1)
    ja = A.ja.copy()
    k0 = A.ia[5]
    ja[k0 + 1], ja[k0 + 2] = ja[k0 + 2], ja[k0 + 1]
    B = jx.BSRMatrix(A.nb, A.b, A.ia, ja, A.a)
    with pytest.raises(RuntimeError):
        jx.bilu0_factor(B)


def test_learned_root_trust_range_matches_the_label_family():
    """the clamp range of predicted roots (jxamd.h) is the candidate family
    the training labels are drawn from (learned.py LO, HI)"""
    learned = pytest.importorskip("jxpamg_all_amd.learned")
    hdr = open(os.path.join(ROOT, "include", "jxamd.h")).read()
    val = {k: float(re.search(rf"#define {k}\s+([0-9.]+)", hdr).group(1))
           for k in ("JXAMD_MLP_LO_MIN", "JXAMD_MLP_LO_MAX", "JXAMD_MLP_HI_MIN", "JXAMD_MLP_HI_MAX")}
    assert (min(learned.LO), max(learned.LO)) == (val["JXAMD_MLP_LO_MIN"], val["JXAMD_MLP_LO_MAX"])
    assert (min(learned.HI), max(learned.HI)) == (val["JXAMD_MLP_HI_MIN"], val["JXAMD_MLP_HI_MAX"])
