"""Pins the oracle (oracle/cpu_ref.c) two ways:

1. against the golden vectors written by the reference itself (FP64 build of
   jxfpamg-0220 driven like src/test_bsr_cprgmres.c): bit-exact;
2. against what the mathematics fixes, independently of the reference:
   dense brute force, triangular solves, LAPACK solves, GMRES invariants.
"""
import numpy as np
import pytest
import scipy.linalg as sla
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from oracle.cpu_ref import Oracle
from tests._common import ranks_system, relerr, system
from tests.conftest import load_fixture

OP_CASES = ["c1", "tiny", "b4", "b2", "lvl3", "nodec", "timpes"]


@pytest.mark.parametrize("name", OP_CASES)
def test_operators_bit_exact_vs_reference(name):
    d, A, rhs, Ad, bd, H, meta = system(name)
    if H is None:
        pytest.skip("fixture has no operator dumps")
    o = Oracle(Ad, H)
    assert np.array_equal(o.relax6(0, d["hsgs_f"], d["hsgs_u_in"]), d["hsgs_u_out"])
    assert np.array_equal(o.relax9(d["coarse_f"]), d["coarse_u"])
    assert np.array_equal(o.vcycle(d["vcycle_f"]), d["vcycle_u"])
    assert np.array_equal(o.bhgs(d["bhgs_x_in"], d["bhgs_b"], 0.7, 0.9), d["bhgs_x_out"])
    p = o.params(stage2_maxit=int(meta["stage2_maxit"]))
    assert np.array_equal(o.cpr(d["cpr_r"], p), d["cpr_x"])


@pytest.mark.parametrize("name", ["c1", "tiny", "b4", "b2", "lvl3", "nodec", "restart", "layered", "timpes"])
def test_gmres_matches_reference(name):
    d, A, rhs, Ad, bd, H, meta = system(name)
    if H is None:
        pytest.skip("needs the reference hierarchy")
    o = Oracle(Ad, H)
    p = o.params(k_dim=int(meta["k_dim"]), max_iter=int(meta["max_iter"]), tol=meta["rtol"],
                 stage2_maxit=int(meta["stage2_maxit"]))
    x, st, hist, rc = o.gmres(bd, p)
    assert rc == 0
    assert st.iterations == int(d["iterations"][0])
    assert np.array_equal(hist, d["res_history"])
    assert st.final_rel_res == d["final_rel_res"][0]
    if "x_final" in d:
        assert np.array_equal(x, d["x_final"])


def _small_pressure_system(n=40, seed=0):
    rng = np.random.default_rng(seed)
    M = sp.random(n, n, density=0.15, random_state=seed, format="csr")
    M = -abs(M) - abs(M.T)
    M = M.tolil()
    M.setdiag(0)
    M = M.tocsr()
    d = -np.asarray(M.sum(axis=1)).ravel() + 0.5 + rng.random(n)
    A = (M + sp.diags(d)).tocsr()
    # diagonal first in every row
    ia, ja, a = [0], [], []
    for i in range(n):
        row = A.getrow(i)
        cols = list(row.indices)
        vals = list(row.data)
        k = cols.index(i)
        cols = [cols[k]] + cols[:k] + cols[k + 1:]
        vals = [vals[k]] + vals[:k] + vals[k + 1:]
        ja += cols
        a += vals
        ia.append(len(ja))
    return A, (np.array(ia, np.int32), np.array(ja, np.int32), np.array(a), n)


def test_relax6_is_forward_then_backward_triangular_solve():
    import jxpamg_all_amd as jx
    A, csr = _small_pressure_system()
    n = A.shape[0]
    H = jx.Hierarchy([csr], [])
    # dummy BSR to build the oracle view
    Ab = jx.BSRMatrix(1, 2, np.array([0, 1], np.int32), np.array([0], np.int32), np.eye(2).ravel())
    o = Oracle(Ab, H)
    rng = np.random.default_rng(1)
    f, u0 = rng.standard_normal(n), rng.standard_normal(n)
    got = o.relax6(0, f, u0)
    L = sp.tril(A, 0).tocsr()
    U = sp.triu(A, 1).tocsr()
    u1 = spla.spsolve_triangular(L, f - U @ u0, lower=True)
    Lo = sp.tril(A, -1).tocsr()
    Ut = sp.triu(A, 0).tocsr()
    u2 = spla.spsolve_triangular(Ut, f - Lo @ u1, lower=False)
    assert relerr(got, u2) < 1e-13


@pytest.mark.parametrize("name,T", [("lvl3_t4", 4), ("tiny_t16", 16), ("tiny_t64", 64)])
def test_threaded_relax6_bit_exact_vs_reference(name, T):
    """OMP_NUM_THREADS=T reference runs: relax 6 splits every level into T
    contiguous chunks (par_relax_6.c:132-223); T = 64 on the 256-row fine
    level is the 4-row chunk of the 60x220x85 benchmark."""
    d, A, rhs, Ad, bd, H, meta = system(name)
    assert int(meta["threads"]) == T
    o = Oracle(Ad, H)
    assert np.array_equal(o.relax6_threads(0, d["hsgs_f"], d["hsgs_u_in"], T), d["hsgs_u_out"])
    # the one-thread sweep gives a different answer on the same inputs
    assert not np.array_equal(o.relax6(0, d["hsgs_f"], d["hsgs_u_in"]), d["hsgs_u_out"])
    o.set_relax_threads([T] * H.num_levels)
    assert np.array_equal(o.vcycle(d["vcycle_f"]), d["vcycle_u"])
    assert np.array_equal(o.cpr(d["cpr_r"]), d["cpr_x"])
    x, st, hist, rc = o.gmres(bd, o.params(tol=meta["rtol"]))
    # the threaded reference also sums its dot products in T parts
    assert st.iterations == int(d["iterations"][0])
    assert relerr(hist, d["res_history"]) < 1e-12


@pytest.mark.parametrize("T", [2, 3, 7, 40])
def test_threaded_relax6_is_block_jacobi_over_chunk_ssor(T):
    """T chunks: forward then backward Gauss-Seidel on each chunk's diagonal
    block, every coupling outside the chunk taken at the entry values; T = n
    reduces to one Jacobi step."""
    import jxpamg_all_amd as jx
    A, csr = _small_pressure_system()
    n = A.shape[0]
    Ab = jx.BSRMatrix(1, 2, np.array([0, 1], np.int32), np.array([0], np.int32), np.eye(2).ravel())
    o = Oracle(Ab, jx.Hierarchy([csr], []))
    rng = np.random.default_rng(T)
    f, u0 = rng.standard_normal(n), rng.standard_normal(n)
    got = o.relax6_threads(0, f, u0, T)
    size, rest = divmod(n, T)
    expect = np.empty(n)
    Ad = A.toarray()
    for c in range(T):
        ns = c * size + min(c, rest)
        ne = ns + size + (1 if c < rest else 0)
        B = Ad[ns:ne, ns:ne]
        g = f[ns:ne] - Ad[ns:ne, :ns] @ u0[:ns] - Ad[ns:ne, ne:] @ u0[ne:]
        u1 = sla.solve_triangular(np.tril(B), g - np.triu(B, 1) @ u0[ns:ne], lower=True)
        expect[ns:ne] = sla.solve_triangular(np.triu(B), g - np.tril(B, -1) @ u1, lower=False)
    assert relerr(got, expect) < 1e-12
    jac = (f - (A - sp.diags(A.diagonal())) @ u0) / A.diagonal()
    assert relerr(o.relax6_threads(0, f, u0, n), jac) < 1e-14


def test_gselim_solves_diagonally_dominant_systems():
    import ctypes
    from oracle.cpu_ref import lib
    rng = np.random.default_rng(2)
    n = 30
    M = rng.standard_normal((n, n)) + n * np.eye(n)
    b = rng.standard_normal(n)
    Mc, x = M.copy(), b.copy()
    assert lib().cref_gselim(ctypes.c_void_p(Mc.ctypes.data), ctypes.c_void_p(x.ctypes.data), n) == 0
    assert relerr(x, np.linalg.solve(M, b)) < 1e-13


@pytest.mark.parametrize("b", [2, 3, 4])
def test_block_kernels(b):
    import ctypes
    from oracle.cpu_ref import lib
    rng = np.random.default_rng(b)
    for trial in range(20):
        M = rng.standard_normal((b, b))
        if trial % 3 == 0:
            M[0, 0] = 1e-3  # forces a row exchange in the pivoted LU
        v = rng.standard_normal(b)
        out = np.zeros(b)
        lib().cref_bsr_block_invmatvec_stable(ctypes.c_void_p(M.ctypes.data), ctypes.c_void_p(v.ctypes.data),
                                              ctypes.c_void_p(out.ctypes.data), b)
        assert relerr(out, np.linalg.solve(M, v)) < 1e-10
        ov = rng.standard_normal(b)
        ref = ov - M @ v
        lib().cref_bsr_block_matvec_stable(-1.0, ctypes.c_void_p(M.ctypes.data), ctypes.c_void_p(v.ctypes.data),
                                           1.0, ctypes.c_void_p(ov.ctypes.data), b)
        assert relerr(ov, ref) < 1e-14


@pytest.mark.parametrize("name", ["tiny", "b4", "b2"])
def test_bsr_matvec_vs_scipy(name):
    d, A, rhs, Ad, bd, H, meta = system(name)
    o = Oracle(Ad)
    S = Ad.to_scipy()
    rng = np.random.default_rng(3)
    x, y = rng.standard_normal(Ad.n), rng.standard_normal(Ad.n)
    assert relerr(o.matvec(1.0, x, 0.0, y), S @ x) < 1e-13
    assert relerr(o.matvec(-1.0, x, 1.0, y), y - S @ x) < 1e-13
    # A * 1 = block row sums
    one = np.ones(Ad.n)
    sums = Ad.a.reshape(-1, Ad.b, Ad.b).sum(axis=2)
    rows = np.repeat(np.arange(Ad.nb), np.diff(Ad.ia))
    expect = np.zeros((Ad.nb, Ad.b))
    np.add.at(expect, rows, sums)
    assert relerr(o.matvec(1.0, one, 0.0, one), expect.ravel()) < 1e-13


@pytest.mark.parametrize("name", ["tiny", "b4", "b2"])
def test_bhgs_unit_weights_is_block_lower_solve(name):
    """w = omega = 1: one forward sweep equals x1 = (D + L)^{-1} (b - U x0)."""
    d, A, rhs, Ad, bd, H, meta = system(name)
    if not int(meta["decouple"]):
        pytest.skip()
    o = Oracle(Ad)
    S = Ad.to_scipy().tocsr()
    b = Ad.b
    rng = np.random.default_rng(4)
    x0, r = rng.standard_normal(Ad.n), rng.standard_normal(Ad.n)
    got = o.bhgs(x0, r, 1.0, 1.0)
    blk_row = np.arange(Ad.n) // b
    S = S.tocoo()
    low = blk_row[S.col] <= blk_row[S.row]
    DL = sp.csr_matrix((S.data[low], (S.row[low], S.col[low])), shape=S.shape)
    Up = sp.csr_matrix((S.data[~low], (S.row[~low], S.col[~low])), shape=S.shape)
    expect = spla.spsolve_triangular(DL, r - Up @ x0, lower=True)
    assert relerr(got, expect) < 1e-12


def test_gmres_invariants_and_true_residual():
    d, A, rhs, Ad, bd, H, meta = system("c1")
    o = Oracle(Ad, H)
    x, st, hist, rc = o.gmres(bd, o.params(tol=1e-8))
    S = Ad.to_scipy()
    true = np.linalg.norm(bd - S @ x) / np.linalg.norm(bd)
    assert st.converged == 1 and true <= 1e-8
    # within one restart cycle the GMRES residual estimate never increases
    assert np.all(np.diff(hist) <= 1e-15 * hist[0])


def test_vcycle_is_linear_and_reduces_error():
    d, A, rhs, Ad, bd, H, meta = system("lvl3")
    o = Oracle(Ad, H)
    rng = np.random.default_rng(5)
    f1, f2 = rng.standard_normal(H.A[0][0].size - 1), rng.standard_normal(H.A[0][0].size - 1)
    u1, u2, u12 = o.vcycle(f1), o.vcycle(f2), o.vcycle(2.0 * f1 - 3.0 * f2)
    assert relerr(u12, 2.0 * u1 - 3.0 * u2) < 1e-10
    ia, ja, a, _ = H.A[0]
    Ap = sp.csr_matrix((a, ja, ia))
    e = np.linalg.norm(f1 - Ap @ u1) / np.linalg.norm(f1)
    assert e < 0.5


def _bsr_parts(A, lu):
    """(L unit block lower, U block upper incl. diagonal) as scipy CSR."""
    b = A.b
    rows = np.repeat(np.arange(A.nb), np.diff(A.ia))
    blocks = lu.reshape(-1, b, b)
    lower = A.ja < rows
    Lb = sp.bsr_matrix((blocks[lower], A.ja[lower], np.concatenate([[0], np.cumsum(np.bincount(rows[lower],
                        minlength=A.nb))])), shape=(A.n, A.n)).tocsr() + sp.identity(A.n, format="csr")
    up = ~lower
    Ub = sp.bsr_matrix((blocks[up], A.ja[up], np.concatenate([[0], np.cumsum(np.bincount(rows[up],
                        minlength=A.nb))])), shape=(A.n, A.n)).tocsr()
    return Lb, Ub


def test_bilu0_is_exact_lu_without_fill():
    """A block-tridiagonal matrix has no fill: BILU(0) is its exact block LU,
    so one BILU(0) solve is A^{-1} r."""
    import jxpamg_all_amd as jx
    rng = np.random.default_rng(3)
    nb, b = 40, 3
    ia, ja, blocks = [0], [], []
    for i in range(nb):
        cols = [i] + [j for j in (i - 1, i + 1) if 0 <= j < nb]
        for j in cols:
            blk = rng.uniform(-1, 1, (b, b))
            if j == i:
                blk += 8 * np.eye(b)
            ja.append(j)
            blocks.append(blk)
        ia.append(len(ja))
    A = jx.BSRMatrix(nb, b, np.array(ia), np.array(ja), np.array(blocks).ravel())
    o = Oracle(A)
    r = rng.standard_normal(A.n)
    z = o.bilu0_solve(r)
    assert relerr(z, spla.spsolve(A.to_scipy().tocsc(), r)) < 1e-12


@pytest.mark.parametrize("name", ["c1", "b4", "b2"])
def test_bilu0_product_matches_a_on_its_pattern(name):
    """ILU(0) defining property: (L U)_ij = A_ij on every stored block."""
    d, A, rhs, Ad, bd, H, meta = system(name)
    As = Ad  # diagonal block first, then ascending block columns
    o = Oracle(As)
    lu, dinv = o.bilu0()
    Lb, Ub = _bsr_parts(As, lu)
    prod = (Lb @ Ub).tocsr()
    S = As.to_scipy().tocsr()
    mask = S.copy()
    mask.data[:] = 1.0
    on_pattern = prod.multiply(mask)
    assert abs(on_pattern - S).max() <= 1e-12 * abs(S).max()
    # dinv holds the inverted diagonal blocks of U
    b = As.b
    Ud = Ub.toarray()
    for i in range(0, As.nb, max(1, As.nb // 17)):
        D = Ud[i * b:(i + 1) * b, i * b:(i + 1) * b]
        assert np.allclose(dinv.reshape(-1, b, b)[i] @ D, np.eye(b), atol=1e-12)
    # the solve applies (L U)^{-1}
    r = np.random.default_rng(1).standard_normal(As.n)
    z = o.bilu0_solve(r)
    assert relerr(Lb @ (Ub @ z), r) < 1e-12


def test_cpr_with_bilu0_stage2_converges():
    d, A, rhs, Ad, bd, H, meta = system("lvl3")
    o = Oracle(Ad, H)
    x, st, hist, rc = o.gmres(bd, o.params(tol=1e-8, stage2_maxit=1, stage2_solver_type=5))
    true = np.linalg.norm(bd - Ad.to_scipy() @ x) / np.linalg.norm(bd)
    assert rc == 0 and st.converged == 1 and true <= 1e-8


def test_poly_smoother_error_propagation():
    """u <- u + w_k D^{-1}(f - A u): the error e = u - A^{-1} f propagates by
    prod_k (I - w_k D^{-1} A), checked with dense matrices."""
    import jxpamg_all_amd as jx
    A, csr = _small_pressure_system()
    n = A.shape[0]
    Ab = jx.BSRMatrix(1, 2, np.array([0, 1], np.int32), np.array([0], np.int32), np.eye(2).ravel())
    o = Oracle(Ab, jx.Hierarchy([csr], []))
    rng = np.random.default_rng(9)
    f, u0 = rng.standard_normal(n), rng.standard_normal(n)
    w = np.array([0.9, 0.35, 1.4, 0.6])
    got = o.poly_smooth(0, f, u0, w)
    Ad = A.toarray()
    Dinv = np.diag(1.0 / np.diag(Ad))
    xs = np.linalg.solve(Ad, f)
    E = np.eye(n)
    for wk in w:
        E = (np.eye(n) - wk * Dinv @ Ad) @ E
    assert relerr(got - xs, E @ (u0 - xs)) < 1e-10


def test_vcycle_with_poly_smoother_is_linear_and_converges():
    d, A, rhs, Ad, bd, H, meta = system("lvl3")
    o = Oracle(Ad, H)
    o.set_poly([[0.6, 1.2, 0.7, 1.0]] * H.num_levels)
    rng = np.random.default_rng(6)
    n = H.A[0][0].size - 1
    f1, f2 = rng.standard_normal(n), rng.standard_normal(n)
    assert relerr(o.vcycle(2.0 * f1 - 3.0 * f2), 2.0 * o.vcycle(f1) - 3.0 * o.vcycle(f2)) < 1e-10
    x, st, hist, rc = o.gmres(bd, o.params(tol=1e-8))
    assert st.converged == 1 and np.linalg.norm(bd - Ad.to_scipy() @ x) / np.linalg.norm(bd) <= 1e-8


@pytest.mark.parametrize("nout", [4, 2])
def test_mlp_checker_gradient_matches_finite_differences(nout):
    """The checker's backward pass (BF16-rounded operands) against central
    differences of the unrounded loss: agreement to the BF16 rounding level
    pins the transposes, signs and tanh derivatives; with nout < 4 the unused
    outputs' weights get exactly zero gradient."""
    from oracle import mlp_ref as M
    rng = np.random.default_rng(0)
    P = rng.normal(0, 0.2, M.NPARAM)
    X, T = rng.normal(0, 1, (20, 16)), rng.normal(0, 1, (20, nout))
    _, g = M.loss_and_grad(P, X, T, nout)
    idx = rng.choice(M.NPARAM, 40, replace=False)
    fd = []
    for i in idx:
        e = np.zeros(M.NPARAM)
        e[i] = 1e-5
        fd.append((M.loss_exact(P + e, X, T, nout) - M.loss_exact(P - e, X, T, nout)) / 2e-5)
    fd = np.array(fd)
    assert np.max(np.abs(fd - g[idx])) < 0.03 * np.max(np.abs(fd))
    assert np.corrcoef(fd, g[idx])[0, 1] > 0.999
    p = M.unpack(g)
    assert not np.any(p["W3"][nout:]) and not np.any(p["b3"][nout:])


@pytest.mark.parametrize("name", ["tiny", "b4", "b2"])
def test_stage2_rank_partitions(name):
    """Stage 2 as run on several ranks (GS inside a rank's contiguous rows,
    other ranks' neighbours at entry values): one partition is the one-rank
    sweep; one partition per row is omega-damped block Jacobi; with
    w = omega = 1 each partition is a block lower-triangular solve with the
    couplings to other partitions on the right-hand side."""
    d, A, rhs, Ad, bd, H, meta = system(name)
    o = Oracle(Ad)
    rng = np.random.default_rng(2)
    x0, r = rng.standard_normal(Ad.n), rng.standard_normal(Ad.n)
    assert np.array_equal(o.bhgs_parts(x0, r, 1), o.bhgs(x0, r))
    S = Ad.to_scipy().tocsr()
    b, nb = Ad.b, Ad.nb
    Dblk = sp.block_diag([S[i * b:(i + 1) * b, i * b:(i + 1) * b].toarray() for i in range(nb)]).tocsr()
    w, om = 0.7, 0.9
    jac = (1 - w * om) * x0 + w * om * spla.spsolve(Dblk.tocsc(), r - (S - Dblk) @ x0)
    assert relerr(o.bhgs_parts(x0, r, nb), jac) < 1e-12
    nparts = 3
    got = o.bhgs_parts(x0, r, nparts, w=1.0, omega=1.0)
    size, rest = divmod(nb, nparts)
    blk_row = np.repeat(np.arange(nb), b)
    for t in range(nparts):
        ns = t * size + min(t, rest)
        ne = ns + size + (1 if t < rest else 0)
        sl = slice(ns * b, ne * b)
        Spp = S[sl, sl].tocoo()
        low = blk_row[ns * b + Spp.col] <= blk_row[ns * b + Spp.row]
        DL = sp.csr_matrix((Spp.data[low], (Spp.row[low], Spp.col[low])), shape=Spp.shape)
        Up = sp.csr_matrix((Spp.data[~low], (Spp.row[~low], Spp.col[~low])), shape=Spp.shape)
        outside = r[sl] - S[sl, :] @ x0 + S[sl, sl] @ x0[sl]
        expect = spla.spsolve_triangular(DL, outside - Up @ x0[sl], lower=True)
        assert relerr(got[sl], expect) < 1e-11


@pytest.mark.parametrize("b", [2, 3, 4])
def test_partitioned_stage2_bit_exact_vs_reference_on_mpi_ranks(b):
    """The oracle's partitioned stage-2 sweep (cref_bsr_hgs_parts) against
    the reference itself run on P MPI ranks (jxf_ParBSRHGSRelax with the
    matrix split over the ranks, ranks.npz written by oracle/ref_bhgs_mpi
    under mpirun): bit for bit for every P from 1 to 24 on 24 block rows,
    including uneven splits (P = 5) and one row per rank (P = 24)."""
    d = load_fixture("ranks")
    A = ranks_system(d, b)
    o = Oracle(A)
    for P in d["ranks"]:
        got = o.bhgs_parts(d[f"b{b}_x0"], d[f"b{b}_rhs"], int(P))
        assert np.array_equal(got, d[f"b{b}_x_p{int(P)}"]), f"P={P}"
