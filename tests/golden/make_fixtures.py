"""Writes the golden fixtures under tests/golden/ by running the reference.

The reference library is compiled from its own sources under
/root/reference in FP64 (`make -C oracle ref`) and driven by
oracle/ref_driver.c, which replays src/test_bsr_cprgmres.c with fixture
dumps.  Run this only where /root/reference exists; the outputs (inputs,
expected outputs and the reference's AMG hierarchy) are committed.

    python tests/golden/make_fixtures.py            # solver fixtures
    python tests/golden/make_fixtures.py decoup     # decoupling types
    python tests/golden/make_fixtures.py ranks      # stage 2 on P MPI ranks
"""
import hashlib
import os
import struct
import subprocess
import sys
import tempfile

import numpy as np

ROOT = os.path.abspath(os.path.join(os.path.dirname(__file__), "..", ".."))
sys.path.insert(0, ROOT)
import jxpamg_all_amd as jx  # noqa: E402

DRIVER = os.path.join(ROOT, "oracle", "_ref", "ref_driver")
OUT = os.path.dirname(os.path.abspath(__file__))

# name: (generator params, rtol, max_iter, stage2_maxit, decouple, dump_level, k_dim)
CASES = {
    "c1": (dict(jx.CONFIGS["C1"]), 1e-6, 100, 3, 1, 2, 30),
    "tiny": (dict(jx.CONFIGS["C1"], nx=8, ny=8, nz=4, seed=7), 1e-6, 100, 3, 1, 2, 30),
    "b4": (dict(jx.CONFIGS["C1"], nx=8, ny=8, nz=6, block_size=4, field=jx.FIELD_SMOOTH, sigma_lnk=2.0,
                corr_len=2.0, seed=4), 1e-8, 100, 3, 1, 2, 30),
    "b2": (dict(jx.CONFIGS["C1"], nx=12, ny=12, nz=6, block_size=2, seed=5), 1e-6, 100, 3, 1, 2, 30),
    "lvl3": (dict(jx.CONFIGS["C1"], nx=20, ny=20, nz=10, seed=3), 1e-6, 100, 3, 1, 1, 30),
    "nodec": (dict(jx.CONFIGS["C1"]), 1e-6, 100, 1, 0, 1, 30),
    "restart": (dict(jx.CONFIGS["C1"], seed=11), 1e-10, 60, 3, 1, 1, 3),
    "layered": (dict(jx.CONFIGS["C2"], nx=12, ny=20, nz=12, n_smooth_layers=5, seed=21), 1e-6, 100, 3, 1, 1, 30),
    # the full 60x220x85 system: scalars and residual history only
    "c2": (dict(jx.CONFIGS["C2"]), 1e-6, 100, 3, 1, 0, 30),
    # lvl3 with OMP_NUM_THREADS=4: the threaded relax-6 (par_relax_6.c) golden vectors
    "lvl3_t4": (dict(jx.CONFIGS["C1"], nx=20, ny=20, nz=10, seed=3), 1e-6, 100, 3, 1, 1, 30),
    # tiny with OMP_NUM_THREADS=16 and 64: relax-6 chunks of 16 and of 4 rows
    # on the fine level (256 rows), the chunk size the 60x220x85 benchmark runs
    "tiny_t16": (dict(jx.CONFIGS["C1"], nx=8, ny=8, nz=4, seed=7), 1e-6, 100, 3, 1, 1, 30),
    "tiny_t64": (dict(jx.CONFIGS["C1"], nx=8, ny=8, nz=4, seed=7), 1e-6, 100, 3, 1, 1, 30),
    # True-IMPES decoupling (type 5) before the solve, as the reference's hybrid drivers use it
    "timpes": (dict(jx.CONFIGS["C1"], nx=10, ny=10, nz=6, seed=8), 1e-6, 100, 3, 5, 2, 30),
    # the full 60x220x85 system decoupled by ABF and by True-IMPES (SHA-256 of
    # the decoupled system only; the reference's True-IMPES solve of this
    # system makes no progress: 100 iterations, relative residual 0.9997)
    "c2_dec_abf": (dict(jx.CONFIGS["C2"]), 1e-6, 100, 3, 1, 3, 30),
    "c2_dec_timpes": (dict(jx.CONFIGS["C2"]), 1e-6, 100, 3, 5, 3, 30),
}
# decoupling-only fixture (decoup.npz): every type on small systems of each
# block size, the thermal variants of ANL/SEM/QI, and True-IMPES on a block
# pattern that is not structurally symmetric
DECOUP_GRIDS = {
    2: dict(jx.CONFIGS["C1"], nx=4, ny=4, nz=3, block_size=2, seed=31),
    3: dict(jx.CONFIGS["C1"], nx=4, ny=4, nz=3, block_size=3, seed=32),
    4: dict(jx.CONFIGS["C1"], nx=4, ny=3, nz=3, block_size=4, field=jx.FIELD_SMOOTH, sigma_lnk=2.0, corr_len=2.0,
            seed=33),
}
DECOUP_RUNS = [(b, t, th) for b in (2, 3, 4) for t in range(1, 7) for th in ((0, 1) if t in (2, 3, 4) else (0,))]
# multi-rank stage-2 fixture (ranks.npz): one forward block hybrid SOR sweep
# of the reference on P MPI ranks (oracle/ref_bhgs_mpi under mpirun), 24 block
# rows, so P ranks own 24 / P rows each as the GPU's stage2_chunk = 24 / P
# partitions do; b = 2, 3 ABF-decoupled, b = 4 as generated
RANKS_GRIDS = {
    2: dict(jx.CONFIGS["C1"], nx=4, ny=3, nz=2, block_size=2, seed=41),
    3: dict(jx.CONFIGS["C1"], nx=4, ny=3, nz=2, block_size=3, seed=42),
    4: dict(jx.CONFIGS["C1"], nx=4, ny=3, nz=2, block_size=4, field=jx.FIELD_SMOOTH, sigma_lnk=2.0, corr_len=2.0,
            seed=43),
}
RANKS = (1, 2, 3, 4, 5, 6, 8, 12, 24)
MPI_DRIVER = os.path.join(ROOT, "oracle", "_ref", "ref_bhgs_mpi")
MPIRUN = "/opt/conda/bin/mpirun"
# OpenMP threads of the reference run (1 unless listed)
THREADS = {"lvl3_t4": 4, "tiny_t16": 16, "tiny_t64": 64}
# operator vectors not kept for these cases (the hierarchy and the solve are)
SOLVE_ONLY = {"restart", "layered"}
OPERATOR_KEYS = ("hsgs_", "coarse_", "vcycle_", "bhgs_", "cpr_", "CF")


def read_dump(path):
    out = {}
    with open(path, "rb") as f:
        data = f.read()
    p = 0
    while p < len(data):
        (ln,) = struct.unpack_from("<i", data, p)
        p += 4
        name = data[p:p + ln].decode()
        p += ln
        kind, cnt = struct.unpack_from("<iq", data, p)
        p += 12
        dt = np.int32 if kind == 0 else np.float64
        out[name] = np.frombuffer(data, dtype=dt, count=cnt, offset=p).copy()
        p += cnt * np.dtype(dt).itemsize
    return out


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def main():
    only = set(sys.argv[1:])
    if not os.path.exists(DRIVER):
        sys.exit("build the reference first: make -C oracle ref")
    for name, (gp, rtol, max_iter, s2, dec, lvl, kdim) in CASES.items():
        if name == "decoup":
            continue
        if only and name not in only:
            continue
        A, rhs = jx.generate(**gp)
        tmp = tempfile.mkdtemp()
        jx.write_bsr_binary(os.path.join(tmp, "A.bin"), A)
        jx.write_rhs_text(os.path.join(tmp, "b.rhs"), rhs)
        dump = os.path.join(tmp, "o.dump")
        env = dict(os.environ, OMP_NUM_THREADS=str(THREADS.get(name, 1)))
        r = subprocess.run([DRIVER, os.path.join(tmp, "A.bin"), os.path.join(tmp, "b.rhs"), dump, repr(rtol),
                            str(max_iter), str(s2), str(dec), str(lvl), str(kdim)], capture_output=True, text=True,
                           env=env, check=True)
        d = read_dump(dump)
        meta = dict(rtol=rtol, max_iter=max_iter, stage2_maxit=s2, decouple=dec, k_dim=kdim,
                    threads=THREADS.get(name, 1))
        rec = {k: v for k, v in d.items()}
        if name in SOLVE_ONLY:
            rec = {k: v for k, v in rec.items() if not k.startswith(OPERATOR_KEYS)}
        if "dec_a" in rec and A.nnzb * A.b * A.b > 50000:
            rec["dec_a_sha256"] = np.frombuffer(sha(rec.pop("dec_a")).encode(), dtype=np.uint8)
        if "dec_b" in rec and A.nb * A.b > 50000:
            rec["dec_b_sha256"] = np.frombuffer(sha(rec.pop("dec_b")).encode(), dtype=np.uint8)
        rec["gen_" + "keys"] = np.array(list(gp.keys()))
        rec["gen_values"] = np.array([float(v) for v in gp.values()])
        rec["meta_keys"] = np.array(list(meta.keys()))
        rec["meta_values"] = np.array([float(v) for v in meta.values()])
        rec["input_a_sha256"] = np.frombuffer(sha(A.a).encode(), dtype=np.uint8)
        rec["input_rhs_sha256"] = np.frombuffer(sha(rhs).encode(), dtype=np.uint8)
        np.savez_compressed(os.path.join(OUT, f"{name}.npz"), **rec)
        print(name, (r.stderr.strip().splitlines() or ["decoupled"])[-1])


def asymmetric(A, rng):
    """A with about a fifth of its off-diagonal blocks dropped (the diagonal
    block stays first), so the block pattern is not structurally symmetric"""
    ia, ja, a, bb = [0], [], [], A.b * A.b
    for i in range(A.nb):
        for k in range(A.ia[i], A.ia[i + 1]):
            if k == A.ia[i] or rng.random() > 0.2:
                ja.append(A.ja[k])
                a.append(A.a[k * bb:(k + 1) * bb])
        ia.append(len(ja))
    return jx.BSRMatrix(A.nb, A.b, np.array(ia), np.array(ja), np.concatenate(a))


def run_decouple(A, rhs, dtype, thermal):
    tmp = tempfile.mkdtemp()
    jx.write_bsr_binary(os.path.join(tmp, "A.bin"), A)
    jx.write_rhs_text(os.path.join(tmp, "b.rhs"), rhs)
    dump = os.path.join(tmp, "o.dump")
    subprocess.run([DRIVER, os.path.join(tmp, "A.bin"), os.path.join(tmp, "b.rhs"), dump, "1e-6", "1", "1",
                    str(dtype), "3", "30", str(thermal)], capture_output=True, text=True,
                   env=dict(os.environ, OMP_NUM_THREADS="1"), check=True)
    d = read_dump(dump)
    return d["dec_a"], d["dec_b"]


def decoup_fixture():
    rec = {}
    for b, gp in DECOUP_GRIDS.items():
        rec[f"b{b}_gen_keys"] = np.array(list(gp.keys()))
        rec[f"b{b}_gen_values"] = np.array([float(v) for v in gp.values()])
    for b, t, th in DECOUP_RUNS:
        A, rhs = jx.generate(**DECOUP_GRIDS[b])
        rec[f"b{b}_t{t}_th{th}_a"], rec[f"b{b}_t{t}_th{th}_b"] = run_decouple(A, rhs, t, th)
    # True-IMPES on a pattern that is not structurally symmetric (host path only)
    A, rhs = jx.generate(**DECOUP_GRIDS[3])
    As = asymmetric(A, np.random.default_rng(7))
    rec["asym_ia"], rec["asym_ja"], rec["asym_a"] = As.ia, As.ja, As.a
    for t in (5, 6):
        rec[f"asym_t{t}_a"], rec[f"asym_t{t}_b"] = run_decouple(As, rhs, t, 0)
    np.savez_compressed(os.path.join(OUT, "decoup.npz"), **rec)
    print("decoup", len(DECOUP_RUNS) + 2, "runs")


def ranks_system(b):
    """the fixture's system of block size b: generated, ABF-decoupled for b < 4"""
    A, rhs = jx.generate(**RANKS_GRIDS[b])
    return jx.decouple_abf(A, rhs) if b < 4 else (A, rhs)


def ranks_fixture():
    rec = {"ranks": np.array(RANKS)}
    for b, gp in RANKS_GRIDS.items():
        rec[f"b{b}_gen_keys"] = np.array(list(gp.keys()))
        rec[f"b{b}_gen_values"] = np.array([float(v) for v in gp.values()])
        A, _ = ranks_system(b)
        rng = np.random.default_rng(100 + b)
        x0, rhs = rng.standard_normal(A.n), rng.standard_normal(A.n)
        rec[f"b{b}_x0"], rec[f"b{b}_rhs"] = x0, rhs
        rec[f"b{b}_a_sha256"] = np.frombuffer(sha(A.a).encode(), dtype=np.uint8)
        with tempfile.TemporaryDirectory() as tmp:
            fa, fx, fb, fo = (os.path.join(tmp, n) for n in ("A.bin", "x0.txt", "b.txt", "out.txt"))
            jx.write_bsr_binary(fa, A)
            jx.write_rhs_text(fx, x0)
            jx.write_rhs_text(fb, rhs)
            for P in RANKS:
                subprocess.run([MPIRUN, "-np", str(P), MPI_DRIVER, fa, fx, fb, fo], capture_output=True, text=True,
                               env=dict(os.environ, OMP_NUM_THREADS="1"), check=True, timeout=300)
                out = jx.read_rhs_text(fo, A.n)
                rec[f"b{b}_x_p{P}"] = out
    np.savez_compressed(os.path.join(OUT, "ranks.npz"), **rec)
    print("ranks", len(RANKS_GRIDS) * len(RANKS), "runs")


if __name__ == "__main__":
    if sys.argv[1:] == ["decoup"]:
        decoup_fixture()
    elif sys.argv[1:] == ["ranks"]:
        ranks_fixture()
    else:
        main()
