"""Writes the golden fixtures under tests/golden/ by running the reference.

The reference library is compiled from its own sources under
/root/reference in FP64 (`make -C oracle ref`) and driven by
oracle/ref_driver.c, which replays src/test_bsr_cprgmres.c with fixture
dumps.  Run this only where /root/reference exists; the outputs (inputs,
expected outputs and the reference's AMG hierarchy) are committed.

    python tests/golden/make_fixtures.py
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
}
# OpenMP threads of the reference run (1 unless listed)
THREADS = {"lvl3_t4": 4}
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
        rec["gen_" + "keys"] = np.array(list(gp.keys()))
        rec["gen_values"] = np.array([float(v) for v in gp.values()])
        rec["meta_keys"] = np.array(list(meta.keys()))
        rec["meta_values"] = np.array([float(v) for v in meta.values()])
        rec["input_a_sha256"] = np.frombuffer(sha(A.a).encode(), dtype=np.uint8)
        rec["input_rhs_sha256"] = np.frombuffer(sha(rhs).encode(), dtype=np.uint8)
        np.savez_compressed(os.path.join(OUT, f"{name}.npz"), **rec)
        print(name, r.stderr.strip().splitlines()[-1])


if __name__ == "__main__":
    main()
