"""Shared helpers for the test modules: rebuild a fixture's system."""
import numpy as np

import jxpamg_all_amd as jx
from tests.conftest import fixture_gen_params, fixture_meta, load_fixture


def system(name):
    """(fixture dict, raw A, raw rhs, decoupled A, decoupled rhs, hierarchy or None, meta)"""
    d = load_fixture(name)
    gp = fixture_gen_params(d)
    A, rhs = jx.generate(**gp)
    meta = fixture_meta(d)
    if int(meta["decouple"]):
        Ad, bd = jx.decouple(A, rhs, int(meta["decouple"]))
    else:
        Ad, bd = A, rhs.copy()
    H = jx.Hierarchy.from_fixture(d) if "A0_ia" in d else None
    return d, A, rhs, Ad, bd, H, meta


def decoup_cases():
    """(key, block size, decoupling type, is_thermal) of every run in decoup.npz"""
    d = load_fixture("decoup")
    out = []
    for k in sorted(d):
        if k.startswith("b") and k.endswith("_a") and "_t" in k:
            b, t, th = k[:-2].split("_")
            out.append((k[:-2], int(b[1:]), int(t[1:]), int(th[2:])))
    return out


def decoup_system(d, b):
    """the undecoupled system of block size b in decoup.npz"""
    keys = [str(k) for k in d[f"b{b}_gen_keys"]]
    gp = {k: (float(v) if k in ("sigma_lnk", "corr_len", "channel_ratio", "kz_ratio", "well_boost") else int(v))
          for k, v in zip(keys, d[f"b{b}_gen_values"])}
    return jx.generate(**gp)


def bits(x):
    return np.ascontiguousarray(x, dtype=np.float64).view(np.uint64)


def relerr(x, y):
    x, y = np.asarray(x), np.asarray(y)
    return float(np.max(np.abs(x - y)) / max(np.max(np.abs(y)), 1e-300))


def ranks_system(d, b):
    """the block-size-b system of ranks.npz (ABF-decoupled for b < 4), its
    matrix checked against the fixture's SHA-256"""
    import hashlib
    A, rhs = decoup_system(d, b)
    A = jx.decouple_abf(A, rhs)[0] if b < 4 else A
    assert hashlib.sha256(np.ascontiguousarray(A.a).tobytes()).hexdigest() == bytes(d[f"b{b}_a_sha256"]).decode()
    return A
