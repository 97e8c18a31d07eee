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
        Ad, bd = jx.decouple_abf(A, rhs)
    else:
        Ad, bd = A, rhs.copy()
    H = jx.Hierarchy.from_fixture(d) if "A0_ia" in d else None
    return d, A, rhs, Ad, bd, H, meta


def relerr(x, y):
    x, y = np.asarray(x), np.asarray(y)
    return float(np.max(np.abs(x - y)) / max(np.max(np.abs(y)), 1e-300))
