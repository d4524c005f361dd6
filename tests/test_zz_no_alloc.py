"""The allocation-free 2x2 renormalisation against the allocating entry point (same kernel,
bit for bit), on two calls in a row (the second smaller, so it runs over staging the first
call left behind), and its refusal of more frequencies than its fixed staging holds."""
import ctypes as C

import numpy as np
import pytest

from courant_amd import lib as L

pytestmark = pytest.mark.gpu


def _case(rng, nf):
    S = 0.6 * (rng.normal(size=(nf, 8))) / np.sqrt(2)
    z1 = rng.uniform(20, 90, nf)
    z2 = rng.uniform(20, 90, nf)
    return np.ascontiguousarray(S), z1, z2


def test_no_alloc_renormalize_matches_the_allocating_entry(cem, gpu_available):
    print("case: s2 renormalize no-alloc vs allocating, nf 801 then 37", flush=True)
    rng = np.random.default_rng(11)
    for nf in (801, 37):
        S, z1, z2 = _case(rng, nf)
        a, b = S.copy(), S.copy()
        assert cem.lib.cem_amd_s2_renormalize(L._dp(a), nf, L._dp(z1), L._dp(z2), 50.0) == 0, cem.last_error()
        assert cem.lib.cem_amd_s2_renormalize_noalloc(L._dp(b), nf, L._dp(z1), L._dp(z2), 50.0) == 0, \
            cem.last_error()
        assert np.array_equal(a, b)
        assert not np.array_equal(b, S)


def test_no_alloc_renormalize_refuses_oversize_input(cem, gpu_available):
    print("case: s2 renormalize no-alloc, nf above the fixed staging", flush=True)
    nf = 65537
    S = np.zeros((nf, 8))
    z = np.full(nf, 50.0)
    assert cem.lib.cem_amd_s2_renormalize_noalloc(L._dp(S), nf, L._dp(z), L._dp(z), 50.0) == -1
    assert "65536" in cem.last_error()
    assert cem.lib.cem_amd_s2_renormalize_noalloc(L._dp(S), 0, L._dp(z), L._dp(z), 50.0) == 0
