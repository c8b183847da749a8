"""The allocation-free J/K (tcg_jk_stored_noalloc) against tcg_jk_stored
(eri == NULL, the resident tensor) through the C ABI: two calls in a row on
one GPU, the second on a smaller system, both storage orders; the size limit
of its static scratch array is refused with an error, not a fault."""
import numpy as np
import pytest

from tests_support import normalised_coef
from william_ci_amd import system as tsys

pytestmark = pytest.mark.gpu


def _one(s, sao, physicist, seed):
    from william_ci_amd.engine import Engine
    nb = s.nbas
    n1 = s.nfuc if sao else nb
    eng = Engine(s, normalised_coef(s))
    try:
        eng.eri_build(sao=sao, physicist=physicist, to_host=False)
        rng = np.random.default_rng(seed)
        P = rng.standard_normal((nb, nb))
        P = P + P.T
        Jn, Kn = eng.jk_stored_noalloc(P, n1=n1)
        J, K = eng.jk_stored(P, n1=n1)
        Jn2, Kn2 = eng.jk_stored_noalloc(P, n1=n1)        # again: the scratch keeps nothing over
    finally:
        eng.close()
    scale = max(np.abs(J).max(), np.abs(K).max())
    for A, B in ((Jn, J), (Kn, K), (Jn2, J), (Kn2, K)):
        assert np.abs(A - B).max() <= 1e-12 * scale
    if n1 < nb:
        assert not Jn[n1:].any() and not Kn[n1:].any()
    return np.abs(Jn - J).max(), np.abs(Kn - K).max()


@pytest.mark.parametrize("physicist", [False, True])
def test_noalloc_matches_jk_stored_big_then_small(physicist):
    big = tsys.chain_1d(32, 1.8, 57.6)                    # 160 functions
    small = tsys.workload("h10_ring")                      # 50 functions
    for name, s, sao in (("h32 SAO", big, True), ("h10 full", small, False)):
        dj, dk = _one(s, sao, physicist, 5)
        print(f"{name} physicist={physicist}: max |dJ| {dj:.2e} |dK| {dk:.2e}", flush=True)


def test_noalloc_refuses_oversize_and_missing_tensor():
    from william_ci_amd.engine import Engine, EngineError
    s = tsys.workload("h10_ring")
    eng = Engine(s, normalised_coef(s))
    try:
        P = np.eye(s.nbas)
        with pytest.raises(EngineError, match="no resident ERI"):
            eng.jk_stored_noalloc(P)
        eng.eri_build(sao=True, to_host=False)
        with pytest.raises(EngineError, match="no resident ERI"):
            eng.jk_stored_noalloc(P, n1=s.nbas)           # the resident tensor is the SAO one
    finally:
        eng.close()


def test_noalloc_size_limit_is_an_error():
    """nbas^2 + 2 n1 nbas above the scratch array's 4194304 doubles: refused
    before any device work (arguments checked on the host)."""
    import ctypes
    from william_ci_amd.engine import Engine, lib
    s = tsys.workload("h10_ring")
    eng = Engine(s, normalised_coef(s))
    try:
        nb = 2048
        P = np.zeros(nb * nb)
        J = np.zeros(nb * nb)
        K = np.zeros(nb * nb)
        rc = lib().tcg_jk_stored_noalloc(eng.h, 5, nb, P, J, K)
        assert rc == -1
        assert "scratch" in lib().tcg_last_error(eng.h).decode()
    finally:
        eng.close()
