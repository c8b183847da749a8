"""The two large workloads at their full size on the GPU.

  * Tori2D 8x8 lattice of H2 (640 functions): direct-SCF RHF on the GPU equals
    the RHF on the stored SAO tensor (both through tcg_rhf_sao), and the
    DF-kMP2 energy with the 41-function even-tempered auxiliary basis is within
    its fitting error of the exact k-MP2 on the same orbitals.
  * Tori3D 4x4x4 lattice of H2 + d (1408 functions, no stored tensor: it would
    take 491 GB): the direct J and K rows of a random density against the same
    rows assembled from shell blocks (tcg_eri_blocks) over every ket (J) and
    every middle pair (K), to the Schwarz screening; the direct RHF converges and its Coulomb energy is
    reproduced by the density-fitted J to the fitting error.
"""
import numpy as np
import pytest

from william_ci_amd import df, scf
from william_ci_amd import system as tsys
from william_ci_amd.basis import normalize

pytestmark = pytest.mark.gpu


def _engine(s):
    from william_ci_amd.engine import Engine
    return Engine(s, normalize(s))


def test_tori2d_8x8_direct_rhf_matches_stored_and_df_mp2():
    s = tsys.workload("h2_8x8")
    print(f"case h2_8x8 nbas={s.nbas}", flush=True)
    eng = _engine(s)
    S, T, V = eng.one_electron()
    eng.eri_build(to_host=False)
    stored = scf.rhf_sao_gpu(s, eng, S, T, V, source="stored")
    eng.direct_init()
    direct = scf.rhf_sao_gpu(s, eng, S, T, V, source="direct")
    print(f"  stored E {stored.energy:.12f} ({stored.iterations} it)  direct E {direct.energy:.12f} "
          f"({direct.iterations} it)", flush=True)
    assert stored.converged and direct.converged
    assert abs(direct.energy - stored.energy) < 1e-8
    exact = eng.kmp2(stored.mo_coef, stored.mo_energies, stored.occupied)
    d = df.DF(eng, df.aux_list(s, df.read_aux(df.AUX_H_ETB41)))
    fitted = d.kmp2(stored.mo_coef, stored.mo_energies, stored.occupied)
    d.close()
    eng.close()
    print(f"  k-MP2 exact {exact['ec']:.10f}  DF {fitted['ec']:.10f}", flush=True)
    assert exact["ec"] < 0.0
    assert abs(fitted["ec"] - exact["ec"]) < 5e-4 * abs(exact["ec"])


@pytest.fixture(scope="module")
def tori3d():
    s = tsys.workload("h2_4x4x4_d")
    eng = _engine(s)
    eng.direct_init()
    yield s, eng
    eng.close()


def _shells(s):
    from william_ci_amd.learned import shell_starts
    st = shell_starts(s)
    return st[:-1], np.diff(st)


def _blocks(eng, q, ncomp):
    """tcg_eri_blocks on the shell quartets q (n, 4), with an output sized to them."""
    from william_ci_amd.learned import _lib
    q = np.ascontiguousarray(q, dtype=np.int32)
    out = np.zeros(int(np.prod(ncomp[q], axis=1).sum()))
    off = np.zeros(len(q) + 1, dtype=np.int64)
    eng._check(_lib().tcg_eri_blocks(eng.h, len(q), q.ravel(), out, off), "tcg_eri_blocks")
    return out, off


def test_tori3d_direct_jk_rows_match_blocks(tori3d):
    s, eng = tori3d
    nb, nf = s.nbas, s.nfuc
    print(f"case h2_4x4x4_d nbas={nb} nfuc={nf}", flush=True)
    first, ncomp = _shells(s)
    nsh = len(first)
    rng = np.random.default_rng(9)
    P = rng.standard_normal((nb, nb)) * 0.01
    P = P + P.T
    J, K = eng.direct_jk(P)
    # an s shell of cell 0 as the first index; J(a, b) over every (c, d), K(a, d) over every (b, c)
    sA = int(np.nonzero((ncomp == 1) & (first < nf))[0][0])
    sB = int(np.nonzero((ncomp == 1) & (first < nf))[0][1])
    cd = np.array([(c, d) for c in range(nsh) for d in range(nsh)], dtype=np.int32)
    a, b = first[sA], first[sB]
    q = np.column_stack([np.full(len(cd), sA), np.full(len(cd), sB), cd[:, 0], cd[:, 1]])
    out, off = _blocks(eng, q, ncomp)
    Jab = 0.0
    for i, (c, d) in enumerate(cd):
        blk = out[off[i]:off[i + 1]].reshape(ncomp[d], ncomp[c])          # [d][c], c fastest
        Jab += np.sum(blk * P[first[c]:first[c] + ncomp[c], first[d]:first[d] + ncomp[d]].T)
    # the direct build drops the elements under the Schwarz screens (1e-13 on the
    # products, as the reference); over the 2e6 kets of this row they add up to
    # ~1e-9 against |P| ~ 1e-2.  A missing or doubled element would show at 1e-4.
    tol = 1e-8
    dJ = abs(J[a, b] - Jab)
    print(f"  J({a},{b}) direct {J[a, b]:.14f} blocks {Jab:.14f} diff {dJ:.2e}", flush=True)
    assert dJ < tol
    sD = sB
    d0 = first[sD]
    q = np.column_stack([np.full(len(cd), sA), cd[:, 0], cd[:, 1], np.full(len(cd), sD)])
    out, off = _blocks(eng, q, ncomp)
    Kad = 0.0
    for i, (bsh, csh) in enumerate(cd):
        blk = out[off[i]:off[i + 1]].reshape(ncomp[csh], ncomp[bsh])      # [c][b]
        Kad += np.sum(blk * P[first[bsh]:first[bsh] + ncomp[bsh], first[csh]:first[csh] + ncomp[csh]].T)
    Kad *= -0.5
    dK = abs(K[a, d0] - Kad)
    print(f"  K({a},{d0}) direct {K[a, d0]:.14f} blocks {Kad:.14f} diff {dK:.2e}", flush=True)
    assert dK < tol
    assert np.all(J[nf:] == 0.0) and np.all(K[nf:] == 0.0)


def test_tori3d_direct_rhf_and_density_fitting(tori3d):
    s, eng = tori3d
    S, T, V = eng.one_electron()
    r = scf.rhf_sao_gpu(s, eng, S, T, V, source="direct")
    print(f"  direct RHF E {r.energy:.10f} ({r.iterations} it, {r.times['jk_s']:.1f} s in J/K)", flush=True)
    assert r.converged
    J, K = eng.direct_jk(r.P)
    d = df.DF(eng, df.aux_list(s, df.read_aux(df.AUX_H_ETB41)))
    Jf, Kf = d.jk(r.P)
    m = d.kmp2(r.mo_coef, r.mo_energies, r.occupied)
    d.close()
    N = s.n_cell
    EJ = 0.5 * N * np.sum(r.P * J.T)
    EJf = 0.5 * N * np.sum(r.P * Jf.T)
    EK = 0.5 * N * np.sum(r.P * K.T)
    EKf = 0.5 * N * np.sum(r.P * Kf.T)
    print(f"  EJ direct {EJ:.10f} DF {EJf:.10f}  EK direct {EK:.10f} DF {EKf:.10f}  DF-kMP2 {m['ec']:.10f}",
          flush=True)
    assert abs(EJf - EJ) < 1e-4 * abs(EJ)
    assert abs(EKf - EK) < 1e-3 * abs(EK)
    assert -0.1 * s.n_occ < m["ec"] < 0.0


def test_tori2d_8x8_learned_direct_matches_converged_quadrature():
    """The 640-function lattice with the reference's start-up rule choice (the
    64-point table): learned per-quartet orders and incremental Fock builds in the
    direct RHF, then the direct k-MP2, against the plain direct path on the
    converged quadrature (the 128-point table): E_total within 1 microhartree, no
    quartet on the fallback rule, and fewer quartets evaluated than the plain
    path's."""
    import os
    from william_ci_amd import learned
    model_path = learned.MODEL_PATH
    if not os.path.exists(model_path):
        pytest.skip("no trained model committed")
    s = tsys.workload("h2_8x8")
    print(f"case h2_8x8 learned direct nbas={s.nbas}", flush=True)
    eng = _engine(s)
    eng.fast_quad_self_check(64)
    S, T, V = eng.one_electron()
    m = learned.Model.load()
    m.install(eng)
    counts = learned.direct_init_learned(eng, m)
    eng.set_direct_screen()
    rl = scf.rhf_sao_gpu(s, eng, S, T, V, source="direct_incremental")
    ml = eng.kmp2_direct(rl.mo_coef, rl.mo_energies, rl.occupied)
    info = eng.direct_screen_info()
    eng.set_quadrature(False, 128)
    eng.direct_init()
    r0 = scf.rhf_sao_gpu(s, eng, S, T, V, source="direct")
    m0 = eng.kmp2_direct(r0.mo_coef, r0.mo_energies, r0.occupied)
    info0 = eng.direct_screen_info()
    eng.close()
    nord = len(m.orders)
    el, e0 = rl.energy + ml["ec"], r0.energy + m0["ec"]
    print(f"  learned E_total {el:.10f} ({rl.iterations} it, {info['kept_total']} quartets)  "
          f"128-point table {e0:.10f} ({r0.iterations} it, {info0['kept_total']} quartets)  diff {el - e0:.2e}")
    assert rl.converged and r0.converged
    assert counts[nord] == 0
    assert abs(el - e0) < 1e-6
    assert info["kept_total"] < info0["kept_total"]
