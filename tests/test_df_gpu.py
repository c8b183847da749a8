"""GPU parity of density fitting (tcg_df_create / tcg_df_jk / tcg_df_kmp2)
against the numpy restatement of df_toroidal (oracle/df.py), whose integrals
come from the C restatement of the quadrature.

Bars: the fitted integrals sum_Q B^Q B^Q within 1e-10 (relative to the
largest); ranks of the pivoted Cholesky as the oracle's zpstrf; J, K within
1e-10; DF-RHF energy within 1e-9 Eh and DF-kMP2 within 1e-10 Eh of the
oracle on the same orbitals.  At the DF-MP2 workload size (H64 chain, 41
auxiliary functions per H): every product's fitted self-repulsion below the
exact one, and the DF correlation energy within 5e-4 relative of the exact
k-MP2 on the same orbitals.
"""
import numpy as np
import pytest

from tests_support import normalised_coef
from william_ci_amd import system as tsys

pytestmark = pytest.mark.gpu


def _ring(n=4, L=7.2):
    return tsys.chain_1d(n, 1.8, L)


def _oracle_fit(s, coef, aux):
    from oracle import df as odf
    from oracle import orc
    cs, ccoef, (nf, nau) = odf.combined(s, coef, aux)
    o = orc.OracleSystem(cs, ccoef, nq=32)
    Vm, T3 = odf.integrals(o, cs, nf, nau)
    return odf.df_B(Vm, T3, s.ncell)


@pytest.fixture(scope="module")
def ring4():
    from william_ci_amd import df as wdf
    from william_ci_amd.engine import Engine
    s = _ring()
    coef = normalised_coef(s)
    aux = wdf.aux_list(s, wdf.etb_aux(tsys.read_basis(tsys.H_BASIS)))
    print("case ring4 DF (etb aux, 25 per H)", flush=True)
    eng = Engine(s, coef)
    d = wdf.DF(eng, aux)
    B, ranks, dmin = _oracle_fit(s, coef, aux)
    yield dict(s=s, coef=coef, aux=aux, eng=eng, df=d, B=B, ranks=ranks, dmin=dmin)
    d.close()
    eng.close()


def test_df_b_matches_oracle(ring4):
    """df_B is fixed up to a unitary rotation of the kept auxiliary functions
    per k (pivots that tie to rounding, e.g. p_y / p_z on a chain, may be taken
    in either order), so the bar is on what B is for: the fitted integrals
    sum_Q B^Q B^Q, and the rank per k."""
    from oracle import df as odf
    d = ring4["df"]
    s = ring4["s"]
    Bg = d.B().transpose(2, 1, 0)                    # [mu0, nu, Q]
    Eg = odf.eri_sao(Bg, s.ncell)
    Eo = odf.eri_sao(ring4["B"], s.ncell)
    err = np.abs(Eg - Eo).max() / np.abs(Eo).max()
    info = d.info()
    print(f"  fitted ERI max rel diff {err:.2e}  |B| diff {np.abs(Bg - ring4['B']).max():.2e}  info {info}  "
          f"times {d.times()}")
    assert err < 1e-10
    assert info["dropped"] == int(d.nau - ring4["ranks"].min())
    assert info["smallest_pivot"] > 1e-10


def test_df_jk_matches_oracle(ring4):
    from oracle import df as odf
    s = ring4["s"]
    nf, nb, N = s.nfuc, s.nbas, s.n_cell
    cdiff, cadd = odf.cell_ops(s.ncell)
    rng = np.random.default_rng(11)
    P0 = rng.normal(size=(nf, nb))
    P = np.zeros((nb, nb))
    for R in range(N):
        for nu in range(nb):
            P[R * nf:(R + 1) * nf, cadd[nu // nf, R] * nf + nu % nf] = P0[:, nu]
    P = 0.5 * (P + P.T)
    Jg, Kg = ring4["df"].jk(P)
    Jo, Ko = odf.jk(ring4["B"], P, s.ncell)
    ej, ek = np.abs(Jg - Jo).max(), np.abs(Kg - Ko).max()
    print(f"  J diff {ej:.2e}  K diff {ek:.2e}")
    assert ej < 1e-10 and ek < 1e-10
    assert np.all(Jg[nf:] == 0) and np.all(Kg[nf:] == 0)


def test_df_rhf_and_kmp2_match_oracle(ring4):
    from oracle import df as odf
    from oracle import kmp2 as okm
    from william_ci_amd import scf
    s, eng, d = ring4["s"], ring4["eng"], ring4["df"]
    S, T, V = eng.one_electron()
    rg = scf.rhf_sao(s, S, T, V, d.jk)
    ro = scf.rhf_sao(s, S, T, V, lambda P: odf.jk(ring4["B"], P, s.ncell))
    print(f"  DF-RHF engine {rg.energy:.12f} oracle {ro.energy:.12f}")
    assert abs(rg.energy - ro.energy) < 1e-9
    got = d.kmp2(rg.mo_coef, rg.mo_energies, rg.occupied)
    occ = np.where(rg.occupied, 2, 0)
    nok = (occ == 2).sum(1)
    o_hi, v_lo = int(nok.max()), int(nok.min())
    ekmo = odf.kmo(ring4["B"], rg.mo_coef, s.ncell, o_hi, v_lo)
    want = okm.kmp2(ekmo, rg.mo_energies, occ, s.ncell, v_lo)
    print(f"  DF-kMP2 engine {got['ec']:.14f} oracle {want[0]:.14f}  gpu {got['gpu_ms']:.2f} ms")
    assert abs(got["ec"] - want[0]) < 1e-10
    assert got["zero_denominators"] == want[3]


def test_df_mp2_h64_chain_etb41():
    """DF-MP2 workload: H64 chain, 41 auxiliary functions per H, RHF with the
    exact stored integrals; DF and exact k-MP2 on the same orbitals."""
    from william_ci_amd import df as wdf
    from william_ci_amd import scf
    from william_ci_amd.engine import Engine
    s = tsys.workload("h64_chain")
    coef = normalised_coef(s)
    aux = wdf.aux_list(s, wdf.read_aux(wdf.AUX_H_ETB41))
    print("case h64_chain DF-MP2 (41 aux per H)", flush=True)
    eng = Engine(s, coef)
    S, T, V = eng.one_electron()
    eng.eri_build(sao=True, to_host=False)
    res = scf.rhf_sao(s, S, T, V, lambda P: eng.jk_stored(P))
    exact = eng.kmp2(res.mo_coef, res.mo_energies, res.occupied)
    d = wdf.DF(eng, aux)
    info = d.info()
    got = d.kmp2(res.mo_coef, res.mo_energies, res.occupied)
    # fitted self-repulsion of every product (mu0 nu) below the exact one
    B = d.B()                                        # [Q, nu, mu0]
    fit_diag = np.einsum("Qnm,Qnm->mn", B, B)
    from william_ci_amd.learned import eri_blocks, shell_starts
    st = shell_starts(s)
    nsh = len(st) - 1
    nsh0 = int(np.sum(st[:-1] < s.nfuc))
    q4 = np.array([(a, b, a, b) for a in range(nsh0) for b in range(nsh)])
    vals, off = eri_blocks(eng, q4)
    exact_diag = np.zeros((s.nfuc, s.nbas))
    for x, (a, b, _, _) in enumerate(q4):
        na, nb_ = st[a + 1] - st[a], st[b + 1] - st[b]
        blk = vals[off[x]:off[x + 1]].reshape(nb_, na, nb_, na)       # column-major (na, nb, na, nb)
        for i in range(na):
            for j in range(nb_):
                exact_diag[st[a] + i, st[b] + j] = blk[j, i, j, i]
    Jx, Kx = eng.jk_stored(res.P)
    Jd, Kd = d.jk(res.P)
    tm = d.times()
    d.close()
    eng.close()
    print(f"  info {info}  times {tm}")
    print(f"  E_c DF {got['ec']:.10f} exact {exact['ec']:.10f}  diff {got['ec'] - exact['ec']:.2e}  "
          f"DF gpu {got['gpu_ms']:.1f} ms  exact gpu {exact['gpu_ms']:.1f} ms")
    print(f"  max fit excess {np.max(fit_diag - exact_diag):.2e}  |dJ| {np.abs(Jd - Jx)[:s.nfuc].max():.2e}  "
          f"|dK| {np.abs(Kd - Kx)[:s.nfuc].max():.2e}")
    assert info["naux"] == 41 * 64
    assert np.all(fit_diag <= exact_diag + 1e-10)
    assert abs(got["ec"] - exact["ec"]) < 5e-4 * abs(exact["ec"])
    nf = s.nfuc
    assert np.abs(Jd[:nf] - Jx[:nf]).max() < 1e-3 and np.abs(Kd[:nf] - Kx[:nf]).max() < 1e-3


def test_df_more_than_64_aux_per_cell():
    """Two H per cell with the 41-function set: 82 auxiliary functions per cell,
    the global-memory metric factorisation; fitted integrals and J/K against
    the oracle."""
    from oracle import df as odf
    from william_ci_amd import df as wdf
    from william_ci_amd.engine import Engine
    s = tsys.build_system("Tori1D\n3 3.6 10.8\nSAO\n$$\nH 0 0 0\nH 1.4 0 0\n$$\n", tsys.H_BASIS)
    coef = normalised_coef(s)
    aux = wdf.aux_list(s, wdf.read_aux(wdf.AUX_H_ETB41))
    print("case H2 chain x3, 82 aux per cell", flush=True)
    eng = Engine(s, coef)
    d = wdf.DF(eng, aux)
    B, ranks, dmin = _oracle_fit(s, coef, aux)
    Bg = d.B().transpose(2, 1, 0)
    Eg, Eo = odf.eri_sao(Bg, s.ncell), odf.eri_sao(B, s.ncell)
    err = np.abs(Eg - Eo).max() / np.abs(Eo).max()
    info = d.info()
    nf, nb, N = s.nfuc, s.nbas, s.n_cell
    cdiff, cadd = odf.cell_ops(s.ncell)
    rng = np.random.default_rng(4)
    P0 = rng.normal(size=(nf, nb))
    P = np.zeros((nb, nb))
    for R in range(N):
        for nu in range(nb):
            P[R * nf:(R + 1) * nf, cadd[nu // nf, R] * nf + nu % nf] = P0[:, nu]
    P = 0.5 * (P + P.T)
    Jg, Kg = d.jk(P)
    Jo, Ko = odf.jk(B, P, s.ncell)
    d.close()
    eng.close()
    print(f"  nau {info['nau']} dropped {info['dropped']} (oracle {int(d.nau - ranks.min())})  fitted ERI rel diff "
          f"{err:.2e}  J {np.abs(Jg - Jo).max():.2e}  K {np.abs(Kg - Ko).max():.2e}")
    assert info["nau"] == 82
    assert info["dropped"] == int(82 - ranks.min())
    assert err < 1e-9
    assert np.abs(Jg - Jo).max() < 1e-9 and np.abs(Kg - Ko).max() < 1e-9


@pytest.mark.parametrize("name", ["h64_chain", "h2_3x3_2d"])
def test_df_mp2_fp32_matrix_cores(name):
    """The FP32 DF-MP2 (FP32 L, (ov|ov) as FP32 matrix-core GEMMs with FP32
    accumulation, compensated FP32 energy sums) against the FP64 DF-MP2 on the
    same orbitals and the same fit.  Bar: the FP32 rounding of the integrals,
    1e-6 relative in E_c (observed ~1e-7), far below 1 microhartree on these
    energies; the same zero-denominator count."""
    from william_ci_amd import df as wdf
    from william_ci_amd import scf
    from william_ci_amd.engine import Engine
    s = tsys.workload(name) if name == "h64_chain" else tsys.lattice_2d_h2(3, 3.6, 10.8)
    print(f"case {name} DF-MP2 FP32 vs FP64", flush=True)
    eng = Engine(s, normalised_coef(s))
    S, T, V = eng.one_electron()
    eng.eri_build(sao=True, to_host=False)
    res = scf.rhf_sao_gpu(s, eng, S, T, V)
    d = wdf.DF(eng, wdf.aux_list(s, wdf.read_aux(wdf.AUX_H_ETB41)))
    e64 = d.kmp2(res.mo_coef, res.mo_energies, res.occupied)
    e32 = d.kmp2(res.mo_coef, res.mo_energies, res.occupied, fp32=True)
    d.close()
    eng.close()
    diff = e32["ec"] - e64["ec"]
    print(f"  E_c FP64 {e64['ec']:.12f}  FP32 {e32['ec']:.12f}  diff {diff:.2e}  "
          f"({e64['gpu_ms']:.2f} / {e32['gpu_ms']:.2f} ms)")
    assert e64["ec"] < 0.0
    assert abs(diff) < 1e-6 * abs(e64["ec"])
    assert abs(e32["direct"] - e64["direct"]) < 1e-6 * abs(e64["direct"])
    assert e32["zero_denominators"] == e64["zero_denominators"]
