"""GPU parity of the k-space SCF algebra (tcg_ao_to_sao, tcg_sao_to_ao,
tcg_heev_blocks) and of the whole RHF_SAO iteration on the GPU (tcg_rhf_sao).

  * The lattice transforms against the reference's own transform_AO_to_SAO /
    transform_SAO_to_AO (src/Tools/SAO_Transfer.f90:1,108, compiled in
    oracle/_ref) on random matrices, to 1e-13 of their scale.
  * The block eigensolver against numpy's eigh of the block after the
    reference's own clean_and_symmetrize (zheev is not in the image):
    eigenvalues, residuals, orthonormality, ascending order.
  * tcg_rhf_sao against the host loop scf.rhf_sao (numpy algebra around the
    GPU J/K, pinned to the reference by test_rhf_pins): the energy after a
    fixed number of iterations (the same sequence of densities) and at
    convergence, orbital energies, density; with stored and direct J/K; in a
    basis with dropped (linearly dependent) vectors; and the k-MP2 energy on
    its orbitals.
"""
import numpy as np
import pytest

from conftest import needs_ref
from william_ci_amd import scf
from william_ci_amd import system as tsys
from william_ci_amd.basis import normalize

pytestmark = pytest.mark.gpu

SYSTEMS = {
    "h10_ring": lambda: tsys.workload("h10_ring"),
    "h2_3x3_2d": lambda: tsys.lattice_2d_h2(3, 3.6, 10.8),
    "h2_2x2x2_d_3d": lambda: tsys.lattice_3d_h2(2, 4.0, 8.0),
}


def _engine(s):
    from william_ci_amd.engine import Engine
    return Engine(s, normalize(s))


def _blocks(M, s):
    nf = s.nfuc
    return [M[k * nf:(k + 1) * nf, k * nf:(k + 1) * nf] for k in range(s.n_cell)]


@needs_ref
@pytest.mark.parametrize("case", list(SYSTEMS))
def test_lattice_transforms_match_reference(case):
    from oracle.ref import RefSystem
    s = SYSTEMS[case]()
    nb, nf = s.nbas, s.nfuc
    print(f"case {case} nbas={nb} nfuc={nf} cells={s.ncell}", flush=True)
    r = RefSystem(s)
    eng = _engine(s)
    rng = np.random.default_rng(5)
    M = rng.standard_normal((nb, nb))            # not translation-symmetric: only unit-cell rows count
    A = eng.ao_to_sao(M)
    Ar = r.ao_to_sao(M)
    d = np.abs(A - Ar).max()
    print(f"  ao_to_sao max |diff| {d:.2e} (scale {np.abs(Ar).max():.2f})", flush=True)
    assert d < 1e-13 * np.abs(Ar).max()
    # k blocks with elements below the 1e-15 cut and exact zeros
    B = np.zeros((nb, nb), dtype=complex)
    for k in range(s.n_cell):
        X = rng.standard_normal((nf, nf)) + 1j * rng.standard_normal((nf, nf))
        X[0, -1] = 3e-16
        X[-1, 0] = 0.0
        B[k * nf:(k + 1) * nf, k * nf:(k + 1) * nf] = X
    Mo = eng.sao_to_ao(B)
    Mr = r.sao_to_ao(B)
    d = np.abs(Mo - Mr).max()
    print(f"  sao_to_ao max |diff| {d:.2e} (scale {np.abs(Mr).max():.2f})", flush=True)
    assert d < 1e-13 * np.abs(Mr).max()
    # a block-circulant matrix comes back from its k blocks
    Mc = eng.sao_to_ao(eng.ao_to_sao(Mr))
    assert np.abs(Mc - Mr).max() < 1e-13 * np.abs(Mr).max()
    eng.close()


@needs_ref
@pytest.mark.parametrize("nf,ncell", [(1, (7, 1, 1)), (5, (4, 1, 1)), (10, (3, 3, 1)), (22, (2, 2, 2)), (37, (2, 1, 1))])
def test_heev_blocks(nf, ncell):
    """Blocks of random Hermitian matrices and special ones (diagonal, zero,
    degenerate, a lower-triangle element below the 1e-15 cut, an upper triangle
    that disagrees with the lower one) against eigh after clean_and_symmetrize."""
    from oracle.ref import clean_and_symmetrize
    kind = {1: "Tori1D", 2: "Tori2D", 3: "Tori3D"}[sum(1 for n in ncell if n > 1) or 1]
    N = ncell[0] * ncell[1] * ncell[2]
    nb = nf * N
    print(f"heev nf={nf} cells={ncell} kind={kind}", flush=True)
    rng = np.random.default_rng(nf)
    blocks = []
    for k in range(N):
        X = rng.standard_normal((nf, nf)) + 1j * rng.standard_normal((nf, nf))
        H = X + X.conj().T
        if k == 1:
            H = np.diag(rng.standard_normal(nf)).astype(complex)
        if k == 2 and N > 2:
            H = np.zeros((nf, nf), dtype=complex)
        if k == 3 and N > 3:
            Q, _ = np.linalg.qr(X)
            d = np.repeat(rng.standard_normal((nf + 1) // 2), 2)[:nf]        # pairs of equal eigenvalues
            H = (Q * d) @ Q.conj().T
        if nf > 2 and k == 0:
            H[2, 0] = 5e-16                      # below the cut
            H[0, 1] += 0.3                       # upper triangle disagrees: the lower one counts
        blocks.append(H)
    A = np.zeros((nb, nb), dtype=complex)
    for k, H in enumerate(blocks):
        A[k * nf:(k + 1) * nf, k * nf:(k + 1) * nf] = H
    A[0, nb - 1] = 7.0 + 1.0j                     # outside the blocks: left alone
    eng = _heev_engine(nf, ncell)
    V, e = eng.heev_blocks(A)
    eng.close()
    assert V[0, nb - 1] == 7.0 + 1.0j
    for k, H in enumerate(blocks):
        Hc = clean_and_symmetrize(H)
        w = np.linalg.eigvalsh(Hc)
        ek = e[k * nf:(k + 1) * nf]
        Vk = V[k * nf:(k + 1) * nf, k * nf:(k + 1) * nf]
        scale = max(1.0, np.abs(w).max())
        assert np.all(np.diff(ek) >= 0.0), k
        assert np.abs(ek - w).max() < 1e-13 * scale, (k, np.abs(ek - w).max())
        res = np.abs(Hc @ Vk - Vk * ek).max()
        orth = np.abs(Vk.conj().T @ Vk - np.eye(nf)).max()
        assert res < 1e-12 * scale, (k, res)
        assert orth < 1e-13, (k, orth)


def _heev_engine(nf, ncell):
    """An engine whose basis has nf s functions per cell on the given lattice
    (the block routines read only nfuc and the cell counts)."""
    from william_ci_amd.engine import Engine
    nx, ny, nz = ncell
    ndp = 3 if nz > 1 else (2 if ny > 1 else 1)
    bas = "a 1\n$ S-TYPE FUNCTIONS\n1 1\n1.0 1.0\n"
    atoms = "\n".join(f"H {0.01 * i} 0 0" for i in range(nf))
    if ndp == 1:
        text = f"Tori1D\n{nx} 3.0 {3.0 * nx}\nSAO\n$$\n{atoms}\n$$\n"
    elif ndp == 2:
        text = f"Tori2D\n{nx} {ny} 3.0 {3.0 * nx} {3.0 * ny}\nSAO\n$$\n{atoms}\n$$\n"
    else:
        text = f"Tori3D\n{nx} {ny} {nz} 3.0 {3.0 * nx} {3.0 * ny} {3.0 * nz}\nSAO\n$$\n{atoms}\n$$\n"
    s = tsys.build_system(text, bas)
    assert s.nfuc == nf and tuple(s.ncell) == tuple(ncell)
    return Engine(s, s.coef.copy())


def _host_and_gpu(s, source="stored", max_scf=scf.MAX_SCF, diis=True):
    eng = _engine(s)
    S, T, V = eng.one_electron()
    if source == "stored":
        eng.eri_build(to_host=False)
        jk = lambda P: eng.jk_stored(P)
    else:
        eng.direct_init()
        jk = lambda P: eng.direct_jk(P)
    host = scf.rhf_sao(s, S, T, V, jk, max_scf=max_scf, diis=diis)
    gpu = scf.rhf_sao_gpu(s, eng, S, T, V, source=source, max_scf=max_scf, diis=diis)
    return eng, host, gpu


@pytest.mark.parametrize("case", list(SYSTEMS))
def test_rhf_sao_gpu_matches_host_loop(case):
    s = SYSTEMS[case]()
    print(f"case {case} nbas={s.nbas} stored J/K", flush=True)
    eng, host, gpu = _host_and_gpu(s)
    print(f"  host E {host.energy:.12f} ({host.iterations} it)  gpu E {gpu.energy:.12f} ({gpu.iterations} it)  "
          f"diff {gpu.energy - host.energy:.2e}", flush=True)
    assert gpu.converged and host.converged
    assert abs(gpu.energy - host.energy) < 1e-9
    assert abs(gpu.iterations - host.iterations) <= 2
    assert np.abs(np.sort(gpu.mo_energies, axis=1) - np.sort(host.mo_energies, axis=1)).max() < 1e-7
    assert np.abs(gpu.P - host.P).max() < 1e-6
    assert gpu.occupied.sum() == s.n_occ
    # k-MP2 on the GPU loop's orbitals
    m_h = eng.kmp2(host.mo_coef, host.mo_energies, host.occupied)
    m_g = eng.kmp2(gpu.mo_coef, gpu.mo_energies, gpu.occupied)
    print(f"  k-MP2 host {m_h['ec']:.12f}  gpu {m_g['ec']:.12f}", flush=True)
    assert abs(m_g["ec"] - m_h["ec"]) < 1e-9
    eng.close()


@pytest.mark.parametrize("max_scf,diis", [(1, True), (3, True), (4, False)])
def test_rhf_sao_gpu_same_iterates(max_scf, diis):
    """After a fixed number of iterations the two loops have built the same
    densities: same energy to round-off (the guess, DIIS and occupations agree
    step by step)."""
    s = tsys.lattice_2d_h2(3, 3.6, 10.8)
    print(f"case h2_3x3_2d max_scf={max_scf} diis={diis}", flush=True)
    eng, host, gpu = _host_and_gpu(s, max_scf=max_scf, diis=diis)
    print(f"  host E {host.energy:.14f}  gpu E {gpu.energy:.14f}", flush=True)
    assert gpu.iterations == host.iterations == max_scf
    assert abs(gpu.energy - host.energy) < 1e-11
    # the DIIS solve amplifies round-off by the condition number of its matrix
    assert np.abs(gpu.P - host.P).max() < 1e-9
    eng.close()


@pytest.mark.parametrize("case", ["h10_ring", "h2_2x2x2_d_3d"])
def test_rhf_sao_gpu_direct(case):
    """Integral-direct J/K inside the GPU loop against the host loop on the
    same direct J/K."""
    s = SYSTEMS[case]()
    print(f"case {case} nbas={s.nbas} direct J/K", flush=True)
    eng, host, gpu = _host_and_gpu(s, source="direct")
    print(f"  host E {host.energy:.12f}  gpu E {gpu.energy:.12f}  diff {gpu.energy - host.energy:.2e}", flush=True)
    assert gpu.converged
    assert abs(gpu.energy - host.energy) < 1e-9
    eng.close()


def test_rhf_sao_gpu_dropped_vectors():
    """H8 ring at 1.0 bohr: two k blocks have S_k eigenvalues below 1e-6
    (canonical X with zero columns, those levels shifted to 1e4).  The host
    loop takes the supercell guess and X here; both span the same space."""
    s = tsys.chain_1d(8, 1.0, 8.0)
    print(f"case h8_ring_1.0 nbas={s.nbas}", flush=True)
    eng, host, gpu = _host_and_gpu(s)
    print(f"  host E {host.energy:.12f} ({host.iterations} it)  gpu E {gpu.energy:.12f} ({gpu.iterations} it)",
          flush=True)
    assert gpu.converged
    assert abs(gpu.energy - host.energy) < 1e-8
    assert np.all(gpu.mo_energies.max(axis=1)[[0, 4]] == pytest.approx(1e4, rel=1e-6))
    eng.close()


def test_rhf_sao_gpu_errors():
    s = tsys.workload("h10_ring")
    eng = _engine(s)
    S, T, V = eng.one_electron()
    with pytest.raises(Exception, match="no resident ERI"):
        eng.rhf_sao(S, T, V, s.n_occ)
    with pytest.raises(Exception, match="nocc"):
        eng.rhf_sao(S, T, V, 0)
    with pytest.raises(Exception, match="direct_init"):
        eng.rhf_sao(S, T, V, s.n_occ, source="direct")
    eng.close()
