"""Host SCF algebra: the k-space forms used by rhf_sao equal the supercell
forms they replace (Lowdin X = S^-1/2 and the core-guess density with the
degenerate level at the cut shared evenly), on translation-invariant
matrices where the k and -k levels are degenerate."""
import numpy as np

from conftest import needs_ref

from william_ci_amd import scf
from william_ci_amd import system as tsys


def _circulant(rng, nf, nb, N, scale):
    row = rng.normal(size=(nf, nb)) * scale
    M = np.zeros((nb, nb))
    for R in range(N):
        M[R * nf:(R + 1) * nf] = np.roll(row, R * nf, axis=1)
    return 0.5 * (M + M.T)


def test_kspace_startup_equals_supercell():
    s = tsys.chain_1d(8, 1.8, 14.4)
    nb, nf, N = s.nbas, s.nfuc, s.n_cell
    rng = np.random.default_rng(1)
    S = np.eye(nb) + 0.05 * _circulant(rng, nf, nb, N, 0.1)
    Hc = _circulant(rng, nf, nb, N, 1.0)
    bl = scf.Bloch(s)
    X1 = scf.canonical_X(S)
    sk, Uk = np.linalg.eigh(bl.to_k(S))
    Xk = (Uk / np.sqrt(sk)[:, None, :]) @ Uk.conj().transpose(0, 2, 1)
    assert np.abs(X1 - bl.to_ao(Xk)).max() < 1e-13
    e0, U0 = np.linalg.eigh(X1.T @ Hc @ X1)
    # a degenerate pair straddles the cut for some occupation: test every cut
    e0k, U0k = np.linalg.eigh(Xk.conj().transpose(0, 2, 1) @ bl.to_k(Hc) @ Xk)
    for nO in range(1, nb // 2):
        P1 = scf._occupy_cut(e0, X1 @ U0, nO)
        P2, Pk = scf._occupy_cut_k(e0k, Xk @ U0k, nO, bl)
        assert np.abs(bl.to_k(P2) - Pk).max() < 1e-12, nO
        assert np.abs(P1 - P2).max() < 1e-12, nO


@needs_ref
def test_bloch_transforms_match_reference():
    """scf.Bloch (numpy FFT) against the reference's own transform_AO_to_SAO and
    transform_SAO_to_AO (src/Tools/SAO_Transfer.f90:1,108) on 1D, 2D and 3D
    lattices, and the reference's clean_and_symmetrize on a planted case."""
    from oracle.ref import RefSystem, clean_and_symmetrize
    rng = np.random.default_rng(3)
    for s in (tsys.chain_1d(6, 1.8, 10.8), tsys.lattice_2d_h2(3, 3.6, 10.8), tsys.lattice_3d_h2(2, 4.0, 8.0)):
        nb, nf = s.nbas, s.nfuc
        r = RefSystem(s)
        bl = scf.Bloch(s)
        M = rng.standard_normal((nb, nb))
        Ms = r.ao_to_sao(M)
        Mk = bl.to_k(M)
        for k in range(s.n_cell):
            blk = Ms[k * nf:(k + 1) * nf, k * nf:(k + 1) * nf]
            assert np.abs(blk - Mk[k]).max() < 1e-13 * np.abs(Mk).max()
            Ms[k * nf:(k + 1) * nf, k * nf:(k + 1) * nf] = 0.0
        assert np.all(Ms == 0.0)                 # nothing off the k blocks
        C = r.sao_to_ao(r.ao_to_sao(M))
        assert np.abs(C - bl.to_ao(Mk)).max() < 1e-13 * np.abs(C).max()
    A = np.array([[1.0, 2.0 + 1.0j, 0.0], [3.0 - 1.0j, 2.0 + 0.5j, 0.0], [4e-16, 1e-3j, -1.0]])
    Ac = clean_and_symmetrize(A)
    expect = np.array([[1.0, 3.0 + 1.0j, 0.0], [3.0 - 1.0j, 2.0, -1e-3j], [0.0, 1e-3j, -1.0]])
    assert np.abs(Ac - expect).max() == 0.0
