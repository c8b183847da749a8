"""The text files of the keywords Details and Read (william_ci_amd/formats.py)
against the reference's own details_integrals / details_integrals_SAO
(src/Tools/print/Details.f90, built into oracle/_ref), on the CPU.

  * OV.dat, KI.dat, NA.dat and the *_matrix.dat files: byte-identical to the
    reference's (fixed Fortran formats es24.16e3, f16.12, i3)
  * ERI.dat: the same lines (indices, order) with the same values read back
    exactly; its spacing follows gfortran's list-directed layout while the
    checker's build is flang's (list-directed layout is compiler-defined)
  * read_integrals (read_integrals_from_file) gives back the arrays exactly
"""
import os

import numpy as np
import pytest

from conftest import needs_ref
from william_ci_amd import formats
from william_ci_amd import system as tsys


def test_fixed_formats():
    assert formats.fortran_es(1.0) == " 1.0000000000000000E+000"
    assert formats.fortran_es(-0.12345678901234568) == "-1.2345678901234568E-001"
    assert formats.fortran_es(0.0) == " 0.0000000000000000E+000"
    assert formats.fortran_es(6.02e23) == " 6.0200000000000000E+023"
    assert formats.fortran_f(1.0) == "  1.000000000000"
    assert formats.fortran_f(-0.123456789012345) == " -0.123456789012"
    # gfortran list-directed REAL(8): G25.17E3 after a separator blank
    assert formats.list_directed_real(1.0) == "   1.0000000000000000     "
    assert formats.list_directed_real(0.5) == "  0.50000000000000000     "
    assert formats.list_directed_real(123.456) == "   123.45600000000000     "
    assert formats.list_directed_real(-1.0) == "  -1.0000000000000000     "
    assert formats.list_directed_real(1.5e-12) == "   1.5000000000000001E-012"   # 17 digits of the double
    for x in (1.0, 0.5, 123.456, -1.0, 1.5e-12, 0.1, 7.3e-5, -9.99999999999999999e16):
        assert float(formats.list_directed_real(x)) == x


def _parse_eri(path):
    rows = []
    with open(path) as f:
        for line in f:
            t = line.split()
            rows.append((int(t[0]), int(t[1]), int(t[2]), int(t[3]), float(t[4])))
    return rows


@needs_ref
@pytest.mark.parametrize("sao", [False, True])
def test_details_files_match_reference(tmp_path, sao):
    from oracle import ref
    s = tsys.chain_1d(2, 1.8, 3.6) if not sao else tsys.chain_1d(3, 1.8, 5.4)
    r = ref.RefSystem(s, sao=sao)
    S, T, V = r.one_electron()
    E = r.eri_full()                                      # E[l,k,j,i] = ERI(i,j,k,l)
    dref, dour = tmp_path / "ref", tmp_path / "ours"
    os.makedirs(dref)
    r.details(str(dref), S, T, V, E)
    formats.write_details(str(dour), S, T, V, E.transpose(3, 2, 1, 0), sao=sao)
    for name in ("OV", "KI", "NA"):
        for suffix in (".dat", "_matrix.dat"):
            a = open(dref / (name + suffix), "rb").read()
            b = open(dour / (name + suffix), "rb").read()
            assert a == b, name + suffix
    er, eo = _parse_eri(dref / "ERI.dat"), _parse_eri(dour / "ERI.dat")
    assert len(er) == len(eo) > 0
    assert er == eo                                       # same indices, order and values
    # Read: read_integrals_from_file on the reference's files
    nb = s.nbas
    S2, T2, V2, H2, G = formats.read_integrals(str(dref), nb)

    def upper(M):
        # what the files hold: the upper triangle (j >= i) mirrored, the full
        # layout without |x| <= 1e-15
        U = np.triu(M)
        if not sao:
            U = np.where(np.abs(U) > 1e-15, U, 0.0)
        return U + np.triu(U, 1).T
    assert np.array_equal(S2, upper(S)) and np.array_equal(T2, upper(T)) and np.array_equal(V2, upper(V))
    assert np.array_equal(H2, upper(T) + upper(V))
    n1 = E.shape[3]
    want = np.zeros((nb, nb, nb, nb))
    Ef = E.transpose(3, 2, 1, 0)
    want[:n1] = np.where(np.abs(Ef) > 1e-15, Ef, 0.0)
    assert np.array_equal(G, want)
