"""MI355X engine for the Clifford-Gaussian (toroidal) integral path of
William_CI: two-electron integrals, Coulomb/exchange and the energies that
consume them, on one GPU through the C ABI in include/tcg.h."""
