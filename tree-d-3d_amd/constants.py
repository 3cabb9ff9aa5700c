"""Physical constants (SI), as in the reference's constants.py:4."""

#: Coulomb constant 1 / (4 pi eps_0), in N m^2 / C^2.
K_E: float = 8.9875517923e9
