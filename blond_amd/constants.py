"""Physical constants, CODATA 2022 (the values scipy.constants 1.15 carries,
which the reference imports: blond/beam/beam.py, blond/impedances/impedance.py:22)."""
c = 299792458.0
e = 1.602176634e-19
m_p = 1.67262192595e-27
m_e = 9.1093837139e-31
m_mu = 1.883531627e-28
epsilon_0 = 8.8541878188e-12
hbar = 1.0545718176461565e-34
