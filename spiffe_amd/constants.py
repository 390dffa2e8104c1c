"""Physical constants (SI, CODATA 2018).

spiffe takes these from the SDDS toolkit's constants header, which the
reference does not vendor; this build fixes them here and passes them to the
device library explicitly, so the host, the device and the CPU restatement all
use the same bits.
"""
PI = 3.141592653589793238462643
c_mks = 299792458.0
e_mks = 1.602176634e-19
me_mks = 9.1093837015e-31
mu_o = 1.25663706212e-6
epsilon_o = 8.8541878128e-12
k_boltzmann_mks = 1.380649e-23
# electron charge over mass times c, src/cathode.c:200
QoMC_electron = -e_mks / (me_mks * c_mks)
