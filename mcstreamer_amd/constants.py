"""Physical constants of the reference (Constants/src/Constants.jl:52-57,151-156),
scipy's CODATA-2014 values."""

electron_mass = 9.10938356e-31
elementary_charge = 1.6021766208e-19
eV = 1.6021766208e-19
epsilon_0 = 8.854187817620389e-12
centi = 0.01
Td = 1e-21
nair = 2.6867811e25
bar = 100000.0
k = 1.38064852e-23
c = 299792458.0
