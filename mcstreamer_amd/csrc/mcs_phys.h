/*
 * mcs_phys.h -- physical constants and integrator coefficients of the
 * reference, as exact double literals shared by host and device code.
 *
 * Constants are the values in Constants/src/Constants.jl:52-57,151-154
 * (scipy's CODATA-2014 table).  The Yoshida coefficients are those of
 * JuMC/src/electron.jl:38-46 evaluated in double precision:
 *   w1 = 1/(2 - 2^(1/3)), w0 = -2^(1/3)/(2 - 2^(1/3)),
 *   c1 = c4 = w1/2, c2 = c3 = (w0 + w1)/2, d1 = d3 = w1, d2 = w0.
 */
#ifndef MCS_PHYS_H
#define MCS_PHYS_H

#define MCS_ELECTRON_MASS 9.10938356e-31
#define MCS_ELEMENTARY_CHARGE 1.6021766208e-19
#define MCS_EV 1.6021766208e-19
#define MCS_EPSILON_0 8.854187817620389e-12
#define MCS_TD 1e-21
#define MCS_NAIR 2.6867811e+25
#define MCS_C 299792458.0   /* speed of light, Constants.jl:39; the photon speed (collisions.jl:136) */

/* charge(p)/mass(p) of an electron, (-e)/m_e rounded once. */
#define MCS_QM (-175882002360.2124)
/* 0.5 * mass(p), the first product of energy() (electron.jl:27). */
#define MCS_HALF_ME 4.55469178e-31

#define MCS_YOSHIDA_W1 1.3512071919596578
#define MCS_YOSHIDA_W0 (-1.7024143839193153)
#define MCS_YOSHIDA_C1 0.6756035959798289
#define MCS_YOSHIDA_C2 (-0.17560359597982877)
#define MCS_YOSHIDA_C3 MCS_YOSHIDA_C2
#define MCS_YOSHIDA_C4 MCS_YOSHIDA_C1
#define MCS_YOSHIDA_D1 MCS_YOSHIDA_W1
#define MCS_YOSHIDA_D2 MCS_YOSHIDA_W0
#define MCS_YOSHIDA_D3 MCS_YOSHIDA_W1

/* |E| above this aborts the reference (gridfields.jl:179); compared squared. */
#define MCS_EFIELD_LIMIT2 1e16

#endif /* MCS_PHYS_H */
