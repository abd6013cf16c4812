"""Goodier (1933): a circular elastic inclusion (radius a) in an infinite
elastic plane under remote uniaxial stress sigma0 along x, plane strain.  The
reference's own known-answer drivers use this solution (Tests/IGFEM_test1.cpp,
Tests/IGFEM_test2.cpp); it is restated here from the Kolosov-Muskhelishvili
potentials (2 mu (u_x + i u_y) = kappa phi - z conj(phi') - conj(psi),
kappa = 3 - 4 nu), with the remote field split into an equibiaxial part
p = sigma0/2 and a pure-shear part S = sigma0/2:

* equibiaxial: u_r = A r inside, u_r = B r + C / r outside, with
  B = p / (2 (lam_m + mu_m)) and C = a^2 B (lam_m + mu_m - lam_i - mu_i) /
  (lam_i + mu_i + mu_m), A = B + C / a^2 (continuity of u_r and sigma_rr);
* pure shear: outside phi = beta a^2 / z, psi = -S z + beta a^4 / z^3; inside
  phi = 0, psi = c z; continuity of the traction resultant and of the
  displacement on |z| = a gives beta = S (mu_m - mu_i) / (mu_i kappa_m + mu_m),
  c = beta - S.
"""
import numpy as np


def lame(E, nu):
    return E * nu / ((1 + nu) * (1 - 2 * nu)), E / (2 * (1 + nu))


def displacement(x, y, xc, yc, a, sigma0, matrix, inclusion):
    """exact (u_x, u_y) at points (x, y) for an inclusion centred at (xc, yc)"""
    lm, mm = lame(*matrix)
    li, mi = lame(*inclusion)
    km = 3 - 4 * matrix[1]
    p = S = 0.5 * sigma0
    B = p / (2 * (lm + mm))
    C = a * a * B * (lm + mm - li - mi) / (li + mi + mm)
    A = B + C / (a * a)
    beta = S * (mm - mi) / (mi * km + mm)
    c = beta - S
    z = (np.asarray(x, float) - xc) + 1j * (np.asarray(y, float) - yc)
    inside = np.abs(z) < a
    zo = np.where(inside, a + 0j, z)  # keeps the outside formula finite inside
    cz = np.conj(zo)
    u_out = B * zo + C / cz + (km * beta * a * a / zo + zo * beta * a * a / cz ** 2 + S * cz
                               - beta * a ** 4 / cz ** 3) / (2 * mm)
    u_in = A * z - c * np.conj(z) / (2 * mi)
    u = np.where(inside, u_in, u_out)
    return u.real, u.imag
