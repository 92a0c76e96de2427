"""Two-particle SPH sums written out by hand from the textbook formulation, independent of the oracle.

Monaghan (1992, ARA&A 30, 543), eqs. (3.3)-(3.4) and (4.1)-(4.2) for the symmetric momentum and energy
equations with his artificial viscosity, in the convention of the engine's CPU step (hydrostatic.rs:401-446):

    rho_i     = m_i W(0, h_i) + sum_j m_j W(r_ij, hbar_ij),          hbar_ij = (h_i + h_j) / 2
    a_i       = - sum_j m_j (P_i / rho_i^2 + P_j / rho_j^2 + Pi_ij) grad_i W_ij
    du_i / dt = 1/2 sum_j m_j (P_i / rho_i^2 + P_j / rho_j^2 + Pi_ij) v_ij . grad_i W_ij
    Pi_ij     = (-alpha cbar mu_ij + beta mu_ij^2) / rhobar   if v_ij . r_ij < 0, else 0
    mu_ij     = hbar v_ij . r_ij / (r_ij^2 + 0.01 hbar^2)

with r_ij = x_i - x_j, v_ij = v_i - v_j, grad_i W_ij = (r_ij / r) dW/dr, the M4 cubic spline of support
radius h (sigma = 8 / (pi h^3); W = sigma (1 - 6 q^2 + 6 q^3) for q < 1/2, 2 sigma (1 - q)^3 for q < 1)
and the Tillotson compressed-branch pressure (used for u <= E_iv at any density):
    P = (a + b / omega) rho u + A mu + B mu^2,   eta = rho / rho0, mu = eta - 1, omega = 1 + u / (E0 eta^2),
whose sound speed c^2 = dP/drho + (P / rho^2) dP/du is differentiated analytically here (the engine takes a
central difference; the two agree to O(step^2)).
"""
import numpy as np


def spline(r, h):
    q = r / h
    s = 8.0 / (np.pi * h ** 3)
    if q < 0.5:
        return s * (1.0 - 6.0 * q * q + 6.0 * q ** 3), s / h * (-12.0 * q + 18.0 * q * q)
    if q < 1.0:
        return 2.0 * s * (1.0 - q) ** 3, -6.0 * s / h * (1.0 - q) ** 2
    return 0.0, 0.0


def tillotson_compressed(t, rho, u):
    """(P, c) of the compressed branch; t = (rho0, a, b, A, B, E0, ...)."""
    rho0, a, b, A, B, E0 = t[:6]
    eta = rho / rho0
    mu = eta - 1.0
    om = 1.0 + u / (E0 * eta * eta)
    p = (a + b / om) * rho * u + A * mu + B * mu * mu
    dom_drho = -2.0 * u / (E0 * eta ** 3 * rho0)
    dp_drho = (a + b / om) * u - b * rho * u / om ** 2 * dom_drho + (A + 2.0 * B * mu) / rho0
    dp_du = (a + b / om) * rho - b * rho * u / om ** 2 / (E0 * eta * eta)
    return p, np.sqrt(dp_drho + p / (rho * rho) * dp_du)


# Plausible restatement mistakes `pair` can make on request (used to show the tests would catch them).
MUTATIONS = ("p_over_rho", "mu_no_eta", "mu_over_r", "dudt_no_half", "pi_rho_i", "grad_sign", "pair_h_max")


def pair(x, v, m, h, u, eos, alpha, beta, mutation=None):
    """Density, acceleration and du/dt of two particles (arrays of two rows / two entries) and the
    intermediate mu and Pi, by the formulas above (`mutation`: one of MUTATIONS, or None)."""
    x = np.asarray(x, np.float64)
    v = np.asarray(v, np.float64)
    rij = x[0] - x[1]
    vij = v[0] - v[1]
    r = float(np.sqrt(rij @ rij))
    hb = max(h[0], h[1]) if mutation == "pair_h_max" else 0.5 * (h[0] + h[1])
    w, dw = spline(r, hb)
    rho = np.array([m[0] * spline(0.0, h[0])[0] + m[1] * w, m[1] * spline(0.0, h[1])[0] + m[0] * w])
    pc = [tillotson_compressed(eos[k], rho[k], u[k]) for k in range(2)]
    p = np.array([q[0] for q in pc])
    c = np.array([q[1] for q in pc])
    vr = float(vij @ rij)
    eta2 = 0.0 if mutation == "mu_no_eta" else 0.01 * hb * hb
    mu = hb * vr / ((r if mutation == "mu_over_r" else r * r) + eta2) if vr < 0.0 else 0.0
    rho_bar = rho[0] if mutation == "pi_rho_i" else 0.5 * (rho[0] + rho[1])
    pi = (-alpha * 0.5 * (c[0] + c[1]) * mu + beta * mu * mu) / rho_bar if vr < 0.0 else 0.0
    k = 1 if mutation == "p_over_rho" else 2
    coeff = p[0] / rho[0] ** k + p[1] / rho[1] ** k + pi
    grad = (-1.0 if mutation == "grad_sign" else 1.0) * rij / r * dw
    acc = np.array([-m[1] * coeff * grad, m[0] * coeff * grad])
    vg = float(vij @ grad)
    half = 1.0 if mutation == "dudt_no_half" else 0.5
    dudt = np.array([half * m[1] * coeff * vg, half * m[0] * coeff * vg])
    return {"rho": rho, "p": p, "c": c, "mu": mu, "pi": pi, "acc": acc, "dudt": dudt}


# The cases: (name, separation as a fraction of hbar, approach speed in m/s (> 0 approaching), alpha, beta,
# materials). Separations 0.4 and 0.7 cover both spline branches.
CASES = [
    ("pressure_receding", 0.4, -800.0, 1.0, 2.0, ("granite", "granite")),
    ("linear_viscosity", 0.4, 900.0, 1.0, 0.0, ("granite", "iron")),
    ("quadratic_viscosity", 0.7, 900.0, 0.0, 2.0, ("iron", "granite")),
    ("full_viscosity", 0.7, 1500.0, 1.0, 2.0, ("granite", "iron")),
]


def case_state(sep, approach, names):
    """Two particles with masses set from their materials' rho0 (after the density sum a mixed-material
    pair may leave one particle in tension, which the compressed branch covers), u below E_iv, the pair off
    every axis so each vector component is exercised, f32-representable inputs. Returns (x, v, m, h, u,
    eos10 rows)."""
    from botheadstudios_amd import materials
    eos = [materials.CATALOGUE[n] for n in names]
    h = np.array([np.float32(9000.0), np.float32(11000.0)], np.float64)
    hb = 0.5 * (h[0] + h[1])
    d = np.array([0.6, -0.48, 0.64])  # unit vector
    x = np.zeros((2, 3))
    x[1] = (sep * hb) * d
    # masses giving rho ~ 1.05 rho0 from the self term and the pair term at this separation
    w = spline(sep * hb, hb)[0]
    m = np.zeros(2)
    for k in range(2):
        m[k] = 1.05 * eos[k][0] / (spline(0.0, h[k])[0] + w)
    v = np.zeros((2, 3))
    v[0] = 0.5 * approach * d + np.array([120.0, 40.0, -75.0])
    v[1] = -0.5 * approach * d
    u = np.array([2.0e5, 3.5e5])
    f32 = lambda a: np.asarray(a, np.float32).astype(np.float64)
    return f32(x), f32(v), f32(m), f32(h), f32(u), eos
