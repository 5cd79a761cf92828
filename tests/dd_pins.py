"""Independent pins of the drift-diffusion (Scharfetter-Gummel) registry entry.

Two facts of the continuous problem fix the discrete flux without restating it:

* Thermodynamic equilibrium.  A carrier of charge number z with density
  c = exp(-z psi) carries no current, so at zero bias (n = e^psi, p = e^-psi for
  z_n = -1, z_p = +1) every carrier flux vanishes, and so does the recombination
  (np - n_i^2)/(n + p + 2 n_i) for n_i = 1.  A swapped B(x) / B(-x) pair or a flipped
  charge number turns the equilibrium into c = exp(+z psi) and leaves O(1) residuals.
* The two-node boundary-value problem.  On an edge of length h with psi linear,
  the current j = -(c' + z c psi') is constant, and c(0) = c_K, c(h) = c_L fix it.
  It is obtained here by integrating the ODE numerically (no Bernoulli function
  involved); the discrete edge flux divided by h must equal it
  (examples/Example160_UnipolarDriftDiffusion1D.jl:43-50 for the flux form,
  src/vfvm_functions.jl:78-90 for fbernoulli_pm).
"""
import numpy as np


def sg_current_ode(cK, cL, psiK, psiL, z, h):
    """Constant current j (direction K -> L) of c' = -j - z psi' c on [0, h] with
    psi linear between psiK and psiL and c(0) = cK, c(h) = cL.  c(h) is affine in j,
    so two integrations (j = 0 and j = 1) give it."""
    from scipy.integrate import solve_ivp

    dpsi = (psiL - psiK) / h

    def end(j, c0):
        sol = solve_ivp(lambda s, c: -j - z * dpsi * c, (0.0, h), [c0], method="DOP853", rtol=1e-13,
                        atol=1e-15)
        return sol.y[0, -1]

    c_hom = end(0.0, cK)
    c_one = end(1.0, cK) - c_hom
    return (cL - c_hom) / c_one


def equilibrium_state(nnodes, rng, psi_range=3.0):
    """(psi, n, p) per node, species fastest, at thermodynamic equilibrium: n = e^psi,
    p = e^-psi with psi random in [-psi_range, psi_range]."""
    psi = rng.uniform(-psi_range, psi_range, nnodes)
    U = np.zeros((3, nnodes), order="F")
    U[0] = psi
    U[1] = np.exp(psi)
    U[2] = np.exp(-psi)
    return U
