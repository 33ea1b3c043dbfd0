"""NumPy FP64 restatement of the flux-marginalised log-likelihood.
TEST INFRASTRUCTURE ONLY.

Follows (fbartolic/caustics):
  marginalized_log_likelihood ... src/caustics/linalg.py:10-70, dense_covariance=False:
      per light curve M = [A, 1], N = M^T diag(C_inv) M, Sigma = solve(N, I),
      beta = Sigma M^T diag(C_inv) f, ll += -1/2 sum (f - M beta)^2 C_inv
      + 1/2 log det(2 pi Sigma); returns (beta_list, ll).
  The products with diag(C_inv) are written as row scalings (M.T * C_inv) @ M, the same
  sums without the n x n matrix.  The dense branch (:72-94) does not run in the
  reference (`jnp.linals` at :80, `M` undefined at :84) and is not restated.

Pins (tests/test_oracle_linalg.py): the flat-prior marginal likelihood
log of the integral of exp(-chi2(beta)/2) over beta, which the diagonal branch equals
exactly, by brute-force quadrature; a hand-solved exact fit; central finite
differences of ll for grad_A.
"""
import numpy as np


def marginalized_log_likelihood(A_list, fobs_list, C_inv_list):
    ll = 0.0
    beta_list = []
    for A, fobs, C_inv in zip(A_list, fobs_list, C_inv_list):
        A = np.asarray(A, float)
        fobs = np.asarray(fobs, float)
        C_inv = np.asarray(C_inv, float)
        M = np.stack([A, np.ones_like(A)]).T
        MTCinvM = (M.T * C_inv) @ M
        Sigma = np.linalg.solve(MTCinvM, np.eye(2))
        beta = Sigma @ ((M.T * C_inv) @ fobs[:, None])
        fpred = (M @ beta).reshape(-1)
        ll += -0.5 * np.sum((fobs - fpred) ** 2 * C_inv) + 0.5 * np.log(
            np.linalg.det(2 * np.pi * Sigma))
        beta_list.append(beta.reshape(-1))
    return beta_list, ll


def grad_A(A_list, fobs_list, C_inv_list):
    """d ll / d A for each light curve: the envelope term C_inv (f - M beta) F_s plus the
    derivative of 1/2 log det(2 pi Sigma) = const - 1/2 log det N, which is
    -1/2 tr(Sigma dN/dA_i) = -C_inv_i (A_i Sigma_00 + Sigma_01)."""
    out = []
    for A, fobs, C_inv in zip(A_list, fobs_list, C_inv_list):
        A = np.asarray(A, float)
        fobs = np.asarray(fobs, float)
        C_inv = np.asarray(C_inv, float)
        M = np.stack([A, np.ones_like(A)]).T
        N = (M.T * C_inv) @ M
        Sigma = np.linalg.solve(N, np.eye(2))
        beta = Sigma @ ((M.T * C_inv) @ fobs)
        r = fobs - M @ beta
        out.append(C_inv * r * beta[0] - C_inv * (A * Sigma[0, 0] + Sigma[0, 1]))
    return out


def marginal_ll_quadrature(A, fobs, C_inv, half_width=9.0, n=801):
    """log of the integral over (F_s, F_b) of exp(-1/2 sum C_inv (f - F_s A - F_b)^2),
    by brute force: chi2 evaluated from the data on an n x n grid spanning +-half_width
    standard deviations around the least-squares point along the principal axes of the
    posterior, summed with the trapezoid rule.  Independent of the closed form except
    for the grid placement."""
    A = np.asarray(A, float)
    fobs = np.asarray(fobs, float)
    C_inv = np.asarray(C_inv, float)
    M = np.stack([A, np.ones_like(A)]).T
    N = (M.T * C_inv) @ M
    b0 = np.linalg.lstsq(M * np.sqrt(C_inv)[:, None], fobs * np.sqrt(C_inv), rcond=None)[0]
    evals, evecs = np.linalg.eigh(N)
    sd = 1.0 / np.sqrt(evals)
    t = np.linspace(-half_width, half_width, n)
    u, v = np.meshgrid(t * sd[0], t * sd[1], indexing="ij")
    B0 = b0[0] + evecs[0, 0] * u + evecs[0, 1] * v
    B1 = b0[1] + evecs[1, 0] * u + evecs[1, 1] * v
    # chi2 from the data, one grid plane per epoch
    chi2 = np.zeros_like(B0)
    for i in range(A.size):
        chi2 += C_inv[i] * (fobs[i] - B0 * A[i] - B1) ** 2
    chi2_min = chi2.min()
    w = np.ones(n)
    w[0] = w[-1] = 0.5
    h = (t[1] - t[0]) ** 2 * sd[0] * sd[1]   # area element; |det evecs| = 1
    integral = h * np.einsum("i,j,ij->", w, w, np.exp(-0.5 * (chi2 - chi2_min)))
    return np.log(integral) - 0.5 * chi2_min
