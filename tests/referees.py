"""Independent referees for the charm first- and second-moment sources.

Test infrastructure. Two kinds:

1. The reference's 1-D (radial) source functions, which its 2+1D gates use as
   the referee of the 2-D sources in the axisymmetric limit:
     hq_consistent_extras  (src/hq_consistent_firstmoment.jl:104-124)
     hq_consistent_m2_rhs  (src/hq_consistent_m2.jl:102-252)
   restated here in numpy with the scaled Bessel functions from scipy (AMOS,
   the library the reference calls).

2. Brute-force covariant tensor algebra in Milne coordinates (tau, x, y, eta),
   metric g = diag(-1, 1, 1, tau^2), boost invariant (u^eta = 0, d_eta = 0): every
   tensor the 2-D sources contract is built from its definition with explicit
   4x4 index sums, so a transposed slot, a wrong sign or a dropped term in the
   generated closed forms of the solver shows up as a mismatch. This is the
   reference's gate Gd3 (test/test_dnmr2d.jl:107-139) carried to every term of
   the second-moment source.
"""
import numpy as np
from scipy.special import kve


# ---- 1-D referees --------------------------------------------------------------
def fm_extras_1d(tau, r, ur, T, dtT, drT, drur, dtur, n, dn_dT, nu, taun, Ds, h, hp):
    """hq_consistent_extras with every term on (src/hq_consistent_firstmoment.jl:111-123)."""
    ut = np.sqrt(1.0 + ur * ur)
    ex = (Ds / T) * (drur * h * n * ur + dtur * h * n * ut + dtT * n * ur * ut + dtT * dn_dT * T * ur * ut
                     + drT * (n + dn_dT * T) * ut ** 2)
    ex += (Ds / T) * nu * (2.0 * drur * h + drT * hp * ur + (2.0 * dtur * h * ur + dtT * hp * ut ** 2) / ut)
    ex += taun * nu * (ut / tau + ur / r)
    return ex


def m2_coeffs_1d(T, n, taun, Ds, h, m):
    """hq_m2_coeffs (src/hq_consistent_m2.jl:102-121)."""
    z = m / T
    K2, K3, K4 = kve(2, z), kve(3, z), kve(4, z)
    K5 = K3 + 8 * K4 / z
    dlnC = z * ((K3 + K5) / (2 * K4) - (K2 + K4) / (2 * K3))
    return dict(dlnC=dlnC, etabar=taun * n * T / 2, lam_a=Ds * (m * m + 6 * T * h) / (2 * T),
                Aco=m * m + 5 * T * h, ABr=5 + z * K2 / K3)


def m2_rhs_1d(tau, r, ur, T, dtT, drT, drur, dtur, p_l, p_phi, PiQ, nu, dta, dra, dtnu, drnu, drp,
              n, taun, Ds, h, hp, tauM, etaM, m):
    """hq_consistent_m2_rhs with every term on (src/hq_consistent_m2.jl:138-207,
    236-251): d_tau of (p_l, p_phi, Pi_Q) in the transported triad."""
    ut = np.sqrt(1.0 + ur * ur)
    p_eta = -(p_l + p_phi)
    Dlur = ur * dtur + ut * drur
    th_l, th_phi, th_eta = Dlur / ut, ur / r, ut / tau
    th = th_l + th_phi + th_eta
    s_l, s_phi, s_eta = th_l - th / 3, th_phi - th / 3, th_eta - th / 3
    a_l = (ut * dtur + ur * drur) / ut
    nu_l = nu / ut
    Dlnu = ur * dtnu + ut * drnu
    gll = Dlnu / ut - ur * nu * Dlur / ut ** 3
    gpp = nu / r
    anu = a_l * nu_l
    thnu = (ur / ut) * dtnu + nu * dtur / ut ** 3 + drnu + ur * nu / (ut * tau) + nu / r
    trnu = thnu - anu
    snl, snp = gll - trnu / 3, gpp - trnu / 3
    etaQ = etaM
    zetaQ = 5.0 / 3.0 * etaQ
    cf = m2_coeffs_1d(T, n, taun, Ds, h, m)
    DlnT = (ut * dtT + ur * drT) / T
    DlnC = cf["dlnC"] * DlnT
    DlTh = (h + T * hp) * (ur * dtT + ut * drT)
    ps = p_l * s_l + p_phi * s_phi + p_eta * s_eta
    geo = tauM * (5.0 / 3.0 * th + DlnC)
    rank1 = 2 * cf["lam_a"] * a_l * nu_l + (Ds / T) * nu_l * DlTh
    src_l = (p_l + 2 * etaQ * snl + geo * p_l + 2 * tauM * (p_l * s_l - ps / 3) + 2 * tauM * PiQ * s_l
             + 2 * cf["etabar"] * s_l + rank1 * (2.0 / 3.0))
    src_p = (p_phi + 2 * etaQ * snp + geo * p_phi + 2 * tauM * (p_phi * s_phi - ps / 3) + 2 * tauM * PiQ * s_phi
             + 2 * cf["etabar"] * s_phi + rank1 * (-1.0 / 3.0))
    src_B = (PiQ + zetaQ * thnu + geo * PiQ + (2.0 / 3.0) * tauM * ps
             + cf["etabar"] * (cf["ABr"] * DlnT + 5.0 / 3.0 * th)
             + (Ds * cf["Aco"] / (3 * T)) * anu + (5.0 / 6.0) * (Ds / T) * nu_l * DlTh)
    src_B += cf["etabar"] * (ut * dta + ur * dra)
    den = tauM * ut
    return (-(tauM * ur * drp[0] + src_l) / den, -(tauM * ur * drp[1] + src_p) / den,
            -(tauM * ur * drp[2] + src_B) / den)


# ---- brute-force Milne tensor algebra -------------------------------------------
class Milne:
    """A boost-invariant flowing state at one point: u^i, nu^i and their first
    derivatives d_mu (mu = tau, x, y). Index order (tau, x, y, eta)."""

    def __init__(self, tau, ux, uy, du, nux=0.0, nuy=0.0, dnu=None):
        """du[mu] = (d_mu u^x, d_mu u^y), dnu likewise, for mu = tau, x, y."""
        self.tau = tau
        ut = np.sqrt(1 + ux * ux + uy * uy)
        self.g = np.diag([-1.0, 1.0, 1.0, tau * tau])
        self.gi = np.diag([-1.0, 1.0, 1.0, 1.0 / (tau * tau)])
        self.u = np.array([ut, ux, uy, 0.0])
        self.Du = self._cov_grad(self.u, du)
        self.theta = np.trace(self.Du)
        self.a = self.u @ self.Du                     # a^nu = u^mu nabla_mu u^nu
        self.Delta = self.gi + np.outer(self.u, self.u)
        self.Dmix = np.eye(4) + np.outer(self.u, self.g @ self.u)   # Delta^mu_nu
        nut = (ux * nux + uy * nuy) / ut
        self.nu = np.array([nut, nux, nuy, 0.0])
        self.Dnu = self._cov_grad(self.nu, dnu if dnu is not None else np.zeros((3, 2)), ortho=True, du=du)

    def _cov_grad(self, v, dv, ortho=False, du=None):
        """nabla_mu v^nu (4x4, row mu): partial derivatives (tau-component from
        normalisation or orthogonality by the chain rule) plus the Milne
        Christoffel Gamma^eta_{eta tau} = 1/tau (only nabla_eta v^eta = v^tau/tau
        survives with v^eta = 0)."""
        ux, uy, ut = self.u[1], self.u[2], self.u[0]
        M = np.zeros((4, 4))
        for mu in range(3):
            dvx, dvy = dv[mu]
            M[mu, 1], M[mu, 2] = dvx, dvy
            if not ortho:
                M[mu, 0] = (ux * dvx + uy * dvy) / ut
            else:
                dux, duy = du[mu]
                dut = (ux * dux + uy * duy) / ut
                num = ux * v[1] + uy * v[2]
                M[mu, 0] = (dux * v[1] + ux * dvx + duy * v[2] + uy * dvy) / ut - num * dut / ut ** 2
        M[3, 3] = v[0] / self.tau
        return M

    def lower(self, v):
        return self.g @ v

    def dot(self, a, b):
        return a @ self.g @ b

    def up_deriv(self, D):
        """g^{alpha alpha} nabla_alpha v^nu: the derivative index raised."""
        return self.gi @ D

    def traceless_sym(self, X):
        """Delta^mu_a Delta^nu_b X^(ab) - Delta^{mu nu} (Delta_ab X^ab) / 3."""
        S = 0.5 * (X + X.T)
        P = self.Dmix @ S @ self.Dmix.T
        tr = np.einsum("ab,ab->", self.g @ self.Delta @ self.g, S)   # Delta_ab S^ab
        return P - self.Delta * tr / 3

    def sigma(self):
        """sigma^{mu nu} of the flow."""
        return self.traceless_sym(self.up_deriv(self.Du))

    def sigma_nu(self):
        """sigma_(nu)^{mu nu}: Delta Delta nabla^(a nu^b) - Delta (theta_nu - a.nu)/3."""
        X = self.up_deriv(self.Dnu)
        P = self.Dmix @ (0.5 * (X + X.T)) @ self.Dmix.T
        thnu = np.trace(self.Dnu)
        return P - self.Delta * (thnu - self.dot(self.a, self.nu)) / 3, thnu

    def omega(self):
        """omega^{mu nu} = (A^{mu nu} - A^{nu mu}) / 2, A^{alpha nu} = nabla_perp^alpha u^nu."""
        A = self.up_deriv(self.Du) + np.outer(self.u, self.a)
        return 0.5 * (A - A.T)

    def rank1(self, A, B):
        """A^<mu B^nu> for u-orthogonal vectors."""
        return self.traceless_sym(np.outer(A, B))

    def ortho_vec(self, vx, vy):
        """(v^tau, v^x, v^y, 0) with v^tau from u.v = 0."""
        ux, uy, ut = self.u[1], self.u[2], self.u[0]
        return np.array([(ux * vx + uy * vy) / ut, vx, vy, 0.0])

    def ortho_tensor(self, pxx, pxy, pyy, peta):
        """pi^{mu nu} from the stored (pixx, pixy, piyy, pi^eta_eta), tau-row from u.pi = 0."""
        ux, uy, ut = self.u[1], self.u[2], self.u[0]
        tx = (ux * pxx + uy * pxy) / ut
        ty = (ux * pxy + uy * pyy) / ut
        tt = (ux * tx + uy * ty) / ut
        return np.array([[tt, tx, ty, 0], [tx, pxx, pxy, 0], [ty, pxy, pyy, 0],
                         [0, 0, 0, peta / self.tau ** 2]])

    def implied_peta(self, pxx, pxy, pyy):
        """pi^eta_eta that makes the stored block traceless (u-orthogonal)."""
        P = self.ortho_tensor(pxx, pxy, pyy, 0.0)
        return P[0, 0] - pxx - pyy

    def contract(self, A, B):
        """A^{mu nu} B_{mu nu}."""
        return np.einsum("ab,ab->", A, self.g @ B @ self.g)

    def single(self, A, B):
        """A^{mu lambda} B_lambda^nu, symmetrised."""
        C = A @ self.g @ B
        return 0.5 * (C + C.T)

    def block(self, X):
        """(xx, xy, yy, tau^2 X^{eta eta}) of a rank-2 tensor."""
        return np.array([X[1, 1], X[1, 2], X[2, 2], X[3, 3] * self.tau ** 2])
