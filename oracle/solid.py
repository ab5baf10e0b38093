"""CPU restatement of the beam solver of SolidSolver.f90 (numpy).

TEST INFRASTRUCTURE ONLY: imported by tests/.  The product package
fsibe_2d_amd never imports this module.

It follows the reference routine by routine, with the reference's dense
algebra: the difference matrices are built as create_diff_matrices builds
them and applied as matrix-vector products (Differentiate1D), the tension
system is solved densely, and the Newton step solves transpose(J) dx = -F with
LAPACK's dense LU (numpy.linalg.solve), as Solve_LGS "DIRECT" does.  The
residual accepts complex iterates, so that the complex-step derivative of
F_nonlinear gives the exact Jacobian to compare with the analytic Jacobi.

Kinematics (mouvement, integrate_position) come from the fluid oracle
(oracle/fsibe_oracle.c), whose versions are pinned bit for bit against the
reference's Fortran.

Pinning (tests/test_solid.py): F_nonlinear against its complex-step
Jacobian vs the analytic Jacobi; the converged step's residual; a straight
unloaded beam stays straight; the static cantilever under gravity against
the Euler-Bernoulli deflection mue g L^4 / (8 eta); the first free-vibration
frequency against (1.8751)^2 sqrt(eta / mue) for implicit Euler,
Crank-Nicolson and BDF2; RHS_beameqn's acceleration of a static equilibrium
is zero; the tension of a beam rotating at constant rate against
T(s) = mue w^2 (L^2 - s^2) / 2.
"""
from __future__ import annotations

import numpy as np

EulerImplicit, CrankNicholson, BDF2 = 1, 2, 5


def diff_matrices(nn: int, dx: float):
    """create_diff_matrices (SolidSolver.f90:1110-1233), dense"""
    D1, D2, D3, D4 = (np.zeros((nn, nn)) for _ in range(4))
    i = np.arange(nn)
    D2[i, i] = -2.0
    D4[i, i] = 6.0
    j = np.arange(nn - 1)
    D1[j, j + 1], D1[j + 1, j] = 1.0, -1.0
    D2[j, j + 1], D2[j + 1, j] = 1.0, 1.0
    D3[j, j + 1], D3[j + 1, j] = -2.0, 2.0
    D4[j, j + 1], D4[j + 1, j] = -4.0, -4.0
    k = np.arange(nn - 2)
    D3[k, k + 2], D3[k + 2, k] = 1.0, -1.0
    D4[k, k + 2], D4[k + 2, k] = 1.0, 1.0
    N = nn - 1
    D1[0, 0:3] = [-3.0, 4.0, -1.0]
    D1[N, N - 2:N + 1] = [1.0, -4.0, 3.0]
    D1 /= 2.0 * dx
    D2[0, 0:4] = [2.0, -5.0, 4.0, -1.0]
    D2[N, N - 3:N + 1] = [-1.0, 4.0, -5.0, 2.0]
    D2 /= dx ** 2
    r3 = [-17.0 / 2.0, 71.0 / 2.0, -59.0, 49.0, -41.0 / 2.0, 7.0 / 2.0]
    D3[0, 0:6] = r3
    D3[1, 0] = 0.0
    D3[1, 1:7] = r3
    D3[N, N - 5:N + 1] = [-v for v in reversed(r3)]
    D3[N - 1, N - 6:N] = [-v for v in reversed(r3)]
    D3[N - 1, N] = 0.0
    D3 /= 2.0 * dx ** 3
    r4 = [35.0 / 6.0, -31.0, 137.0 / 2.0, -242.0 / 3.0, 107.0 / 2.0, -19.0, 17.0 / 6.0]
    D4[0, 0:7] = r4
    D4[1, 0] = 0.0
    D4[1, 1:8] = r4
    D4[N, N - 6:N + 1] = list(reversed(r4))
    D4[N - 1, N - 7:N] = list(reversed(r4))
    D4[N - 1, N] = 0.0
    D4 /= dx ** 4
    return D1, D2, D3, D4


class SolidOracle:
    """One beam (type(solid)) and the solver state.  `kin` is a fluid oracle
    (oracle.Oracle) whose parameters and beam carry the kinematics."""

    def __init__(self, kin, mue=1.0, eta=0.1, sigma=0.0, grav=0.0, TimeMethodSolid=BDF2, iImpulse=0,
                 T_release=None):
        self.kin = kin
        p = kin.p
        self.ns = ns = p.ns
        self.ds = p.length / (ns - 1)
        self.mue, self.eta, self.sigma, self.grav = mue, eta, sigma, grav
        self.method, self.iImpulse = TimeMethodSolid, iImpulse
        self.T_release = p.T_release if T_release is None else T_release
        self.y0 = p.y0
        self.Dn = diff_matrices(ns, self.ds)
        self.De = diff_matrices(ns + 3, self.ds)
        b = kin.beam_arrays()  # init_beams (init_beam.f90:80-120) done by the fluid oracle
        z = np.zeros(ns)
        self.x, self.y, self.vx, self.vy = b["x"], b["y"], b["vx"], b["vy"]
        self.theta, self.theta_dot = b["theta"], b["theta_dot"]
        self.ax, self.ay = z.copy(), z.copy()
        self.pressure_old, self.pressure_new, self.tau_old, self.tau_new = z.copy(), z.copy(), z.copy(), z.copy()
        self.oldold = np.zeros((6, ns))
        self.dt_old = 0.0
        self.StartupStep = True
        self.iter = 0
        self.brake = False

    # ---- kinematics of the fluid oracle -----------------------------------
    def motion(self, t):
        return self.kin.mouvement(t)

    def integrate(self, t):
        self.kin.set_beam_arrays(theta=self.theta, theta_dot=self.theta_dot)
        self.kin.integrate_position(t)
        b = self.kin.beam_arrays()
        self.x, self.y, self.vx, self.vy = b["x"], b["y"], b["vx"], b["vy"]

    # ---- GravityImpulse (SolidSolver.f90:874-911) -------------------------
    def gravity_impulse(self, time):
        if self.iImpulse in (1, 2):
            T0, T1 = 0.75, 0.85
            self.grav = (1.0 if self.iImpulse == 1 else 4.0) * np.sin(np.pi * (time - T0) / (T1 - T0))
            if time < T0 or time > T1:
                self.grav = 0.0
        elif self.iImpulse == 3:
            T0, T1 = 2.0, 3.0
            if time <= T0:
                k = 0.0
            elif time <= T1:
                t = (time - T0) / (T1 - T0)
                k = -20.0 * t ** 7 + 70.0 * t ** 6 - 84.0 * t ** 5 + 35.0 * t ** 4
            else:
                k = 1.0
            self.grav = 0.7 * k

    # ---- Tension (SolidSolver.f90:1000-1081) ------------------------------
    def tension(self, time, th, thd, pr, tau):
        a, at, att, LE = self.motion(time)
        D1, D2, D3, _ = self.Dn
        ns, ds, eta, mue = self.ns, self.ds, self.eta, self.mue
        th_s, th_ss, th_sss, tau_s = D1 @ th, D2 @ th, D3 @ th, D1 @ tau
        K1 = mue * (LE[4] * np.cos(a) + LE[5] * np.sin(a) + self.grav * np.sin(a)) - tau[0]
        T_s0 = -eta * th_ss[0] * th_s[0] + K1
        A = np.zeros((ns, ns))
        i = np.arange(ns)
        A[i, i] = -2.0 / ds ** 2 - th_s ** 2
        j = np.arange(ns - 1)
        A[j, j + 1] = 1.0 / ds ** 2
        A[j + 1, j] = 1.0 / ds ** 2
        A[ns - 1, ns - 1] = 1.0  # Dirichlet T(ns-1) = 0
        A[ns - 1, ns - 2] = 0.0
        A[0, 1] = 2.0 / ds ** 2  # Neumann at the leading edge (ghost T(-1) eliminated)
        rhs = -(pr * th_s) - (2.0 * eta * th_s * th_sss) - (eta * th_ss ** 2) - mue * (thd + at) ** 2 - tau_s
        rhs[ns - 1] = 0.0
        rhs[0] = rhs[0] + T_s0 * 2.0 / ds
        # the reference scales each row by its diagonal before the solve
        dg = np.diag(A).copy()
        T = np.linalg.solve(A / dg[:, None], rhs / dg)
        T_s = D1 @ T
        T_s[0] = T_s0
        return T, T_s

    # ---- RHS_beameqn (SolidSolver.f90:918-994) ----------------------------
    def rhs_beameqn(self, time, theta, theta_dot, pr, tau):
        """returns (theta_t, theta_tt, T)"""
        a, at, att, LE = self.motion(time)
        th, thd = np.array(theta, float), np.array(theta_dot, float)
        th[0], thd[0] = 0.0, 0.0
        T, T_s = self.tension(time, th, thd, pr, tau)
        ns, ds, eta, mue = self.ns, self.ds, self.eta, self.mue
        K2 = pr[0] + mue * (LE[5] * np.cos(a) - LE[4] * np.sin(a) + self.grav * np.cos(a))
        C2 = 2.0 * ds * K2 - T[0] * th[1] + (eta / ds ** 2) * (10.0 * th[0] - 12.0 * th[1] + 6.0 * th[2] - th[3])
        ext = np.zeros(ns + 3)
        ext[0] = C2 / ((3.0 * eta / ds ** 2) - T[0])
        ext[1:ns + 1] = th
        A1 = th[ns - 3] / 12.0 - 2.0 * th[ns - 2] / 3.0
        A2 = -th[ns - 3] / 12.0 + 4.0 * th[ns - 2] / 3.0 - 5.0 * th[ns - 1] / 2.0
        ext[ns + 1] = 3.0 * (A1 - A2) / 2.0
        ext[ns + 2] = 12.0 * (2.0 * A1 - A2)
        E1, E2, _, E4 = self.De
        s1, s2, s4 = (E1 @ ext)[1:ns + 1], (E2 @ ext)[1:ns + 1], (E4 @ ext)[1:ns + 1]
        p_s = self.Dn[0] @ pr
        acc = (-p_s - eta * s4 + s2 * (T + eta * s1 ** 2) + 2.0 * T_s * s1 - mue * att - self.sigma * thd
               + tau * s1) / mue
        acc[0] = 0.0
        return thd, acc, T

    # ---- scheme constants (SolidSolver.f90:577-594) -----------------------
    @staticmethod
    def scheme(m, dt, dt_old):
        if m == EulerImplicit:
            return 1.0, 0.0, 1.0, 0.0
        if m == CrankNicholson:
            return 2.0, 1.0, 1.0, 0.0
        R = dt / dt_old
        return (1.0 + 2.0 * R) / (1.0 + R), 0.0, (1.0 + R) ** 2 / (1.0 + 2.0 * R), -R ** 2 / (1.0 + 2.0 * R)

    def step_method(self):
        return CrankNicholson if (self.StartupStep and self.method == BDF2) else self.method

    def context(self, time, dt, m=None):
        """the fixed data of the Newton system of the step time -> time + dt"""
        m = self.step_method() if m is None else m
        a, at, att, LE = self.motion(time + dt)
        old_rhs = np.zeros(self.ns)
        if m == CrankNicholson:
            _, old_rhs, _ = self.rhs_beameqn(time, self.theta, self.theta_dot, self.pressure_old, self.tau_old)
        return dict(m=m, dt=dt, C=self.scheme(m, dt, self.dt_old), alpha=a, at=at, att=att, LE=LE,
                    th_old=self.theta.copy(), thd_old=self.theta_dot.copy(), th_oo=self.oldold[4].copy(),
                    thd_oo=self.oldold[5].copy(), p=self.pressure_new.copy(), tau=self.tau_new.copy(),
                    tau_s=self.Dn[0] @ self.tau_new, old_rhs=old_rhs)

    # ---- F_nonlinear (SolidSolver.f90:552-667) ----------------------------
    def residual(self, x, c):
        ns, ds, eta, mue, sigma = self.ns, self.ds, self.eta, self.mue, self.sigma
        th = lambda j: x[j + 1]              # theta(-1 .. ns+1)
        T = lambda j: x[ns + 4 + j]          # T(-1 .. ns-1)
        thv = x[1:ns + 1]
        C1, C2, C3, C4 = c["C"]
        dt, a, at, att, LE = c["dt"], c["alpha"], c["at"], c["att"], c["LE"]
        p, tau, tau_s = c["p"], c["tau"], c["tau_s"]
        tdn = (C1 / dt) * (thv - C3 * c["th_old"] - C4 * c["th_oo"]) - C2 * c["thd_old"]
        K1 = mue * (LE[4] * np.cos(a) + LE[5] * np.sin(a) + self.grav * np.sin(a)) - tau[0]
        K2 = p[0] + mue * (LE[5] * np.cos(a) - LE[4] * np.sin(a) + self.grav * np.cos(a))
        F = np.zeros(2 * ns + 4, dtype=x.dtype)
        F[0] = th(0)
        F[1] = (T(1) - T(-1)) / (2 * ds) + (eta / (2 * ds ** 3)) * (th(-1) - 2 * th(0) + th(1)) * (th(1) - th(-1)) - K1
        F[2] = (T(0) * (th(1) - th(-1)) / (2 * ds)
                - (eta / (2 * ds ** 3)) * (-3 * th(-1) + 10 * th(0) - 12 * th(1) + 6 * th(2) - th(3)) - K2)
        n = ns
        F[3] = (th(n - 3) / 12 - 2 * th(n - 2) / 3 + 2 * th(n) / 3 - th(n + 1) / 12) / ds
        F[4] = (-th(n - 3) / 12 + 4 * th(n - 2) / 3 - 2.5 * th(n - 1) + 4 * th(n) / 3 - th(n + 1) / 12) / ds ** 2
        F[5] = T(n - 1)
        q0 = (th(1) - th(-1)) / (2 * ds)
        F[6] = ((T(-1) - 2 * T(0) + T(1)) / ds ** 2 - T(0) * q0 ** 2 + p[0] * q0
                + (eta / ds ** 4) * (th(1) - th(-1)) * (-2.5 * th(0) + 9 * th(1) - 12 * th(2) + 7 * th(3) - 1.5 * th(4))
                + eta * ((th(1) - 2 * th(0) + th(-1)) / ds ** 2) ** 2 + mue * (tdn[0] + at) ** 2 - tau_s[0])
        e = n - 1
        qe = (th(n) - th(n - 2)) / (2 * ds)
        F[7] = (C3 * c["thd_old"][e] + C4 * c["thd_oo"][e] - tdn[e] + (dt / (C1 * mue)) * (
            -(3 * p[e] - 4 * p[e - 1] + p[e - 2]) / (2 * ds)
            - eta * (th(n + 1) - 4 * th(n) + 6 * th(e) - 4 * th(n - 2) + th(n - 3)) / ds ** 4
            + (T(e) + eta * qe ** 2) * (th(n) - 2 * th(e) + th(n - 2)) / ds ** 2
            + (3 * T(e) - 4 * T(e - 1) + T(e - 2)) * (th(n) - th(n - 2)) / (2 * ds ** 2)
            - att * mue ** 2 - sigma * tdn[e] + tau[e] * (th(n) - th(n - 2)) / (2 * ds) + C2 * mue * c["old_rhs"][e]))
        i = np.arange(1, n - 1)
        ti = lambda o: x[i + o + 1]
        Ti = lambda o: x[n + 4 + i + o]
        q = (ti(1) - ti(-1)) / (2 * ds)
        F[8 + i - 1] = (C3 * c["thd_old"][i] + C4 * c["thd_oo"][i] - tdn[i] + (dt / (C1 * mue)) * (
            -(p[i + 1] - p[i - 1]) / (2 * ds)
            - eta * (ti(2) - 4 * ti(1) + 6 * ti(0) - 4 * ti(-1) + ti(-2)) / ds ** 4
            + (Ti(0) + eta * q ** 2) * (ti(1) - 2 * ti(0) + ti(-1)) / ds ** 2
            + (Ti(1) - Ti(-1)) * (ti(1) - ti(-1)) / (2 * ds ** 2)
            - att * mue ** 2 - sigma * tdn[i] + tau[i] * (ti(1) - ti(-1)) / (2 * ds) + C2 * mue * c["old_rhs"][i]))
        F[8 + (n - 2) + i - 1] = ((Ti(-1) - 2 * Ti(0) + Ti(1)) / ds ** 2 - Ti(0) * q ** 2
                                  + p[i] * (ti(1) - ti(-1)) / (2 * ds)
                                  + eta * (ti(1) - ti(-1)) * (ti(2) - 2 * ti(1) + 2 * ti(-1) - ti(-2)) / (2 * ds ** 4)
                                  + eta * ((ti(1) - 2 * ti(0) + ti(-1)) / ds ** 2) ** 2 + mue * (tdn[i] + at) ** 2
                                  - tau_s[i])
        return F

    # ---- Jacobi (SolidSolver.f90:671-863): returns dF_k/dx_l as [k, l] ----
    def jacobian(self, x, c):
        ns, ds, eta, mue, sigma = self.ns, self.ds, self.eta, self.mue, self.sigma
        n = ns
        th = lambda j: x[j + 1]
        T = lambda j: x[ns + 4 + j]
        C1, C2, C3, C4 = c["C"]
        dt, at = c["dt"], c["at"]
        p, tau = c["p"], c["tau"]
        thv = x[1:ns + 1]
        tdn = (C1 / dt) * (thv - C3 * c["th_old"] - C4 * c["th_oo"]) - C2 * c["thd_old"]
        D = C1 / dt
        J = np.zeros((2 * n + 4, 2 * n + 4))

        def s(l, k, v):  # the reference's J(l, k)
            J[k - 1, l - 1] = v

        T0 = n + 4
        s(2, 1, 1.0)
        s(1, 2, (eta / (2 * ds ** 3)) * (-2 * th(-1) + 2 * th(0)))
        s(2, 2, -(eta / ds ** 3) * (th(1) - th(-1)))
        s(3, 2, (eta / (2 * ds ** 3)) * (-2 * th(0) + 2 * th(1)))
        s(T0, 2, -0.5 / ds)
        s(T0 + 2, 2, 0.5 / ds)
        s(1, 3, (-0.5 * eta / ds ** 3) * (-3.0) - T(0) / (2 * ds))
        s(2, 3, (-0.5 * eta / ds ** 3) * 10.0)
        s(3, 3, (-0.5 * eta / ds ** 3) * (-12.0) + T(0) / (2 * ds))
        s(4, 3, (-0.5 * eta / ds ** 3) * 6.0)
        s(5, 3, (-0.5 * eta / ds ** 3) * (-1.0))
        s(T0 + 1, 3, (th(1) - th(-1)) / (2 * ds))
        s(n + 3, 4, (-1.0 / 12.0) / ds)
        s(n + 2, 4, (2.0 / 3.0) / ds)
        s(n, 4, (-2.0 / 3.0) / ds)
        s(n - 1, 4, (1.0 / 12.0) / ds)
        s(n + 3, 5, (-1.0 / 12.0) / ds ** 2)
        s(n + 2, 5, (4.0 / 3.0) / ds ** 2)
        s(n + 1, 5, (-5.0 / 2.0) / ds ** 2)
        s(n, 5, (4.0 / 3.0) / ds ** 2)
        s(n - 1, 5, (-1.0 / 12.0) / ds ** 2)
        s(2 * n + 4, 6, 1.0)
        st = -2.5 * th(0) + 9 * th(1) - 12 * th(2) + 7 * th(3) - 1.5 * th(4)
        cc = th(1) - 2 * th(0) + th(-1)
        dd = th(1) - th(-1)
        s(1, 7, T(0) * dd / (2 * ds ** 2) - p[0] / (2 * ds) - (eta / ds ** 4) * st + 2 * eta * cc / ds ** 4)
        s(2, 7, -2.5 * eta * dd / ds ** 4 - 4 * eta * cc / ds ** 4 + 2 * mue * (tdn[0] + at) * D)
        s(3, 7, -T(0) * dd / (2 * ds ** 2) + p[0] / (2 * ds) + (eta / ds ** 4) * st + 9 * eta * dd / ds ** 4
          + 2 * eta * cc / ds ** 4)
        s(4, 7, -12 * eta * dd / ds ** 4)
        s(5, 7, 7 * eta * dd / ds ** 4)
        s(6, 7, -1.5 * eta * dd / ds ** 4)
        s(n + 4, 7, 1 / ds ** 2)
        s(n + 5, 7, (-2 - dd ** 2 / 4) / ds ** 2)
        s(n + 6, 7, 1 / ds ** 2)
        e = n - 1
        g = dt / (C1 * mue)
        dd = th(n) - th(n - 2)
        cc = th(n) - 2 * th(e) + th(n - 2)
        Tq = T(e) + eta * (dd / (2 * ds)) ** 2
        tg = 3 * T(e) - 4 * T(e - 1) + T(e - 2)
        s(2 * n + 4, 8, (dt / (C1 * mue * ds ** 2)) * (cc + 1.5 * dd))
        s(2 * n + 3, 8, -4 * g * dd / (2 * ds ** 2))
        s(2 * n + 2, 8, g * dd / (2 * ds ** 2))
        s(n + 3, 8, -dt * eta / (C1 * mue * ds ** 4))
        s(n + 2, 8, (dt / (C1 * mue * ds ** 2)) * (4 * eta / ds ** 2 + Tq + tau[e] * ds / 2
                                                   + eta * cc * dd / (2 * ds ** 2) + tg / 2))
        s(n, 8, (dt / (C1 * mue * ds ** 2)) * (4 * eta / ds ** 2 + Tq - tau[e] * ds / 2
                                               - eta * cc * dd / (2 * ds ** 2) - tg / 2))
        s(n + 1, 8, -D + g * (-6 * eta / ds ** 4 - 2 * Tq / ds ** 2 - D * sigma))
        s(n - 1, 8, -dt * eta / (C1 * mue * ds ** 4))
        for i in range(1, n - 1):
            k, l, m = 8 + i, i + 2, i + n + 5
            dd = th(i + 1) - th(i - 1)
            cc = th(i + 1) - 2 * th(i) + th(i - 1)
            Tq = T(i) + eta * (dd / (2 * ds)) ** 2
            s(l, k, -D + g * (-6 * eta / ds ** 4 - (2 / ds ** 2) * Tq - D * sigma))
            s(l - 1, k, g * (4 * eta / ds ** 4 + Tq / ds ** 2 - tau[i] / (2 * ds) - eta * cc * dd / (2 * ds ** 4)
                             - (T(i + 1) - T(i - 1)) / (2 * ds ** 2)))
            s(l + 1, k, g * (4 * eta / ds ** 4 + Tq / ds ** 2 + tau[i] / (2 * ds) + eta * cc * dd / (2 * ds ** 4)
                             + (T(i + 1) - T(i - 1)) / (2 * ds ** 2)))
            s(l - 2, k, -eta * dt / (C1 * mue * ds ** 4))
            s(l + 2, k, -eta * dt / (C1 * mue * ds ** 4))
            s(m, k, g * cc / ds ** 2)
            s(m - 1, k, -g * dd / (2 * ds ** 2))
            s(m + 1, k, g * dd / (2 * ds ** 2))
            k = 8 + n - 2 + i
            st = th(i + 2) - 2 * th(i + 1) + 2 * th(i - 1) - th(i - 2)
            s(l, k, -4 * eta * cc / ds ** 4 + 2 * mue * (tdn[i] + at) * D)
            s(l - 1, k, T(i) * dd / (2 * ds ** 2) - p[i] / (2 * ds) + eta * dd / ds ** 4 - eta * st / (2 * ds ** 4)
              + 2 * eta * cc / ds ** 4)
            s(l + 1, k, -T(i) * dd / (2 * ds ** 2) + p[i] / (2 * ds) - eta * dd / ds ** 4 + eta * st / (2 * ds ** 4)
              + 2 * eta * cc / ds ** 4)
            s(l - 2, k, -eta * dd / (2 * ds ** 4))
            s(l + 2, k, eta * dd / (2 * ds ** 4))
            s(m, k, -2 / ds ** 2 - (dd / (2 * ds)) ** 2)
            s(m - 1, k, 1 / ds ** 2)
            s(m + 1, k, 1 / ds ** 2)
        return J

    def complex_step_jacobian(self, x, c, h=1e-30):
        """exact dF_k/dx_l of F_nonlinear by complex steps"""
        n = x.size
        J = np.zeros((n, n))
        for l in range(n):
            xc = x.astype(complex)
            xc[l] += 1j * h
            J[:, l] = self.residual(xc, c).imag / h
        return J

    # ---- predictor (SolidSolver.f90:167-209) -------------------------------
    def guess(self, time, dt):
        ns, ds, eta, mue = self.ns, self.ds, self.eta, self.mue
        a, at, att, LE = self.motion(time + dt)
        th1, thd1, T = self.rhs_beameqn(time, self.theta, self.theta_dot, self.pressure_old, self.tau_old)
        tg = self.theta + dt * th1
        x = np.zeros(2 * ns + 4)
        x[1:ns + 1] = tg
        K2 = self.pressure_new[0] + mue * (LE[5] * np.cos(a) - LE[4] * np.sin(a) + self.grav * np.cos(a))
        C2 = 2.0 * ds * K2 - T[0] * tg[1] + (eta / ds ** 2) * (10.0 * tg[0] - 12.0 * tg[1] + 6.0 * tg[2] - tg[3])
        x[0] = C2 / ((3.0 * eta / ds ** 2) - T[0])
        A1 = tg[ns - 3] / 12.0 - 2.0 * tg[ns - 2] / 3.0
        A2 = -tg[ns - 3] / 12.0 + 4.0 * tg[ns - 2] / 3.0 - 5.0 * tg[ns - 1] / 2.0
        x[ns + 1] = 3.0 * (A1 - A2) / 2.0
        x[ns + 2] = 12.0 * (2.0 * A1 - A2)
        K1 = mue * (LE[4] * np.cos(a) + LE[5] * np.sin(a) + self.grav * np.sin(a)) - self.tau_new[0]
        th_ss0 = (2.0 * tg[0] - 5.0 * tg[1] + 4.0 * tg[2] - 1.0 * tg[3]) / ds ** 2
        th_s0 = (-1.5 * tg[0] + 2.0 * tg[1] - 0.5 * tg[2]) / ds
        T_s0 = -eta * th_ss0 * th_s0 + K1
        x[ns + 3] = T[1] - 2.0 * ds * T_s0
        x[ns + 4:] = T
        return x

    # ---- IBES_solver (SolidSolver.f90:99-414) -----------------------------
    def ibes(self, time, dt):
        m = self.step_method()
        self.StartupStep = False
        old = np.stack([self.x, self.y, self.vx, self.vy, self.theta, self.theta_dot])
        x = self.guess(time, dt)
        c = self.context(time, dt, m)
        it = 0
        while True:
            F = -self.residual(x, c)
            J = self.jacobian(x, c)
            dx = np.linalg.solve(J, F)
            it += 1
            x = x + dx
            err = np.sqrt(np.sum(dx ** 2))
            with np.errstate(invalid="ignore", divide="ignore"):  # x = 0: no relative criterion
                err_rel = abs(err / np.sqrt(np.sum(x ** 2)))
            if it > 49:
                raise RuntimeError("IBES did not converge")
            if (err < 1e-7 or err_rel < 1e-7) and it > 2:
                break
        self.iter = it
        C1, C2, C3, C4 = self.scheme(m, dt, self.dt_old)
        self.theta = x[1:self.ns + 1].copy()
        self.theta_dot = (C1 / dt) * (self.theta - C3 * old[4] - C4 * self.oldold[4]) - C2 * old[5]
        self.integrate(time + dt)
        self.ax = (self.vx - old[2]) / dt
        self.ay = (self.vy - old[3]) / dt
        self.brake = np.max(np.abs(self.theta_dot - old[5])) > 100.0
        self.oldold = old
        self.dt_old = dt

    # ---- SolidEnergies (SolidSolver.f90:76-92) ----------------------------
    def energies(self):
        th_s = self.Dn[0] @ self.theta
        ds, mue = self.ds, self.mue
        return dict(E_kinetic=mue * 0.5 * ds * np.sum(self.vx ** 2 + self.vy ** 2),
                    E_pot=mue * self.grav * ds * np.sum(self.y - self.y0),
                    E_elastic=self.eta * 0.5 * ds * np.sum(th_s ** 2),
                    Inertial_Force=(mue * ds * np.sum(self.ax), mue * ds * np.sum(self.ay)))

    # ---- one step of the loop (time_step.f90:209-228) ---------------------
    def step(self, time, dt):
        self.iter, self.brake = 0, False
        self.gravity_impulse(time)
        if time > self.T_release:
            self.ibes(time, dt)
        else:
            self.integrate(time + dt)
        self.pressure_old = self.pressure_new.copy()
        self.tau_old = self.tau_new.copy()
