"""Time evolution of a variational state with the TDVP equations.

The reference computes S and F (source/network_functions/TDVP.cu.template:181-334)
and leaves the integration to user scripts.  This module is that loop, on the
device path: every right-hand side is one TDVP evaluation (sampling, E_loc, O_k,
S/F on the GPU) and one shifted solve (S + shift) x = F in fp64.

    real time:       d theta / dt = -i (S + shift)^-1 F
    imaginary time:  d theta / dt =   -(S + shift)^-1 F   (stochastic reconfiguration)

``RK4`` integrates either with the classical fourth-order Runge-Kutta scheme;
``Euler`` is the first-order step used for ground-state searches.
"""
import numpy as np

from .api import TDVP, ExpectationValue, ExactSummationSpins


class _Integrator:
    def __init__(self, psi, op, ensemble, dt, shift=1e-3, method="cg", real_time=True, rtol=1e-6, max_iter=500):
        self.psi = psi
        self.op = op
        self.ensemble = ensemble
        self.dt = float(dt)
        self.shift = float(shift)
        self.method = method
        self.real_time = bool(real_time)
        self.rtol = float(rtol)
        self.max_iter = int(max_iter)
        self.tdvp = TDVP(psi.num_params)
        self.t = 0.0
        self.evaluations = 0
        self.energies = []

    def rhs(self, params):
        """d params / dt at params (the wave function is left at params).  With
        exact sums the state is renormalised first, so that the weights |psi|^2
        entering S and F sum to one."""
        self.psi.params = params
        if isinstance(self.ensemble, ExactSummationSpins):
            self.psi.normalize(self.ensemble)
        if self.method == "cholesky":
            self.tdvp.eval(self.op, self.psi, self.ensemble)
        else:
            self.tdvp.eval_F(self.op, self.psi, self.ensemble)
        x = self.tdvp.solve(self.shift, self.method, rtol=self.rtol, max_iter=self.max_iter)
        self.evaluations += 1
        self.energies.append(self.tdvp.E_local)
        return -1j * x if self.real_time else -x


class Euler(_Integrator):
    def step(self):
        p = self.psi.params
        p = p + self.dt * self.rhs(p)
        self.psi.params = p
        self.t += self.dt
        return p


class RK4(_Integrator):
    """classical Runge-Kutta: four TDVP evaluations per step"""

    def step(self):
        p0 = self.psi.params
        h = self.dt
        k1 = self.rhs(p0)
        k2 = self.rhs(p0 + 0.5 * h * k1)
        k3 = self.rhs(p0 + 0.5 * h * k2)
        k4 = self.rhs(p0 + h * k3)
        p = p0 + (h / 6.0) * (k1 + 2 * k2 + 2 * k3 + k4)
        self.psi.params = p
        self.t += h
        return p

    def run(self, steps, observables=(), callback=None):
        """advance ``steps`` steps; returns times and the expectation values of
        ``observables`` (Operators) after every step"""
        ev = ExpectationValue()
        times, values = [], []
        for _ in range(int(steps)):
            self.step()
            times.append(self.t)
            values.append([ev(o, self.psi, self.ensemble) for o in observables])
            if callback is not None:
                callback(self)
        return np.array(times), np.array(values, dtype=complex)
