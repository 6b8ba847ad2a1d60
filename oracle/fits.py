"""ORACLE — test infrastructure only (see oracle/__init__.py).

numpy restatement of the kinetic-theory species properties and their fits, transport.rs:14-137:
collision integrals by quadratic interpolation in ln T* between three rows of the Monchick-Mason tables
(transport.rs:82-96), viscosity, binary diffusion and the Kee/Warnatz conductivity (transport.rs:97-116),
and the weighted least-squares fits in ln(T/T0) on 50 nodes in [300, 3000] K with row weights 1/y²
(transport.rs:24-36, 125-136), solved here by numpy's SVD least squares (the reference uses nalgebra's SVD).
"""
import json
import math
import os

import numpy as np

KB = 1.380649e-23
NA = 6.02214076e23
R = KB * NA
C_LIGHT = 299792458.0
MU0 = 2.0 * 7.2973525693e-3 * 6.62607015e-34 / (1.602176634e-19 ** 2 * C_LIGHT)
EPS0 = 1.0 / (C_LIGHT * C_LIGHT * MU0)

_TABLES = json.load(open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "..", "combustion_amd", "data",
                                      "collision_integrals.json")))
HEADER_T = np.array(_TABLES["header_T_star"])
HEADER_DELTA = np.array(_TABLES["header_delta_star"])
OMEGA22 = np.array(_TABLES["omega22"]["rows"])
ASTAR = np.array(_TABLES["a_star"]["rows"])


def _delta_fits(table):
    V = np.stack([HEADER_DELTA ** k for k in range(7)], axis=1)
    return np.array([np.linalg.lstsq(V, row, rcond=None)[0][:6] for row in table])


OMEGA22_FITS = _delta_fits(OMEGA22)
ASTAR_FITS = _delta_fits(ASTAR)


class Physics:
    def __init__(self, mech):
        self.m = mech

    def reduced_mass(self, a, b):
        W = self.m["W"]
        return (W[a] / NA) * (W[b] / NA) / (W[a] / NA + W[b] / NA)

    def chi(self, a, b):
        mu, m = self.m["dipole"], self.m
        if (mu[a] > 0) == (mu[b] > 0):
            return 1.0
        p, n = (a, b) if mu[a] != 0 else (b, a)
        return 1.0 + 0.25 * m["polarizability"][n] / m["diameter"][n] ** 3 * (
            mu[p] / math.sqrt(4 * math.pi * EPS0 * m["well_depth_J"][p] * m["diameter"][p] ** 3)) ** 2 * math.sqrt(
            m["well_depth_J"][p] / m["well_depth_J"][n])

    def T_star(self, a, b, T):
        eps = self.m["well_depth_J"]
        return T * KB / (math.sqrt(eps[a] * eps[b]) * self.chi(a, b) ** 2)

    def delta_star(self, a, b):
        m = self.m
        return m["dipole"][a] * m["dipole"][b] / (8 * math.pi * EPS0 * math.sqrt(m["well_depth_J"][a] * m["well_depth_J"][b]) *
                                                  ((m["diameter"][a] + m["diameter"][b]) / 2) ** 3)

    def collision(self, table, fits, first, a, b, T):
        x0 = math.log(self.T_star(a, b, T))
        h = np.log(HEADER_T)
        above = np.nonzero(x0 < h[1:])[0]
        assert len(above), "T* above the table"
        start = min(int(above[0]), first + len(table) - 3)
        assert start - first >= 0, "T* below the table"
        x = h[start:start + 3]
        assert x[0] <= x0 <= x[2]
        d = self.delta_star(a, b)
        rows = slice(start - first, start - first + 3)
        y = table[rows, 0] if d == 0 else np.array([sum(p[k] * d ** k for k in range(6)) for p in fits[rows]])
        # Lagrange form of the parabola through the three nodes (same polynomial as transport.rs:14-17)
        v = sum(y[i] * np.prod([(x0 - x[j]) / (x[i] - x[j]) for j in range(3) if j != i]) for i in range(3))
        assert v > 0
        return v

    def omega22(self, a, b, T):
        return self.collision(OMEGA22, OMEGA22_FITS, 1, a, b, T)

    def omega11(self, a, b, T):
        return self.omega22(a, b, T) / self.collision(ASTAR, ASTAR_FITS, 0, a, b, T)

    def viscosity(self, k, T):
        m = self.m
        return 5 / 16 * math.sqrt(math.pi * m["W"][k] / NA * KB * T) / (self.omega22(k, k, T) * math.pi * m["diameter"][k] ** 2)

    def diffusivity(self, a, b, T):
        m = self.m
        d = (m["diameter"][a] + m["diameter"][b]) / 2 * self.chi(a, b) ** (-1 / 6)
        return 3 / 16 * math.sqrt(2 * math.pi / self.reduced_mass(a, b)) * (KB * T) ** 1.5 / (math.pi * d * d * self.omega11(a, b, T))

    def cp_R(self, k, T):
        m = self.m
        a = m["nasa"][k][0 if T <= m["split"][k] else 1]
        return a[0] + a[1] * T + a[2] * T ** 2 + a[3] * T ** 3 + a[4] * T ** 4

    def conductivity(self, k, T):
        m = self.m
        f_int = m["W"][k] / NA / (KB * T) * self.diffusivity(k, k, T) / self.viscosity(k, T)
        fz = lambda t: 1 + math.pi ** 1.5 / math.sqrt(t) * (0.5 + 1 / t) + (0.25 * math.pi ** 2 + 2) / t
        dof = m["internal_dof"][k]
        c1 = 2 / math.pi * (2.5 - f_int) / (m["rotational_relaxation"][k] * fz(self.T_star(k, k, 298.0)) / fz(self.T_star(k, k, T)) +
                                            2 / math.pi * (5 / 3 * dof + f_int))
        f_trans = 2.5 * (1 - c1 * dof / 1.5)
        f_rot = f_int * (1 + c1)
        cv_int = self.cp_R(k, T) - 2.5 - dof
        return self.viscosity(k, T) / (m["W"][k] / NA) * KB * (1.5 * f_trans + f_rot * dof + f_int * cv_int)


def fits(mech, D=4, T0=1000.0):
    """(conductivity[K][D], viscosity[K][D], diffusivity[K][K][D]) like transport.rs:125-136."""
    ph = Physics(mech)
    K = mech["K"]
    T = 300.0 + np.arange(50) / 49.0 * 2700.0
    x = np.log(T / T0)
    V = np.stack([x ** k for k in range(D)], axis=1)

    def fit(y):
        w = 1.0 / y ** 2
        return np.linalg.lstsq(V * w[:, None], y * w, rcond=None)[0]

    c = np.array([fit(np.array([ph.conductivity(k, t) / math.sqrt(t) for t in T])) for k in range(K)])
    v = np.array([fit(np.array([math.sqrt(ph.viscosity(k, t) / math.sqrt(t)) for t in T])) for k in range(K)])
    d = np.zeros((K, K, D))
    for k in range(K):
        for j in range(k):
            d[k, j] = d[j, k] = fit(np.array([ph.diffusivity(k, j, t) / (t * math.sqrt(t)) for t in T]))
    return c, v, d
