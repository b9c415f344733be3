"""Lattice definitions (reference: src/dynamics.py).

``D2Q9(tau, delta_t)`` and ``D2Q13(tau, delta_t)`` keep the reference's class
constants (``DIM``, ``NUM_QUIVERS``, ``KSI``, ``W``, ``C``) and constructor.
The constants are FP32 values, the ones the reference's JAX path (x64
disabled) computes with: W is the f32 rounding of 4/9, 1/9, 1/36 (D2Q13:
3/8, 1/12, 1/16, 1/96) and C = 1/sqrt(3) (D2Q13: 1/sqrt(2)) in f32. The host
``calc_eq``/``calc_macro`` here are conveniences for initial conditions and
inspection; the time step itself runs in the HIP library.
"""
from __future__ import annotations

import numpy as np


class Dynamics:
    """Base class (src/dynamics.py:12-47)."""
    DIM: int
    NUM_QUIVERS: int
    KSI: np.ndarray
    W: np.ndarray
    C: np.float32

    def __init__(self):
        pass

    def ones_pdf(self):
        return np.ones(self.NUM_QUIVERS, np.float32)

    def density(self, pdf):
        return np.sum(np.asarray(pdf, np.float32), axis=-1, keepdims=True)

    def velocity(self, pdf, rho):
        pdf = np.asarray(pdf, np.float32)
        return (pdf @ self.KSI.astype(np.float32)) / rho

    def calc_macro(self, pdf):
        rho = self.density(pdf)
        return rho, self.velocity(pdf, rho)


class D2Q9(Dynamics):
    """D2Q9 lattice (src/dynamics.py:50-74)."""
    DIM = 2
    NUM_QUIVERS = 9
    KSI = np.array([[0, 0], [1, 0], [0, 1], [-1, 0], [0, -1],
                    [1, 1], [-1, 1], [-1, -1], [1, -1]], dtype=np.int32)
    W = np.array([4 / 9, 1 / 9, 1 / 9, 1 / 9, 1 / 9,
                  1 / 36, 1 / 36, 1 / 36, 1 / 36]).astype(np.float32)
    C = np.float32(1) / np.sqrt(np.float32(3))
    LATTICE_ID = 0

    def __init__(self, tau, delta_t):
        super().__init__()
        self.tau = tau
        self.delta_t = delta_t

    def calc_eq(self, rho, vel):
        """W*rho*(1 + (ksi.u)/C^2 + (ksi.u)^2/(2C^4) - (u.u)/(2C^2)) for
        rho[...,1], vel[...,2] (vectorised; FP32)."""
        rho = np.asarray(rho, np.float32)
        vel = np.asarray(vel, np.float32)
        C2 = self.C * self.C
        C4 = C2 * C2
        cu = vel @ self.KSI.T.astype(np.float32)
        uu = np.sum(vel * vel, axis=-1, keepdims=True)
        return (self.W * rho) * (((np.float32(1) + cu / C2) + (cu * cu) / (np.float32(2) * C4))
                                 - uu / (np.float32(2) * C2))


class D2Q13(Dynamics):
    """D2Q13 lattice (src/dynamics.py:77-102): the D2Q9 velocities plus the
    four double-length axis velocities; the equilibrium adds cubic terms."""
    DIM = 2
    NUM_QUIVERS = 13
    KSI = np.array([[0, 0], [1, 0], [0, 1], [-1, 0], [0, -1],
                    [1, 1], [-1, 1], [-1, -1], [1, -1],
                    [2, 0], [0, 2], [-2, 0], [0, -2]], dtype=np.int32)
    W = np.array([3 / 8, 1 / 12, 1 / 12, 1 / 12, 1 / 12, 1 / 16, 1 / 16, 1 / 16, 1 / 16,
                  1 / 96, 1 / 96, 1 / 96, 1 / 96]).astype(np.float32)
    C = np.float32(1) / np.sqrt(np.float32(2))
    LATTICE_ID = 1

    def __init__(self, tau, delta_t):
        super().__init__()
        self.tau = tau
        self.delta_t = delta_t

    def calc_eq(self, rho, vel):
        """W*rho*(1 + cu/C^2 + cu^2/(2C^4) - uu/(2C^2) + cu^3/(2C^6)
        - 3 cu uu/(2C^4)) with cu = ksi.u, uu = u.u, summed left to right
        (vectorised; FP32)."""
        rho = np.asarray(rho, np.float32)
        vel = np.asarray(vel, np.float32)
        C2 = self.C * self.C
        C4 = C2 * C2
        C6 = C2 * C4
        two = np.float32(2)
        cu = vel @ self.KSI.T.astype(np.float32)
        uu = np.sum(vel * vel, axis=-1, keepdims=True)
        br = ((np.float32(1) + cu / C2) + (cu * cu) / (two * C4)) - uu / (two * C2)
        br = (br + (cu * (cu * cu)) / (two * C6)) - ((np.float32(3) * cu) * uu) / (two * C4)
        return (self.W * rho) * br
