"""Input generators: the fiber/pulse builders of the reference's drivers and the
named workloads (2-mode Gaussian, 6-mode and 10-mode GRIN) used for parity and
benchmarking.

Harness conventions follow the reference drivers: the Blow-Wood Raman response
sampled on t_j = j*dt (simulations/main.cpp:67-75, matlab/GRIN2.m:59-63), the
field grid t_j = (j - NT/2)*dt with dt = T_w/NT (simulations/main.cpp:114-127,
matlab/GRIN2.m:44-45), and the column-major MATLAB coupling tensor ingestion
Q(p,k,l,m) -> COO (matlab/mm_propagation.cpp:197-226).
"""
from __future__ import annotations

import math
from typing import List, Sequence

import numpy as np

from .problem import NonZeroValue, SimulationParameters, Sparse3DArray  # noqa: F401 (re-exported)

SR = 6.4341e12  # m^-2, coupling scale of the reference examples (simulations/main.cpp:26)
GAMMA = 1.4e-13  # W^-1 m
F_RAMAN = 0.18
TAU1 = 12.2e-3  # ps
TAU2 = 32e-3  # ps


def blow_wood_response(nt: int, time_window: float, tau1: float = TAU1, tau2: float = TAU2) -> np.ndarray:
    """h_R(t_j) = ((tau1^2 + tau2^2) / (tau1 tau2^2)) exp(-t/tau2) sin(t/tau1), t_j = j*dt
    (simulations/main.cpp:67-75: a = (1/tau1^2 + 1/tau2^2)*tau1, b = exp(-t/tau2)*sin(t/tau1))."""
    delta = time_window / nt
    a = (1.0 / (tau1 * tau1) + 1.0 / (tau2 * tau2)) * tau1
    out = np.empty(nt, dtype=np.float64)
    for i in range(nt):
        t = i * delta
        out[i] = a * (math.exp(-t / tau2) * math.sin(t / tau1))
    return out


def time_grid(nt: int, time_window: float) -> np.ndarray:
    """t_j = (j - NT/2) * dt, dt = T_w / NT (matlab/GRIN2.m:44-45)."""
    dt = time_window / nt
    return (np.arange(nt, dtype=np.float64) - nt // 2) * dt


def gaussian(nt: int, time_window: float, t0: float, peak_power: float, phase: float = 0.0) -> np.ndarray:
    """sqrt(P) exp(-t^2 / (2 T0^2)) exp(i phase) on the field grid."""
    t = time_grid(nt, time_window)
    return (math.sqrt(peak_power) * np.exp(-t * t / (2.0 * t0 * t0)) * np.exp(1j * phase)).astype(np.complex128)


def soliton(t_final: float, nt: int, t0: float, c: float, q: float) -> np.ndarray:
    """build_soliton of simulations/main.cpp:114-127 (t stepped by repeated addition)."""
    phi = np.empty(nt, dtype=np.complex128)
    h = t_final / nt
    t = -t_final / 2
    for j in range(nt):
        phi[j] = 0.5 * math.sqrt(q / 2) * (1.0 / math.cosh(q / 4 * (t - t0))) * complex(
            math.cos(c * ((t - t0) / 2)), math.sin(c * ((t - t0) / 2)))
        t += h
    return phi


def coupling_from_matlab(q4: np.ndarray) -> List[Sparse3DArray]:
    """Column-major M x M x M x M array Q(p,k,l,m) -> per-p COO lists in the order the MEX
    gateway inserts them (matlab/mm_propagation.cpp:197-226): p fastest, then k, l, m;
    zeros skipped."""
    q4 = np.asarray(q4, dtype=np.float64)
    M = q4.shape[0]
    if q4.shape != (M, M, M, M):
        raise ValueError("coupling array must be M x M x M x M")
    out = [Sparse3DArray() for _ in range(M)]
    for m in range(M):
        for l in range(M):
            for k in range(M):
                for p in range(M):
                    v = q4[p, k, l, m]
                    if v != 0.0:
                        out[p].addNonZeroValue(NonZeroValue((k, l, m), float(v)))
    return out


def grin_coupling(M: int, seed: int, mirror: bool = False, rng: np.random.Generator | None = None) -> List[Sparse3DArray]:
    """GRIN-like coupling rule: Q^(p) has a nonzero at every (k, l, m) with k + l - m - p even
    and k <= l, value SR*u with u ~ U(-0.5, 1) drawn from numpy's default_rng(seed) (PCG64)
    in insertion order p, k, l >= k, m. With mirror=True the (l, k, m) partner of every k < l
    entry is inserted right after it with the same value."""
    g = rng if rng is not None else np.random.default_rng(seed)
    out = []
    for p in range(M):
        q = Sparse3DArray()
        for k in range(M):
            for l in range(k, M):
                for m in range(M):
                    if (k + l - m - p) % 2 == 0:
                        v = SR * float(g.uniform(-0.5, 1.0))
                        q.addNonZeroValue(NonZeroValue((k, l, m), v))
                        if mirror and k < l:
                            q.addNonZeroValue(NonZeroValue((l, k, m), v))
        out.append(q)
    return out


def grin_coupling_symmetric(M: int, rng: np.random.Generator) -> List[Sparse3DArray]:
    """Fully symmetric GRIN-like coupling: Q^(p)_{k,l,m} = SR*u(sorted(p,k,l,m)) at every
    (k, l, m) with k + l - m - p even (k and l both ways, i.e. stored symmetric in k, l).
    One u ~ U(-0.5, 1) per index multiset, drawn from `rng` in first-encounter order of the
    insertion loop p, k, l, m. A tensor symmetric under every index permutation is what
    overlap integrals of real mode profiles give; with the physical pair decode it makes
    the Kerr and Raman terms conserve the total energy sum_p ||A_p||^2."""
    vals = {}
    out = []
    for p in range(M):
        q = Sparse3DArray()
        for k in range(M):
            for l in range(M):
                for m in range(M):
                    if (k + l - m - p) % 2 == 0:
                        key = tuple(sorted((p, k, l, m)))
                        if key not in vals:
                            vals[key] = SR * float(rng.uniform(-0.5, 1.0))
                        q.addNonZeroValue(NonZeroValue((k, l, m), vals[key]))
        out.append(q)
    return out


def two_mode_gaussian(nt: int = 4096, n_steps: int = 5000, fiber_length: float = 0.5,
                      time_window: float = 10.0, method_order: int = 4) -> SimulationParameters:
    """2-mode Gaussian workload: M=2, NT=4096, T_w=10 ps, L=0.5 m, 5000 RK4 steps,
    beta0={0,-2e3} 1/m, beta1={0,1e-3} ps/m, beta2=0.02 ps^2/m, the 3+3 coupling list below,
    gamma=1.4e-13, fR=0.18, Blow-Wood h_R, Gaussians T0=0.1 ps of 100 W and 10 W."""
    sp = SimulationParameters()
    sp.n_modes = 2
    sp.fiber_length = fiber_length
    sp.dispersion_coefficients = [[0.0, 0.0, 0.02], [-2.0e3, 1.0e-3, 0.02]]
    sp.coupling_coefficients = [
        Sparse3DArray([(0, 0, 0, SR), (0, 0, 1, SR / 3), (0, 1, 1, SR / 2)]),
        Sparse3DArray([(1, 1, 1, -SR / 3), (0, 0, 1, SR / 3), (0, 1, 1, SR / 2)]),
    ]
    sp.nt = nt
    sp.time_window = time_window
    sp.pulse_width = 0.1
    sp.initial_fields = np.stack([gaussian(nt, time_window, 0.1, 100.0), gaussian(nt, time_window, 0.1, 10.0)])
    sp.n_steps = n_steps
    sp.method_order = method_order
    sp.nonlinearity_const = GAMMA
    sp.raman_proportion = F_RAMAN
    sp.raman_response = blow_wood_response(nt, time_window)
    return sp


def grin_fiber(M: int = 6, nt: int = 16384, n_steps: int = 20000, fiber_length: float = 1.0,
               time_window: float = 20.0, seed: int = 0, method_order: int = 4,
               mirror: bool = False, t0: float = 0.1, power_range=(10.0, 1000.0),
               symmetric: bool = False) -> SimulationParameters:
    """GRIN multimode workload: beta0_p = -p*1.5e3 1/m, beta1_p = p*5e-4 ps/m, beta2 = 0.02 ps^2/m,
    beta3 = 3e-5 ps^3/m; coupling by `grin_coupling`; Blow-Wood h_R; Gaussians of width t0 with
    per-mode peak power log-uniform in power_range and phase U(0, 2 pi). One default_rng(seed)
    stream draws, in order, the coupling values, the M powers and the M phases. With
    symmetric=True the coupling is `grin_coupling_symmetric` (energy-conserving with the
    physical pair decode); otherwise `grin_coupling` (k <= l only, or mirrored)."""
    rng = np.random.default_rng(seed)
    sp = SimulationParameters()
    sp.n_modes = M
    sp.fiber_length = fiber_length
    sp.dispersion_coefficients = [[-p * 1.5e3, p * 5e-4, 0.02, 3e-5] for p in range(M)]
    if symmetric:
        sp.coupling_coefficients = grin_coupling_symmetric(M, rng)
    else:
        sp.coupling_coefficients = grin_coupling(M, seed, mirror=mirror, rng=rng)
    lo, hi = power_range
    powers = np.exp(rng.uniform(math.log(lo), math.log(hi), size=M))
    phases = rng.uniform(0.0, 2.0 * math.pi, size=M)
    sp.nt = nt
    sp.time_window = time_window
    sp.pulse_width = t0
    sp.initial_fields = np.stack([gaussian(nt, time_window, t0, float(powers[p]), float(phases[p])) for p in range(M)])
    sp.n_steps = n_steps
    sp.method_order = method_order
    sp.nonlinearity_const = GAMMA
    sp.raman_proportion = F_RAMAN
    sp.raman_response = blow_wood_response(nt, time_window)
    return sp


BATCH_MODES = (3, 6, 10)
BATCH_NT = (4096, 8192, 16384)


def batched_fiber(instance: int, n_steps: int = 10000) -> SimulationParameters:
    """One fiber of the batched workload: a default_rng(instance) stream draws, in order, M from
    {3, 6, 10}, NT from {4096, 8192, 16384}, the length U(0.2, 1) m and T0 U(0.05, 0.3) ps; the
    GRIN fiber itself (coupling values, powers log-uniform in [10 W, 10 kW], phases) is
    `grin_fiber(..., seed=instance, symmetric=True)`."""
    rng = np.random.default_rng(instance)
    M = int(rng.choice(BATCH_MODES))
    nt = int(rng.choice(BATCH_NT))
    length = float(rng.uniform(0.2, 1.0))
    t0 = float(rng.uniform(0.05, 0.3))
    return grin_fiber(M=M, nt=nt, n_steps=n_steps, fiber_length=length, seed=instance, t0=t0,
                      power_range=(10.0, 1.0e4), symmetric=True)


def grin2_soliton_pair(nt: int = 2 ** 15, n_steps: int = 1024) -> SimulationParameters:
    """The reference's 2-mode GRIN soliton-pair example (simulations/main.cpp:10-96)."""
    t_final = 160.0
    t0 = -t_final / 10
    c = 1.0
    sp = SimulationParameters()
    sp.n_modes = 2
    sp.dispersion_coefficients = [[0.0, 0.0, 18.938e-3], [0.0, 0.0, 18.866e-3]]
    sp.coupling_coefficients = [
        Sparse3DArray([(0, 0, 0, SR), (0, 0, 1, SR / 3), (0, 1, 1, SR / 2)]),
        Sparse3DArray([(1, 0, 0, SR / 5), (1, 1, 0, SR / 3), (1, 1, 1, 0.8 * SR)]),
    ]
    sp.fiber_length = 16.0
    sp.n_steps = n_steps
    sp.method_order = 4
    lam = 1030e-9
    f0 = 2.99792458e-4 / lam
    w0 = 2 * math.pi * f0
    n2 = 2.3 * 1e-20
    sp.nonlinearity_const = (n2 * w0) / 2.99792458e-4
    sp.raman_proportion = 0.18
    sp.raman_response = blow_wood_response(nt, t_final)
    beta_2 = sp.dispersion_coefficients[0][2]
    sp.nt = nt
    sp.pulse_width = math.sqrt(beta_2) / (SR * sp.nonlinearity_const)
    sp.time_window = t_final
    q = SR * sp.nonlinearity_const
    s = math.sqrt(beta_2 / 2)
    sp.initial_fields = np.stack([
        soliton(t_final / s, nt, t0 / s, c * math.sqrt(2 / beta_2), q),
        soliton(t_final / s, nt, -t0 / s, -c * math.sqrt(2 / beta_2), 0.8 * q),
    ])
    return sp


def rel_l2_per_mode(a: np.ndarray, ref: np.ndarray) -> np.ndarray:
    """||a_p - ref_p||_2 / ||ref_p||_2 for each mode p."""
    a = np.asarray(a)
    ref = np.asarray(ref)
    num = np.linalg.norm((a - ref).reshape(ref.shape[0], -1), axis=1)
    den = np.linalg.norm(ref.reshape(ref.shape[0], -1), axis=1)
    return num / np.where(den == 0, 1.0, den)


def energy_per_mode(fields: np.ndarray, time_window: float) -> np.ndarray:
    """E_p = sum_j |A_p,j|^2 * dt (W ps), dt = T_w / NT."""
    f = np.asarray(fields)
    return (np.abs(f) ** 2).sum(axis=1) * (time_window / f.shape[1])
