"""Initial states for the benchmark and test workloads (packed FP64 cells).

Cells are packed in the reference's per-cell order
``partial_mass[nmat], mom[dim], E, alpha[nmat-1]`` (src/app/dim_mpi_main.cpp:255-273),
grid x-fastest (src/dim/solver/advance/geometry.hpp:25-55).  Every builder
returns a :class:`Case` holding the grid, spacing, materials, boundary names
and the packed state; the state is an input, so it only needs to be
deterministic, not bit-identical to the reference's own initialiser.

The shock-bubble profile follows the reference's diffuse shock-bubble
initialiser (src/setup/dim_initial_conditions.hpp:331-404): tanh-smoothed
bubble weight, volume-weighted densities per material, weight-averaged
velocity and pressure.
"""
from __future__ import annotations

import dataclasses
import math

import numpy as np

WATER = ("stiffened_gas", 4.4, 6.0e8)
AIR = ("ideal_gas", 1.4, 0.0)
HELIUM = ("ideal_gas", 1.648, 0.0)
# noble-Abel air with co-volume b = 1e-3 m^3/kg
# (configs/DIM/bubble_collapse/gorsse_tc9_water_air_bubble_2d.txt:33-34)
NA_AIR = ("noble_abel", 1.4, 0.0, 1.0e-3)


def material_energy_density(mat, alpha, rho, p):
    """alpha_k rho_k e_k(rho_k, p) of one material: ideal, stiffened
    (eos.hpp:482-489, 587-595) or noble-Abel (eos.hpp:692-702)."""
    kind, g, pinf = mat[:3]
    if kind in ("noble_abel", "noble_abel_gas", "real_gas"):
        b = mat[3] if len(mat) > 3 else 0.0
        return alpha * p * (1.0 - b * rho) / (g - 1.0)
    add = g * pinf if kind == "stiffened_gas" else 0.0
    return alpha * (p + add) / (g - 1.0)


@dataclasses.dataclass
class Case:
    name: str
    dim: int
    N: tuple
    dx: tuple
    materials: list
    bc_lo: tuple
    bc_hi: tuple
    U: np.ndarray  # (ncells, nv) float64, or None when only `ic` is given
    cfl: float = 0.8
    lambda_model: str = "kapila"
    alpha_source_floor: float = 0.0
    ic: dict = None  # keyword arguments of dimhip.make_ic describing the same initial state

    @property
    def nmat(self):
        return len(self.materials)

    @property
    def nv(self):
        return 2 * self.nmat + self.dim

    @property
    def ncells(self):
        return int(np.prod(self.N))


def prim_to_packed(materials, alpha, rho, vel, p):
    """Closed-form primitive -> packed conservative (alpha: (n,nmat), rho: (n,nmat),
    vel: (n,dim), p: (n,))."""
    n, nmat = alpha.shape
    dim = vel.shape[1]
    m = alpha * rho
    rt = m.sum(axis=1)
    e = np.zeros(n)
    for k, mat in enumerate(materials):
        e += material_energy_density(mat, alpha[:, k], rho[:, k], p)
    U = np.zeros((n, 2 * nmat + dim))
    U[:, :nmat] = m
    U[:, nmat:nmat + dim] = rt[:, None] * vel
    U[:, nmat + dim] = e + 0.5 * rt * np.sum(vel * vel, axis=1)
    U[:, nmat + dim + 1:] = alpha[:, :nmat - 1]
    return U


def _centres(N, lo, dx):
    axes = [lo[d] + (np.arange(N[d]) + 0.5) * dx[d] for d in range(len(N))]
    grids = np.meshgrid(*axes[::-1], indexing="ij")  # slowest axis first
    return [g.reshape(-1) for g in grids[::-1]]  # x-fastest flattening


def water_air_tube_1d(n=1000, alpha_floor=1e-6, air=AIR):
    """Two-material shock tube on [0,1], interface at 0.5: water (stiffened gas,
    gamma=4.4, p_inf=6e8, rho=1000, p=1e9) left, air (ideal, gamma=1.4, rho=50,
    p=1e5) right, both at rest, volume fractions floored at alpha_floor."""
    dx = 1.0 / n
    (x,) = _centres((n,), (0.0,), (dx,))
    left = x < 0.5
    a0 = np.where(left, 1.0 - alpha_floor, alpha_floor)
    alpha = np.stack([a0, 1.0 - a0], axis=1)
    rho = np.stack([np.full(n, 1000.0), np.full(n, 50.0)], axis=1)
    p = np.where(left, 1.0e9, 1.0e5)
    vel = np.zeros((n, 1))
    mats = [WATER, air]
    U = prim_to_packed(mats, alpha, rho, vel, p)
    return Case("water_air_tube_1d", 1, (n,), (dx,), mats, ("transmissive",), ("transmissive",), U)


def water_na_air_tube_1d(n=1000, alpha_floor=1e-6):
    """The water/air tube of water_air_tube_1d with noble-Abel air (b = 1e-3)."""
    c = water_air_tube_1d(n, alpha_floor, air=NA_AIR)
    c.name = "water_na_air_tube_1d"
    return c


def rankine_hugoniot(gamma, mach, rho1, p1):
    """Post-shock (rho2, p2, |u2|) behind a shock of Mach `mach` running into gas at rest."""
    m2 = mach * mach
    p2 = p1 * (1.0 + 2.0 * gamma / (gamma + 1.0) * (m2 - 1.0))
    rho2 = rho1 * (gamma + 1.0) * m2 / ((gamma - 1.0) * m2 + 2.0)
    c1 = math.sqrt(gamma * p1 / rho1)
    w = mach * c1
    u2 = w * (1.0 - rho1 / rho2)
    return rho2, p2, u2


def shock_bubble(N, lo, hi, mats, air_slot, bubble_slot, pre, post, bubble, centre, radius,
                 shock_axis, shock_position, thickness, sharpness=2.0, bc_lo=None, bc_hi=None,
                 name="shock_bubble", bubble_shape=None, build_state=True):
    """Diffuse shock-bubble state (dim_initial_conditions.hpp:331-404).

    pre/post/bubble: (rho, vel tuple, p).  Post-shock gas sits at x >= shock_position.
    bubble_shape: optional callable(coords) -> signed distance (negative inside),
    replacing the sphere (no device description then).  With build_state=False
    the packed state is not built on the host: the case carries only the
    description for the device initialiser.
    """
    dim = len(N)
    dx = tuple((hi[d] - lo[d]) / N[d] for d in range(dim))
    bc_lo = bc_lo or tuple(["transmissive"] * dim)
    bc_hi = bc_hi or tuple(["transmissive"] * dim)
    ic = None
    if bubble_shape is None:
        ic = dict(kind="shock_bubble", domain_min=tuple(lo), domain_max=tuple(hi), thickness=thickness,
                  sharpness=sharpness, shock_axis=shock_axis, shock_position=shock_position,
                  left=(pre[0], tuple(pre[1]), pre[2], air_slot),
                  right=(post[0], tuple(post[1]), post[2], air_slot),
                  bubble=(bubble[0], tuple(bubble[1]), bubble[2], bubble_slot),
                  centre=tuple(centre), radius=radius)
    if not build_state:
        if ic is None:
            raise ValueError("a custom bubble shape needs the host-built state")
        return Case(name, dim, tuple(N), dx, list(mats), tuple(bc_lo), tuple(bc_hi), None, ic=ic)
    X = _centres(N, lo, dx)
    if bubble_shape is None:
        r2 = sum((X[d] - centre[d]) ** 2 for d in range(dim))
        sd = np.sqrt(r2) - radius
    else:
        sd = bubble_shape(X)
    post_mask = X[shock_axis] >= shock_position
    rho_air = np.where(post_mask, post[0], pre[0])
    vel_air = np.stack([np.where(post_mask, post[1][d], pre[1][d]) for d in range(dim)], axis=1)
    p_air = np.where(post_mask, post[2], pre[2])
    if thickness <= 0.0:
        bw = (sd <= 0.0).astype(np.float64)
    else:
        bw = 0.5 * (1.0 - np.tanh(max(sharpness, 1e-12) * sd / thickness))
    aw = 1.0 - bw
    n = sd.shape[0]
    nmat = len(mats)
    alpha = np.zeros((n, nmat))
    rho = np.ones((n, nmat))
    air_on = aw > 1e-14
    bub_on = bw > 1e-14
    alpha[:, air_slot] += np.where(air_on, aw, 0.0)
    alpha[:, bubble_slot] += np.where(bub_on, bw, 0.0)
    rho[:, air_slot] = np.where(air_on, rho_air, 1.0)
    rho[:, bubble_slot] = np.where(bub_on, bubble[0], 1.0)
    vel = aw[:, None] * vel_air + bw[:, None] * np.asarray(bubble[1])[None, :]
    p = aw * p_air + bw * bubble[2]
    alpha = np.maximum(alpha, 0.0)
    alpha /= alpha.sum(axis=1, keepdims=True)
    U = prim_to_packed(mats, alpha, rho, vel, p)
    return Case(name, dim, tuple(N), dx, list(mats), tuple(bc_lo), tuple(bc_hi), U, ic=ic)


def helium_shock_bubble_2d(nx=2048, ny=1024):
    """Helium bubble hit by a Mach 1.22 planar air shock on [0,0.325]x[0,0.089] m.

    Air gamma=1.4, rho=1.29, p=101325 at rest; post-shock state from
    Rankine-Hugoniot at x >= 0.225, moving -x; helium gamma=1.648, rho=0.214,
    r=0.025 at (0.175, 0.0445); transmissive in x, reflective in y; interface
    thickness 3 dx, tanh sharpness 2.
    """
    rho1, p1 = 1.29, 101325.0
    rho2, p2, u2 = rankine_hugoniot(1.4, 1.22, rho1, p1)
    lo, hi = (0.0, 0.0), (0.325, 0.089)
    thickness = 3.0 * (hi[0] - lo[0]) / nx
    return shock_bubble((nx, ny), lo, hi, [AIR, HELIUM], 0, 1,
                        pre=(rho1, (0.0, 0.0), p1), post=(rho2, (-u2, 0.0), p2),
                        bubble=(0.214, (0.0, 0.0), p1), centre=(0.175, 0.0445), radius=0.025,
                        shock_axis=0, shock_position=0.225, thickness=thickness,
                        bc_lo=("transmissive", "reflective"), bc_hi=("transmissive", "reflective"),
                        name=f"helium_shock_bubble_2d_{nx}x{ny}")


def stiffened_hugoniot(gamma, pinf, rho1, p1, p2):
    """Post-shock density and particle speed for a stiffened-gas shock from (rho1,p1) at rest
    to pressure p2 (Rankine-Hugoniot with the stiffened-gas energy)."""
    a = (p2 + pinf) / (p1 + pinf)
    g = gamma
    rho2 = rho1 * ((g + 1.0) * a + (g - 1.0)) / ((g - 1.0) * a + (g + 1.0))
    u2 = math.sqrt((p2 - p1) * (1.0 / rho1 - 1.0 / rho2))
    return rho2, u2


def bubble_collapse_3d(n=512, build_state=True):
    """Air bubble (ideal, gamma=1.4, rho=1, r=0.15, centre of the unit cube) in
    water (stiffened gas gamma=4.4, p_inf=6e8, rho=1000, p=1e5), struck by a
    planar 1e9 Pa liquid shock entering from x=0; transmissive on every face."""
    rho1, p1, p2 = 1000.0, 1.0e5, 1.0e9
    rho2, u2 = stiffened_hugoniot(4.4, 6.0e8, rho1, p1, p2)
    h = 1.0 / n
    # the reference initialiser puts its "right" state at x >= shock_position;
    # post-shock liquid on the left is the "left" state with positive velocity
    case = shock_bubble((n, n, n), (0.0, 0.0, 0.0), (1.0, 1.0, 1.0), [WATER, AIR], 0, 1,
                        pre=(rho2, (u2, 0.0, 0.0), p2), post=(rho1, (0.0, 0.0, 0.0), p1),
                        bubble=(1.0, (0.0, 0.0, 0.0), p1), centre=(0.5, 0.5, 0.5), radius=0.15,
                        shock_axis=0, shock_position=0.2, thickness=2.0 * h,
                        name=f"bubble_collapse_3d_{n}", build_state=build_state)
    return case


def gorsse_tc9_2d(nx=480, ny=400, build_state=True):
    """Gorsse et al. TC9 water/air bubble collapse
    (configs/DIM/bubble_collapse/gorsse_tc9_water_air_bubble_2d.txt):
    [-0.2, 1.0] x [0, 1]; stiffened-gas water (gamma 4.4, p_inf 6e8) at
    rho 1000, p 1e5 left of x = 0.75 and rho 1230, u = -432.69, p 1e9 right of
    it; a noble-Abel air bubble (gamma 1.4, b = 1e-3) of radius 0.2 at
    (0.5, 0.5), rho 1.2, p 1e5; interface thickness 0.015; CFL 0.45;
    zero-gradient x and reflective y boundaries."""
    case = shock_bubble((nx, ny), (-0.2, 0.0), (1.0, 1.0), [WATER, NA_AIR], 0, 1,
                        pre=(1000.0, (0.0, 0.0), 1.0e5), post=(1230.0, (-432.69, 0.0), 1.0e9),
                        bubble=(1.2, (0.0, 0.0), 1.0e5), centre=(0.5, 0.5), radius=0.2,
                        shock_axis=0, shock_position=0.75, thickness=0.015,
                        bc_lo=("transmissive", "reflective"), bc_hi=("transmissive", "reflective"),
                        name=f"gorsse_tc9_{nx}x{ny}", build_state=build_state)
    case.cfl = 0.45
    return case


def batch_instances(count=64, seed=20261203, nx_range=(256, 1024), ny_range=(256, 512)):
    """Seeded shock-interface instances of uneven size (instance i uses
    numpy's PCG64 generator seeded with seed + i).

    Per instance: Nx ~ U{nx_range}, Ny ~ U{ny_range}, height H = 1 and
    length Nx/Ny (square cells); ambient gas a and interface gas b ideal with
    gamma ~ U[1.3, 1.7]; density ratio rho_b/rho_a ~ U[0.1, 10]; shock Mach
    ~ U[1.1, 3.0] from Rankine-Hugoniot in gas a, entering from the right;
    interface radius ~ U[0.1, 0.3] H; half the instances a circular bubble,
    half a sinusoidally perturbed planar interface x = x0 + A sin(2 pi k y/H);
    transmissive in x, reflective in y; interface thickness 2 cells.
    """
    cases = []
    for i in range(count):
        rng = np.random.default_rng(seed + i)
        nx = int(rng.integers(nx_range[0], nx_range[1] + 1))
        ny = int(rng.integers(ny_range[0], ny_range[1] + 1))
        ga, gb = rng.uniform(1.3, 1.7, 2)
        ratio = rng.uniform(0.1, 10.0)
        mach = rng.uniform(1.1, 3.0)
        r = rng.uniform(0.1, 0.3)
        k = 0 if i % 2 == 0 else int(rng.integers(1, 4))
        cases.append(shock_interface(nx, ny, float(ga), float(gb), ratio, mach, r, k, name=f"batch_{i}_{nx}x{ny}"))
    return cases


def shock_interface(nx, ny, gamma_a, gamma_b, ratio, mach, radius, wavenumber=0, name=None):
    """One shock-interface instance of the batched sweep on [0, nx/ny] x [0, 1]
    (square cells): ambient ideal gas a (rho 1, p 1e5, at rest) with a
    region of ideal gas b (density ratio*1, same pressure) whose interface is
    a circle of the given radius centred at (0.4 L, 0.5) (wavenumber 0), or
    the sinusoid x = 0.4 L + (radius/2) sin(2 pi k y) with k = wavenumber;
    a Mach `mach` shock in gas a (Rankine-Hugoniot) at
    x = min(0.4 L + radius + 0.1 L, 0.9 L) moving -x; transmissive in x,
    reflective in y; interface thickness 2 cells, tanh sharpness 2."""
    H = 1.0
    Lx = nx / ny * H
    r = radius * H
    rho1, p1 = 1.0, 1.0e5
    rho2, p2, u2 = rankine_hugoniot(gamma_a, mach, rho1, p1)
    mats = [("ideal_gas", float(gamma_a), 0.0), ("ideal_gas", float(gamma_b), 0.0)]
    x0 = 0.4 * Lx
    shock_x = min(x0 + r + 0.1 * Lx, 0.9 * Lx)
    if wavenumber == 0:
        shape = None
    else:
        amp = 0.5 * r
        shape = (lambda X, x0=x0, amp=amp, k=wavenumber: X[0] - (x0 + amp * np.sin(2.0 * np.pi * k * X[1] / H)))
    return shock_bubble((nx, ny), (0.0, 0.0), (Lx, H), mats, 0, 1,
                        pre=(rho1, (0.0, 0.0), p1), post=(rho2, (-u2, 0.0), p2),
                        bubble=(rho1 * ratio, (0.0, 0.0), p1), centre=(x0, 0.5 * H), radius=r,
                        shock_axis=0, shock_position=shock_x, thickness=2.0 * H / ny,
                        bc_lo=("transmissive", "reflective"), bc_hi=("transmissive", "reflective"),
                        name=name or f"shock_interface_{nx}x{ny}", bubble_shape=shape)
