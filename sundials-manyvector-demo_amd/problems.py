"""Synthetic inputs for the benchmark configurations (numpy, host side).

All generators return host numpy arrays in the ManyVector layout of include/mvd_rhs.h:
fluid fields as flat (nz, ny, nx) C-order arrays (x fastest), species AoS (cell, species).
Random draws use numpy's PCG64 with the given seed, in (k, j, i) C order.
"""
import numpy as np

MH = 1.67e-24            # hydrogen mass used by the network (raja_primordial_network.cpp:914)
KBOLTZ = 1.3806488e-16   # Boltzmann constant of the initial conditions (primordial_blast.cpp:173)
M_AMU = 1.66053904e-24   # primordial_ode_main.cpp:256
# species order H2_1, H2_2, H_1, H_2, H_m0, He_1, He_2, He_3 (raja_primordial_network.cpp:920-929)
WEIGHTS = np.array([2.0, 2.0, 1.00794, 1.00794, 1.00794, 4.002602, 4.002602, 4.002602])


def cell_centres(n, lo=0.0, hi=1.0):
    h = (hi - lo) / n
    return lo + (np.arange(n) + 0.5) * h


def eos_inv(rho, mx, my, mz, p, gamma):
    """EulerData::eos_inv (euler3D.hpp:1393-1398)."""
    return p / (gamma - 1.0) + (mx * mx + my * my + mz * mz) * 0.5 / rho


def blast3d(n, seed=1234, gamma=1.4):
    """3D fluid blast on [0,1]^3, nchem = 0:
    rho = 1 + 0.01*U(-1,1); v_x, v_y, v_z ~ N(0, sigma=0.01); p = 1 + 10*exp(-|x-0.5|^2/0.01)."""
    rng = np.random.default_rng(seed)
    shape = (n, n, n)
    rho = 1.0 + 0.01 * rng.uniform(-1.0, 1.0, size=shape)
    v = rng.normal(0.0, 0.01, size=(3,) + shape)
    x = cell_centres(n)
    zz, yy, xx = np.meshgrid(x, x, x, indexing="ij")
    r2 = (xx - 0.5) ** 2 + (yy - 0.5) ** 2 + (zz - 0.5) ** 2
    p = 1.0 + 10.0 * np.exp(-r2 / 0.01)
    mx, my, mz = rho * v[0], rho * v[1], rho * v[2]
    et = eos_inv(rho, mx, my, mz, p, gamma)
    return [np.ascontiguousarray(a.reshape(-1)) for a in (rho, mx, my, mz, et)]


def sod(nx, ny=3, nz=3, gamma=1.4):
    """Sod shock tube along x (sod.cpp:135-147): x<0.5 -> (1,0,1), else (0.125,0,0.1)."""
    x = cell_centres(nx)
    rho1 = np.where(x < 0.5, 1.0, 0.125)
    p1 = np.where(x < 0.5, 1.0, 0.1)
    rho = np.broadcast_to(rho1, (nz, ny, nx)).copy()
    p = np.broadcast_to(p1, (nz, ny, nx)).copy()
    z = np.zeros_like(rho)
    et = eos_inv(rho, z, z, z, p, gamma)
    return [np.ascontiguousarray(a.reshape(-1)) for a in (rho, z, z.copy(), z.copy(), et)]


def smooth_state(nx, ny, nz, nchem=0, seed=7, gamma=1.4, amp=0.2, noise=0.0):
    """Smooth positive random-phase state used by parity tests (plus optional noise)."""
    rng = np.random.default_rng(seed)
    shape = (nz, ny, nx)
    zz, yy, xx = np.meshgrid(cell_centres(nz), cell_centres(ny), cell_centres(nx), indexing="ij")
    ph = rng.uniform(0, 2 * np.pi, size=8)

    def wave(a, b, c, k):
        return np.sin(2 * np.pi * xx + ph[a]) * np.cos(2 * np.pi * yy + ph[b]) * \
            np.sin(2 * np.pi * zz + ph[c] + k)

    rho = 1.0 + amp * wave(0, 1, 2, 0.3) + noise * rng.uniform(-1, 1, shape)
    vx = 0.3 * wave(1, 2, 3, 0.1) + noise * rng.normal(size=shape)
    vy = -0.2 * wave(2, 3, 4, 0.7) + noise * rng.normal(size=shape)
    vz = 0.25 * wave(3, 4, 5, 1.1) + noise * rng.normal(size=shape)
    p = 1.0 + amp * wave(4, 5, 6, 0.9) + noise * rng.uniform(-1, 1, shape)
    mx, my, mz = rho * vx, rho * vy, rho * vz
    et = eos_inv(rho, mx, my, mz, p, gamma)
    out = [np.ascontiguousarray(a.reshape(-1)) for a in (rho, mx, my, mz, et)]
    if nchem > 0:
        c = 0.5 + 0.4 * rng.uniform(size=(nz * ny * nx, nchem))
        out.append(np.ascontiguousarray(c.reshape(-1)))
    else:
        out.append(np.zeros(1))
    return out


def chem_cells(n, seed=4321, alpha=(1.0,) * 8, ntot=(1e-2, 1e2), temp=(10.0, 1e4),
               gamma=5.0 / 3.0, with_rho=False):
    """Independent primordial-chemistry cells (AoS, 10 per cell), physical units:
    total number density n_tot log-uniform in [1e-2,1e2] cm^-3 with rho = n_tot*mH;
    mass fractions of the 8 heavy species ~ Dirichlet(alpha); H- capped at half the positive
    charge so the charge-balanced electron density de = (H+ + He+ + 2He++ - H- + H2+)
    is positive; T log-uniform in [10,1e4] K, ge = kB*T*n/(rho*(gamma-1)).
    Species are number densities; de is stored divided by m_amu/mH as in
    primordial_ode_main.cpp:357-376. with_rho: also return the mass density drawn per cell."""
    rng = np.random.default_rng(seed)
    lo, hi = np.log(ntot[0]), np.log(ntot[1])
    n_tot = np.exp(rng.uniform(lo, hi, size=n))
    rho = n_tot * MH
    X = rng.dirichlet(np.asarray(alpha, dtype=np.float64), size=n)
    nd = X * rho[:, None] / (WEIGHTS[None, :] * MH)
    pos = nd[:, 3] + nd[:, 6] + 2.0 * nd[:, 7] + nd[:, 1]
    cap = 0.5 * pos
    excess = np.maximum(nd[:, 4] - cap, 0.0)
    nd[:, 4] -= excess
    nd[:, 2] += excess * WEIGHTS[4] / WEIGHTS[2]
    de = (nd[:, 3] + nd[:, 6] + 2.0 * nd[:, 7] - nd[:, 4] + nd[:, 1]) * MH
    T = np.exp(rng.uniform(np.log(temp[0]), np.log(temp[1]), size=n))
    ndens = nd.sum(axis=1)
    ge = KBOLTZ * T * ndens / (rho * (gamma - 1.0))
    y = np.empty((n, 10))
    y[:, :8] = nd
    y[:, 8] = de / M_AMU
    y[:, 9] = ge
    if with_rho:
        return np.ascontiguousarray(y.reshape(-1)), T, rho
    return np.ascontiguousarray(y.reshape(-1)), T


def primordial_blast_mri(nx, ny=None, nz=None, bump=100.0, width=0.05, t_in=1e4, t_amb=10.0,
                         gamma=5.0 / 3.0, units=(3.0e70, 3.0857e30, 1.0e11), core_trace=True):
    """Multirate primordial blast (configuration 5): gas at rest on [0,1]^3 with a Gaussian
    overdensity of `bump` times the background density0 = 1e2*mH and width `width` centred
    in the box, temperature t_amb outside rising to t_in at the centre, both with the
    reference's profile shape exp(-2 r^2 / width^2) (primordial_blast.cpp:218-231); species
    composition as primordial_blast.cpp:238-271 (trace 1e-3 fractions outside r^2/w^2 < 2,
    tiny = 1e-40 / small = 1e-12 inside, Hfrac = 0.76; core_trace=False uses the outside
    fractions everywhere), ge = kB T n / (rho (gamma - 1)).
    Units (mass, length, time) of tests/primordial_blast/input_primordial_blast_mr.txt:34-37.
    Returns (fluid: 5 flat code-unit arrays, chem: AoS physical, 10 per cell, as
    initial_conditions stores them before prepare_Dengo_structures)."""
    ny = nx if ny is None else ny
    nz = nx if nz is None else nz
    mu, lu, tu = units
    density_units = mu / lu / lu / lu
    energy_units = mu / lu / tu / tu
    tiny, small, hfrac = 1e-40, 1e-12, 0.76
    hi_w, he_w = 1.00794 * MH, 4.002602 * MH
    density0 = 1e2 * MH
    zz, yy, xx = np.meshgrid(cell_centres(nz), cell_centres(ny), cell_centres(nx), indexing="ij")
    rsq = (xx - 0.5) ** 2 + (yy - 0.5) ** 2 + (zz - 0.5) ** 2
    prof = np.exp(-2.0 * rsq / width / width)
    density = density0 * (1.0 + (bump - 1.0) * prof)
    T = t_amb + (t_in - t_amb) * prof
    inside = (rsq / width / width < 2.0) & core_trace
    H2I = np.where(inside, tiny, 1e-3) * density
    H2II = np.where(inside, tiny, 1e-3) * density
    HII = np.where(inside, small, 1e-3) * density
    HM = np.where(inside, tiny, 1e-3) * density
    HeII = np.where(inside, small, 1e-3) * density
    HeIII = np.where(inside, small, 1e-3) * density
    HeI = (1.0 - hfrac) * density - HeII - HeIII
    HI = density - (H2I + H2II + HII + HM + HeI + HeII + HeIII)
    n = np.stack([H2I / (2 * hi_w), H2II / (2 * hi_w), HI / hi_w, HII / hi_w, HM / hi_w,
                  HeI / he_w, HeII / he_w, HeIII / he_w], axis=-1)
    ndens = n.sum(axis=-1)
    de = (n[..., 3] + n[..., 6] + 2.0 * n[..., 7] - n[..., 4] + n[..., 1]) * MH
    ge = KBOLTZ * T * ndens / (density * (gamma - 1.0))
    chem = np.empty(density.shape + (10,))
    chem[..., :8] = n
    chem[..., 8] = de / M_AMU
    chem[..., 9] = ge
    rho = density / density_units
    zero = np.zeros_like(rho)
    et = ge / energy_units                   # initial_conditions: (ge + 0.5 rho v^2)/EnergyUnits, v = 0
    fluid = [np.ascontiguousarray(a.reshape(-1)) for a in (rho, zero, zero, zero, et)]
    return fluid, np.ascontiguousarray(chem.reshape(-1))
