"""Seeded synthetic systems for the benchmark and parity workloads.

SPC water (all workloads): O -0.82 e / H +0.41 e, masses 15.9994 / 1.008,
O-H 0.1 nm, H-O-H 109.47 deg, LJ on O-O only (C6 = 2.6173456e-3,
C12 = 2.634129e-6 kJ mol^-1 nm^6 / nm^12), every water its own charge group
and solvent molecule (crates/gromos-core/src/topology.rs:1204-1263), reaction
field eps_rf = 61, kappa = 0, four_pi_eps_i = 138.9354.

Waters sit on a cubic lattice with uniform +-jitter per coordinate, uniformly
random rigid orientations and Maxwell-Boltzmann velocities projected onto
rigid-body motion (so SETTLE starts from consistent velocities).
"""
import math

import numpy as np

from .system import FOUR_PI_EPS_I, K_BOLTZMANN, Perturbation, System, csr_from_lists

SPC_C6 = 2.6173456e-3
SPC_C12 = 2.634129e-6
M_O, M_H = 15.9994, 1.008
Q_O, Q_H = -0.82, 0.41
D_OH = 0.1
THETA_HOH = math.radians(109.47)
D_HH = 2.0 * D_OH * math.sin(0.5 * THETA_HOH)
AVOGADRO = 6.02214076e23


def box_for_density(n_waters, density_g_cm3=0.997):
    m = n_waters * (M_O + 2 * M_H) / AVOGADRO        # g
    vol_nm3 = m / density_g_cm3 * 1e21
    return vol_nm3 ** (1.0 / 3.0)


def random_rotations(rng, n):
    """Uniform random rotation matrices (Shoemake quaternions)."""
    u1, u2, u3 = rng.random(n), rng.random(n), rng.random(n)
    q = np.stack([np.sqrt(1 - u1) * np.sin(2 * np.pi * u2), np.sqrt(1 - u1) * np.cos(2 * np.pi * u2),
                  np.sqrt(u1) * np.sin(2 * np.pi * u3), np.sqrt(u1) * np.cos(2 * np.pi * u3)], axis=1)
    x, y, z, w = q.T
    R = np.empty((n, 3, 3))
    R[:, 0, 0] = 1 - 2 * (y * y + z * z); R[:, 0, 1] = 2 * (x * y - z * w); R[:, 0, 2] = 2 * (x * z + y * w)
    R[:, 1, 0] = 2 * (x * y + z * w); R[:, 1, 1] = 1 - 2 * (x * x + z * z); R[:, 1, 2] = 2 * (y * z - x * w)
    R[:, 2, 0] = 2 * (x * z - y * w); R[:, 2, 1] = 2 * (y * z + x * w); R[:, 2, 2] = 1 - 2 * (x * x + y * y)
    return R


def water_template():
    s, c = math.sin(0.5 * THETA_HOH), math.cos(0.5 * THETA_HOH)
    return np.array([[0.0, 0.0, 0.0], [D_OH * s, D_OH * c, 0.0], [-D_OH * s, D_OH * c, 0.0]])


def rigid_mb_velocities(rng, pos, masses, temperature):
    """Per-atom Maxwell-Boltzmann velocities projected on rigid-body motion per water."""
    n = len(pos) // 3
    m = np.asarray(masses[:3])
    sig = np.sqrt(K_BOLTZMANN * temperature / np.tile(m, n))[:, None]
    v = rng.normal(size=(3 * n, 3)) * sig
    x = pos.reshape(n, 3, 3)
    v = v.reshape(n, 3, 3)
    M = m.sum()
    com = (x * m[None, :, None]).sum(1) / M
    vcom = (v * m[None, :, None]).sum(1) / M
    r = x - com[:, None, :]
    L = (np.cross(r, v) * m[None, :, None]).sum(1)
    I = np.zeros((n, 3, 3))
    for a in range(3):
        rr = r[:, a, :]
        I += m[a] * (np.einsum("ni,ni->n", rr, rr)[:, None, None] * np.eye(3)[None] - rr[:, :, None] * rr[:, None, :])
    omega = np.linalg.solve(I, L[..., None])[..., 0]
    vr = vcom[:, None, :] + np.cross(omega[:, None, :], r)
    return vr.reshape(-1, 3)


def spc_water_box(n_waters, box=None, density=0.997, jitter=0.02, seed=0, temperature=300.0,
                  rcutp=1.4, rcutl=1.4, nsnb=5, epsrf=61.0, dt=0.002, name="spc"):
    rng = np.random.default_rng(seed)
    L = box_for_density(n_waters, density) if box is None else float(box)
    side = int(math.ceil(round(n_waters ** (1.0 / 3.0), 9)))
    while side ** 3 < n_waters:
        side += 1
    sites = np.stack(np.meshgrid(np.arange(side), np.arange(side), np.arange(side), indexing="ij"),
                     axis=-1).reshape(-1, 3)[:, ::-1]
    if len(sites) > n_waters:
        sites = sites[np.sort(rng.choice(len(sites), n_waters, replace=False))]
    spacing = L / side
    o = (sites + 0.5) * spacing + rng.uniform(-jitter, jitter, size=(n_waters, 3))
    R = random_rotations(rng, n_waters)
    tmpl = water_template()
    pos = (o[:, None, :] + np.einsum("nij,aj->nai", R, tmpl)).reshape(-1, 3)
    mass = np.tile([M_O, M_H, M_H], n_waters)
    vel = rigid_mb_velocities(rng, pos, mass, temperature) if temperature > 0 else np.zeros_like(pos)
    n = 3 * n_waters
    charge = np.tile([Q_O, Q_H, Q_H], n_waters)
    iac = np.tile(np.array([0, 1, 1], np.int32), n_waters)
    lj = np.zeros((2, 2, 4))
    lj[0, 0] = (SPC_C6, SPC_C12, SPC_C6, SPC_C12)
    excl = []
    for w in range(n_waters):
        b = 3 * w
        excl += [[b + 1, b + 2], [b, b + 2], [b, b + 1]]
    excl_off, excl_idx = csr_from_lists(excl, n)
    o14_off = np.zeros(n + 1, np.int32)
    return System(n_atoms=n, n_solute_atoms=0, cg_start=np.arange(0, n + 1, 3, dtype=np.int32), n_solute_cg=0,
                  charge=charge, iac=iac, mass=mass, n_types=2, lj=lj.reshape(-1), excl_off=excl_off,
                  excl=excl_idx, o14_off=o14_off, o14=np.zeros(0, np.int32), atoms_per_solvent=3, pos=pos,
                  vel=vel, box=np.array([L, L, L]), boundary=1, rcutp=rcutp, rcutl=rcutl, nsnb=nsnb,
                  epsrf=epsrf, kappa=0.0, fpepsi=FOUR_PI_EPS_I, dt=dt, dist_oh=D_OH, dist_hh=D_HH,
                  name=name).validate()


def water_216(seed=1):
    """216 SPC on a 6^3 lattice, L = 1.8621 nm, twin range 0.8 / 1.4 (cutoff > L/2)."""
    return spc_water_box(216, box=1.8621, seed=seed, rcutp=0.8, rcutl=1.4, nsnb=5, name="water_216")


def water_32768(seed=2):
    """32,768 SPC at 0.997 g/cm^3 (L ~ 9.94 nm), single range 1.4 nm."""
    return spc_water_box(32768, seed=seed, rcutp=1.4, rcutl=1.4, nsnb=5, name="water_32768")


def water_replica(seed, temperature, n_waters=4000, rc=1.2):
    """12,000-atom SPC replica box (L ~ 4.93 nm), cutoff 1.2 nm."""
    return spc_water_box(n_waters, seed=seed, temperature=temperature, rcutp=rc, rcutl=rc, nsnb=5,
                         name=f"replica_{temperature:.1f}K")
