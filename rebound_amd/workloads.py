"""Seeded synthetic workloads built with the library's own generators
(reb_random_*, tools.c:64-105), drawing in the same order as a C program that
makes the same calls against the reference library."""
import ctypes
import math

from . import _abi, lib, Sim


def shearing_sheet(n, root_size, n_root, seed, omega=1.0, collisions="tree"):
    """Shearing sheet of hard spheres (the ring recipe of
    examples/shearing_sheet/problem.c): n_root x n_root root boxes of
    root_size, SHEAR boundary with 2 ghost boxes in x and y, TREE gravity at
    opening_angle2 0.25 with G = 1.546/400 and softening 0.1, SEI with
    dt = 1e-3 * 2 pi; radii power-law dN/dR ~ R^-3 in [1,4], m = 400 * 4/3 pi R^3,
    positions uniform in x, y and normal in z, velocity dispersion 3.2 (= 2 <R>
    Omega) on top of the shear -1.5 Omega x; hard-sphere collisions with a
    constant coefficient of restitution 0.5.  Returns a configured Sim."""
    L = lib()
    sim = Sim()
    s = sim.s
    s.rand_seed = seed
    s.OMEGA = omega
    s.G = 1.546 / 400.0
    s.softening = 0.1
    s.dt = 1e-3 * 2.0 * math.pi
    s.root_size = root_size
    s.N_root_x = n_root
    s.N_root_y = n_root
    s.N_root_z = 1
    s.N_ghost_x = 2
    s.N_ghost_y = 2
    s.N_ghost_z = 0
    s.boundary = _abi.BOUNDARY["shear"]
    s.gravity = _abi.GRAVITY["tree"]
    s.opening_angle2 = 0.25
    sim.set_integrator("sei")
    bx = root_size * n_root
    by = root_size * n_root
    r = sim.ptr
    uni, nor, pw = L.reb_random_uniform, L.reb_random_normal, L.reb_random_powerlaw
    add = L.reb_simulation_add
    p = _abi.Particle()
    four_thirds_pi_rho = 400.0 * 4.0 / 3.0 * math.pi
    for _ in range(n):
        p.x = uni(r, -bx / 2.0, bx / 2.0)
        p.y = uni(r, -by / 2.0, by / 2.0)
        p.z = nor(r, 1.0)
        R = pw(r, 1.0, 4.0, -3.0)
        p.vx = nor(r, 3.2 * 3.2)
        p.vy = -1.5 * p.x * omega + nor(r, 3.2 * 3.2)
        p.vz = nor(r, 3.2 * 3.2) * 0.5
        p.r = R
        p.m = four_thirds_pi_rho * R * R * R
        add(r, p)
    L.reb_amd_set_coefficient_of_restitution(r, 0.5)
    if collisions:
        s.collision = _abi.COLLISION[collisions]
        hs = ctypes.cast(L.reb_collision_resolve_hardsphere, ctypes.c_void_p).value
        s.collision_resolve = ctypes.cast(hs, _abi.RESOLVE)
    return sim
