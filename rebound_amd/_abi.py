"""ctypes mirror of include/rebound.h (the drop-in ABI).

The layouts are those of REBOUND 5.0.0 on LP64 (rebound.h:86-413); the loader
checks sizeof(Simulation) against reb_simulation_struct_size() of the library.
This is the binding a Python caller (or REBOUND's own ctypes package,
rebound/simulation.py:1365-1470) uses against librebound_amd.so.
"""
import ctypes
from ctypes import (CFUNCTYPE, POINTER, Structure, c_char_p, c_double, c_int, c_int64, c_size_t, c_uint,
                    c_uint32, c_uint64, c_void_p)

SIZE_MAX = ctypes.c_size_t(-1).value


class Particle(Structure):
    _fields_ = [("x", c_double), ("y", c_double), ("z", c_double),
                ("vx", c_double), ("vy", c_double), ("vz", c_double),
                ("ax", c_double), ("ay", c_double), ("az", c_double),
                ("m", c_double), ("r", c_double),
                ("name", c_char_p), ("ap", c_void_p), ("sim", c_void_p)]


class Vec3d(Structure):
    _fields_ = [("x", c_double), ("y", c_double), ("z", c_double)]


class Vec6d(Structure):
    _fields_ = [("x", c_double), ("y", c_double), ("z", c_double),
                ("vx", c_double), ("vy", c_double), ("vz", c_double)]


class Collision(Structure):
    _fields_ = [("p1", c_size_t), ("p2", c_size_t), ("gb", Vec6d), ("ri", c_size_t)]


class Integrator(Structure):
    _fields_ = [("documentation", c_char_p), ("step", c_void_p), ("synchronize", c_void_p),
                ("create", c_void_p), ("free", c_void_p), ("did_add_particle", c_void_p),
                ("will_remove_particle", c_void_p), ("field_descriptor_list", c_void_p)]


class IntegratorConfiguration(Structure):
    _fields_ = [("name", c_char_p), ("callbacks", Integrator), ("state", c_void_p)]


class LeapfrogState(Structure):
    _fields_ = [("order", c_uint)]


class WHFastState(Structure):
    _fields_ = [("corrector", c_uint), ("corrector2", c_uint), ("kernel", c_int), ("coordinates", c_int),
                ("safe_mode", c_uint), ("keep_unsynchronized", c_uint), ("N_allocated", c_size_t),
                ("p_jh", POINTER(Particle)), ("N_allocated_var", c_size_t), ("p_jh_var", POINTER(Particle)),
                ("recalculate_coordinates_but_not_synchronized_warning", c_uint)]


class Simulation(Structure):
    pass


HEARTBEAT = CFUNCTYPE(None, POINTER(Simulation))
STEP = CFUNCTYPE(None, POINTER(Simulation), c_void_p)
COR = CFUNCTYPE(c_double, POINTER(Simulation), c_double)
RESOLVE = CFUNCTYPE(c_int, POINTER(Simulation), Collision)

Simulation._fields_ = [
    ("t", c_double), ("G", c_double), ("softening", c_double), ("OMEGA", c_double), ("OMEGAZ", c_double),
    ("dt", c_double), ("dt_last_done", c_double), ("steps_done", c_uint64),
    ("is_synchronized", c_uint), ("did_modify_particles", c_uint),
    ("N", c_size_t), ("N_allocated", c_size_t), ("particles", POINTER(Particle)),
    ("N_map", c_size_t), ("map", POINTER(c_size_t)),
    ("N_var", c_size_t), ("particles_var", POINTER(Particle)),
    ("N_var_config", c_size_t), ("var_config", c_void_p),
    ("N_active", c_size_t), ("testparticle_type", c_int), ("testparticle_hidewarnings", c_int),
    ("name_list", c_void_p), ("N_name_list", c_size_t), ("name_hash_table", c_void_p),
    ("gravity_cs", c_void_p), ("N_allocated_gravity_cs", c_size_t), ("tree_root", c_void_p),
    ("opening_angle2", c_double), ("status", c_int), ("exact_finish_time", c_int),
    ("force_is_velocity_dependent", c_int), ("gravity_ignore_terms", c_int),
    ("output_timing_last", c_double),
    ("save_messages", c_int), ("messages", POINTER(c_char_p)),
    ("messages_var_rescale_warning", c_int), ("messages_timestep_warning", c_int),
    ("exit_max_distance", c_double), ("exit_min_distance", c_double), ("usleep", c_double),
    ("display_settings", c_void_p), ("display_data", c_void_p), ("server_data", c_void_p),
    ("track_energy_offset", c_int), ("energy_offset", c_double),
    ("walltime", c_double), ("walltime_last_step", c_double), ("walltime_last_steps", c_double),
    ("walltime_last_steps_sum", c_double), ("walltime_last_steps_N", c_int),
    ("python_unit_l", c_uint32), ("python_unit_m", c_uint32), ("python_unit_t", c_uint32),
    ("root_size", c_double), ("N_root_x", c_size_t), ("N_root_y", c_size_t), ("N_root_z", c_size_t),
    ("N_ghost_x", c_int), ("N_ghost_y", c_int), ("N_ghost_z", c_int),
    ("collisions", POINTER(Collision)), ("N_allocated_collisions", c_size_t), ("N_collisions", c_size_t),
    ("N_targets", c_size_t), ("minimum_collision_velocity", c_double), ("collisions_plog", c_double),
    ("collisions_log_n", c_int64),
    ("calculate_megno", c_int), ("megno_Ys", c_double), ("megno_Yss", c_double), ("megno_cov_Yt", c_double),
    ("megno_var_t", c_double), ("megno_mean_t", c_double), ("megno_mean_Y", c_double),
    ("megno_initial_t", c_double), ("megno_n", c_int64),
    ("rand_seed", c_uint),
    ("simulationarchive_version", c_int), ("simulationarchive_auto_interval", c_double),
    ("simulationarchive_auto_walltime", c_double), ("simulationarchive_auto_step", c_uint64),
    ("simulationarchive_next", c_double), ("simulationarchive_next_step", c_uint64),
    ("simulationarchive_filename", c_char_p),
    ("collision", c_int), ("integrator", IntegratorConfiguration), ("boundary", c_int), ("gravity", c_int),
    ("gravity_custom", c_void_p),
    ("odes", c_void_p), ("N_odes", c_size_t), ("N_allocated_odes", c_size_t),
    ("additional_forces", c_void_p), ("pre_timestep_modifications", c_void_p),
    ("post_timestep_modifications", c_void_p), ("heartbeat", HEARTBEAT), ("key_callback", c_void_p),
    ("coefficient_of_restitution", COR), ("collision_resolve", RESOLVE),
    ("free_particle_ap", c_void_p), ("extras_cleanup", c_void_p), ("extras", c_void_p),
]

class VarConfig(Structure):
    """struct reb_variational_configuration (rebound.h:702-710)."""
    _fields_ = [("sim", c_void_p), ("order", c_int), ("index", c_int), ("testparticle", c_int),
                ("index_1st_order_a", c_int), ("index_1st_order_b", c_int), ("lrescale", c_double)]


# enum values (rebound.h:209-233, 372-393; integrator_whfast.h:31-45)
JOB_STATE = {-2: "unknown", -1: "failed", 0: "pending", 1: "running", 2: "done"}
GRAVITY = {"none": 0, "basic": 1, "compensated": 2, "tree": 3, "jacobi": 5, "custom": 7}
COLLISION = {"none": 0, "direct": 1, "tree": 2, "line": 4, "linetree": 5}
BOUNDARY = {"none": 0, "open": 1, "periodic": 2, "shear": 3}
IGNORE = {"none": 0, "between_0_and_1": 1, "involving_0": 2}
WHFAST_COORDINATES = {"jacobi": 0, "democraticheliocentric": 1, "whds": 2, "barycentric": 3}
PRECISION = {"fp64": 0, "exact": 1, "split": 2}
STATUS = {-1: "RUNNING", 0: "SUCCESS", 1: "GENERIC_ERROR", 2: "NO_PARTICLES", 3: "ENCOUNTER", 4: "ESCAPE",
          5: "USER", 6: "SIGINT", 7: "COLLISION"}
RESOLVE_HALT = 0


def declare(lib):
    """Attach argument/return types to the library's entry points."""
    S = POINTER(Simulation)
    sig = {
        "reb_simulation_create": (S, []),
        "reb_simulation_free": (None, [S]),
        "reb_simulation_copy": (S, [S]),
        "reb_simulation_create_from_file": (S, [c_char_p, c_int64]),
        "reb_simulation_save_to_file": (None, [S, c_char_p]),
        "reb_simulation_save_to_file_interval": (None, [S, c_char_p, c_double]),
        "reb_simulation_save_to_file_walltime": (None, [S, c_char_p, c_double]),
        "reb_simulation_save_to_file_step": (None, [S, c_char_p, ctypes.c_uint64]),
        "reb_simulation_struct_size": (c_size_t, []),
        "reb_simulation_set_integrator": (c_void_p, [S, c_char_p]),
        "reb_simulation_add": (None, [S, Particle]),
        "reb_simulation_add_plummer": (None, [S, c_size_t, c_double, c_double]),
        "reb_particle_from_orbit": (Particle, [c_double, Particle, c_double, c_double, c_double, c_double,
                                               c_double, c_double, c_double]),
        "reb_particle_from_orbit_err": (Particle, [c_double, Particle, c_double, c_double, c_double, c_double,
                                                   c_double, c_double, c_double, POINTER(c_int)]),
        "reb_simulation_remove_particle": (c_int, [S, c_size_t]),
        "reb_simulation_steps": (None, [S, c_size_t]),
        "reb_simulation_integrate": (c_int, [S, c_double]),
        "reb_simulation_synchronize": (None, [S]),
        "reb_simulation_update_acceleration": (None, [S]),
        "reb_simulation_energy": (c_double, [S]),
        "reb_simulation_com": (Particle, [S]),
        "reb_simulation_move_to_com": (None, [S]),
        "reb_simulation_add_variation_1st_order": (c_int, [S, c_int]),
        "reb_simulation_add_variation_2nd_order": (c_int, [S, c_int, c_size_t, c_size_t]),
        "reb_simulation_rescale_var": (None, [S]),
        "reb_simulation_init_megno": (None, [S]),
        "reb_simulation_init_megno_seed": (None, [S, c_uint]),
        "reb_simulation_megno": (c_double, [S]),
        "reb_simulation_lyapunov": (c_double, [S]),
        "reb_simulation_move_to_hel": (None, [S]),
        "reb_simulation_stop": (None, [S]),
        "reb_simulation_error": (None, [S, c_char_p]),
        "reb_simulation_warning": (None, [S, c_char_p]),
        "reb_pop_message": (c_int, [POINTER(c_char_p), ctypes.c_char * 256]),
        "reb_random_uniform": (c_double, [S, c_double, c_double]),
        "reb_random_powerlaw": (c_double, [S, c_double, c_double, c_double]),
        "reb_random_normal": (c_double, [S, c_double]),
        "reb_random_rayleigh": (c_double, [S, c_double]),
        "reb_boundary_get_ghostbox": (Vec6d, [S, c_int, c_int, c_int]),
        "reb_collision_resolve_hardsphere": (c_int, [S, Collision]),
        "reb_collision_resolve_halt": (c_int, [S, Collision]),
        "reb_collision_resolve_merge": (c_int, [S, Collision]),
        "reb_integrator_register": (None, [Integrator, c_char_p]),
        "reb_integrator_whfast_kepler_solver": (None, [POINTER(Particle), c_double, c_double, S]),
        "reb_simulation_diff": (c_int, [S, S, c_int]),
        "reb_simulation_diff_char": (c_void_p, [S, S]),
        "reb_binarydata_simulation_to_stream": (None, [S, POINTER(c_void_p), POINTER(c_size_t)]),
        "reb_simulation_copy_with_messages": (None, [S, S, POINTER(c_int)]),
        "reb_free": (None, [c_void_p]),
        "reb_amd_set_precision": (c_int, [S, c_int]),
        "reb_amd_get_precision": (c_int, [S]),
        "reb_amd_device_count": (c_int, []),
        "reb_amd_device_name": (c_int, [c_char_p, c_size_t]),
        "reb_amd_version": (c_char_p, []),
        "reb_amd_last_error": (c_char_p, []),
        "reb_amd_gravity_workspace_bytes": (c_size_t, [c_size_t, c_int]),
        "reb_amd_gravity_direct": (c_int, [c_void_p] * 7 + [c_size_t, c_size_t, c_int, c_int, c_double, c_double,
                                                           c_int, c_int, c_void_p, c_void_p]),
        "reb_amd_gravity_noalloc_max_bytes": (c_size_t, []),
        "reb_amd_gravity_direct_noalloc": (c_int, [c_void_p] * 7 + [c_size_t, c_size_t, c_int, c_int, c_double,
                                                                   c_double, c_int, c_int, c_void_p]),
        "reb_amd_kepler_solver_batch": (c_int, [POINTER(c_double), POINTER(c_double), POINTER(c_double), c_size_t]),
        "reb_amd_tree_dump": (ctypes.c_long, [S, POINTER(c_double), ctypes.c_long]),
        "reb_amd_tree_interactions": (ctypes.c_longlong, [S]),
        "reb_amd_set_profiling": (None, [S, c_int]),
        "reb_amd_kernel_time": (c_int, [S, POINTER(c_double), POINTER(c_int)]),
        "reb_amd_kernel_pairs": (c_double, [S]),
        "reb_amd_set_resident": (None, [S, c_int]),
        "reb_amd_set_coefficient_of_restitution": (None, [S, c_double]),
        "reb_amd_coefficient_of_restitution_constant": (c_double, [S, c_double]),
        "reb_amd_sync_host": (c_int, [S]),
        "reb_amd_mark_host_modified": (None, [S]),
        "reb_amd_synchronize_device": (c_int, [S]),
        "reb_amd_ensemble_create": (c_void_p, [c_size_t, POINTER(c_size_t), POINTER(Particle), POINTER(c_double),
                                               c_double, c_int, c_int]),
        "reb_amd_ensemble_free": (None, [c_void_p]),
        "reb_amd_set_device_forces": (None, [S, c_void_p, c_void_p]),
        "reb_amd_ensemble_set_whfast": (c_int, [c_void_p, c_int, ctypes.c_uint, ctypes.c_uint]),
        "reb_amd_ensemble_set_whfast_kernel": (c_int, [c_void_p, ctypes.c_uint]),
        "reb_amd_ensemble_steps": (c_int, [c_void_p, ctypes.c_uint64]),
        "reb_amd_ensemble_wait": (c_int, [c_void_p]),
        "reb_amd_ensemble_steps_done": (ctypes.c_uint64, [c_void_p]),
        "reb_amd_ensemble_n_systems": (c_size_t, [c_void_p]),
        "reb_amd_ensemble_n_bodies": (c_size_t, [c_void_p]),
        "reb_amd_ensemble_get_particles": (c_int, [c_void_p, POINTER(Particle)]),
        "reb_amd_ensemble_init_megno": (c_int, [c_void_p, POINTER(Particle)]),
        "reb_amd_ensemble_megno": (c_int, [c_void_p, POINTER(c_double), POINTER(c_double)]),
        "reb_amd_ensemble_get_var_particles": (c_int, [c_void_p, POINTER(Particle)]),
        "reb_amd_ensemble_energy": (c_int, [c_void_p, POINTER(c_double)]),
        "reb_amd_ensemble_flags": (c_int, [c_void_p, POINTER(ctypes.c_uint)]),
        "reb_amd_ensemble_set_launch": (c_int, [c_void_p, c_int, ctypes.c_long]),
        "reb_amd_ensemble_set_packing": (c_int, [c_void_p, c_int]),
        "reb_amd_ensemble_get_launch": (c_int, [c_void_p, POINTER(c_int), POINTER(ctypes.c_long), POINTER(c_int),
                                                POINTER(c_int)]),
        "reb_amd_ensemble_tune": (c_int, [c_void_p, ctypes.c_long, POINTER(c_double)]),
        "reb_amd_ensemble_error": (c_char_p, [c_void_p]),
        "reb_amd_server_create": (c_void_p, [c_size_t, c_int, ctypes.c_long]),
        "reb_amd_server_free": (None, [c_void_p]),
        "reb_amd_server_submit_whfast": (ctypes.c_long, [c_void_p, c_size_t, POINTER(Particle), c_double, c_double,
                                                         c_int, ctypes.c_uint64]),
        "reb_amd_server_submit_nbody": (ctypes.c_long, [c_void_p, c_size_t, POINTER(Particle), c_double, c_double,
                                                        c_double, ctypes.c_uint64, c_int, c_int]),
        "reb_amd_server_status": (c_int, [c_void_p, ctypes.c_long]),
        "reb_amd_server_submit_simulation": (ctypes.c_long, [c_void_p, S, ctypes.c_uint64]),
        "reb_amd_server_checkpoint": (c_int, [c_void_p, ctypes.c_long, c_char_p]),
        "reb_amd_server_whfast_bytes": (c_size_t, [c_size_t]),
        "reb_amd_server_nbody_bytes": (c_size_t, [c_size_t, c_int]),
        "reb_amd_server_wait": (c_int, [c_void_p, ctypes.c_long, c_double]),
        "reb_amd_server_result": (c_int, [c_void_p, ctypes.c_long, POINTER(Particle), c_size_t, POINTER(c_double),
                                          POINTER(c_double)]),
        "reb_amd_server_job_error": (c_char_p, [c_void_p, ctypes.c_long]),
        "reb_amd_server_stats": (c_int, [c_void_p, POINTER(c_double), c_int]),
    }
    for name, (res, args) in sig.items():
        if hasattr(lib, name):
            f = getattr(lib, name)
            f.restype = res
            f.argtypes = args
    return lib
