"""The drop-in boundary: the library loads, exports exactly what include/*.h
declares, and its structs have REBOUND's layout (no GPU needed)."""
import ctypes
import os
import re
import shutil
import subprocess

import pytest

from conftest import ROOT

HEADERS = [os.path.join(ROOT, "include", h) for h in ("rebound.h", "rebound_amd.h")]


def declared_functions():
    names = set()
    for h in HEADERS:
        txt = open(h).read()
        for m in re.finditer(r"REB_API\s+(?:extern\s+)?[^;(]*?\b(reb_\w+)\s*[(;]", txt):
            names.add(m.group(1))
    return names


def exported_symbols():
    out = subprocess.check_output(["nm", "-D", "--defined-only", os.path.join(ROOT, "rebound_amd",
                                                                              "librebound_amd.so")], text=True)
    return {ln.split()[-1] for ln in out.splitlines() if ln.strip()}


def test_library_loads(amd):
    assert amd.lib().reb_simulation_struct_size() == 816
    assert b"gfx950" in amd.lib().reb_amd_version()


def test_every_declared_symbol_is_exported(amd):
    decl = declared_functions()
    exp = exported_symbols()
    assert len(decl) > 40
    missing = sorted(decl - exp)
    assert not missing, missing


def test_only_reb_symbols_exported(amd):
    # REBOUND's CI rule (.github/workflows/c.yml): every exported symbol starts with reb_
    bad = sorted(s for s in exported_symbols() if not s.startswith("reb_"))
    assert not bad, bad


def test_struct_sizes(amd):
    from rebound_amd import _abi
    assert ctypes.sizeof(_abi.Particle) == 112
    assert ctypes.sizeof(_abi.Collision) == 72
    assert ctypes.sizeof(_abi.Simulation) == 816
    assert ctypes.sizeof(_abi.WHFastState) == 64


PROBE = r"""
#include <stdio.h>
#include <stddef.h>
#include "rebound.h"
#define F(f) printf("sim.%s %zu\n", #f, offsetof(struct reb_simulation, f));
#define W(f) printf("whfast.%s %zu\n", #f, offsetof(struct reb_integrator_whfast_state, f));
#define E(v) printf("enum.%s %d\n", #v, (int)(v));
int main(void){
  printf("size %zu %zu %zu %zu\n", sizeof(struct reb_particle), sizeof(struct reb_collision),
         sizeof(struct reb_simulation), sizeof(struct reb_integrator_whfast_state));
  F(t) F(G) F(softening) F(OMEGA) F(OMEGAZ) F(dt) F(dt_last_done) F(steps_done) F(is_synchronized)
  F(did_modify_particles) F(N) F(N_allocated) F(particles) F(N_map) F(map) F(N_var) F(particles_var)
  F(N_var_config) F(var_config) F(N_active) F(testparticle_type) F(testparticle_hidewarnings) F(name_list)
  F(N_name_list) F(name_hash_table) F(gravity_cs) F(N_allocated_gravity_cs) F(tree_root) F(opening_angle2)
  F(status) F(exact_finish_time) F(force_is_velocity_dependent) F(gravity_ignore_terms) F(output_timing_last)
  F(save_messages) F(messages) F(messages_var_rescale_warning) F(messages_timestep_warning) F(exit_max_distance)
  F(exit_min_distance) F(usleep) F(display_settings) F(display_data) F(server_data) F(track_energy_offset)
  F(energy_offset) F(walltime) F(walltime_last_step) F(walltime_last_steps) F(walltime_last_steps_sum)
  F(walltime_last_steps_N) F(python_unit_l) F(python_unit_m) F(python_unit_t) F(root_size) F(N_root_x)
  F(N_root_y) F(N_root_z) F(N_ghost_x) F(N_ghost_y) F(N_ghost_z) F(collisions) F(N_allocated_collisions)
  F(N_collisions) F(N_targets) F(minimum_collision_velocity) F(collisions_plog) F(collisions_log_n)
  F(calculate_megno) F(megno_Ys) F(megno_Yss) F(megno_cov_Yt) F(megno_var_t) F(megno_mean_t) F(megno_mean_Y)
  F(megno_initial_t) F(megno_n) F(rand_seed) F(simulationarchive_version) F(simulationarchive_auto_interval)
  F(simulationarchive_auto_walltime) F(simulationarchive_auto_step) F(simulationarchive_next)
  F(simulationarchive_next_step) F(simulationarchive_filename) F(collision) F(integrator) F(boundary) F(gravity)
  F(gravity_custom) F(odes) F(N_odes) F(N_allocated_odes) F(additional_forces) F(pre_timestep_modifications)
  F(post_timestep_modifications) F(heartbeat) F(key_callback) F(coefficient_of_restitution) F(collision_resolve)
  F(free_particle_ap) F(extras_cleanup) F(extras)
  W(corrector) W(corrector2) W(kernel) W(coordinates) W(safe_mode) W(keep_unsynchronized) W(N_allocated)
  W(p_jh) W(N_allocated_var) W(p_jh_var) W(recalculate_coordinates_but_not_synchronized_warning)
  E(REB_GRAVITY_BASIC) E(REB_GRAVITY_COMPENSATED) E(REB_GRAVITY_TREE) E(REB_COLLISION_DIRECT) E(REB_COLLISION_TREE)
  E(REB_BOUNDARY_SHEAR) E(REB_GRAVITY_IGNORE_TERMS_INVOLVING_0) E(REB_STATUS_COLLISION)
  E(REB_INTEGRATOR_WHFAST_COORDINATES_DEMOCRATICHELIOCENTRIC)
  return 0;
}
"""


def _probe(tmp_path, incdir, tag):
    src = tmp_path / ("probe_%s.c" % tag)
    src.write_text(PROBE)
    exe = tmp_path / ("probe_%s" % tag)
    subprocess.check_call(["gcc", "-std=c99", "-I" + incdir, str(src), "-o", str(exe)])
    return subprocess.check_output([str(exe)], text=True)


def test_layout_identical_to_reference(tmp_path):
    """Field offsets and enum values equal those of REBOUND's own rebound.h
    (compiled with gcc against both headers; needs the reference sources)."""
    ref_inc = "/root/reference/src"
    if not os.path.isdir(ref_inc) or not shutil.which("gcc"):
        pytest.skip("reference headers not present on this machine")
    ours = _probe(tmp_path, os.path.join(ROOT, "include"), "ours")
    theirs = _probe(tmp_path, ref_inc, "ref")
    assert ours == theirs


def test_offsets_match_python_mirror(tmp_path):
    if not shutil.which("gcc"):
        pytest.skip("no C compiler")
    from rebound_amd import _abi
    out = _probe(tmp_path, os.path.join(ROOT, "include"), "mirror")
    for ln in out.splitlines():
        key, val = ln.split(" ", 1)
        if key.startswith("sim."):
            assert getattr(_abi.Simulation, key[4:]).offset == int(val), key
        if key.startswith("whfast."):
            assert getattr(_abi.WHFastState, key[7:]).offset == int(val), key
