// Shared descriptors of the WHFast group kernel (whfast.hip), used by the
// per-simulation plugin and by the batched ensemble (ensemble.hip).
#pragma once

#include <cstddef>

namespace rbamd {

// One planetary system inside a body array (structure of arrays, `stride`
// doubles per field): bodies body0 .. body0+N-1, body0 is the central body.
struct WHSys {
    int body0;
    int N;           // bodies including the central one
    int Nact;        // N_active used by the WHFast sums (N unless test particles of type 0)
    int Nact_grav;   // N_active used by the gravity loop
    int tp_type;     // testparticle_type
    int steps;       // steps in the current launch (-1: the launch's count)
    int req;         // requests for this launch: bit0 synchronize at its end,
                     // bit1 particles modified (recompute the heliocentric state),
                     // bit2 safe mode for this system
    int pad;
    double dt;
    double G;
};

// A wavefront's unit of work: `nsys` systems of one group size G (lanes per
// system), taken from order[sys0 .. sys0+nsys).
struct WHTask {
    int G;
    int sys0;
    int nsys;
};

// Coordinate systems of the group kernel (integrator_whfast.h:39-46).
enum { WH_DH = 0, WH_JACOBI = 1 };

struct WHParams {
    const WHSys* sys;
    const WHTask* tasks;
    const int* order;
    unsigned int* flags;      // per system: bit0 synchronized, bit1 period warning, bit2 resync warning
    unsigned int* counter;    // work counter of the persistent grid
    int ntasks;
    int grid;                 // persistent workgroups (one wavefront each)
    double* body;             // fields x y z vx vy vz ax ay az m (r unused), stride doubles apart
    double* ph;               // heliocentric fields x y z vx vy vz m
    size_t stride;
    double G;
    double soft2;
    int compensated;
    long nsteps;
    int safe_mode;
    int did_modify;
    int sync_at_end;
    int coords = WH_DH;       // WH_DH or WH_JACOBI
    int corrector = 0;        // first symplectic corrector order (Jacobi): 0, 3, 5, 7, 11, 17
    int corrector2 = 0;       // second corrector (Jacobi): 0 or 1
    int kernel = 0;           // 0 default, 1 modified kick, 2 composition, 3 lazy (Jacobi)
    int jacgrav = 0;          // accelerations by the Jacobi gravity (kernels 1 and 3)
    // First-order variational equations (Jacobi, one system per launch).
    // Variational particle k of the simulation is entry k of `pv` (inertial,
    // fields x y z vx vy vz ax ay az m, `vstride` doubles apart) and of `jv`
    // (its Jacobi copy, the same fields); set v covers entries
    // var_index[v] .. var_index[v] + N - 1.
    int nvar = 0;
    const int* var_index = nullptr;
    double* pv = nullptr;
    double* jv = nullptr;
    size_t vstride = 0;
    int megno = 0;                // MEGNO: <deltad, delta>/|delta|^2 of the step into megno_out
    double* megno_out = nullptr;
};

int whfast_launch(const WHParams& P, hipStream_t st);
int whfast_energy_launch(const WHParams& P, double* out, int nsys, hipStream_t st);
int whfast_group_size(int K);
int whfast_max_resident(int device);

}  // namespace rbamd
