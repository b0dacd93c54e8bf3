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
    // First-order variational equations (Jacobi).  `pv` holds the inertial
    // variational particles (fields x y z vx vy vz ax ay az m, `vstride`
    // doubles apart) and `jv` their Jacobi copies (the same fields); set v of
    // the system whose bodies start at body0 covers entries
    // body0 + var_index[v] .. body0 + var_index[v] + N - 1.
    int nvar = 0;
    const int* var_index = nullptr;
    double* pv = nullptr;
    double* jv = nullptr;
    size_t vstride = 0;
    // MEGNO: 1 = the step's <deltad, delta>/|delta|^2 into megno_out (one
    // system, the host keeps the sums); 2 = each system's time and sums in
    // megno_sys (kMegnoFields per system: t, initial t, Ys, Yss, cov_Yt,
    // var_t, mean_t, mean_Y, n), updated on the device every step
    int megno = 0;
    double* megno_out = nullptr;
    double* megno_sys = nullptr;
};
constexpr int kMegnoFields = 9;

int whfast_launch(const WHParams& P, hipStream_t st);
int whfast_energy_launch(const WHParams& P, double* out, int nsys, hipStream_t st);
int whfast_group_size(int K);
int whfast_max_resident(int device);

}  // namespace rbamd
