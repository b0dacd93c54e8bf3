// Internal declarations shared by the host runtime and the HIP kernels of the
// MI355X force-and-step engine.  Nothing here is part of the public ABI.
#pragma once

#include <hip/hip_runtime.h>
#include <cstddef>
#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>

#include "rebound.h"
#include "rebound_amd.h"

// The enums declared inside struct reb_simulation / reb_integrator_whfast_state
// are file-scope in C but member-scope in C++; bring their names into scope.
#define RB_SIM_ENUM(name) constexpr auto name = reb_simulation::name;
RB_SIM_ENUM(REB_GRAVITY_IGNORE_TERMS_NONE)
RB_SIM_ENUM(REB_GRAVITY_IGNORE_TERMS_BETWEEN_0_AND_1)
RB_SIM_ENUM(REB_GRAVITY_IGNORE_TERMS_INVOLVING_0)
RB_SIM_ENUM(REB_COLLISION_NONE)
RB_SIM_ENUM(REB_COLLISION_DIRECT)
RB_SIM_ENUM(REB_COLLISION_TREE)
RB_SIM_ENUM(REB_COLLISION_LINE)
RB_SIM_ENUM(REB_COLLISION_LINETREE)
RB_SIM_ENUM(REB_BOUNDARY_NONE)
RB_SIM_ENUM(REB_BOUNDARY_OPEN)
RB_SIM_ENUM(REB_BOUNDARY_PERIODIC)
RB_SIM_ENUM(REB_BOUNDARY_SHEAR)
RB_SIM_ENUM(REB_GRAVITY_NONE)
RB_SIM_ENUM(REB_GRAVITY_BASIC)
RB_SIM_ENUM(REB_GRAVITY_COMPENSATED)
RB_SIM_ENUM(REB_GRAVITY_TREE)
RB_SIM_ENUM(REB_GRAVITY_JACOBI)
RB_SIM_ENUM(REB_GRAVITY_CUSTOM)
#undef RB_SIM_ENUM
#define RB_WH_ENUM(name) constexpr auto name = reb_integrator_whfast_state::name;
RB_WH_ENUM(REB_INTEGRATOR_WHFAST_KERNEL_DEFAULT)
RB_WH_ENUM(REB_INTEGRATOR_WHFAST_COORDINATES_JACOBI)
RB_WH_ENUM(REB_INTEGRATOR_WHFAST_COORDINATES_DEMOCRATICHELIOCENTRIC)
#undef RB_WH_ENUM

namespace rbamd {

// ------------------------------------------------------------------ errors
// Record a HIP failure on the simulation (turns into REB_STATUS_GENERIC_ERROR
// at the next exit check, as reb_simulation_error does in the reference).
bool hip_check(struct reb_simulation* r, hipError_t e, const char* what);
#define RB_HIP(r, call) ::rbamd::hip_check((r), (call), #call)

void set_last_error(const std::string& msg);

// ------------------------------------------------------------------ device buffers
// A growable device allocation.  Freed explicitly (never in a destructor that
// could run after the HIP runtime has shut down).
struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
    bool ensure(size_t need);            // grow (contents not preserved)
    void release();
    template <class T> T* as() const { return static_cast<T*>(p); }
};

// Field order of the structure-of-arrays particle block on the device.
enum Field { FX = 0, FY, FZ, FVX, FVY, FVZ, FAX, FAY, FAZ, FM, FR, NFIELD };

// Where the authoritative copy of the particle state lives.
enum Loc { LOC_HOST = 0, LOC_DEVICE = 1, LOC_BOTH = 2 };

struct TreeBuffers;          // tree.hip
struct ResolveScratch;       // collision_host.hip

// Collisions of the last device search in detection order, in pinned host
// memory: bodies, ghost-box index, root box (the device record's layout),
// and the ghost boxes of that search.  r->collisions is filled from it when
// the resolve needs the reference's struct list.
struct CollRecH {
    int p1, p2, gb, ri;
};
struct CollList {
    CollRecH* rec = nullptr;
    size_t cap = 0, n = 0;
    std::vector<reb_vec6d> gb;
    bool pending = false;     // r->collisions not yet written from rec
    bool rec_host = false;    // rec holds the records (else only the device does)
};
struct WHFastDevice;         // whfast.hip

// Device-side state of one simulation.  Allocated next to the public struct
// by reb_simulation_create; never visible through the ABI.
struct DeviceState {
    int device = -1;
    hipStream_t stream = nullptr;
    size_t cap = 0;                 // particle capacity of `soa`
    DevBuf soa;                     // NFIELD * cap doubles
    DevBuf partial;                 // j-split partial sums for direct gravity
    DevBuf scratch;                 // small reductions (energy, flags)
    DevBuf aos;                     // AoS staging for uploads/downloads
    DevBuf flags;                   // device error words (boundary wrap)
    double* host_stage = nullptr;   // pinned staging, NFIELD * host_cap doubles
    size_t host_cap = 0;
    int loc = LOC_HOST;
    int precision = REB_AMD_PRECISION_FP64;
    int gravity_tile = 0;           // autotuned j-split (0 = pick)
    TreeBuffers* tree = nullptr;
    WHFastDevice* whfast = nullptr;
    // kernel timing: when `profiling` is set, every launch of the dominant
    // pair kernel of direct gravity is bracketed by an event pair from `kev`
    bool profiling = false;
    std::vector<hipEvent_t> kev;
    size_t kev_used = 0;
    // device residency: when set, steps()/integrate() neither upload nor
    // download r->particles; reb_amd_sync_host / reb_amd_mark_host_modified
    // move the state explicitly
    bool resident = false;
    // the stream was handed in by the library's serving loop (not destroyed here)
    bool owns_stream = true;
    // value returned by reb_amd_coefficient_of_restitution_constant
    double cor_constant = 1.0;
    // host work arrays of the collision resolve, kept between steps
    ResolveScratch* resolve = nullptr;
    // collisions of the last device search (pinned host copy)
    CollList* coll = nullptr;
    // device-side additional forces (reb_amd_set_device_forces)
    reb_amd_device_forces_fn dev_forces = nullptr;
    void* dev_forces_user = nullptr;
    // the EXACT energy fell back to the compensated sum once (warned)
    bool warned_energy_exact = false;
    // task list of the symmetric direct-gravity path: for N bodies, at this
    // device address inside `partial`
    std::vector<int2> sym_tasks_host;
    const void* sym_tasks_dev = nullptr;
    int sym_tasks_n = -1;
    // ordered pairs evaluated by the timed pair kernels since the last read
    double kpairs = 0.;
    double* field(int f) const { return soa.as<double>() + (size_t)f * cap; }
};

// The allocation returned by reb_simulation_create: public struct first.
struct SimBlock {
    struct reb_simulation pub;
    DeviceState dev;
};
inline DeviceState* dev_of(struct reb_simulation* r) { return &reinterpret_cast<SimBlock*>(r)->dev; }
inline const DeviceState* dev_of(const struct reb_simulation* r) { return &reinterpret_cast<const SimBlock*>(r)->dev; }

// ------------------------------------------------------------------ residency
bool device_ready(struct reb_simulation* r);            // pick device, create stream
bool ensure_device(struct reb_simulation* r);            // host -> device if host is newer
bool ensure_host(struct reb_simulation* r);              // device -> host if device is newer
bool prepare_host(struct reb_simulation* r);             // ensure_host, and WHFast's p_jh current
void mark_host_newer(struct reb_simulation* r);
void mark_device_newer(struct reb_simulation* r);
bool stream_sync(struct reb_simulation* r);
void device_release(struct reb_simulation* r);

// ------------------------------------------------------------------ ghost boxes
// All ghost-box offsets of the (2Nx+1)(2Ny+1)(2Nz+1) grid in the reference's
// loop order (x outer, z inner).  gravity.c:84-86, boundary.c:145-201.
std::vector<struct reb_vec6d> ghostboxes(struct reb_simulation* r, int nx, int ny, int nz);

// ------------------------------------------------------------------ kernels (launchers)
// Direct O(N^2) gravity (gravity.c:167-531) on the simulation's device state.
bool gravity_direct(struct reb_simulation* r, bool compensated, const double* kick_drift = nullptr);
// The pair-symmetric direct-gravity path (gravity.hip) for N bodies, all
// active, no ghost boxes, nothing ignored; its workspace in bytes.
bool gravity_symmetric_path(size_t N, int precision, bool compensated);
size_t gravity_symmetric_bytes(size_t N);

// Batched order-2 leapfrog steps with direct gravity for many independent
// simulations (gravity.hip k_seg_*).  A SegJob holds one simulation's device
// arrays, the chunking it would use alone (seg_plan) and its constants.
struct SegJob {
    double *x, *y, *z, *vx, *vy, *vz, *ax, *ay, *az;
    const double* m;
    double* partial;
    int n, chunk, nchunk, mode;
    double negG, soft2, kick, drift;
};
struct SegTile {     // one pair workgroup: job, target block, source chunk
    int job, bx, by;
};
struct SegRange {    // 256 bodies of one job
    int job, first;
};
void seg_plan(int n, int* chunk, int* used, size_t* workspace);
int seg_tiles_of(int n, int used);
int seg_step_launch(const SegJob* jobs, const SegRange* ranges, int nranges, const SegTile* tiles, int ntiles,
                    hipStream_t st);
int seg_fast_mode(int precision, int compensated);
// Energy (tools.c:108-162) of device arrays into out[0] (kinetic) and out[1]
// (potential) on the stream; `scratch` holds 2 n + 16 doubles.
int energy_launch(const double* x, const double* y, const double* z, const double* vx, const double* vy,
                  const double* vz, const double* m, int n, int n_active, int n_interact, double G, bool exact,
                  double* scratch, hipStream_t st, double* terms = nullptr, size_t terms_cap = 0);
bool gravity_zero(struct reb_simulation* r);
bool gravity_jacobi(struct reb_simulation* r);               // gravity.c:108-163
bool gravity_tree(struct reb_simulation* r);                 // tree.hip
bool collision_search_device(struct reb_simulation* r);      // tree.hip / collision
double energy_device(struct reb_simulation* r);              // tools.c:108-162

// Streams (integrator_leapfrog.c:72-93, integrator_sei.c:126-174, boundary.c:35-141)
bool launch_drift(struct reb_simulation* r, double dt);
bool launch_kick(struct reb_simulation* r, double dt);
bool launch_sei_h012(struct reb_simulation* r, double dt, double sindt, double tandt,
                     double sindtz, double tandtz, int with_kick);
bool boundary_device(struct reb_simulation* r);              // returns false on error
bool check_device_flags(struct reb_simulation* r);           // error words set by kernels

// WHFast (whfast.hip)
void whfast_step_device(struct reb_simulation* r, struct reb_integrator_whfast_state* s);
void whfast_synchronize_device(struct reb_simulation* r, struct reb_integrator_whfast_state* s);
void whfast_release(struct reb_simulation* r);
void megno_update(struct reb_simulation* r, double dY, double dt_done);   // host_api.cpp
void tree_release(struct reb_simulation* r);
void resolve_release(struct reb_simulation* r);
// The built-in hard-sphere resolve with a constant restitution runs on compact
// records the device prepares right after a search (collision_host.hip):
// compact_resolve_mark numbers the bodies of the nc device records; after the
// next stream synchronization compact_resolve_gather packs them and starts
// their copy to the host.  False on an allocation failure.
bool compact_resolve_eligible(const struct reb_simulation* r, int n_ghostboxes);
// Draws `expect` rand_r results from r->rand_seed ahead of the shuffle (while
// the device searches), without advancing r->rand_seed.
void compact_resolve_draw_ahead(struct reb_simulation* r, size_t expect);
bool compact_resolve_mark(struct reb_simulation* r, const CollRecH* rec, size_t nc);
bool compact_resolve_gather(struct reb_simulation* r, const CollRecH* rec, size_t nc);
// r->collisions lives in pinned host memory (the device copies the shuffled
// list into it): grown to at least nc entries (contents not kept), and freed.
bool collisions_reserve(struct reb_simulation* r, size_t nc);
void collisions_free(struct reb_simulation* r);
// Writes r->collisions from the last device search (detection order) when it
// is still pending; false on an allocation failure.
bool materialize_collisions(struct reb_simulation* r);
void coll_release(struct reb_simulation* r);
// simulationarchive.c:377-409: automatic snapshots by interval, step or walltime
void archive_heartbeat(struct reb_simulation* r);

// Host-side driver pieces shared between files
void collision_resolve_host(struct reb_simulation* r);       // collision.c:336-404
void run_host_callback(struct reb_simulation* r, void (*cb)(struct reb_simulation* const));

// Monotonic clock (seconds)
double now_seconds();

}  // namespace rbamd
