// Internal host/device structures of libpypat_hip.so (not part of the C ABI).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "pypat_hip.h"

// Device view of one field map.  Node arrays g[a] and the raw value table are
// what the strict FP64 gather reads; coef64/coef32 are the per-cell trilinear
// coefficient tables of the fast paths.
struct MapDesc {
    int32_t n[3];       // nodes per axis as given (1 = inactive axis)
    int32_t ng[3];      // length of g[a] (z gets one extra node when periodic)
    int32_t nc[3];      // cells per axis in the coefficient tables (1 on inactive axes)
    int32_t ncomp;      // components stored (3 or 6)
    int32_t periodic;   // z wraps with period L
    int32_t uniform;    // every active axis is close to uniform: guess-and-correct lookup is exact
    double lo[3];       // g[a][0]
    double hi[3];       // g[a][ng-1]
    double inv_d[3];    // (ng-1) / (hi-lo) on active axes, 0 otherwise
    double L;           // z period (periodic maps)
    double fill;        // value outside the grid
    double inv_nz;      // 1 / nc[2] (periodic maps)
    const double* g[3];
    const double* raw;      // [comp][nx][ny][ng_z] float64
    const double* coef64;   // [cell][comp][8] float64, cell = (i*nc1 + j)*nc2 + k
    const float* coef32;    // same layout, float32
};

// Device form of one pp_field term.
struct FieldDev {
    int32_t kind;       // pp_field_kind
    int32_t comp0;
    int32_t has_scale2;
    int32_t has_offset;
    int32_t rf;         // pp_rf_mode
    int32_t _pad;
    double scale1, scale2, offset;
    double value[3];
    const double* array;    // PP_FIELD_ARRAY: (n, 3)
    MapDesc map;
};

struct PushParams {
    int64_t n;
    double* x; double* y; double* z;
    double* vx; double* vy; double* vz;
    uint8_t* alive;
    const double* dt_pp;
    double dt, qm;
    int64_t nsteps, step0;
    FieldDev E, B;
    int32_t fused;          // E and B read the same map: one cell lookup for both
    int32_t kick_only;
    const double* rf_table;
    int64_t rf_step0;
    double* r_hist; double* v_hist;
    int64_t rec_every, n_records;
    int64_t hist_rows;      // particles per history record (>= n)
    const int32_t* ids;     // original particle index per slot (sorted state), or null
    // electrode mesh of the loss test inside the step loop (device pp_bvh::MeshDev); null: no test
    const void* mesh;
    // constants computed on the host with the reference's Python expressions
    double clight, c2, vclamp;
    double yc[3], yd[3];    // Yoshida coefficients, pusher.py:176-181
};

// Hot path: E and B from one uniform 3-D map (components 0..2 = E, 3..5 = B),
// no RF / offset terms.  Compact, so the step loop keeps it in SGPRs.
struct HotMap {
    const void* coef;       // [cell][6][8] float32 or float64 coefficient table
    uint32_t bytes;         // table size (buffer range: loads beyond it return 0)
    int32_t nc0, nc1, nc2;  // cells per axis
    double c0, c1, c2;      // -lo[a] * inv_d[a]
    double id0, id1, id2;   // inv_d[a]
    double nc2d, nc12d;     // nc2 and nc1 * nc2 as doubles: cell = i nc12d + j nc2d + k, exact in FP64
    double sE, sB;          // scale1 * scale2 of E and B
    double fill;            // value outside the grid: 0 (the hot kernels take only such maps)
    uint32_t ncells;        // cells in the table (nc0 nc1 nc2)
    uint32_t zmagic;        // floor((2^32 - 1) / nc2): quotient estimate for the periodic z wrap
    uint32_t zbias;         // nc2 * floor(2^30 / nc2): makes the z cell index non-negative before the wrap
    uint32_t zmask;         // nc2 - 1 when nc2 is a power of two, else 0
    double zinv;            // 1 / nc2 (exact when nc2 is a power of two)
};

// RF factors of the hot kernels: table [step - step0][stage][cos, sin]; e, b = pp_rf_mode of E and B
struct HotRF {
    const double* table;   // NULL: static fields
    int64_t step0;
    int32_t e, b;
};

struct HotParams {
    double* x; double* y; double* z;
    double* vx; double* vy; double* vz;
    const int32_t* ids;
    int64_t n, nsteps, step0;
    double dt, qm, c2, clight, vclamp;
    double* r_hist; double* v_hist;
    int64_t rec_every, n_records;
    int64_t hist_rows;       // rows of a history record (>= n: a compacted state covers a live prefix)
    uint8_t* alive;          // alive mask, or null (every particle live)
    const void* mesh;        // device pp_bvh::MeshDev of the loss test, or null (with alive: no test)
    int32_t kick_only;
    int32_t _pad;
    HotMap m;
    HotRF rf;
};

struct BatchParams {
    double* x; double* y; double* z;
    double* vx; double* vy; double* vz;
    const pp_bunch* bunches;
    const int32_t* tile_bunch;
    const int64_t* tile_start;
    int64_t ntiles;
    int64_t n;              // particles in the state (bounds)
    int64_t step0, nsteps;
    const int64_t* step0_dev;
    double c2, clight, vclamp;
    int32_t nmaps;
    int32_t nbunch;
    HotMap maps[PP_MAX_BATCH_MAPS];
};

// Host-side map object behind the opaque pp_map handle.
struct pp_map {
    MapDesc d;
    size_t bytes;
    void* allocs[6];
    int nalloc;
    int device;
};

namespace pp_strict {
hipError_t launch_push(int alg, const PushParams& p, int block, hipStream_t s);
hipError_t launch_field_eval(const FieldDev& f, const double* pts, int64_t m, double rf, double* out,
                             hipStream_t s);
struct BatchArgs {
    int which;
    const double* a; const double* b; const double* c; const double* d; const double* e;
    const double* s_pp;
    int64_t m;
    double s1, s2, qm;
    int32_t f1, f2;
    double c2, clight, vclamp;
    double* out;
};
hipError_t launch_batch_kernel(const BatchArgs& args, hipStream_t s);
hipError_t launch_util(int which, double* a, double* b, double* c, double* aos, int64_t m, double c2,
                       hipStream_t s);
hipError_t launch_fill(double* p, double v, int64_t n, hipStream_t s);
hipError_t launch_add_i64(int64_t* p, int64_t v, hipStream_t s);
hipError_t launch_build_coef(const MapDesc& d, double* c64, float* c32, hipStream_t s);
}  // namespace pp_strict

namespace pp_fast {
hipError_t launch_push(int alg, int prec, const PushParams& p, int block, hipStream_t s);
hipError_t launch_hot(int alg, int prec, bool periodic, const HotParams& p, int block, hipStream_t s);
hipError_t launch_batched(int prec, bool periodic, int tile_class, const BatchParams& p, int block, hipStream_t s);
hipError_t launch_field_eval(const FieldDev& f, int prec, const double* pts, int64_t m, double rf, double* out,
                             hipStream_t s);
hipError_t launch_gather_probe(const float* table, uint32_t bytes, uint32_t ncell, int mode, int iters,
                               int64_t nthreads, float* out, hipStream_t s);
}  // namespace pp_fast

// SoA state viewed by the beam-moment reductions
struct MomentArgs {
    const double* x; const double* y; const double* z;
    const double* vx; const double* vy; const double* vz;
    const uint8_t* alive;       // NULL: every slot counts
    int64_t n;
    int32_t u_state;            // velocity slots hold u = gamma v
    int32_t z_energy;           // kinetic energy from p_z only (Z_ENERGY setting)
    double mass_mev, clight, beta2_max;
};

constexpr int kMomentBlocks = 1024;   // partials per reduction pass

// spline interpolators (pp_spline.hip)
namespace pp_splk {
struct SplineDesc {
    const double* coef;        // [ncomp][N0][N1][N2], last axis fastest (N = n + 2 pad)
    const double* axis[3];     // node coordinates, n[a] each
    int64_t n[3];              // nodes per axis (1 for unused axes)
    int64_t N[3];              // coefficient extent per axis
    int64_t ncomp;
    int32_t ndim, order, nearest, rgi, pad;
    double cval;
};
}  // namespace pp_splk

struct pp_spline {
    pp_splk::SplineDesc d;
    void* allocs[4];
    int nalloc;
    int device;
};

namespace pp_spl {
void destroy(pp_spline* s);
hipError_t create(const double* values, int64_t ncomp, int ndim, const int64_t* n, const double* const* axes, int order,
                  int nearest, int rgi, double cval, int device, pp_spline** out);
hipError_t eval(const pp_spline* s, const double* pts, int64_t m, int32_t stride, int32_t col0, double* out,
                hipStream_t st);
}  // namespace pp_spl

namespace pp_beam {
size_t workspace_bytes();
hipError_t moments(const MomentArgs& a, double* sums, double* central, void* ws, hipStream_t s);
hipError_t ellipse_counts(const MomentArgs& a, const double* twiss_dev, double* counts, void* ws, hipStream_t s);
// fixed per-device scratch of pp_beam_moments_noalloc (see pp_moments.hip)
size_t noalloc_scratch_bytes();
hipError_t init_noalloc_scratch(int device);   // with `device` current; once per device
void* noalloc_scratch(int device);             // device address, or null before init_noalloc_scratch
}  // namespace pp_beam

// box mesh of the space-charge solve: n cell-centre nodes per axis, spacing h = L / n
struct PoissonDesc {
    int32_t n[3];
    int32_t _pad;
    double h[3];
    double L[3];
};

constexpr double kEps0 = 8.8541878188e-12;   // vacuum permittivity (scipy.constants.epsilon_0, CODATA 2022)

struct pp_poisson;
namespace pp_poisson_impl {
int create(const PoissonDesc& d, pp_poisson** out, char* err, size_t errn);
void destroy(pp_poisson* p);
const PoissonDesc& desc(const pp_poisson* p);
int device_of(const pp_poisson* p);
hipError_t deposit(const pp_poisson* p, const double* x, const double* y, const double* z, const double* q,
                   double qc, const uint8_t* alive, int64_t n, double* rho, hipStream_t s);
hipError_t solve(const pp_poisson* p, const double* rho, double* phi, hipStream_t s);
hipError_t gradient(const pp_poisson* p, const double* phi, double* ex, double* ey, double* ez, hipStream_t s);
// electrode-bounded domains: classification and Shortley-Weller operator from a closed triangle mesh
int set_conductors(pp_poisson* p, const double* vertices, int64_t nv, const int32_t* triangles, int64_t nt, char* err,
                   size_t errn);
int set_ray_data(pp_poisson* p, const double* dist, const int32_t* count, char* err, size_t errn);
bool has_conductors(const pp_poisson* p);
void conductor_data(const pp_poisson* p, int32_t* type, double* dist, int32_t* count);
void set_tolerance(pp_poisson* p, double rtol, int32_t max_iter);
void stats(const pp_poisson* p, int32_t* iterations, double* residual);
}  // namespace pp_poisson_impl

struct pp_mesh;
namespace pp_collide {
int create(const double* vertices, int64_t nv, const int32_t* triangles, int64_t nt, pp_mesh** out, char* err,
           size_t errn);
void destroy(pp_mesh* m);
void info(const pp_mesh* m, int64_t* ntris, int64_t* nnodes);
int device_of(const pp_mesh* m);
const void* device_view(const pp_mesh* m);   // device pp_bvh::MeshDev, or null for an empty mesh
hipError_t collide(const pp_mesh* m, const double* x0, const double* y0, const double* z0, const double* x1,
                   const double* y1, const double* z1, int64_t n, uint8_t* alive, uint8_t* hit, hipStream_t s);
}  // namespace pp_collide

namespace pp_sort {
hipError_t workspace_bytes(int64_t n, size_t* bytes);
hipError_t sort_state(const MapDesc& m, double* arrs[6], int32_t* ids, int64_t n, void* ws, size_t ws_bytes,
                      hipStream_t s, const int32_t* group = nullptr, int group_shift = 0, int refine_axis = 2,
                      int refine_bits = 0);
// alive (optional) is scattered back with the state
hipError_t unsort_state(double* arrs[6], int32_t* ids, int64_t n, void* ws, size_t ws_bytes, hipStream_t s,
                        uint8_t* alive = nullptr);
hipError_t iota(int32_t* ids, int64_t n, hipStream_t s);
// Stable partition of the state into live slots (front) and dead slots (back), each
// part ordered by map cell when m is given; alive is rewritten to 1...1 0...0 and the
// live count is left in *n_live_dev (device memory).
hipError_t compact_state(const MapDesc* m, double* arrs[6], int32_t* ids, uint8_t* alive, int64_t n, void* ws,
                         size_t ws_bytes, hipStream_t s, int refine_axis, int refine_bits, int64_t* n_live_dev);
int key_bits_of(const MapDesc& m);              // bits of the cell key of map m
int64_t* result_slot(void* ws, int64_t n);      // 256-byte result area inside a sort workspace of n slots
}  // namespace pp_sort

enum {
    PP_BK_BORIS = 0, PP_BK_VAY = 1, PP_BK_RK4REL = 2, PP_BK_POS = 3, PP_BK_LORENTZ = 4, PP_BK_AXPY = 5,
    PP_BK_RK4COMB = 6, PP_BK_CLAMP = 7, PP_BK_SCALEOFF = 8
};
enum { PP_UT_AOS2SOA = 0, PP_UT_SOA2AOS = 1, PP_UT_V2U = 2, PP_UT_U2V = 3 };
