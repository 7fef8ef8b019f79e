/*
 * pypat_hip.h -- C ABI of the MI355X particle push + field gather engine
 * (libpypat_hip.so).
 *
 * The reference (PyPATools) is pure Python; its only "native" functions are the
 * Numba-JIT batch kernels and the Pusher methods that drive them.  Each entry
 * point below names the reference function it replaces (file:line, relative to
 * the PyPATools checkout).
 *
 * Conventions
 *   - Every function returns 0 (PP_OK) or a PP_ERR_* code; pp_last_error()
 *     gives the message of the last failure on the calling thread.  No C++
 *     exception crosses this boundary.
 *   - Compute entry points take DEVICE pointers (allocate with pp_malloc, move
 *     data with pp_memcpy_*) and a stream handle (NULL = the library's default
 *     stream).  They enqueue work and return without synchronising unless
 *     documented as blocking.
 *   - Particle state is SoA float64: x, y, z [m] and vx, vy, vz.  For the
 *     u-state algorithms (PP_ALG_BORIS_REL, PP_ALG_VAY) vx..vz hold
 *     u = gamma*v [m/s]; convert with pp_v_to_u / pp_u_to_v.
 *   - Point and field arrays at the gather boundary are float64 (M, 3) C-order,
 *     as Field.__call__ takes and returns them (field.py:213-244).
 *   - The library owns pp_map objects and frees them in pp_map_destroy.
 */
#ifndef PYPAT_HIP_H
#define PYPAT_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PP_ABI_VERSION 3

enum pp_status {
    PP_OK = 0,
    PP_ERR_INVALID = 1,      /* bad argument (ValueError on the Python side) */
    PP_ERR_HIP = 2,          /* HIP runtime failure */
    PP_ERR_NOMEM = 3,        /* device allocation refused */
    PP_ERR_UNSUPPORTED = 4   /* valid request this build does not implement */
};

/* Integrators, Pusher.ALGORITHMS (pusher.py:465-466) + two extensions. */
enum pp_algorithm {
    PP_ALG_LEAPFROG = 0,     /* pusher.py:113-121, 730-745 */
    PP_ALG_BORIS = 1,        /* pusher.py:124-165, 369-379 (non-relativistic) */
    PP_ALG_RK4 = 2,          /* pusher.py:803-892 */
    PP_ALG_YOSHIDA = 3,      /* pusher.py:184-209 */
    PP_ALG_VAY_REL = 4,      /* pusher.py:216-279, 382-392 */
    PP_ALG_RK4_REL = 5,      /* pusher.py:282-329, 803-892 */
    PP_ALG_YOSHIDA_REL = 6,  /* pusher.py:332-362 */
    PP_ALG_BORIS_REL = 7,    /* extension: relativistic Boris on u = gamma v */
    PP_ALG_VAY = 8           /* extension: Vay, Phys. Plasmas 15, 056701 (2008) */
};

/* Arithmetic of a launch. */
enum pp_precision {
    PP_PREC_FP64 = 0,        /* IEEE FP64 in the reference's NumPy operation order, no FMA
                                contraction: bitwise equal to the FP64 oracle */
    PP_PREC_FP64_FAST = 1,   /* FP64 with FMA contraction and precomputed cell coefficients */
    PP_PREC_FP32_MIXED = 2   /* FP32 map + FP32 increments, FP64 position/velocity accumulation */
};

enum pp_field_kind { PP_FIELD_ZERO = 0, PP_FIELD_CONST = 1, PP_FIELD_MAP = 2, PP_FIELD_ARRAY = 3 };
enum pp_rf_mode { PP_RF_NONE = 0, PP_RF_COS = 1, PP_RF_SIN = 2 };
enum pp_map_flags { PP_MAP_PERIODIC_Z = 1, PP_MAP_NO_FP32 = 2, PP_MAP_NO_COEF = 4, PP_MAP_NO_FP64 = 8 };

typedef struct pp_map pp_map;          /* device-resident field map (opaque) */
typedef void* pp_stream;               /* hipStream_t */

/*
 * One field term, the device form of a Field / ScaledField (field.py:101-690):
 *   value = ((scale1 * F) [* scale2] [+ offset]) [* rf(t)]
 * F is the constant `value` (kind CONST, already multiplied by Field.scaling),
 * the map components comp0..comp0+2 (kind MAP), or per-particle values from a
 * device (n, 3) array (kind ARRAY: fields evaluated elsewhere, e.g. by a Python
 * callable; single-step pushes of single-gather algorithms only).  rf(t) is read
 * from the launch's rf table (cos or sin column), see pp_push_args.
 */
typedef struct pp_field {
    int32_t kind;            /* enum pp_field_kind */
    int32_t comp0;           /* first map component used (MAP) */
    const pp_map* map;       /* MAP */
    const double* array;     /* ARRAY: device (n, 3) */
    double scale1;           /* Field.scaling (field.py:500-502) */
    double scale2;           /* ScaledField.scale (field.py:690), used if has_scale2 */
    double offset;           /* ScaledField.offset, used if has_offset */
    int32_t has_scale2;
    int32_t has_offset;
    int32_t rf;              /* enum pp_rf_mode */
    int32_t _pad;
    double value[3];         /* CONST */
} pp_field;

/*
 * A fused multi-step push of n particles (Pusher.track_batch's step loop,
 * pusher.py:966-990, with push_batch pusher.py:699-801 inside).
 */
typedef struct pp_push_args {
    int32_t algorithm;       /* enum pp_algorithm */
    int32_t precision;       /* enum pp_precision */
    int64_t n;               /* particles */
    double* x; double* y; double* z;
    double* vx; double* vy; double* vz;
    uint8_t* alive;          /* optional; 0 = inactive: not pushed, not recorded */
    const double* dt_pp;     /* optional per-particle dt (RK4 family, pusher.py:815-816) */
    const int32_t* ids;      /* optional: original particle index of each slot after pp_sort_state;
                                history rows are written at ids[i] */
    double dt;               /* [s] */
    double q_over_m;         /* [C/kg] */
    int64_t nsteps;          /* steps in this call */
    int64_t step0;           /* global index of the first step */
    pp_field efield;
    pp_field bfield;
    /* RF table: rf_table[(step - rf_step0) * 8 + stage * 2 + {0: cos, 1: sin}],
       stage 0..3 at t_n + (0, 1/2, 1/2, 1) dt; required if a field has rf != NONE */
    const double* rf_table;
    int64_t rf_step0;
    int64_t rf_nsteps;       /* rows in rf_table (bounds check) */
    /* history (optional): device AoS (n_records, n, 3); record k holds the state
       after global step k*rec_every - 1, k = 1 .. n_records-1 (pusher.py:975-980) */
    double* r_hist;
    double* v_hist;
    int64_t rec_every;
    int64_t n_records;
    int32_t kick_only;       /* 1: velocity-only kick with this dt (Boris staggering,
                                pusher.py:946-958, 992-1007); nsteps must be 1 */
    int32_t block_size;      /* threads per block: 64, 128, 192 or 256; 0 = library default (256) */
    int64_t hist_rows;       /* particles per history record (row stride; rows are ids[i] when ids is
                                given): 0 = n.  Set it to the original count when n covers a
                                compacted live prefix (pp_compact_state) */
} pp_push_args;

/* ---- runtime ------------------------------------------------------------ */
int pp_init(int device);                                   /* select device, create default stream */
const char* pp_last_error(void);
int pp_abi_version(void);
/* sha256 (hex) of the fast-path kernel sources the library was built from (no counterpart in the
   reference): profiler counters recorded with it are matched to the binary that runs. */
const char* pp_kernel_digest(void);
/* The constants the kernels use, as the reference's Python expressions give them:
   out[0] CLIGHT, out[1] CLIGHT**2 (pusher.py:90), out[2] 0.9999*CLIGHT (:276),
   out[3..5] YOSHIDA_C1..C3, out[6..8] YOSHIDA_D1..D3 (:176-181).  n >= 9. */
int pp_get_constants(double* out, int32_t n);
/* sizeof(pp_field), sizeof(pp_push_args) as compiled (binding layout check). */
int pp_struct_sizes(int64_t* field_size, int64_t* push_args_size);
/* sizeof(pp_bunch), sizeof(pp_batch_args). */
int pp_batch_struct_sizes(int64_t* bunch_size, int64_t* batch_args_size);
int pp_device_count(int* count);
int pp_mem_info(size_t* free_bytes, size_t* total_bytes);  /* hipMemGetInfo */
int pp_malloc(void** dptr, size_t bytes);
int pp_free(void* dptr);
int pp_memcpy_h2d(void* dst, const void* src, size_t bytes, pp_stream s);
int pp_memcpy_d2h(void* dst, const void* src, size_t bytes, pp_stream s);
int pp_memcpy_d2d(void* dst, const void* src, size_t bytes, pp_stream s);
int pp_fill_f64(double* dptr, double value, int64_t count, pp_stream s);   /* kernel fill, graph-safe */
int pp_stream_create(pp_stream* s);
int pp_stream_destroy(pp_stream s);
int pp_stream_sync(pp_stream s);
int pp_device_sync(void);
/* events for timing: pp_event is a hipEvent_t */
typedef void* pp_event;
int pp_event_create(pp_event* e);
int pp_event_destroy(pp_event e);
int pp_event_record(pp_event e, pp_stream s);
int pp_event_elapsed_ms(float* ms, pp_event start, pp_event stop);   /* blocking on stop */
/* Asynchronous transfers for the serving loop (no counterpart in the reference, whose
   pushers run on the host).  pp_host_alloc returns page-locked host memory
   (hipHostMalloc); the *_async copies enqueue on the stream and return at once (host
   buffers must be page-locked and stay untouched until an event recorded after the copy
   has completed); pp_event_query sets *done = 1 once the event has completed, without
   blocking; pp_event_sync blocks until it has. */
int pp_host_alloc(void** hptr, size_t bytes);
int pp_host_free(void* hptr);
int pp_memcpy_h2d_async(void* dst, const void* src, size_t bytes, pp_stream s);
int pp_memcpy_d2h_async(void* dst, const void* src, size_t bytes, pp_stream s);
int pp_event_query(pp_event e, int32_t* done);
int pp_event_sync(pp_event e);

/* ---- field maps (Field.from_arrays + get_interpolator, field.py:345-387,
        interpolators.py:533-610) ------------------------------------------ */
/*
 * values: host float64 [ncomp][nx][ny][nz] (each component 'ij'-ordered, z
 * fastest, as NumbaInterpolator stores it, interpolators.py:314-315).
 * gx, gy, gz: host node coordinates, strictly increasing; an axis with one node
 * is inactive (1-D fields use the z axis, 2-D fields x and y, as
 * Field._get_field_1d/_2d evaluate them, field.py:457-489).
 * fill_value: value of every component outside the grid (bounds_error=False,
 * interpolators.py:94-97; Field uses 0.0, field.py:380-383).
 * flags: PP_MAP_PERIODIC_Z wraps z with period nz*dz (extension).  PP_MAP_NO_FP32 /
 * PP_MAP_NO_FP64 skip the FP32 (192 B per 6-component cell) / FP64 (384 B) coefficient
 * table, PP_MAP_NO_COEF both: maps for FP32-mixed pushes only need the FP32 table, and
 * launches of a precision whose table is missing fail with PP_ERR_UNSUPPORTED.
 * Blocking (uploads and builds the cell-coefficient tables).
 */
int pp_map_create(const double* values, int32_t ncomp, int32_t nx, int32_t ny, int32_t nz,
                  const double* gx, const double* gy, const double* gz, double fill_value, int32_t flags,
                  pp_map** out);
int pp_map_destroy(pp_map* map);
int pp_map_bytes(const pp_map* map, size_t* bytes);

/* Field.__call__ for one field term at M points (field.py:213-244 -> _get_field_0d/1d/2d/3d
   -> NumbaInterpolator._interp3d_batch, interpolators.py:259-267).
   pts, out: device float64 (M, 3).  rf terms are evaluated with factor rf_value. */
int pp_field_eval(const pp_field* field, int32_t precision, const double* pts, int64_t m,
                  double rf_value, double* out, pp_stream s);

/* Spline interpolation of gridded values: CoordinateMapper.__call__ with the orders of its
   methods 'nearest' 0, 'linear' 1, 'cubic' 3, 'quintic' 5 (interpolators.py:450-460, 469-510:
   np.interp to fractional node indices, then scipy.ndimage.map_coordinates), and the
   'nearest' method of the 'scipy' backend (RegularGridInterpolator, interpolators.py:585).
   values: host float64 [ncomp][n0]..[n_{ndim-1}] (last axis fastest); axis0..2: host node
   coordinates of the first ndim axes, strictly increasing (n >= 2 each for PP_SPLINE_RGI).
   flags: PP_SPLINE_NEAREST_MODE = map_coordinates mode 'nearest' (the reference's
   bounds_error=True case: callers reject points outside the grid first); otherwise mode
   'constant': points outside the grid on any axis get cval.  PP_SPLINE_RGI: order 0 with
   RegularGridInterpolator's rule (node i+1 only past the cell's midpoint, ties to i).
   Coefficients of orders 3 and 5 are computed on the device at creation (blocking).
   pp_spline_eval: pts device, point i's coordinates at pts[i*stride + col0 + a], a < ndim;
   out device [m][ncomp].  Stream-ordered. */
typedef struct pp_spline pp_spline;
enum pp_spline_flags { PP_SPLINE_NEAREST_MODE = 1, PP_SPLINE_RGI = 2 };
int pp_spline_create(const double* values, int32_t ncomp, int32_t ndim, const int64_t* shape, const double* axis0,
                     const double* axis1, const double* axis2, int32_t order, int32_t flags, double cval,
                     pp_spline** out);
int pp_spline_destroy(pp_spline* s);
int pp_spline_eval(const pp_spline* s, const double* pts, int64_t m, int32_t stride, int32_t col0, double* out,
                   pp_stream st);

/* ---- pushers -------------------------------------------------------------- */
/* Fused multi-step push: Pusher.track_batch step loop (pusher.py:966-990) /
   push_batch (pusher.py:699-801) when nsteps == 1. */
int pp_push(const pp_push_args* args, pp_stream s);

/* Per-step kernels with fields supplied per particle (the reference's Numba batch
   kernels).  v, E, B, r and outputs: device float64 (M, 3). */
int pp_boris_push_batch(const double* v, const double* E, const double* B, int64_t m,
                        double dt, double q_over_m, double* v_out, pp_stream s);          /* pusher.py:369-379 */
int pp_vay_push_batch(const double* v, const double* E, const double* B, int64_t m,
                      double dt, double q_over_m, double* v_out, pp_stream s);            /* pusher.py:382-392 */
int pp_rk4_rel_dbetagamma_dt_batch(const double* v, const double* E, const double* B, int64_t m,
                                   double q_over_m, double* out, pp_stream s);            /* pusher.py:301-329 */
int pp_position_update_batch(const double* r, const double* v, int64_t m, double dt,
                             double* r_out, pp_stream s);                                  /* pusher.py:395-404 */

/* Elementwise stage kernels on (M,3) arrays, one IEEE operation per NumPy
   operation of _rk4_step_batch (pusher.py:818-890) and the field algebra
   (field.py:641-653, 689-690).  s_pp / w_pp: optional per-particle factors
   replacing the scalar. */
int pp_lorentz_batch(const double* v, const double* E, const double* B, int64_t m, double q_over_m,
                     double* out, pp_stream s);                         /* q/m (E + v x B), pusher.py:828-829 */
int pp_axpy_batch(const double* a, double s_, const double* s_pp, const double* b, int64_t m,
                  double* out, pp_stream s);                            /* a + s*b */
int pp_rk4_combine_batch(const double* a, double w, const double* w_pp, const double* k1, const double* k2,
                         const double* k3, const double* k4, int64_t m, double* out, pp_stream s);
                                                                        /* a + w*(((k1 + 2k2) + 2k3) + k4) */
int pp_clamp_speed_batch(const double* v, int64_t m, double* out, pp_stream s);   /* pusher.py:886-890 */
int pp_scale_offset_batch(const double* a, double scale, int32_t has_scale, double offset, int32_t has_offset,
                          int64_t m, double* out, pp_stream s);          /* scale*a + offset, field.py:690 */

/* ---- batched push of many independent bunches (extension: the reference
        tracks one particle array per call) ------------------------------------ */
/* Staggering flags of a bunch (pp_bunch.flags), for the staggered algorithms BORIS,
   BORIS_REL and VAY: Pusher.track_batch's half-step velocity kicks (pusher.py:946-958,
   992-1007).  KICK_START: in the launch whose window starts at the bunch's step 0,
   kick the velocity by -dt/2 before the first step.  KICK_END: in the launch whose
   window reaches the bunch's last step, kick it by +dt/2 after that step.  A bunch
   with both flags ends with the velocity track_batch returns in v_hist[-1]. */
enum pp_bunch_flags { PP_BUNCH_KICK_START = 1, PP_BUNCH_KICK_END = 2 };

/* One bunch: particles [offset, offset+count) of the shared SoA state. */
typedef struct pp_bunch {
    int64_t offset;
    int64_t count;
    int32_t algorithm;       /* BORIS, VAY, VAY_REL, BORIS_REL, RK4, RK4_REL or LEAPFROG */
    int32_t nsteps;          /* total steps of this bunch */
    int32_t map_index;       /* which of the launch's maps (0..nmaps-1) */
    int32_t flags;           /* enum pp_bunch_flags (ignored by the other algorithms) */
    double dt;
    double q_over_m;
    double escale;           /* multiplies the map's E components (0..2) */
    double bscale;           /* multiplies the map's B components (3..5) */
} pp_bunch;

#define PP_MAX_BATCH_MAPS 4

typedef struct pp_batch_args {
    int32_t precision;       /* PP_PREC_FP64_FAST or PP_PREC_FP32_MIXED */
    int32_t block_size;      /* = tile size used by pp_batch_tiles */
    int64_t n;               /* particles in the state */
    double* x; double* y; double* z;
    double* vx; double* vy; double* vz;  /* v, or u = gamma v for BORIS_REL / VAY bunches */
    const pp_bunch* bunches; /* device array */
    int32_t nbunch;
    int32_t nmaps;
    const int32_t* tile_bunch;   /* device, from pp_batch_tiles */
    const int64_t* tile_start;   /* device */
    int64_t ntiles;
    int64_t step0;           /* steps [step0, step0+nsteps) of every bunch (a bunch stops at its nsteps) */
    int64_t nsteps;
    const int64_t* step0_dev;  /* optional device scalar read instead of step0 (graph replays) */
    const pp_map* maps[PP_MAX_BATCH_MAPS];   /* 6-component (E then B) uniform 3-D maps */
    int32_t tile_class;      /* enum pp_tile_class: which algorithms the launch's kernel holds; tiles of
                                bunches outside the class are skipped (their state is left as it is) */
    int32_t _pad;
} pp_batch_args;
/* PP_TILES_ALL: one kernel for every algorithm (allocated for 4 wavefronts per SIMD);
   PP_TILES_SINGLE_GATHER: boris, vay, vay_rel, boris_rel, leapfrog (7 wavefronts per SIMD in
   FP32-mixed, 6 in FP64-fast);
   PP_TILES_RK4: rk4, rk4_rel (4 wavefronts per SIMD).  Callers with both kinds of bunches pass
   each launch the tiles of one class. */
enum pp_tile_class { PP_TILES_ALL = 0, PP_TILES_SINGLE_GATHER = 1, PP_TILES_RK4 = 2 };

/* Host helper: split the bunches into tiles of `tile` particles that never
   straddle two bunches.  tile_bunch/tile_start: host arrays with room for
   sum(ceil(count/tile)) entries (query with NULL outputs). */
int pp_batch_tiles(const pp_bunch* bunches, int32_t nbunch, int32_t tile, int32_t* tile_bunch,
                   int64_t* tile_start, int64_t* ntiles);
/* One launch advancing every bunch over [step0, step0+nsteps) (fused, state in registers). */
int pp_push_batched(const pp_batch_args* a, pp_stream s);

/* ---- HIP graphs: capture a stream's launches once, replay them ---------- */
typedef void* pp_graph;
int pp_graph_begin(pp_stream s);                    /* stream capture (the stream must not be NULL) */
int pp_graph_end(pp_stream s, pp_graph* g);         /* end capture, instantiate */
int pp_graph_launch(pp_graph g, pp_stream s);
int pp_graph_destroy(pp_graph g);
int pp_add_i64(int64_t* dptr, int64_t value, pp_stream s);   /* *dptr += value on the device (kernel) */

/* Cell sort (no reference counterpart: the reference's loops have no memory
   locality to manage).  Reorders the SoA state of `a` (x..vz) by the map cell of
   each particle -- efield's map, else bfield's -- and permutes `ids` alongside,
   so a wavefront's gathers hit the same cache lines.  pp_unsort_state restores
   the original order (ids -> 0..n-1), of args->alive too when it is given.
   Workspace: pp_sort_workspace_bytes(n).
   No-op when neither field is a map. */
int pp_sort_workspace_bytes(int64_t n, size_t* bytes);
int pp_sort_state(const pp_push_args* a, int32_t* ids, void* workspace, size_t ws_bytes, pp_stream s);
/* Same, for a batched state: the key is (group[i] << group_shift) | cell, so each
   bunch keeps its slot range; the map is `map`. */
int pp_sort_state_grouped(const pp_push_args* a, const pp_map* map, const int32_t* group, int32_t group_shift,
                          int32_t* ids, void* workspace, size_t ws_bytes, pp_stream s);
/* General form: key = ((cell << refine_bits) | sub-cell bin along refine_axis) | (group << group_shift).
   Sub-cell bins along the direction of motion keep a wavefront's particles in one
   cell for more steps.  map: NULL = efield's map, else bfield's; group may be NULL.
   The key must fit 32 bits. */
int pp_sort_state_ex(const pp_push_args* a, const pp_map* map, const int32_t* group, int32_t group_shift,
                     int32_t refine_axis, int32_t refine_bits, int32_t* ids, void* workspace, size_t ws_bytes,
                     pp_stream s);
int pp_unsort_state(const pp_push_args* a, int32_t* ids, void* workspace, size_t ws_bytes, pp_stream s);
int pp_iota_i32(int32_t* ids, int64_t n, pp_stream s);

/* AoS (M,3) <-> SoA and v <-> u conversions (state staging for pp_push). */
int pp_aos_to_soa(const double* aos, int64_t m, double* a, double* b, double* c, pp_stream s);
int pp_soa_to_aos(const double* a, const double* b, const double* c, int64_t m, double* aos, pp_stream s);
int pp_v_to_u(double* vx, double* vy, double* vz, int64_t m, pp_stream s);   /* u = gamma(v) v, pusher.py:74-96 clamp */
int pp_u_to_v(double* ux, double* uy, double* uz, int64_t m, pp_stream s);   /* v = u / sqrt(1 + u^2/c^2) */

/* Beam moments of a device-resident state (ParticleDistribution, particles.py:693-843).
   state: n, x..vz and optionally alive (other fields unused); u_state = 1 when the
   velocity slots hold u = gamma v.  Momenta are beta*gamma = gamma v / c (physical,
   not the reference's per-component conversion, particles.py:72-77); angles are
   v / v_z with |v_z| < 1e-10 replaced by 1e-10 (particles.py:463-467); kinetic energy
   is m (gamma - 1) from |p| (or p_z when z_energy), mass in MeV.
   sums_dev[PP_MOMENT_SUMS]: n, sum x,y,z, sum px,py,pz, sum |p|, sum E, sum x',y',z',
   min x,y,z, max x,y,z.  central_dev[PP_MOMENT_CENTRAL]: sums of (x-<x>)^2 (3),
   (x'-<x'>)^2 (3), then (x-<x>)(x'-<x'>), (y-<y>)(y'-<y'>), (x-<x>)(y'-<y'>),
   (y-<y>)(x'-<x'>), (E-<E>)^2.  Stream-ordered; workspace from
   pp_beam_workspace_bytes. */
#define PP_MOMENT_SUMS 18
#define PP_MOMENT_CENTRAL 11
int pp_beam_workspace_bytes(size_t* bytes);
int pp_beam_moments(const pp_push_args* state, int32_t u_state, double mass_mev, int32_t z_energy, double* sums_dev,
                    double* central_dev, void* workspace, size_t ws_bytes, pp_stream s);
/* Particles inside the 4-rms ellipses gamma x^2 + 2 alpha x x' + beta x'^2 < 4 eps, for x and y
   (particles.py:821-835).  twiss_dev[8] = alpha_x, beta_x, gamma_x, alpha_y, beta_y, gamma_y,
   eps_x, eps_y; counts_dev[2]. */
int pp_beam_ellipse_counts(const pp_push_args* state, int32_t u_state, const double* twiss_dev, double* counts_dev,
                           void* workspace, size_t ws_bytes, pp_stream s);
/* pp_beam_moments without a workspace and without any allocation.  No counterpart in the
   reference: it is for callers that cannot allocate device memory (code under HIP graph
   capture, tight serving loops, harnesses that forbid hipMalloc).  The block partials go to a
   fixed device scratch of pp_beam_noalloc_scratch_bytes() (147456 B: 1024 blocks x 18
   doubles) that each device keeps from pp_init until the process ends; any particle count
   fits in it.  Same kernels and results as pp_beam_moments.  pp_init must have been called
   (PP_ERR_INVALID otherwise).  Every call on one device shares that device's scratch and
   leaves its contents behind: calls that overlap on one device (two streams, two threads)
   race and must be ordered by the caller. */
int pp_beam_moments_noalloc(const pp_push_args* state, int32_t u_state, double mass_mev, int32_t z_energy,
                            double* sums_dev, double* central_dev, pp_stream s);
int pp_beam_noalloc_scratch_bytes(size_t* bytes);

/* Electrode surfaces as a triangle mesh for particle loss (replaces the
   electrode_assembly.segment_intersects_surface call of Pusher.track_batch,
   pusher.py:982-990; the reference's py_electrodes meshes are an external package).
   pp_mesh_create copies host vertices (nv, 3) and vertex-index triangles (nt, 3) to
   the device with a bounding-volume hierarchy.  pp_collide_segments tests the
   segments (x0,y0,z0) -> (x1,y1,z1) of slots i < n: a segment that meets a triangle
   (closed: edges and end points count; a segment in the triangle's plane does not)
   sets hit[i] = 1 and alive[i] = 0.  Slots with alive[i] == 0 are skipped (hit 0).
   alive and hit may each be NULL.  Stream-ordered. */
typedef struct pp_mesh pp_mesh;
int pp_mesh_create(const double* vertices, int64_t nv, const int32_t* triangles, int64_t nt, pp_mesh** out);
int pp_mesh_destroy(pp_mesh* m);
int pp_mesh_info(const pp_mesh* m, int64_t* ntriangles, int64_t* nnodes);
int pp_collide_segments(const pp_mesh* m, const double* x0, const double* y0, const double* z0, const double* x1,
                        const double* y1, const double* z1, int64_t n, uint8_t* alive, uint8_t* hit, pp_stream s);

/* Fused push with the loss test inside the step loop: Pusher.track_batch with an
   electrode assembly (pusher.py:966-990, push -> record -> segment test -> active
   update).  Each live slot (args->alive[i] != 0, required) takes up to nsteps steps;
   after every step, and after its history record, the segment r_old -> r_new is
   tested against the mesh as in pp_collide_segments, and a hit clears alive[i] and
   ends that slot's steps (its state is the post-step state, as in the reference).
   Same arguments and results as pp_push otherwise; never takes the hot kernels. */
int pp_push_lossy(const pp_push_args* args, const pp_mesh* mesh, pp_stream s);

/* Alive-mask compaction (the live subset r_current[active] of pusher.py:972-973,
   kept resident): a stable partition of slots [0, args->n) into live slots first and
   dead slots after, each part ordered by map cell (with sub-cell bins as in
   pp_sort_state_ex) when `map`, or a map field of args, is given.  The six state
   arrays and ids are permuted together; args->alive (required) is rewritten to
   1 on [0, *n_live) and 0 after; *n_live (host) receives the live count, and the
   call waits for the stream.  Later pushes may then run over the first *n_live
   slots only; pp_unsort_state over all args->n slots restores the original order.
   Workspace: pp_sort_workspace_bytes(args->n). */
int pp_compact_state(const pp_push_args* args, const pp_map* map, int32_t refine_axis, int32_t refine_bits,
                     int32_t* ids, int64_t* n_live, void* workspace, size_t ws_bytes, pp_stream s);

/* Space charge on a box mesh without electrodes (PyAMGPoissonSolver.solve,
   poisson_amg.py:616-797): nx*ny*nz cell-centre nodes spanning the extent
   (Lx, Ly, Lz) centred on the origin, spacing h = L / n, phi = 0 one spacing
   outside the box.  Device arrays are (nx, ny, nz) float64, z fastest.
   pp_cic_deposit adds the cloud-in-cell charge density of particles (positions
   x, y, z; charge q[i], or q_const when q is NULL; slots with alive[i] == 0
   skipped) to rho [C/m^3] (poisson_amg.py:849-911; rho is not cleared).
   pp_poisson_solve: phi with -lap(phi) = rho / eps0 for the reference's 7-point
   operator (poisson_amg.py:293-444 without conductors), solved directly by sine
   transforms.  pp_efield_from_potential: E = -grad(phi) by central differences,
   0 on the outer node layer (poisson_amg.py:98-130).  All stream-ordered. */
typedef struct pp_poisson pp_poisson;
int pp_poisson_create(int32_t nx, int32_t ny, int32_t nz, double lx, double ly, double lz, pp_poisson** out);
int pp_poisson_destroy(pp_poisson* p);
int pp_cic_deposit(const pp_poisson* p, const double* x, const double* y, const double* z, const double* q,
                   double q_const, const uint8_t* alive, int64_t n, double* rho, pp_stream s);
int pp_poisson_solve(const pp_poisson* p, const double* rho, double* phi, pp_stream s);
int pp_efield_from_potential(const pp_poisson* p, const double* phi, double* ex, double* ey, double* ez,
                             pp_stream s);

/* Electrode-bounded domains (PyAMGPoissonSolver with an electrode assembly,
   poisson_amg.py:241-444, 616-735).  pp_poisson_set_conductors takes host
   vertices (nv, 3) and vertex-index triangles (nt, 3) of closed surfaces (solid,
   grounded electrodes; nt == 0 removes them).  Along the six axis rays of every
   node it counts surface crossings and measures the distance to the nearest one,
   which stands in for py_electrodes' compute_axis_aligned_surface_intersections:
   projected edge functions with a top-left rule count each crossing of a closed
   surface once.  Nodes with an odd count on any ray are CONDUCTOR (2); others with
   a distance at most the axis spacing (rounded to float32) are BOUNDARY (1), the
   rest INTERIOR (0).  Non-conductor nodes get the reference's Shortley-Weller
   rows (arm max(distance, 1e-3 h) towards a conductor neighbour).  From then on
   pp_poisson_solve solves that system with phi = 0 on conductors, by BiCGSTAB
   preconditioned with the box solve, to a relative residual rtol (default 1e-10,
   at most 500 iterations; pp_poisson_set_tolerance); it synchronises the stream
   once per iteration.  pp_poisson_conductor_data copies the cell types (N),
   distances (N, 6: +x, -x, +y, -y, +z, -z; inf without a crossing) and crossing
   counts (N, 6) to host arrays (each may be NULL).  pp_poisson_solve_stats
   reports the iterations and relative residual of the last solve. */
int pp_poisson_set_conductors(pp_poisson* p, const double* vertices, int64_t nv, const int32_t* triangles,
                              int64_t nt);
/* The same from ray data a caller computed (the arrays py_electrodes'
   compute_axis_aligned_surface_intersections returns, poisson_amg.py:250-251):
   host distances (N, 6) and crossing counts (N, 6), node order (nx, ny, nz). */
int pp_poisson_set_ray_data(pp_poisson* p, const double* distances, const int32_t* counts);
int pp_poisson_conductor_data(const pp_poisson* p, int32_t* cell_type, double* distances, int32_t* counts);
int pp_poisson_set_tolerance(pp_poisson* p, double rtol, int32_t max_iterations);
int pp_poisson_solve_stats(const pp_poisson* p, int32_t* iterations, double* relative_residual);

/* Gather-bandwidth probe (measurement utility; no counterpart in the reference).
   nthreads threads each read `iters` cells of 192 B from the FP32 coefficient table
   of `map` through the same buffer-load path as the FP32 hot kernels and fold them
   into out[thread].  mode 0: the cell is uniform per 64-lane wavefront; mode 1: each
   lane reads its own cell.  Cells follow a hashed sequence over the first `ncell`
   cells (a power of two, at most the map's cell count), so the footprint sets the
   cache level that serves the reads.  Bytes read = nthreads * iters * 192. */
int pp_gather_probe(const pp_map* map, int32_t mode, uint32_t ncell, int32_t iters, int64_t nthreads, float* out,
                    pp_stream s);

#ifdef __cplusplus
}
#endif
#endif /* PYPAT_HIP_H */
