// Host runtime of libpypat_hip.so: the C ABI declared in include/pypat_hip.h.
// Validation happens here, on the host, before any launch: every kernel can then
// assume its indices and table rows are in range.
#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <string.h>

#include <mutex>
#include <vector>

#include "pp_internal.h"

namespace {

thread_local char g_err[512] = "";
std::mutex g_mu;
int g_device = -1;
hipStream_t g_stream = nullptr;

// Constants from the reference's Python expressions (global_variables.py:36,
// pusher.py:90, :276, :176-181), evaluated identically in IEEE double.
const double kClight = 299792458.0;
// YOSHIDA_C1..C3, D1..D3 as the Python expressions of pusher.py:176-181 evaluate
// them (1/(2 - 2^(1/3)), -2^(1/3)/(2 - 2^(1/3)), ...), written out exactly.
const double kYoshidaC[3] = {0x1.59e8b6eb96339p+0, -0x1.b3d16dd72c671p+0, 0x1.59e8b6eb96339p+0};
const double kYoshidaD[3] = {0x1.59e8b6eb96339p+1, -0x1.86dd1261614d6p+1, 0x1.59e8b6eb96339p+1};

int fail(int code, const char* fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
    return code;
}

int hip_fail(hipError_t e, const char* what) {
    return fail(e == hipErrorOutOfMemory ? PP_ERR_NOMEM : PP_ERR_HIP, "%s: %s", what, hipGetErrorString(e));
}

#define HIP_TRY(expr, what)                        \
    do {                                           \
        hipError_t _e = (expr);                    \
        if (_e != hipSuccess) return hip_fail(_e, what); \
    } while (0)

hipStream_t stream_of(pp_stream s) { return s ? static_cast<hipStream_t>(s) : g_stream; }

int ensure_init() {
    if (g_device >= 0) return PP_OK;
    return pp_init(0);
}

}  // namespace

extern "C" {

int pp_abi_version(void) { return PP_ABI_VERSION; }

#ifndef PP_KERNEL_DIGEST
#define PP_KERNEL_DIGEST "unknown"
#endif
const char* pp_kernel_digest(void) { return PP_KERNEL_DIGEST; }

int pp_struct_sizes(int64_t* field_size, int64_t* push_args_size) {
    if (!field_size || !push_args_size) return fail(PP_ERR_INVALID, "null argument");
    *field_size = (int64_t)sizeof(pp_field);
    *push_args_size = (int64_t)sizeof(pp_push_args);
    return PP_OK;
}

int pp_batch_struct_sizes(int64_t* bunch_size, int64_t* batch_args_size) {
    if (!bunch_size || !batch_args_size) return fail(PP_ERR_INVALID, "null argument");
    *bunch_size = (int64_t)sizeof(pp_bunch);
    *batch_args_size = (int64_t)sizeof(pp_batch_args);
    return PP_OK;
}

int pp_get_constants(double* out, int32_t n) {
    if (!out || n < 9) return fail(PP_ERR_INVALID, "need room for 9 constants");
    out[0] = kClight;
    out[1] = kClight * kClight;
    out[2] = 0.9999 * kClight;
    for (int k = 0; k < 3; ++k) {
        out[3 + k] = kYoshidaC[k];
        out[6 + k] = kYoshidaD[k];
    }
    return PP_OK;
}

const char* pp_last_error(void) { return g_err; }

int pp_init(int device) {
    std::lock_guard<std::mutex> lk(g_mu);
    int count = 0;
    HIP_TRY(hipGetDeviceCount(&count), "hipGetDeviceCount");
    if (device < 0 || device >= count) return fail(PP_ERR_INVALID, "device %d out of range (%d devices)", device, count);
    HIP_TRY(hipSetDevice(device), "hipSetDevice");
    if (g_device != device) {
        if (g_stream) (void)hipStreamDestroy(g_stream);
        g_stream = nullptr;
        // default priority (no priority streams on this engine)
        HIP_TRY(hipStreamCreateWithFlags(&g_stream, hipStreamNonBlocking), "hipStreamCreate");
        g_device = device;
    }
    // device address of the fixed scratch of pp_beam_moments_noalloc, resolved once per device
    HIP_TRY(pp_beam::init_noalloc_scratch(device), "hipGetSymbolAddress(scratch)");
    return PP_OK;
}

int pp_device_count(int* count) {
    if (!count) return fail(PP_ERR_INVALID, "null count");
    HIP_TRY(hipGetDeviceCount(count), "hipGetDeviceCount");
    return PP_OK;
}

int pp_mem_info(size_t* free_bytes, size_t* total_bytes) {
    int rc = ensure_init();
    if (rc) return rc;
    HIP_TRY(hipMemGetInfo(free_bytes, total_bytes), "hipMemGetInfo");
    return PP_OK;
}

int pp_malloc(void** dptr, size_t bytes) {
    int rc = ensure_init();
    if (rc) return rc;
    if (!dptr) return fail(PP_ERR_INVALID, "null output pointer");
    *dptr = nullptr;
    if (bytes == 0) bytes = 256;
    HIP_TRY(hipMalloc(dptr, bytes), "hipMalloc");
    return PP_OK;
}

int pp_free(void* dptr) {
    if (!dptr) return PP_OK;
    HIP_TRY(hipFree(dptr), "hipFree");
    return PP_OK;
}

int pp_memcpy_h2d(void* dst, const void* src, size_t bytes, pp_stream s) {
    if (!bytes) return PP_OK;
    HIP_TRY(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, stream_of(s)), "hipMemcpy H2D");
    HIP_TRY(hipStreamSynchronize(stream_of(s)), "hipStreamSynchronize");
    return PP_OK;
}

int pp_memcpy_d2h(void* dst, const void* src, size_t bytes, pp_stream s) {
    if (!bytes) return PP_OK;
    HIP_TRY(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, stream_of(s)), "hipMemcpy D2H");
    HIP_TRY(hipStreamSynchronize(stream_of(s)), "hipStreamSynchronize");
    return PP_OK;
}

int pp_memcpy_d2d(void* dst, const void* src, size_t bytes, pp_stream s) {
    if (!bytes) return PP_OK;
    HIP_TRY(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, stream_of(s)), "hipMemcpy D2D");
    return PP_OK;
}

int pp_fill_f64(double* dptr, double value, int64_t count, pp_stream s) {
    if (count <= 0) return PP_OK;
    if (!dptr) return fail(PP_ERR_INVALID, "null pointer");
    HIP_TRY(pp_strict::launch_fill(dptr, value, count, stream_of(s)), "fill kernel");
    return PP_OK;
}

int pp_stream_create(pp_stream* s) {
    int rc = ensure_init();
    if (rc) return rc;
    hipStream_t h;
    HIP_TRY(hipStreamCreateWithFlags(&h, hipStreamNonBlocking), "hipStreamCreate");
    *s = h;
    return PP_OK;
}

int pp_stream_destroy(pp_stream s) {
    if (s) HIP_TRY(hipStreamDestroy(static_cast<hipStream_t>(s)), "hipStreamDestroy");
    return PP_OK;
}

int pp_stream_sync(pp_stream s) {
    int rc = ensure_init();
    if (rc) return rc;
    HIP_TRY(hipStreamSynchronize(stream_of(s)), "hipStreamSynchronize");
    return PP_OK;
}

int pp_device_sync(void) {
    HIP_TRY(hipDeviceSynchronize(), "hipDeviceSynchronize");
    return PP_OK;
}

int pp_event_create(pp_event* e) {
    int rc = ensure_init();
    if (rc) return rc;
    hipEvent_t h;
    HIP_TRY(hipEventCreate(&h), "hipEventCreate");
    *e = h;
    return PP_OK;
}

int pp_event_destroy(pp_event e) {
    if (e) HIP_TRY(hipEventDestroy(static_cast<hipEvent_t>(e)), "hipEventDestroy");
    return PP_OK;
}

int pp_event_record(pp_event e, pp_stream s) {
    HIP_TRY(hipEventRecord(static_cast<hipEvent_t>(e), stream_of(s)), "hipEventRecord");
    return PP_OK;
}

int pp_host_alloc(void** hptr, size_t bytes) {
    int rc = ensure_init();
    if (rc) return rc;
    if (!hptr) return fail(PP_ERR_INVALID, "null output pointer");
    *hptr = nullptr;
    HIP_TRY(hipHostMalloc(hptr, bytes ? bytes : 256, hipHostMallocDefault), "hipHostMalloc");
    return PP_OK;
}

int pp_host_free(void* hptr) {
    if (!hptr) return PP_OK;
    HIP_TRY(hipHostFree(hptr), "hipHostFree");
    return PP_OK;
}

int pp_memcpy_h2d_async(void* dst, const void* src, size_t bytes, pp_stream s) {
    if (!bytes) return PP_OK;
    if (!dst || !src) return fail(PP_ERR_INVALID, "null pointer");
    HIP_TRY(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, stream_of(s)), "hipMemcpyAsync H2D");
    return PP_OK;
}

int pp_memcpy_d2h_async(void* dst, const void* src, size_t bytes, pp_stream s) {
    if (!bytes) return PP_OK;
    if (!dst || !src) return fail(PP_ERR_INVALID, "null pointer");
    HIP_TRY(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, stream_of(s)), "hipMemcpyAsync D2H");
    return PP_OK;
}

int pp_event_query(pp_event e, int32_t* done) {
    if (!e || !done) return fail(PP_ERR_INVALID, "null argument");
    const hipError_t r = hipEventQuery(static_cast<hipEvent_t>(e));
    if (r == hipSuccess) {
        *done = 1;
        return PP_OK;
    }
    if (r == hipErrorNotReady) {
        *done = 0;
        return PP_OK;
    }
    return hip_fail(r, "hipEventQuery");
}

int pp_event_sync(pp_event e) {
    if (!e) return fail(PP_ERR_INVALID, "null event");
    HIP_TRY(hipEventSynchronize(static_cast<hipEvent_t>(e)), "hipEventSynchronize");
    return PP_OK;
}

int pp_event_elapsed_ms(float* ms, pp_event start, pp_event stop) {
    HIP_TRY(hipEventSynchronize(static_cast<hipEvent_t>(stop)), "hipEventSynchronize");
    HIP_TRY(hipEventElapsedTime(ms, static_cast<hipEvent_t>(start), static_cast<hipEvent_t>(stop)),
            "hipEventElapsedTime");
    return PP_OK;
}

// ---------------------------------------------------------------------------
// Field maps
// ---------------------------------------------------------------------------

int pp_map_destroy(pp_map* m) {
    if (!m) return PP_OK;
    for (int k = 0; k < m->nalloc; ++k) (void)hipFree(m->allocs[k]);
    delete m;
    return PP_OK;
}

static int map_alloc(pp_map* m, size_t bytes, void** out) {
    void* p = nullptr;
    hipError_t e = hipMalloc(&p, bytes ? bytes : 256);
    if (e != hipSuccess) return hip_fail(e, "hipMalloc (field map)");
    m->allocs[m->nalloc++] = p;
    m->bytes += bytes;
    *out = p;
    return PP_OK;
}

int pp_map_create(const double* values, int32_t ncomp, int32_t nx, int32_t ny, int32_t nz, const double* gx,
                  const double* gy, const double* gz, double fill_value, int32_t flags, pp_map** out) {
    int rc = ensure_init();
    if (rc) return rc;
    if (!out) return fail(PP_ERR_INVALID, "null output handle");
    *out = nullptr;
    if (!values || !gx || !gy || !gz) return fail(PP_ERR_INVALID, "null map input");
    if (ncomp < 1 || ncomp > 6) return fail(PP_ERR_INVALID, "ncomp must be 1..6, got %d", ncomp);
    const int32_t n[3] = {nx, ny, nz};
    const double* g[3] = {gx, gy, gz};
    for (int a = 0; a < 3; ++a) {
        if (n[a] < 1) return fail(PP_ERR_INVALID, "axis %d has %d nodes", a, n[a]);
        for (int i = 1; i < n[a]; ++i)
            if (!(g[a][i] > g[a][i - 1]) || !isfinite(g[a][i]))
                return fail(PP_ERR_INVALID, "axis %d nodes must be finite and strictly increasing", a);
    }
    const bool periodic = (flags & PP_MAP_PERIODIC_Z) != 0;
    if (periodic && nz < 2) return fail(PP_ERR_INVALID, "periodic z needs at least 2 z nodes");
    if ((int64_t)nx * ny * nz > (int64_t)1 << 31) return fail(PP_ERR_INVALID, "map too large");

    pp_map* m = new pp_map();
    memset(m, 0, sizeof(*m));
    m->device = g_device;
    MapDesc& d = m->d;
    d.ncomp = ncomp;
    d.fill = fill_value;
    d.periodic = periodic ? 1 : 0;
    d.uniform = 1;
    std::vector<double> gext[3];
    for (int a = 0; a < 3; ++a) {
        d.n[a] = n[a];
        gext[a].assign(g[a], g[a] + n[a]);
        if (a == 2 && periodic) {
            // extra node z0 + L, L = nz * dz, dz = (z_last - z0)/(nz - 1), repeats node 0
            const double dz = (g[2][nz - 1] - g[2][0]) / (double)(nz - 1);
            d.L = (double)nz * dz;
            gext[2].push_back(g[2][0] + d.L);
        }
        d.ng[a] = (int32_t)gext[a].size();
        d.lo[a] = gext[a].front();
        d.hi[a] = gext[a].back();
        d.nc[a] = d.n[a] > 1 ? d.ng[a] - 1 : 1;
        if (d.n[a] > 1) {
            const double span = d.hi[a] - d.lo[a];
            const double h = span / (double)(d.ng[a] - 1);
            d.inv_d[a] = (double)(d.ng[a] - 1) / span;
            for (int i = 0; i < d.ng[a]; ++i) {
                const double ideal = d.lo[a] + h * i;
                if (fabs(gext[a][i] - ideal) > 0.25 * h) d.uniform = 0;
            }
        }
    }
    if (periodic) d.inv_nz = 1.0 / (double)d.nc[2];

    // raw table [comp][nx][ny][ngz] (periodic: node nz repeats node 0)
    const int64_t ngz = d.ng[2];
    const int64_t per_comp = (int64_t)nx * ny * ngz;
    std::vector<double> raw((size_t)(per_comp * ncomp));
    for (int c = 0; c < ncomp; ++c)
        for (int64_t i = 0; i < nx; ++i)
            for (int64_t j = 0; j < ny; ++j) {
                const double* src = values + ((int64_t)c * nx * ny + i * ny + j) * nz;
                double* dst = raw.data() + c * per_comp + (i * ny + j) * ngz;
                memcpy(dst, src, sizeof(double) * nz);
                if (periodic) dst[nz] = src[0];
            }
    void* p;
    hipStream_t s = g_stream;
    for (int a = 0; a < 3; ++a) {
        if ((rc = map_alloc(m, sizeof(double) * gext[a].size(), &p))) { pp_map_destroy(m); return rc; }
        d.g[a] = static_cast<const double*>(p);
        if (hipMemcpyAsync(p, gext[a].data(), sizeof(double) * gext[a].size(), hipMemcpyHostToDevice, s) != hipSuccess) {
            pp_map_destroy(m);
            return fail(PP_ERR_HIP, "map upload failed");
        }
    }
    if ((rc = map_alloc(m, sizeof(double) * raw.size(), &p))) { pp_map_destroy(m); return rc; }
    d.raw = static_cast<const double*>(p);
    if (hipMemcpyAsync(p, raw.data(), sizeof(double) * raw.size(), hipMemcpyHostToDevice, s) != hipSuccess) {
        pp_map_destroy(m);
        return fail(PP_ERR_HIP, "map upload failed");
    }
    const int64_t ncell = (int64_t)d.nc[0] * d.nc[1] * d.nc[2];
    double* c64 = nullptr;
    float* c32 = nullptr;
    if (!(flags & PP_MAP_NO_COEF) && (flags & PP_MAP_NO_FP64) && (flags & PP_MAP_NO_FP32))
        return fail(PP_ERR_INVALID, "PP_MAP_NO_FP64 with PP_MAP_NO_FP32 leaves no coefficient table: use PP_MAP_NO_COEF");
    if (!(flags & PP_MAP_NO_COEF)) {
        if (!(flags & PP_MAP_NO_FP64)) {
            if ((rc = map_alloc(m, sizeof(double) * ncell * ncomp * 8, &p))) { pp_map_destroy(m); return rc; }
            c64 = static_cast<double*>(p);
        }
        if (!(flags & PP_MAP_NO_FP32)) {
            if ((rc = map_alloc(m, sizeof(float) * ncell * ncomp * 8, &p))) { pp_map_destroy(m); return rc; }
            c32 = static_cast<float*>(p);
        }
        hipError_t e = pp_strict::launch_build_coef(d, c64, c32, s);
        if (e != hipSuccess) { pp_map_destroy(m); return hip_fail(e, "coefficient build"); }
    }
    d.coef64 = c64;
    d.coef32 = c32;
    hipError_t e = hipStreamSynchronize(s);
    if (e != hipSuccess) { pp_map_destroy(m); return hip_fail(e, "map build"); }
    *out = m;
    return PP_OK;
}

int pp_map_bytes(const pp_map* m, size_t* bytes) {
    if (!m || !bytes) return fail(PP_ERR_INVALID, "null argument");
    *bytes = m->bytes;
    return PP_OK;
}

// ---------------------------------------------------------------------------
// Field terms and pushes
// ---------------------------------------------------------------------------

static int make_field(const pp_field& f, int precision, FieldDev& out) {
    memset(&out, 0, sizeof(out));
    if (f.kind != PP_FIELD_ZERO && f.kind != PP_FIELD_CONST && f.kind != PP_FIELD_MAP && f.kind != PP_FIELD_ARRAY)
        return fail(PP_ERR_INVALID, "unknown field kind %d", f.kind);
    if (f.kind == PP_FIELD_ARRAY && !f.array) return fail(PP_ERR_INVALID, "array field without an array");
    if (f.rf < PP_RF_NONE || f.rf > PP_RF_SIN) return fail(PP_ERR_INVALID, "unknown rf mode %d", f.rf);
    out.kind = f.kind == PP_FIELD_ZERO ? PP_FIELD_CONST : f.kind;
    out.comp0 = f.comp0;
    out.has_scale2 = f.has_scale2;
    out.has_offset = f.has_offset;
    out.rf = f.rf;
    out.scale1 = f.scale1;
    out.scale2 = f.scale2;
    out.offset = f.offset;
    for (int c = 0; c < 3; ++c) out.value[c] = f.kind == PP_FIELD_ZERO ? 0.0 : f.value[c];
    out.array = f.kind == PP_FIELD_ARRAY ? f.array : nullptr;
    if (f.kind == PP_FIELD_MAP) {
        if (!f.map) return fail(PP_ERR_INVALID, "map field without a map");
        if (f.map->device != g_device) return fail(PP_ERR_INVALID, "map lives on device %d, current is %d",
                                                   f.map->device, g_device);
        if (f.comp0 < 0 || f.comp0 + 3 > f.map->d.ncomp)
            return fail(PP_ERR_INVALID, "components %d..%d outside a %d-component map", f.comp0, f.comp0 + 2,
                        f.map->d.ncomp);
        out.map = f.map->d;
        if (precision != PP_PREC_FP64) {
            if (!out.map.uniform && (out.map.n[0] > 1 || out.map.n[1] > 1 || out.map.n[2] > 1))
                return fail(PP_ERR_UNSUPPORTED, "precision %d needs a uniform grid; use PP_PREC_FP64 for "
                            "non-uniform grids", precision);
            if (precision == PP_PREC_FP64_FAST && !out.map.coef64)
                return fail(PP_ERR_UNSUPPORTED, "map was built without FP64 coefficient tables");
            if (precision == PP_PREC_FP32_MIXED && !out.map.coef32)
                return fail(PP_ERR_UNSUPPORTED, "map was built without FP32 tables");
        }
    }
    return PP_OK;
}

static void fill_hot_map(const MapDesc& m, const void* coef, uint64_t bytes, HotMap& h) {
    memset(&h, 0, sizeof(h));
    h.coef = coef;
    h.bytes = (uint32_t)bytes;
    h.nc0 = m.nc[0];
    h.nc1 = m.nc[1];
    h.nc2 = m.nc[2];
    h.id0 = m.inv_d[0];
    h.id1 = m.inv_d[1];
    h.id2 = m.inv_d[2];
    h.c0 = -m.lo[0] * m.inv_d[0];
    h.c1 = -m.lo[1] * m.inv_d[1];
    h.c2 = -m.lo[2] * m.inv_d[2];
    h.nc2d = (double)m.nc[2];
    h.nc12d = (double)m.nc[1] * (double)m.nc[2];
    if (m.nc[2] > 0) {
        h.zmagic = (uint32_t)(0xFFFFFFFFull / (uint64_t)m.nc[2]);
        h.zbias = (uint32_t)m.nc[2] * ((1u << 30) / (uint32_t)m.nc[2]);
        if ((m.nc[2] & (m.nc[2] - 1)) == 0) h.zmask = (uint32_t)m.nc[2] - 1u;
        h.zinv = 1.0 / (double)m.nc[2];
    }
    h.sE = 1.0;
    h.sB = 1.0;
    h.fill = m.fill;
    h.ncells = (uint32_t)((uint64_t)m.nc[0] * m.nc[1] * m.nc[2]);
}

// The hot kernels cover the common case: E and B from one uniform 3-D map built
// with E in components 0..2 and B in 3..5 (each optionally with an RF factor), no
// offset, every particle active, one dt, a single-gather algorithm or RK4, fast
// or mixed precision.
static bool hot_eligible(const pp_push_args* a, const PushParams& p, HotParams& hp) {
    if (a->precision == PP_PREC_FP64 || !p.fused || a->dt_pp) return false;
    if (a->algorithm == PP_ALG_YOSHIDA || a->algorithm == PP_ALG_YOSHIDA_REL) return false;
    const FieldDev& E = p.E;
    const FieldDev& B = p.B;
    const MapDesc& m = E.map;
    if ((E.rf || B.rf) && !p.rf_table) return false;
    if ((E.rf || B.rf) && a->alive) return false;   // the hot loss builds carry no RF factors
    if (E.has_offset || B.has_offset || E.comp0 != 0 || B.comp0 != 3 || m.ncomp != 6) return false;
    if (!m.uniform || m.n[0] < 2 || m.n[1] < 2 || m.n[2] < 2) return false;
    if (m.fill != 0.0) return false;        // the hot gathers read 0 outside the map
    const bool mixed = a->precision == PP_PREC_FP32_MIXED;
    const void* coef = mixed ? (const void*)m.coef32 : (const void*)m.coef64;
    if (!coef) return false;
    const uint64_t bytes = (uint64_t)m.nc[0] * m.nc[1] * m.nc[2] * 48 * (mixed ? 4 : 8);
    if (bytes >= 0x7FFFFFFFull) return false;
    memset(&hp, 0, sizeof(hp));
    hp.x = p.x; hp.y = p.y; hp.z = p.z;
    hp.vx = p.vx; hp.vy = p.vy; hp.vz = p.vz;
    hp.ids = p.ids;
    hp.n = p.n;
    hp.nsteps = p.nsteps;
    hp.step0 = p.step0;
    hp.dt = p.dt;
    hp.qm = p.qm;
    hp.c2 = p.c2;
    hp.clight = p.clight;
    hp.vclamp = p.vclamp;
    hp.r_hist = p.r_hist;
    hp.v_hist = p.v_hist;
    hp.rec_every = p.rec_every;
    hp.n_records = p.n_records;
    hp.hist_rows = p.hist_rows;
    hp.alive = p.alive;
    hp.mesh = p.mesh;
    hp.kick_only = p.kick_only;
    fill_hot_map(m, coef, bytes, hp.m);
    if (E.rf || B.rf) {
        hp.rf.table = p.rf_table;
        hp.rf.step0 = p.rf_step0;
        hp.rf.e = E.rf;
        hp.rf.b = B.rf;
    }
    hp.m.sE = E.has_scale2 ? E.scale1 * E.scale2 : E.scale1;
    hp.m.sB = B.has_scale2 ? B.scale1 * B.scale2 : B.scale1;
    return true;
}

int pp_spline_create(const double* values, int32_t ncomp, int32_t ndim, const int64_t* shape, const double* axis0,
                     const double* axis1, const double* axis2, int32_t order, int32_t flags, double cval,
                     pp_spline** out) {
    int rc = ensure_init();
    if (rc) return rc;
    if (!values || !shape || !out) return fail(PP_ERR_INVALID, "null argument");
    *out = nullptr;
    if (ncomp < 1) return fail(PP_ERR_INVALID, "ncomp must be positive");
    if (ndim < 1 || ndim > 3) return fail(PP_ERR_INVALID, "Only 1D, 2D, 3D supported. Got %dD", ndim);
    if (order != 0 && order != 1 && order != 3 && order != 5)
        return fail(PP_ERR_INVALID, "spline order must be 0, 1, 3 or 5 (got %d)", order);
    if (flags & ~(PP_SPLINE_NEAREST_MODE | PP_SPLINE_RGI)) return fail(PP_ERR_INVALID, "unknown spline flags");
    const bool rgi = (flags & PP_SPLINE_RGI) != 0;
    if (rgi && order != 0) return fail(PP_ERR_INVALID, "PP_SPLINE_RGI is the order-0 rule of RegularGridInterpolator");
    const double* axes[3] = {axis0, axis1, axis2};
    int64_t total = ncomp;
    for (int a = 0; a < ndim; ++a) {
        if (!axes[a]) return fail(PP_ERR_INVALID, "axis %d is null", a);
        if (shape[a] < 1 || shape[a] > ((int64_t)1 << 31)) return fail(PP_ERR_INVALID, "axis %d: bad node count", a);
        if (rgi && shape[a] < 2) return fail(PP_ERR_INVALID, "axis %d: RegularGridInterpolator needs 2 nodes", a);
        for (int64_t k = 1; k < shape[a]; ++k)
            if (!(axes[a][k] > axes[a][k - 1]))
                return fail(PP_ERR_INVALID, "axis %d: node coordinates must be strictly increasing", a);
        total *= shape[a];
        if (total > ((int64_t)1 << 33)) return fail(PP_ERR_INVALID, "grid too large");
    }
    hipError_t e = pp_spl::create(values, ncomp, ndim, shape, axes, order, (flags & PP_SPLINE_NEAREST_MODE) != 0, rgi,
                                  cval, g_device, out);
    if (e != hipSuccess) return hip_fail(e, "spline creation");
    return PP_OK;
}

int pp_spline_destroy(pp_spline* s) {
    pp_spl::destroy(s);
    return PP_OK;
}

int pp_spline_eval(const pp_spline* s, const double* pts, int64_t m, int32_t stride, int32_t col0, double* out,
                   pp_stream st) {
    int rc = ensure_init();
    if (rc) return rc;
    if (!s) return fail(PP_ERR_INVALID, "null spline");
    if (m < 0) return fail(PP_ERR_INVALID, "negative point count");
    if (m == 0) return PP_OK;
    if (!pts || !out) return fail(PP_ERR_INVALID, "null device pointer");
    if (col0 < 0 || stride < 1 || col0 + s->d.ndim > stride)
        return fail(PP_ERR_INVALID, "columns %d..%d do not fit a row of %d", col0, col0 + s->d.ndim - 1, stride);
    if (s->device != g_device) return fail(PP_ERR_INVALID, "spline lives on another device");
    HIP_TRY(pp_spl::eval(s, pts, m, stride, col0, out, stream_of(st)), "spline eval");
    return PP_OK;
}

int pp_field_eval(const pp_field* field, int32_t precision, const double* pts, int64_t m, double rf_value,
                  double* out, pp_stream s) {
    int rc = ensure_init();
    if (rc) return rc;
    if (!field || m < 0) return fail(PP_ERR_INVALID, "bad arguments");
    if (m == 0) return PP_OK;
    if (!pts || !out) return fail(PP_ERR_INVALID, "null device pointer");
    if (precision < PP_PREC_FP64 || precision > PP_PREC_FP32_MIXED)
        return fail(PP_ERR_INVALID, "unknown precision %d", precision);
    FieldDev f;
    if ((rc = make_field(*field, precision, f))) return rc;
    if (f.kind == PP_FIELD_ARRAY) return fail(PP_ERR_INVALID, "array fields are not evaluated by pp_field_eval");
    hipError_t e = precision == PP_PREC_FP64 ? pp_strict::launch_field_eval(f, pts, m, rf_value, out, stream_of(s))
                                             : pp_fast::launch_field_eval(f, precision, pts, m, rf_value, out,
                                                                          stream_of(s));
    if (e != hipSuccess) return hip_fail(e, "field eval launch");
    return PP_OK;
}

static int push_impl(const pp_push_args* a, const pp_mesh* mesh, pp_stream s) {
    int rc = ensure_init();
    if (rc) return rc;
    if (!a) return fail(PP_ERR_INVALID, "null args");
    if (mesh) {
        if (!a->alive) return fail(PP_ERR_INVALID, "a loss test needs the alive mask");
        if (a->kick_only) return fail(PP_ERR_INVALID, "kick_only moves no particle: no loss test");
        if (pp_collide::device_of(mesh) != g_device) return fail(PP_ERR_INVALID, "mesh lives on another device");
    }
    if (a->algorithm < PP_ALG_LEAPFROG || a->algorithm > PP_ALG_VAY)
        return fail(PP_ERR_INVALID, "unknown algorithm %d", a->algorithm);
    if (a->precision < PP_PREC_FP64 || a->precision > PP_PREC_FP32_MIXED)
        return fail(PP_ERR_INVALID, "unknown precision %d", a->precision);
    if (a->n < 0 || a->nsteps < 0 || a->step0 < 0) return fail(PP_ERR_INVALID, "negative size");
    if (a->n == 0 || a->nsteps == 0) return PP_OK;
    if (a->n > ((int64_t)1 << 40)) return fail(PP_ERR_INVALID, "too many particles");
    if (!a->x || !a->y || !a->z || !a->vx || !a->vy || !a->vz) return fail(PP_ERR_INVALID, "null state pointer");
    if (!isfinite(a->dt) || !isfinite(a->q_over_m)) return fail(PP_ERR_INVALID, "dt and q_over_m must be finite");
    const bool kick_alg = a->algorithm == PP_ALG_BORIS || a->algorithm == PP_ALG_BORIS_REL || a->algorithm == PP_ALG_VAY;
    if (a->kick_only && (!kick_alg || a->nsteps != 1))
        return fail(PP_ERR_INVALID, "kick_only needs a Boris/Vay-type algorithm and nsteps == 1");
    if (a->dt_pp && a->algorithm != PP_ALG_RK4 && a->algorithm != PP_ALG_RK4_REL)
        return fail(PP_ERR_INVALID, "per-particle dt is supported by the RK4 family only (pusher.py:815-816)");
    if (a->r_hist || a->v_hist) {
        if (!a->r_hist || !a->v_hist) return fail(PP_ERR_INVALID, "r_hist and v_hist go together");
        if (a->rec_every < 1 || a->n_records < 1) return fail(PP_ERR_INVALID, "rec_every and n_records must be >= 1");
    }
    if (a->block_size < 0 || a->block_size > 256 || (a->block_size % 64) != 0)
        return fail(PP_ERR_INVALID, "block_size must be 64, 128, 192 or 256 (or 0)");
    if (a->hist_rows != 0 && a->hist_rows < a->n)
        return fail(PP_ERR_INVALID, "hist_rows (%lld) must be 0 or at least n (%lld)", (long long)a->hist_rows,
                    (long long)a->n);
    PushParams p;
    memset(&p, 0, sizeof(p));
    if ((rc = make_field(a->efield, a->precision, p.E))) return rc;
    if ((rc = make_field(a->bfield, a->precision, p.B))) return rc;
    const bool rf = p.E.rf != PP_RF_NONE || p.B.rf != PP_RF_NONE;
    if (rf) {
        if (!a->rf_table) return fail(PP_ERR_INVALID, "time-dependent field without an rf table");
        const int64_t last = a->step0 + (a->kick_only ? 0 : a->nsteps - 1);
        if (a->step0 < a->rf_step0 || last >= a->rf_step0 + a->rf_nsteps)
            return fail(PP_ERR_INVALID, "rf table rows [%lld, %lld) do not cover steps [%lld, %lld]",
                        (long long)a->rf_step0, (long long)(a->rf_step0 + a->rf_nsteps), (long long)a->step0,
                        (long long)last);
    }
    const bool has_array = a->efield.kind == PP_FIELD_ARRAY || a->bfield.kind == PP_FIELD_ARRAY;
    if (has_array && (a->algorithm == PP_ALG_RK4 || a->algorithm == PP_ALG_RK4_REL || (a->nsteps != 1)))
        return fail(PP_ERR_INVALID, "array fields need a single-gather algorithm and nsteps == 1");
    p.fused = (a->efield.kind == PP_FIELD_MAP && a->bfield.kind == PP_FIELD_MAP && a->efield.map == a->bfield.map) ? 1 : 0;
    p.n = a->n;
    p.x = a->x; p.y = a->y; p.z = a->z;
    p.vx = a->vx; p.vy = a->vy; p.vz = a->vz;
    p.alive = a->alive;
    p.dt_pp = a->dt_pp;
    p.dt = a->dt;
    p.qm = a->q_over_m;
    p.nsteps = a->kick_only ? 1 : a->nsteps;
    p.step0 = a->step0;
    p.kick_only = a->kick_only;
    p.rf_table = a->rf_table;
    p.rf_step0 = a->rf_step0;
    p.r_hist = a->r_hist;
    p.v_hist = a->v_hist;
    p.ids = a->ids;
    p.rec_every = a->r_hist ? a->rec_every : 1;
    p.n_records = a->r_hist ? a->n_records : 0;
    p.hist_rows = a->hist_rows > 0 ? a->hist_rows : a->n;
    p.clight = kClight;
    p.c2 = kClight * kClight;     // CLIGHT ** 2 (Python float pow == c*c for this value)
    p.vclamp = 0.9999 * kClight;  // pusher.py:276
    for (int k = 0; k < 3; ++k) {
        p.yc[k] = kYoshidaC[k];
        p.yd[k] = kYoshidaD[k];
    }
    if (mesh) p.mesh = pp_collide::device_view(mesh);   // null for an empty mesh: nothing to hit
    HotParams hp;
    hipError_t e;
    if (hot_eligible(a, p, hp))
        e = pp_fast::launch_hot(a->algorithm, a->precision, p.E.map.periodic != 0, hp, a->block_size, stream_of(s));
    else if (a->precision == PP_PREC_FP64)
        e = pp_strict::launch_push(a->algorithm, p, a->block_size, stream_of(s));
    else
        e = pp_fast::launch_push(a->algorithm, a->precision, p, a->block_size, stream_of(s));
    if (e != hipSuccess) return hip_fail(e, "push launch");
    return PP_OK;
}

int pp_push(const pp_push_args* a, pp_stream s) { return push_impl(a, nullptr, s); }

int pp_push_lossy(const pp_push_args* a, const pp_mesh* mesh, pp_stream s) {
    if (!mesh) return fail(PP_ERR_INVALID, "null mesh");
    return push_impl(a, mesh, s);
}

int pp_batch_tiles(const pp_bunch* bunches, int32_t nbunch, int32_t tile, int32_t* tile_bunch, int64_t* tile_start,
                   int64_t* ntiles) {
    if (!bunches || nbunch < 0 || tile < 64 || tile > 1024 || (tile % 64) || !ntiles)
        return fail(PP_ERR_INVALID, "bad tile arguments");
    int64_t t = 0;
    for (int32_t b = 0; b < nbunch; ++b) {
        if (bunches[b].count < 0 || bunches[b].offset < 0) return fail(PP_ERR_INVALID, "bunch %d: bad range", b);
        for (int64_t o = 0; o < bunches[b].count; o += tile, ++t) {
            if (tile_bunch) tile_bunch[t] = b;
            if (tile_start) tile_start[t] = bunches[b].offset + o;
        }
    }
    *ntiles = t;
    return PP_OK;
}

int pp_push_batched(const pp_batch_args* a, pp_stream s) {
    int rc = ensure_init();
    if (rc) return rc;
    if (!a) return fail(PP_ERR_INVALID, "null args");
    if (a->precision != PP_PREC_FP64_FAST && a->precision != PP_PREC_FP32_MIXED)
        return fail(PP_ERR_INVALID, "batched pushes run in PP_PREC_FP64_FAST or PP_PREC_FP32_MIXED");
    if (a->nmaps < 1 || a->nmaps > PP_MAX_BATCH_MAPS) return fail(PP_ERR_INVALID, "nmaps must be 1..%d",
                                                                  PP_MAX_BATCH_MAPS);
    if (a->ntiles < 0 || a->nsteps < 0 || a->step0 < 0 || a->n < 0) return fail(PP_ERR_INVALID, "negative size");
    if (a->ntiles == 0 || a->nsteps == 0) return PP_OK;
    if (a->ntiles > 0x7FFFFFFF) return fail(PP_ERR_INVALID, "too many tiles");
    if (a->block_size != 64 && a->block_size != 128 && a->block_size != 256)
        return fail(PP_ERR_INVALID, "block_size must be 64, 128 or 256 (the tile size)");
    if (!a->x || !a->y || !a->z || !a->vx || !a->vy || !a->vz || !a->bunches || !a->tile_bunch || !a->tile_start)
        return fail(PP_ERR_INVALID, "null device pointer");
    BatchParams p;
    memset(&p, 0, sizeof(p));
    const bool mixed = a->precision == PP_PREC_FP32_MIXED;
    int periodic = -1;
    for (int k = 0; k < a->nmaps; ++k) {
        const pp_map* mp = a->maps[k];
        if (!mp) return fail(PP_ERR_INVALID, "map %d is null", k);
        const MapDesc& m = mp->d;
        if (mp->device != g_device) return fail(PP_ERR_INVALID, "map %d lives on another device", k);
        if (m.ncomp != 6 || !m.uniform || m.n[0] < 2 || m.n[1] < 2 || m.n[2] < 2)
            return fail(PP_ERR_UNSUPPORTED, "map %d: batched pushes need 6-component uniform 3-D maps", k);
        if (m.fill != 0.0) return fail(PP_ERR_UNSUPPORTED, "map %d: batched pushes need maps that fill 0 outside", k);
        const void* coef = mixed ? (const void*)m.coef32 : (const void*)m.coef64;
        if (!coef) return fail(PP_ERR_UNSUPPORTED, "map %d has no coefficient table for this precision", k);
        const uint64_t bytes = (uint64_t)m.nc[0] * m.nc[1] * m.nc[2] * 48 * (mixed ? 4 : 8);
        if (bytes >= 0x7FFFFFFFull) return fail(PP_ERR_UNSUPPORTED, "map %d too large for 32-bit offsets", k);
        if (periodic < 0) periodic = m.periodic;
        else if (periodic != m.periodic) return fail(PP_ERR_UNSUPPORTED, "maps must all be periodic or all not");
        fill_hot_map(m, coef, bytes, p.maps[k]);
    }
    p.x = a->x; p.y = a->y; p.z = a->z;
    p.vx = a->vx; p.vy = a->vy; p.vz = a->vz;
    p.bunches = a->bunches;
    p.tile_bunch = a->tile_bunch;
    p.tile_start = a->tile_start;
    p.ntiles = a->ntiles;
    p.n = a->n;
    p.step0 = a->step0;
    p.step0_dev = a->step0_dev;
    p.nsteps = a->nsteps;
    p.c2 = kClight * kClight;
    p.clight = kClight;
    p.vclamp = 0.9999 * kClight;
    p.nmaps = a->nmaps;
    p.nbunch = a->nbunch;
    if (a->tile_class < PP_TILES_ALL || a->tile_class > PP_TILES_RK4) return fail(PP_ERR_INVALID, "bad tile_class");
    HIP_TRY(pp_fast::launch_batched(a->precision, periodic == 1, a->tile_class, p, a->block_size, stream_of(s)),
            "batched launch");
    return PP_OK;
}

int pp_graph_begin(pp_stream s) {
    int rc = ensure_init();
    if (rc) return rc;
    if (!s) return fail(PP_ERR_INVALID, "graph capture needs a created stream");
    HIP_TRY(hipStreamBeginCapture(static_cast<hipStream_t>(s), hipStreamCaptureModeThreadLocal), "begin capture");
    return PP_OK;
}

int pp_graph_end(pp_stream s, pp_graph* g) {
    if (!s || !g) return fail(PP_ERR_INVALID, "null argument");
    hipGraph_t graph;
    HIP_TRY(hipStreamEndCapture(static_cast<hipStream_t>(s), &graph), "end capture");
    hipGraphExec_t exec;
    hipError_t e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (e != hipSuccess) return hip_fail(e, "graph instantiate");
    *g = exec;
    return PP_OK;
}

int pp_graph_launch(pp_graph g, pp_stream s) {
    if (!g) return fail(PP_ERR_INVALID, "null graph");
    HIP_TRY(hipGraphLaunch(static_cast<hipGraphExec_t>(g), stream_of(s)), "graph launch");
    return PP_OK;
}

int pp_graph_destroy(pp_graph g) {
    if (g) HIP_TRY(hipGraphExecDestroy(static_cast<hipGraphExec_t>(g)), "graph destroy");
    return PP_OK;
}

int pp_add_i64(int64_t* dptr, int64_t value, pp_stream s) {
    int rc = ensure_init();
    if (rc) return rc;
    if (!dptr) return fail(PP_ERR_INVALID, "null pointer");
    HIP_TRY(pp_strict::launch_add_i64(dptr, value, stream_of(s)), "add kernel");
    return PP_OK;
}

static const MapDesc* sort_map(const pp_push_args* a);

int pp_sort_state_ex(const pp_push_args* a, const pp_map* map, const int32_t* group, int32_t group_shift,
                     int32_t refine_axis, int32_t refine_bits, int32_t* ids, void* ws, size_t ws_bytes, pp_stream s) {
    int rc = ensure_init();
    if (rc) return rc;
    if (!a || !ids) return fail(PP_ERR_INVALID, "null argument");
    if (a->n < 0 || a->n > INT32_MAX) return fail(PP_ERR_INVALID, "sort supports 0 <= n < 2^31");
    if (a->n == 0) return PP_OK;
    if (!a->x || !a->y || !a->z || !a->vx || !a->vy || !a->vz) return fail(PP_ERR_INVALID, "null state pointer");
    const MapDesc* mp = map ? &map->d : sort_map(a);
    if (!mp) return PP_OK;
    const MapDesc& m = *mp;
    if (refine_axis < 0 || refine_axis > 2 || refine_bits < 0 || refine_bits > 8)
        return fail(PP_ERR_INVALID, "refine_axis must be 0..2 and refine_bits 0..8");
    const uint64_t ncell = (uint64_t)m.nc[0] * m.nc[1] * m.nc[2];
    const uint64_t span = ncell << refine_bits;
    if (group && (group_shift < 0 || group_shift > 31 || ((uint64_t)1 << group_shift) < span))
        return fail(PP_ERR_INVALID, "group_shift %d leaves no room for %llu keys", group_shift,
                    (unsigned long long)span);
    if (span > 0xFFFFFFFFull) return fail(PP_ERR_UNSUPPORTED, "map too large for 32-bit sort keys");
    size_t need = 0;
    HIP_TRY(pp_sort::workspace_bytes(a->n, &need), "sort workspace query");
    if (!ws || ws_bytes < need) return fail(PP_ERR_INVALID, "sort workspace of %zu bytes needed", need);
    double* arrs[6] = {a->x, a->y, a->z, a->vx, a->vy, a->vz};
    HIP_TRY(pp_sort::sort_state(m, arrs, ids, a->n, ws, ws_bytes, stream_of(s), group, group_shift, refine_axis,
                                refine_bits),
            "sort");
    return PP_OK;
}

int pp_sort_state_grouped(const pp_push_args* a, const pp_map* map, const int32_t* group, int32_t group_shift,
                          int32_t* ids, void* ws, size_t ws_bytes, pp_stream s) {
    if (!map || !group) return fail(PP_ERR_INVALID, "null argument");
    return pp_sort_state_ex(a, map, group, group_shift, 2, 0, ids, ws, ws_bytes, s);
}

static const MapDesc* sort_map(const pp_push_args* a) {
    if (a->efield.kind == PP_FIELD_MAP && a->efield.map) return &a->efield.map->d;
    if (a->bfield.kind == PP_FIELD_MAP && a->bfield.map) return &a->bfield.map->d;
    return nullptr;
}

int pp_sort_workspace_bytes(int64_t n, size_t* bytes) {
    int rc = ensure_init();
    if (rc) return rc;
    if (n < 0 || !bytes) return fail(PP_ERR_INVALID, "bad arguments");
    HIP_TRY(pp_sort::workspace_bytes(n, bytes), "sort workspace query");
    return PP_OK;
}

static int sort_common(const pp_push_args* a, int32_t* ids, void* ws, size_t ws_bytes, pp_stream s, bool unsort) {
    int rc = ensure_init();
    if (rc) return rc;
    if (!a || !ids) return fail(PP_ERR_INVALID, "null argument");
    if (a->n < 0 || a->n > INT32_MAX) return fail(PP_ERR_INVALID, "sort supports 0 <= n < 2^31");
    if (a->n == 0) return PP_OK;
    if (!a->x || !a->y || !a->z || !a->vx || !a->vy || !a->vz) return fail(PP_ERR_INVALID, "null state pointer");
    size_t need = 0;
    HIP_TRY(pp_sort::workspace_bytes(a->n, &need), "sort workspace query");
    if (!ws || ws_bytes < need) return fail(PP_ERR_INVALID, "sort workspace of %zu bytes needed", need);
    double* arrs[6] = {a->x, a->y, a->z, a->vx, a->vy, a->vz};
    if (unsort) {
        HIP_TRY(pp_sort::unsort_state(arrs, ids, a->n, ws, ws_bytes, stream_of(s), a->alive), "unsort");
        return PP_OK;
    }
    const MapDesc* m = sort_map(a);
    if (!m) return PP_OK;
    if ((uint64_t)m->nc[0] * m->nc[1] * m->nc[2] > 0xFFFFFFFFull) return fail(PP_ERR_UNSUPPORTED, "map too large");
    HIP_TRY(pp_sort::sort_state(*m, arrs, ids, a->n, ws, ws_bytes, stream_of(s)), "sort");
    return PP_OK;
}

int pp_sort_state(const pp_push_args* a, int32_t* ids, void* ws, size_t ws_bytes, pp_stream s) {
    return sort_common(a, ids, ws, ws_bytes, s, false);
}

int pp_unsort_state(const pp_push_args* a, int32_t* ids, void* ws, size_t ws_bytes, pp_stream s) {
    return sort_common(a, ids, ws, ws_bytes, s, true);
}

int pp_compact_state(const pp_push_args* a, const pp_map* map, int32_t refine_axis, int32_t refine_bits,
                     int32_t* ids, int64_t* n_live, void* ws, size_t ws_bytes, pp_stream s) {
    int rc = ensure_init();
    if (rc) return rc;
    if (!a || !ids || !n_live) return fail(PP_ERR_INVALID, "null argument");
    if (!a->alive) return fail(PP_ERR_INVALID, "compaction needs the alive mask");
    if (a->n < 0 || a->n > INT32_MAX) return fail(PP_ERR_INVALID, "compaction supports 0 <= n < 2^31");
    *n_live = 0;
    if (a->n == 0) return PP_OK;
    if (!a->x || !a->y || !a->z || !a->vx || !a->vy || !a->vz) return fail(PP_ERR_INVALID, "null state pointer");
    if (refine_axis < 0 || refine_axis > 2 || refine_bits < 0 || refine_bits > 8)
        return fail(PP_ERR_INVALID, "refine_axis must be 0..2 and refine_bits 0..8");
    const MapDesc* m = map ? &map->d : sort_map(a);
    if (m && pp_sort::key_bits_of(*m) + refine_bits > 31)
        return fail(PP_ERR_UNSUPPORTED, "map too large for a cell key with the dead flag in 32 bits");
    size_t need = 0;
    HIP_TRY(pp_sort::workspace_bytes(a->n, &need), "sort workspace query");
    if (!ws || ws_bytes < need) return fail(PP_ERR_INVALID, "sort workspace of %zu bytes needed", need);
    double* arrs[6] = {a->x, a->y, a->z, a->vx, a->vy, a->vz};
    int64_t* slot = pp_sort::result_slot(ws, a->n);
    const hipStream_t st = stream_of(s);
    HIP_TRY(pp_sort::compact_state(m, arrs, ids, a->alive, a->n, ws, ws_bytes, st, refine_axis, refine_bits, slot),
            "compaction");
    HIP_TRY(hipMemcpyAsync(n_live, slot, sizeof(int64_t), hipMemcpyDeviceToHost, st), "live count copy");
    HIP_TRY(hipStreamSynchronize(st), "compaction sync");
    return PP_OK;
}

int pp_iota_i32(int32_t* ids, int64_t n, pp_stream s) {
    int rc = ensure_init();
    if (rc) return rc;
    if (n < 0 || n > INT32_MAX) return fail(PP_ERR_INVALID, "bad size");
    if (n == 0) return PP_OK;
    if (!ids) return fail(PP_ERR_INVALID, "null pointer");
    HIP_TRY(pp_sort::iota(ids, n, stream_of(s)), "iota");
    return PP_OK;
}

// ---------------------------------------------------------------------------
// Per-step kernels and conversions
// ---------------------------------------------------------------------------

static int batch_launch(pp_strict::BatchArgs q, pp_stream s) {
    int rc = ensure_init();
    if (rc) return rc;
    if (q.m < 0) return fail(PP_ERR_INVALID, "negative size");
    if (q.m == 0) return PP_OK;
    const bool need_b = q.which != PP_BK_CLAMP && q.which != PP_BK_SCALEOFF;
    const bool need_c = q.which == PP_BK_BORIS || q.which == PP_BK_VAY || q.which == PP_BK_RK4REL ||
                        q.which == PP_BK_LORENTZ || q.which == PP_BK_RK4COMB;
    const bool need_de = q.which == PP_BK_RK4COMB;
    if (!q.a || !q.out || (need_b && !q.b) || (need_c && !q.c) || (need_de && (!q.d || !q.e)))
        return fail(PP_ERR_INVALID, "null device pointer");
    q.c2 = kClight * kClight;
    q.clight = kClight;
    q.vclamp = 0.9999 * kClight;
    hipError_t e = pp_strict::launch_batch_kernel(q, stream_of(s));
    if (e != hipSuccess) return hip_fail(e, "batch kernel launch");
    return PP_OK;
}

static pp_strict::BatchArgs bargs(int which, const double* a, const double* b, const double* c, int64_t m,
                                  double* out) {
    pp_strict::BatchArgs q;
    memset(&q, 0, sizeof(q));
    q.which = which;
    q.a = a;
    q.b = b;
    q.c = c;
    q.m = m;
    q.out = out;
    return q;
}

int pp_boris_push_batch(const double* v, const double* E, const double* B, int64_t m, double dt, double q_over_m,
                        double* v_out, pp_stream s) {
    auto q = bargs(PP_BK_BORIS, v, E, B, m, v_out);
    q.s1 = dt;
    q.qm = q_over_m;
    return batch_launch(q, s);
}

int pp_vay_push_batch(const double* v, const double* E, const double* B, int64_t m, double dt, double q_over_m,
                      double* v_out, pp_stream s) {
    auto q = bargs(PP_BK_VAY, v, E, B, m, v_out);
    q.s1 = dt;
    q.qm = q_over_m;
    return batch_launch(q, s);
}

int pp_rk4_rel_dbetagamma_dt_batch(const double* v, const double* E, const double* B, int64_t m, double q_over_m,
                                   double* out, pp_stream s) {
    auto q = bargs(PP_BK_RK4REL, v, E, B, m, out);
    q.qm = q_over_m;
    return batch_launch(q, s);
}

int pp_position_update_batch(const double* r, const double* v, int64_t m, double dt, double* r_out, pp_stream s) {
    auto q = bargs(PP_BK_POS, r, v, nullptr, m, r_out);
    q.s1 = dt;
    return batch_launch(q, s);
}

int pp_lorentz_batch(const double* v, const double* E, const double* B, int64_t m, double q_over_m, double* out,
                     pp_stream s) {
    auto q = bargs(PP_BK_LORENTZ, v, E, B, m, out);
    q.qm = q_over_m;
    return batch_launch(q, s);
}

int pp_axpy_batch(const double* a, double s_, const double* s_pp, const double* b, int64_t m, double* out,
                  pp_stream s) {
    auto q = bargs(PP_BK_AXPY, a, b, nullptr, m, out);
    q.s1 = s_;
    q.s_pp = s_pp;
    return batch_launch(q, s);
}

int pp_rk4_combine_batch(const double* a, double w, const double* w_pp, const double* k1, const double* k2,
                         const double* k3, const double* k4, int64_t m, double* out, pp_stream s) {
    auto q = bargs(PP_BK_RK4COMB, a, k1, k2, m, out);
    q.d = k3;
    q.e = k4;
    q.s1 = w;
    q.s_pp = w_pp;
    return batch_launch(q, s);
}

int pp_clamp_speed_batch(const double* v, int64_t m, double* out, pp_stream s) {
    return batch_launch(bargs(PP_BK_CLAMP, v, nullptr, nullptr, m, out), s);
}

int pp_scale_offset_batch(const double* a, double scale, int32_t has_scale, double offset, int32_t has_offset,
                          int64_t m, double* out, pp_stream s) {
    auto q = bargs(PP_BK_SCALEOFF, a, nullptr, nullptr, m, out);
    q.s1 = scale;
    q.f1 = has_scale;
    q.s2 = offset;
    q.f2 = has_offset;
    return batch_launch(q, s);
}

static int util_common(int which, double* a, double* b, double* c, double* aos, int64_t m, pp_stream s) {
    int rc = ensure_init();
    if (rc) return rc;
    if (m < 0) return fail(PP_ERR_INVALID, "negative size");
    if (m == 0) return PP_OK;
    if (!a || !b || !c || ((which == PP_UT_AOS2SOA || which == PP_UT_SOA2AOS) && !aos))
        return fail(PP_ERR_INVALID, "null device pointer");
    hipError_t e = pp_strict::launch_util(which, a, b, c, aos, m, kClight * kClight, stream_of(s));
    if (e != hipSuccess) return hip_fail(e, "conversion kernel launch");
    return PP_OK;
}

int pp_poisson_create(int32_t nx, int32_t ny, int32_t nz, double lx, double ly, double lz, pp_poisson** out) {
    int rc = ensure_init();
    if (rc) return rc;
    if (!out) return fail(PP_ERR_INVALID, "null output");
    *out = nullptr;
    if (nx < 3 || ny < 3 || nz < 3 || (int64_t)nx * ny * nz > ((int64_t)1 << 28))
        return fail(PP_ERR_INVALID, "mesh cells must be >= 3 per axis and at most 2^28 in total");
    if (!(lx > 0.0) || !(ly > 0.0) || !(lz > 0.0) || !std::isfinite(lx) || !std::isfinite(ly) || !std::isfinite(lz))
        return fail(PP_ERR_INVALID, "domain extent must be positive and finite");
    PoissonDesc d;
    memset(&d, 0, sizeof(d));
    d.n[0] = nx; d.n[1] = ny; d.n[2] = nz;
    d.L[0] = lx; d.L[1] = ly; d.L[2] = lz;
    d.h[0] = lx / nx; d.h[1] = ly / ny; d.h[2] = lz / nz;     // poisson_amg.py:181-183
    char err[256] = {0};
    rc = pp_poisson_impl::create(d, out, err, sizeof(err));
    if (rc) return fail(rc, "%s", err);
    return PP_OK;
}

int pp_poisson_destroy(pp_poisson* p) {
    pp_poisson_impl::destroy(p);
    return PP_OK;
}

static int poisson_check(const pp_poisson* p) {
    if (!p) return fail(PP_ERR_INVALID, "null solver");
    if (pp_poisson_impl::device_of(p) != g_device) return fail(PP_ERR_INVALID, "solver lives on another device");
    return PP_OK;
}

int pp_cic_deposit(const pp_poisson* p, const double* x, const double* y, const double* z, const double* q,
                   double q_const, const uint8_t* alive, int64_t n, double* rho, pp_stream s) {
    int rc = ensure_init();
    if (rc) return rc;
    if ((rc = poisson_check(p))) return rc;
    if (n < 0) return fail(PP_ERR_INVALID, "negative particle count");
    if (!rho) return fail(PP_ERR_INVALID, "null rho");
    if (n == 0) return PP_OK;
    if (!x || !y || !z) return fail(PP_ERR_INVALID, "null device pointer");
    HIP_TRY(pp_poisson_impl::deposit(p, x, y, z, q, q_const, alive, n, rho, stream_of(s)), "CIC deposit");
    return PP_OK;
}

int pp_poisson_solve(const pp_poisson* p, const double* rho, double* phi, pp_stream s) {
    int rc = ensure_init();
    if (rc) return rc;
    if ((rc = poisson_check(p))) return rc;
    if (!rho || !phi) return fail(PP_ERR_INVALID, "null device pointer");
    if (rho == phi) return fail(PP_ERR_INVALID, "rho and phi must be distinct arrays");
    HIP_TRY(pp_poisson_impl::solve(p, rho, phi, stream_of(s)), "Poisson solve");
    return PP_OK;
}

int pp_poisson_set_conductors(pp_poisson* p, const double* vertices, int64_t nv, const int32_t* triangles,
                              int64_t nt) {
    int rc = ensure_init();
    if (rc) return rc;
    if ((rc = poisson_check(p))) return rc;
    if (nv < 0 || nt < 0 || nt > ((int64_t)1 << 30)) return fail(PP_ERR_INVALID, "vertex / triangle counts out of range");
    if ((nv > 0 && !vertices) || (nt > 0 && !triangles)) return fail(PP_ERR_INVALID, "null host array");
    char err[256] = {0};
    rc = pp_poisson_impl::set_conductors(p, vertices, nv, triangles, nt, err, sizeof(err));
    if (rc) return fail(rc, "%s", err);
    return PP_OK;
}

int pp_poisson_set_ray_data(pp_poisson* p, const double* distances, const int32_t* counts) {
    int rc = ensure_init();
    if (rc) return rc;
    if ((rc = poisson_check(p))) return rc;
    if (!distances || !counts) return fail(PP_ERR_INVALID, "null host array");
    char err[256] = {0};
    rc = pp_poisson_impl::set_ray_data(p, distances, counts, err, sizeof(err));
    if (rc) return fail(rc, "%s", err);
    return PP_OK;
}

int pp_poisson_conductor_data(const pp_poisson* p, int32_t* cell_type, double* distances, int32_t* counts) {
    int rc = ensure_init();
    if (rc) return rc;
    if ((rc = poisson_check(p))) return rc;
    if (!pp_poisson_impl::has_conductors(p)) return fail(PP_ERR_INVALID, "no conductors are set");
    pp_poisson_impl::conductor_data(p, cell_type, distances, counts);
    return PP_OK;
}

int pp_poisson_set_tolerance(pp_poisson* p, double rtol, int32_t max_iterations) {
    int rc = ensure_init();
    if (rc) return rc;
    if ((rc = poisson_check(p))) return rc;
    if (!(rtol > 0.0) || !std::isfinite(rtol) || max_iterations < 1)
        return fail(PP_ERR_INVALID, "rtol must be positive and max_iterations at least 1");
    pp_poisson_impl::set_tolerance(p, rtol, max_iterations);
    return PP_OK;
}

int pp_poisson_solve_stats(const pp_poisson* p, int32_t* iterations, double* relative_residual) {
    int rc = ensure_init();
    if (rc) return rc;
    if ((rc = poisson_check(p))) return rc;
    pp_poisson_impl::stats(p, iterations, relative_residual);
    return PP_OK;
}

int pp_efield_from_potential(const pp_poisson* p, const double* phi, double* ex, double* ey, double* ez,
                             pp_stream s) {
    int rc = ensure_init();
    if (rc) return rc;
    if ((rc = poisson_check(p))) return rc;
    if (!phi || !ex || !ey || !ez) return fail(PP_ERR_INVALID, "null device pointer");
    HIP_TRY(pp_poisson_impl::gradient(p, phi, ex, ey, ez, stream_of(s)), "field from potential");
    return PP_OK;
}

int pp_mesh_create(const double* vertices, int64_t nv, const int32_t* triangles, int64_t nt, pp_mesh** out) {
    int rc = ensure_init();
    if (rc) return rc;
    if (!out) return fail(PP_ERR_INVALID, "null output");
    *out = nullptr;
    if (nv < 0 || nt < 0 || nt > ((int64_t)1 << 30) || nv > ((int64_t)1 << 31) - 1)
        return fail(PP_ERR_INVALID, "vertex / triangle counts out of range");
    if ((nv > 0 && !vertices) || (nt > 0 && !triangles)) return fail(PP_ERR_INVALID, "null host array");
    char err[256] = {0};
    rc = pp_collide::create(vertices, nv, triangles, nt, out, err, sizeof(err));
    if (rc) return fail(rc, "%s", err);
    return PP_OK;
}

int pp_mesh_destroy(pp_mesh* m) {
    pp_collide::destroy(m);
    return PP_OK;
}

int pp_mesh_info(const pp_mesh* m, int64_t* ntriangles, int64_t* nnodes) {
    if (!m) return fail(PP_ERR_INVALID, "null mesh");
    pp_collide::info(m, ntriangles, nnodes);
    return PP_OK;
}

int pp_collide_segments(const pp_mesh* m, const double* x0, const double* y0, const double* z0, const double* x1,
                        const double* y1, const double* z1, int64_t n, uint8_t* alive, uint8_t* hit, pp_stream s) {
    int rc = ensure_init();
    if (rc) return rc;
    if (!m) return fail(PP_ERR_INVALID, "null mesh");
    if (pp_collide::device_of(m) != g_device) return fail(PP_ERR_INVALID, "mesh lives on another device");
    if (n < 0) return fail(PP_ERR_INVALID, "negative segment count");
    if (n == 0) return PP_OK;
    if (!x0 || !y0 || !z0 || !x1 || !y1 || !z1) return fail(PP_ERR_INVALID, "null device pointer");
    HIP_TRY(pp_collide::collide(m, x0, y0, z0, x1, y1, z1, n, alive, hit, stream_of(s)), "collision kernel");
    return PP_OK;
}

static int moment_args(const pp_push_args* a, int32_t u_state, double mass_mev, int32_t z_energy, MomentArgs& m) {
    if (!a) return fail(PP_ERR_INVALID, "null state");
    if (a->n < 0) return fail(PP_ERR_INVALID, "negative particle count");
    if (a->n > 0 && (!a->x || !a->y || !a->z || !a->vx || !a->vy || !a->vz))
        return fail(PP_ERR_INVALID, "null device pointer");
    if (!(mass_mev > 0.0)) return fail(PP_ERR_INVALID, "mass_mev must be positive");
    memset(&m, 0, sizeof(m));
    m.x = a->x; m.y = a->y; m.z = a->z;
    m.vx = a->vx; m.vy = a->vy; m.vz = a->vz;
    m.alive = a->alive;
    m.n = a->n;
    m.u_state = u_state != 0;
    m.z_energy = z_energy != 0;
    m.mass_mev = mass_mev;
    m.clight = kClight;
    m.beta2_max = 0.9999;
    return PP_OK;
}

int pp_beam_workspace_bytes(size_t* bytes) {
    if (!bytes) return fail(PP_ERR_INVALID, "null argument");
    *bytes = pp_beam::workspace_bytes();
    return PP_OK;
}

int pp_beam_moments(const pp_push_args* state, int32_t u_state, double mass_mev, int32_t z_energy, double* sums_dev,
                    double* central_dev, void* workspace, size_t ws_bytes, pp_stream s) {
    int rc = ensure_init();
    if (rc) return rc;
    MomentArgs m;
    if ((rc = moment_args(state, u_state, mass_mev, z_energy, m))) return rc;
    if (!sums_dev || !central_dev || !workspace) return fail(PP_ERR_INVALID, "null output or workspace");
    if (ws_bytes < pp_beam::workspace_bytes()) return fail(PP_ERR_INVALID, "workspace too small");
    HIP_TRY(pp_beam::moments(m, sums_dev, central_dev, workspace, stream_of(s)), "beam moment kernels");
    return PP_OK;
}

int pp_beam_moments_noalloc(const pp_push_args* state, int32_t u_state, double mass_mev, int32_t z_energy,
                            double* sums_dev, double* central_dev, pp_stream s) {
    // no allocation of any kind here: the library must already be initialised on this thread's device
    if (g_device < 0) return fail(PP_ERR_INVALID, "pp_beam_moments_noalloc: call pp_init first");
    MomentArgs m;
    int rc = moment_args(state, u_state, mass_mev, z_energy, m);
    if (rc) return rc;
    if (!sums_dev || !central_dev) return fail(PP_ERR_INVALID, "null output");
    if (pp_beam::workspace_bytes() > pp_beam::noalloc_scratch_bytes())
        return fail(PP_ERR_UNSUPPORTED, "pp_beam_moments_noalloc: %zu B of block partials exceed the %zu B scratch",
                    pp_beam::workspace_bytes(), pp_beam::noalloc_scratch_bytes());
    void* ws = pp_beam::noalloc_scratch(g_device);
    if (!ws) return fail(PP_ERR_INVALID, "pp_beam_moments_noalloc: no scratch on device %d (call pp_init)", g_device);
    HIP_TRY(pp_beam::moments(m, sums_dev, central_dev, ws, stream_of(s)), "beam moment kernels");
    return PP_OK;
}

int pp_beam_noalloc_scratch_bytes(size_t* bytes) {
    if (!bytes) return fail(PP_ERR_INVALID, "null argument");
    *bytes = pp_beam::noalloc_scratch_bytes();
    return PP_OK;
}

int pp_beam_ellipse_counts(const pp_push_args* state, int32_t u_state, const double* twiss_dev, double* counts_dev,
                           void* workspace, size_t ws_bytes, pp_stream s) {
    int rc = ensure_init();
    if (rc) return rc;
    MomentArgs m;
    if ((rc = moment_args(state, u_state, 1.0, 0, m))) return rc;
    if (!twiss_dev || !counts_dev || !workspace) return fail(PP_ERR_INVALID, "null argument");
    if (ws_bytes < pp_beam::workspace_bytes()) return fail(PP_ERR_INVALID, "workspace too small");
    HIP_TRY(pp_beam::ellipse_counts(m, twiss_dev, counts_dev, workspace, stream_of(s)), "ellipse count kernels");
    return PP_OK;
}

int pp_gather_probe(const pp_map* map, int32_t mode, uint32_t ncell, int32_t iters, int64_t nthreads, float* out,
                    pp_stream s) {
    int rc = ensure_init();
    if (rc) return rc;
    if (!map || !out) return fail(PP_ERR_INVALID, "null argument");
    if (map->device != g_device) return fail(PP_ERR_INVALID, "map lives on another device");
    const MapDesc& m = map->d;
    if (!m.coef32 || m.ncomp != 6) return fail(PP_ERR_UNSUPPORTED, "probe needs a 6-component map with an FP32 table");
    const uint64_t cells = (uint64_t)m.nc[0] * m.nc[1] * m.nc[2];
    const uint64_t bytes = cells * 192u;
    if (bytes >= 0x7FFFFFFFull) return fail(PP_ERR_UNSUPPORTED, "map too large for 32-bit offsets");
    if (mode != 0 && mode != 1) return fail(PP_ERR_INVALID, "mode must be 0 or 1");
    if (ncell == 0 || (ncell & (ncell - 1)) || ncell > cells)
        return fail(PP_ERR_INVALID, "ncell must be a power of two no larger than the map's %llu cells",
                    (unsigned long long)cells);
    if (iters < 0 || nthreads < 0 || nthreads > ((int64_t)1 << 31)) return fail(PP_ERR_INVALID, "bad sizes");
    if (nthreads == 0) return PP_OK;
    hipError_t e = pp_fast::launch_gather_probe(m.coef32, (uint32_t)bytes, ncell, mode, iters, nthreads, out,
                                                stream_of(s));
    if (e != hipSuccess) return hip_fail(e, "gather probe launch");
    return PP_OK;
}

int pp_aos_to_soa(const double* aos, int64_t m, double* a, double* b, double* c, pp_stream s) {
    return util_common(PP_UT_AOS2SOA, a, b, c, const_cast<double*>(aos), m, s);
}

int pp_soa_to_aos(const double* a, const double* b, const double* c, int64_t m, double* aos, pp_stream s) {
    return util_common(PP_UT_SOA2AOS, const_cast<double*>(a), const_cast<double*>(b), const_cast<double*>(c), aos, m, s);
}

int pp_v_to_u(double* vx, double* vy, double* vz, int64_t m, pp_stream s) {
    return util_common(PP_UT_V2U, vx, vy, vz, nullptr, m, s);
}

int pp_u_to_v(double* ux, double* uy, double* uz, int64_t m, pp_stream s) {
    return util_common(PP_UT_U2V, ux, uy, uz, nullptr, m, s);
}

}  // extern "C"
