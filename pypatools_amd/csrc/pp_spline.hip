// B-spline interpolation of gridded values as the reference's CoordinateMapper
// computes it (src/PyPATools/field_src/interpolators.py:415-526), and the
// 'nearest' method of its 'scipy' backend (RegularGridInterpolator):
//
//   - per axis, a point maps to a fractional node index the way np.interp(x,
//     grid, arange(n), left=-1, right=n) does it (interpolators.py:515-520):
//     the node j with g[j] <= x < g[j+1] found by binary search, then
//     (1 / (g[j+1] - g[j])) * (x - g[j]) + j, rounded after each operation
//     (this translation unit is built without FMA contraction);
//   - the spline of order 0, 1, 3 or 5 is evaluated as
//     scipy.ndimage.map_coordinates evaluates it (interpolators.py:504-510):
//     order 0 takes the node floor(c + 0.5); orders 3 and 5 use coefficients
//     from the recursive B-spline prefilter (poles of the B-spline's
//     z-transform, Unser 1993), with taps beyond the edges folded back by
//     whole-sample symmetry (mode 'constant', points with c outside [0, n-1]
//     on any axis get the fill value) or, for mode 'nearest', from an array
//     padded with 12 copies of the edge node on each side and filtered with
//     half-sample symmetric boundaries, taps clamped to it;
//   - RGI 'nearest': node i or i+1 of the cell [g[i], g[i+1]] holding x, i+1
//     only when (x - g[i]) / (g[i+1] - g[i]) > 0.5; outside the grid, fill.
//
// The coefficients are built on the device: a padding kernel, then one
// filtering pass per axis with one thread per grid line.
#include <math.h>
#include <string.h>

#include "pp_internal.h"

namespace pp_splk {

constexpr int kPad = 12;   // edge copies per side for mode 'nearest' (orders 3, 5)

__device__ __forceinline__ int64_t stride_of(const SplineDesc& d, int a) {
    int64_t s = 1;
    for (int b = a + 1; b < 3; ++b) s *= d.N[b];
    return s;
}

// ---------------------------------------------------------------------------
// Coefficients
// ---------------------------------------------------------------------------

// coef[c][i] = values[c][clamp(i - pad)] (edge replication of mode 'nearest')
__global__ void __launch_bounds__(256) k_pad(const double* __restrict__ vals, double* __restrict__ coef, int64_t nc,
                                             int64_t n0, int64_t n1, int64_t n2, int64_t N0, int64_t N1, int64_t N2,
                                             int64_t pad) {
    const int64_t total = nc * N0 * N1 * N2;
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < total;
         t += (int64_t)gridDim.x * blockDim.x) {
        int64_t r = t;
        const int64_t k = r % N2;
        r /= N2;
        const int64_t j = r % N1;
        r /= N1;
        const int64_t i = r % N0;
        const int64_t c = r / N0;
        const int64_t si = min(max(i - pad, (int64_t)0), n0 - 1);
        const int64_t sj = min(max(j - pad, (int64_t)0), n1 - 1);
        const int64_t sk = min(max(k - pad, (int64_t)0), n2 - 1);
        coef[t] = vals[((c * n0 + si) * n1 + sj) * n2 + sk];
    }
}

// One pole of the prefilter on every line along `axis`: causal then anticausal recursion.
// reflect = 0: whole-sample symmetric boundaries (mirror); 1: half-sample symmetric.
__global__ void __launch_bounds__(128) k_filter(double* __restrict__ coef, int64_t nlines, int64_t len,
                                                int64_t stride, int64_t inner, double z, double gain, int reflect) {
    const int64_t line = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (line >= nlines) return;
    // line -> base: lines enumerate (outer, inner) with the axis between them
    const int64_t outer = line / inner, in = line % inner;
    double* c = coef + outer * len * stride + in;
    if (len == 1) return;
    if (gain != 1.0)
        for (int64_t k = 0; k < len; ++k) c[k * stride] *= gain;
    // causal initial value: the recursion's steady state over the symmetric periodic extension
    double c0;
    if (!reflect) {
        const int64_t period = 2 * len - 2;
        double acc = 0.0, p = 1.0;
        for (int64_t k = 0; k < period; ++k) {
            const int64_t src = k < len ? k : period - k;
            acc += p * c[src * stride];
            p *= z;
        }
        c0 = acc / (1.0 - p);
    } else {
        const int64_t period = 2 * len;
        double acc = 0.0, p = 1.0;
        for (int64_t k = 0; k < period; ++k) {
            const int64_t src = k < len ? k : period - 1 - k;
            acc += p * c[src * stride];
            p *= z;
        }
        c0 = c[0] + z * (acc / (1.0 - p));
    }
    c[0] = c0;
    for (int64_t k = 1; k < len; ++k) c[k * stride] += z * c[(k - 1) * stride];
    const int64_t l = len - 1;
    if (!reflect) c[l * stride] = (z / (z * z - 1.0)) * (c[l * stride] + z * c[(l - 1) * stride]);
    else c[l * stride] = (z / (z - 1.0)) * c[l * stride];
    for (int64_t k = l - 1; k >= 0; --k) c[k * stride] = z * (c[(k + 1) * stride] - c[k * stride]);
}

// ---------------------------------------------------------------------------
// Evaluation
// ---------------------------------------------------------------------------

// largest j with g[j] <= x, for g[0] <= x < g[n-1] (n >= 2)
__device__ __forceinline__ int64_t bsearch_le(const double* g, int64_t n, double x) {
    int64_t lo = 0, hi = n - 1;
    while (hi - lo > 1) {
        const int64_t mid = (lo + hi) >> 1;
        if (g[mid] <= x) lo = mid;
        else hi = mid;
    }
    return lo;
}

// np.interp(x, g, arange(n), left=-1, right=n)
__device__ double interp_index(const double* g, int64_t n, double x) {
    if (!(x >= g[0])) return x < g[0] ? -1.0 : (double)n;      // NaN maps outside
    if (x > g[n - 1]) return (double)n;
    if (x == g[n - 1]) return (double)(n - 1);
    const int64_t j = bsearch_le(g, n, x);
    if (g[j] == x) return (double)j;
    const double slope = 1.0 / (g[j + 1] - g[j]);
    return slope * (x - g[j]) + (double)j;
}

__device__ __forceinline__ double bspline3(double t) {
    t = fabs(t);
    if (t < 1.0) return 2.0 / 3.0 - t * t + 0.5 * t * t * t;
    if (t < 2.0) {
        const double u = 2.0 - t;
        return u * u * u / 6.0;
    }
    return 0.0;
}

__device__ __forceinline__ double bspline5(double t) {
    t = fabs(t);
    const double t2 = t * t;
    if (t < 1.0) return 11.0 / 20.0 - 0.5 * t2 + 0.25 * t2 * t2 - t2 * t2 * t / 12.0;
    if (t < 2.0)
        return 17.0 / 40.0 + t * (5.0 / 8.0 + t * (-7.0 / 4.0 + t * (5.0 / 4.0 + t * (-3.0 / 8.0 + t / 24.0))));
    if (t < 3.0) {
        const double u = 3.0 - t;
        const double u2 = u * u;
        return u2 * u2 * u / 120.0;
    }
    return 0.0;
}

// whole-sample symmetric fold of a tap index into [0, n)
__device__ __forceinline__ int64_t fold_mirror(int64_t k, int64_t n) {
    if (n == 1) return 0;
    const int64_t p = 2 * n - 2;
    k %= p;
    if (k < 0) k += p;
    return k < n ? k : p - k;
}

constexpr int kMaxTaps = 6;

// taps (coefficient indices along the axis) and weights of one axis; returns the tap count, 0 = outside
__device__ int axis_taps(const SplineDesc& d, int a, double x, int64_t (&idx)[kMaxTaps], double (&w)[kMaxTaps]) {
    const double* g = d.axis[a];
    const int64_t n = d.n[a];
    if (d.rgi) {   // RegularGridInterpolator 'nearest'
        if (!(x >= g[0] && x <= g[n - 1])) return 0;
        if (n == 1) {
            idx[0] = 0;
            w[0] = 1.0;
            return 1;
        }
        int64_t i = 0, hi = n;   // np.searchsorted(g, x) - 1 = (count of nodes < x) - 1, clipped to [0, n-2]
        while (i < hi) {
            const int64_t mid = (i + hi) >> 1;
            if (g[mid] < x) i = mid + 1;
            else hi = mid;
        }
        i = min(max(i - 1, (int64_t)0), n - 2);
        const double t = (x - g[i]) / (g[i + 1] - g[i]);
        idx[0] = t <= 0.5 ? i : i + 1;
        w[0] = 1.0;
        return 1;
    }
    double c = interp_index(g, n, x);
    if (!d.nearest) {
        if (!(c >= 0.0 && c <= (double)(n - 1))) return 0;
    } else {
        c = fmin(fmax(c, 0.0), (double)(n - 1));
    }
    const int64_t N = d.N[a];
    const double u = c + (double)d.pad;
    if (d.order == 0) {
        const int64_t k = (int64_t)floor(u + 0.5);
        idx[0] = min(max(k, (int64_t)0), N - 1);
        w[0] = 1.0;
        return 1;
    }
    const double fl = floor(u);
    const int ntap = d.order + 1;
    const int64_t s0 = (int64_t)fl - d.order / 2;
    for (int q = 0; q < ntap; ++q) {
        const int64_t k = s0 + q;
        const double t = u - (double)k;
        w[q] = d.order == 1 ? fmax(0.0, 1.0 - fabs(t)) : (d.order == 3 ? bspline3(t) : bspline5(t));
        idx[q] = d.nearest ? min(max(k, (int64_t)0), N - 1) : fold_mirror(k, N);
    }
    return ntap;
}

// out[i*ncomp + c] for the points pts[i*stride + col0 + a], a < ndim
__global__ void __launch_bounds__(256) k_eval(const SplineDesc d, const double* __restrict__ pts, int64_t m,
                                              int32_t stride, int32_t col0, double* __restrict__ out) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m) return;
    int64_t idx[3][kMaxTaps];
    double w[3][kMaxTaps];
    int nt[3] = {1, 1, 1};
    bool inside = true;
    for (int a = 0; a < 3; ++a) {
        if (a < d.ndim) {
            nt[a] = axis_taps(d, a, pts[i * stride + col0 + a], idx[a], w[a]);
            inside = inside && nt[a] > 0;
        } else {
            idx[a][0] = 0;
            w[a][0] = 1.0;
        }
    }
    const int64_t s0 = stride_of(d, 0), s1 = stride_of(d, 1);
    const int64_t per = d.N[0] * d.N[1] * d.N[2];
    for (int64_t c = 0; c < d.ncomp; ++c) {
        double acc = d.cval;
        if (inside) {
            const double* q = d.coef + c * per;
            acc = 0.0;
            for (int p0 = 0; p0 < nt[0]; ++p0) {
                double a1 = 0.0;
                for (int p1 = 0; p1 < nt[1]; ++p1) {
                    double a2 = 0.0;
                    const double* r = q + idx[0][p0] * s0 + idx[1][p1] * s1;
                    for (int p2 = 0; p2 < nt[2]; ++p2) a2 += w[2][p2] * r[idx[2][p2]];
                    a1 += w[1][p1] * a2;
                }
                acc += w[0][p0] * a1;
            }
        }
        out[i * d.ncomp + c] = acc;
    }
}

}  // namespace pp_splk

// ---------------------------------------------------------------------------
// Host side (called from the C ABI in pp_host.hip)
// ---------------------------------------------------------------------------

namespace pp_spl {

static int grid_blocks(int64_t n, int block) {
    const int64_t b = (n + block - 1) / block;
    return (int)(b < 1 ? 1 : (b > 65535 ? 65535 : b));
}

// poles of the B-spline prefilter of `order` (|z| < 1) and their count
static int poles(int order, double* z) {
    if (order == 3) {
        z[0] = sqrt(3.0) - 2.0;
        return 1;
    }
    if (order == 5) {   // roots of z^4 + 26 z^3 + 66 z^2 + 26 z + 1 inside the unit circle
        z[0] = sqrt(135.0 / 2.0 - sqrt(17745.0 / 4.0)) + sqrt(105.0 / 4.0) - 13.0 / 2.0;
        z[1] = sqrt(135.0 / 2.0 + sqrt(17745.0 / 4.0)) - sqrt(105.0 / 4.0) - 13.0 / 2.0;
        return 2;
    }
    return 0;
}

void destroy(pp_spline* s) {
    if (!s) return;
    for (int k = 0; k < s->nalloc; ++k) (void)hipFree(s->allocs[k]);
    delete s;
}

// values (host, [ncomp][n0][n1][n2]) and axes (host) are validated by the caller
hipError_t create(const double* values, int64_t ncomp, int ndim, const int64_t* n, const double* const* axes, int order,
                  int nearest, int rgi, double cval, int device, pp_spline** out) {
    pp_spline* s = new pp_spline();
    memset(&s->d, 0, sizeof(s->d));
    s->nalloc = 0;
    s->device = device;
    pp_splk::SplineDesc& d = s->d;
    d.ndim = ndim;
    d.order = order;
    d.nearest = nearest;
    d.rgi = rgi;
    d.cval = cval;
    d.ncomp = ncomp;
    d.pad = (nearest && order >= 3) ? pp_splk::kPad : 0;
    int64_t nvals = ncomp;
    for (int a = 0; a < 3; ++a) {
        d.n[a] = a < ndim ? n[a] : 1;
        d.N[a] = a < ndim ? d.n[a] + 2 * d.pad : 1;
        nvals *= d.n[a];
    }
    const int64_t ncoef = ncomp * d.N[0] * d.N[1] * d.N[2];
    hipError_t e;
    auto alloc = [&](size_t bytes, void** p) {
        e = hipMalloc(p, bytes ? bytes : 256);
        if (e == hipSuccess) s->allocs[s->nalloc++] = *p;
        return e == hipSuccess;
    };
    void *vals = nullptr, *coef = nullptr, *ax = nullptr;
    int64_t naxis = 0;
    for (int a = 0; a < ndim; ++a) naxis += d.n[a];
    if (!alloc(sizeof(double) * nvals, &vals) || !alloc(sizeof(double) * ncoef, &coef) ||
        !alloc(sizeof(double) * naxis, &ax)) {
        destroy(s);
        return e;
    }
    int64_t off = 0;
    for (int a = 0; a < ndim; ++a) {
        if ((e = hipMemcpy(static_cast<double*>(ax) + off, axes[a], sizeof(double) * d.n[a], hipMemcpyHostToDevice)))
            break;
        d.axis[a] = static_cast<double*>(ax) + off;
        off += d.n[a];
    }
    if (e == hipSuccess) e = hipMemcpy(vals, values, sizeof(double) * nvals, hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        destroy(s);
        return e;
    }
    d.coef = static_cast<double*>(coef);
    hipLaunchKernelGGL(pp_splk::k_pad, dim3(grid_blocks(ncoef, 256)), dim3(256), 0, 0, static_cast<double*>(vals),
                       static_cast<double*>(coef), ncomp, d.n[0], d.n[1], d.n[2], d.N[0], d.N[1], d.N[2],
                       (int64_t)d.pad);
    double z[2];
    const int np = poles(order, z);
    if (np > 0) {
        double gain = 1.0;
        for (int q = 0; q < np; ++q) gain *= (1.0 - z[q]) * (1.0 - 1.0 / z[q]);
        for (int a = 0; a < ndim; ++a) {
            const int64_t len = d.N[a];
            int64_t inner = 1;
            for (int b = a + 1; b < 3; ++b) inner *= d.N[b];
            const int64_t nlines = ncoef / len;
            for (int q = 0; q < np; ++q)
                hipLaunchKernelGGL(pp_splk::k_filter, dim3((unsigned)((nlines + 127) / 128)), dim3(128), 0, 0,
                                   static_cast<double*>(coef), nlines, len, inner, inner, z[q], q == 0 ? gain : 1.0,
                                   d.nearest);
        }
    }
    e = hipGetLastError();
    if (e == hipSuccess) e = hipDeviceSynchronize();
    (void)hipFree(vals);
    s->allocs[0] = s->allocs[--s->nalloc];   // vals was the first allocation
    if (e != hipSuccess) {
        destroy(s);
        return e;
    }
    *out = s;
    return hipSuccess;
}

hipError_t eval(const pp_spline* s, const double* pts, int64_t m, int32_t stride, int32_t col0, double* out,
                hipStream_t st) {
    if (m <= 0) return hipSuccess;
    hipLaunchKernelGGL(pp_splk::k_eval, dim3((unsigned)((m + 255) / 256)), dim3(256), 0, st, s->d, pts, m, stride, col0,
                       out);
    return hipGetLastError();
}

}  // namespace pp_spl
