// Beam moments of the device-resident particle state: the collective
// quantities of ParticleDistribution (src/PyPATools/particles.py:693-843) as
// reductions over the SoA state, without copying it to the host.
//
// Pass 1 sums positions, momenta (beta*gamma), kinetic energy and the angles
// v / v_z, and takes the position extremes.  Pass 2 takes the means from pass 1
// and sums centred second moments (no cancellation of sum-of-squares forms).
// Each pass writes one partial per block; one block per quantity then combines
// the partials in a fixed order, so a result does not depend on scheduling.  A third pass counts
// the particles inside the 4-rms Twiss ellipses.
#include <mutex>

#include "pp_internal.h"

namespace pp_beamk {

constexpr int kBlock = 256;
constexpr double kEpsilon = 1e-10;   // global_variables.py:33, the v_z guard of particles.py:464

struct Particle {
    double r[3], v[3], p[3], pabs, ekin, ang[3];
};

// momentum (beta*gamma) and velocity of one slot; u-state slots hold gamma*v
__device__ __forceinline__ Particle load_particle(const MomentArgs& a, int64_t i) {
    Particle q;
    q.r[0] = a.x[i];
    q.r[1] = a.y[i];
    q.r[2] = a.z[i];
    const double w0 = a.vx[i], w1 = a.vy[i], w2 = a.vz[i];
    const double ic = 1.0 / a.clight;
    double g;
    if (a.u_state) {
        const double u2 = (w0 * w0 + w1 * w1 + w2 * w2) * ic * ic;
        g = sqrt(1.0 + u2);
        q.p[0] = w0 * ic;
        q.p[1] = w1 * ic;
        q.p[2] = w2 * ic;
        q.v[0] = w0 / g;
        q.v[1] = w1 / g;
        q.v[2] = w2 / g;
    } else {
        double b2 = (w0 * w0 + w1 * w1 + w2 * w2) * ic * ic;
        b2 = b2 < a.beta2_max ? b2 : a.beta2_max;
        g = 1.0 / sqrt(1.0 - b2);
        q.v[0] = w0;
        q.v[1] = w1;
        q.v[2] = w2;
        q.p[0] = g * w0 * ic;
        q.p[1] = g * w1 * ic;
        q.p[2] = g * w2 * ic;
    }
    const double p2 = q.p[0] * q.p[0] + q.p[1] * q.p[1] + q.p[2] * q.p[2];
    q.pabs = sqrt(p2);
    // E = m (gamma - 1) with pr = |p| (or p_z): m pr^2 / (sqrt(pr^2 + 1) + 1), no cancellation
    const double pr = a.z_energy ? q.p[2] : q.pabs;
    q.ekin = a.mass_mev * (pr * pr) / (sqrt(pr * pr + 1.0) + 1.0);
    const double vz = fabs(q.v[2]) < kEpsilon ? kEpsilon : q.v[2];
    q.ang[0] = q.v[0] / vz;
    q.ang[1] = q.v[1] / vz;
    q.ang[2] = q.v[2] / vz;
    return q;
}

__device__ __forceinline__ bool counted(const MomentArgs& a, int64_t i) { return !a.alive || a.alive[i]; }

// block-wide reduction of NQ doubles (sum, or min / max for the flagged ones) into out[NQ]
template <int NQ>
__device__ void block_reduce(double (&v)[NQ], const int (&op)[NQ], double* out) {
    __shared__ double red[kBlock / 64][NQ];
    for (int k = 0; k < NQ; ++k) {
        for (int off = 32; off > 0; off >>= 1) {
            const double o = __shfl_xor(v[k], off, 64);
            v[k] = op[k] == 0 ? v[k] + o : (op[k] < 0 ? fmin(v[k], o) : fmax(v[k], o));
        }
    }
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0)
        for (int k = 0; k < NQ; ++k) red[wave][k] = v[k];
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int k = 0; k < NQ; ++k) {
            double s = red[0][k];
            for (int w = 1; w < (int)(blockDim.x >> 6); ++w)
                s = op[k] == 0 ? s + red[w][k] : (op[k] < 0 ? fmin(s, red[w][k]) : fmax(s, red[w][k]));
            out[k] = s;
        }
    }
}

// pass 1: n, sum r, sum p, sum |p|, sum E, sum angles, min r, max r
constexpr int kQ1 = 18;
__constant__ int kOp1[kQ1] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, -1, -1, -1, 1, 1, 1};

__global__ void __launch_bounds__(kBlock) k_pass1(const MomentArgs a, double* __restrict__ partial) {
    double v[kQ1];
    int op[kQ1];
    for (int k = 0; k < kQ1; ++k) {
        op[k] = kOp1[k];
        v[k] = op[k] == 0 ? 0.0 : (op[k] < 0 ? INFINITY : -INFINITY);
    }
    for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < a.n; i += (int64_t)gridDim.x * kBlock) {
        if (!counted(a, i)) continue;
        const Particle q = load_particle(a, i);
        v[0] += 1.0;
        for (int d = 0; d < 3; ++d) {
            v[1 + d] += q.r[d];
            v[4 + d] += q.p[d];
            v[9 + d] += q.ang[d];
            v[12 + d] = fmin(v[12 + d], q.r[d]);
            v[15 + d] = fmax(v[15 + d], q.r[d]);
        }
        v[7] += q.pabs;
        v[8] += q.ekin;
    }
    block_reduce<kQ1>(v, op, partial + (size_t)blockIdx.x * kQ1);
}

// pass 2: centred second moments about the pass-1 means
constexpr int kQ2 = 11;

__global__ void __launch_bounds__(kBlock) k_pass2(const MomentArgs a, const double* __restrict__ sums,
                                                  double* __restrict__ partial) {
    const double n = sums[0] > 0.0 ? sums[0] : 1.0;
    const double mr[3] = {sums[1] / n, sums[2] / n, sums[3] / n};
    const double ma[3] = {sums[9] / n, sums[10] / n, sums[11] / n};
    const double me = sums[8] / n;
    double v[kQ2];
    int op[kQ2];
    for (int k = 0; k < kQ2; ++k) {
        v[k] = 0.0;
        op[k] = 0;
    }
    for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < a.n; i += (int64_t)gridDim.x * kBlock) {
        if (!counted(a, i)) continue;
        const Particle q = load_particle(a, i);
        double dr[3], da[3];
        for (int d = 0; d < 3; ++d) {
            dr[d] = q.r[d] - mr[d];
            da[d] = q.ang[d] - ma[d];
            v[d] += dr[d] * dr[d];
            v[3 + d] += da[d] * da[d];
        }
        v[6] += dr[0] * da[0];
        v[7] += dr[1] * da[1];
        v[8] += dr[0] * da[1];
        v[9] += dr[1] * da[0];
        const double de = q.ekin - me;
        v[10] += de * de;
    }
    block_reduce<kQ2>(v, op, partial + (size_t)blockIdx.x * kQ2);
}

// pass 3: particles with gamma x^2 + 2 alpha x x' + beta x'^2 < 4 eps (x and y), particles.py:821-835
__global__ void __launch_bounds__(kBlock) k_ellipse(const MomentArgs a, const double* __restrict__ tw,
                                                    double* __restrict__ partial) {
    double v[2] = {0.0, 0.0};
    int op[2] = {0, 0};
    for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < a.n; i += (int64_t)gridDim.x * kBlock) {
        if (!counted(a, i)) continue;
        const Particle q = load_particle(a, i);
        // tw = alpha_x, beta_x, gamma_x, alpha_y, beta_y, gamma_y, eps_x, eps_y
        const double ex = tw[2] * q.r[0] * q.r[0] + 2.0 * tw[0] * q.r[0] * q.ang[0] + tw[1] * q.ang[0] * q.ang[0];
        const double ey = tw[5] * q.r[1] * q.r[1] + 2.0 * tw[3] * q.r[1] * q.ang[1] + tw[4] * q.ang[1] * q.ang[1];
        v[0] += ex < 4.0 * tw[6] ? 1.0 : 0.0;
        v[1] += ey < 4.0 * tw[7] ? 1.0 : 0.0;
    }
    block_reduce<2>(v, op, partial + (size_t)blockIdx.x * 2);
}

// combines the nblocks partials of quantity blockIdx.x: thread t takes blocks t, t+256, ...
// in order, then a fixed-shape tree over the 256 threads (the result does not depend on timing)
__global__ void __launch_bounds__(kBlock) k_finish(const double* __restrict__ partial, int nblocks, int nq, int which,
                                                   double* __restrict__ out) {
    __shared__ double sh[kBlock];
    const int k = blockIdx.x, t = threadIdx.x;
    const int op = which == 1 ? kOp1[k] : 0;
    double s = op == 0 ? 0.0 : (op < 0 ? INFINITY : -INFINITY);
    for (int b = t; b < nblocks; b += kBlock) {
        const double o = partial[(size_t)b * nq + k];
        s = op == 0 ? s + o : (op < 0 ? fmin(s, o) : fmax(s, o));
    }
    sh[t] = s;
    __syncthreads();
    for (int w = kBlock / 2; w > 0; w >>= 1) {
        if (t < w) sh[t] = op == 0 ? sh[t] + sh[t + w] : (op < 0 ? fmin(sh[t], sh[t + w]) : fmax(sh[t], sh[t + w]));
        __syncthreads();
    }
    if (t == 0) out[k] = sh[0];
}

}  // namespace pp_beamk

namespace pp_beam {

static int blocks_for(int64_t n) {
    const int64_t b = (n + pp_beamk::kBlock - 1) / pp_beamk::kBlock;
    return (int)(b < 1 ? 1 : (b > kMomentBlocks ? kMomentBlocks : b));
}

size_t workspace_bytes() { return sizeof(double) * (size_t)kMomentBlocks * pp_beamk::kQ1; }

// ---------------------------------------------------------------------------
// Fixed device scratch for the allocation-free moments (pp_beam_moments_noalloc):
// the block partials of one call, kMomentBlocks x kQ1 doubles (147456 B), held by
// every device the library initialises until the process ends.  In host code the
// array's name is only a host-side stand-in; kernels get the device address that
// hipGetSymbolAddress returns for the current device, resolved once per device.
// Every call on one device shares its array: calls that overlap on one device race.
// ---------------------------------------------------------------------------

static __device__ __align__(256) double scratch[(size_t)kMomentBlocks * pp_beamk::kQ1];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};
}  // namespace

size_t noalloc_scratch_bytes() { return sizeof(scratch); }

hipError_t init_noalloc_scratch(int device) {
    if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (g_scratch_ptr[device]) return hipSuccess;
    void* ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    if (!ptr) return hipErrorInvalidSymbol;
    g_scratch_ptr[device] = ptr;
    return hipSuccess;
}

void* noalloc_scratch(int device) {
    if (device < 0 || device >= kMaxDevices) return nullptr;
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    return g_scratch_ptr[device];
}

hipError_t moments(const MomentArgs& a, double* sums, double* central, void* ws, hipStream_t s) {
    double* partial = static_cast<double*>(ws);
    const int nb = blocks_for(a.n);
    hipLaunchKernelGGL(pp_beamk::k_pass1, dim3(nb), dim3(pp_beamk::kBlock), 0, s, a, partial);
    hipLaunchKernelGGL(pp_beamk::k_finish, dim3(pp_beamk::kQ1), dim3(pp_beamk::kBlock), 0, s, partial, nb,
                       pp_beamk::kQ1, 1, sums);
    hipLaunchKernelGGL(pp_beamk::k_pass2, dim3(nb), dim3(pp_beamk::kBlock), 0, s, a, sums, partial);
    hipLaunchKernelGGL(pp_beamk::k_finish, dim3(pp_beamk::kQ2), dim3(pp_beamk::kBlock), 0, s, partial, nb,
                       pp_beamk::kQ2, 2, central);
    return hipGetLastError();
}

hipError_t ellipse_counts(const MomentArgs& a, const double* twiss_dev, double* counts, void* ws, hipStream_t s) {
    double* partial = static_cast<double*>(ws);
    const int nb = blocks_for(a.n);
    hipLaunchKernelGGL(pp_beamk::k_ellipse, dim3(nb), dim3(pp_beamk::kBlock), 0, s, a, twiss_dev, partial);
    hipLaunchKernelGGL(pp_beamk::k_finish, dim3(2), dim3(pp_beamk::kBlock), 0, s, partial, nb, 2, 3, counts);
    return hipGetLastError();
}

}  // namespace pp_beam
