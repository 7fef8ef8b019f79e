// Space-charge field on a box mesh: cloud-in-cell charge deposit, a direct
// Poisson solve and E = -grad(phi), the path of PyAMGPoissonSolver.solve
// (src/PyPATools/poisson_amg.py:616-797) for a domain without electrodes.
//
// The reference discretises -lap(phi) = rho / eps0 on the nx*ny*nz cell-centre
// nodes with the 7-point stencil and phi = 0 one spacing outside the box
// (poisson_amg.py:293-444 with no conductor cells) and solves it iteratively.
// That operator separates: along each axis it is tridiag(-1, 2, -1) / h^2 with
// zero ends, whose eigenvectors are the discrete sine transform (DST-I) rows
// S[k][i] = sin(pi (k+1)(i+1) / (n+1)) with eigenvalues (4/h^2) sin^2(pi (k+1) /
// (2 (n+1))).  So phi = S_x S_y S_z [ (S_x S_y S_z b) / lambda ] * prod 2/(n+1),
// an exact solve in six axis transforms and one division.
//
// Deposit and gradient follow the reference's CPU forms operation by operation
// (poisson_amg.py:849-911, 98-130); FMA contraction is off in this file.
#include <cmath>
#include <vector>

#include "pp_internal.h"

#pragma clang fp contract(off)

namespace pp_poissonk {

// CIC deposit, poisson_amg.py:849-911: u = clip((x + L/2) / h, 0, n - 1), i = floor(u),
// f = u - i, then i = clip(i, 0, n - 2) (f is kept), density q w / (hx hy hz) on the 8 corners
// (the reference sums q w and divides by the cell volume once).
__global__ void __launch_bounds__(256) k_deposit(const PoissonDesc d, const double* __restrict__ x,
                                                 const double* __restrict__ y, const double* __restrict__ z,
                                                 const double* __restrict__ q, double qc, const uint8_t* alive,
                                                 int64_t n, double* rho) {
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    if (alive && !alive[p]) return;
    const double pos[3] = {x[p], y[p], z[p]};
    int ix[3];
    double f[3];
    for (int a = 0; a < 3; ++a) {
        double u = (pos[a] + d.L[a] / 2.0) / d.h[a];
        const double top = (double)(d.n[a] - 1);
        u = u < 0.0 ? 0.0 : (u > top ? top : u);
        u = u == u ? u : 0.0;            // NaN position -> node 0 (in bounds)
        int i = (int)floor(u);
        f[a] = u - (double)i;
        i = i < 0 ? 0 : (i > d.n[a] - 2 ? d.n[a] - 2 : i);
        ix[a] = i;
    }
    const double c = q ? q[p] : qc;
    const double vol = (d.h[0] * d.h[1]) * d.h[2];   // density: charge / cell volume
    for (int dx = 0; dx < 2; ++dx)
        for (int dy = 0; dy < 2; ++dy)
            for (int dz = 0; dz < 2; ++dz) {
                const double w = ((dx == 0 ? 1.0 - f[0] : f[0]) * (dy == 0 ? 1.0 - f[1] : f[1])) *
                                 (dz == 0 ? 1.0 - f[2] : f[2]);
                const int64_t idx = ((int64_t)(ix[0] + dx) * d.n[1] + (ix[1] + dy)) * d.n[2] + (ix[2] + dz);
                atomicAdd(rho + idx, c * w / vol);
            }
}

// out(l, k) = scale * sum_i S[k][i] in(l, i) along the contiguous axis (length n, lines of n)
__global__ void __launch_bounds__(256) k_dst_inner(const double* __restrict__ in, const double* __restrict__ ST,
                                                   int n, int64_t lines, double scale, double* __restrict__ out) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= lines * n) return;
    const int64_t l = t / n;
    const int k = (int)(t - l * n);
    const double* src = in + l * n;
    double s = 0.0;
    for (int i = 0; i < n; ++i) s += ST[(int64_t)i * n + k] * src[i];
    out[t] = scale * s;
}

// out(a, k, b) = scale * sum_i S[k][i] in(a, i, b): axis of length n with stride `inner`, `outer` blocks
__global__ void __launch_bounds__(256) k_dst_outer(const double* __restrict__ in, const double* __restrict__ S,
                                                   int n, int64_t inner, int64_t outer, double scale,
                                                   double* __restrict__ out) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= outer * n * inner) return;
    const int64_t b = t % inner;
    const int64_t r = t / inner;
    const int k = (int)(r % n);
    const int64_t a = r / n;
    const double* src = in + a * n * inner + b;
    const double* row = S + (int64_t)k * n;
    double s = 0.0;
    for (int i = 0; i < n; ++i) s += row[i] * src[(int64_t)i * inner];
    out[t] = scale * s;
}

// bhat / (lambda_x[i] + lambda_y[j] + lambda_z[k]) * scale, with b = rho / eps0 folded into scale
__global__ void __launch_bounds__(256) k_divide(double* __restrict__ v, const double* __restrict__ lx,
                                                const double* __restrict__ ly, const double* __restrict__ lz, int nx,
                                                int ny, int nz, double scale) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)nx * ny * nz) return;
    const int k = (int)(t % nz);
    const int64_t r = t / nz;
    const int j = (int)(r % ny);
    const int i = (int)(r / ny);
    v[t] = v[t] * scale / ((lx[i] + ly[j]) + lz[k]);
}

// E = -grad(phi) by central differences on interior nodes, 0 on the outer layer (poisson_amg.py:98-130)
__global__ void __launch_bounds__(256) k_gradient(const PoissonDesc d, const double* __restrict__ phi,
                                                  double* __restrict__ ex, double* __restrict__ ey,
                                                  double* __restrict__ ez) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int nx = d.n[0], ny = d.n[1], nz = d.n[2];
    if (t >= (int64_t)nx * ny * nz) return;
    const int k = (int)(t % nz);
    const int64_t r = t / nz;
    const int j = (int)(r % ny);
    const int i = (int)(r / ny);
    if (i == 0 || j == 0 || k == 0 || i == nx - 1 || j == ny - 1 || k == nz - 1) {
        ex[t] = 0.0;
        ey[t] = 0.0;
        ez[t] = 0.0;
        return;
    }
    const int64_t sx = (int64_t)ny * nz, sy = nz;
    ex[t] = -(phi[t + sx] - phi[t - sx]) / (2.0 * d.h[0]);
    ey[t] = -(phi[t + sy] - phi[t - sy]) / (2.0 * d.h[1]);
    ez[t] = -(phi[t + 1] - phi[t - 1]) / (2.0 * d.h[2]);
}

}  // namespace pp_poissonk

struct pp_poisson {
    PoissonDesc d;
    double* S[3];      // S[a][k][i] (symmetric, so also S^T)
    double* lam[3];    // eigenvalues per axis
    double* work;      // 2 n^3
    int device;
};

namespace pp_poisson_impl {

static unsigned blocks(int64_t n) { return (unsigned)((n + 255) / 256); }

int create(const PoissonDesc& d, pp_poisson** out, char* err, size_t errn) {
    pp_poisson* P = new pp_poisson();
    P->d = d;
    hipGetDevice(&P->device);
    const int64_t cells = (int64_t)d.n[0] * d.n[1] * d.n[2];
    bool ok = true;
    for (int a = 0; a < 3; ++a) {
        const int n = d.n[a];
        std::vector<double> S((size_t)n * n), lam((size_t)n);
        const double pi = 3.14159265358979323846;
        for (int k = 0; k < n; ++k) {
            for (int i = 0; i < n; ++i) S[(size_t)k * n + i] = sin(pi * (double)(k + 1) * (double)(i + 1) / (double)(n + 1));
            const double s = sin(pi * (double)(k + 1) / (2.0 * (double)(n + 1)));
            lam[k] = 4.0 / (d.h[a] * d.h[a]) * (s * s);
        }
        ok = ok && hipMalloc(&P->S[a], sizeof(double) * n * n) == hipSuccess &&
             hipMalloc(&P->lam[a], sizeof(double) * n) == hipSuccess &&
             hipMemcpy(P->S[a], S.data(), sizeof(double) * n * n, hipMemcpyHostToDevice) == hipSuccess &&
             hipMemcpy(P->lam[a], lam.data(), sizeof(double) * n, hipMemcpyHostToDevice) == hipSuccess;
    }
    ok = ok && hipMalloc(&P->work, sizeof(double) * 2 * cells) == hipSuccess;
    if (!ok) {
        snprintf(err, errn, "device allocation for the Poisson solver failed");
        for (int a = 0; a < 3; ++a) {
            if (P->S[a]) (void)hipFree(P->S[a]);
            if (P->lam[a]) (void)hipFree(P->lam[a]);
        }
        if (P->work) (void)hipFree(P->work);
        delete P;
        return PP_ERR_NOMEM;
    }
    *out = P;
    return PP_OK;
}

void destroy(pp_poisson* P) {
    if (!P) return;
    for (int a = 0; a < 3; ++a) {
        (void)hipFree(P->S[a]);
        (void)hipFree(P->lam[a]);
    }
    (void)hipFree(P->work);
    delete P;
}

const PoissonDesc& desc(const pp_poisson* P) { return P->d; }
int device_of(const pp_poisson* P) { return P->device; }

hipError_t deposit(const pp_poisson* P, const double* x, const double* y, const double* z, const double* q,
                   double qc, const uint8_t* alive, int64_t n, double* rho, hipStream_t s) {
    hipLaunchKernelGGL(pp_poissonk::k_deposit, dim3(blocks(n)), dim3(256), 0, s, P->d, x, y, z, q, qc, alive, n,
                       rho);
    return hipGetLastError();
}

// applies S along every axis of `in` (nx, ny, nz), result in `out`; `tmp` is scratch (may alias in)
static void transform3(const pp_poisson* P, const double* in, double* tmp, double* out, double scale, hipStream_t s) {
    const int nx = P->d.n[0], ny = P->d.n[1], nz = P->d.n[2];
    const int64_t cells = (int64_t)nx * ny * nz;
    hipLaunchKernelGGL(pp_poissonk::k_dst_inner, dim3(blocks(cells)), dim3(256), 0, s, in, P->S[2], nz,
                       (int64_t)nx * ny, 1.0, out);
    hipLaunchKernelGGL(pp_poissonk::k_dst_outer, dim3(blocks(cells)), dim3(256), 0, s, out, P->S[1], ny, (int64_t)nz,
                       (int64_t)nx, 1.0, tmp);
    hipLaunchKernelGGL(pp_poissonk::k_dst_outer, dim3(blocks(cells)), dim3(256), 0, s, tmp, P->S[0], nx,
                       (int64_t)ny * nz, (int64_t)1, scale, out);
}

hipError_t solve(const pp_poisson* P, const double* rho, double* phi, hipStream_t s) {
    const int nx = P->d.n[0], ny = P->d.n[1], nz = P->d.n[2];
    const int64_t cells = (int64_t)nx * ny * nz;
    double* a = P->work;
    double* b = P->work + cells;
    transform3(P, rho, a, b, 1.0, s);                       // b = S rho
    const double norm = (2.0 / (nx + 1)) * (2.0 / (ny + 1)) * (2.0 / (nz + 1));
    hipLaunchKernelGGL(pp_poissonk::k_divide, dim3(blocks(cells)), dim3(256), 0, s, b, P->lam[0], P->lam[1], P->lam[2],
                       nx, ny, nz, 1.0 / kEps0);             // b = S rho / eps0 / lambda
    transform3(P, b, a, phi, norm, s);                      // phi = S b * prod 2/(n+1)
    return hipGetLastError();
}

hipError_t gradient(const pp_poisson* P, const double* phi, double* ex, double* ey, double* ez, hipStream_t s) {
    const int64_t cells = (int64_t)P->d.n[0] * P->d.n[1] * P->d.n[2];
    hipLaunchKernelGGL(pp_poissonk::k_gradient, dim3(blocks(cells)), dim3(256), 0, s, P->d, phi, ex, ey, ez);
    return hipGetLastError();
}

}  // namespace pp_poisson_impl
