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

// CIC weights of one particle, poisson_amg.py:849-911: u = clip((x + L/2) / h, 0, n - 1),
// i = floor(u), f = u - i, then i = clip(i, 0, n - 2) (f is kept); density q w / (hx hy hz) on the
// 8 corners (the reference sums q w and divides by the cell volume once).
__device__ __forceinline__ int64_t cic_weights(const PoissonDesc& d, double px, double py, double pz, double c,
                                               double (&a)[8]) {
    const double pos[3] = {px, py, pz};
    int ix[3];
    double f[3];
    for (int k = 0; k < 3; ++k) {
        double u = (pos[k] + d.L[k] / 2.0) / d.h[k];
        const double top = (double)(d.n[k] - 1);
        u = u < 0.0 ? 0.0 : (u > top ? top : u);
        u = u == u ? u : 0.0;            // NaN position -> node 0 (in bounds)
        int i = (int)floor(u);
        f[k] = u - (double)i;
        i = i < 0 ? 0 : (i > d.n[k] - 2 ? d.n[k] - 2 : i);
        ix[k] = i;
    }
    const double vol = (d.h[0] * d.h[1]) * d.h[2];
    for (int dx = 0; dx < 2; ++dx)
        for (int dy = 0; dy < 2; ++dy)
            for (int dz = 0; dz < 2; ++dz) {
                const double w = ((dx == 0 ? 1.0 - f[0] : f[0]) * (dy == 0 ? 1.0 - f[1] : f[1])) *
                                 (dz == 0 ? 1.0 - f[2] : f[2]);
                a[dx * 4 + dy * 2 + dz] = c * w / vol;
            }
    return ((int64_t)ix[0] * d.n[1] + ix[1]) * d.n[2] + ix[2];
}

// Deposit with a per-block table in LDS: a dense bunch puts thousands of particles on the same
// nodes, and same-address atomics serialise in L2.  Each block takes a contiguous chunk of
// particles, adds their contributions into an open-addressed LDS table (node index -> sum, LDS
// atomics), and then adds each table entry once to rho.  A contribution that finds no free slot
// within kProbe probes goes straight to rho.
constexpr int kTable = 4096;      // slots (power of two): 48 KB of LDS
constexpr int kProbe = 16;
constexpr int kDepositBlock = 1024;

__global__ void __launch_bounds__(kDepositBlock) k_deposit(const PoissonDesc d, const double* __restrict__ x,
                                                          const double* __restrict__ y, const double* __restrict__ z,
                                                          const double* __restrict__ q, double qc,
                                                          const uint8_t* alive, int64_t n, int64_t chunk,
                                                          double* rho) {
    __shared__ int64_t keys[kTable];
    __shared__ double vals[kTable];
    for (int t = threadIdx.x; t < kTable; t += kDepositBlock) {
        keys[t] = -1;
        vals[t] = 0.0;
    }
    __syncthreads();
    const int64_t sx = (int64_t)d.n[1] * d.n[2], sy = d.n[2];
    const int64_t begin = (int64_t)blockIdx.x * chunk;
    const int64_t end = begin + chunk < n ? begin + chunk : n;
    for (int64_t p = begin + threadIdx.x; p < end; p += kDepositBlock) {
        if (alive && !alive[p]) continue;
        double a[8];
        const int64_t base = cic_weights(d, x[p], y[p], z[p], q ? q[p] : qc, a);
        for (int k = 0; k < 8; ++k) {
            const int64_t node = base + (k >> 2) * sx + ((k >> 1) & 1) * sy + (k & 1);
            uint32_t h = (uint32_t)(node * 2654435761ull) & (kTable - 1);
            bool placed = false;
            for (int probe = 0; probe < kProbe && !placed; ++probe) {
                const int64_t prev = atomicCAS((unsigned long long*)&keys[h], (unsigned long long)-1ll,
                                               (unsigned long long)node);
                if (prev == -1 || prev == node) {
                    atomicAdd(&vals[h], a[k]);
                    placed = true;
                } else {
                    h = (h + 1) & (kTable - 1);
                }
            }
            if (!placed) atomicAdd(rho + node, a[k]);
        }
    }
    __syncthreads();
    for (int t = threadIdx.x; t < kTable; t += kDepositBlock)
        if (keys[t] >= 0 && vals[t] != 0.0) atomicAdd(rho + keys[t], vals[t]);
}

// Batched row-major FP64 GEMM, C[b] = alpha A[b] B[b], for the axis transforms (each one is a
// product with the n x n sine matrix).  64 x 64 output tiles, 16-deep k slices staged in LDS,
// 256 threads with a 4 x 4 register block each; edges are zero-filled.
struct Gemm {
    const double* A; const double* B; double* C;
    int64_t M, N, K, lda, ldb, ldc, sa, sb, sc;
    double alpha;
};

constexpr int kBM = 64, kBN = 64, kBK = 16;

__global__ void __launch_bounds__(256) k_gemm(const Gemm g) {
    __shared__ double As[kBK][kBM + 1];
    __shared__ double Bs[kBK][kBN + 1];
    const int64_t b = blockIdx.z;
    const double* A = g.A + b * g.sa;
    const double* B = g.B + b * g.sb;
    double* C = g.C + b * g.sc;
    const int64_t m0 = (int64_t)blockIdx.y * kBM, n0 = (int64_t)blockIdx.x * kBN;
    const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;
    double acc[4][4] = {};
    for (int64_t k0 = 0; k0 < g.K; k0 += kBK) {
        for (int e = threadIdx.x; e < kBM * kBK; e += 256) {
            const int r = e / kBK, c = e % kBK;            // A tile: rows m0 + r, columns k0 + c
            const int64_t gm = m0 + r, gk = k0 + c;
            As[c][r] = (gm < g.M && gk < g.K) ? A[gm * g.lda + gk] : 0.0;
        }
        for (int e = threadIdx.x; e < kBK * kBN; e += 256) {
            const int r = e / kBN, c = e % kBN;            // B tile: rows k0 + r, columns n0 + c
            const int64_t gk = k0 + r, gn = n0 + c;
            Bs[r][c] = (gk < g.K && gn < g.N) ? B[gk * g.ldb + gn] : 0.0;
        }
        __syncthreads();
#pragma unroll
        for (int kk = 0; kk < kBK; ++kk) {
            double av[4], bv[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                av[u] = As[kk][ty + 16 * u];
                bv[u] = Bs[kk][tx + 16 * u];
            }
#pragma unroll
            for (int u = 0; u < 4; ++u)
#pragma unroll
                for (int v = 0; v < 4; ++v) acc[u][v] += av[u] * bv[v];
        }
        __syncthreads();
    }
#pragma unroll
    for (int u = 0; u < 4; ++u)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const int64_t gm = m0 + ty + 16 * u, gn = n0 + tx + 16 * v;
            if (gm < g.M && gn < g.N) C[gm * g.ldc + gn] = g.alpha * acc[u][v];
        }
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
    // about 2048 blocks, each a contiguous chunk of at least 16 particles per thread
    const int64_t per = (int64_t)pp_poissonk::kDepositBlock * 16;
    int64_t chunk = (n + 2047) / 2048;
    chunk = chunk < per ? per : chunk;
    const unsigned nb = (unsigned)((n + chunk - 1) / chunk);
    hipLaunchKernelGGL(pp_poissonk::k_deposit, dim3(nb), dim3(pp_poissonk::kDepositBlock), 0, s, P->d, x, y, z, q,
                       qc, alive, n, chunk, rho);
    return hipGetLastError();
}

static void gemm(const pp_poissonk::Gemm& g, int64_t batch, hipStream_t s) {
    const dim3 grid((unsigned)((g.N + pp_poissonk::kBN - 1) / pp_poissonk::kBN),
                    (unsigned)((g.M + pp_poissonk::kBM - 1) / pp_poissonk::kBM), (unsigned)batch);
    hipLaunchKernelGGL(pp_poissonk::k_gemm, grid, dim3(256), 0, s, g);
}

// applies S along every axis of `in` (nx, ny, nz) into `out`, via `tmp`; the last product is scaled
static void transform3(const pp_poisson* P, const double* in, double* tmp, double* out, double scale, hipStream_t s) {
    const int64_t nx = P->d.n[0], ny = P->d.n[1], nz = P->d.n[2];
    pp_poissonk::Gemm g;
    // z: out (nx ny x nz) = in (nx ny x nz) . S_z  (S symmetric)
    g = {in, P->S[2], out, nx * ny, nz, nz, nz, nz, nz, 0, 0, 0, 1.0};
    gemm(g, 1, s);
    // y: tmp[x] (ny x nz) = S_y (ny x ny) . out[x] (ny x nz), one product per x
    g = {P->S[1], out, tmp, ny, nz, ny, ny, nz, nz, 0, ny * nz, ny * nz, 1.0};
    gemm(g, nx, s);
    // x: out (nx x ny nz) = S_x (nx x nx) . tmp (nx x ny nz)
    g = {P->S[0], tmp, out, nx, ny * nz, nx, nx, ny * nz, ny * nz, 0, 0, 0, scale};
    gemm(g, 1, s);
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
