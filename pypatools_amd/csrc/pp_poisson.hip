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
#include <algorithm>
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
// product with the n x n sine matrix), on the FP64 matrix cores: BM x BN output tiles, BK-deep k slices
// staged in LDS (edges zero-filled), four wavefronts (2 x 2) with 16 x 16 tiles of v_mfma_f64_16x16x4.
// 32 x 32 x 32 tiles give products up to 128^3 enough workgroups to fill the chip; 64 x 64 x 16 serve larger ones.
// Lane l of a 16x16x4 step holds A[l % 16][l / 16] and B[l / 16][l % 16]; its four results are
// D[l / 16 + 4 r][l % 16], r = 0..3 (rows interleaved across the lane groups).
struct Gemm {
    const double* A; const double* B; double* C;
    int64_t M, N, K, lda, ldb, ldc, sa, sb, sc;
    double alpha;
};

typedef double double4_t __attribute__((ext_vector_type(4)));

template <int kBM, int kBN, int kBK>
__global__ void __launch_bounds__(256) k_gemm_mfma(const Gemm g) {
    constexpr int kTU = kBM / 32, kTV = kBN / 32;             // 16 x 16 MFMA tiles per wavefront
    constexpr int kLA = kBM * kBK / 256, kLB = kBK * kBN / 256;   // tile elements each thread moves per k slice
    static_assert(kLA * 256 == kBM * kBK && kLB * 256 == kBK * kBN, "tile must split evenly over 256 threads");
    // two LDS buffers: the next k slice is read from global memory into registers while the matrix cores work
    // on the current one, then written to the other buffer (one barrier per slice)
    __shared__ double As[2][kBK][kBM + 1];
    __shared__ double Bs[2][kBK][kBN + 1];
    const int64_t b = blockIdx.z;
    const double* A = g.A + b * g.sa;
    const double* B = g.B + b * g.sb;
    double* C = g.C + b * g.sc;
    const int64_t m0 = (int64_t)blockIdx.y * kBM, n0 = (int64_t)blockIdx.x * kBN;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int wm = (wave >> 1) * (kBM / 2), wn = (wave & 1) * (kBN / 2);   // this wavefront's quarter of the tile
    const int li = lane & 15, lk = lane >> 4;
    double4_t acc[kTU][kTV];
#pragma unroll
    for (int u = 0; u < kTU; ++u)
#pragma unroll
        for (int v = 0; v < kTV; ++v) acc[u][v] = double4_t{0.0, 0.0, 0.0, 0.0};
    double ra[kLA], rb[kLB];
    auto load = [&](int64_t k0) {
#pragma unroll
        for (int q = 0; q < kLA; ++q) {
            const int e = threadIdx.x + 256 * q, r = e / kBK, c = e % kBK;   // A tile: rows m0 + r, columns k0 + c
            const int64_t gm = m0 + r, gk = k0 + c;
            ra[q] = (gm < g.M && gk < g.K) ? A[gm * g.lda + gk] : 0.0;
        }
#pragma unroll
        for (int q = 0; q < kLB; ++q) {
            const int e = threadIdx.x + 256 * q, r = e / kBN, c = e % kBN;   // B tile: rows k0 + r, columns n0 + c
            const int64_t gk = k0 + r, gn = n0 + c;
            rb[q] = (gk < g.K && gn < g.N) ? B[gk * g.ldb + gn] : 0.0;
        }
    };
    auto store = [&](int buf) {
#pragma unroll
        for (int q = 0; q < kLA; ++q) {
            const int e = threadIdx.x + 256 * q;
            As[buf][e % kBK][e / kBK] = ra[q];
        }
#pragma unroll
        for (int q = 0; q < kLB; ++q) {
            const int e = threadIdx.x + 256 * q;
            Bs[buf][e / kBN][e % kBN] = rb[q];
        }
    };
    load(0);
    store(0);
    __syncthreads();
    int buf = 0;
    for (int64_t k0 = 0; k0 < g.K; k0 += kBK) {
        const bool more = k0 + kBK < g.K;
        if (more) load(k0 + kBK);
#pragma unroll
        for (int kk = 0; kk < kBK; kk += 4) {
            double a[kTU], bb[kTV];
#pragma unroll
            for (int u = 0; u < kTU; ++u) a[u] = As[buf][kk + lk][wm + 16 * u + li];
#pragma unroll
            for (int v = 0; v < kTV; ++v) bb[v] = Bs[buf][kk + lk][wn + 16 * v + li];
#pragma unroll
            for (int u = 0; u < kTU; ++u)
#pragma unroll
                for (int v = 0; v < kTV; ++v) acc[u][v] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[u], bb[v], acc[u][v], 0, 0, 0);
        }
        if (more) store(buf ^ 1);   // the other buffer's readers finished before the previous barrier
        __syncthreads();
        buf ^= 1;
    }
#pragma unroll
    for (int u = 0; u < kTU; ++u)
#pragma unroll
        for (int v = 0; v < kTV; ++v)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int64_t gm = m0 + wm + 16 * u + lk + 4 * r, gn = n0 + wn + 16 * v + li;
                if (gm < g.M && gn < g.N) C[gm * g.ldc + gn] = g.alpha * acc[u][v][r];
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

// ---- electrode-bounded domains: Shortley-Weller operator and BiCGSTAB ----------

// Operator rows: coef[0] = diagonal, coef[1 + d] = neighbour d (+x, -x, +y, -y, +z, -z); 0 where there is no
// entry (outside the box or a conductor).  Conductor rows hold only the box diagonal (phi = 0 there).
struct SwOperator {
    const double* coef[7];
    int nx, ny, nz;
};

__global__ void __launch_bounds__(256) k_sw_apply(const SwOperator A, const double* __restrict__ x,
                                                  double* __restrict__ y) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t n = (int64_t)A.nx * A.ny * A.nz;
    if (t >= n) return;
    const int k = (int)(t % A.nz);
    const int64_t r = t / A.nz;
    const int j = (int)(r % A.ny);
    const int i = (int)(r / A.ny);
    const int64_t sx = (int64_t)A.ny * A.nz, sy = A.nz;
    double acc = A.coef[0][t] * x[t];
    if (i + 1 < A.nx) acc += A.coef[1][t] * x[t + sx];
    if (i > 0) acc += A.coef[2][t] * x[t - sx];
    if (j + 1 < A.ny) acc += A.coef[3][t] * x[t + sy];
    if (j > 0) acc += A.coef[4][t] * x[t - sy];
    if (k + 1 < A.nz) acc += A.coef[5][t] * x[t + 1];
    if (k > 0) acc += A.coef[6][t] * x[t - 1];
    y[t] = acc;
}

constexpr int kDotBlocks = 512;

// partial sums of up to three dot products: part[block][q] = sum over the block's slice of a_q . b_q
struct Dots {
    const double* a[3];
    const double* b[3];
    int nq;
};

__global__ void __launch_bounds__(256) k_dots(const Dots D, int64_t n, double* __restrict__ part) {
    __shared__ double red[4][3];
    double s[3] = {0.0, 0.0, 0.0};
    for (int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x; t < n; t += (int64_t)gridDim.x * 256)
        for (int q = 0; q < D.nq; ++q) s[q] += D.a[q][t] * D.b[q][t];
    for (int q = 0; q < 3; ++q)
        for (int off = 32; off > 0; off >>= 1) s[q] += __shfl_xor(s[q], off, 64);
    if ((threadIdx.x & 63) == 0)
        for (int q = 0; q < 3; ++q) red[threadIdx.x >> 6][q] = s[q];
    __syncthreads();
    if (threadIdx.x < 3) part[(size_t)blockIdx.x * 3 + threadIdx.x] =
        ((red[0][threadIdx.x] + red[1][threadIdx.x]) + red[2][threadIdx.x]) + red[3][threadIdx.x];
}

// out[slot + q] = sum of the kDotBlocks partials of dot q, in a fixed order
__global__ void __launch_bounds__(256) k_dots_finish(const double* __restrict__ part, int nq, double* __restrict__ sc,
                                                     int slot) {
    __shared__ double sh[256];
    for (int q = 0; q < nq; ++q) {
        double v = 0.0;
        for (int b = threadIdx.x; b < kDotBlocks; b += 256) v += part[(size_t)b * 3 + q];
        sh[threadIdx.x] = v;
        __syncthreads();
        for (int w = 128; w > 0; w >>= 1) {
            if ((int)threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
            __syncthreads();
        }
        if (threadIdx.x == 0) sc[slot + q] = sh[0];
        __syncthreads();
    }
}

// BiCGSTAB scalars in device memory (no host round trip inside an iteration)
// (b, b) and (r, r) are adjacent, as are (t, s) and (t, t): one two-dot reduction fills each pair
enum { kRho = 0, kRhoOld, kAlpha, kOmega, kBeta, kR0V, kTS, kTT, kBB, kRR, kScalars = 16 };

__global__ void k_scalars(double* sc, int stage) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    if (stage == 0) {          // start: rho_old = alpha = omega = 1
        sc[kRhoOld] = 1.0;
        sc[kAlpha] = 1.0;
        sc[kOmega] = 1.0;
    } else if (stage == 1) {   // beta from the new rho
        sc[kBeta] = (sc[kRho] / sc[kRhoOld]) * (sc[kAlpha] / sc[kOmega]);
    } else if (stage == 2) {   // alpha = rho / (r0, v)
        sc[kAlpha] = sc[kRho] / sc[kR0V];
    } else if (stage == 3) {   // omega = (t, s) / (t, t); rho_old = rho
        sc[kOmega] = sc[kTT] != 0.0 ? sc[kTS] / sc[kTT] : 0.0;
        sc[kRhoOld] = sc[kRho];
    }
}

// which: 0 p = r + beta (p - omega v); 1 s = r - alpha v; 2 x += alpha ph + omega sh, r = s - omega t;
// 3 r = b - r (residual from A x in r); 4 u0 = w0; 5 u0 = 0
__global__ void __launch_bounds__(256) k_bicg_update(int which, int64_t n, const double* __restrict__ sc, double* u0,
                                                     double* u1, const double* w0, const double* w1, const double* w2,
                                                     const double* w3) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    if (which == 0) {
        u0[t] = w0[t] + sc[kBeta] * (u0[t] - sc[kOmega] * w1[t]);                  // p
    } else if (which == 1) {
        u0[t] = w0[t] - sc[kAlpha] * w1[t];                                        // s
    } else if (which == 2) {
        u0[t] = u0[t] + sc[kAlpha] * w0[t] + sc[kOmega] * w1[t];                   // x
        u1[t] = w2[t] - sc[kOmega] * w3[t];                                        // r
    } else if (which == 3) {
        u0[t] = w0[t] - u0[t];
    } else if (which == 4) {
        u0[t] = w0[t];
    } else {
        u0[t] = 0.0;
    }
}

// b = rho / eps0 on non-conductor nodes, 0 on conductors; and phi = 0 on conductors (which == 1)
__global__ void __launch_bounds__(256) k_mask(int which, int64_t n, const int32_t* __restrict__ type,
                                              const double* __restrict__ in, double* __restrict__ out) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const bool cond = type[t] == 2;
    if (which == 0) out[t] = cond ? 0.0 : in[t] / kEps0;
    else if (cond) out[t] = 0.0;
}

}  // namespace pp_poissonk

struct pp_poisson {
    PoissonDesc d;
    double* S[3];      // S[a][k][i] (symmetric, so also S^T)
    double* lam[3];    // eigenvalues per axis
    double* work;      // 2 n^3
    int device;
    // electrode-bounded domain (null: box domain, direct solve)
    std::vector<int32_t> h_type;         // cell types (0 interior, 1 boundary, 2 conductor)
    std::vector<double> h_dist;          // (N, 6) distances along the axis rays
    std::vector<int32_t> h_count;        // (N, 6) crossing counts
    int32_t* type;                       // device cell types
    double* coef;                        // 7 N operator coefficients
    double* vec;                         // BiCGSTAB vectors: x r r0 p v s t ph sh (9 N)
    double* part;                        // dot partials
    double* sc;                          // scalars
    double rtol;
    int32_t max_iter;
    mutable int32_t last_iter;           // statistics of the last solve
    mutable double last_res;
};

namespace pp_poisson_impl {

static unsigned blocks(int64_t n) { return (unsigned)((n + 255) / 256); }

int create(const PoissonDesc& d, pp_poisson** out, char* err, size_t errn) {
    pp_poisson* P = new pp_poisson();
    P->d = d;
    P->type = nullptr;
    P->coef = P->vec = P->part = P->sc = nullptr;
    P->rtol = 1e-10;
    P->max_iter = 500;
    P->last_iter = 0;
    P->last_res = 0.0;
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

static void free_conductors(pp_poisson* P) {
    for (void* q : {(void*)P->type, (void*)P->coef, (void*)P->vec, (void*)P->part, (void*)P->sc})
        if (q) (void)hipFree(q);
    P->type = nullptr;
    P->coef = P->vec = P->part = P->sc = nullptr;
    P->h_type.clear();
    P->h_dist.clear();
    P->h_count.clear();
}

void destroy(pp_poisson* P) {
    if (!P) return;
    for (int a = 0; a < 3; ++a) {
        (void)hipFree(P->S[a]);
        (void)hipFree(P->lam[a]);
    }
    (void)hipFree(P->work);
    free_conductors(P);
    delete P;
}

const PoissonDesc& desc(const pp_poisson* P) { return P->d; }
int device_of(const pp_poisson* P) { return P->device; }

hipError_t deposit(const pp_poisson* P, const double* x, const double* y, const double* z, const double* q,
                   double qc, const uint8_t* alive, int64_t n, double* rho, hipStream_t s) {
    // about 2048 blocks, each a contiguous chunk of at least 4 particles per thread: 1e6 particles make
    // 245 blocks, about one per CU (16 per thread left most CUs idle; 1 per thread flushes the LDS table
    // too often)
    const int64_t per = (int64_t)pp_poissonk::kDepositBlock * 4;
    int64_t chunk = (n + 2047) / 2048;
    chunk = chunk < per ? per : chunk;
    const unsigned nb = (unsigned)((n + chunk - 1) / chunk);
    hipLaunchKernelGGL(pp_poissonk::k_deposit, dim3(nb), dim3(pp_poissonk::kDepositBlock), 0, s, P->d, x, y, z, q,
                       qc, alive, n, chunk, rho);
    return hipGetLastError();
}

static void gemm(const pp_poissonk::Gemm& g, int64_t batch, hipStream_t s) {
    const int64_t tiles64 = ((g.N + 63) / 64) * ((g.M + 63) / 64) * batch;
    if (tiles64 < 2048) {
        const dim3 grid((unsigned)((g.N + 31) / 32), (unsigned)((g.M + 31) / 32), (unsigned)batch);
        hipLaunchKernelGGL((pp_poissonk::k_gemm_mfma<32, 32, 32>), grid, dim3(256), 0, s, g);
    } else {
        const dim3 grid((unsigned)((g.N + 63) / 64), (unsigned)((g.M + 63) / 64), (unsigned)batch);
        hipLaunchKernelGGL((pp_poissonk::k_gemm_mfma<64, 64, 16>), grid, dim3(256), 0, s, g);
    }
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

// out = A_box^-1 in * scale, A_box the 7-point operator of the box without conductors (exact, by sine transforms)
static void box_inverse(const pp_poisson* P, const double* in, double* out, double scale, hipStream_t s) {
    const int nx = P->d.n[0], ny = P->d.n[1], nz = P->d.n[2];
    const int64_t cells = (int64_t)nx * ny * nz;
    double* a = P->work;
    double* b = P->work + cells;
    transform3(P, in, a, b, 1.0, s);                        // b = S in
    const double norm = (2.0 / (nx + 1)) * (2.0 / (ny + 1)) * (2.0 / (nz + 1));
    hipLaunchKernelGGL(pp_poissonk::k_divide, dim3(blocks(cells)), dim3(256), 0, s, b, P->lam[0], P->lam[1], P->lam[2],
                       nx, ny, nz, scale);                  // b = S in * scale / lambda
    transform3(P, b, a, out, norm, s);                      // out = S b * prod 2/(n+1)
}

// ---- electrode-bounded domains ------------------------------------------------------------------

// Cell-centre coordinates of axis a as numpy.linspace(-L/2 + h/2, L/2 - h/2, n) computes them
// (i * step + start, the last one exactly stop): the classification sees the reference's nodes.
static std::vector<double> axis_nodes(const PoissonDesc& d, int a) {
    const int n = d.n[a];
    const double start = -d.L[a] / 2.0 + d.h[a] / 2.0, stop = d.L[a] / 2.0 - d.h[a] / 2.0;
    std::vector<double> x((size_t)n);
    if (n == 1) {
        x[0] = start;
        return x;
    }
    const double step = (stop - start) / (double)(n - 1);
    for (int i = 0; i < n; ++i) x[(size_t)i] = (double)i * step + start;
    x[(size_t)n - 1] = stop;
    return x;
}

static inline bool owned(double du, double dv) { return dv > 0.0 || (dv == 0.0 && du < 0.0); }

// Crossings of the six axis rays of every node with the mesh (counts and distance to the nearest one).
// For axis a the mesh is projected onto the plane of the other two axes (u, v) relative to each node line;
// edge functions E_ij = P_i.u P_j.v - P_i.v P_j.u are exact negatives for the two triangles sharing an edge,
// and an edge passing exactly through the line belongs to one of them by the top-left rule, so every
// crossing of a closed surface counts once.  Depth = (E_12 V_0[a] + E_20 V_1[a] + E_01 V_2[a]) / A.
static void axis_rays(const PoissonDesc& d, const double* V, int64_t nt, const int32_t* T, std::vector<double>& dist,
                      std::vector<int32_t>& count) {
    const int64_t N = (int64_t)d.n[0] * d.n[1] * d.n[2];
    dist.assign((size_t)N * 6, INFINITY);
    count.assign((size_t)N * 6, 0);
    std::vector<double> ax[3] = {axis_nodes(d, 0), axis_nodes(d, 1), axis_nodes(d, 2)};
    double ext = 0.0;
    for (int a = 0; a < 3; ++a) ext = std::max(ext, d.L[a]);
    const double pad = 1e-9 * ext;
    const int64_t stride[3] = {(int64_t)d.n[1] * d.n[2], d.n[2], 1};
    for (int a = 0; a < 3; ++a) {
        const int u = (a + 1) % 3, v = (a + 2) % 3;
        const int nu = d.n[u], nv = d.n[v];
        std::vector<std::vector<double>> depths((size_t)nu * nv);
        for (int64_t t = 0; t < nt; ++t) {
            const double* P[3] = {V + 3 * (int64_t)T[3 * t], V + 3 * (int64_t)T[3 * t + 1], V + 3 * (int64_t)T[3 * t + 2]};
            const double umin = std::min(P[0][u], std::min(P[1][u], P[2][u])) - pad;
            const double umax = std::max(P[0][u], std::max(P[1][u], P[2][u])) + pad;
            const double vmin = std::min(P[0][v], std::min(P[1][v], P[2][v])) - pad;
            const double vmax = std::max(P[0][v], std::max(P[1][v], P[2][v])) + pad;
            const int iu0 = (int)(std::lower_bound(ax[u].begin(), ax[u].end(), umin) - ax[u].begin());
            const int iu1 = (int)(std::upper_bound(ax[u].begin(), ax[u].end(), umax) - ax[u].begin());
            const int iv0 = (int)(std::lower_bound(ax[v].begin(), ax[v].end(), vmin) - ax[v].begin());
            const int iv1 = (int)(std::upper_bound(ax[v].begin(), ax[v].end(), vmax) - ax[v].begin());
            for (int iu = iu0; iu < iu1; ++iu)
                for (int iv = iv0; iv < iv1; ++iv) {
                    const double ou = ax[u][(size_t)iu], ov = ax[v][(size_t)iv];
                    double pu[3], pv[3];
                    for (int q = 0; q < 3; ++q) {
                        pu[q] = P[q][u] - ou;
                        pv[q] = P[q][v] - ov;
                    }
                    const double e01 = pu[0] * pv[1] - pv[0] * pu[1];
                    const double e12 = pu[1] * pv[2] - pv[1] * pu[2];
                    const double e20 = pu[2] * pv[0] - pv[2] * pu[0];
                    const double A = (e01 + e12) + e20;
                    if (A == 0.0) continue;
                    const double sg = A > 0.0 ? 1.0 : -1.0;
                    const double E[3] = {e01, e12, e20};
                    bool inside = true;
                    for (int q = 0; q < 3 && inside; ++q) {
                        const int i = q, j = (q + 1) % 3;
                        const double se = sg * E[q];
                        const double du = sg * (P[j][u] - P[i][u]), dv = sg * (P[j][v] - P[i][v]);
                        inside = se > 0.0 || (se == 0.0 && owned(du, dv));
                    }
                    if (!inside) continue;
                    depths[(size_t)iu * nv + iv].push_back(((e12 * P[0][a] + e20 * P[1][a]) + e01 * P[2][a]) / A);
                }
        }
        for (int iu = 0; iu < nu; ++iu)
            for (int iv = 0; iv < nv; ++iv) {
                std::vector<double>& dl = depths[(size_t)iu * nv + iv];
                if (dl.empty()) continue;
                std::sort(dl.begin(), dl.end());
                for (int ia = 0; ia < d.n[a]; ++ia) {
                    const double x = ax[a][(size_t)ia];
                    const size_t nb = (size_t)(std::lower_bound(dl.begin(), dl.end(), x) - dl.begin());  // depth < x
                    int64_t idx[3];
                    idx[a] = ia;
                    idx[u] = iu;
                    idx[v] = iv;
                    const int64_t node = idx[0] * stride[0] + idx[1] * stride[1] + idx[2] * stride[2];
                    count[(size_t)node * 6 + 2 * a] = (int32_t)(dl.size() - nb);
                    count[(size_t)node * 6 + 2 * a + 1] = (int32_t)nb;
                    if (nb < dl.size()) dist[(size_t)node * 6 + 2 * a] = dl[nb] - x;
                    if (nb > 0) dist[(size_t)node * 6 + 2 * a + 1] = x - dl[nb - 1];
                }
            }
    }
}

static int build_operator(pp_poisson* P, char* err, size_t errn);

int set_conductors(pp_poisson* P, const double* V, int64_t nv, const int32_t* T, int64_t nt, char* err, size_t errn) {
    free_conductors(P);
    if (nt == 0) return PP_OK;
    for (int64_t k = 0; k < 3 * nv; ++k)
        if (!std::isfinite(V[k])) {
            snprintf(err, errn, "vertex %lld is not finite", (long long)(k / 3));
            return PP_ERR_INVALID;
        }
    for (int64_t k = 0; k < 3 * nt; ++k)
        if (T[k] < 0 || T[k] >= nv) {
            snprintf(err, errn, "triangle %lld has a vertex index outside [0, %lld)", (long long)(k / 3), (long long)nv);
            return PP_ERR_INVALID;
        }
    axis_rays(P->d, V, nt, T, P->h_dist, P->h_count);
    return build_operator(P, err, errn);
}

int set_ray_data(pp_poisson* P, const double* dist, const int32_t* count, char* err, size_t errn) {
    free_conductors(P);
    const size_t n6 = (size_t)6 * P->d.n[0] * P->d.n[1] * P->d.n[2];
    for (size_t k = 0; k < n6; ++k)
        if (!(dist[k] >= 0.0) || count[k] < 0) {
            snprintf(err, errn, "ray data entry %zu: distances must be >= 0 (inf allowed), counts >= 0", k);
            return PP_ERR_INVALID;
        }
    P->h_dist.assign(dist, dist + n6);
    P->h_count.assign(count, count + n6);
    return build_operator(P, err, errn);
}

// cell types (poisson_amg.py:241-291) and Shortley-Weller rows (:293-444) from the ray data
static int build_operator(pp_poisson* P, char* err, size_t errn) {
    const PoissonDesc& d = P->d;
    const int nx = d.n[0], ny = d.n[1], nz = d.n[2];
    const int64_t N = (int64_t)nx * ny * nz;
    // poisson_amg.py:241-291: conductor where a ray has an odd crossing count; boundary where a distance
    // is at most its axis spacing rounded to float32
    P->h_type.assign((size_t)N, 0);
    for (int64_t t = 0; t < N; ++t) {
        bool cond = false, bnd = false;
        for (int q = 0; q < 6; ++q) {
            cond = cond || (P->h_count[(size_t)t * 6 + q] % 2 != 0);
            bnd = bnd || P->h_dist[(size_t)t * 6 + q] <= (double)(float)d.h[q / 2];
        }
        P->h_type[(size_t)t] = cond ? 2 : (bnd ? 1 : 0);
    }
    // poisson_amg.py:293-444: Shortley-Weller rows (conductor rows become diagonal rows here)
    std::vector<double> coef((size_t)7 * N, 0.0);
    const double box_diag = (2.0 / (d.h[0] * d.h[0]) + 2.0 / (d.h[1] * d.h[1])) + 2.0 / (d.h[2] * d.h[2]);
    const int64_t stride[3] = {(int64_t)ny * nz, nz, 1};
    for (int i = 0; i < nx; ++i)
        for (int j = 0; j < ny; ++j)
            for (int k = 0; k < nz; ++k) {
                const int64_t t = (int64_t)i * stride[0] + (int64_t)j * stride[1] + k;
                if (P->h_type[(size_t)t] == 2) {   // phi = 0, scaled like the box rows (keeps the system balanced)
                    coef[(size_t)t] = box_diag;
                    continue;
                }
                const int ijk[3] = {i, j, k};
                double diag = 0.0;
                for (int a = 0; a < 3; ++a) {
                    const double h = d.h[a];
                    const bool has_p = ijk[a] + 1 < d.n[a], has_m = ijk[a] > 0;
                    const bool cond_p = has_p && P->h_type[(size_t)(t + stride[a])] == 2;
                    const bool cond_m = has_m && P->h_type[(size_t)(t - stride[a])] == 2;
                    const double hp = cond_p ? std::max(P->h_dist[(size_t)t * 6 + 2 * a], h * 1e-3) : h;
                    const double hm = cond_m ? std::max(P->h_dist[(size_t)t * 6 + 2 * a + 1], h * 1e-3) : h;
                    diag += 2.0 / (hp * hm);
                    if (has_p && !cond_p) coef[(size_t)(1 + 2 * a) * N + t] = -2.0 / (hp * (hp + hm));
                    if (has_m && !cond_m) coef[(size_t)(2 + 2 * a) * N + t] = -2.0 / (hm * (hp + hm));
                }
                coef[(size_t)t] = diag;
            }
    const bool ok = hipMalloc(&P->type, sizeof(int32_t) * N) == hipSuccess &&
                    hipMalloc(&P->coef, sizeof(double) * 7 * N) == hipSuccess &&
                    hipMalloc(&P->vec, sizeof(double) * 9 * N) == hipSuccess &&
                    hipMalloc(&P->part, sizeof(double) * 3 * pp_poissonk::kDotBlocks) == hipSuccess &&
                    hipMalloc(&P->sc, sizeof(double) * pp_poissonk::kScalars) == hipSuccess &&
                    hipMemcpy(P->type, P->h_type.data(), sizeof(int32_t) * N, hipMemcpyHostToDevice) == hipSuccess &&
                    hipMemcpy(P->coef, coef.data(), sizeof(double) * 7 * N, hipMemcpyHostToDevice) == hipSuccess;
    if (!ok) {
        free_conductors(P);
        snprintf(err, errn, "device allocation for the conductor operator failed");
        return PP_ERR_NOMEM;
    }
    return PP_OK;
}

bool has_conductors(const pp_poisson* P) { return P->coef != nullptr; }

void conductor_data(const pp_poisson* P, int32_t* type, double* dist, int32_t* count) {
    const size_t N = P->h_type.size();
    if (type) std::copy(P->h_type.begin(), P->h_type.end(), type);
    if (dist) std::copy(P->h_dist.begin(), P->h_dist.begin() + 6 * N, dist);
    if (count) std::copy(P->h_count.begin(), P->h_count.begin() + 6 * N, count);
}

void set_tolerance(pp_poisson* P, double rtol, int32_t max_iter) {
    P->rtol = rtol;
    P->max_iter = max_iter;
}

void stats(const pp_poisson* P, int32_t* iterations, double* residual) {
    if (iterations) *iterations = P->last_iter;
    if (residual) *residual = P->last_res;
}

// BiCGSTAB with right preconditioning, for A phi = rho / eps0 with the Shortley-Weller rows.  The
// preconditioner is the box solve with its result set to 0 on conductor nodes, so every iterate is 0 there
// (and the residual too, since conductor rows are diagonal with b = 0); it starts from that preconditioned b.
// Reads one residual norm per iteration to the host.
static hipError_t solve_conductors(const pp_poisson* P, const double* rho, double* phi, hipStream_t s) {
    using namespace pp_poissonk;
    const PoissonDesc& d = P->d;
    const int64_t N = (int64_t)d.n[0] * d.n[1] * d.n[2];
    const unsigned g = blocks(N);
    double* x = phi;
    double* r = P->vec;
    double* r0 = P->vec + N;
    double* p = P->vec + 2 * N;
    double* v = P->vec + 3 * N;
    double* sv = P->vec + 4 * N;
    double* tv = P->vec + 5 * N;
    double* ph = P->vec + 6 * N;
    double* sh = P->vec + 7 * N;
    double* b = P->vec + 8 * N;
    SwOperator A;
    for (int q = 0; q < 7; ++q) A.coef[q] = P->coef + (size_t)q * N;
    A.nx = d.n[0];
    A.ny = d.n[1];
    A.nz = d.n[2];
    auto dots = [&](const double* a0, const double* b0, const double* a1, const double* b1, const double* a2,
                    const double* b2, int nq, int slot) {
        Dots D{{a0, a1, a2}, {b0, b1, b2}, nq};
        hipLaunchKernelGGL(k_dots, dim3(kDotBlocks), dim3(256), 0, s, D, N, P->part);
        hipLaunchKernelGGL(k_dots_finish, dim3(1), dim3(256), 0, s, P->part, nq, P->sc, slot);
    };
    auto upd = [&](int which, double* u0, double* u1, const double* w0, const double* w1, const double* w2,
                   const double* w3) {
        hipLaunchKernelGGL(k_bicg_update, dim3(g), dim3(256), 0, s, which, N, P->sc, u0, u1, w0, w1, w2, w3);
    };
    auto precond = [&](const double* in, double* out) {
        box_inverse(P, in, out, 1.0, s);
        hipLaunchKernelGGL(k_mask, dim3(g), dim3(256), 0, s, 1, N, P->type, nullptr, out);
    };
    hipLaunchKernelGGL(k_mask, dim3(g), dim3(256), 0, s, 0, N, P->type, rho, b);     // b = rho / eps0, 0 on conductors
    precond(b, x);                                                                  // x0 = K^-1 b
    hipLaunchKernelGGL(k_sw_apply, dim3(g), dim3(256), 0, s, A, x, r);
    upd(3, r, nullptr, b, nullptr, nullptr, nullptr);                               // r = b - A x0
    upd(4, r0, nullptr, r, nullptr, nullptr, nullptr);
    upd(4, p, nullptr, r, nullptr, nullptr, nullptr);
    upd(5, v, nullptr, nullptr, nullptr, nullptr, nullptr);                         // v = 0
    hipLaunchKernelGGL(k_scalars, dim3(1), dim3(64), 0, s, P->sc, 0);
    dots(b, b, r, r, nullptr, nullptr, 2, kBB);
    double host[2];
    hipError_t e = hipMemcpyAsync(host, P->sc + kBB, sizeof(double), hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipMemcpyAsync(host + 1, P->sc + kRR, sizeof(double), hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) return e;
    const double bnorm = std::sqrt(host[0]);
    double res = bnorm > 0.0 ? std::sqrt(host[1]) / bnorm : 0.0;
    int it = 0;
    while (res > P->rtol && it < P->max_iter) {
        dots(r0, r, nullptr, nullptr, nullptr, nullptr, 1, kRho);
        hipLaunchKernelGGL(k_scalars, dim3(1), dim3(64), 0, s, P->sc, 1);
        if (it > 0) upd(0, p, nullptr, r, v, nullptr, nullptr);                     // p = r + beta (p - omega v)
        precond(p, ph);
        hipLaunchKernelGGL(k_sw_apply, dim3(g), dim3(256), 0, s, A, ph, v);
        dots(r0, v, nullptr, nullptr, nullptr, nullptr, 1, kR0V);
        hipLaunchKernelGGL(k_scalars, dim3(1), dim3(64), 0, s, P->sc, 2);
        upd(1, sv, nullptr, r, v, nullptr, nullptr);                                 // s = r - alpha v
        precond(sv, sh);
        hipLaunchKernelGGL(k_sw_apply, dim3(g), dim3(256), 0, s, A, sh, tv);
        dots(tv, sv, tv, tv, nullptr, nullptr, 2, kTS);
        hipLaunchKernelGGL(k_scalars, dim3(1), dim3(64), 0, s, P->sc, 3);
        upd(2, x, r, ph, sh, sv, tv);                                               // x += alpha ph + omega sh; r = s - omega t
        dots(r, r, nullptr, nullptr, nullptr, nullptr, 1, kRR);
        e = hipMemcpyAsync(host + 1, P->sc + kRR, sizeof(double), hipMemcpyDeviceToHost, s);
        if (e == hipSuccess) e = hipStreamSynchronize(s);
        if (e != hipSuccess) return e;
        res = std::sqrt(host[1]) / bnorm;
        ++it;
        if (!(res == res)) break;                                                   // breakdown (NaN)
    }
    hipLaunchKernelGGL(k_mask, dim3(g), dim3(256), 0, s, 1, N, P->type, nullptr, x);  // phi = 0 on conductors
    P->last_iter = it;
    P->last_res = res;
    return hipGetLastError();
}

hipError_t solve(const pp_poisson* P, const double* rho, double* phi, hipStream_t s) {
    if (P->coef) return solve_conductors(P, rho, phi, s);
    box_inverse(P, rho, phi, 1.0 / kEps0, s);
    return hipGetLastError();
}

hipError_t gradient(const pp_poisson* P, const double* phi, double* ex, double* ey, double* ez, hipStream_t s) {
    const int64_t cells = (int64_t)P->d.n[0] * P->d.n[1] * P->d.n[2];
    hipLaunchKernelGGL(pp_poissonk::k_gradient, dim3(blocks(cells)), dim3(256), 0, s, P->d, phi, ex, ey, ez);
    return hipGetLastError();
}

}  // namespace pp_poisson_impl
