// Strict FP64 kernels: no FMA contraction, one IEEE operation per NumPy
// operation of the reference, so every result equals the FP64 oracle bit for bit.
#include "pp_internal.h"
#include "pp_bvh.inc"

#pragma clang fp contract(off)

#define PP_STRICT 1

namespace pp_strict {

#include "pp_fp64.inc"

// ---- per-step kernels with per-particle fields (the reference's Numba batch
// kernels) and the elementwise stages of _rk4_step_batch / the field algebra.

__global__ void __launch_bounds__(256) k_batch(const BatchArgs q) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= q.m) return;
    const d3 A = load3(q.a, i);
    d3 r;
    switch (q.which) {
        case PP_BK_POS:                                   // position_update_batch, pusher.py:402
            r = A + load3(q.b, i) * q.s1;
            break;
        case PP_BK_BORIS:                                 // boris_push_batch, pusher.py:376
            r = boris_kick(A, load3(q.b, i), load3(q.c, i), q.s1, q.qm);
            break;
        case PP_BK_VAY:                                   // vay_push_batch, pusher.py:389
            r = vay_rel_kick(A, load3(q.b, i), load3(q.c, i), q.s1, q.qm, q.c2, q.clight, q.vclamp);
            break;
        case PP_BK_RK4REL:                                // rk4_rel_dbetagamma_dt_batch, pusher.py:326
            r = rk4_rel_deriv(A, load3(q.b, i), load3(q.c, i), q.qm, q.c2);
            break;
        case PP_BK_LORENTZ:                               // pusher.py:828-829
            r = lorentz(A, load3(q.b, i), load3(q.c, i), q.qm);
            break;
        case PP_BK_AXPY: {                                // r + 0.5*dt*k  (pusher.py:834-835 ...)
            const double sc = q.s_pp ? q.s_pp[i] : q.s1;
            r = A + sc * load3(q.b, i);
            break;
        }
        case PP_BK_RK4COMB: {                             // pusher.py:882-883
            const double w = q.s_pp ? q.s_pp[i] : q.s1;
            r = A + w * (((load3(q.b, i) + 2.0 * load3(q.c, i)) + 2.0 * load3(q.d, i)) + load3(q.e, i));
            break;
        }
        case PP_BK_CLAMP:                                 // pusher.py:886-890
            r = clamp_speed(A, q.vclamp);
            break;
        default: {                                        // PP_BK_SCALEOFF, field.py:690
            r = A;
            if (q.f1) r = q.s1 * r;
            if (q.f2) r = r + mk3(q.s2, q.s2, q.s2);
            break;
        }
    }
    q.out[3 * i] = r.x;
    q.out[3 * i + 1] = r.y;
    q.out[3 * i + 2] = r.z;
}

__global__ void __launch_bounds__(256) k_util(int which, double* a, double* b, double* c, double* aos, int64_t m,
                                              double c2) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m) return;
    if (which == PP_UT_AOS2SOA) {
        a[i] = aos[3 * i];
        b[i] = aos[3 * i + 1];
        c[i] = aos[3 * i + 2];
    } else if (which == PP_UT_SOA2AOS) {
        aos[3 * i] = a[i];
        aos[3 * i + 1] = b[i];
        aos[3 * i + 2] = c[i];
    } else if (which == PP_UT_V2U) {
        const d3 v = mk3(a[i], b[i], c[i]);
        const d3 u = rgamma(v, c2) * v;
        a[i] = u.x;
        b[i] = u.y;
        c[i] = u.z;
    } else {
        const d3 v = v_from_u(mk3(a[i], b[i], c[i]), c2);
        a[i] = v.x;
        b[i] = v.y;
        c[i] = v.z;
    }
}

__global__ void k_add_i64(int64_t* p, int64_t v) {
    if (threadIdx.x == 0 && blockIdx.x == 0) *p += v;
}

__global__ void __launch_bounds__(256) k_fill(double* p, double v, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// Per-cell trilinear coefficients of every component from the raw node table.
__global__ void __launch_bounds__(256) k_build_coef(const MapDesc d, double* __restrict__ c64, float* __restrict__ c32) {
    const int64_t ncell = (int64_t)d.nc[0] * d.nc[1] * d.nc[2];
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= ncell * d.ncomp) return;
    const int comp = (int)(t % d.ncomp);
    const int64_t cell = t / d.ncomp;
    const int k = (int)(cell % d.nc[2]);
    const int j = (int)((cell / d.nc[2]) % d.nc[1]);
    const int i = (int)(cell / ((int64_t)d.nc[2] * d.nc[1]));
    const int64_t ngz = d.ng[2], ngy = d.ng[1];
    const int i1 = d.n[0] > 1 ? i + 1 : i;
    const int j1 = d.n[1] > 1 ? j + 1 : j;
    const int k1 = d.n[2] > 1 ? k + 1 : k;   // periodic maps carry the wrap node at k = nz
    const double* v = d.raw + (int64_t)comp * d.ng[0] * ngy * ngz;
    auto at = [&](int a, int b, int c) { return v[((int64_t)a * ngy + b) * ngz + c]; };
    const double c000 = at(i, j, k), c100 = at(i1, j, k), c010 = at(i, j1, k), c110 = at(i1, j1, k);
    const double c001 = at(i, j, k1), c101 = at(i1, j, k1), c011 = at(i, j1, k1), c111 = at(i1, j1, k1);
    double q[8];
    q[0] = c000;
    q[1] = c100 - c000;
    q[2] = c010 - c000;
    q[3] = c001 - c000;
    q[4] = (c110 - c100) - (c010 - c000);
    q[5] = (c101 - c100) - (c001 - c000);
    q[6] = (c011 - c010) - (c001 - c000);
    q[7] = ((c111 - c110) - (c101 - c100)) - ((c011 - c010) - (c001 - c000));
    const int64_t o = (cell * d.ncomp + comp) * 8;
    for (int e = 0; e < 8; ++e) {
        if (c64) c64[o + e] = q[e];
        if (c32) c32[o + e] = (float)q[e];
    }
}

static inline unsigned grid_for(int64_t n, int block) { return (unsigned)((n + block - 1) / block); }

#define PP_STRICT_CASES(L, U)                                                                                \
    switch (alg) {                                                                                        \
        case PP_ALG_LEAPFROG: hipLaunchKernelGGL((k_push64<PP_ALG_LEAPFROG, L, U>), g, b, 0, s, p); break;   \
        case PP_ALG_BORIS: hipLaunchKernelGGL((k_push64<PP_ALG_BORIS, L, U>), g, b, 0, s, p); break;         \
        case PP_ALG_RK4: hipLaunchKernelGGL((k_push64<PP_ALG_RK4, L, U>), g, b, 0, s, p); break;             \
        case PP_ALG_YOSHIDA: hipLaunchKernelGGL((k_push64<PP_ALG_YOSHIDA, L, U>), g, b, 0, s, p); break;     \
        case PP_ALG_VAY_REL: hipLaunchKernelGGL((k_push64<PP_ALG_VAY_REL, L, U>), g, b, 0, s, p); break;     \
        case PP_ALG_RK4_REL: hipLaunchKernelGGL((k_push64<PP_ALG_RK4_REL, L, U>), g, b, 0, s, p); break;     \
        case PP_ALG_YOSHIDA_REL: hipLaunchKernelGGL((k_push64<PP_ALG_YOSHIDA_REL, L, U>), g, b, 0, s, p); break; \
        case PP_ALG_BORIS_REL: hipLaunchKernelGGL((k_push64<PP_ALG_BORIS_REL, L, U>), g, b, 0, s, p); break; \
        case PP_ALG_VAY: hipLaunchKernelGGL((k_push64<PP_ALG_VAY, L, U>), g, b, 0, s, p); break;             \
        default: return hipErrorInvalidValue;                                                             \
    }

hipError_t launch_push(int alg, const PushParams& p, int block, hipStream_t s) {
    if (block <= 0) block = 256;
    const dim3 g(grid_for(p.n, block)), b(block);
    const int fields = fields_case(p);
    if (p.mesh) {
        if (fields == kFieldsFusedStatic) PP_STRICT_CASES(true, kFieldsFusedStatic)
        else if (fields == kFieldsFusedRF) PP_STRICT_CASES(true, kFieldsFusedRF)
        else PP_STRICT_CASES(true, kFieldsGeneral)
    } else if (fields == kFieldsUniform) {
        PP_STRICT_CASES(false, kFieldsUniform)
    } else if (fields == kFieldsFusedStatic) {
        PP_STRICT_CASES(false, kFieldsFusedStatic)
    } else if (fields == kFieldsFusedRF) {
        PP_STRICT_CASES(false, kFieldsFusedRF)
    } else {
        PP_STRICT_CASES(false, kFieldsGeneral)
    }
    return hipGetLastError();
}

hipError_t launch_field_eval(const FieldDev& f, const double* pts, int64_t m, double rf, double* out, hipStream_t s) {
    hipLaunchKernelGGL(k_field_eval64, dim3(grid_for(m, 256)), dim3(256), 0, s, f, pts, m, rf, out);
    return hipGetLastError();
}

hipError_t launch_batch_kernel(const BatchArgs& args, hipStream_t s) {
    hipLaunchKernelGGL(k_batch, dim3(grid_for(args.m, 256)), dim3(256), 0, s, args);
    return hipGetLastError();
}

hipError_t launch_util(int which, double* a, double* b, double* c, double* aos, int64_t m, double c2, hipStream_t s) {
    hipLaunchKernelGGL(k_util, dim3(grid_for(m, 256)), dim3(256), 0, s, which, a, b, c, aos, m, c2);
    return hipGetLastError();
}

hipError_t launch_fill(double* p, double v, int64_t n, hipStream_t s) {
    hipLaunchKernelGGL(k_fill, dim3(grid_for(n, 256)), dim3(256), 0, s, p, v, n);
    return hipGetLastError();
}

hipError_t launch_add_i64(int64_t* p, int64_t v, hipStream_t s) {
    hipLaunchKernelGGL(k_add_i64, dim3(1), dim3(64), 0, s, p, v);
    return hipGetLastError();
}

hipError_t launch_build_coef(const MapDesc& d, double* c64, float* c32, hipStream_t s) {
    const int64_t work = (int64_t)d.nc[0] * d.nc[1] * d.nc[2] * d.ncomp;
    hipLaunchKernelGGL(k_build_coef, dim3(grid_for(work, 256)), dim3(256), 0, s, d, c64, c32);
    return hipGetLastError();
}

}  // namespace pp_strict
