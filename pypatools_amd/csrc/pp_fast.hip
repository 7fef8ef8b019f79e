// Fast kernels: FP64 with FMA contraction and the cell-coefficient gather
// (PP_PREC_FP64_FAST), and the FP32-mixed path (PP_PREC_FP32_MIXED):
//   - map stored and interpolated in FP32 (per-cell trilinear coefficients,
//     48 floats = 192 B per cell for 6 components, one base address per gather);
//   - Boris / relativistic Boris / Vay rotations and RK4 derivatives computed as
//     FP32 increments from FP32-rounded state;
//   - position, velocity (or u = gamma v) accumulated in FP64; gamma - 1 from
//     (u^2/c^2) / (1 + gamma) so it keeps full relative precision at low energy.
#include "pp_internal.h"

#define PP_STRICT 0

namespace pp_fast {

#include "pp_fp64.inc"

struct f3 {
    float x, y, z;
};

DEV f3 mkf(float a, float b, float c) { return f3{a, b, c}; }
DEV f3 tof(d3 a) { return f3{(float)a.x, (float)a.y, (float)a.z}; }
DEV d3 tod(f3 a) { return d3{(double)a.x, (double)a.y, (double)a.z}; }
DEV f3 operator+(f3 a, f3 b) { return f3{a.x + b.x, a.y + b.y, a.z + b.z}; }
DEV f3 operator-(f3 a, f3 b) { return f3{a.x - b.x, a.y - b.y, a.z - b.z}; }
DEV f3 operator*(float s, f3 a) { return f3{s * a.x, s * a.y, s * a.z}; }
DEV f3 crossf(f3 a, f3 b) { return f3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
DEV float norm2f(f3 a) { return a.x * a.x + a.y * a.y + a.z * a.z; }
DEV float dotf(f3 a, f3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
DEV d3 addf(d3 a, f3 b) { return d3{a.x + (double)b.x, a.y + (double)b.y, a.z + (double)b.z}; }
DEV float rcpf(float x) { return __builtin_amdgcn_rcpf(x); }

DEV float coef_eval32(const float* __restrict__ q, float tx, float ty, float tz) {
    const float4 lo = *reinterpret_cast<const float4*>(q);      // a b c d
    const float4 hi = *reinterpret_cast<const float4*>(q + 4);  // e f g h
    const float p1 = fmaf(ty, hi.x, lo.y);
    const float p2 = fmaf(tx, hi.w, hi.z);
    const float p3 = fmaf(tx, hi.y, lo.w);
    const float p4 = fmaf(ty, p2, p3);
    const float r = fmaf(ty, lo.z, fmaf(tx, p1, lo.x));
    return fmaf(tz, p4, r);
}

DEV f3 finish_term32(const FieldDev& f, f3 raw, float rfv) {
    f3 v;
    if (f.kind == PP_FIELD_MAP) {
        v = (float)(f.has_scale2 ? f.scale1 * f.scale2 : f.scale1) * raw;
    } else if (f.kind == PP_FIELD_ARRAY) {
        v = f.has_scale2 ? (float)f.scale2 * raw : raw;
    } else {
        d3 c = mk3(f.value[0], f.value[1], f.value[2]);
        if (f.has_scale2) c = f.scale2 * c;
        v = tof(c);
    }
    if (f.has_offset) v = v + mkf((float)f.offset, (float)f.offset, (float)f.offset);
    if (f.rf != PP_RF_NONE) v = rfv * v;
    return v;
}

DEV f3 map_components32(const FieldDev& f, double px, double py, double pz) {
    const FastCell c = fast_locate(f.map, f.map.coef32, px, py, pz);
    if (!c.inside) return mkf((float)f.map.fill, (float)f.map.fill, (float)f.map.fill);
    const float* q = static_cast<const float*>(c.blk) + f.comp0 * 8;
    const float tx = (float)c.tx, ty = (float)c.ty, tz = (float)c.tz;
    return mkf(coef_eval32(q, tx, ty, tz), coef_eval32(q + 8, tx, ty, tz), coef_eval32(q + 16, tx, ty, tz));
}

DEV void eval_fields32(const PushParams& p, int64_t i, double px, double py, double pz, int64_t gs, int stage, f3& E,
                       f3& B) {
    const float rfE = p.E.rf != PP_RF_NONE ? (float)rf_factor(p, gs, stage, p.E.rf) : 1.f;
    const float rfB = p.B.rf != PP_RF_NONE ? (float)rf_factor(p, gs, stage, p.B.rf) : 1.f;
    f3 e0 = mkf(0.f, 0.f, 0.f), b0 = mkf(0.f, 0.f, 0.f);
    if (p.fused) {
        const FastCell c = fast_locate(p.E.map, p.E.map.coef32, px, py, pz);
        const float fl = (float)p.E.map.fill;
        e0 = mkf(fl, fl, fl);
        b0 = e0;
        if (c.inside) {
            const float* q = static_cast<const float*>(c.blk);
            const float tx = (float)c.tx, ty = (float)c.ty, tz = (float)c.tz;
            const float* qe = q + p.E.comp0 * 8;
            const float* qb = q + p.B.comp0 * 8;
            e0 = mkf(coef_eval32(qe, tx, ty, tz), coef_eval32(qe + 8, tx, ty, tz), coef_eval32(qe + 16, tx, ty, tz));
            b0 = mkf(coef_eval32(qb, tx, ty, tz), coef_eval32(qb + 8, tx, ty, tz), coef_eval32(qb + 16, tx, ty, tz));
        }
    } else {
        if (p.E.kind == PP_FIELD_MAP) e0 = map_components32(p.E, px, py, pz);
        else if (p.E.kind == PP_FIELD_ARRAY) e0 = tof(load3(p.E.array, i));
        if (p.B.kind == PP_FIELD_MAP) b0 = map_components32(p.B, px, py, pz);
        else if (p.B.kind == PP_FIELD_ARRAY) b0 = tof(load3(p.B.array, i));
    }
    E = finish_term32(p.E, e0, rfE);
    B = finish_term32(p.B, b0, rfB);
}

// gamma - 1 = w / (1 + sqrt(1 + w)), w = u^2/c^2: no cancellation at low energy.
DEV float gamma_m1_from_u(f3 u, float inv_c2) {
    const float w = norm2f(u) * inv_c2;
    return w * rcpf(1.f + sqrtf(1.f + w));
}

// v = u / gamma with gamma = 1 + (gamma - 1) assembled in FP64.
DEV d3 v_from_u_mixed(d3 u, float inv_c2) {
    const double g = 1.0 + (double)gamma_m1_from_u(tof(u), inv_c2);
    return (1.0 / g) * u;
}

DEV f3 lorentz32(f3 v, f3 E, f3 B, float qm) { return qm * (E + crossf(v, B)); }

DEV f3 rk4_rel_deriv32(f3 v, f3 E, f3 B, float qm, float inv_c2) {
    float beta_sq = norm2f(v) * inv_c2;
    beta_sq = beta_sq >= 0.9999f ? 0.9999f : beta_sq;
    const float ig = sqrtf(1.f - beta_sq);            // 1 / gamma
    const f3 force = qm * (E + crossf(v, B));
    const float vdf = dotf(v, force);
    return ig * force - (vdf * ig * inv_c2) * v;
}

template <int ALG>
DEV void step_mixed(int64_t i, d3& r, d3& v, d3& vrec, double dt, float h, int64_t gs, const PushParams& p) {
    const float inv_c2 = (float)(1.0 / p.c2);
    f3 E, B;
    if constexpr (ALG == PP_ALG_RK4 || ALG == PP_ALG_RK4_REL) {
        constexpr bool REL = (ALG == PP_ALG_RK4_REL);
        const float qm = (float)p.qm;
        auto deriv = [&](d3 vv, f3 ee, f3 bb) {
            if constexpr (REL) return rk4_rel_deriv32(tof(vv), ee, bb, qm, inv_c2);
            else return lorentz32(tof(vv), ee, bb, qm);
        };
        const double hdt = 0.5 * dt;
        eval_fields32(p, i, r.x, r.y, r.z, gs, 0, E, B);
        const f3 k1v = deriv(v, E, B);
        const d3 r2 = r + hdt * v;
        const d3 v2 = v + hdt * tod(k1v);
        eval_fields32(p, i, r2.x, r2.y, r2.z, gs, 1, E, B);
        const f3 k2v = deriv(v2, E, B);
        const d3 r3 = r + hdt * v2;
        const d3 v3 = v + hdt * tod(k2v);
        eval_fields32(p, i, r3.x, r3.y, r3.z, gs, 2, E, B);
        const f3 k3v = deriv(v3, E, B);
        const d3 r4 = r + dt * v3;
        const d3 v4 = v + dt * tod(k3v);
        eval_fields32(p, i, r4.x, r4.y, r4.z, gs, 3, E, B);
        const f3 k4v = deriv(v4, E, B);
        const double w = dt / 6.0;
        r = r + w * (((v + 2.0 * v2) + 2.0 * v3) + v4);
        const f3 ksum = (k1v + 2.f * k2v) + (2.f * k3v + k4v);
        v = v + w * tod(ksum);
        if constexpr (REL) v = clamp_speed(v, p.vclamp);
        vrec = v;
    } else {
        eval_fields32(p, i, r.x, r.y, r.z, gs, 0, E, B);
        if constexpr (ALG == PP_ALG_BORIS) {
            const f3 dE = h * E;
            const d3 vm = addf(v, dE);
            const f3 vmf = tof(vm);
            const f3 t = h * B;
            const f3 s = (2.f * rcpf(1.f + norm2f(t))) * t;
            const f3 vp = vmf + crossf(vmf, t);
            const f3 dv = crossf(vp, s) + dE;
            v = addf(vm, dv);
            r = r + v * dt;
            vrec = v;
        } else if constexpr (ALG == PP_ALG_BORIS_REL) {
            const f3 dE = h * E;
            const d3 um = addf(v, dE);
            const f3 umf = tof(um);
            const float gm = 1.f + gamma_m1_from_u(umf, inv_c2);
            const f3 t = (h * rcpf(gm)) * B;
            const f3 s = (2.f * rcpf(1.f + norm2f(t))) * t;
            const f3 up = umf + crossf(umf, t);
            const f3 du = crossf(up, s) + dE;
            v = addf(um, du);
            vrec = v_from_u_mixed(v, inv_c2);
            r = r + vrec * dt;
        } else if constexpr (ALG == PP_ALG_VAY) {
            const f3 uf = tof(v);
            const float g = 1.f + gamma_m1_from_u(uf, inv_c2);
            const f3 tau = h * B;
            const f3 du1 = (2.f * h) * E + crossf(rcpf(g) * uf, tau);
            const d3 up = addf(v, du1);
            const f3 upf = tof(up);
            const float tau2 = norm2f(tau);
            const float inv_c = (float)(1.0 / p.clight);
            const float ustar = dotf(upf, tau) * inv_c;
            const float sigma = (1.f + norm2f(upf) * inv_c2) - tau2;
            const float gnew = sqrtf(0.5f * (sigma + sqrtf(sigma * sigma + 4.f * (tau2 + ustar * ustar))));
            const f3 t = rcpf(gnew) * tau;
            const float t2 = norm2f(t);
            // u_new - u' = [(u'.t) t + u' x t - t^2 u'] / (1 + t^2)
            const f3 du2 = rcpf(1.f + t2) * ((dotf(upf, t) * t + crossf(upf, t)) - t2 * upf);
            v = addf(up, du2);
            vrec = v_from_u_mixed(v, inv_c2);
            r = r + vrec * dt;
        } else {
            // leapfrog, vay_rel, yoshida(_rel): FP32 map, FP64 integrator arithmetic
            const d3 Ed = tod(E), Bd = tod(B);
            const double qm = p.qm;
            if constexpr (ALG == PP_ALG_LEAPFROG) {
                v = v + lorentz(v, Ed, Bd, qm) * dt;
                r = r + v * dt;
            } else if constexpr (ALG == PP_ALG_VAY_REL) {
                v = vay_rel_kick(v, Ed, Bd, dt, qm, p.c2, p.clight, p.vclamp);
                r = r + v * dt;
            } else {
                constexpr bool REL = (ALG == PP_ALG_YOSHIDA_REL);
#pragma unroll
                for (int st = 0; st < 3; ++st) {
                    d3 acc;
                    if constexpr (REL) acc = rk4_rel_deriv(v, Ed, Bd, qm, p.c2);
                    else acc = lorentz(v, Ed, Bd, qm);
                    v = v + (p.yc[st] * dt) * acc;
                    r = r + (p.yd[st] * dt) * v;
                }
                if constexpr (REL) v = clamp_speed(v, p.vclamp);
            }
            vrec = v;
        }
    }
}

template <int ALG>
DEV d3 kick_mixed(int64_t i, d3 r, d3 v, double dt, int64_t gs, const PushParams& p) {
    f3 E, B;
    eval_fields32(p, i, r.x, r.y, r.z, gs, 0, E, B);
    const d3 Ed = tod(E), Bd = tod(B);
    if constexpr (ALG == PP_ALG_BORIS) return boris_kick(v, Ed, Bd, dt, p.qm);
    else if constexpr (ALG == PP_ALG_BORIS_REL) return boris_rel_kick(v, Ed, Bd, dt, p.qm, p.c2);
    else return vay_kick(v, Ed, Bd, dt, p.qm, p.c2, p.clight);
}

template <int ALG>
__global__ void __launch_bounds__(256) k_push_mixed(const PushParams p) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= p.n) return;
    if (p.alive && !p.alive[i]) return;
    d3 r = mk3(p.x[i], p.y[i], p.z[i]);
    d3 v = mk3(p.vx[i], p.vy[i], p.vz[i]);
    const double dt = p.dt_pp ? p.dt_pp[i] : p.dt;
    if (p.kick_only) {
        if constexpr (ALG == PP_ALG_BORIS || ALG == PP_ALG_BORIS_REL || ALG == PP_ALG_VAY) {
            v = kick_mixed<ALG>(i, r, v, dt, p.step0, p);
            p.vx[i] = v.x;
            p.vy[i] = v.y;
            p.vz[i] = v.z;
        }
        return;
    }
    const float h = (float)(0.5 * p.qm * dt);
    const bool rec_on = p.r_hist != nullptr;
    int64_t rec_left = p.rec_every - (p.step0 % p.rec_every);
    int64_t rec_idx = p.step0 / p.rec_every + 1;
    d3 vrec = v;
    for (int64_t s = 0; s < p.nsteps; ++s) {
        step_mixed<ALG>(i, r, v, vrec, dt, h, p.step0 + s, p);
        if (rec_on && --rec_left == 0) {
            rec_left = p.rec_every;
            write_record(p, rec_idx, i, r, vrec);
            ++rec_idx;
        }
    }
    p.x[i] = r.x;
    p.y[i] = r.y;
    p.z[i] = r.z;
    p.vx[i] = v.x;
    p.vy[i] = v.y;
    p.vz[i] = v.z;
}

__global__ void __launch_bounds__(256) k_field_eval32(const FieldDev f, const double* __restrict__ pts, int64_t m,
                                                       double rfv, double* __restrict__ out) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m) return;
    const double px = pts[3 * i], py = pts[3 * i + 1], pz = pts[3 * i + 2];
    f3 raw = mkf(0.f, 0.f, 0.f);
    if (f.kind == PP_FIELD_MAP) raw = map_components32(f, px, py, pz);
    const f3 v = finish_term32(f, raw, (float)rfv);
    out[3 * i] = v.x;
    out[3 * i + 1] = v.y;
    out[3 * i + 2] = v.z;
}

static inline unsigned grid_for(int64_t n, int block) { return (unsigned)((n + block - 1) / block); }

#define PP_FAST_CASES(KERNEL)                                                                             \
    switch (alg) {                                                                                        \
        case PP_ALG_LEAPFROG: hipLaunchKernelGGL(KERNEL<PP_ALG_LEAPFROG>, g, b, 0, s, p); break;          \
        case PP_ALG_BORIS: hipLaunchKernelGGL(KERNEL<PP_ALG_BORIS>, g, b, 0, s, p); break;                \
        case PP_ALG_RK4: hipLaunchKernelGGL(KERNEL<PP_ALG_RK4>, g, b, 0, s, p); break;                    \
        case PP_ALG_YOSHIDA: hipLaunchKernelGGL(KERNEL<PP_ALG_YOSHIDA>, g, b, 0, s, p); break;            \
        case PP_ALG_VAY_REL: hipLaunchKernelGGL(KERNEL<PP_ALG_VAY_REL>, g, b, 0, s, p); break;            \
        case PP_ALG_RK4_REL: hipLaunchKernelGGL(KERNEL<PP_ALG_RK4_REL>, g, b, 0, s, p); break;            \
        case PP_ALG_YOSHIDA_REL: hipLaunchKernelGGL(KERNEL<PP_ALG_YOSHIDA_REL>, g, b, 0, s, p); break;    \
        case PP_ALG_BORIS_REL: hipLaunchKernelGGL(KERNEL<PP_ALG_BORIS_REL>, g, b, 0, s, p); break;        \
        case PP_ALG_VAY: hipLaunchKernelGGL(KERNEL<PP_ALG_VAY>, g, b, 0, s, p); break;                    \
        default: return hipErrorInvalidValue;                                                             \
    }

hipError_t launch_push(int alg, int prec, const PushParams& p, int block, hipStream_t s) {
    if (block <= 0) block = 256;
    const dim3 g(grid_for(p.n, block)), b(block);
    if (prec == PP_PREC_FP64_FAST) {
        PP_FAST_CASES(k_push64)
    } else {
        PP_FAST_CASES(k_push_mixed)
    }
    return hipGetLastError();
}

hipError_t launch_field_eval(const FieldDev& f, int prec, const double* pts, int64_t m, double rf, double* out,
                             hipStream_t s) {
    if (prec == PP_PREC_FP64_FAST)
        hipLaunchKernelGGL(k_field_eval64, dim3(grid_for(m, 256)), dim3(256), 0, s, f, pts, m, rf, out);
    else
        hipLaunchKernelGGL(k_field_eval32, dim3(grid_for(m, 256)), dim3(256), 0, s, f, pts, m, rf, out);
    return hipGetLastError();
}

}  // namespace pp_fast
