// Fast kernels: FP64 with FMA contraction and the cell-coefficient gather
// (PP_PREC_FP64_FAST), and the FP32-mixed path (PP_PREC_FP32_MIXED):
//   - map stored and interpolated in FP32 (per-cell trilinear coefficients,
//     48 floats = 192 B per cell for 6 components, one base address per gather);
//   - Boris / relativistic Boris / Vay rotations and RK4 derivatives computed as
//     FP32 increments from FP32-rounded state;
//   - position, velocity (or u = gamma v) accumulated in FP64; gamma - 1 from
//     (u^2/c^2) / (1 + gamma) so it keeps full relative precision at low energy.
#include "pp_internal.h"
#include "pp_bvh.inc"

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

// E and B of a fused six-component map at one position (one cell lookup for both).
DEV void fused_components32(const PushParams& p, double px, double py, double pz, f3& e0, f3& b0) {
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
}

DEV void eval_fields32(const PushParams& p, int64_t i, double px, double py, double pz, int64_t gs, int stage, f3& E,
                       f3& B) {
    const float rfE = p.E.rf != PP_RF_NONE ? (float)rf_factor(p, gs, stage, p.E.rf) : 1.f;
    const float rfB = p.B.rf != PP_RF_NONE ? (float)rf_factor(p, gs, stage, p.B.rf) : 1.f;
    f3 e0 = mkf(0.f, 0.f, 0.f), b0 = mkf(0.f, 0.f, 0.f);
    if (p.fused) {
        fused_components32(p, px, py, pz, e0, b0);
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

struct GenericFE32 {
    const PushParams& p;
    DEV void operator()(int64_t i, double x, double y, double z, int64_t gs, int stage, f3& E, f3& B) const {
        eval_fields32(p, i, x, y, z, gs, stage, E, B);
    }
};

// Field evaluators with raw() (kRawE) hand E and B over before their multipliers.
template <class FE, class = void>
struct HasRawE {
    static constexpr bool value = false;
};
template <class FE>
struct HasRawE<FE, decltype((void)FE::kRawE)> {
    static constexpr bool value = FE::kRawE;
};

// PRE: the field evaluator returns h E and h B (h = q/m dt / 2) for the
// Boris-family steps, folding the scale into its field multipliers.
template <int ALG, bool PRE = false, class FE>
DEV void step_mixed(int64_t i, d3& r, d3& v, d3& vrec, double dt, float h, int64_t gs, const FE& fe,
                    const AlgConst& p) {
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
        fe(i, r.x, r.y, r.z, gs, 0, E, B);
        const f3 k1v = deriv(v, E, B);
        const d3 r2 = r + hdt * v;
        const d3 v2 = v + hdt * tod(k1v);
        fe(i, r2.x, r2.y, r2.z, gs, 1, E, B);
        const f3 k2v = deriv(v2, E, B);
        const d3 r3 = r + hdt * v2;
        const d3 v3 = v + hdt * tod(k2v);
        fe(i, r3.x, r3.y, r3.z, gs, 2, E, B);
        const f3 k3v = deriv(v3, E, B);
        const d3 r4 = r + dt * v3;
        const d3 v4 = v + dt * tod(k3v);
        fe(i, r4.x, r4.y, r4.z, gs, 3, E, B);
        const f3 k4v = deriv(v4, E, B);
        const double w = dt / 6.0;
        r = r + w * (((v + 2.0 * v2) + 2.0 * v3) + v4);
        const f3 ksum = (k1v + 2.f * k2v) + (2.f * k3v + k4v);
        v = v + w * tod(ksum);
        if constexpr (REL) v = clamp_speed(v, p.vclamp);
        vrec = v;
    } else if constexpr (ALG == PP_ALG_BORIS && PRE && HasRawE<FE>::value) {
        // as below, with the field multipliers (h times the scales) folded in: E's kE into two FMAs per
        // component; B's kB into the rotation, t = kB b:  s = 2 t / (1 + t^2) = (2 kB / (1 + kB^2 b^2)) b
        // and v- x t = kB (v- x b), so t itself is never formed
        float kE, kB;
        fe.raw(i, r.x, r.y, r.z, gs, 0, E, B, kE, kB);
        const f3 vf = tof(v);
        const f3 vmf = mkf(fmaf(kE, E.x, vf.x), fmaf(kE, E.y, vf.y), fmaf(kE, E.z, vf.z));
        const f3 s = ((2.f * kB) * rcpf(fmaf(kB * kB, norm2f(B), 1.f))) * B;
        const f3 cb = crossf(vmf, B);
        const f3 vp = mkf(fmaf(kB, cb.x, vmf.x), fmaf(kB, cb.y, vmf.y), fmaf(kB, cb.z, vmf.z));
        const f3 c = crossf(vp, s);
        const float k2 = 2.f * kE;
        v = addf(v, mkf(fmaf(k2, E.x, c.x), fmaf(k2, E.y, c.y), fmaf(k2, E.z, c.z)));
        r = r + v * dt;
        vrec = v;
    } else {
        fe(i, r.x, r.y, r.z, gs, 0, E, B);
        if constexpr (ALG == PP_ALG_BORIS) {
            // v_new = v + (2 dE + v' x s) with v' = v- + v- x t, v- = v + dE: one FP64 accumulation
            const f3 dE = PRE ? E : h * E;
            const f3 vmf = tof(v) + dE;
            const f3 t = PRE ? B : h * B;
            const f3 s = (2.f * rcpf(1.f + norm2f(t))) * t;
            const f3 vp = vmf + crossf(vmf, t);
            const f3 dv = crossf(vp, s) + 2.f * dE;
            v = addf(v, dv);
            r = r + v * dt;
            vrec = v;
        } else if constexpr (ALG == PP_ALG_BORIS_REL) {
            const f3 dE = PRE ? E : h * E;
            const d3 um = addf(v, dE);
            const f3 umf = tof(um);
            const float gm = 1.f + gamma_m1_from_u(umf, inv_c2);
            const f3 t = (PRE ? rcpf(gm) : h * rcpf(gm)) * B;
            const f3 s = (2.f * rcpf(1.f + norm2f(t))) * t;
            const f3 up = umf + crossf(umf, t);
            const f3 du = crossf(up, s) + dE;
            v = addf(um, du);
            vrec = v_from_u_mixed(v, inv_c2);
            r = r + vrec * dt;
        } else if constexpr (ALG == PP_ALG_VAY) {
            const f3 uf = tof(v);
            const float g = 1.f + gamma_m1_from_u(uf, inv_c2);
            const f3 tau = PRE ? B : h * B;
            const f3 du1 = (PRE ? 2.f : 2.f * h) * E + crossf(rcpf(g) * uf, tau);
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

template <int ALG, class FE>
DEV d3 kick_mixed(int64_t i, d3 r, d3 v, double dt, int64_t gs, const FE& fe, const AlgConst& p) {
    f3 E, B;
    fe(i, r.x, r.y, r.z, gs, 0, E, B);
    const d3 Ed = tod(E), Bd = tod(B);
    if constexpr (ALG == PP_ALG_BORIS) return boris_kick(v, Ed, Bd, dt, p.qm);
    else if constexpr (ALG == PP_ALG_BORIS_REL) return boris_rel_kick(v, Ed, Bd, dt, p.qm, p.c2);
    else return vay_kick(v, Ed, Bd, dt, p.qm, p.c2, p.clight);
}

// E and B of uniform terms without an RF factor (uniform_term): the eval_fields32 values, taken once.
struct ConstFE32 {
    f3 E0, B0;
    DEV void operator()(int64_t, double, double, double, int64_t, int, f3& E, f3& B) const {
        E = E0;
        B = B0;
    }
};

template <int ALG, bool LOSS, class FE>
DEV void push_mixed_particle(const PushParams& p, int64_t i, const FE& fe) {
    d3 r = mk3(p.x[i], p.y[i], p.z[i]);
    d3 v = mk3(p.vx[i], p.vy[i], p.vz[i]);
    const double dt = p.dt_pp ? p.dt_pp[i] : p.dt;
    const AlgConst ac = alg_const(p);
    if (p.kick_only) {
        if constexpr (ALG == PP_ALG_BORIS || ALG == PP_ALG_BORIS_REL || ALG == PP_ALG_VAY) {
            v = kick_mixed<ALG>(i, r, v, dt, p.step0, fe, ac);
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
        const d3 r0 = r;
        step_mixed<ALG>(i, r, v, vrec, dt, h, p.step0 + s, fe, ac);
        if (rec_on && --rec_left == 0) {
            rec_left = p.rec_every;
            write_record(p, rec_idx, i, r, vrec);
            ++rec_idx;
        }
        if constexpr (LOSS) {
            if (lost_on_segment(p, r0, r)) {
                p.alive[i] = 0;
                break;
            }
        }
    }
    p.x[i] = r.x;
    p.y[i] = r.y;
    p.z[i] = r.z;
    p.vx[i] = v.x;
    p.vy[i] = v.y;
    p.vz[i] = v.z;
}

// E and B of one fused map with no RF factor on either term (eval_fields32 without the other kinds).
struct FusedFE32 {
    const PushParams& p;
    DEV void operator()(int64_t, double x, double y, double z, int64_t, int, f3& E, f3& B) const {
        f3 e0, b0;
        fused_components32(p, x, y, z, e0, b0);
        E = finish_term32(p.E, e0, 1.f);
        B = finish_term32(p.B, b0, 1.f);
    }
};

// The same with the RF factors of the terms (eval_fields32 for a fused map).
struct FusedRFFE32 {
    const PushParams& p;
    DEV void operator()(int64_t, double x, double y, double z, int64_t gs, int stage, f3& E, f3& B) const {
        const float rfE = p.E.rf != PP_RF_NONE ? (float)rf_factor(p, gs, stage, p.E.rf) : 1.f;
        const float rfB = p.B.rf != PP_RF_NONE ? (float)rf_factor(p, gs, stage, p.B.rf) : 1.f;
        f3 e0, b0;
        fused_components32(p, x, y, z, e0, b0);
        E = finish_term32(p.E, e0, rfE);
        B = finish_term32(p.B, b0, rfB);
    }
};

// FIELDS (fields_case): uniform terms are taken once per particle; a fused map keeps
// only its own code in the step loop.
template <int ALG, bool LOSS, int FIELDS = kFieldsGeneral>
__global__ void __launch_bounds__(256) k_push_mixed(const PushParams p) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= p.n) return;
    if (p.alive && !p.alive[i]) return;
    if constexpr (FIELDS == kFieldsUniform) {
        const f3 zero = mkf(0.f, 0.f, 0.f);
        push_mixed_particle<ALG, LOSS>(p, i, ConstFE32{finish_term32(p.E, zero, 1.f), finish_term32(p.B, zero, 1.f)});
    } else if constexpr (FIELDS == kFieldsFusedStatic) {
        push_mixed_particle<ALG, LOSS>(p, i, FusedFE32{p});
    } else if constexpr (FIELDS == kFieldsFusedRF) {
        push_mixed_particle<ALG, LOSS>(p, i, FusedRFFE32{p});
    } else {
        push_mixed_particle<ALG, LOSS>(p, i, GenericFE32{p});
    }
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

// ---------------------------------------------------------------------------
// Hot path: one fused E+B map, uniform grid, no RF/offset.  The map is read
// through a buffer resource with 32-bit offsets (one base in SGPRs, loads past
// the table return 0).
// ---------------------------------------------------------------------------

typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
constexpr int kBufferDword3 = 0x00020000;   // gfx9 raw buffer descriptor word 3

DEV float4 ld4f(__amdgpu_buffer_rsrc_t r, uint32_t off) {
    const u32x4 v = __builtin_amdgcn_raw_buffer_load_b128(r, (int)off, 0, 0);
    return make_float4(__uint_as_float(v.x), __uint_as_float(v.y), __uint_as_float(v.z), __uint_as_float(v.w));
}

DEV double2 ld2d(__amdgpu_buffer_rsrc_t r, uint32_t off) {
    const u32x4 v = __builtin_amdgcn_raw_buffer_load_b128(r, (int)off, 0, 0);
    double2 d;
    d.x = __hiloint2double((int)v.y, (int)v.x);
    d.y = __hiloint2double((int)v.w, (int)v.z);
    return d;
}

// One component from its 8 coefficients (of 1, x, y, z, xy, xz, yz, xyz) in nested form,
// f = (c0 + ty (cy + tz cyz) + tz cz) + tx (cx + ty (cxy + tz cxyz) + tz cxz): 7 FMAs with no
// products of the weights, and only two of them take two coefficients (an FMA reads one scalar
// register, so the scalar-cache path below moves one coefficient of each of those two into a
// vector register).  Every route to the coefficients evaluates this same chain.
DEV float ceval(float4 lo, float4 hi, float tx, float ty, float tz) {
    const float a = fmaf(ty, fmaf(tz, hi.w, hi.x), lo.y);
    const float b = fmaf(tz, hi.y, a);
    const float d = fmaf(tz, lo.w, fmaf(ty, fmaf(tz, hi.z, lo.z), lo.x));
    return fmaf(tx, b, d);
}

DEV double cevald(double2 ab, double2 cd, double2 ef, double2 gh, double tx, double ty, double tz) {
    const double p1 = fma(ty, ef.x, ab.y);
    const double p2 = fma(tx, gh.y, gh.x);
    const double p3 = fma(tx, ef.y, cd.y);
    const double p4 = fma(ty, p2, p3);
    const double r = fma(ty, cd.x, fma(tx, p1, ab.x));
    return fma(tz, p4, r);
}

// cell index (an integer-valued double) and weight on one non-periodic axis, u-space bounds test.
// Only points inside (ok) use the index, and there floor(u) >= 0: it is clamped at the top alone
// (u = nc lies in the last cell).  Inside, u - fu is exact and lies in [0, 1]; the clamp (the
// subtraction's output modifier, no instruction of its own) keeps the weights of points outside
// finite, where the coefficients read as zero.
DEV double hot_axis(double x, double id, double c, int nc, double& t, bool& ok) {
    const double u = fma(x, id, c);
    ok = ok && (u >= 0.0) && (u <= (double)nc);
    const double fu = fmin(floor(u), (double)(nc - 1));
    t = fmin(fmax(u - fu, 0.0), 1.0);
    return fu;
}

// periodic axis: k = floor(u) mod nc, as an integer-valued double.
// A power-of-two nc (wave-uniform test) wraps in FP64: fu / nc = fu * 2^-b is exact, and so are its
// floor and fu - nc floor(fu / nc), which lies in [0, nc) for every |fu| < 2^53; a NaN position
// maps to cell 0 (its weight is NaN).
// Other nc wrap in integer arithmetic: floor(u) is clamped to +-2^29 (NaN goes to -2^29) and offset
// by bias = nc * floor(2^30 / nc) (>= 2^29 for nc <= 2^29) into [0, 2^31); q = mulhi(ku,
// floor((2^32-1)/nc)) is the quotient or one less, so one conditional subtraction leaves k in
// [0, nc) for every input, and k = floor(u) mod nc whenever |floor(u)| <= 2^29.
DEV double hot_axis_periodic(double z, const HotMap& m, double& t) {
    const double u = fma(z, m.id2, m.c2);
    const double fu = floor(u);
    t = u - fu;
    if (m.zmask != 0u) return fmax(fma(-m.nc2d, floor(fu * m.zinv), fu), 0.0);
    const uint32_t nc = (uint32_t)m.nc2;
    const int32_t ki = (int32_t)fmin(fmax(fu, -536870912.0), 536870912.0);
    const uint32_t ku = (uint32_t)ki + m.zbias;
    uint32_t k = ku - __umulhi(ku, m.zmagic) * nc;
    if (k >= nc) k -= nc;
    return (double)k;
}

// The hot paths take coefficient tables below 2^31 bytes (192 or 384 per cell), so every cell index is
// below 2^24: the index i nc1 nc2 + j nc2 + k is exact as an FP64 fma chain on the integer-valued axis
// indices, and the byte offsets cell * 192 (or 384) are exact as 24-bit multiplies (full rate, where
// 32-bit integer multiplies run at quarter rate).
template <bool PERIODIC>
DEV uint32_t hot_locate(const HotMap& m, double x, double y, double z, double& tx, double& ty, double& tz,
                        bool& ok) {
    ok = true;
    const double i = hot_axis(x, m.id0, m.c0, m.nc0, tx, ok);
    const double j = hot_axis(y, m.id1, m.c1, m.nc1, ty, ok);
    double k;
    if constexpr (PERIODIC) k = hot_axis_periodic(z, m, tz);
    else k = hot_axis(z, m.id2, m.c2, m.nc2, tz, ok);
    return (uint32_t)fma(i, m.nc12d, fma(j, m.nc2d, k));
}

// LDS staging of a wavefront's cells: up to kStageSlots distinct cells per gather,
// each cell as PIECES x 16 B (12 for the FP32 table, 24 for FP64), padded to
// PIECES + 1 so that slots start in different LDS banks.
// 16 slots cover the distinct cells of nearly every wavefront of a cell-sorted C2 bunch up to the next
// sort 500 steps later (a CPU replay of the bunch: 2.2 distinct cells per wavefront one step after the
// sort, 7.4 after 250 steps, 9.5 after 500 steps with 35 % of wavefronts above 10 and the 90th
// percentile at 13); a wavefront over more cells gathers per lane.
constexpr int kStageSlots = 16;
constexpr int kStageWaves = 4;      // wavefronts per block (the hot kernels launch <= 256 threads)

// Staging roles of the lanes of a full wavefront, fixed for the whole launch.  In each
// pass, lane l loads piece l % PIECES of the cell in slot l / PIECES (CPP = 64 / PIECES
// cells per pass; lanes beyond CPP * PIECES stage nothing).
template <int PIECES>
struct StageLanes {
    static constexpr int kCPP = 64 / PIECES;
    static constexpr int kPitch = PIECES + 1;   // 16 B units per slot
    uint32_t slot;   // l / PIECES, 64 for lanes that stage nothing
    uint32_t src;    // byte offset of the lane's piece within its cell
    uint32_t dst;    // 16 B unit of the lane's piece in the wavefront's LDS area, pass 0
};

template <int PIECES>
DEV StageLanes<PIECES> stage_lanes() {
    const uint32_t lane = __builtin_amdgcn_mbcnt_hi(~0u, __builtin_amdgcn_mbcnt_lo(~0u, 0u));
    const uint32_t sl = lane / PIECES, part = lane - sl * PIECES;
    StageLanes<PIECES> s;
    s.slot = sl < (uint32_t)StageLanes<PIECES>::kCPP ? sl : 64u;
    s.src = part * 16u;
    s.dst = sl * StageLanes<PIECES>::kPitch + part;
    return s;
}

#if defined(__HIP_DEVICE_COMPILE__)
// The distinct cells of a full wavefront, at most kStageSlots of them: each lane's slot in
// `myslot`.  Returns their number, or 0 when the wavefront spans more cells than that.  The scan
// is unrolled at compile time (template recursion): per distinct cell one readlane, one compare,
// one select, and scalar mask updates with no loop counter (measured: the rolled loop ran C2
// at 1.57e11 particle-steps/s, this form at 1.89e11).
template <int IT>
DEV void distinct_step(uint32_t cell, uint64_t& todo, uint32_t& myslot, int& ns) {
    if constexpr (IT < kStageSlots) {
        if (todo == 0) return;
        const uint32_t c = __builtin_amdgcn_readlane(cell, (int)__builtin_ctzll(todo));
        const bool other = cell != c;
        todo &= __builtin_amdgcn_ballot_w64(other);
        myslot = other ? myslot : (uint32_t)IT;
        ns = IT + 1;
        distinct_step<IT + 1>(cell, todo, myslot, ns);
    }
}

DEV int distinct_cells(uint32_t cell, uint32_t& myslot) {
    uint64_t todo = ~0ull;
    int ns = 0;
    distinct_step<0>(cell, todo, myslot, ns);
    return todo == 0 ? ns : 0;
}

// Loads the wavefront's ns distinct cells (cell_bytes each) into its LDS slots: every lane
// writes its cell to its slot's entry of the wavefront's cell list (lanes of one slot write the
// same value), then each staging lane reads the cell of its slot from the list and loads its
// piece.  No block barrier is needed: the slots and the list are private to the wavefront and
// its LDS operations execute in order (the wavefront fence and wave_barrier keep the compiler
// from reordering them).
template <int PIECES>
DEV void stage_cells(float4* wv, uint32_t* list, __amdgpu_buffer_rsrc_t rs, uint32_t cell_bytes, int ns,
                     uint32_t cell, uint32_t myslot, const StageLanes<PIECES>& L) {
    constexpr int kCPP = StageLanes<PIECES>::kCPP;
    constexpr int kPitch = StageLanes<PIECES>::kPitch;
    list[myslot] = cell;
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
#pragma unroll
    for (int p = 0; p * kCPP < kStageSlots; ++p) {
        if (p * kCPP >= ns) break;
        if (L.slot + (uint32_t)(kCPP * p) < (uint32_t)ns) {
            const uint32_t c = list[L.slot + kCPP * p];
            wv[L.dst + kCPP * kPitch * p] = ld4f(rs, __umul24(c, cell_bytes) + L.src);
        }
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// The wavefront's LDS slots and cell list.  One area per element size in each kernel, shared by the
// gather functors a kernel instantiates (a staging with reuse and the kicks outside its step loop
// never interleave).
DEV float4* stage_area32() {
    __shared__ float4 lds[kStageWaves][kStageSlots * StageLanes<12>::kPitch];
    return lds[threadIdx.x >> 6];
}
DEV float4* stage_area64() {
    __shared__ float4 lds[kStageWaves][kStageSlots * StageLanes<24>::kPitch];
    return lds[threadIdx.x >> 6];
}
DEV uint32_t* stage_list() {
    __shared__ uint32_t list[kStageWaves][kStageSlots];
    return list[threadIdx.x >> 6];
}

// Constant-address-space view of the coefficient table: loads at a wave-uniform
// address become scalar loads (s_load_dwordx16) into SGPRs.
typedef const __attribute__((address_space(4))) float* cfloat_p;

DEV float ceval_s(cfloat_p q, float tx, float ty, float tz) {
    const float a = fmaf(ty, fmaf(tz, q[7], q[4]), q[1]);
    const float b = fmaf(tz, q[5], a);
    const float d = fmaf(tz, q[3], fmaf(ty, fmaf(tz, q[6], q[2]), q[0]));
    return fmaf(tx, b, d);
}
#endif

DEV double2 as_d2(float4 v) {
    double2 d;
    d.x = __hiloint2double((int)__float_as_uint(v.y), (int)__float_as_uint(v.x));
    d.y = __hiloint2double((int)__float_as_uint(v.w), (int)__float_as_uint(v.z));
    return d;
}

// Wave issue priority of the hot gathers: raised from the cell lookup until the cell data is staged (or
// the gather is done), so that wavefronts about to issue their staging loads and scans go ahead of those in
// their arithmetic and the loads of more wavefronts overlap.  Measured on C2: 2.19e11 particle-steps/s
// against 2.11e11 with every wavefront at the same priority; priorities 1, 2 and 3, and dropping it only
// after the evaluation, ran alike.
DEV void gather_priority(bool raised) {
#if defined(__HIP_DEVICE_COMPILE__)
    if (raised) __builtin_amdgcn_s_setprio(1);
    else __builtin_amdgcn_s_setprio(0);
#else
    (void)raised;
#endif
}

// Field multipliers of the FP32 hot gather: E = kE * e and B = kB * b.  kE, kB are the
// field scales, times h = q/m dt / 2 for the Boris-family steps (step_mixed<ALG, true>).
struct HotScale {
    float kE, kB;
};

DEV HotScale hot_scale(const HotMap& m, double h) {
    HotScale s;
    s.kE = (float)(m.sE * h);
    s.kB = (float)(m.sB * h);
    return s;
}

// Points outside the map gather from a cell index past the end of the table: its buffer loads
// return 0, which is the field there (the hot kernels take only maps that fill 0 outside).  The
// byte offsets of these indices (24-bit products) exceed every table the hot paths accept
// (below 2^31 bytes: fewer than 0xFFFFFF cells of 192 B, 0x7FFFFF of 384 B).
constexpr uint32_t kOutside32 = 0xFFFFFFu;
constexpr uint32_t kOutside64 = 0x7FFFFFu;

// RF factor of one field at global step gs and RK stage (wave-uniform: a scalar load)
DEV double hot_rf(const HotRF& rf, int mode, int64_t gs, int stage) {
    return rf.table[(gs - rf.step0) * 8 + stage * 2 + (mode == PP_RF_SIN ? 1 : 0)];
}

// REUSE (the RK4 kernels, whose four stages and successive steps mostly stay in the same cells): the
// wavefront remembers the cells it staged through LDS last; while every lane is still in its cell, the
// coefficients are read from the same slots without a new staging (same values, same result).
template <bool PERIODIC, bool RF = false, bool REUSE = false>
struct HotFE32 {
    const HotMap& m;
    __amdgpu_buffer_rsrc_t rs;
    HotScale k;
    HotRF rf;
    StageLanes<12> lanes = stage_lanes<12>();
    mutable uint32_t last_cell = 0;   // REUSE: this lane's cell and LDS slot at the last staging
    mutable uint32_t last_slot = 0;
    mutable bool staged = false;      // REUSE: the wavefront has staged at least once (wave-uniform)
    static constexpr bool kRawE = true;   // raw(): E and B before their multipliers (step_mixed<BORIS> folds them)

    // field multipliers at global step gs and RK stage (RF factors applied)
    DEV void scales(int64_t gs, int stage, float& kE, float& kB) const {
        kE = k.kE;
        kB = k.kB;
        if constexpr (RF) {
            if (rf.e != PP_RF_NONE) kE *= (float)hot_rf(rf, rf.e, gs, stage);
            if (rf.b != PP_RF_NONE) kB *= (float)hot_rf(rf, rf.b, gs, stage);
        }
    }

    DEV void operator()(int64_t, double x, double y, double z, int64_t gs, int stage, f3& E, f3& B) const {
        float v[6], kE, kB;
        gather(x, y, z, v);
        scales(gs, stage, kE, kB);
        E = mkf(kE * v[0], kE * v[1], kE * v[2]);
        B = mkf(kB * v[3], kB * v[4], kB * v[5]);
    }

    // E and B unscaled, with their multipliers kE and kB
    DEV void raw(int64_t, double x, double y, double z, int64_t gs, int stage, f3& E, f3& B, float& kE,
                 float& kB) const {
        float v[6];
        gather(x, y, z, v);
        scales(gs, stage, kE, kB);
        E = mkf(v[0], v[1], v[2]);
        B = mkf(v[3], v[4], v[5]);
    }

    // the six map components (E then B) at a point, unscaled
    DEV void gather(double x, double y, double z, float (&v)[6]) const {
        gather_priority(true);
        double tx, ty, tz;
        bool ok;
        uint32_t cell = hot_locate<PERIODIC>(m, x, y, z, tx, ty, tz, ok);
        cell = ok ? cell : kOutside32;
        const float wx = (float)tx, wy = (float)ty, wz = (float)tz;
        bool done = false;
#if defined(__HIP_DEVICE_COMPILE__)
        constexpr int kPitch = StageLanes<12>::kPitch;
        float4* wv = stage_area32();
        if constexpr (REUSE) {
            if (staged && __builtin_amdgcn_ballot_w64(cell != last_cell) == 0) {
                const float4* q = wv + __umul24(last_slot, (uint32_t)kPitch);
#pragma unroll
                for (int c = 0; c < 6; ++c) v[c] = ceval(q[2 * c], q[2 * c + 1], wx, wy, wz);
                done = true;
            }
        }
        // Three ways to the cell coefficients, chosen per wavefront:
        //  - all lanes in one cell (common right after a cell sort): the cell's
        //    192 B are read once through the scalar cache;
        //  - a full wavefront over at most kStageSlots distinct cells: each
        //    distinct cell is loaded once (12 lanes x 16 B) into the wavefront's
        //    LDS slots and every lane reads its cell from LDS;
        //  - otherwise every lane gathers its cell through the vector memory path.
        // All three evaluate the same FMA chain on the same values.
        const uint32_t c0 = __builtin_amdgcn_readfirstlane(cell);
        if (done) {
        } else if (__builtin_amdgcn_ballot_w64(cell != c0) == 0 && c0 < m.ncells) {
            const cfloat_p q = (cfloat_p)m.coef + (size_t)c0 * 48u;
#pragma unroll
            for (int c = 0; c < 6; ++c) v[c] = ceval_s(q + c * 8, wx, wy, wz);
            done = true;
        } else if (__builtin_amdgcn_ballot_w64(true) == ~0ull) {
            uint32_t myslot = 0;
            const int ns = distinct_cells(cell, myslot);
            if (ns > 0) {
                stage_cells<12>(wv, stage_list(), rs, 192u, ns, cell, myslot, lanes);
                gather_priority(false);
                const float4* q = wv + __umul24(myslot, (uint32_t)kPitch);
#pragma unroll
                for (int c = 0; c < 6; ++c) v[c] = ceval(q[2 * c], q[2 * c + 1], wx, wy, wz);
                __builtin_amdgcn_wave_barrier();
                done = true;
                if constexpr (REUSE) {
                    last_cell = cell;
                    last_slot = myslot;
                    staged = true;
                }
            }
        }
#endif
        if (!done) {
            const uint32_t base = __umul24(cell, 48u * 4u);
#pragma unroll
            for (int c = 0; c < 6; ++c)
                v[c] = ceval(ld4f(rs, base + c * 32u), ld4f(rs, base + c * 32u + 16u), wx, wy, wz);
        }
        gather_priority(false);
    }
};

// FP64-fast gather: 384 B cells ([comp][8 doubles]).  A full wavefront over at most
// kStageSlots distinct cells stages them through LDS (24 lanes x 16 B per cell, two cells
// per load instruction); other wavefronts gather per lane.  Both read the same doubles.
template <bool PERIODIC, bool RF = false, bool REUSE = false>
struct HotFE64 {
    const HotMap& m;
    __amdgpu_buffer_rsrc_t rs;
    HotRF rf;
    StageLanes<24> lanes = stage_lanes<24>();
    mutable uint32_t last_cell = 0;
    mutable uint32_t last_slot = 0;
    mutable bool staged = false;
    DEV void operator()(int64_t, double x, double y, double z, int64_t gs, int stage, d3& E, d3& B) const {
        gather_priority(true);
        double tx, ty, tz;
        bool ok;
        uint32_t cell = hot_locate<PERIODIC>(m, x, y, z, tx, ty, tz, ok);
        cell = ok ? cell : kOutside64;
        double v[6];
        bool done = false;
#if defined(__HIP_DEVICE_COMPILE__)
        constexpr int kPitch = StageLanes<24>::kPitch;
        float4* wv = stage_area64();
        uint32_t slot = 0;
        if constexpr (REUSE) {
            if (staged && __builtin_amdgcn_ballot_w64(cell != last_cell) == 0) {
                slot = last_slot;
                done = true;
            }
        }
        if (!done && __builtin_amdgcn_ballot_w64(true) == ~0ull) {
            const int ns = distinct_cells(cell, slot);
            if (ns > 0) {
                stage_cells<24>(wv, stage_list(), rs, 384u, ns, cell, slot, lanes);
                gather_priority(false);
                done = true;
                if constexpr (REUSE) {
                    last_cell = cell;
                    last_slot = slot;
                    staged = true;
                }
            }
        }
        if (done) {
            const float4* q = wv + __umul24(slot, (uint32_t)kPitch);
#pragma unroll
            for (int c = 0; c < 6; ++c)
                v[c] = cevald(as_d2(q[4 * c]), as_d2(q[4 * c + 1]), as_d2(q[4 * c + 2]), as_d2(q[4 * c + 3]), tx, ty,
                              tz);
            __builtin_amdgcn_wave_barrier();
        }
#endif
        if (!done) {
            const uint32_t base = __umul24(cell, 48u * 8u);
#pragma unroll
            for (int c = 0; c < 6; ++c) {
                const uint32_t o = base + c * 64u;
                v[c] = cevald(ld2d(rs, o), ld2d(rs, o + 16u), ld2d(rs, o + 32u), ld2d(rs, o + 48u), tx, ty, tz);
            }
        }
        gather_priority(false);
        double sE = m.sE, sB = m.sB;
        if constexpr (RF) {
            if (rf.e != PP_RF_NONE) sE *= hot_rf(rf, rf.e, gs, stage);
            if (rf.b != PP_RF_NONE) sB *= hot_rf(rf, rf.b, gs, stage);
        }
        E = mk3(sE * v[0], sE * v[1], sE * v[2]);
        B = mk3(sB * v[3], sB * v[4], sB * v[5]);
    }
};

DEV AlgConst hot_const(const HotParams& p) {
    AlgConst a;
    a.qm = p.qm;
    a.c2 = p.c2;
    a.clight = p.clight;
    a.vclamp = p.vclamp;
    for (int k = 0; k < 3; ++k) {
        a.yc[k] = 0.0;
        a.yd[k] = 0.0;
    }
    return a;
}

// Boris-family steps take h E and h B straight from the FP32 hot gather
template <int ALG>
constexpr bool kPrescaled = ALG == PP_ALG_BORIS || ALG == PP_ALG_BORIS_REL || ALG == PP_ALG_VAY;

// 4 wavefronts per SIMD (<= 128 VGPRs) keeps enough loads and LDS reads in flight;
// the compiler would otherwise spend registers for a deeper load pipeline at 3.
// Four wavefronts per SIMD also for the RK4 steps: measured on C3, the spills this costs are cheaper than
// halving the occupancy (2 wavefronts per SIMD ran 1.6x slower without staging reuse).
// Wavefronts per SIMD the register allocation of k_hot aims for (512 VGPRs per SIMD lane): 6 for the
// single-gather steps (80 VGPRs; C2 measured 2.06e11 particle-steps/s at 6, 1.95e11 at 5, 1.93e11 at 4,
// where the 4-wavefront allocation took 116 VGPRs), 4 for the RK4 steps (C3 the same at 4, 5 and 6: their
// spills grow with the occupancy).
template <int ALG>
constexpr int kHotWaves = (ALG == PP_ALG_RK4 || ALG == PP_ALG_RK4_REL) ? 4 : 6;
// The loss builds without a mesh hold the frozen state as well: 5 wavefronts per SIMD (88 VGPRs, no
// spills) for the FP32-mixed single-gather steps, 4 for FP64-fast and the RK4 steps.  The mesh builds
// (k_hot_mesh) take 4 everywhere: at 5 the FP32-mixed one spilled 20 VGPRs (C2 bunch, 10 mm pipe 8.4e10
// particle-steps/s against 8.7e10 at 4; 3 mm pipe 3.85e10 at both).
template <int ALG, bool MIXED, bool LOSS>
constexpr int kHotLossWaves = !LOSS ? kHotWaves<ALG> : (MIXED && kHotWaves<ALG> > 4 ? 5 : 4);
constexpr int kHotMeshWaves = 4;

// LOSS: the state carries an alive mask (p.alive); MESH (k_hot_mesh): every step's segment is also tested
// against the electrode mesh p.mesh inside the step loop (pp_push_lossy).  A particle dead at the start of the launch, or
// lost in it, keeps its wavefront's lane: it repeats the step from its frozen state so the wavefront stays
// full for the staged gathers, but it writes nothing; a particle lost in a step stores that step's state and
// its record, as the generic kernels do (the reference marks it inactive after the step, pusher.py:982-990).
// Every launch with an alive mask takes a loss build, k_hot_mesh with a mesh and k_hot<..., true> without:
// both run the same step arithmetic, so per-step launches tested on the host and fused launches agree bit
// for bit, and the mesh test's registers stay out of the build without it.
template <int ALG, bool MIXED, bool PERIODIC, bool RF, bool LOSS, bool MESH>
DEV void hot_body(const HotParams& p) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= p.n) return;
    bool live = true;
    if constexpr (LOSS) live = p.alive[i] != 0;
    const __amdgpu_buffer_rsrc_t rs =
        __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(p.m.coef), (short)0, (int)p.m.bytes, kBufferDword3);
    const AlgConst ac = hot_const(p);
    d3 r = mk3(p.x[i], p.y[i], p.z[i]);
    d3 v = mk3(p.vx[i], p.vy[i], p.vz[i]);
    const double dt = p.dt;
    if (p.kick_only) {
        if constexpr (ALG == PP_ALG_BORIS || ALG == PP_ALG_BORIS_REL || ALG == PP_ALG_VAY) {
            if (!live) return;
            if constexpr (MIXED)
                v = kick_mixed<ALG>(i, r, v, dt, p.step0, HotFE32<PERIODIC, RF>{p.m, rs, hot_scale(p.m, 1.0), p.rf},
                                    ac);
            else v = kick64<ALG>(i, r, v, dt, p.step0, HotFE64<PERIODIC, RF>{p.m, rs, p.rf}, ac);
            p.vx[i] = v.x;
            p.vy[i] = v.y;
            p.vz[i] = v.z;
        }
        return;
    }
    const float h = (float)(0.5 * p.qm * dt);
    constexpr bool PRE = kPrescaled<ALG>;
    constexpr bool REUSE = ALG == PP_ALG_RK4 || ALG == PP_ALG_RK4_REL;
    const HotFE32<PERIODIC, RF, REUSE> fe32{p.m, rs, hot_scale(p.m, PRE ? 0.5 * p.qm * dt : 1.0), p.rf};
    const HotFE64<PERIODIC, RF, REUSE> fe64{p.m, rs, p.rf};
    const bool rec_on = p.r_hist != nullptr;
    int64_t rec_left = p.rec_every - (p.step0 % p.rec_every);
    int64_t rec_idx = p.step0 / p.rec_every + 1;
    d3 vrec = v;
    const pp_bvh::MeshDev* mesh = static_cast<const pp_bvh::MeshDev*>(p.mesh);
    for (int64_t s = 0; s < p.nsteps; ++s) {
        const d3 r0 = r, v0 = v;
        pp_bvh::ClearProbe probe{0, false, false};   // read before the step so its latency is hidden
        if constexpr (MESH) probe = pp_bvh::mesh_clearance(mesh, r.x, r.y, r.z);
        if constexpr (MIXED) step_mixed<ALG, PRE>(i, r, v, vrec, dt, h, p.step0 + s, fe32, ac);
        else step64<ALG>(i, r, v, vrec, dt, p.step0 + s, fe64, ac);
        if constexpr (LOSS) {
            if (!live) {
                r = r0;
                v = v0;
            }
        }
        if (rec_on && --rec_left == 0) {
            rec_left = p.rec_every;
            if (live) write_record_to(p.r_hist, p.v_hist, p.hist_rows, p.n_records, p.ids, rec_idx, i, r, vrec);
            ++rec_idx;
        }
        if constexpr (MESH) {
            if (live && pp_bvh::mesh_hits_fetched(mesh, probe, r0.x, r0.y, r0.z, r.x, r.y, r.z)) {
                live = false;
                p.alive[i] = 0;
                p.x[i] = r.x;
                p.y[i] = r.y;
                p.z[i] = r.z;
                p.vx[i] = v.x;
                p.vy[i] = v.y;
                p.vz[i] = v.z;
            }
        }
    }
    if (live) {
        p.x[i] = r.x;
        p.y[i] = r.y;
        p.z[i] = r.z;
        p.vx[i] = v.x;
        p.vy[i] = v.y;
        p.vz[i] = v.z;
    }
}

template <int ALG, bool MIXED, bool PERIODIC, bool RF, bool LOSS>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(kHotLossWaves<ALG, MIXED, LOSS>)))
k_hot(const HotParams p) {
    hot_body<ALG, MIXED, PERIODIC, RF, LOSS, false>(p);
}

template <int ALG, bool MIXED, bool PERIODIC>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(kHotMeshWaves)))
k_hot_mesh(const HotParams p) {
    hot_body<ALG, MIXED, PERIODIC, false, true, true>(p);
}

// ---------------------------------------------------------------------------
// Batched push: one block = one tile of one bunch; the bunch's algorithm,
// q/m, dt, step count, field scales and map come from a device table (uniform
// per block, so the branches below are scalar).
// ---------------------------------------------------------------------------

// Half-step velocity kick of a staggered bunch (the kick_only launches of the single-bunch path).
template <int ALG, bool MIXED, bool PERIODIC>
DEV d3 batch_kick(int64_t i, d3 r, d3 v, double dtk, const HotMap& m, __amdgpu_buffer_rsrc_t rs,
                  const AlgConst& ac) {
    if constexpr (MIXED) return kick_mixed<ALG>(i, r, v, dtk, 0, HotFE32<PERIODIC>{m, rs, hot_scale(m, 1.0), HotRF{}}, ac);
    else return kick64<ALG>(i, r, v, dtk, 0, HotFE64<PERIODIC>{m, rs, HotRF{}}, ac);
}

// ns steps from step0; kick_start / kick_end: the staggering half kicks before the first and after
// the last step (pp_bunch_flags), for the staggered algorithms only.
template <int ALG, bool MIXED, bool PERIODIC>
DEV void batch_steps(int64_t i, d3& r, d3& v, int64_t step0, int64_t ns, double dt, const HotMap& m,
                     __amdgpu_buffer_rsrc_t rs, const AlgConst& ac, bool kick_start, bool kick_end) {
    constexpr bool STAGGERED = ALG == PP_ALG_BORIS || ALG == PP_ALG_BORIS_REL || ALG == PP_ALG_VAY;
    if constexpr (STAGGERED) {
        if (kick_start) v = batch_kick<ALG, MIXED, PERIODIC>(i, r, v, -0.5 * dt, m, rs, ac);
    }
    d3 vrec = v;
    const float h = (float)(0.5 * ac.qm * dt);
    constexpr bool PRE = kPrescaled<ALG>;
    constexpr bool REUSE = ALG == PP_ALG_RK4 || ALG == PP_ALG_RK4_REL;
    const HotFE32<PERIODIC, false, REUSE> fe32{m, rs, hot_scale(m, PRE ? 0.5 * ac.qm * dt : 1.0), HotRF{}};
    const HotFE64<PERIODIC, false, REUSE> fe64{m, rs, HotRF{}};
    for (int64_t s = 0; s < ns; ++s) {
        if constexpr (MIXED) step_mixed<ALG, PRE>(i, r, v, vrec, dt, h, step0 + s, fe32, ac);
        else step64<ALG>(i, r, v, vrec, dt, step0 + s, fe64, ac);
    }
    if constexpr (STAGGERED) {
        if (kick_end) v = batch_kick<ALG, MIXED, PERIODIC>(i, r, v, 0.5 * dt, m, rs, ac);
    }
}

// Kernel per tile class (pp_tile_class): the single-gather algorithms alone fit 7 wavefronts per SIMD in
// FP32-mixed (72 VGPRs with no scratch access in the step loops; 6 in FP64-fast); a kernel holding the RK4
// steps keeps 4.  Measured on C4 (Boris and Vay bunches only): 1.62e11 particle-steps/s with every algorithm
// in one kernel at 4, 1.90e11 at 6 (whose allocation spilled inside the step loops), 2.03e11 at 7, 1.98e11
// at 8; the C5 mix with its RK4 jobs ran worse with the RK4 steps at 6.
template <bool MIXED, int CLASS>
constexpr int kBatchWaves = CLASS != PP_TILES_SINGLE_GATHER ? 4 : (MIXED ? 7 : 6);

template <int CLASS>
constexpr bool batch_class_has(int alg) {
    return CLASS == PP_TILES_ALL ||
           (CLASS == PP_TILES_RK4) == (alg == PP_ALG_RK4 || alg == PP_ALG_RK4_REL);
}

template <bool MIXED, bool PERIODIC, int CLASS>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(kBatchWaves<MIXED, CLASS>)))
k_batched(const BatchParams p) {
    const int64_t t = blockIdx.x;
    if (t >= p.ntiles) return;
    const int32_t b = p.tile_bunch[t];
    if (b < 0 || b >= p.nbunch) return;
    const pp_bunch bd = p.bunches[b];
    if (bd.map_index < 0 || bd.map_index >= p.nmaps) return;
    const int64_t i = p.tile_start[t] + threadIdx.x;
    if (i < bd.offset || i >= bd.offset + bd.count || i >= p.n) return;
    const int64_t step0 = p.step0_dev ? *p.step0_dev : p.step0;
    const int64_t end = bd.nsteps < step0 + p.nsteps ? (int64_t)bd.nsteps : step0 + p.nsteps;
    const int64_t ns = end > step0 ? end - step0 : 0;
    // staggering kicks: in the window holding the bunch's step 0, and in the one reaching its last step
    // (a bunch of 0 steps takes both in the first window, as track_batch does with nsteps = 0)
    const bool kick_start = (bd.flags & PP_BUNCH_KICK_START) && step0 == 0;
    const bool kick_end = (bd.flags & PP_BUNCH_KICK_END) && (ns > 0 ? end == bd.nsteps : step0 == 0 && bd.nsteps == 0);
    if (ns == 0 && !kick_start && !kick_end) return;
    HotMap m = p.maps[bd.map_index];
    m.sE *= bd.escale;
    m.sB *= bd.bscale;
    const __amdgpu_buffer_rsrc_t rs =
        __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(m.coef), (short)0, (int)m.bytes, kBufferDword3);
    AlgConst ac;
    ac.qm = bd.q_over_m;
    ac.c2 = p.c2;
    ac.clight = p.clight;
    ac.vclamp = p.vclamp;
    for (int k = 0; k < 3; ++k) {
        ac.yc[k] = 0.0;
        ac.yd[k] = 0.0;
    }
    d3 r = mk3(p.x[i], p.y[i], p.z[i]);
    d3 v = mk3(p.vx[i], p.vy[i], p.vz[i]);
    switch (bd.algorithm) {
        case PP_ALG_BORIS:
            if constexpr (batch_class_has<CLASS>(PP_ALG_BORIS))
                batch_steps<PP_ALG_BORIS, MIXED, PERIODIC>(i, r, v, step0, ns, bd.dt, m, rs, ac, kick_start, kick_end);
            else
                return;
            break;
        case PP_ALG_VAY:
            if constexpr (batch_class_has<CLASS>(PP_ALG_VAY))
                batch_steps<PP_ALG_VAY, MIXED, PERIODIC>(i, r, v, step0, ns, bd.dt, m, rs, ac, kick_start, kick_end);
            else
                return;
            break;
        case PP_ALG_VAY_REL:
            if constexpr (batch_class_has<CLASS>(PP_ALG_VAY_REL))
                batch_steps<PP_ALG_VAY_REL, MIXED, PERIODIC>(i, r, v, step0, ns, bd.dt, m, rs, ac, kick_start, kick_end);
            else
                return;
            break;
        case PP_ALG_BORIS_REL:
            if constexpr (batch_class_has<CLASS>(PP_ALG_BORIS_REL))
                batch_steps<PP_ALG_BORIS_REL, MIXED, PERIODIC>(i, r, v, step0, ns, bd.dt, m, rs, ac, kick_start, kick_end);
            else
                return;
            break;
        case PP_ALG_RK4:
            if constexpr (batch_class_has<CLASS>(PP_ALG_RK4))
                batch_steps<PP_ALG_RK4, MIXED, PERIODIC>(i, r, v, step0, ns, bd.dt, m, rs, ac, kick_start, kick_end);
            else
                return;
            break;
        case PP_ALG_RK4_REL:
            if constexpr (batch_class_has<CLASS>(PP_ALG_RK4_REL))
                batch_steps<PP_ALG_RK4_REL, MIXED, PERIODIC>(i, r, v, step0, ns, bd.dt, m, rs, ac, kick_start, kick_end);
            else
                return;
            break;
        case PP_ALG_LEAPFROG:
            if constexpr (batch_class_has<CLASS>(PP_ALG_LEAPFROG))
                batch_steps<PP_ALG_LEAPFROG, MIXED, PERIODIC>(i, r, v, step0, ns, bd.dt, m, rs, ac, kick_start, kick_end);
            else
                return;
            break;
        default: return;
    }
    p.x[i] = r.x;
    p.y[i] = r.y;
    p.z[i] = r.z;
    p.vx[i] = v.x;
    p.vy[i] = v.y;
    p.vz[i] = v.z;
}

// ---------------------------------------------------------------------------
// Gather-bandwidth probe: the hot kernels' 12 x 16 B buffer loads per cell with
// no push arithmetic, to measure the rate the vector memory path sustains for
// wave-uniform (mode 0) and per-lane (mode 1) cells.
// ---------------------------------------------------------------------------

__global__ void __launch_bounds__(256) k_gather_probe(const float* table, uint32_t bytes, uint32_t mask, int mode,
                                                      int iters, int64_t nthreads, float* __restrict__ out) {
    const int64_t gid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (gid >= nthreads) return;
    const __amdgpu_buffer_rsrc_t rs =
        __builtin_amdgcn_make_buffer_rsrc(const_cast<float*>(table), (short)0, (int)bytes, kBufferDword3);
    uint32_t h = (uint32_t)(mode == 0 ? (gid >> 6) : gid) * 2654435761u + 12345u;
    float acc = 0.f;
    for (int it = 0; it < iters; ++it) {
        const uint32_t base = ((h >> 7) & mask) * 192u;
        h = h * 1664525u + 1013904223u;
#pragma unroll
        for (int c = 0; c < 12; ++c) {
            const float4 q = ld4f(rs, base + 16u * c);
            acc += (q.x + q.y) + (q.z + q.w);
        }
    }
    out[gid] = acc;
}

static inline unsigned grid_for(int64_t n, int block) { return (unsigned)((n + block - 1) / block); }

hipError_t launch_gather_probe(const float* table, uint32_t bytes, uint32_t ncell, int mode, int iters,
                               int64_t nthreads, float* out, hipStream_t s) {
    hipLaunchKernelGGL(k_gather_probe, dim3(grid_for(nthreads, 256)), dim3(256), 0, s, table, bytes, ncell - 1,
                       mode, iters, nthreads, out);
    return hipGetLastError();
}

template <int CLASS>
static void launch_batched_class(bool mixed, bool periodic, const BatchParams& p, dim3 g, dim3 b, hipStream_t s) {
    if (mixed) {
        if (periodic) hipLaunchKernelGGL((k_batched<true, true, CLASS>), g, b, 0, s, p);
        else hipLaunchKernelGGL((k_batched<true, false, CLASS>), g, b, 0, s, p);
    } else {
        if (periodic) hipLaunchKernelGGL((k_batched<false, true, CLASS>), g, b, 0, s, p);
        else hipLaunchKernelGGL((k_batched<false, false, CLASS>), g, b, 0, s, p);
    }
}

hipError_t launch_batched(int prec, bool periodic, int tile_class, const BatchParams& p, int block, hipStream_t s) {
    const dim3 g((unsigned)p.ntiles), b(block);
    const bool mixed = prec == PP_PREC_FP32_MIXED;
    if (tile_class == PP_TILES_SINGLE_GATHER) launch_batched_class<PP_TILES_SINGLE_GATHER>(mixed, periodic, p, g, b, s);
    else if (tile_class == PP_TILES_RK4) launch_batched_class<PP_TILES_RK4>(mixed, periodic, p, g, b, s);
    else launch_batched_class<PP_TILES_ALL>(mixed, periodic, p, g, b, s);
    return hipGetLastError();
}

#define PP_FAST_CASES(KERNEL, L, U)                                                                          \
    switch (alg) {                                                                                        \
        case PP_ALG_LEAPFROG: hipLaunchKernelGGL((KERNEL<PP_ALG_LEAPFROG, L, U>), g, b, 0, s, p); break;     \
        case PP_ALG_BORIS: hipLaunchKernelGGL((KERNEL<PP_ALG_BORIS, L, U>), g, b, 0, s, p); break;           \
        case PP_ALG_RK4: hipLaunchKernelGGL((KERNEL<PP_ALG_RK4, L, U>), g, b, 0, s, p); break;               \
        case PP_ALG_YOSHIDA: hipLaunchKernelGGL((KERNEL<PP_ALG_YOSHIDA, L, U>), g, b, 0, s, p); break;       \
        case PP_ALG_VAY_REL: hipLaunchKernelGGL((KERNEL<PP_ALG_VAY_REL, L, U>), g, b, 0, s, p); break;       \
        case PP_ALG_RK4_REL: hipLaunchKernelGGL((KERNEL<PP_ALG_RK4_REL, L, U>), g, b, 0, s, p); break;       \
        case PP_ALG_YOSHIDA_REL: hipLaunchKernelGGL((KERNEL<PP_ALG_YOSHIDA_REL, L, U>), g, b, 0, s, p); break; \
        case PP_ALG_BORIS_REL: hipLaunchKernelGGL((KERNEL<PP_ALG_BORIS_REL, L, U>), g, b, 0, s, p); break;   \
        case PP_ALG_VAY: hipLaunchKernelGGL((KERNEL<PP_ALG_VAY, L, U>), g, b, 0, s, p); break;               \
        default: return hipErrorInvalidValue;                                                             \
    }

hipError_t launch_push(int alg, int prec, const PushParams& p, int block, hipStream_t s) {
    if (block <= 0) block = 256;
    const dim3 g(grid_for(p.n, block)), b(block);
    const bool loss = p.mesh != nullptr;
    const int fields = fields_case(p);
    // the loss test runs with the general or a fused-map evaluator
    if (prec == PP_PREC_FP64_FAST) {
        if (loss) {
            if (fields == kFieldsFusedStatic) PP_FAST_CASES(k_push64, true, kFieldsFusedStatic)
            else if (fields == kFieldsFusedRF) PP_FAST_CASES(k_push64, true, kFieldsFusedRF)
            else PP_FAST_CASES(k_push64, true, kFieldsGeneral)
        } else if (fields == kFieldsUniform) {
            PP_FAST_CASES(k_push64, false, kFieldsUniform)
        } else if (fields == kFieldsFusedStatic) {
            PP_FAST_CASES(k_push64, false, kFieldsFusedStatic)
        } else if (fields == kFieldsFusedRF) {
            PP_FAST_CASES(k_push64, false, kFieldsFusedRF)
        } else {
            PP_FAST_CASES(k_push64, false, kFieldsGeneral)
        }
    } else {
        if (loss) {
            if (fields == kFieldsFusedStatic) PP_FAST_CASES(k_push_mixed, true, kFieldsFusedStatic)
            else if (fields == kFieldsFusedRF) PP_FAST_CASES(k_push_mixed, true, kFieldsFusedRF)
            else PP_FAST_CASES(k_push_mixed, true, kFieldsGeneral)
        } else if (fields == kFieldsUniform) {
            PP_FAST_CASES(k_push_mixed, false, kFieldsUniform)
        } else if (fields == kFieldsFusedStatic) {
            PP_FAST_CASES(k_push_mixed, false, kFieldsFusedStatic)
        } else if (fields == kFieldsFusedRF) {
            PP_FAST_CASES(k_push_mixed, false, kFieldsFusedRF)
        } else {
            PP_FAST_CASES(k_push_mixed, false, kFieldsGeneral)
        }
    }
    return hipGetLastError();
}

#define PP_HOT_LAUNCH(A, RFV, LV)                                                                       \
    if (mixed) {                                                                                        \
        if (periodic) hipLaunchKernelGGL((k_hot<A, true, true, RFV, LV>), g, b, 0, s, p);              \
        else hipLaunchKernelGGL((k_hot<A, true, false, RFV, LV>), g, b, 0, s, p);                      \
    } else {                                                                                            \
        if (periodic) hipLaunchKernelGGL((k_hot<A, false, true, RFV, LV>), g, b, 0, s, p);             \
        else hipLaunchKernelGGL((k_hot<A, false, false, RFV, LV>), g, b, 0, s, p);                     \
    }

#define PP_HOT_MESH_LAUNCH(A)                                                                           \
    if (mixed) {                                                                                        \
        if (periodic) hipLaunchKernelGGL((k_hot_mesh<A, true, true>), g, b, 0, s, p);                  \
        else hipLaunchKernelGGL((k_hot_mesh<A, true, false>), g, b, 0, s, p);                          \
    } else {                                                                                            \
        if (periodic) hipLaunchKernelGGL((k_hot_mesh<A, false, true>), g, b, 0, s, p);                 \
        else hipLaunchKernelGGL((k_hot_mesh<A, false, false>), g, b, 0, s, p);                         \
    }

// launches with an alive mask (hot_eligible takes them without RF factors only) run a loss build
#define PP_HOT_CASE(A)                                                                                  \
    case A:                                                                                             \
        if (loss && p.mesh) {                                                                           \
            PP_HOT_MESH_LAUNCH(A)                                                                       \
        } else if (loss) {                                                                              \
            PP_HOT_LAUNCH(A, false, true)                                                               \
        } else if (rf) {                                                                                \
            PP_HOT_LAUNCH(A, true, false)                                                               \
        } else {                                                                                        \
            PP_HOT_LAUNCH(A, false, false)                                                              \
        }                                                                                               \
        break;

hipError_t launch_hot(int alg, int prec, bool periodic, const HotParams& p, int block, hipStream_t s) {
    if (block <= 0) block = 256;
    const dim3 g(grid_for(p.n, block)), b(block);
    const bool mixed = prec == PP_PREC_FP32_MIXED;
    const bool rf = p.rf.table != nullptr;
    const bool loss = p.alive != nullptr;
    if (loss && rf) return hipErrorInvalidValue;
    switch (alg) {
        PP_HOT_CASE(PP_ALG_LEAPFROG)
        PP_HOT_CASE(PP_ALG_BORIS)
        PP_HOT_CASE(PP_ALG_RK4)
        PP_HOT_CASE(PP_ALG_VAY_REL)
        PP_HOT_CASE(PP_ALG_RK4_REL)
        PP_HOT_CASE(PP_ALG_BORIS_REL)
        PP_HOT_CASE(PP_ALG_VAY)
        default: return hipErrorInvalidValue;
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
