// Particle loss on electrode surfaces: the segment each particle moved along in
// one step (r_old -> r_new) is tested against a triangle mesh, the GPU form of
// the reference's electrode_assembly.segment_intersects_surface call in
// Pusher.track_batch (src/PyPATools/pusher.py:982-990).
//
// The mesh is held as a bounding-volume hierarchy built on the host: nodes in
// depth-first order, each with the index that follows its subtree ("escape"),
// so the device walk needs no stack and every step moves to a higher node index
// (it always ends).  Node boxes are grown by a margin far above rounding, so they
// only cull; the hit decision is the FP64 segment-triangle test (Moller-Trumbore,
// closed intervals, no FMA contraction) and matches a brute-force test over every
// triangle bit for bit.
#include <algorithm>
#include <cmath>
#include <numeric>
#include <vector>

#include "pp_internal.h"

#pragma clang fp contract(off)

namespace pp_collidek {

struct Node {
    double lo[3], hi[3];
    int32_t first, count;   // leaf: triangles [first, first + count); inner node: count == 0
    int32_t escape, _pad;   // next node index after this subtree
};

struct Tri {
    double v0[3], e1[3], e2[3];   // e1 = v1 - v0, e2 = v2 - v0
};

__device__ __forceinline__ bool box_hit(const Node& b, const double p[3], const double d[3]) {
    double t0 = 0.0, t1 = 1.0;
    for (int a = 0; a < 3; ++a) {
        if (d[a] == 0.0) {
            if (p[a] < b.lo[a] || p[a] > b.hi[a]) return false;
        } else {
            const double inv = 1.0 / d[a];
            double ta = (b.lo[a] - p[a]) * inv, tb = (b.hi[a] - p[a]) * inv;
            if (ta > tb) {
                const double s = ta;
                ta = tb;
                tb = s;
            }
            t0 = ta > t0 ? ta : t0;
            t1 = tb < t1 ? tb : t1;
            if (!(t0 <= t1)) return false;
        }
    }
    return true;
}

// segment p + t d, t in [0, 1], against one triangle (edges included)
__device__ __forceinline__ bool tri_hit(const Tri& t, const double p[3], const double d[3]) {
    const double px = d[1] * t.e2[2] - d[2] * t.e2[1];
    const double py = d[2] * t.e2[0] - d[0] * t.e2[2];
    const double pz = d[0] * t.e2[1] - d[1] * t.e2[0];
    const double det = t.e1[0] * px + t.e1[1] * py + t.e1[2] * pz;
    if (det == 0.0) return false;
    const double inv = 1.0 / det;
    const double sx = p[0] - t.v0[0], sy = p[1] - t.v0[1], sz = p[2] - t.v0[2];
    const double u = (sx * px + sy * py + sz * pz) * inv;
    if (u < 0.0 || u > 1.0) return false;
    const double qx = sy * t.e1[2] - sz * t.e1[1];
    const double qy = sz * t.e1[0] - sx * t.e1[2];
    const double qz = sx * t.e1[1] - sy * t.e1[0];
    const double w = (d[0] * qx + d[1] * qy + d[2] * qz) * inv;
    if (w < 0.0 || u + w > 1.0) return false;
    const double s = (t.e2[0] * qx + t.e2[1] * qy + t.e2[2] * qz) * inv;
    return s >= 0.0 && s <= 1.0;
}

__global__ void __launch_bounds__(256) k_collide(const Node* __restrict__ nodes, int32_t nnodes,
                                                 const Tri* __restrict__ tris, const double* __restrict__ x0,
                                                 const double* __restrict__ y0, const double* __restrict__ z0,
                                                 const double* __restrict__ x1, const double* __restrict__ y1,
                                                 const double* __restrict__ z1, int64_t n, uint8_t* alive,
                                                 uint8_t* hit) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    if (alive && !alive[i]) {
        if (hit) hit[i] = 0;
        return;
    }
    const double p[3] = {x0[i], y0[i], z0[i]};
    const double d[3] = {x1[i] - p[0], y1[i] - p[1], z1[i] - p[2]};
    bool h = false;
    int32_t node = 0;
    while (node < nnodes) {
        const Node& b = nodes[node];
        if (!box_hit(b, p, d)) {
            node = b.escape;
        } else if (b.count == 0) {
            node = node + 1;
        } else {
            for (int32_t k = 0; k < b.count && !h; ++k) h = tri_hit(tris[b.first + k], p, d);
            if (h) break;
            node = b.escape;
        }
    }
    if (hit) hit[i] = h ? 1 : 0;
    if (h && alive) alive[i] = 0;
}

}  // namespace pp_collidek

struct pp_mesh {
    pp_collidek::Node* nodes;
    pp_collidek::Tri* tris;
    int32_t nnodes;
    int64_t ntris;
    int device;
};

namespace pp_collide {

using pp_collidek::Node;
using pp_collidek::Tri;

namespace {

constexpr int kLeaf = 4;

struct Builder {
    const std::vector<Tri>& tri;
    std::vector<double> lo, hi, cen;   // per triangle, 3 each
    std::vector<int32_t> order;
    std::vector<Node> nodes;
    double margin;

    explicit Builder(const std::vector<Tri>& t) : tri(t) {}

    int32_t build(int32_t b, int32_t e) {
        const int32_t me = (int32_t)nodes.size();
        nodes.push_back(Node{});
        Node nd{};
        for (int a = 0; a < 3; ++a) {
            nd.lo[a] = INFINITY;
            nd.hi[a] = -INFINITY;
        }
        double clo[3] = {INFINITY, INFINITY, INFINITY}, chi[3] = {-INFINITY, -INFINITY, -INFINITY};
        for (int32_t k = b; k < e; ++k) {
            const int32_t t = order[k];
            for (int a = 0; a < 3; ++a) {
                nd.lo[a] = std::min(nd.lo[a], lo[3 * t + a]);
                nd.hi[a] = std::max(nd.hi[a], hi[3 * t + a]);
                clo[a] = std::min(clo[a], cen[3 * t + a]);
                chi[a] = std::max(chi[a], cen[3 * t + a]);
            }
        }
        for (int a = 0; a < 3; ++a) {
            nd.lo[a] -= margin;
            nd.hi[a] += margin;
        }
        if (e - b <= kLeaf) {
            nd.first = b;
            nd.count = e - b;
        } else {
            int axis = 0;
            for (int a = 1; a < 3; ++a)
                if (chi[a] - clo[a] > chi[axis] - clo[axis]) axis = a;
            const int32_t mid = b + (e - b) / 2;
            std::nth_element(order.begin() + b, order.begin() + mid, order.begin() + e,
                             [&](int32_t u, int32_t v) { return cen[3 * u + axis] < cen[3 * v + axis]; });
            nd.first = 0;
            nd.count = 0;
            build(b, mid);
            build(mid, e);
        }
        nd.escape = (int32_t)nodes.size();
        nodes[me] = nd;
        return me;
    }
};

}  // namespace

int create(const double* v, int64_t nv, const int32_t* idx, int64_t nt, pp_mesh** out, char* err, size_t errn) {
    std::vector<Tri> tri((size_t)nt);
    double glo[3] = {INFINITY, INFINITY, INFINITY}, ghi[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (int64_t k = 0; k < nv; ++k)
        for (int a = 0; a < 3; ++a) {
            const double x = v[3 * k + a];
            if (!std::isfinite(x)) {
                snprintf(err, errn, "vertex %lld is not finite", (long long)k);
                return PP_ERR_INVALID;
            }
            glo[a] = std::min(glo[a], x);
            ghi[a] = std::max(ghi[a], x);
        }
    Builder bld(tri);
    bld.lo.resize(3 * (size_t)nt);
    bld.hi.resize(3 * (size_t)nt);
    bld.cen.resize(3 * (size_t)nt);
    for (int64_t t = 0; t < nt; ++t) {
        const int32_t a = idx[3 * t], b = idx[3 * t + 1], c = idx[3 * t + 2];
        if (a < 0 || b < 0 || c < 0 || a >= nv || b >= nv || c >= nv) {
            snprintf(err, errn, "triangle %lld has a vertex index outside [0, %lld)", (long long)t, (long long)nv);
            return PP_ERR_INVALID;
        }
        for (int q = 0; q < 3; ++q) {
            tri[t].v0[q] = v[3 * a + q];
            tri[t].e1[q] = v[3 * b + q] - v[3 * a + q];
            tri[t].e2[q] = v[3 * c + q] - v[3 * a + q];
            const double x0 = v[3 * a + q], x1 = v[3 * b + q], x2 = v[3 * c + q];
            bld.lo[3 * t + q] = std::min(x0, std::min(x1, x2));
            bld.hi[3 * t + q] = std::max(x0, std::max(x1, x2));
            bld.cen[3 * t + q] = (x0 + x1 + x2) / 3.0;
        }
    }
    double ext = 0.0;
    for (int a = 0; a < 3; ++a) ext = std::max(ext, nv ? ghi[a] - glo[a] : 0.0);
    bld.margin = 1e-9 * ext + 1e-300;
    bld.order.resize((size_t)nt);
    std::iota(bld.order.begin(), bld.order.end(), 0);
    if (nt > 0) bld.build(0, (int32_t)nt);
    std::vector<Tri> sorted((size_t)nt);
    for (int64_t k = 0; k < nt; ++k) sorted[k] = tri[bld.order[k]];
    pp_mesh* m = new pp_mesh();
    m->nnodes = (int32_t)bld.nodes.size();
    m->ntris = nt;
    m->nodes = nullptr;
    m->tris = nullptr;
    hipGetDevice(&m->device);
    if (m->nnodes > 0) {
        if (hipMalloc(&m->nodes, sizeof(Node) * m->nnodes) != hipSuccess ||
            hipMalloc(&m->tris, sizeof(Tri) * (size_t)nt) != hipSuccess ||
            hipMemcpy(m->nodes, bld.nodes.data(), sizeof(Node) * m->nnodes, hipMemcpyHostToDevice) != hipSuccess ||
            hipMemcpy(m->tris, sorted.data(), sizeof(Tri) * (size_t)nt, hipMemcpyHostToDevice) != hipSuccess) {
            if (m->nodes) (void)hipFree(m->nodes);
            if (m->tris) (void)hipFree(m->tris);
            delete m;
            snprintf(err, errn, "device allocation for the mesh failed");
            return PP_ERR_NOMEM;
        }
    }
    *out = m;
    return PP_OK;
}

void destroy(pp_mesh* m) {
    if (!m) return;
    if (m->nodes) (void)hipFree(m->nodes);
    if (m->tris) (void)hipFree(m->tris);
    delete m;
}

void info(const pp_mesh* m, int64_t* ntris, int64_t* nnodes) {
    if (ntris) *ntris = m->ntris;
    if (nnodes) *nnodes = m->nnodes;
}

int device_of(const pp_mesh* m) { return m->device; }

hipError_t collide(const pp_mesh* m, const double* x0, const double* y0, const double* z0, const double* x1,
                   const double* y1, const double* z1, int64_t n, uint8_t* alive, uint8_t* hit, hipStream_t s) {
    const unsigned grid = (unsigned)((n + 255) / 256);
    hipLaunchKernelGGL(pp_collidek::k_collide, dim3(grid), dim3(256), 0, s, m->nodes, m->nnodes, m->tris, x0, y0, z0,
                       x1, y1, z1, n, alive, hit);
    return hipGetLastError();
}

}  // namespace pp_collide
