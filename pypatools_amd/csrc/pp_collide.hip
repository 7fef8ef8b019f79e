// Particle loss on electrode surfaces: the segment each particle moved along in
// one step (r_old -> r_new) is tested against a triangle mesh, the GPU form of
// the reference's electrode_assembly.segment_intersects_surface call in
// Pusher.track_batch (src/PyPATools/pusher.py:982-990).
//
// The host builds the bounding-volume hierarchy (median split on the longest
// centroid axis, leaves of at most four triangles) and a pyramid of candidate
// grids that names, for each small block of space, the subtree holding every
// triangle near it; pp_bvh.inc holds the device layouts, the candidate lookup and
// the walk, shared with the push kernels that test losses inside the step loop.
#include <algorithm>
#include <cmath>
#include <numeric>
#include <vector>

#include "pp_internal.h"
#include "pp_bvh.inc"

#pragma clang fp contract(off)

namespace pp_collidek {

__global__ void __launch_bounds__(256) k_collide(const pp_bvh::MeshDev* md, const double* __restrict__ x0,
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
    const bool h = md ? pp_bvh::mesh_hits(md, x0[i], y0[i], z0[i], x1[i], y1[i], z1[i]) : false;
    if (hit) hit[i] = h ? 1 : 0;
    if (h && alive) alive[i] = 0;
}

}  // namespace pp_collidek

struct pp_mesh {
    pp_bvh::Node* nodes;
    pp_bvh::Tri* tris;
    pp_bvh::Range* ranges;  // candidate grids, all levels
    uint8_t* clear;         // level-0 clearance (MeshDev::clear)
    int32_t* cells;         // level-0 triangle lists (MeshDev::cells, MeshDev::cell_tris), or null
    int32_t* cell_tris;
    pp_bvh::MeshDev* dev;   // device copy of the descriptor
    int32_t nnodes;
    int64_t ntris;
    int64_t range_cells;
    int device;
};

namespace pp_collide {

using pp_bvh::Node;
using pp_bvh::Tri;

namespace {

constexpr int kLeaf = 4;

struct Builder {
    const std::vector<Tri>& tri;
    std::vector<double> lo, hi, cen;   // per triangle, 3 each
    std::vector<int32_t> order;
    std::vector<Node> nodes;
    std::vector<int32_t> parent, depth;
    double margin;

    explicit Builder(const std::vector<Tri>& t) : tri(t) {}

    int32_t build(int32_t b, int32_t e, int32_t up = -1, int32_t dep = 0) {
        const int32_t me = (int32_t)nodes.size();
        nodes.push_back(Node{});
        parent.push_back(up);
        depth.push_back(dep);
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
            build(b, mid, me, dep + 1);
            build(mid, e, me, dep + 1);
        }
        nd.escape = (int32_t)nodes.size();
        nodes[me] = nd;
        return me;
    }
};

// Candidate grids over the mesh box [glo, ghi].  Level 0 has about 64 cells per
// triangle (at most 2^22 cells, at most 4096 per axis); level L halves level L-1 on
// every axis (cell index >> L).  A cell's node is the lowest common ancestor of the
// leaves of every triangle whose box, grown by the margin, overlaps the cell; each
// level stores, per cell, the node range [s, e) of that ancestor's subtree, or an
// empty range.
struct Candidates {
    int32_t n0[3] = {1, 1, 1};
    double lo[3] = {0, 0, 0}, inv[3] = {0, 0, 0};
    int32_t nlevels = 0;
    int64_t off[pp_bvh::kMaxLevels] = {};
    std::vector<pp_bvh::Range> ranges;
    std::vector<uint8_t> clear;   // level 0: Chebyshev distance to the nearest non-empty cell, capped at 255
    // Level-0 triangle lists (MeshDev::cells, MeshDev::cell_tris): the triangles (positions in the
    // sorted order) whose grown box overlaps each cell.  A cell with more than kCellTris of them has
    // no list (start word with bit 31 set); all lists are dropped when they would hold more than
    // kListRefs entries.
    std::vector<int32_t> cells, cell_tris;
    static constexpr int32_t kCellTris = 32;
    static constexpr int64_t kListRefs = (int64_t)1 << 24;

    // c0, c1: the level-0 cell box of each triangle (3 each, in sorted order)
    void build_lists(const std::vector<int32_t>& c0, const std::vector<int32_t>& c1, int64_t nt) {
        const int64_t total = (int64_t)n0[0] * n0[1] * n0[2];
        std::vector<int32_t> cnt((size_t)total, 0);
        for (int64_t k = 0; k < nt; ++k)
            for (int32_t i = c0[3 * k]; i <= c1[3 * k]; ++i)
                for (int32_t j = c0[3 * k + 1]; j <= c1[3 * k + 1]; ++j)
                    for (int32_t q = c0[3 * k + 2]; q <= c1[3 * k + 2]; ++q) {
                        int32_t& c = cnt[((size_t)i * n0[1] + j) * n0[2] + q];
                        if (c <= kCellTris) ++c;
                    }
        cells.assign((size_t)total + 1, 0);
        int64_t refs = 0;
        for (int64_t c = 0; c < total; ++c) {
            const bool listed = cnt[(size_t)c] <= kCellTris;
            cells[(size_t)c] = (int32_t)refs | (listed ? 0 : (int32_t)0x80000000u);
            if (listed) refs += cnt[(size_t)c];
            if (refs > kListRefs) {
                cells.clear();
                return;
            }
        }
        cells[(size_t)total] = (int32_t)refs;
        cell_tris.assign((size_t)refs, 0);
        std::vector<int32_t> fill((size_t)total, 0);
        for (int64_t k = 0; k < nt; ++k)
            for (int32_t i = c0[3 * k]; i <= c1[3 * k]; ++i)
                for (int32_t j = c0[3 * k + 1]; j <= c1[3 * k + 1]; ++j)
                    for (int32_t q = c0[3 * k + 2]; q <= c1[3 * k + 2]; ++q) {
                        const size_t c = ((size_t)i * n0[1] + j) * n0[2] + q;
                        if (cells[c] >= 0) cell_tris[(size_t)(cells[c] + fill[c]++)] = (int32_t)k;
                    }
    }

    // Multi-source breadth-first search over the 26 neighbours of each cell: a
    // search front that advances one step per ring reaches a cell at exactly its
    // Chebyshev distance from the nearest non-empty cell.
    void build_clearance(const std::vector<int32_t>& level0) {
        const int64_t total = (int64_t)n0[0] * n0[1] * n0[2];
        clear.assign((size_t)total, 255);
        std::vector<int64_t> front, next;
        for (int64_t c = 0; c < total; ++c)
            if (level0[(size_t)c] >= 0) {
                clear[(size_t)c] = 0;
                front.push_back(c);
            }
        for (int d = 1; d < 255 && !front.empty(); ++d) {
            next.clear();
            for (const int64_t c : front) {
                const int32_t i = (int32_t)(c / ((int64_t)n0[1] * n0[2]));
                const int32_t j = (int32_t)((c / n0[2]) % n0[1]);
                const int32_t q = (int32_t)(c % n0[2]);
                for (int32_t di = -1; di <= 1; ++di)
                    for (int32_t dj = -1; dj <= 1; ++dj)
                        for (int32_t dq = -1; dq <= 1; ++dq) {
                            const int32_t a = i + di, b = j + dj, e = q + dq;
                            if (a < 0 || b < 0 || e < 0 || a >= n0[0] || b >= n0[1] || e >= n0[2]) continue;
                            const int64_t t = ((int64_t)a * n0[1] + b) * n0[2] + e;
                            if (clear[(size_t)t] == 255) {
                                clear[(size_t)t] = (uint8_t)d;
                                next.push_back(t);
                            }
                        }
            }
            front.swap(next);
        }
    }

    int32_t cell(int a, double x) const {
        const double f = std::floor((x - lo[a]) * inv[a]);
        if (!(f >= 0.0)) return 0;
        return f > (double)(n0[a] - 1) ? n0[a] - 1 : (int32_t)f;
    }

    static int32_t lca(const Builder& bld, int32_t a, int32_t b) {
        if (a < 0) return b;
        if (b < 0) return a;
        while (a != b) {
            if (bld.depth[a] > bld.depth[b]) a = bld.parent[a];
            else if (bld.depth[b] > bld.depth[a]) b = bld.parent[b];
            else {
                a = bld.parent[a];
                b = bld.parent[b];
            }
        }
        return a;
    }

    void build(const double glo[3], const double ghi[3], const Builder& bld, int64_t nt) {
        double ext[3], emax = 0.0;
        for (int a = 0; a < 3; ++a) {
            ext[a] = ghi[a] - glo[a];
            emax = std::max(emax, ext[a]);
        }
        const double target = (double)std::min<int64_t>(std::max<int64_t>(64 * nt, 4096), (int64_t)1 << 22);
        double vol = 1.0;
        for (int a = 0; a < 3; ++a) vol *= std::max(ext[a], 1e-6 * emax);
        double h = std::cbrt(vol / target);
        for (int tries = 0; tries < 64; ++tries, h *= 1.25) {
            int64_t total = 1;
            for (int a = 0; a < 3; ++a) {
                n0[a] = ext[a] > 0.0 && h > 0.0 ? (int32_t)std::min(4096.0, std::max(1.0, std::ceil(ext[a] / h))) : 1;
                total *= n0[a];
            }
            if (total <= ((int64_t)1 << 23)) break;
        }
        for (int a = 0; a < 3; ++a) {
            lo[a] = glo[a];
            inv[a] = ext[a] > 0.0 ? (double)n0[a] / ext[a] : 0.0;
        }
        // leaf of each triangle (positions in the sorted triangle order)
        std::vector<int32_t> leaf_of((size_t)nt, 0);
        for (int32_t v = 0; v < (int32_t)bld.nodes.size(); ++v)
            if (bld.nodes[v].count > 0)
                for (int32_t k = 0; k < bld.nodes[v].count; ++k) leaf_of[bld.nodes[v].first + k] = v;
        // level 0: common ancestor of the triangles near each cell
        std::vector<int32_t> cur((size_t)n0[0] * n0[1] * n0[2], -1);
        std::vector<int32_t> tc0(3 * (size_t)nt), tc1(3 * (size_t)nt);
        for (int64_t k = 0; k < nt; ++k) {
            const int32_t t = bld.order[k];
            int32_t* c0 = &tc0[3 * (size_t)k];
            int32_t* c1 = &tc1[3 * (size_t)k];
            for (int a = 0; a < 3; ++a) {
                c0[a] = cell(a, bld.lo[3 * t + a] - bld.margin);
                c1[a] = cell(a, bld.hi[3 * t + a] + bld.margin);
            }
            for (int32_t i = c0[0]; i <= c1[0]; ++i)
                for (int32_t j = c0[1]; j <= c1[1]; ++j)
                    for (int32_t q = c0[2]; q <= c1[2]; ++q) {
                        int32_t& c = cur[((size_t)i * n0[1] + j) * n0[2] + q];
                        c = lca(bld, c, leaf_of[k]);
                    }
        }
        build_clearance(cur);
        build_lists(tc0, tc1, nt);
        int32_t n[3] = {n0[0], n0[1], n0[2]};
        ranges.clear();
        for (nlevels = 0; nlevels < pp_bvh::kMaxLevels; ++nlevels) {
            off[nlevels] = (int64_t)ranges.size();
            for (const int32_t v : cur) ranges.push_back(v >= 0 ? pp_bvh::Range{v, bld.nodes[v].escape}
                                                                : pp_bvh::Range{0, 0});
            if (n[0] == 1 && n[1] == 1 && n[2] == 1) {
                ++nlevels;
                break;
            }
            // next level: cell index >> 1 on every axis
            int32_t m[3];
            for (int a = 0; a < 3; ++a) m[a] = ((n[a] - 1) >> 1) + 1;
            std::vector<int32_t> nxt((size_t)m[0] * m[1] * m[2], -1);
            for (int32_t i = 0; i < n[0]; ++i)
                for (int32_t j = 0; j < n[1]; ++j)
                    for (int32_t q = 0; q < n[2]; ++q) {
                        int32_t& c = nxt[((size_t)(i >> 1) * m[1] + (j >> 1)) * m[2] + (q >> 1)];
                        c = lca(bld, c, cur[((size_t)i * n[1] + j) * n[2] + q]);
                    }
            cur.swap(nxt);
            for (int a = 0; a < 3; ++a) n[a] = m[a];
        }
    }
};

}  // namespace

void destroy(pp_mesh* m);

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
    Candidates cand;
    if (nt > 0) cand.build(glo, ghi, bld, nt);
    pp_mesh* m = new pp_mesh();
    m->nnodes = (int32_t)bld.nodes.size();
    m->ntris = nt;
    m->nodes = nullptr;
    m->tris = nullptr;
    m->ranges = nullptr;
    m->clear = nullptr;
    m->cells = nullptr;
    m->cell_tris = nullptr;
    m->dev = nullptr;
    m->range_cells = nt > 0 ? (int64_t)cand.ranges.size() : 0;
    hipGetDevice(&m->device);
    if (m->nnodes > 0) {
        pp_bvh::MeshDev d{};
        const size_t rbytes = sizeof(pp_bvh::Range) * cand.ranges.size();
        bool ok = hipMalloc(&m->nodes, sizeof(Node) * m->nnodes) == hipSuccess &&
                  hipMalloc(&m->tris, sizeof(Tri) * (size_t)nt) == hipSuccess &&
                  hipMalloc(&m->ranges, rbytes) == hipSuccess &&
                  hipMalloc(&m->clear, cand.clear.size()) == hipSuccess &&
                  hipMalloc(&m->dev, sizeof(pp_bvh::MeshDev)) == hipSuccess;
        const bool lists = !cand.cells.empty();
        if (ok && lists)
            ok = hipMalloc(&m->cells, sizeof(int32_t) * cand.cells.size()) == hipSuccess &&
                 (cand.cell_tris.empty() ||
                  hipMalloc(&m->cell_tris, sizeof(int32_t) * cand.cell_tris.size()) == hipSuccess);
        if (ok) {
            d.nodes = m->nodes;
            d.tris = m->tris;
            d.ranges = m->ranges;
            d.clear = m->clear;
            d.cells = m->cells;
            d.cell_tris = m->cell_tris;
            d.lists = lists ? 1 : 0;
            d.nnodes = m->nnodes;
            d.nlevels = cand.nlevels;
            // FP32 cell estimate of the pre-test (clear_ahead): its error, in cells, is at most
            // 2^-24 (2 X inv + |lo| inv + n0) with X the largest |coordinate| of the grid box
            // (rounding of x, of inv, of -lo inv, of the fma); the device takes twice that as
            // ferr.  Where it is not below a quarter cell the offset is set so that no point
            // passes the pre-test's range check.
            double ferr = 0.0;
            for (int a = 0; a < 3; ++a) {
                const double X = std::max(std::fabs(glo[a]), std::fabs(ghi[a]));
                ferr = std::max(ferr, std::ldexp(2.0 * X * cand.inv[a] + std::fabs(cand.lo[a]) * cand.inv[a] +
                                                 (double)cand.n0[a], -23));
            }
            const bool fp32_ok = ferr < 0.25;
            d.ferr = fp32_ok ? (float)ferr : 0.25f;
            for (int a = 0; a < 3; ++a) {
                d.n0[a] = cand.n0[a];
                d.lo[a] = cand.lo[a];
                d.inv[a] = cand.inv[a];
                d.sf[a] = (float)cand.inv[a];
                d.of[a] = fp32_ok ? (float)(-cand.lo[a] * cand.inv[a]) : -1e30f;
            }
            for (int l = 0; l < pp_bvh::kMaxLevels; ++l) d.off[l] = cand.off[l];
            d.margin = bld.margin;
            ok = hipMemcpy(m->nodes, bld.nodes.data(), sizeof(Node) * m->nnodes, hipMemcpyHostToDevice) == hipSuccess &&
                 hipMemcpy(m->tris, sorted.data(), sizeof(Tri) * (size_t)nt, hipMemcpyHostToDevice) == hipSuccess &&
                 hipMemcpy(m->ranges, cand.ranges.data(), rbytes, hipMemcpyHostToDevice) == hipSuccess &&
                 hipMemcpy(m->clear, cand.clear.data(), cand.clear.size(), hipMemcpyHostToDevice) == hipSuccess &&
                 hipMemcpy(m->dev, &d, sizeof(d), hipMemcpyHostToDevice) == hipSuccess &&
                 (!lists || hipMemcpy(m->cells, cand.cells.data(), sizeof(int32_t) * cand.cells.size(),
                                      hipMemcpyHostToDevice) == hipSuccess) &&
                 (!m->cell_tris || hipMemcpy(m->cell_tris, cand.cell_tris.data(),
                                             sizeof(int32_t) * cand.cell_tris.size(),
                                             hipMemcpyHostToDevice) == hipSuccess);
        }
        if (!ok) {
            destroy(m);
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
    if (m->ranges) (void)hipFree(m->ranges);
    if (m->clear) (void)hipFree(m->clear);
    if (m->cells) (void)hipFree(m->cells);
    if (m->cell_tris) (void)hipFree(m->cell_tris);
    if (m->dev) (void)hipFree(m->dev);
    delete m;
}

void info(const pp_mesh* m, int64_t* ntris, int64_t* nnodes) {
    if (ntris) *ntris = m->ntris;
    if (nnodes) *nnodes = m->nnodes;
}

int device_of(const pp_mesh* m) { return m->device; }

const void* device_view(const pp_mesh* m) { return m->dev; }

hipError_t collide(const pp_mesh* m, const double* x0, const double* y0, const double* z0, const double* x1,
                   const double* y1, const double* z1, int64_t n, uint8_t* alive, uint8_t* hit, hipStream_t s) {
    const unsigned grid = (unsigned)((n + 255) / 256);
    // an empty mesh has no descriptor (md == null): nothing is hit
    hipLaunchKernelGGL(pp_collidek::k_collide, dim3(grid), dim3(256), 0, s, m->dev, x0, y0, z0, x1, y1, z1, n, alive,
                       hit);
    return hipGetLastError();
}

}  // namespace pp_collide
