// Nested-dissection direct factorisation of a 2D FDFD operator (complex64),
// the preconditioner of FP64-refined FGMRES for grids too wide for the
// block-tridiagonal path.  It replaces, for those grids, the sparse direct
// factorisation of ceviche's default solve (solvers.py:50-62 behind
// primitives.py:103-119, fdfd.py:172-192 / :211-244).
//
// The grid is bisected recursively (nested dissection): a box of more than
// kLeafCells cells is cut by the middle line of its longer side, which is
// the node's separator; the two halves are its children.  A node's
// variables I are all cells of a leaf box or the separator cells of an
// internal node; its boundary B is the ring of cells around its box (the
// stencil neighbours outside it, without the periodic wrap, which the
// preconditioner drops like the block-tridiagonal path drops the x-wrap),
// all of them on separators of its ancestors.  Each node eliminates I from a
// dense frontal matrix F over [I; B], assembled as the sum of its children's
// Schur complements (extend-add) and the couplings of A that involve its own
// variables (each entry of A added once, at the first node that eliminates
// one of its two cells),
// then forms Q = F_II^-1 and X = Q F_IB, passes S = F_BB - F_BI X up to its
// parent and forms G = F_BI Q.  Fronts of at most 64 interior cells do this
// by a Gauss-Jordan sweep (k_nd_gj, k_nd_gj_panel); larger interior blocks
// up to kLuMaxRows cells by a blocked Gauss-Jordan on [F_II | I F_IB] with
// 16-column LU panels, which leaves [Q X] directly; above that F_II = P L U
// by rocsolver and [Q X] by the triangular solves of [I F_IB] (fronts of at
// most 64 interior cells that the sweep kernels do not hold: the LU, then
// the inverse in place and one product).
// F keeps Q, X and G in place of its blocks:
//   F = [[Q, X], [G, S]].
// Nodes of equal height (distance to the leaves) are processed together as
// strided batches padded to the level's largest |I| (kmax, identity padding)
// and |B| (bmax, zero padding): dense LU, inversion or triangular solves
// (rocsolver getrf, getri / getrs) and the products (rocblas gemm) run on the
// matrix cores, one launch sequence per level.  The symbolic analysis
// (tree, index sets, maps) is host C++ done once per grid shape.
//
// Solve of A x = b (transpose 0), one matrix-vector product per node and sweep:
//   up the tree:   [y_I; t] = [Q; G] b_I,  b_B -= t
//   down the tree: x_I = y_I - X x_B
// and of A^T x = b (transpose 1):
//   up:   [y_I; t] = [Q X]^T b_I,  b_B -= t
//   down: x_I = y_I - G^T x_B
// (A^T = [[F_II^T, F_BI^T], [F_IB^T, F_BB^T]], F_IB^T F_II^-T = X^T and
// F_II^-T F_BI^T = G^T).  Each product is one launch for all nodes of a
// level (a workgroup per node up to 256 rows and columns, else 64 x 64 tiles
// with partial sums added by a second launch), the input gathered from the
// work vector.  Nodes of a level share boundary cells (the two sides of a
// separator), so their b_B contributions are summed per cell in a fixed
// order by one more launch.
#include "fdfd_stencil.h"

#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include <algorithm>
#include <mutex>
#include <vector>

namespace cv {
namespace {

constexpr int kLeafCells = 32;
constexpr int kMvTile = 64;        // rows and columns of one tile of the application products
constexpr int kNdSolveMin = 64;   // interior blocks above this size: [Q X] directly, not the inverse and a product
constexpr int kLuPanel = 16;      // columns per panel of the blocked Gauss-Jordan
constexpr int kLuMin = 64;        // interior blocks above this size take the blocked Gauss-Jordan ...
constexpr int kLuMaxRows = 2048;  // ... up to this size (the panel kernels hold 1024 threads x 2 rows)

struct NdGroup {               // (parent, child) pairs whose child sits at one level, one child slot
    int child_level = 0, slot = 0;
    std::vector<int32_t> parent, child, map;   // map: per pair, the child level's bmax entries
    int64_t off_parent = 0, off_child = 0, off_map = 0;
};

struct NdLevel {
    int nb = 0, kmax = 0, bmax = 0, m = 0;
    std::vector<int32_t> icells, bcells;   // nb * kmax, nb * bmax (-1 = padding)
    // per interior position the [I; B] positions of its 4 stencil neighbours
    // whose couplings this node assembles (-1: outside the grid or eliminated
    // below this node)
    std::vector<int32_t> anbr;
    std::vector<NdGroup> groups;           // children's Schur complements, extend-added slot by slot
    // the level's boundary cells, each with its contributions e * bmax + j
    // (several nodes of a level can border the same separator cell)
    std::vector<int32_t> bs_cells, bs_ptr, bs_idx;
    // byte offsets of the device copies (plan buffer) and of the numeric data (factor buffer)
    int64_t off_icells = 0, off_bcells = 0, off_anbr = 0, off_bs_cells = 0, off_bs_ptr = 0, off_bs_idx = 0;
    int64_t off_F = 0, off_piv = 0, info_first = 0;
};

}  // namespace
}  // namespace cv

struct cv_nd_plan {
    int nx = 0, ny = 0;
    std::vector<cv::NdLevel> L;
    int64_t plan_bytes = 0, factor_bytes = 0;
    int64_t off_w = 0, off_part = 0, off_tb = 0, off_info = 0, off_r = 0, off_r2 = 0;
    int64_t n_nodes = 0;
    double flops = 0.0;   // real flops of one numeric factorisation (unpadded sizes)
};

namespace cv {
namespace {

int64_t al256(int64_t x) { return (x + 255) & ~int64_t(255); }

struct Box {
    int x0, x1, y0, y1;
    int c0 = -1, c1 = -1, height = 0;
    int sx = -1, sy = -1;   // the separator: column x = sx or row y = sy (internal nodes)
    std::vector<int32_t> I, B;
};

// Recursive bisection: a box of more than kLeafCells cells (and a side of
// at least 3) is cut by the middle line of its longer side, which becomes
// the node's separator; the two remaining halves are its children.
int build_tree(std::vector<Box>& nodes, int x0, int x1, int y0, int y1) {
    const int idx = (int)nodes.size();
    nodes.push_back(Box{x0, x1, y0, y1});
    const int w = x1 - x0, h = y1 - y0;
    if ((int64_t)w * h <= kLeafCells || (w < 3 && h < 3)) return idx;
    int c0, c1;
    if (w >= h) {
        const int xm = (x0 + x1) / 2;
        nodes[idx].sx = xm;
        c0 = build_tree(nodes, x0, xm, y0, y1);
        c1 = build_tree(nodes, xm + 1, x1, y0, y1);
    } else {
        const int ym = (y0 + y1) / 2;
        nodes[idx].sy = ym;
        c0 = build_tree(nodes, x0, x1, y0, ym);
        c1 = build_tree(nodes, x0, x1, ym + 1, y1);
    }
    nodes[idx].c0 = c0;
    nodes[idx].c1 = c1;
    nodes[idx].height = 1 + std::max(nodes[c0].height, nodes[c1].height);
    return idx;
}

// the cells outside box b that are stencil neighbours of its cells (its
// 5-point ring, clipped to the grid): all of them lie on separators of
// b's ancestors; sorted
void box_ring(const Box& b, int nx, int ny, std::vector<int32_t>& out) {
    out.clear();
    auto cell = [&](int i, int j) { return (int32_t)((int64_t)i * ny + j); };
    if (b.x0 > 0)
        for (int j = b.y0; j < b.y1; ++j) out.push_back(cell(b.x0 - 1, j));
    if (b.y0 > 0)
        for (int i = b.x0; i < b.x1; ++i) out.push_back(cell(i, b.y0 - 1));
    if (b.y1 < ny)
        for (int i = b.x0; i < b.x1; ++i) out.push_back(cell(i, b.y1));
    if (b.x1 < nx)
        for (int j = b.y0; j < b.y1; ++j) out.push_back(cell(b.x1, j));
    std::sort(out.begin(), out.end());
}

// stencil neighbours of cell c in the order xm, xp, ym, yp (-1 outside the grid; no wrap)
void neighbours(int32_t c, int nx, int ny, int32_t (&nb)[4]) {
    const int i = c / ny, j = c % ny;
    nb[0] = (i > 0) ? c - ny : -1;
    nb[1] = (i < nx - 1) ? c + ny : -1;
    nb[2] = (j > 0) ? c - 1 : -1;
    nb[3] = (j < ny - 1) ? c + 1 : -1;
}

}  // namespace
}  // namespace cv

using namespace cv;

extern "C" cv_nd_plan* cv_nd_plan_create(int nx, int ny) {
    if (nx <= 0 || ny <= 0 || (int64_t)nx * ny >= (int64_t)0x7fffffff) {
        set_error("nd plan: bad grid %d x %d", nx, ny);
        return nullptr;
    }
    cv_nd_plan* P = new cv_nd_plan();
    P->nx = nx;
    P->ny = ny;
    std::vector<Box> nodes;
    nodes.reserve(4 * ((int64_t)nx * ny / kLeafCells + 4));
    build_tree(nodes, 0, nx, 0, ny);
    const int64_t N = (int64_t)nx * ny;
    int H = 0;
    for (auto& b : nodes) H = std::max(H, b.height);
    // variables I (all cells of a leaf, the separator of an internal node)
    // and boundary B (the box's ring)
    for (auto& b : nodes) {
        box_ring(b, nx, ny, b.B);
        if (b.c0 < 0) {
            for (int i = b.x0; i < b.x1; ++i)
                for (int j = b.y0; j < b.y1; ++j) b.I.push_back((int32_t)((int64_t)i * ny + j));
        } else if (b.sx >= 0) {
            for (int j = b.y0; j < b.y1; ++j) b.I.push_back((int32_t)((int64_t)b.sx * ny + j));
        } else {
            for (int i = b.x0; i < b.x1; ++i) b.I.push_back((int32_t)((int64_t)i * ny + b.sy));
        }
    }
    // algorithmic work of one factorisation, real flops of complex arithmetic
    // (a complex multiply-add is 8): LU k^3/3, inverse from the LU factors
    // 2k^3/3, X = Q F_IB k^2 b, S b^2 k and G = F_BI Q b k^2 complex
    // multiply-adds per node (the triangular-solve route for large interior
    // blocks does k^3/3 more)
    for (const auto& b : nodes) {
        const double k = (double)b.I.size(), bb = (double)b.B.size();
        P->flops += 8.0 * (k * k * k + 2.0 * k * k * bb + bb * bb * k);
    }
    // group by height
    P->L.resize(H + 1);
    std::vector<int> level_index(nodes.size());
    for (size_t q = 0; q < nodes.size(); ++q) {
        NdLevel& L = P->L[nodes[q].height];
        level_index[q] = L.nb++;
        L.kmax = std::max<int>(L.kmax, (int)nodes[q].I.size());
        L.bmax = std::max<int>(L.bmax, (int)nodes[q].B.size());
    }
    for (auto& L : P->L) {
        L.m = L.kmax + L.bmax;
        if ((int64_t)L.nb * L.m * L.m > ((int64_t)1 << 40)) {
            set_error("nd plan: %d x %d is too large", nx, ny);
            delete P;
            return nullptr;
        }
    }
    std::vector<std::vector<int>> members(H + 1);
    for (size_t q = 0; q < nodes.size(); ++q) members[nodes[q].height].push_back((int)q);
    // position of a cell in the current node's [I; B] ordering
    std::vector<int32_t> pos(N, -1);
    for (int h = 0; h <= H; ++h) {
        NdLevel& L = P->L[h];
        L.icells.assign((size_t)L.nb * L.kmax, -1);
        L.bcells.assign((size_t)L.nb * L.bmax, -1);
        L.anbr.assign((size_t)L.nb * L.kmax * 4, -1);
        std::vector<std::pair<int32_t, int32_t>> contrib;   // (boundary cell, e * bmax + j)
        for (int q : members[h]) {
            const Box& b = nodes[q];
            const int e = level_index[q];
            for (size_t p = 0; p < b.I.size(); ++p) {
                L.icells[(size_t)e * L.kmax + p] = b.I[p];
                pos[b.I[p]] = (int32_t)p;
            }
            for (size_t p = 0; p < b.B.size(); ++p) {
                L.bcells[(size_t)e * L.bmax + p] = b.B[p];
                pos[b.B[p]] = (int32_t)(L.kmax + p);
                contrib.emplace_back(b.B[p], (int32_t)((int64_t)e * L.bmax + p));
            }
            for (size_t p = 0; p < b.I.size(); ++p) {
                int32_t nbv[4];
                neighbours(b.I[p], nx, ny, nbv);
                for (int d = 0; d < 4; ++d)
                    L.anbr[((size_t)e * L.kmax + p) * 4 + d] = nbv[d] >= 0 ? pos[nbv[d]] : -1;
            }
            if (b.c0 >= 0) {
                for (int side = 0; side < 2; ++side) {
                    const int cq = side == 0 ? b.c0 : b.c1;
                    const int ch = nodes[cq].height;
                    const NdLevel& CL = P->L[ch];
                    NdGroup* g = nullptr;
                    for (auto& gg : L.groups)
                        if (gg.child_level == ch && gg.slot == side) g = &gg;
                    if (!g) {
                        L.groups.push_back(NdGroup());
                        g = &L.groups.back();
                        g->child_level = ch;
                        g->slot = side;
                    }
                    g->parent.push_back(e);
                    g->child.push_back(level_index[cq]);
                    const size_t base = g->map.size();
                    g->map.resize(base + CL.bmax, -1);
                    for (size_t p = 0; p < nodes[cq].B.size(); ++p) {
                        const int32_t ps = pos[nodes[cq].B[p]];
                        if (ps < 0) {   // a child's ring lies on its parent's separator or ring
                            set_error("nd plan: inconsistent tree at node %d", q);
                            delete P;
                            return nullptr;
                        }
                        g->map[base + p] = ps;
                    }
                }
            }
            for (int32_t c : b.I) pos[c] = -1;
            for (int32_t c : b.B) pos[c] = -1;
        }
        // the groups of one child level: slot 0 before slot 1 (extend-add order)
        std::stable_sort(L.groups.begin(), L.groups.end(), [](const NdGroup& x, const NdGroup& y) {
            return x.child_level != y.child_level ? x.child_level < y.child_level : x.slot < y.slot;
        });
        std::sort(contrib.begin(), contrib.end());
        L.bs_ptr.push_back(0);
        for (size_t i = 0; i < contrib.size(); ++i) {
            if (i == 0 || contrib[i].first != contrib[i - 1].first) {
                if (i > 0) L.bs_ptr.push_back((int32_t)i);
                L.bs_cells.push_back(contrib[i].first);
            }
            L.bs_idx.push_back(contrib[i].second);
        }
        if (!contrib.empty()) L.bs_ptr.push_back((int32_t)contrib.size());
    }
    // device layouts: index arrays in the plan buffer, numeric data in the factor buffer
    int64_t po = 0, fo = 0, pmax = 0, rmax = 0, r2max = 0, tbmax = 0, info_first = 0;
    auto take = [](int64_t& o, int64_t bytes) {
        const int64_t r = o;
        o += al256(bytes);
        return r;
    };
    for (auto& L : P->L) {
        L.off_icells = take(po, 4 * (int64_t)L.icells.size());
        L.off_bcells = take(po, 4 * (int64_t)L.bcells.size());
        L.off_anbr = take(po, 4 * (int64_t)L.anbr.size());
        for (auto& g : L.groups) {
            g.off_parent = take(po, 4 * (int64_t)g.parent.size());
            g.off_child = take(po, 4 * (int64_t)g.child.size());
            g.off_map = take(po, 4 * (int64_t)g.map.size());
        }
        L.off_bs_cells = take(po, 4 * (int64_t)L.bs_cells.size());
        L.off_bs_ptr = take(po, 4 * (int64_t)L.bs_ptr.size());
        L.off_bs_idx = take(po, 4 * (int64_t)L.bs_idx.size());
        L.off_F = take(fo, 8 * (int64_t)L.nb * L.m * L.m);
        L.off_piv = take(fo, 4 * (int64_t)L.nb * std::max(1, L.kmax));
        L.info_first = info_first;
        info_first += L.nb;
        // partial sums of the tiled products: per node ceil(k/64) tiles of m
        // outputs (up sweeps) or ceil(b/64) tiles of k outputs (down sweeps)
        pmax = std::max<int64_t>(pmax, (int64_t)L.nb * std::max<int64_t>((int64_t)((L.kmax + kMvTile - 1) / kMvTile) * L.m,
                                                                        (int64_t)((L.bmax + kMvTile - 1) / kMvTile) * L.kmax));
        rmax = std::max<int64_t>(rmax, (int64_t)L.nb * L.kmax * L.m);
        if (L.kmax > kLuMin && L.kmax <= kLuMaxRows)   // the blocked Gauss-Jordan's second buffer
            r2max = std::max<int64_t>(r2max, (int64_t)L.nb * L.kmax * (L.kmax + L.m));
        tbmax = std::max<int64_t>(tbmax, (int64_t)L.nb * L.bmax);
    }
    P->n_nodes = info_first;
    P->off_info = take(fo, 4 * info_first);
    P->off_w = take(fo, 8 * N);
    P->off_part = take(fo, 8 * std::max<int64_t>(1, pmax));
    P->off_tb = take(fo, 8 * std::max<int64_t>(1, tbmax));
    P->off_r = take(fo, 8 * std::max<int64_t>(1, rmax));
    P->off_r2 = take(fo, 8 * std::max<int64_t>(1, r2max));
    P->plan_bytes = std::max<int64_t>(po, 256);
    P->factor_bytes = fo;
    return P;
}

extern "C" void cv_nd_plan_destroy(cv_nd_plan* p) { delete p; }

extern "C" int64_t cv_nd_plan_bytes(const cv_nd_plan* p) { return p ? p->plan_bytes : -1; }
extern "C" int64_t cv_nd_factor_bytes(const cv_nd_plan* p) { return p ? p->factor_bytes : -1; }
extern "C" double cv_nd_plan_flops(const cv_nd_plan* p) { return p ? p->flops : -1.0; }

extern "C" int cv_nd_plan_stats(const cv_nd_plan* p, int* levels, int* nb, int* kmax, int* bmax, int max_levels) {
    CV_CHECK(p, "null plan");
    const int n = (int)p->L.size();
    if (levels) *levels = n;
    for (int h = 0; h < std::min(n, max_levels); ++h) {
        if (nb) nb[h] = p->L[h].nb;
        if (kmax) kmax[h] = p->L[h].kmax;
        if (bmax) bmax[h] = p->L[h].bmax;
    }
    return CV_OK;
}

extern "C" int cv_nd_plan_upload(const cv_nd_plan* p, void* plan_dev, int64_t plan_bytes, void* stream) {
    CV_CHECK(p && plan_dev, "null argument");
    CV_CHECK(plan_bytes >= p->plan_bytes, "plan buffer too small: %lld < %lld", (long long)plan_bytes,
             (long long)p->plan_bytes);
    hipStream_t s = as_stream(stream);
    char* base = (char*)plan_dev;
    auto up = [&](int64_t off, const std::vector<int32_t>& v) -> int {
        if (!v.empty()) CV_HIP(hipMemcpyAsync(base + off, v.data(), 4 * v.size(), hipMemcpyHostToDevice, s));
        return CV_OK;
    };
    for (const auto& L : p->L) {
        int rc = CV_OK;
        if ((rc = up(L.off_icells, L.icells)) || (rc = up(L.off_bcells, L.bcells)) || (rc = up(L.off_anbr, L.anbr)) ||
            (rc = up(L.off_bs_cells, L.bs_cells)) || (rc = up(L.off_bs_ptr, L.bs_ptr)) ||
            (rc = up(L.off_bs_idx, L.bs_idx)))
            return rc;
        for (const auto& g : L.groups)
            if ((rc = up(g.off_parent, g.parent)) || (rc = up(g.off_child, g.child)) || (rc = up(g.off_map, g.map)))
                return rc;
    }
    // the host vectors are read by the copies above: wait for them before returning
    CV_HIP(hipStreamSynchronize(s));
    return CV_OK;
}

namespace cv {
namespace {

// ------------------------------------------------------------------ kernels
__device__ __forceinline__ cx<float> row_dir(const Row5<float>& r, int d) {
    return d == 0 ? r.xm : d == 1 ? r.xp : d == 2 ? r.ym : r.yp;
}

// The couplings of A a node adds to its frontal matrix (after its
// children's Schur complements): for every variable s (row p) the diagonal
// and, for each stencil neighbour t at position q >= 0 of the node's
// [I; B] ordering, A[s, t] at (p, q) and, when t is a boundary cell
// (q >= kmax), A[t, s] at (q, p).  Couplings to cells eliminated below the
// node (q < 0) were added there.  Each entry has a single writer.  Padded
// interior positions take the identity.
__global__ void __launch_bounds__(256)
k_nd_assemble(const cv_fdfd_desc* __restrict__ desc, const cx<float>* __restrict__ prof,
              const float* __restrict__ m0, const float* __restrict__ m1, const int32_t* __restrict__ icells,
              const int32_t* __restrict__ anbr, cx<float>* __restrict__ F, int nb, int kmax, int m) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)nb * kmax) return;
    const int e = (int)(t / kmax), p = (int)(t - (int64_t)e * kmax);
    cx<float>* Fe = F + (int64_t)e * m * m;
    const int32_t c = icells[t];
    if (c < 0) {
        Fe[(int64_t)p * m + p] = Fe[(int64_t)p * m + p] + mk<float>(1.f, 0.f);
        return;
    }
    const cv_fdfd_desc d = desc[0];
    const Row5<float> r = row_coefs<float>(d, prof, m0, m1, c / d.ny, c % d.ny);
    Fe[(int64_t)p * m + p] = Fe[(int64_t)p * m + p] + r.c;
#pragma unroll
    for (int dd = 0; dd < 4; ++dd) {
        const int q = anbr[t * 4 + dd];
        if (q < 0) continue;
        Fe[(int64_t)q * m + p] = Fe[(int64_t)q * m + p] + row_dir(r, dd);
        if (q >= kmax) {
            const int32_t tc = dd == 0 ? c - d.ny : dd == 1 ? c + d.ny : dd == 2 ? c - 1 : c + 1;
            const Row5<float> rt = row_coefs<float>(d, prof, m0, m1, tc / d.ny, tc % d.ny);
            Fe[(int64_t)p * m + q] = Fe[(int64_t)p * m + q] + row_dir(rt, dd ^ 1);
        }
    }
}

// children's Schur complements (the bottom-right bc x bc block of their F)
// added into the parents' F at the mapped positions (one child slot per
// launch, so every parent entry has one writer); one thread per entry of a
// child's block, rows fastest (the reads of a wave are contiguous)
__global__ void __launch_bounds__(256)
k_nd_extend(const cx<float>* __restrict__ Fc, int mc, int kc, int bc, int npairs, const int32_t* __restrict__ parent,
            const int32_t* __restrict__ child, const int32_t* __restrict__ map, cx<float>* __restrict__ F, int m) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t per = (int64_t)bc * bc;
    if (t >= (int64_t)npairs * per) return;
    const int pr = (int)(t / per);
    const int r = (int)(t - (int64_t)pr * per);
    const int jj = r / bc, ii = r - jj * bc;
    const int32_t* mp = map + (int64_t)pr * bc;
    const int pj = mp[jj], pi = mp[ii];
    if (pj < 0 || pi < 0) return;
    const cx<float> v = Fc[(int64_t)child[pr] * mc * mc + (int64_t)(kc + jj) * mc + kc + ii];
    cx<float>* dst = F + (int64_t)parent[pr] * m * m + (int64_t)pj * m + pi;
    *dst = *dst + v;
}

// Pivot search keys: one 64-bit word per candidate whose maximum is the
// pivot, largest |a|^2 first and ties to the lower position.  High word: the
// bits of |a|^2 plus one (non-negative floats order as their bit patterns;
// 0 means no candidate), low word: the complement of the position.  A NaN
// magnitude never becomes a candidate.
__device__ __forceinline__ uint64_t piv_key(float v, int pos) {
    return (v == v) ? ((uint64_t)(__float_as_uint(v) + 1u) << 32) | (uint32_t)(~pos) : 0ull;
}

template <int CTRL>
__device__ __forceinline__ uint64_t dpp_u64(uint64_t v) {
    const int lo = __builtin_amdgcn_update_dpp(0, (int)(uint32_t)v, CTRL, 0xF, 0xF, false);
    const int hi = __builtin_amdgcn_update_dpp(0, (int)(uint32_t)(v >> 32), CTRL, 0xF, 0xF, false);
    return ((uint64_t)(uint32_t)hi << 32) | (uint32_t)lo;
}

__device__ __forceinline__ uint64_t umax64(uint64_t a, uint64_t b) { return a > b ? a : b; }

__device__ __forceinline__ uint64_t readlane64(uint64_t v, int l) {
    const uint32_t lo = __builtin_amdgcn_readlane((int)(uint32_t)v, l);
    const uint32_t hi = __builtin_amdgcn_readlane((int)(uint32_t)(v >> 32), l);
    return ((uint64_t)hi << 32) | lo;
}

// Maximum over each row of 16 lanes (every lane of the row gets it) by DPP
// lane exchanges within the row: neighbours, pairs, half-row and row mirrors.
// All lanes must be active.
__device__ __forceinline__ uint64_t row16_max(uint64_t k) {
    k = umax64(k, dpp_u64<0xB1>(k));    // quad_perm [1,0,3,2]
    k = umax64(k, dpp_u64<0x4E>(k));    // quad_perm [2,3,0,1]
    k = umax64(k, dpp_u64<0x141>(k));   // row_half_mirror
    k = umax64(k, dpp_u64<0x140>(k));   // row_mirror
    return k;
}

// Maximum over the wave, uniform.
__device__ __forceinline__ uint64_t wave_max64(uint64_t k) {
    k = row16_max(k);
    return umax64(umax64(readlane64(k, 0), readlane64(k, 16)), umax64(readlane64(k, 32), readlane64(k, 48)));
}

// ---- small fronts (k <= 64 interior cells, m <= 16 MB rows): the whole
// elimination of the interior block in one kernel, one workgroup per node.
// A swap-free Gauss-Jordan sweep over the k interior pivots (partial
// pivoting among the interior rows not yet used: largest |a|^2, ties to the
// lower row) turns F in place into the sweep [[Q, X], [-G, S]] with the
// interior rows in pivot order: pivot p, taken in column p at row r_p,
// leaves in row r_p what row p of a row-exchanging sweep holds, and that
// sweep's Q and -G columns are those of the exchanged rows.  The store
// undoes both: element (x, y) goes to row pi(x) (the pivot that took row x)
// for x < k, to column r_y for y < k, negated in the G block.  The result is
// what the LU + inverse + products path stores: Q = F_II^-1, X = Q F_IB,
// G = F_BI Q, S = F_BB - F_BI X.  Thread t holds rows r0 + 16a and columns
// c0 + 16c (r0 = t % 16, c0 = t / 16) in registers; per pivot the pivot
// column goes through LDS (its owners write it, every wave finds the pivot
// row by the same butterfly argmax), then the scaled pivot row, then every
// thread updates its MB x MB entries: a - col_i row_j, -col_i / a_pp in the
// pivot column, the scaled row in the pivot row.  Two barriers per pivot,
// the LDS vectors double-buffered.  info: first zero pivot (1-based), the
// rocSOLVER convention.
constexpr int kGjMaxK = 64;
constexpr int kGjMaxM = 112;   // fronts of up to 7 x 16 rows (registers: 49 complex per thread)

template <int MB>
__global__ void __launch_bounds__(256) k_nd_gj(cx<float>* __restrict__ F, int m, int k,
                                               rocblas_int* __restrict__ info) {
    __shared__ cx<float> colp[2][16 * MB];
    __shared__ cx<float> rowp[2][16 * MB];
    __shared__ int s_row_of[kGjMaxK];   // pivot p -> its row r_p
    __shared__ int s_pi[kGjMaxK];       // row x -> the pivot that took it
    const int t = threadIdx.x, r0 = t & 15, c0 = t >> 4, lane = t & 63;
    cx<float>* Fe = F + (size_t)blockIdx.x * m * m;
    cx<float> v[MB][MB];
#pragma unroll
    for (int a = 0; a < MB; ++a)
#pragma unroll
        for (int c = 0; c < MB; ++c) {
            const int i = r0 + 16 * a, j = c0 + 16 * c;
            v[a][c] = (i < m && j < m) ? Fe[(size_t)j * m + i] : czero<float>();
        }
    uint64_t used = 0;   // rows taken as pivots (uniform)
    int bad = 0;         // first zero pivot, 1-based (uniform)
#pragma unroll
    for (int pc = 0; pc < kGjMaxK / 16; ++pc) {
        if (pc * 16 >= k) break;
        for (int pp = 0; pp < 16; ++pp) {
            const int p = pc * 16 + pp;
            if (p >= k) break;
            const int buf = p & 1;
            if (c0 == pp) {   // owners of column p
#pragma unroll
                for (int a = 0; a < MB; ++a) colp[buf][r0 + 16 * a] = v[a][pc];
            }
            __syncthreads();
            // pivot row: the same butterfly argmax in every wave
            // the largest unused |a|^2 of column p, ties to the lower row: one
            // key reduction, the same in every wave.  No comparable entry (a
            // NaN column): the first unused row, reported as a zero pivot.
            const uint64_t best = wave_max64((lane < k && !((used >> lane) & 1ull))
                                                 ? piv_key(norm2(colp[buf][lane]), lane) : 0ull);
            const int r = best ? (int)~(uint32_t)best : __builtin_ctzll(~used);
            const cx<float> apiv = colp[buf][r];
            const bool zero_piv = best == 0 || norm2(apiv) == 0.f;
            const cx<float> inv = zero_piv ? czero<float>() : crecip(apiv);
            if (zero_piv && bad == 0) bad = p + 1;
            const int rhi = r >> 4, rlo = r & 15;
            if (r0 == rlo) {   // owners of row r: the scaled pivot row
#pragma unroll
                for (int c = 0; c < MB; ++c) {
                    cx<float> x = v[0][c];
#pragma unroll
                    for (int a = 1; a < MB; ++a)
                        if (a == rhi) x = v[a][c];
                    rowp[buf][c0 + 16 * c] = (c == pc && c0 == pp) ? inv : x * inv;
                }
            }
            if (t == 0) s_row_of[p] = r;
            used |= 1ull << r;
            __syncthreads();
            cx<float> cr[MB], rr[MB];
#pragma unroll
            for (int a = 0; a < MB; ++a) cr[a] = colp[buf][r0 + 16 * a];
#pragma unroll
            for (int c = 0; c < MB; ++c) rr[c] = rowp[buf][c0 + 16 * c];
#pragma unroll
            for (int a = 0; a < MB; ++a) {
                const bool in_r = (a == rhi) && (r0 == rlo);
#pragma unroll
                for (int c = 0; c < MB; ++c) {
                    cx<float> u = v[a][c] - cr[a] * rr[c];
                    if (c == pc && c0 == pp) u = -(cr[a] * inv);
                    if (in_r) u = rr[c];
                    v[a][c] = u;
                }
            }
        }
    }
    __syncthreads();
    if (t < k) s_pi[s_row_of[t]] = t;
    if (t == 0) info[blockIdx.x] = bad;
    __syncthreads();
#pragma unroll
    for (int a = 0; a < MB; ++a)
#pragma unroll
        for (int c = 0; c < MB; ++c) {
            const int x = r0 + 16 * a, y = c0 + 16 * c;
            if (x < m && y < m) {
                const int orow = x < k ? s_pi[x] : x;
                const int ocol = y < k ? s_row_of[y] : y;
                Fe[(size_t)ocol * m + orow] = (x >= k && y < k) ? -v[a][c] : v[a][c];
            }
        }
}

// ---- fronts with at most 64 interior cells and up to 448 rows: the same
// swap-free sweep on the left m x k block only (the sweep's first k columns
// evolve on their own), which gives Q (rows in pivot order) over F_II and
// -G over F_BI; the store undoes the permutations and the sign as k_nd_gj's
// does, and X = Q F_IB, S = F_BB - G F_IB follow by two GEMMs on the
// untouched F_IB.  1024 threads: thread t holds rows r0 + 64a (r0 = t % 64)
// and columns c0 + 16c (c0 = t / 64) of the block.
constexpr int kGjPanelMaxM = 448;

template <int MR>
__global__ void __launch_bounds__(1024) k_nd_gj_panel(cx<float>* __restrict__ F, int m, int k,
                                                      rocblas_int* __restrict__ info) {
    constexpr int MC = kGjMaxK / 16;
    __shared__ cx<float> colp[2][64 * MR];
    __shared__ cx<float> rowp[2][kGjMaxK];
    __shared__ int s_row_of[kGjMaxK];
    __shared__ int s_pi[kGjMaxK];
    const int t = threadIdx.x, r0 = t & 63, c0 = t >> 6, lane = t & 63;
    cx<float>* Fe = F + (size_t)blockIdx.x * m * m;
    cx<float> v[MR][MC];
#pragma unroll
    for (int a = 0; a < MR; ++a)
#pragma unroll
        for (int c = 0; c < MC; ++c) {
            const int i = r0 + 64 * a, j = c0 + 16 * c;
            v[a][c] = (i < m && j < k) ? Fe[(size_t)j * m + i] : czero<float>();
        }
    uint64_t used = 0;
    int bad = 0;
#pragma unroll
    for (int pc = 0; pc < MC; ++pc) {
        if (pc * 16 >= k) break;
        for (int pp = 0; pp < 16; ++pp) {
            const int p = pc * 16 + pp;
            if (p >= k) break;
            const int buf = p & 1;
            if (c0 == pp) {   // owners of column p
#pragma unroll
                for (int a = 0; a < MR; ++a) colp[buf][r0 + 64 * a] = v[a][pc];
            }
            __syncthreads();
            // the largest unused |a|^2 of column p, ties to the lower row: one
            // key reduction, the same in every wave.  No comparable entry (a
            // NaN column): the first unused row, reported as a zero pivot.
            const uint64_t best = wave_max64((lane < k && !((used >> lane) & 1ull))
                                                 ? piv_key(norm2(colp[buf][lane]), lane) : 0ull);
            const int r = best ? (int)~(uint32_t)best : __builtin_ctzll(~used);   // < k <= 64: row r sits at a = 0
            const cx<float> apiv = colp[buf][r];
            const bool zero_piv = best == 0 || norm2(apiv) == 0.f;
            const cx<float> inv = zero_piv ? czero<float>() : crecip(apiv);
            if (zero_piv && bad == 0) bad = p + 1;
            if (r0 == r) {   // owners of row r (its entries are v[0][*])
#pragma unroll
                for (int c = 0; c < MC; ++c)
                    rowp[buf][c0 + 16 * c] = (c == pc && c0 == pp) ? inv : v[0][c] * inv;
            }
            if (t == 0) s_row_of[p] = r;
            used |= 1ull << r;
            __syncthreads();
            cx<float> cr[MR], rr[MC];
#pragma unroll
            for (int a = 0; a < MR; ++a) cr[a] = colp[buf][r0 + 64 * a];
#pragma unroll
            for (int c = 0; c < MC; ++c) rr[c] = rowp[buf][c0 + 16 * c];
#pragma unroll
            for (int a = 0; a < MR; ++a) {
                const bool in_r = (a == 0) && (r0 == r);
#pragma unroll
                for (int c = 0; c < MC; ++c) {
                    cx<float> u = v[a][c] - cr[a] * rr[c];
                    if (c == pc && c0 == pp) u = -(cr[a] * inv);
                    if (in_r) u = rr[c];
                    v[a][c] = u;
                }
            }
        }
    }
    __syncthreads();
    if (t < k) s_pi[s_row_of[t]] = t;
    if (t == 0) info[blockIdx.x] = bad;
    __syncthreads();
#pragma unroll
    for (int a = 0; a < MR; ++a)
#pragma unroll
        for (int c = 0; c < MC; ++c) {
            const int x = r0 + 64 * a, y = c0 + 16 * c;
            if (x < m && y < k) {
                const int orow = x < k ? s_pi[x] : x;
                Fe[(size_t)s_row_of[y] * m + orow] = (x >= k) ? -v[a][c] : v[a][c];
            }
        }
}

// ---- panels of the large interior blocks (kLuMin < k <= kLuMaxRows): LU
// with partial pivoting of kLuPanel columns at a time, L (unit, strictly
// lower) and U in place of the panel, ipiv 1-based sequential row
// interchanges (the LAPACK form); the blocked Gauss-Jordan below carries
// each panel to the rest of the augmented matrix.

// LU of the panel rows j0..k-1, columns j0..j0+nbw-1 of every node
// (blockIdx.x; node e's matrix at F + e nstride, leading dimension ld), NT
// threads holding RPT rows each in registers.  Rows are not
// moved while the panel is factorised: each thread keeps the current
// (interchanged) position of its rows, and per column every wave publishes
// its candidate pivot (largest |a|^2, ties to the lower position) with the
// candidate row to LDS; after one barrier all waves pick the same pivot,
// exchange the two positions and apply the rank-1 update to the rows below
// it.  The LDS slots alternate between even and odd columns, so one barrier
// per column suffices: a wave writes column c + 2's slots only after the
// barrier of column c + 1, which every wave reaches after reading column
// c's.  The rows are stored at their final positions, then ipiv (the
// sequential interchanges) and the first zero pivot -- the same values,
// operation for operation, as exchanging the rows at every column.
template <int NT, int RPT>
__global__ void __launch_bounds__(NT)
k_lu_panel(cx<float>* __restrict__ F, int ld, int64_t nstride, int k, int j0, int nbw, rocblas_int* __restrict__ ipiv,
           rocblas_int* __restrict__ info) {
    constexpr int NW = NT / 64;
    static_assert(NW <= 16, "the wave keys are reduced within one row of 16 lanes");
    __shared__ uint64_t s_key[2][NW];
    __shared__ cx<float> s_row[2][NW][kLuPanel];
    const int e = blockIdx.x, tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const int nr = k - j0;
    cx<float>* Fe = F + (int64_t)e * nstride;
    cx<float> a[RPT][kLuPanel];
    int pos[RPT];   // current position of each held row (-1: no row)
#pragma unroll
    for (int q = 0; q < RPT; ++q) {
        const int r = tid + q * NT;
        const int rr = min(r, nr - 1);
        pos[q] = r < nr ? r : -1;
#pragma unroll
        for (int c = 0; c < kLuPanel; ++c) {
            const cx<float> v = Fe[(int64_t)(j0 + min(c, nbw - 1)) * ld + j0 + rr];
            a[q][c] = (r < nr && c < nbw) ? v : czero<float>();
        }
    }
    int zero_at = 0;
    int pv[kLuPanel];   // position exchanged with c at column c (thread 0's record)
#pragma unroll
    for (int c = 0; c < kLuPanel; ++c) {
        if (c < nbw) {
            const int par = c & 1;
            uint64_t key = 0;
#pragma unroll
            for (int q = 0; q < RPT; ++q)
                if (pos[q] >= c) key = umax64(key, piv_key(norm2(a[q][c]), pos[q]));
            const uint64_t wkey = wave_max64(key);
            if (wkey != 0) {
                const int wpos = (int)~(uint32_t)wkey;
#pragma unroll
                for (int q = 0; q < RPT; ++q)
                    if (pos[q] == wpos)
#pragma unroll
                        for (int c2 = c; c2 < kLuPanel; ++c2) s_row[par][w][c2] = a[q][c2];
            }
            if (lane == 0) s_key[par][w] = wkey;
            __syncthreads();
            // the workgroup's pivot: the largest of the NW wave keys (unique
            // unless zero: positions differ), found by every wave
            const uint64_t k0 = lane < NW ? s_key[par][lane] : 0ull;
            const uint64_t gkey = readlane64(row16_max(k0), 0);
            int p, ws;
            if (gkey == 0) {
                // no comparable entry (a NaN column): keep the row at position
                // c; the zero-pivot test flags it.  Uniform: every wave read
                // the same keys.
#pragma unroll
                for (int q = 0; q < RPT; ++q)
                    if (pos[q] == c)
#pragma unroll
                        for (int c2 = c; c2 < kLuPanel; ++c2) s_row[par][0][c2] = a[q][c2];
                __syncthreads();
                p = c;
                ws = 0;
            } else {
                p = (int)~(uint32_t)gkey;
                ws = __builtin_ctzll(__ballot(k0 == gkey));
            }
            cx<float> pr[kLuPanel];
#pragma unroll
            for (int c2 = c; c2 < kLuPanel; ++c2) pr[c2] = s_row[par][ws][c2];
            pv[c] = p;
#pragma unroll
            for (int q = 0; q < RPT; ++q) {
                if (pos[q] == p)
                    pos[q] = c;
                else if (pos[q] == c)
                    pos[q] = p;
            }
            const cx<float> d = pr[c];
            if (norm2(d) > 0.f) {
                const cx<float> inv = crecip(d);
#pragma unroll
                for (int q = 0; q < RPT; ++q)
                    if (pos[q] > c) {
                        const cx<float> l = a[q][c] * inv;
                        a[q][c] = l;
#pragma unroll
                        for (int c2 = c + 1; c2 < kLuPanel; ++c2) a[q][c2] = a[q][c2] - l * pr[c2];
                    }
            } else if (zero_at == 0) {
                zero_at = j0 + c + 1;
            }
        }
    }
#pragma unroll
    for (int q = 0; q < RPT; ++q)
        if (pos[q] >= 0)
#pragma unroll
            for (int c = 0; c < kLuPanel; ++c)
                if (c < nbw) Fe[(int64_t)(j0 + c) * ld + j0 + pos[q]] = a[q][c];
    if (tid == 0) {
#pragma unroll
        for (int c = 0; c < kLuPanel; ++c)
            if (c < nbw) ipiv[(int64_t)e * k + j0 + c] = j0 + pv[c] + 1;
        if (zero_at && info[e] == 0) info[e] = zero_at;
    }
}

// ---- blocked Gauss-Jordan on the augmented matrix [F_II | R], R = [I F_IB]
// (k x m per node, leading dimension k, node stride k m).  Each step reads
// the augmented matrix from one of two buffers and writes every column
// right of its panel (all k rows) to the other: buffer 0 is F_II in the
// front (leading dimension m, node stride m^2) beside R, buffer 1 one
// k x (k + m) matrix per node (leading dimension k).  Step p factorises its
// 16-column panel in buffer p mod 2 (k_lu_panel), then k_gj_step carries the
// panel's interchanges and L11^-1 to the right columns and eliminates the
// panel's columns from every other row (rows below with L21, rows above
// with F[row, panel] U11^-1).  F_II ends block diagonal with the panels'
// U11 (panel q's in buffer q mod 2), and k_gj_final solves each block,
// leaving [Q X] = F_II^-1 [I F_IB] in the right part of the last buffer.
// Columns left of a panel are not needed again and are not carried.

// Column j (0 <= j < k + m) of node e's augmented matrix in buffer b.
__device__ __forceinline__ cx<float>* aug_col(cx<float>* F, cx<float>* R, cx<float>* B1, int m, int k, int e, int b,
                                              int j) {
    if (b) return B1 + (int64_t)e * k * (k + m) + (int64_t)j * k;
    return j < k ? F + (int64_t)e * m * m + (int64_t)j * m : R + (int64_t)e * k * m + (int64_t)(j - k) * k;
}

// One step of the blocked Gauss-Jordan for the panel j0..j0+nbw-1 (factorised
// in buffer sb), writing the augmented columns j0 + nbw .. k + m - 1 of rows
// 0..k-1 to buffer 1 - sb.  One workgroup per 64 x 64 tile (grid.x row tiles,
// grid.y column tiles, grid.z nodes).  Every workgroup derives the panel
// rows' new values for its columns itself, from the unmodified source: the
// interchanges touch at most 2 nbw positions, and lane i (mod 32) traces
// entry i -- position j0 + i (i < nbw) or the target of interchange i - nbw
// (when below the panel and not an earlier target) -- back through the
// interchanges to the position its value comes from; lanes gather their
// entry of 8 columns, one thread per column applies L11^-1 to the panel
// rows (U12), and another wave forms the multipliers of the tile's rows
// above the panel (w U11 = F[row, panel]).  Outputs: U12 in the panel rows;
// elsewhere the row's value after the interchanges (the moved value for a
// target row) minus its multiplier times U12.  The tile's outputs are loaded
// first so that their latency overlaps the rest; each thread accumulates a
// 4 x 4 block (rows tr + 16i, columns tc + 16j).
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(1, 4)))
k_gj_step(cx<float>* __restrict__ F, cx<float>* __restrict__ R, cx<float>* __restrict__ B1, int m, int k, int j0,
          int nbw, int sb, const rocblas_int* __restrict__ ipiv) {
    constexpr int NE = 2 * kLuPanel;
    __shared__ cx<float> Ls[kLuPanel][64];        // multipliers of the tile's rows
    __shared__ cx<float> G[NE][64 + 1];           // gathered entries of the tile's columns (U12, moved values)
    __shared__ cx<float> PB[kLuPanel][kLuPanel + 1];   // the panel block: L11 below, U11 on and above the diagonal
    __shared__ cx<float> rdiag[kLuPanel];
    __shared__ int tmap[64];                      // tile row -> target entry, -1 if none
    const int t = threadIdx.x, tr = t & 15, tc = t >> 4, lane = t & 63, li = lane & 31;
    const int r0 = blockIdx.x * 64;
    const int cb = j0 + nbw + blockIdx.y * 64;
    const int e = blockIdx.z, db = 1 - sb;
    const int ncol = k + m;
    cx<float> cv[4][4];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const cx<float>* col = aug_col(F, R, B1, m, k, e, sb, min(cb + tc + 16 * j, ncol - 1));
#pragma unroll
        for (int i = 0; i < 4; ++i) cv[i][j] = col[min(r0 + tr + 16 * i, k - 1)];
    }
    // the interchanges, traced per entry (every wave holds the same entries)
    const int tl = li < kLuPanel ? li : li - kLuPanel;
    const int my_pv = (tl < nbw) ? ipiv[(int64_t)e * k + j0 + tl] - 1 : -1;
    int pvu[kLuPanel];
#pragma unroll
    for (int u = 0; u < kLuPanel; ++u) pvu[u] = __builtin_amdgcn_readlane(my_pv, u);
    int dst = -1, src = -1;
    if (tl < nbw) {
        const int x = li < kLuPanel ? j0 + li : my_pv;
        bool keep = li < kLuPanel || x >= j0 + nbw;
        int y = x;
#pragma unroll
        for (int u = kLuPanel - 1; u >= 0; --u)
            if (u < nbw) {
                if (li >= kLuPanel && u < tl && pvu[u] == x) keep = false;
                if (y == j0 + u)
                    y = pvu[u];
                else if (y == pvu[u])
                    y = j0 + u;
            }
        if (keep) {
            dst = x;
            src = y;
        }
    }
    int tdst[kLuPanel];   // target positions below the panel, uniform; -1: none
#pragma unroll
    for (int i = 0; i < kLuPanel; ++i) tdst[i] = __builtin_amdgcn_readlane(dst, kLuPanel + i);
    const int cg = t >> 5;
#pragma unroll
    for (int q = 0; q < 8; ++q) {
        const int c = cg + 8 * q, j = cb + c;
        G[li][c] = (dst >= 0 && j < ncol) ? aug_col(F, R, B1, m, k, e, sb, j)[src] : czero<float>();
    }
    {
        const int a = t & 15, bc = t >> 4;
        PB[a][bc] = bc < nbw && a < nbw ? aug_col(F, R, B1, m, k, e, sb, j0 + bc)[j0 + a] : czero<float>();
    }
#pragma unroll
    for (int q = 0; q < kLuPanel * 64 / 256; ++q) {
        const int idx = t + 256 * q;
        const int kk = idx >> 6, r = idx & 63;
        Ls[kk][r] = (r0 + r < k && kk < nbw) ? aug_col(F, R, B1, m, k, e, sb, j0 + kk)[r0 + r] : czero<float>();
    }
    if (t < 64) {
        int f = -1;
#pragma unroll
        for (int i = 0; i < kLuPanel; ++i)
            if (tdst[i] == r0 + t) f = i;
        tmap[t] = f;
    }
    if (t < kLuPanel) rdiag[t] = t < nbw ? crecip(aug_col(F, R, B1, m, k, e, sb, j0 + t)[j0 + t]) : czero<float>();
    __syncthreads();
    if (t < 64) {   // U12 of column t: L11^-1 on the gathered panel rows
        cx<float> x[kLuPanel];
#pragma unroll
        for (int a = 0; a < kLuPanel; ++a) x[a] = G[a][t];
#pragma unroll
        for (int a = 0; a < kLuPanel; ++a)
#pragma unroll
            for (int u = 0; u < kLuPanel; ++u)
                if (u < a && a < nbw) x[a] = x[a] - PB[a][u] * x[u];
#pragma unroll
        for (int a = 0; a < kLuPanel; ++a) G[a][t] = x[a];
    } else if (t < 128 && r0 + (t - 64) < j0) {   // rows above the panel (second wave): w U11 = F[row, panel]
        const int rr = t - 64;   // w replaces the row's entries in Ls as it is solved
        for (int a = 0; a < kLuPanel; ++a) {
            cx<float> x = Ls[a][rr];
            for (int u = 0; u < a; ++u) x = x - Ls[u][rr] * PB[u][a];
            Ls[a][rr] = a < nbw ? x * rdiag[a] : czero<float>();
        }
    }
    __syncthreads();
    cx<float> acc[4][4];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] = czero<float>();
#pragma unroll 2
    for (int kk = 0; kk < kLuPanel; ++kk) {
        cx<float> a[4], b[4];
#pragma unroll
        for (int i = 0; i < 4; ++i) a[i] = Ls[kk][tr + 16 * i];
#pragma unroll
        for (int j = 0; j < 4; ++j) b[j] = G[kk][tc + 16 * j];
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
            for (int j = 0; j < 4; ++j) acc[i][j] = acc[i][j] + a[i] * b[j];
    }
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const int cl = tc + 16 * j, c = cb + cl;
        if (c < ncol) {
            cx<float>* col = aug_col(F, R, B1, m, k, e, db, c);
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const int rl = tr + 16 * i, r = r0 + rl;
                if (r < k) {
                    // the panel row's U12, or the (moved) value minus the update; both
                    // LDS reads are unconditional and opaque, so that the choice stays a
                    // choice of values and cv stays in registers
                    const int f = tmap[rl];
                    cx<float> u12 = G[min(max(r - j0, 0), kLuPanel - 1)][cl];
                    cx<float> mv = G[kLuPanel + max(f, 0)][cl];
                    asm volatile("" : "+v"(u12.re), "+v"(u12.im), "+v"(mv.re), "+v"(mv.im));
                    const cx<float> base = f >= 0 ? mv : cv[i][j];
                    col[r] = (r >= j0 && r < j0 + nbw) ? u12 : base - acc[i][j];
                }
            }
        }
    }
}

// [Q X] = U^-1 Z per panel's block of rows once every panel is eliminated
// (F_II is then block diagonal with the panels' U11; panel q's in buffer
// q mod 2, the right-hand sides in buffer fin): one thread per (block,
// column of [Q X]), grid.x column groups, grid.y blocks, grid.z nodes.
__global__ void __launch_bounds__(256)
k_gj_final(cx<float>* __restrict__ F, cx<float>* __restrict__ R, cx<float>* __restrict__ B1, int m, int k, int fin) {
    __shared__ cx<float> U[kLuPanel][kLuPanel + 1];
    __shared__ cx<float> rd[kLuPanel];
    const int q = blockIdx.y, j0 = q * kLuPanel, nbw = min(kLuPanel, k - j0);
    const int e = blockIdx.z;
    const int t = threadIdx.x;
    {
        const int a = t & 15, bc = t >> 4;
        U[a][bc] = (a <= bc && bc < nbw) ? aug_col(F, R, B1, m, k, e, q & 1, j0 + bc)[j0 + a] : czero<float>();
    }
    __syncthreads();
    if (t < kLuPanel) rd[t] = t < nbw ? crecip(U[t][t]) : czero<float>();
    __syncthreads();
    const int c = blockIdx.x * 256 + t;
    if (c >= m) return;
    cx<float>* col = aug_col(F, R, B1, m, k, e, fin, k + c) + j0;
    cx<float> y[kLuPanel];
#pragma unroll
    for (int a = 0; a < kLuPanel; ++a) y[a] = a < nbw ? col[a] : czero<float>();
#pragma unroll
    for (int a = kLuPanel - 1; a >= 0; --a)
        if (a < nbw) {
            cx<float> x = y[a];
#pragma unroll
            for (int u = 0; u < kLuPanel; ++u)
                if (u > a && u < nbw) x = x - U[a][u] * y[u];
            y[a] = x * rd[a];
        }
#pragma unroll
    for (int a = 0; a < kLuPanel; ++a)
        if (a < nbw) col[a] = y[a];
}

// R = [I F_IB] (k x m per node, column-major, leading dimension k)
__global__ void __launch_bounds__(256)
k_nd_rhs(const cx<float>* __restrict__ F, cx<float>* __restrict__ R, int nb, int k, int m) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)nb * k * m) return;
    const int64_t km = (int64_t)k * m;
    const int e = (int)(t / km);
    const int64_t r = t - (int64_t)e * km;
    const int j = (int)(r / k), i = (int)(r - (int64_t)j * k);
    R[t] = j < k ? (i == j ? mk<float>(1.f, 0.f) : czero<float>()) : F[(int64_t)e * m * m + (int64_t)j * m + i];
}

// copy a rows x cols block between per-node column-major matrices
__global__ void __launch_bounds__(256)
k_nd_block_copy(const cx<float>* __restrict__ src, int64_t src_stride, int src_ld, cx<float>* __restrict__ dst,
                int64_t dst_stride, int dst_ld, int nb, int rows, int cols) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t per = (int64_t)rows * cols;
    if (t >= (int64_t)nb * per) return;
    const int e = (int)(t / per);
    const int64_t r = t - (int64_t)e * per;
    const int j = (int)(r / rows), i = (int)(r - (int64_t)j * rows);
    dst[(int64_t)e * dst_stride + (int64_t)j * dst_ld + i] = src[(int64_t)e * src_stride + (int64_t)j * src_ld + i];
}

// One 64 x 64 tile of a per-node product over a block of F (rows row0..,
// columns col0.., M x N, column-major with leading dimension m).  The input
// vector is gathered from the work vector w through xcells (xn entries per
// node, -1 = padding).  TRANS = 0: y = A x, the four waves take 16 columns
// each, lanes the rows, the wave sums combined through LDS; partial sums per
// column tile in part[(e * tn + tj) * M + i].  TRANS = 1: y = A^T x, lanes
// the rows, each wave 16 output columns reduced over the lanes; partial sums
// per row tile in part[(e * tm + ti) * N + j].  Workgroups are node-major.
template <bool TRANS>
__global__ void __launch_bounds__(256)
k_nd_mv(const cx<float>* __restrict__ F, int m, int row0, int col0, int M, int N, int tm, int tn,
        const int32_t* __restrict__ xcells, int xn, const cx<float>* __restrict__ w, cx<float>* __restrict__ part) {
    const int wid = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int per = tm * tn;
    const int e = blockIdx.x / per;
    const int rem = blockIdx.x - e * per;
    const int ti = rem / tn, tj = rem - ti * tn;
    const int i = ti * kMvTile + lane;
    const cx<float>* A = F + (int64_t)e * m * m + row0 + (int64_t)col0 * m;
    const int32_t* xc = xcells + (int64_t)e * xn;
    const int j0 = tj * kMvTile + wid * 16;
    if (!TRANS) {
        __shared__ cx<float> red[4][kMvTile];
        cx<float> acc = czero<float>();
        if (i < M) {
#pragma unroll
            for (int c = 0; c < 16; ++c) {
                const int j = j0 + c;
                if (j < N) {
                    const int32_t cell = xc[j];
                    const cx<float> xj = cell >= 0 ? w[cell] : czero<float>();
                    acc = acc + A[i + (int64_t)j * m] * xj;
                }
            }
        }
        red[wid][lane] = acc;
        __syncthreads();
        if (wid == 0 && i < M) {
            const cx<float> v = red[0][lane] + red[1][lane] + red[2][lane] + red[3][lane];
            part[((int64_t)e * tn + tj) * M + i] = v;
        }
    } else {
        cx<float> xi = czero<float>();
        if (i < M) {
            const int32_t cell = xc[i];
            if (cell >= 0) xi = w[cell];
        }
#pragma unroll
        for (int c = 0; c < 16; ++c) {
            const int j = j0 + c;
            cx<float> v = (i < M && j < N) ? A[i + (int64_t)j * m] * xi : czero<float>();
#pragma unroll
            for (int off = 32; off > 0; off >>= 1) {
                v.re += __shfl_xor(v.re, off);
                v.im += __shfl_xor(v.im, off);
            }
            if (lane == 0 && j < N) part[((int64_t)e * tm + ti) * N + j] = v;
        }
    }
}

// The same product for nodes of at most kMvNode rows and columns, one
// workgroup per node, fused with the gather and the scatter (a node's input
// and output cells belong to that node alone within a level): the input is
// gathered into LDS, then one thread per output, 16 independent loads in
// flight: TRANS = 0 an output row (the loads of a wave are coalesced), TRANS
// = 1 an output column (each thread reads its column contiguously).
// Outputs o < k go to the interior cells (w = y when set_interior, else
// w -= y), o >= k to the contribution buffer tb (e * b + o - k), which
// k_nd_bsum subtracts from the boundary cells.
constexpr int kMvNode = 256;
template <bool TRANS>
__global__ void __launch_bounds__(256)
k_nd_mv_node(const cx<float>* __restrict__ F, int m, int row0, int col0, int M, int N,
             const int32_t* __restrict__ xcells, int xn, const int32_t* __restrict__ icells, int k, int b,
             cx<float>* __restrict__ w, cx<float>* __restrict__ tb, int set_interior) {
    __shared__ cx<float> xs[kMvNode];
    const int e = blockIdx.x, tid = threadIdx.x;
    const int32_t* xc = xcells + (int64_t)e * xn;
    for (int q = tid; q < xn; q += blockDim.x) {
        const int32_t c = xc[q];
        xs[q] = c >= 0 ? w[c] : czero<float>();
    }
    __syncthreads();
    const int o = tid;
    const int nout = TRANS ? N : M, nin = TRANS ? M : N;
    if (o >= nout) return;
    const cx<float>* A = F + (int64_t)e * m * m + row0 + (int64_t)col0 * m;
    // element q of output o's row (TRANS = 0) or column (TRANS = 1)
    const cx<float>* base = TRANS ? A + (int64_t)o * m : A + o;
    const int64_t step = TRANS ? 1 : m;
    cx<float> acc = czero<float>();
    int q = 0;
    for (; q + 16 <= nin; q += 16) {
        cx<float> a[16];
#pragma unroll
        for (int u = 0; u < 16; ++u) a[u] = base[(int64_t)(q + u) * step];
#pragma unroll
        for (int u = 0; u < 16; ++u) acc = acc + a[u] * xs[q + u];
    }
    for (; q + 4 <= nin; q += 4) {
        cx<float> a[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) a[u] = base[(int64_t)(q + u) * step];
#pragma unroll
        for (int u = 0; u < 4; ++u) acc = acc + a[u] * xs[q + u];
    }
    for (; q < nin; ++q) acc = acc + base[(int64_t)q * step] * xs[q];
    if (o < k) {
        const int32_t c = icells[(int64_t)e * k + o];
        if (c >= 0) w[c] = set_interior ? acc : w[c] - acc;
    } else {
        tb[(int64_t)e * b + (o - k)] = acc;
    }
}

// The non-transposed product for nodes with at most 32 outputs (the down
// sweep's x_I = y_I - X x_B on the small levels), one 64-thread workgroup per
// node: output o = t mod Mp (Mp the power of two >= M) and slice t / Mp of
// the inputs, so every lane has loads in flight; the slices are summed by
// lane exchanges.  Outputs as in k_nd_mv_node (o < k: interior cells, else
// the contribution buffer).
template <int MP>
__global__ void __launch_bounds__(64)
k_nd_mv_node_split(const cx<float>* __restrict__ F, int m, int row0, int col0, int M, int N,
                   const int32_t* __restrict__ xcells, int xn, const int32_t* __restrict__ icells, int k, int b,
                   cx<float>* __restrict__ w, cx<float>* __restrict__ tb, int set_interior) {
    constexpr int S = 64 / MP;
    __shared__ cx<float> xs[kMvNode];
    const int e = blockIdx.x, tid = threadIdx.x;
    const int32_t* xc = xcells + (int64_t)e * xn;
    for (int q = tid; q < xn; q += 64) {
        const int32_t c = xc[q];
        xs[q] = c >= 0 ? w[c] : czero<float>();
    }
    __syncthreads();
    const int o = tid % MP, sl = tid / MP;
    const cx<float>* A = F + (int64_t)e * m * m + row0 + (int64_t)col0 * m;
    cx<float> acc = czero<float>();
    if (o < M) {
        int q = sl;
        for (; q + 3 * S < N; q += 4 * S) {
            cx<float> a[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) a[u] = A[o + (int64_t)(q + u * S) * m];
#pragma unroll
            for (int u = 0; u < 4; ++u) acc = acc + a[u] * xs[q + u * S];
        }
        for (; q < N; q += S) acc = acc + A[o + (int64_t)q * m] * xs[q];
    }
#pragma unroll
    for (int off = MP; off < 64; off <<= 1) {
        acc.re += __shfl_xor(acc.re, off);
        acc.im += __shfl_xor(acc.im, off);
    }
    if (sl != 0 || o >= M) return;
    if (o < k) {
        const int32_t c = icells[(int64_t)e * k + o];
        if (c >= 0) w[c] = set_interior ? acc : w[c] - acc;
    } else {
        tb[(int64_t)e * b + (o - k)] = acc;
    }
}

// Sum of the nt partial-sum tiles of each output o (Mout per node): o < k
// to the interior cells (w = sum when set_interior, else w -= sum), o >= k
// to the contribution buffer tb (e * b + o - k).
__global__ void __launch_bounds__(256)
k_nd_mv_reduce(const cx<float>* __restrict__ part, int nt, int Mout, int nb, const int32_t* __restrict__ icells,
               int k, int b, cx<float>* __restrict__ w, cx<float>* __restrict__ tb, int set_interior) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)nb * Mout) return;
    const int e = (int)(t / Mout), o = (int)(t - (int64_t)e * Mout);
    cx<float> sum = czero<float>();
    for (int q = 0; q < nt; ++q) sum = sum + part[((int64_t)e * nt + q) * Mout + o];
    if (o < k) {
        const int32_t c = icells[(int64_t)e * k + o];
        if (c >= 0) w[c] = set_interior ? sum : w[c] - sum;
    } else {
        tb[(int64_t)e * b + (o - k)] = sum;
    }
}

// w[cell] -= the sum of its contributions, for each boundary cell of a level
// (CSR over the level's nodes, fixed order)
__global__ void __launch_bounds__(256)
k_nd_bsum(const int32_t* __restrict__ cells, const int32_t* __restrict__ ptr, const int32_t* __restrict__ idx, int n,
          const cx<float>* __restrict__ tb, cx<float>* __restrict__ w) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    cx<float> sum = czero<float>();
    for (int q = ptr[t]; q < ptr[t + 1]; ++q) sum = sum + tb[idx[q]];
    const int32_t c = cells[t];
    w[c] = w[c] - sum;
}

__global__ void __launch_bounds__(256) k_nd_copy(const cx<float>* __restrict__ a, cx<float>* __restrict__ b, int64_t n) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) b[t] = a[t];
}

int nblk(int64_t n) { return (int)((n + 255) / 256); }

// ------------------------------------------------------- library handle
// One rocBLAS handle (rocSOLVER uses rocBLAS handles) per (device, stream),
// created on first use, bound to its stream once and kept for the process:
// threads factorising on different streams never share a handle, so a
// handle's stream cannot be switched under another thread's calls.  Callers
// use a few long-lived streams (the per-instance workers keep theirs).
std::mutex g_handle_mu;
struct HandleSlot {
    int dev;
    hipStream_t stream;
    rocblas_handle h;
};
std::vector<HandleSlot> g_handles;

int get_handle(hipStream_t s, rocblas_handle* out) {
    int dev = 0;
    CV_HIP(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lock(g_handle_mu);
    for (const auto& slot : g_handles)
        if (slot.dev == dev && slot.stream == s) {
            *out = slot.h;
            return CV_OK;
        }
    CV_CHECK(g_handles.size() < 256, "too many streams with rocBLAS handles");
    rocblas_handle h = nullptr;
    CV_CHECK(rocblas_create_handle(&h) == rocblas_status_success, "rocblas_create_handle failed");
    CV_CHECK(rocblas_set_stream(h, s) == rocblas_status_success, "rocblas_set_stream failed");
    g_handles.push_back({dev, s, h});
    *out = h;
    return CV_OK;
}

#define CV_RB(expr)                                                                     \
    do {                                                                                \
        const rocblas_status _st = (expr);                                              \
        if (_st != rocblas_status_success) {                                            \
            ::cv::set_error("%s:%d %s: rocblas status %d", __FILE__, __LINE__, #expr, (int)_st); \
            return CV_ERR_HIP;                                                          \
        }                                                                               \
    } while (0)

using rc_t = rocblas_float_complex;

}  // namespace

// M^-1 v (transpose 0) or M^-T v (transpose 1) with the nested-dissection
// factor; v and z complex64 over the grid (may alias).
int nd_apply(const cv_nd_plan* P, const void* plan_dev, const void* factor, const cx<float>* v, cx<float>* z,
             int transpose, hipStream_t s) {
    const char* pb = (const char*)plan_dev;
    char* fb = (char*)factor;
    const int64_t N = (int64_t)P->nx * P->ny;
    cx<float>* w = (cx<float>*)(fb + P->off_w);
    cx<float>* part = (cx<float>*)(fb + P->off_part);
    cx<float>* tb = (cx<float>*)(fb + P->off_tb);
    k_nd_copy<<<nblk(N), 256, 0, s>>>(v, w, N);
    CV_LAUNCH_CHECK();
    auto tiles = [](int n) { return (n + kMvTile - 1) / kMvTile; };
    // one product per level and sweep: tiles (launch 1), sum + scatter (launch 2)
    auto product = [&](const NdLevel& L, bool trans, int row0, int col0, int M, int N, const int32_t* xcells,
                       int xn, int Mout, int set_interior) -> int {
        const int32_t* ic = (const int32_t*)(pb + L.off_icells);
        const cx<float>* F = (const cx<float>*)(fb + L.off_F);
        if (L.m <= kMvNode) {   // small nodes: one fused launch, one workgroup per node
            if (!trans && M <= 32) {   // few outputs: the inputs split over the lanes
                auto* kern = M <= 8 ? k_nd_mv_node_split<8> : M <= 16 ? k_nd_mv_node_split<16>
                                                                     : k_nd_mv_node_split<32>;
                kern<<<L.nb, 64, 0, s>>>(F, L.m, row0, col0, M, N, xcells, xn, ic, L.kmax, L.bmax, w, tb,
                                         set_interior);
            } else if (!trans)
                k_nd_mv_node<false><<<L.nb, 64 * ((M + 63) / 64), 0, s>>>(F, L.m, row0, col0, M, N, xcells, xn, ic,
                                                                         L.kmax, L.bmax, w, tb, set_interior);
            else
                k_nd_mv_node<true><<<L.nb, 64 * ((N + 63) / 64), 0, s>>>(F, L.m, row0, col0, M, N, xcells, xn, ic,
                                                                        L.kmax, L.bmax, w, tb, set_interior);
            CV_LAUNCH_CHECK();
            return CV_OK;
        }
        const int tm = tiles(M), tn = tiles(N);
        const int64_t nwg = (int64_t)L.nb * tm * tn;
        if (!trans)
            k_nd_mv<false><<<(unsigned)nwg, 256, 0, s>>>(F, L.m, row0, col0, M, N, tm, tn, xcells, xn, w, part);
        else
            k_nd_mv<true><<<(unsigned)nwg, 256, 0, s>>>(F, L.m, row0, col0, M, N, tm, tn, xcells, xn, w, part);
        CV_LAUNCH_CHECK();
        const int nt = trans ? tm : tn;
        k_nd_mv_reduce<<<nblk((int64_t)L.nb * Mout), 256, 0, s>>>(part, nt, Mout, L.nb, ic, L.kmax, L.bmax, w, tb,
                                                                   set_interior);
        CV_LAUNCH_CHECK();
        return CV_OK;
    };
    const int H = (int)P->L.size();
    for (int l = 0; l < H; ++l) {   // up the tree: [y_I; t] = [Q; G] b_I (or [Q X]^T b_I), b_B -= t
        const NdLevel& L = P->L[l];
        if (L.kmax == 0) continue;
        const int32_t* ic = (const int32_t*)(pb + L.off_icells);
        const int rc = !transpose ? product(L, false, 0, 0, L.m, L.kmax, ic, L.kmax, L.m, 1)
                                  : product(L, true, 0, 0, L.kmax, L.m, ic, L.kmax, L.m, 1);
        if (rc != CV_OK) return rc;
        const int nbs = (int)L.bs_cells.size();
        if (nbs > 0) {   // b_B -= t, the contributions of the level's nodes summed per boundary cell
            k_nd_bsum<<<nblk(nbs), 256, 0, s>>>((const int32_t*)(pb + L.off_bs_cells),
                                                (const int32_t*)(pb + L.off_bs_ptr), (const int32_t*)(pb + L.off_bs_idx),
                                                nbs, tb, w);
            CV_LAUNCH_CHECK();
        }
    }
    for (int l = H - 1; l >= 0; --l) {   // down the tree: x_I -= X x_B (or G^T x_B)
        const NdLevel& L = P->L[l];
        if (L.kmax == 0 || L.bmax == 0) continue;
        const int32_t* bc = (const int32_t*)(pb + L.off_bcells);
        const int rc = !transpose ? product(L, false, 0, L.kmax, L.kmax, L.bmax, bc, L.bmax, L.kmax, 0)
                                  : product(L, true, L.kmax, 0, L.bmax, L.kmax, bc, L.bmax, L.kmax, 0);
        if (rc != CV_OK) return rc;
    }
    k_nd_copy<<<nblk(N), 256, 0, s>>>(w, z, N);
    CV_LAUNCH_CHECK();
    return CV_OK;
}

}  // namespace cv

extern "C" int cv_nd_factor(const cv_nd_plan* P, const void* plan_dev, const cv_fdfd_desc* desc,
                            const cv_fdfd_desc* desc_host, const void* prof32, const void* m0_32, const void* m1_32,
                            void* factor, int64_t factor_bytes, void* stream) {
    CV_CHECK(P && plan_dev && desc && desc_host && prof32 && m0_32 && factor, "null pointer argument");
    const cv_fdfd_desc& dh = desc_host[0];
    CV_CHECK(dh.nx == P->nx && dh.ny == P->ny && dh.cell_off == 0, "descriptor does not match the %d x %d plan",
             P->nx, P->ny);
    CV_CHECK(dh.nblk == 1 && dh.n_cpl == 0, "frequency-coupled systems are not supported");
    CV_CHECK(dh.pol == CV_POL_EZ || m1_32, "Hz operator needs m1");
    CV_CHECK(factor_bytes >= P->factor_bytes, "factor buffer too small: %lld < %lld", (long long)factor_bytes,
             (long long)P->factor_bytes);
    hipStream_t s = as_stream(stream);
    rocblas_handle h;
    {
        const int rc = get_handle(s, &h);
        if (rc != CV_OK) return rc;
    }
    const char* pb = (const char*)plan_dev;
    char* fb = (char*)factor;
    const cx<float>* prof = (const cx<float>*)prof32;
    const float* m0 = (const float*)m0_32;
    const float* m1 = (const float*)m1_32;
    rocblas_int* info = (rocblas_int*)(fb + P->off_info);
    const rc_t minus_one{-1.f, 0.f}, one{1.f, 0.f}, zero_c{0.f, 0.f};
    CV_HIP(hipMemsetAsync(info, 0, 4 * (size_t)P->n_nodes, s));
    for (size_t l = 0; l < P->L.size(); ++l) {
        const NdLevel& L = P->L[l];
        rc_t* F = (rc_t*)(fb + L.off_F);
        const rocblas_stride sF = (rocblas_stride)L.m * L.m;
        CV_HIP(hipMemsetAsync(F, 0, 8 * (size_t)L.nb * L.m * L.m, s));
        // frontal matrix: the children's Schur complements (extend-add, one
        // child slot per launch), then the node's own couplings of A
        for (const auto& g : L.groups) {
            const NdLevel& C = P->L[g.child_level];
            if (C.bmax == 0 || g.parent.empty()) continue;
            const int npairs = (int)g.parent.size();
            k_nd_extend<<<nblk((int64_t)npairs * C.bmax * C.bmax), 256, 0, s>>>(
                (const cx<float>*)(fb + C.off_F), C.m, C.kmax, C.bmax, npairs, (const int32_t*)(pb + g.off_parent),
                (const int32_t*)(pb + g.off_child), (const int32_t*)(pb + g.off_map), (cx<float>*)F, L.m);
            CV_LAUNCH_CHECK();
        }
        if (L.kmax > 0)
            k_nd_assemble<<<nblk((int64_t)L.nb * L.kmax), 256, 0, s>>>(
                desc, prof, m0, m1, (const int32_t*)(pb + L.off_icells), (const int32_t*)(pb + L.off_anbr),
                (cx<float>*)F, L.nb, L.kmax, L.m);
        CV_LAUNCH_CHECK();
        if (L.kmax == 0) continue;
        const int k = L.kmax, b = L.bmax, m = L.m;
        rocblas_int* piv = (rocblas_int*)(fb + L.off_piv);
        rc_t* R = (rc_t*)(fb + P->off_r);
        if (k <= kGjMaxK && m <= kGjMaxM) {   // the whole interior elimination in one launch
            const int mb = (m + 15) / 16;
            cx<float>* Fc = (cx<float>*)F;
            if (mb <= 4) k_nd_gj<4><<<L.nb, 256, 0, s>>>(Fc, m, k, info + L.info_first);
            else if (mb == 5) k_nd_gj<5><<<L.nb, 256, 0, s>>>(Fc, m, k, info + L.info_first);
            else if (mb == 6) k_nd_gj<6><<<L.nb, 256, 0, s>>>(Fc, m, k, info + L.info_first);
            else k_nd_gj<7><<<L.nb, 256, 0, s>>>(Fc, m, k, info + L.info_first);
            CV_LAUNCH_CHECK();
            continue;
        }
        if (k <= kGjMaxK && m <= kGjPanelMaxM) {   // the left block by the sweep, X and S by two GEMMs
            const int mr = (m + 63) / 64;
            cx<float>* Fc = (cx<float>*)F;
            if (mr <= 3) k_nd_gj_panel<3><<<L.nb, 1024, 0, s>>>(Fc, m, k, info + L.info_first);
            else if (mr == 4) k_nd_gj_panel<4><<<L.nb, 1024, 0, s>>>(Fc, m, k, info + L.info_first);
            else if (mr == 5) k_nd_gj_panel<5><<<L.nb, 1024, 0, s>>>(Fc, m, k, info + L.info_first);
            else if (mr == 6) k_nd_gj_panel<6><<<L.nb, 1024, 0, s>>>(Fc, m, k, info + L.info_first);
            else k_nd_gj_panel<7><<<L.nb, 1024, 0, s>>>(Fc, m, k, info + L.info_first);
            CV_LAUNCH_CHECK();
            if (b > 0) {   // X = Q F_IB (into R), S = F_BB - G F_IB, then X over F_IB
                const rocblas_stride sR = (rocblas_stride)k * b;
                CV_RB(rocblas_cgemm_strided_batched(h, rocblas_operation_none, rocblas_operation_none, k, b, k, &one,
                                                    F, m, sF, F + (int64_t)k * m, m, sF, &zero_c, R, k, sR, L.nb));
                CV_RB(rocblas_cgemm_strided_batched(h, rocblas_operation_none, rocblas_operation_none, b, b, k,
                                                    &minus_one, F + k, m, sF, F + (int64_t)k * m, m, sF, &one,
                                                    F + (int64_t)k * m + k, m, sF, L.nb));
                k_nd_block_copy<<<nblk((int64_t)L.nb * k * b), 256, 0, s>>>((const cx<float>*)R, sR, k,
                                                                            (cx<float>*)F + (int64_t)k * m, sF, m,
                                                                            L.nb, k, b);
                CV_LAUNCH_CHECK();
            }
            continue;
        }
        bool gj = false;   // [Q X] left by the blocked Gauss-Jordan in Rq (node stride sRq, leading dimension k)
        rc_t* Rq = R;
        rocblas_stride sRq = (rocblas_stride)k * m;
        if (k > kLuMin && k <= kLuMaxRows) {
            gj = true;
            cx<float>* Fc = (cx<float>*)F;
            cx<float>* Rc = (cx<float>*)R;
            cx<float>* B1 = (cx<float>*)(fb + P->off_r2);
            k_nd_rhs<<<nblk((int64_t)L.nb * k * m), 256, 0, s>>>(Fc, Rc, L.nb, k, m);
            CV_LAUNCH_CHECK();
            const int ncol = k + m;
            int sb = 0;   // buffer of the current panel
            for (int j0 = 0; j0 < k; j0 += kLuPanel, sb ^= 1) {
                const int nbw = std::min(kLuPanel, k - j0);
                const int nr = k - j0;   // panel rows: 512 threads x 1 or 2 rows, 1024 x 2 above 1024 rows
                cx<float>* Fp = sb ? B1 : Fc;
                const int ld = sb ? k : m;
                const int64_t ns = sb ? (int64_t)k * ncol : (int64_t)m * m;
                if (nr <= 512)
                    k_lu_panel<512, 1><<<L.nb, 512, 0, s>>>(Fp, ld, ns, k, j0, nbw, piv, info + L.info_first);
                else if (nr <= 1024)
                    k_lu_panel<512, 2><<<L.nb, 512, 0, s>>>(Fp, ld, ns, k, j0, nbw, piv, info + L.info_first);
                else
                    k_lu_panel<1024, 2><<<L.nb, 1024, 0, s>>>(Fp, ld, ns, k, j0, nbw, piv, info + L.info_first);
                CV_LAUNCH_CHECK();
                const int cols = ncol - j0 - nbw;   // >= m > 0
                k_gj_step<<<dim3((unsigned)((k + 63) / 64), (unsigned)((cols + 63) / 64), (unsigned)L.nb), 256, 0, s>>>(
                    Fc, Rc, B1, m, k, j0, nbw, sb, piv);
                CV_LAUNCH_CHECK();
            }
            // sb is now the buffer the last step wrote
            k_gj_final<<<dim3((unsigned)((m + 255) / 256), (unsigned)((k + kLuPanel - 1) / kLuPanel), (unsigned)L.nb),
                         256, 0, s>>>(Fc, Rc, B1, m, k, sb);
            CV_LAUNCH_CHECK();
            if (sb) {
                Rq = (rc_t*)(B1 + (int64_t)k * k);
                sRq = (rocblas_stride)k * ncol;
            }
        } else {
            CV_RB(rocsolver_cgetrf_strided_batched(h, k, k, F, m, sF, piv, k, info + L.info_first, L.nb));
        }
        if (k > kNdSolveMin) {
            // [Q X] = F_II^-1 [I F_IB] (from the blocked Gauss-Jordan, or by
            // the LU triangular solves above kLuMaxRows; both faster than the
            // in-place inverse plus a product for these k), into the top k
            // rows of F; then S = F_BB - F_BI X and G = F_BI Q
            const rocblas_stride sR = (rocblas_stride)k * m;
            if (!gj) {
                k_nd_rhs<<<nblk((int64_t)L.nb * k * m), 256, 0, s>>>((const cx<float>*)F, (cx<float>*)R, L.nb, k,
                                                                     m);
                CV_LAUNCH_CHECK();
                CV_RB(rocsolver_cgetrs_strided_batched(h, rocblas_operation_none, k, m, F, m, sF, piv, k, R, k, sR,
                                                       L.nb));
            }
            const rc_t* Rx = gj ? Rq : R;
            const rocblas_stride sRx = gj ? sRq : sR;
            if (b > 0)
                CV_RB(rocblas_cgemm_strided_batched(h, rocblas_operation_none, rocblas_operation_none, b, b, k,
                                                    &minus_one, F + k, m, sF, Rx + (int64_t)k * k, k, sRx, &one,
                                                    F + (int64_t)k * m + k, m, sF, L.nb));
            k_nd_block_copy<<<nblk((int64_t)L.nb * k * m), 256, 0, s>>>((const cx<float>*)Rx, sRx, k, (cx<float>*)F,
                                                                        sF, m, L.nb, k, m);
            CV_LAUNCH_CHECK();
        } else {
            // Q = F_II^-1 in place: LU with partial pivoting, then the inverse
            // from the factors; X = Q F_IB (into R), S = F_BB - F_BI X, then X
            // over F_IB
            CV_RB(rocsolver_cgetri_strided_batched(h, k, F, m, sF, piv, k, info + L.info_first, L.nb));
            if (b > 0) {
                const rocblas_stride sR = (rocblas_stride)k * b;
                CV_RB(rocblas_cgemm_strided_batched(h, rocblas_operation_none, rocblas_operation_none, k, b, k, &one,
                                                    F, m, sF, F + (int64_t)k * m, m, sF, &zero_c, R, k, sR, L.nb));
                CV_RB(rocblas_cgemm_strided_batched(h, rocblas_operation_none, rocblas_operation_none, b, b, k,
                                                    &minus_one, F + k, m, sF, R, k, sR, &one,
                                                    F + (int64_t)k * m + k, m, sF, L.nb));
                k_nd_block_copy<<<nblk((int64_t)L.nb * k * b), 256, 0, s>>>((const cx<float>*)R, sR, k,
                                                                            (cx<float>*)F + (int64_t)k * m, sF, m,
                                                                            L.nb, k, b);
                CV_LAUNCH_CHECK();
            }
        }
        if (b > 0) {   // G = F_BI Q (into R), then G over F_BI
            const rocblas_stride sG = (rocblas_stride)b * k;
            CV_RB(rocblas_cgemm_strided_batched(h, rocblas_operation_none, rocblas_operation_none, b, k, k, &one,
                                                F + k, m, sF, F, m, sF, &zero_c, R, b, sG, L.nb));
            k_nd_block_copy<<<nblk((int64_t)L.nb * b * k), 256, 0, s>>>((const cx<float>*)R, sG, b,
                                                                        (cx<float>*)F + k, sF, m, L.nb, b, k);
            CV_LAUNCH_CHECK();
        }
    }
    std::vector<rocblas_int> hinfo(P->n_nodes);
    CV_HIP(hipMemcpyAsync(hinfo.data(), info, 4 * (size_t)P->n_nodes, hipMemcpyDeviceToHost, s));
    CV_HIP(hipStreamSynchronize(s));
    for (int64_t q = 0; q < P->n_nodes; ++q)
        if (hinfo[q] != 0) {
            set_error("nested dissection: zero pivot %d in the frontal matrix of node %lld", (int)hinfo[q],
                      (long long)q);
            return CV_ERR_SINGULAR;
        }
    return CV_OK;
}

extern "C" int cv_nd_apply(const cv_nd_plan* P, const void* plan_dev, const void* factor, const void* v, void* z,
                           int transpose, void* stream) {
    CV_CHECK(P && plan_dev && factor && v && z, "null pointer argument");
    return nd_apply(P, plan_dev, factor, (const cx<float>*)v, (cx<float>*)z, transpose, as_stream(stream));
}
