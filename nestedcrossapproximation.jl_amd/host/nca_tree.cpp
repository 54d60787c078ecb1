// Host-side cluster trees, admissibility partition and geometric far-pivot
// selection (Phase 1 of the batched bottom-up compressor).
//
// These are the inputs of the GPU hot path. In the reference they come from
// third-party Julia packages that are not vendored (H2Trees.jl for the trees,
// AdaptiveCrossApproximation.jl for `interactions`/`isnear` and
// `TreeMimicryPivoting`); the functions here are deterministic stand-ins with
// the same contracts:
//   * a node's values are a contiguous range of a point permutation, children
//     partition their parent, node ids are assigned level by level;
//   * far/near CSR by test node (reference FarData, farinteractions.jl:16-19);
//   * far field of a node = own far nodes followed by those of every ancestor
//     (farinteractions.jl:38-60);
//   * the number of far pivots follows _selectfarpivots/_clustersize
//     (batchedbottomup.jl:54-72, pivots.jl:15-25).
// All indices in this file are 0-based.

#include <algorithm>
#include <stdexcept>
#include <memory>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <vector>

#include <omp.h>

namespace {

struct Tree {
    int64_t n = 0;
    std::vector<int64_t> parent, level, lo, hi, childptr, children, perm;
    std::vector<double> center, radius;
};

struct Partition {
    std::vector<int64_t> nearptr, nearnodes, farptr, farnodes;
};

struct Csr {
    std::vector<int64_t> ptr, idx;
};

inline double dist2(const double* a, const double* b) {
    double dx = a[0] - b[0], dy = a[1] - b[1], dz = a[2] - b[2];
    return dx * dx + dy * dy + dz * dz;
}

void finish_children(Tree& t) {
    const int64_t nn = (int64_t)t.parent.size();
    std::vector<int64_t> count(nn + 1, 0);
    for (int64_t i = 1; i < nn; ++i) count[t.parent[i] + 1]++;
    t.childptr.assign(nn + 1, 0);
    for (int64_t i = 0; i < nn; ++i) t.childptr[i + 1] = t.childptr[i] + count[i + 1];
    t.children.assign(t.childptr[nn], 0);
    std::vector<int64_t> cur(t.childptr.begin(), t.childptr.end() - 1);
    for (int64_t i = 1; i < nn; ++i) t.children[cur[t.parent[i]]++] = i;
}

// Regular 2^3 tree: the bounding cube is halved while a box holds more than
// `minvalues` points and its half size exceeds `minhalfsize`; empty octants are
// dropped. Radius of a box = half size * sqrt(3) (the ball around the cube).
Tree build_twon(const double* pts, int64_t n, int64_t minvalues, double minhalfsize) {
    Tree t;
    t.n = n;
    t.perm.resize(n);
    for (int64_t i = 0; i < n; ++i) t.perm[i] = i;
    double mn[3] = {std::numeric_limits<double>::max(), std::numeric_limits<double>::max(),
                    std::numeric_limits<double>::max()};
    double mx[3] = {-mn[0], -mn[1], -mn[2]};
    for (int64_t i = 0; i < n; ++i)
        for (int d = 0; d < 3; ++d) {
            mn[d] = std::min(mn[d], pts[3 * i + d]);
            mx[d] = std::max(mx[d], pts[3 * i + d]);
        }
    if (n == 0) {
        for (int d = 0; d < 3; ++d) mn[d] = mx[d] = 0.0;
    }
    double h = 0.0;
    for (int d = 0; d < 3; ++d) h = std::max(h, 0.5 * (mx[d] - mn[d]));
    if (h <= 0.0) h = 1e-300;
    std::vector<double> half;
    t.parent.push_back(-1);
    t.level.push_back(0);
    t.lo.push_back(0);
    t.hi.push_back(n);
    for (int d = 0; d < 3; ++d) t.center.push_back(0.5 * (mn[d] + mx[d]));
    half.push_back(h);
    std::vector<int64_t> tmp(n);
    for (int64_t node = 0; node < (int64_t)t.parent.size(); ++node) {
        const int64_t lo = t.lo[node], hi = t.hi[node];
        const double hs = half[node];
        if (hi - lo <= minvalues || hs <= minhalfsize) continue;
        const double c[3] = {t.center[3 * node], t.center[3 * node + 1], t.center[3 * node + 2]};
        int64_t cnt[8] = {0};
        auto octant = [&](int64_t p) {
            const double* x = pts + 3 * p;
            return (x[0] >= c[0] ? 1 : 0) | (x[1] >= c[1] ? 2 : 0) | (x[2] >= c[2] ? 4 : 0);
        };
        for (int64_t k = lo; k < hi; ++k) cnt[octant(t.perm[k])]++;
        int nonempty = 0;
        for (int o = 0; o < 8; ++o) nonempty += cnt[o] > 0;
        if (nonempty <= 1 && hs <= 1e-12) continue;
        int64_t start[8];
        int64_t s = lo;
        for (int o = 0; o < 8; ++o) {
            start[o] = s;
            s += cnt[o];
        }
        int64_t pos[8];
        std::memcpy(pos, start, sizeof(pos));
        for (int64_t k = lo; k < hi; ++k) tmp[pos[octant(t.perm[k])]++] = t.perm[k];
        std::memcpy(&t.perm[lo], &tmp[lo], sizeof(int64_t) * (hi - lo));
        for (int o = 0; o < 8; ++o) {
            if (cnt[o] == 0) continue;
            t.parent.push_back(node);
            t.level.push_back(t.level[node] + 1);
            t.lo.push_back(start[o]);
            t.hi.push_back(start[o] + cnt[o]);
            const double q = 0.5 * hs;
            t.center.push_back(c[0] + ((o & 1) ? q : -q));
            t.center.push_back(c[1] + ((o & 2) ? q : -q));
            t.center.push_back(c[2] + ((o & 4) ? q : -q));
            half.push_back(q);
        }
    }
    const int64_t nn = (int64_t)t.parent.size();
    t.radius.resize(nn);
    for (int64_t i = 0; i < nn; ++i) t.radius[i] = half[i] * std::sqrt(3.0);
    finish_children(t);
    return t;
}

// Deterministic k-means splitting: seeds by farthest-point traversal from the
// point farthest from the centroid, Lloyd iterations with ties to the lower
// cluster. Center = centroid, radius = largest distance to it (bounding sphere).
// Sums over a node's points are fixed-order so that a parallel builder can
// produce the same bits: slot s accumulates the terms q = s, s + 256, ... in
// order, then the slots are added by a halving tree (slot i += slot i + w for
// w = 128, 64, ..., 1).
constexpr int kSumSlots = 256;

double slot_tree(double* part) {
    for (int w = kSumSlots / 2; w > 0; w >>= 1)
        for (int i = 0; i < w; ++i) part[i] += part[i + w];
    return part[0];
}

void centroid_radius(const double* pts, const int64_t* ids, int64_t m, double* c, double* r) {
    double s[3];
    for (int d = 0; d < 3; ++d) {
        double part[kSumSlots] = {0};
        for (int64_t k = 0; k < m; ++k) part[k % kSumSlots] += pts[3 * ids[k] + d];
        s[d] = slot_tree(part);
    }
    for (int d = 0; d < 3; ++d) c[d] = m > 0 ? s[d] / (double)m : 0.0;
    double r2 = 0.0;
    for (int64_t k = 0; k < m; ++k) r2 = std::max(r2, dist2(pts + 3 * ids[k], c));
    *r = std::sqrt(r2);
}

Tree build_kmeans(const double* pts, int64_t n, int64_t nclusters, int64_t minvalues) {
    Tree t;
    t.n = n;
    t.perm.resize(n);
    for (int64_t i = 0; i < n; ++i) t.perm[i] = i;
    t.parent.push_back(-1);
    t.level.push_back(0);
    t.lo.push_back(0);
    t.hi.push_back(n);
    double c[3], r;
    centroid_radius(pts, t.perm.data(), n, c, &r);
    for (int d = 0; d < 3; ++d) t.center.push_back(c[d]);
    t.radius.push_back(r);
    const int k = (int)std::max<int64_t>(2, nclusters);
    std::vector<int64_t> tmp(n);
    std::vector<int> assign;
    for (int64_t node = 0; node < (int64_t)t.parent.size(); ++node) {
        const int64_t lo = t.lo[node], hi = t.hi[node], m = hi - lo;
        if (m <= minvalues || m == 0) continue;
        const int64_t* ids = &t.perm[lo];
        std::vector<double> cent(3 * k);
        const double* c0 = &t.center[3 * node];
        int64_t first = 0;
        double best = -1.0;
        for (int64_t q = 0; q < m; ++q) {
            double d2 = dist2(pts + 3 * ids[q], c0);
            if (d2 > best) { best = d2; first = q; }
        }
        for (int d = 0; d < 3; ++d) cent[d] = pts[3 * ids[first] + d];
        std::vector<double> mind(m, std::numeric_limits<double>::max());
        int kk = 1;
        for (; kk < k; ++kk) {
            int64_t arg = -1;
            double bd = 0.0;
            for (int64_t q = 0; q < m; ++q) {
                mind[q] = std::min(mind[q], dist2(pts + 3 * ids[q], &cent[3 * (kk - 1)]));
                if (mind[q] > bd) { bd = mind[q]; arg = q; }
            }
            if (arg < 0) break;
            for (int d = 0; d < 3; ++d) cent[3 * kk + d] = pts[3 * ids[arg] + d];
        }
        const int kc = kk;
        if (kc < 2) continue;  // all points coincide
        assign.assign(m, -1);
        for (int it = 0; it < 100; ++it) {
            bool changed = false;
            for (int64_t q = 0; q < m; ++q) {
                int a = 0;
                double bd = dist2(pts + 3 * ids[q], &cent[0]);
                for (int j = 1; j < kc; ++j) {
                    double d2 = dist2(pts + 3 * ids[q], &cent[3 * j]);
                    if (d2 < bd) { bd = d2; a = j; }
                }
                if (a != assign[q]) { assign[q] = a; changed = true; }
            }
            if (!changed) break;
            std::vector<double> s(3 * kc, 0.0);
            std::vector<int64_t> cnt(kc, 0);
            std::vector<double> part((size_t)kSumSlots * 3 * kc, 0.0);
            for (int64_t q = 0; q < m; ++q) {
                cnt[assign[q]]++;
                for (int d = 0; d < 3; ++d)
                    part[(size_t)(3 * assign[q] + d) * kSumSlots + q % kSumSlots] += pts[3 * ids[q] + d];
            }
            for (int j = 0; j < 3 * kc; ++j) s[j] = slot_tree(&part[(size_t)j * kSumSlots]);
            for (int j = 0; j < kc; ++j)
                if (cnt[j] > 0)
                    for (int d = 0; d < 3; ++d) cent[3 * j + d] = s[3 * j + d] / (double)cnt[j];
        }
        std::vector<int64_t> cnt(kc, 0);
        for (int64_t q = 0; q < m; ++q) cnt[assign[q]]++;
        int nonempty = 0;
        for (int j = 0; j < kc; ++j) nonempty += cnt[j] > 0;
        if (nonempty < 2) continue;
        std::vector<int64_t> start(kc), pos(kc);
        int64_t s = lo;
        for (int j = 0; j < kc; ++j) { start[j] = s; pos[j] = s; s += cnt[j]; }
        for (int64_t q = 0; q < m; ++q) tmp[pos[assign[q]]++] = ids[q];
        std::memcpy(&t.perm[lo], &tmp[lo], sizeof(int64_t) * m);
        for (int j = 0; j < kc; ++j) {
            if (cnt[j] == 0) continue;
            t.parent.push_back(node);
            t.level.push_back(t.level[node] + 1);
            t.lo.push_back(start[j]);
            t.hi.push_back(start[j] + cnt[j]);
            centroid_radius(pts, &t.perm[start[j]], cnt[j], c, &r);
            for (int d = 0; d < 3; ++d) t.center.push_back(c[d]);
            t.radius.push_back(r);
        }
    }
    finish_children(t);
    return t;
}

struct TreeView {
    int64_t nn;
    const int64_t* childptr;
    const int64_t* children;
    const double* center;
    const double* radius;
    bool leaf(int64_t i) const { return childptr[i + 1] == childptr[i]; }
};

// Admissibility of a pair of bounding balls (radius = half size * sqrt(3)
// for boxes), dist = |c_t - c_s| - (r_t + r_s).
//   standard (wideband = 0): near iff max(diam_t, diam_s) > eta * dist;
//   wideband (IsNearWidebandFunctor, nearinteractions.jl:83-103): with
//   sep = max(dist, 0), a pair whose smaller diameter d has k d <= gamma is
//   low frequency, near iff 2 max(r) > eta_lf sep; otherwise near iff
//   4 k max(r)^2 > eta_hf sep.
struct Criterion {
    int wideband = 0;
    double eta = 1.0, k = 0.0, etalf = 1.0, etahf = 5.0, gamma = 1.0;
};

inline bool isnear_pair(const TreeView& a, const TreeView& b, int64_t t, int64_t s, const Criterion& c) {
    const double ra = a.radius[t], rb = b.radius[s];
    const double d = std::sqrt(dist2(a.center + 3 * t, b.center + 3 * s)) - (ra + rb);
    if (!c.wideband) return 2.0 * std::max(ra, rb) > c.eta * d;
    const double sep = std::max(d, 0.0);
    if (c.k * std::min(2.0 * ra, 2.0 * rb) <= c.gamma) return 2.0 * std::max(ra, rb) > c.etalf * sep;
    return 4.0 * c.k * std::max(ra * ra, rb * rb) > c.etahf * sep;
}

void traverse(const TreeView& a, const TreeView& b, int64_t t, int64_t s, const Criterion& eta,
              std::vector<std::pair<int64_t, int64_t>>& nearp,
              std::vector<std::pair<int64_t, int64_t>>& farp) {
    if (!isnear_pair(a, b, t, s, eta)) {
        farp.emplace_back(t, s);
        return;
    }
    const bool lt = a.leaf(t), ls = b.leaf(s);
    if (lt && ls) {
        nearp.emplace_back(t, s);
    } else if (lt) {
        for (int64_t k = b.childptr[s]; k < b.childptr[s + 1]; ++k)
            traverse(a, b, t, b.children[k], eta, nearp, farp);
    } else if (ls) {
        for (int64_t k = a.childptr[t]; k < a.childptr[t + 1]; ++k)
            traverse(a, b, a.children[k], s, eta, nearp, farp);
    } else {
        for (int64_t i = a.childptr[t]; i < a.childptr[t + 1]; ++i)
            for (int64_t k = b.childptr[s]; k < b.childptr[s + 1]; ++k)
                traverse(a, b, a.children[i], b.children[k], eta, nearp, farp);
    }
}

void to_csr(int64_t nrows, std::vector<std::pair<int64_t, int64_t>>& pairs, std::vector<int64_t>& ptr,
            std::vector<int64_t>& idx) {
    ptr.assign(nrows + 1, 0);
    for (auto& p : pairs) ptr[p.first + 1]++;
    for (int64_t i = 0; i < nrows; ++i) ptr[i + 1] += ptr[i];
    idx.assign(pairs.size(), 0);
    std::vector<int64_t> cur(ptr.begin(), ptr.end() - 1);
    for (auto& p : pairs) idx[cur[p.first]++] = p.second;
#pragma omp parallel for schedule(dynamic, 256)
    for (int64_t i = 0; i < nrows; ++i) std::sort(idx.begin() + ptr[i], idx.begin() + ptr[i + 1]);
}

Partition interactions(const TreeView& a, const TreeView& b, const Criterion& eta) {
    // Expand the pair traversal breadth-first to a frontier, then finish every
    // frontier pair depth-first in parallel; the CSR is sorted afterwards, so
    // the result does not depend on the schedule.
    std::vector<std::pair<int64_t, int64_t>> nearp, farp, frontier{{0, 0}};
    while (!frontier.empty() && frontier.size() < 8192) {
        std::vector<std::pair<int64_t, int64_t>> next;
        bool expanded = false;
        for (auto& pr : frontier) {
            int64_t t = pr.first, s = pr.second;
            if (!isnear_pair(a, b, t, s, eta)) { farp.push_back(pr); continue; }
            const bool lt = a.leaf(t), ls = b.leaf(s);
            if (lt && ls) { nearp.push_back(pr); continue; }
            expanded = true;
            if (lt) {
                for (int64_t k = b.childptr[s]; k < b.childptr[s + 1]; ++k) next.emplace_back(t, b.children[k]);
            } else if (ls) {
                for (int64_t k = a.childptr[t]; k < a.childptr[t + 1]; ++k) next.emplace_back(a.children[k], s);
            } else {
                for (int64_t i = a.childptr[t]; i < a.childptr[t + 1]; ++i)
                    for (int64_t k = b.childptr[s]; k < b.childptr[s + 1]; ++k)
                        next.emplace_back(a.children[i], b.children[k]);
            }
        }
        frontier.swap(next);
        if (!expanded) break;
    }
    const int64_t nf = (int64_t)frontier.size();
    int nth = omp_get_max_threads();
    std::vector<std::vector<std::pair<int64_t, int64_t>>> ln(nth), lf(nth);
#pragma omp parallel for schedule(dynamic, 16)
    for (int64_t i = 0; i < nf; ++i) {
        int tid = omp_get_thread_num();
        traverse(a, b, frontier[i].first, frontier[i].second, eta, ln[tid], lf[tid]);
    }
    for (int i = 0; i < nth; ++i) {
        nearp.insert(nearp.end(), ln[i].begin(), ln[i].end());
        farp.insert(farp.end(), lf[i].begin(), lf[i].end());
    }
    Partition p;
    to_csr(a.nn, nearp, p.nearptr, p.nearnodes);
    to_csr(a.nn, farp, p.farptr, p.farnodes);
    return p;
}

// Phase 1 stand-in for geometric (tree-mimicry) pivoting. For node t:
//   F(t) = own far nodes, then the far nodes of every ancestor;
//   n = min(maxrank, ceil(oversample |t|)) capped by the points F(t) holds
//   (oversample = 1 is the reference's count rule, _clustersize);
//   candidates = an even subsample of every far node's points;
//   first pivot = candidate nearest to the center of t, every next pivot
//   maximizes d_min^2 / d_c^(2 weight), d_min = distance to the chosen
//   pivots, d_c = distance to the center of t.
// Ties go to the first candidate in far-field order. The defaults of the host
// mirror (weight 1.5, oversample 1) were measured against dense direct
// summation on 400 rows (oracle operator): 65,536-point sphere, k = 1, tol
// 1e-6: 1.8 tol (weight 1: 21 tol, weight 1.2: 10 tol; weight 1 with
// oversample 2: 2.6 tol); 1M-point sphere, k = 2, tol 1e-4: 1.9 tol with
// 7.8e8 coupling entries (weight 1.2: 2.1 tol, 6.9e8; weight 1 with
// oversample 2: 2.5 tol, 6.2e8; weight 1: 6.9 tol).
// farfield(t, F) fills F with the far nodes of item t; nvals(t) = |t|;
// center(t) = its center (3 doubles).
template <class FarField, class NVals, class Center>
Csr select_farpivots(int64_t nitems, FarField farfield, NVals nvals, Center center, const int64_t* other_lo,
                     const int64_t* other_hi, const int64_t* other_perm, const double* other_pts, int64_t maxrank,
                     double weight, double oversample) {
    std::vector<std::vector<int64_t>> result(nitems);
#pragma omp parallel
    {
        std::vector<int64_t> F, pool;
        std::vector<double> d2c, mind;
#pragma omp for schedule(dynamic, 64)
        for (int64_t t = 0; t < nitems; ++t) {
            F.clear();
            farfield(t, F);
            if (F.empty()) continue;
            const int64_t limit = std::min<int64_t>(maxrank, (int64_t)std::ceil(oversample * (double)nvals(t)));
            int64_t n = 0, total = 0;
            for (int64_t f : F) {
                const int64_t sz = other_hi[f] - other_lo[f];
                total += sz;
                if (n < limit) n += std::min(sz, limit - n);
            }
            if (n == 0) continue;
            int64_t per = std::max<int64_t>(1, (std::max<int64_t>(8 * n, 512) + (int64_t)F.size() - 1) /
                                                   (int64_t)F.size());
            for (;;) {
                pool.clear();
                for (int64_t f : F) {
                    const int64_t sz = other_hi[f] - other_lo[f];
                    const int64_t k = std::min(sz, per);
                    for (int64_t j = 0; j < k; ++j) pool.push_back(other_perm[other_lo[f] + (j * sz) / k]);
                }
                if ((int64_t)pool.size() >= n || (int64_t)pool.size() >= total) break;
                per *= 2;
            }
            const int64_t np = (int64_t)pool.size();
            const double* ct = center(t);
            d2c.resize(np);
            mind.assign(np, std::numeric_limits<double>::infinity());
            int64_t first = 0;
            for (int64_t q = 0; q < np; ++q) {
                d2c[q] = std::max(dist2(other_pts + 3 * pool[q], ct), 1e-300);
                if (d2c[q] < d2c[first]) first = q;
            }
            if (weight == 1.5)  // exactly as the device computes it
                for (int64_t q = 0; q < np; ++q) d2c[q] = d2c[q] * std::sqrt(d2c[q]);
            else if (weight != 1.0)
                for (int64_t q = 0; q < np; ++q) d2c[q] = std::pow(d2c[q], weight);
            std::vector<int64_t>& out = result[t];
            out.reserve(n);
            int64_t cur = first;
            for (int64_t sel = 0; sel < n; ++sel) {
                out.push_back(pool[cur]);
                mind[cur] = -1.0;  // chosen
                const double* pc = other_pts + 3 * pool[cur];
                int64_t arg = -1;
                double best = -1.0;
                for (int64_t q = 0; q < np; ++q) {
                    if (mind[q] < 0.0) continue;
                    double d2 = dist2(other_pts + 3 * pool[q], pc);
                    if (d2 < mind[q]) mind[q] = d2;
                    double score = mind[q] / d2c[q];
                    if (score > best) { best = score; arg = q; }
                }
                if (arg < 0) break;
                cur = arg;
            }
        }
    }
    Csr c;
    c.ptr.assign(nitems + 1, 0);
    for (int64_t t = 0; t < nitems; ++t) c.ptr[t + 1] = c.ptr[t] + (int64_t)result[t].size();
    c.idx.resize(c.ptr[nitems]);
    for (int64_t t = 0; t < nitems; ++t)
        std::copy(result[t].begin(), result[t].end(), c.idx.begin() + c.ptr[t]);
    return c;
}

// Far field of a node: its own far nodes, then those of every ancestor.
Csr farpivots(int64_t nn_self, const int64_t* parent, const int64_t* lo, const int64_t* hi,
              const double* center, const int64_t* other_lo, const int64_t* other_hi,
              const int64_t* other_perm, const double* other_pts, const int64_t* farptr,
              const int64_t* farnodes, int64_t maxrank, double weight, double oversample) {
    return select_farpivots(
        nn_self,
        [&](int64_t t, std::vector<int64_t>& F) {
            for (int64_t node = t; node >= 0; node = parent[node])
                for (int64_t k = farptr[node]; k < farptr[node + 1]; ++k) F.push_back(farnodes[k]);
        },
        [&](int64_t t) { return hi[t] - lo[t]; }, [&](int64_t t) { return center + 3 * t; }, other_lo, other_hi,
        other_perm, other_pts, maxrank, weight, oversample);
}

// ---------------------------------------------------------------------------
// Directional far data (the host-side subdivision of the far interactions into
// directions): twondirections.jl:259-308 for TwoN trees with a direction tree,
// boundingballdirections.jl:44-152 for bounding-ball trees with per-cluster
// Fibonacci direction sets. Direction labels: 0 = low frequency, positive =
// a direction-tree node (1-based) or an index into the cluster's set (1-based).
struct DirData {
    std::vector<int64_t> dirptr, dirs, diridxtofarptr, chdirptr, chdirmap;
};

inline double dot3(const double* a, const double* b) { return (a[0] * b[0] + a[1] * b[1]) + a[2] * b[2]; }

// directions(Val(3), diam, k, gamma): Fibonacci points, n from the aperture
std::vector<double> fib_directions(double diam, double k, double gamma) {
    const double kd = std::max(k * diam, std::numeric_limits<double>::epsilon());
    const double a = std::asin(std::min(1.0, gamma / kd));
    const int64_t n = std::max<int64_t>(
        6, (int64_t)std::ceil(6.0 * std::pow(4.0, std::log(std::acos(1.0 / std::sqrt(3.0)) / a) / std::log(2.0))));
    std::vector<double> v(3 * n);
    const double ga = M_PI * (3.0 - std::sqrt(5.0));
    for (int64_t i = 0; i < n; ++i) {
        const double z = 1.0 - 2.0 * ((double)i + 0.5) / (double)n;
        const double r = std::sqrt(std::max(0.0, 1.0 - z * z));
        v[3 * i] = r * std::cos((double)i * ga);
        v[3 * i + 1] = r * std::sin((double)i * ga);
        v[3 * i + 2] = z;
    }
    return v;
}

int64_t nearest(const std::vector<double>& ev, const double* vec) {
    int64_t best = 0;
    double bd = dot3(&ev[0], vec);
    for (int64_t i = 1; i < (int64_t)ev.size() / 3; ++i) {
        const double d = dot3(&ev[3 * i], vec);
        if (d > bd) { bd = d; best = i; }
    }
    return best;
}

// Direction tree of 3D TwoN trees: 6 roots (+x -x +y -y +z -z), 4 nested patches per level.
struct DirTree {
    int maxlevel = 1, dirroot = 0;
    std::vector<double> dir;  // 3 per node, index 0 unused
    std::vector<int64_t> parent, firstchild;
    DirTree(int maxlevel_, int dirroot_) : maxlevel(maxlevel_), dirroot(dirroot_) {
        int64_t total = 6 * (((int64_t)1 << (2 * maxlevel)) - 1) / 3;
        dir.assign(3 * (total + 1), 0.0);
        parent.assign(total + 1, 0);
        firstchild.assign(total + 1, 0);
        std::vector<int> axis(total + 1, 0);
        int64_t nxt = 1;
        for (int ax = 0; ax < 3; ++ax)
            for (double sg : {1.0, -1.0}) {
                dir[3 * nxt + ax] = sg;
                axis[nxt] = ax;
                ++nxt;
            }
        int64_t start = 1, count = 6;
        double scale = 0.5;
        for (int lev = 0; lev + 1 < maxlevel; ++lev) {
            for (int64_t node = start; node < start + count; ++node) {
                firstchild[node] = nxt;
                const int ax = axis[node];
                int others[2], k = 0;
                for (int d = 0; d < 3; ++d)
                    if (d != ax) others[k++] = d;
                for (int mask = 0; mask < 4; ++mask) {
                    for (int d = 0; d < 3; ++d) dir[3 * nxt + d] = dir[3 * node + d];
                    for (int b = 0; b < 2; ++b) dir[3 * nxt + others[b]] += scale * (((mask >> b) & 1) ? 1.0 : -1.0);
                    axis[nxt] = ax;
                    parent[nxt] = node;
                    ++nxt;
                }
            }
            start += count;
            count *= 4;
            scale /= 2.0;
        }
        for (int64_t i = 1; i <= total; ++i) {
            const double nr = std::sqrt((dir[3 * i] * dir[3 * i] + dir[3 * i + 1] * dir[3 * i + 1]) +
                                        dir[3 * i + 2] * dir[3 * i + 2]);
            for (int d = 0; d < 3; ++d) dir[3 * i + d] /= nr;
        }
    }
    int64_t direction(const double* v, int clusterlevel) const {
        const int depth = std::min(dirroot - clusterlevel + 1, maxlevel);
        int64_t cur = 0;
        for (int lev = 0; lev < depth; ++lev) {
            const int64_t c0 = lev == 0 ? 1 : firstchild[cur];
            const int nc = lev == 0 ? 6 : 4;
            int64_t best = c0;
            double bd = -std::numeric_limits<double>::infinity();
            for (int64_t c = c0; c < c0 + nc; ++c) {
                const double d = dot3(&dir[3 * c], v);
                if (d > bd) { bd = d; best = c; }
            }
            cur = best;
        }
        return cur;
    }
};

DirData directional_fardata(int twon, int64_t nn, const int64_t* parent, const int64_t* level,
                            const int64_t* childptr, const int64_t* children, const double* center,
                            const double* radius, const int64_t* farptr, int64_t* farnodes, const double* fcenter,
                            double k, double gamma, double etahf) {
    // low-frequency classification: TwoN per level, bounding balls per node (generalized radius)
    auto lf_level = [&](int64_t lev) {
        const double h0 = radius[0] / std::sqrt(3.0);
        return k * (2.0 * std::sqrt(3.0) * h0 / std::ldexp(1.0, (int)lev)) <= gamma;
    };
    std::vector<double> grad(nn, 0.0);
    if (!twon && nn > 0) {
        grad[0] = radius[0];
        for (int64_t p = 0; p < nn; ++p) {
            const int64_t nc = childptr[p + 1] - childptr[p];
            if (!nc) continue;
            double sum = 0.0;
            for (int64_t c = childptr[p]; c < childptr[p + 1]; ++c) sum += radius[children[c]];
            const double g = std::min(grad[p], sum / (double)nc);
            for (int64_t c = childptr[p]; c < childptr[p + 1]; ++c) grad[children[c]] = g;
        }
    }
    auto lf_node = [&](int64_t node) { return twon ? lf_level(level[node]) : k * (2.0 * grad[node]) <= gamma; };
    std::unique_ptr<DirTree> dt;
    if (twon) {
        int64_t lflevel = 0;
        while (!lf_level(lflevel)) ++lflevel;
        const int64_t dirroot = lflevel - 1;
        int64_t first = dirroot;
        for (int64_t t = 0; t < nn; ++t)
            if (farptr[t + 1] > farptr[t] && !lf_level(level[t]) && level[t] < first) first = level[t];
        dt.reset(new DirTree((int)(dirroot - first + 1), (int)dirroot));
    }
    std::vector<std::vector<int64_t>> dirs(nn), dcount(nn), chmap(nn);
    std::vector<std::vector<double>> dvec(nn);
    std::vector<char> hasvec(nn, 0);
    std::vector<int64_t> order(nn);
    for (int64_t i = 0; i < nn; ++i) order[i] = i;
    std::stable_sort(order.begin(), order.end(), [&](int64_t a, int64_t b) { return level[a] < level[b]; });
    // level by level; the nodes of a level only read their parents' results
    std::vector<int64_t> lstart(1, 0);
    for (int64_t i = 1; i < nn; ++i)
        if (level[order[i]] != level[order[i - 1]]) lstart.push_back(i);
    lstart.push_back(nn);
    for (size_t L = 0; L + 1 < lstart.size(); ++L)
#pragma omp parallel for schedule(dynamic, 4)
    for (int64_t oi = lstart[L]; oi < lstart[L + 1]; ++oi) {
        const int64_t node = order[oi];
        const int64_t pnode = parent[node];
        const bool hasfars = farptr[node + 1] > farptr[node];
        const bool hasparentdirs = pnode >= 0 && !dirs[pnode].empty();
        if (lf_node(node)) {
            const bool inherit = hasparentdirs && level[node] > 0 && lf_node(pnode);
            if (hasfars || inherit) dirs[node] = {0};
            if (inherit) chmap[node] = {0};
            continue;
        }
        const bool hasparent = twon ? hasparentdirs : (pnode >= 0 && hasvec[pnode]);
        if (!hasfars && !hasparent) continue;
        // (a direction tree is only needed for TwoN trees; dt is shared read-only)
        const int64_t a = farptr[node], b = farptr[node + 1];
        std::vector<double> total;
        if (!twon) total = fib_directions(2.0 * radius[node], k, gamma);
        std::vector<int64_t> nd(b - a);
        for (int64_t i = a; i < b; ++i) {
            double v[3];
            for (int d = 0; d < 3; ++d) v[d] = center[3 * node + d] - fcenter[3 * farnodes[i] + d];
            nd[i - a] = twon ? dt->direction(v, (int)level[node]) : nearest(total, v) + 1;
        }
        std::vector<int64_t> perm(b - a);
        for (int64_t i = 0; i < b - a; ++i) perm[i] = i;
        std::stable_sort(perm.begin(), perm.end(), [&](int64_t x, int64_t y) { return nd[x] < nd[y]; });
        std::vector<int64_t> fn(b - a), sd(b - a);
        for (int64_t i = 0; i < b - a; ++i) {
            fn[i] = farnodes[a + perm[i]];
            sd[i] = nd[perm[i]];
        }
        std::copy(fn.begin(), fn.end(), farnodes + a);
        std::vector<int64_t> uniq(sd);
        uniq.erase(std::unique(uniq.begin(), uniq.end()), uniq.end());
        const size_t noriginal = uniq.size();
        std::vector<int64_t> pm(pnode >= 0 ? dirs[pnode].size() : 0, 0);
        auto addmap = [&](size_t idx, int64_t c) {
            auto it = std::find(uniq.begin(), uniq.end(), c);
            if (it == uniq.end()) {
                uniq.push_back(c);
                pm[idx] = (int64_t)uniq.size();
            } else {
                pm[idx] = (int64_t)(it - uniq.begin()) + 1;
            }
        };
        if (twon) {
            if (hasparentdirs)
                for (size_t i = 0; i < dirs[pnode].size(); ++i) addmap(i, dt->parent[dirs[pnode][i]]);
        } else if (pnode >= 0 && hasvec[pnode] && !dvec[pnode].empty()) {
            double ctopc[3];
            for (int d = 0; d < 3; ++d) ctopc[d] = center[3 * node + d] - center[3 * pnode + d];
            const double mult = k * (2.0 * radius[pnode]) / etahf + 2.0 * radius[pnode];
            const size_t np = dvec[pnode].size() / 3;
            std::vector<int64_t> cd(np);
            for (size_t i = 0; i < np; ++i) {
                double tgt[3];
                for (int d = 0; d < 3; ++d) tgt[d] = ctopc[d] + mult * dvec[pnode][3 * i + d];
                cd[i] = nearest(total, tgt) + 1;
            }
            for (size_t i = 0; i < np; ++i) addmap(i, cd[i]);
        }
        std::vector<int64_t> cnt(uniq.size(), 0);
        for (size_t u = 0; u < noriginal; ++u)
            cnt[u] = (int64_t)std::count(sd.begin(), sd.end(), uniq[u]);
        if (!twon) {
            dvec[node].resize(3 * uniq.size());
            for (size_t u = 0; u < uniq.size(); ++u)
                for (int d = 0; d < 3; ++d) dvec[node][3 * u + d] = total[3 * (uniq[u] - 1) + d];
            hasvec[node] = 1;
        }
        std::vector<int64_t> m0;
        for (int64_t v : pm) m0.push_back(v - 1);
        chmap[node] = m0;
        dirs[node] = uniq;
        dcount[node] = cnt;
    }
    DirData D;
    D.dirptr.assign(nn + 1, 0);
    D.chdirptr.assign(nn + 1, 0);
    for (int64_t t = 0; t < nn; ++t) {
        if (!dirs[t].empty() && dcount[t].empty()) dcount[t] = {farptr[t + 1] - farptr[t]};
        int64_t nz = 0;
        for (int64_t v : chmap[t]) nz += v >= 0;
        D.dirptr[t + 1] = D.dirptr[t] + (int64_t)dirs[t].size();
        D.chdirptr[t + 1] = D.chdirptr[t] + (dirs[t].size() == 1 && dirs[t][0] == 0 ? (int64_t)chmap[t].size() : nz);
    }
    const int64_t ne = D.dirptr[nn];
    D.dirs.resize(ne);
    D.diridxtofarptr.assign(ne + 1, 0);
    D.chdirmap.resize(D.chdirptr[nn]);
    for (int64_t t = 0; t < nn; ++t) {
        int64_t off = farptr[t];
        for (size_t i = 0; i < dirs[t].size(); ++i) {
            D.dirs[D.dirptr[t] + i] = dirs[t][i];
            D.diridxtofarptr[D.dirptr[t] + i] = off;
            off += dcount[t][i];
        }
        if (!dirs[t].empty() && off != farptr[t + 1])
            throw std::runtime_error("directional far data: direction counts do not sum to the far nodes");
        std::copy(chmap[t].begin(), chmap[t].begin() + (D.chdirptr[t + 1] - D.chdirptr[t]),
                  D.chdirmap.begin() + D.chdirptr[t]);
    }
    D.diridxtofarptr[ne] = farptr[nn];
    return D;
}

}  // namespace

extern "C" {

typedef struct ncat_dirdata ncat_dirdata;

// directional_fardata over one tree; farnodes (0-based CSR of far nodes of
// the other tree, fcenter its centers) is reordered by direction in place.
// Returns null on inconsistent input.
ncat_dirdata* ncat_directional_fardata(int32_t twon, int64_t nn, const int64_t* parent, const int64_t* level,
                                       const int64_t* childptr, const int64_t* children, const double* center,
                                       const double* radius, const int64_t* farptr, int64_t* farnodes,
                                       const double* fcenter, double k, double gamma, double etahf) {
    try {
        return reinterpret_cast<ncat_dirdata*>(new DirData(directional_fardata(
            twon, nn, parent, level, childptr, children, center, radius, farptr, farnodes, fcenter, k, gamma, etahf)));
    } catch (...) {
        return nullptr;
    }
}

void ncat_dirdata_sizes(const ncat_dirdata* h, int64_t* ne, int64_t* nmap) {
    const DirData& d = *reinterpret_cast<const DirData*>(h);
    *ne = (int64_t)d.dirs.size();
    *nmap = (int64_t)d.chdirmap.size();
}

void ncat_dirdata_export(const ncat_dirdata* h, int64_t* dirptr, int64_t* dirs, int64_t* diridxtofarptr,
                         int64_t* chdirptr, int64_t* chdirmap) {
    const DirData& d = *reinterpret_cast<const DirData*>(h);
    std::copy(d.dirptr.begin(), d.dirptr.end(), dirptr);
    std::copy(d.dirs.begin(), d.dirs.end(), dirs);
    std::copy(d.diridxtofarptr.begin(), d.diridxtofarptr.end(), diridxtofarptr);
    std::copy(d.chdirptr.begin(), d.chdirptr.end(), chdirptr);
    std::copy(d.chdirmap.begin(), d.chdirmap.end(), chdirmap);
}

void ncat_dirdata_free(ncat_dirdata* h) { delete reinterpret_cast<DirData*>(h); }


typedef struct ncat_tree ncat_tree;
typedef struct ncat_partition ncat_partition;
typedef struct ncat_csr ncat_csr;

ncat_tree* ncat_tree_twon(const double* pts, int64_t n, int64_t minvalues, double minhalfsize) {
    return reinterpret_cast<ncat_tree*>(new Tree(build_twon(pts, n, minvalues, minhalfsize)));
}

ncat_tree* ncat_tree_kmeans(const double* pts, int64_t n, int64_t nclusters, int64_t minvalues) {
    return reinterpret_cast<ncat_tree*>(new Tree(build_kmeans(pts, n, nclusters, minvalues)));
}

int64_t ncat_tree_nnodes(const ncat_tree* h) {
    return (int64_t)reinterpret_cast<const Tree*>(h)->parent.size();
}

// parent/level/lo/hi/radius: nnodes; childptr: nnodes+1; children: nnodes-1;
// perm: npoints; center: 3*nnodes.
void ncat_tree_export(const ncat_tree* h, int64_t* parent, int64_t* level, int64_t* lo, int64_t* hi,
                      int64_t* childptr, int64_t* children, int64_t* perm, double* center, double* radius) {
    const Tree& t = *reinterpret_cast<const Tree*>(h);
    const size_t nn = t.parent.size();
    std::copy(t.parent.begin(), t.parent.end(), parent);
    std::copy(t.level.begin(), t.level.end(), level);
    std::copy(t.lo.begin(), t.lo.end(), lo);
    std::copy(t.hi.begin(), t.hi.end(), hi);
    std::copy(t.childptr.begin(), t.childptr.end(), childptr);
    std::copy(t.children.begin(), t.children.end(), children);
    std::copy(t.perm.begin(), t.perm.end(), perm);
    std::copy(t.center.begin(), t.center.begin() + 3 * nn, center);
    std::copy(t.radius.begin(), t.radius.end(), radius);
}

void ncat_tree_free(ncat_tree* h) { delete reinterpret_cast<Tree*>(h); }

ncat_partition* ncat_interactions(int64_t nn_a, const int64_t* childptr_a, const int64_t* children_a,
                                  const double* center_a, const double* radius_a, int64_t nn_b,
                                  const int64_t* childptr_b, const int64_t* children_b,
                                  const double* center_b, const double* radius_b, double eta) {
    TreeView a{nn_a, childptr_a, children_a, center_a, radius_a};
    TreeView b{nn_b, childptr_b, children_b, center_b, radius_b};
    Criterion c;
    c.eta = eta;
    return reinterpret_cast<ncat_partition*>(new Partition(interactions(a, b, c)));
}

// interactions with the wideband admissibility isnearwideband(k; etalf, etahf, gamma)
ncat_partition* ncat_interactions_wideband(int64_t nn_a, const int64_t* childptr_a, const int64_t* children_a,
                                           const double* center_a, const double* radius_a, int64_t nn_b,
                                           const int64_t* childptr_b, const int64_t* children_b,
                                           const double* center_b, const double* radius_b, double k, double etalf,
                                           double etahf, double gamma) {
    TreeView a{nn_a, childptr_a, children_a, center_a, radius_a};
    TreeView b{nn_b, childptr_b, children_b, center_b, radius_b};
    Criterion c;
    c.wideband = 1;
    c.k = k;
    c.etalf = etalf;
    c.etahf = etahf;
    c.gamma = gamma;
    return reinterpret_cast<ncat_partition*>(new Partition(interactions(a, b, c)));
}

void ncat_partition_sizes(const ncat_partition* h, int64_t* nnear, int64_t* nfar) {
    const Partition& p = *reinterpret_cast<const Partition*>(h);
    *nnear = (int64_t)p.nearnodes.size();
    *nfar = (int64_t)p.farnodes.size();
}

void ncat_partition_export(const ncat_partition* h, int64_t* nearptr, int64_t* nearnodes, int64_t* farptr,
                           int64_t* farnodes) {
    const Partition& p = *reinterpret_cast<const Partition*>(h);
    std::copy(p.nearptr.begin(), p.nearptr.end(), nearptr);
    std::copy(p.nearnodes.begin(), p.nearnodes.end(), nearnodes);
    std::copy(p.farptr.begin(), p.farptr.end(), farptr);
    std::copy(p.farnodes.begin(), p.farnodes.end(), farnodes);
}

void ncat_partition_free(ncat_partition* h) { delete reinterpret_cast<Partition*>(h); }

ncat_csr* ncat_farpivots(int64_t nn_self, const int64_t* parent, const int64_t* lo, const int64_t* hi,
                         const double* center, const int64_t* other_lo, const int64_t* other_hi,
                         const int64_t* other_perm, const double* other_pts, const int64_t* farptr,
                         const int64_t* farnodes, int64_t maxrank, double weight, double oversample) {
    return reinterpret_cast<ncat_csr*>(new Csr(farpivots(nn_self, parent, lo, hi, center, other_lo, other_hi,
                                                         other_perm, other_pts, farptr, farnodes, maxrank, weight,
                                                         oversample)));
}

// Phase 1 over explicit far fields (one list of far nodes of the other tree
// per item, 0-based CSR ffptr/ffnodes), item t of node itemnode[t]: the
// directional form, one item per (cluster, direction) with its dirfarfield.
ncat_csr* ncat_farpivots_lists(int64_t nitems, const int64_t* itemnode, const int64_t* lo, const int64_t* hi,
                               const double* center, const int64_t* other_lo, const int64_t* other_hi,
                               const int64_t* other_perm, const double* other_pts, const int64_t* ffptr,
                               const int64_t* ffnodes, int64_t maxrank, double weight, double oversample) {
    return reinterpret_cast<ncat_csr*>(new Csr(select_farpivots(
        nitems,
        [&](int64_t t, std::vector<int64_t>& F) { F.assign(ffnodes + ffptr[t], ffnodes + ffptr[t + 1]); },
        [&](int64_t t) { return hi[itemnode[t]] - lo[itemnode[t]]; },
        [&](int64_t t) { return center + 3 * itemnode[t]; }, other_lo, other_hi, other_perm, other_pts, maxrank,
        weight, oversample)));
}

// dirfarfield of every direction index (abstractdirections.jl:205-243), as a
// 0-based CSR of far nodes. Inputs (0-based): parent (-1 at the root),
// dirptr (nn+1) direction indices of each node, dirs (direction labels, 0 =
// low frequency), chdirptr (nn+1) / chdirmap: for each direction of the parent
// the local direction of the node, diridxtofarptr (ne+1) and farnodes.
// Label 0: the node's own far nodes of direction 0, then those of each
// ancestor while it holds direction 0. Other labels: the own far nodes, then
// level by level upwards the far nodes of the parent directions that map onto
// the current set, in increasing local order.
ncat_csr* ncat_dirfarfield(int64_t nn, const int64_t* parent, const int64_t* dirptr, const int64_t* dirs,
                           const int64_t* chdirptr, const int64_t* chdirmap, const int64_t* diridxtofarptr,
                           const int64_t* farnodes) {
    const int64_t ne = dirptr[nn];
    std::vector<int64_t> node_of(ne);
    for (int64_t t = 0; t < nn; ++t)
        for (int64_t e = dirptr[t]; e < dirptr[t + 1]; ++e) node_of[e] = t;
    std::vector<std::vector<int64_t>> out(ne);
#pragma omp parallel
    {
        std::vector<int64_t> S, next;
        std::vector<char> mark;
#pragma omp for schedule(dynamic, 256)
        for (int64_t e = 0; e < ne; ++e) {
            std::vector<int64_t>& F = out[e];
            auto own = [&](int64_t d) {
                for (int64_t k = diridxtofarptr[d]; k < diridxtofarptr[d + 1]; ++k) F.push_back(farnodes[k]);
            };
            const int64_t t = node_of[e];
            if (dirs[e] == 0) {
                for (int64_t cur = t; cur >= 0; cur = parent[cur]) {
                    int64_t z = -1;
                    for (int64_t d = dirptr[cur]; d < dirptr[cur + 1]; ++d)
                        if (dirs[d] == 0) { z = d; break; }
                    if (z < 0) break;
                    own(z);
                }
                continue;
            }
            own(e);
            S.assign(1, e - dirptr[t]);
            for (int64_t cur = t; parent[cur] >= 0 && !S.empty(); cur = parent[cur]) {
                const int64_t nd = dirptr[cur + 1] - dirptr[cur];
                mark.assign((size_t)nd, 0);
                for (int64_t l : S) mark[l] = 1;
                next.clear();
                for (int64_t j = 0; j < chdirptr[cur + 1] - chdirptr[cur]; ++j) {
                    const int64_t l = chdirmap[chdirptr[cur] + j];
                    if (l >= 0 && l < nd && mark[l]) next.push_back(j);
                }
                const int64_t P = parent[cur];
                for (int64_t j : next) own(dirptr[P] + j);
                S.swap(next);
            }
        }
    }
    Csr* c = new Csr();
    c->ptr.assign(ne + 1, 0);
    for (int64_t e = 0; e < ne; ++e) c->ptr[e + 1] = c->ptr[e] + (int64_t)out[e].size();
    c->idx.resize(c->ptr[ne]);
    for (int64_t e = 0; e < ne; ++e) std::copy(out[e].begin(), out[e].end(), c->idx.begin() + c->ptr[e]);
    return reinterpret_cast<ncat_csr*>(c);
}

// farfield of every node as a 0-based CSR: its own far nodes, then those of
// every ancestor (farinteractions.jl:38-60).
ncat_csr* ncat_farfield_csr(int64_t nn, const int64_t* parent, const int64_t* farptr, const int64_t* farnodes) {
    Csr* c = new Csr();
    c->ptr.assign(nn + 1, 0);
    for (int64_t t = 0; t < nn; ++t) {
        int64_t k = 0;
        for (int64_t a = t; a >= 0; a = parent[a]) k += farptr[a + 1] - farptr[a];
        c->ptr[t + 1] = c->ptr[t] + k;
    }
    c->idx.resize(c->ptr[nn]);
#pragma omp parallel for schedule(dynamic, 256)
    for (int64_t t = 0; t < nn; ++t) {
        int64_t o = c->ptr[t];
        for (int64_t a = t; a >= 0; a = parent[a])
            for (int64_t k = farptr[a]; k < farptr[a + 1]; ++k) c->idx[o++] = farnodes[k];
    }
    return reinterpret_cast<ncat_csr*>(c);
}

// Packs nb complex blocks (interleaved doubles): block b is rows[b] x cols[b]
// column-major at dst + 2 dst_off[b], read from src at src_off[b] with leading
// dimension ld[b], as is (transpose = 0: dst(i, j) = src(i + j ld)) or
// transposed (transpose = 1: dst(i, j) = src(j + i ld)).
void ncat_gather_blocks(int64_t nb, const int64_t* src_off, const int64_t* ld, const int64_t* rows,
                        const int64_t* cols, int32_t transpose, const double* src, const int64_t* dst_off,
                        double* dst) {
#pragma omp parallel for schedule(dynamic, 256)
    for (int64_t b = 0; b < nb; ++b) {
        const int64_t m = rows[b], n = cols[b], l = ld[b];
        const double* s0 = src + 2 * src_off[b];
        double* d0 = dst + 2 * dst_off[b];
        for (int64_t j = 0; j < n; ++j)
            for (int64_t i = 0; i < m; ++i) {
                const int64_t k = transpose ? j + i * l : i + j * l;
                d0[2 * (i + j * m)] = s0[2 * k];
                d0[2 * (i + j * m) + 1] = s0[2 * k + 1];
            }
    }
}

// CSR transpose with the rows of every column in increasing order (a
// counting sort; farinteractions.jl:66-92).
ncat_csr* ncat_csr_transpose(int64_t nrows, const int64_t* ptr, const int64_t* idx, int64_t ncols) {
    Csr* c = new Csr();
    c->ptr.assign(ncols + 1, 0);
    const int64_t nnz = ptr[nrows] - ptr[0];
    for (int64_t k = 0; k < nnz; ++k) c->ptr[idx[ptr[0] + k] + 1]++;
    for (int64_t j = 0; j < ncols; ++j) c->ptr[j + 1] += c->ptr[j];
    c->idx.resize(nnz);
    std::vector<int64_t> cur(c->ptr.begin(), c->ptr.end() - 1);
    for (int64_t r = 0; r < nrows; ++r)
        for (int64_t k = ptr[r]; k < ptr[r + 1]; ++k) c->idx[cur[idx[k]]++] = r;
    return reinterpret_cast<ncat_csr*>(c);
}

int64_t ncat_csr_nnz(const ncat_csr* h) { return (int64_t)reinterpret_cast<const Csr*>(h)->idx.size(); }

void ncat_csr_export(const ncat_csr* h, int64_t* ptr, int64_t* idx) {
    const Csr& c = *reinterpret_cast<const Csr*>(h);
    std::copy(c.ptr.begin(), c.ptr.end(), ptr);
    std::copy(c.idx.begin(), c.idx.end(), idx);
}

void ncat_csr_free(ncat_csr* h) { delete reinterpret_cast<Csr*>(h); }

}  // extern "C"
