// Cluster trees built on the device: the regular 2^3 tree (TwoN) and the
// deterministic k-means tree. Both are level-synchronous: one level of nodes is
// split per step, node ids are handed out level by level in the order of the
// parents (an exclusive scan over the level's child counts), and the point
// permutation is reordered inside every split node only, stably. The host
// reads back one count per level (the number of new nodes).
//
// The trees are the same, bit for bit, as the host builders of
// host/nca_tree.cpp (build_twon, build_kmeans): same comparisons, same
// floating-point operations in the same order (sums are fixed-order: slot
// s = q mod 256 accumulates its terms in order, then a halving tree), and
// -ffp-contract=off keeps products and sums apart.

#include <hipcub/hipcub.hpp>

#include <cmath>
#include <limits>
#include <vector>

#include "nca_common.hpp"

struct nca_tree {
    int64_t n = 0;
    // 0-based; children of the node i are the ids childptr[i] + 1 .. childptr[i + 1]
    std::vector<int64_t> parent, level, lo, hi, childptr, perm;
    std::vector<double> center, radius;
};

namespace nca {
namespace {

constexpr int kTreeThreads = 256;
constexpr int kMaxClusters = 8;

inline unsigned tblocks(int64_t n) { return (unsigned)((n + 255) / 256); }

__device__ __forceinline__ double tdist2(double ax, double ay, double az, double bx, double by, double bz) {
    const double dx = ax - bx, dy = ay - by, dz = az - bz;
    return dx * dx + dy * dy + dz * dz;
}

// node records of a level-synchronous build (0-based)
struct NodeArrays {
    DevBuf<int32_t> parent, level, lo, hi, nch;
    DevBuf<double> center, half;  // center 3 per node; half: box half size (TwoN) or radius (k-means)
    int64_t cap = 0;

    template <class T>
    static void grow_one(DevBuf<T>& b, int64_t have, int64_t cap, int per, hipStream_t st) {
        DevBuf<T> nb;
        nb.alloc((size_t)(cap * per));
        if (have > 0) NCA_HIP(hipMemcpyAsync(nb.p, b.p, sizeof(T) * have * per, hipMemcpyDeviceToDevice, st));
        b = std::move(nb);
    }
    void reserve(int64_t have, int64_t need, hipStream_t st) {
        if (need <= cap) return;
        const int64_t c = std::max<int64_t>(need, 2 * cap);
        grow_one(parent, have, c, 1, st);
        grow_one(level, have, c, 1, st);
        grow_one(lo, have, c, 1, st);
        grow_one(hi, have, c, 1, st);
        grow_one(nch, have, c, 1, st);
        grow_one(center, have, c, 3, st);
        grow_one(half, have, c, 1, st);
        cap = c;
    }
};

// ---- regular 2^3 tree ------------------------------------------------------

// partial bounding boxes: block b reduces a grid-stride share to 6 values
__global__ void k_bbox_partial(const double* __restrict__ xyz, int64_t n, double* __restrict__ part) {
    __shared__ double s[6][kTreeThreads];
    double v[6] = {INFINITY, INFINITY, INFINITY, -INFINITY, -INFINITY, -INFINITY};
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        for (int d = 0; d < 3; ++d) {
            const double x = xyz[3 * i + d];
            v[d] = fmin(v[d], x);
            v[3 + d] = fmax(v[3 + d], x);
        }
    for (int d = 0; d < 6; ++d) s[d][threadIdx.x] = v[d];
    __syncthreads();
    for (int w = kTreeThreads / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w)
            for (int d = 0; d < 3; ++d) {
                s[d][threadIdx.x] = fmin(s[d][threadIdx.x], s[d][threadIdx.x + w]);
                s[3 + d][threadIdx.x] = fmax(s[3 + d][threadIdx.x], s[3 + d][threadIdx.x + w]);
            }
        __syncthreads();
    }
    if (threadIdx.x < 6) part[6 * blockIdx.x + threadIdx.x] = s[threadIdx.x][0];
}

// key of position i: (lo of its node) * 8 + octant, the octant only in a
// node of the current level that may split (else 0, so the position keeps
// its place under the stable sort)
__global__ void k_twon_keys(const double* __restrict__ xyz, const int32_t* __restrict__ perm,
                            const int32_t* __restrict__ pnode, int64_t n, int32_t first,
                            const int32_t* __restrict__ lo, const int32_t* __restrict__ hi,
                            const double* __restrict__ center, const double* __restrict__ half, int64_t minvalues,
                            double minhalfsize, uint64_t* __restrict__ keys, int32_t* __restrict__ idx) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int t = pnode[i];
    uint64_t key = (uint64_t)lo[t] << 3;
    if (t >= first && (int64_t)(hi[t] - lo[t]) > minvalues && half[t] > minhalfsize) {
        const double* x = xyz + 3 * (int64_t)perm[i];
        const double* c = center + 3 * (int64_t)t;
        key |= (uint64_t)((x[0] >= c[0] ? 1 : 0) | (x[1] >= c[1] ? 2 : 0) | (x[2] >= c[2] ? 4 : 0));
    }
    keys[i] = key;
    idx[i] = (int32_t)i;
}

__global__ void k_tree_gather(const int32_t* __restrict__ idx, int64_t n, const int32_t* __restrict__ perm,
                              const int32_t* __restrict__ pnode, int32_t* __restrict__ perm2,
                              int32_t* __restrict__ pnode2) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int32_t j = idx[i];
    perm2[i] = perm[j];
    pnode2[i] = pnode[j];
}

__device__ __forceinline__ int32_t lower_bound_key(const uint64_t* __restrict__ keys, int32_t a, int32_t b,
                                                   uint64_t v) {
    while (a < b) {
        const int32_t m = a + (b - a) / 2;
        if (keys[m] < v) a = m + 1;
        else b = m;
    }
    return a;
}

// per node of the level: octant counts (from the sorted keys of its range),
// split decision and number of children
__global__ void k_twon_split(int32_t first, int32_t nl, const uint64_t* __restrict__ keys,
                             const int32_t* __restrict__ lo, const int32_t* __restrict__ hi,
                             const double* __restrict__ half, int64_t minvalues, double minhalfsize,
                             int32_t* __restrict__ cnt8, int32_t* __restrict__ nch) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= nl) return;
    const int t = first + j;
    const int32_t l = lo[t], h = hi[t];
    const double hs = half[t];
    int c = 0;
    if ((int64_t)(h - l) > minvalues && hs > minhalfsize) {
        const uint64_t base = (uint64_t)l << 3;
        int nonempty = 0;
        int32_t s = l;
        for (int o = 0; o < 8; ++o) {
            const int32_t e = lower_bound_key(keys, s, h, base + o + 1);
            cnt8[8 * j + o] = e - s;
            nonempty += e > s;
            s = e;
        }
        c = (nonempty <= 1 && hs <= 1e-12) ? 0 : nonempty;
    }
    nch[j] = c;
}

// child records of every split node of the level, ids first_new + off[j] + ...
__global__ void k_twon_emit(int32_t first, int32_t nl, int32_t first_new, const int32_t* __restrict__ off,
                            const int32_t* __restrict__ ncl, const int32_t* __restrict__ cnt8,
                            int32_t* __restrict__ parent, int32_t* __restrict__ level, int32_t* __restrict__ lo,
                            int32_t* __restrict__ hi, double* __restrict__ center, double* __restrict__ half,
                            int32_t* __restrict__ nch, int32_t* __restrict__ childmap) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= nl) return;
    const int t = first + j;
    nch[t] = ncl[j];
    if (ncl[j] == 0) return;
    int id = first_new + off[j];
    int32_t s = lo[t];
    const double q = 0.5 * half[t];
    const double cx = center[3 * t], cy = center[3 * t + 1], cz = center[3 * t + 2];
    for (int o = 0; o < 8; ++o) {
        const int32_t c = cnt8[8 * j + o];
        if (c == 0) continue;
        parent[id] = t;
        level[id] = level[t] + 1;
        lo[id] = s;
        hi[id] = s + c;
        center[3 * id] = cx + ((o & 1) ? q : -q);
        center[3 * id + 1] = cy + ((o & 2) ? q : -q);
        center[3 * id + 2] = cz + ((o & 4) ? q : -q);
        half[id] = q;
        childmap[8 * j + o] = id;
        s += c;
        ++id;
    }
}

__global__ void k_twon_move(const uint64_t* __restrict__ keys, int64_t n, int32_t first,
                            const int32_t* __restrict__ ncl, const int32_t* __restrict__ childmap,
                            int32_t* __restrict__ pnode) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int t = pnode[i];
    if (t < first || ncl[t - first] == 0) return;
    pnode[i] = childmap[8 * (t - first) + (int)(keys[i] & 7)];
}

// ---- k-means tree ------------------------------------------------------------

__device__ __forceinline__ void halving_sum(double* red, int tid) {
    for (int w = kTreeThreads / 2; w > 0; w >>= 1) {
        if (tid < w) red[tid] += red[tid + w];
        __syncthreads();
    }
}

// argmax over the workgroup: larger value wins, ties go to the lower index
__device__ __forceinline__ void argmax_reduce(double* rv, int32_t* ri, int tid) {
    for (int w = kTreeThreads / 2; w > 0; w >>= 1) {
        if (tid < w) {
            const double a = rv[tid], b = rv[tid + w];
            const int32_t ia = ri[tid], ib = ri[tid + w];
            if (b > a || (b == a && ib >= 0 && (ia < 0 || ib < ia))) {
                rv[tid] = b;
                ri[tid] = ib;
            }
        }
        __syncthreads();
    }
}

// points in permutation order: xs[i] = xyz[perm[i]]
__global__ void k_km_gather(const double* __restrict__ xyz, const int32_t* __restrict__ perm, int64_t n,
                            double* __restrict__ xs) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int64_t p = perm[i];
    xs[3 * i] = xyz[3 * p];
    xs[3 * i + 1] = xyz[3 * p + 1];
    xs[3 * i + 2] = xyz[3 * p + 2];
}

// One workgroup per node of the level: its center (centroid) and radius
// (largest distance to it), then, for a node above minvalues, the seeds by
// farthest-point traversal and the Lloyd iterations; writes the assignment of
// every position, the cluster counts and the number of children.
__global__ void __launch_bounds__(kTreeThreads) k_km_node(
    int32_t first, const double* __restrict__ xs, const int32_t* __restrict__ lo, const int32_t* __restrict__ hi,
    int32_t k, int64_t minvalues, double* __restrict__ center, double* __restrict__ radius,
    int8_t* __restrict__ asg, int32_t* __restrict__ cntk, int32_t* __restrict__ ncl) {
    __shared__ double red[kTreeThreads];
    __shared__ int32_t redi[kTreeThreads];
    __shared__ double cent[3 * kMaxClusters];
    __shared__ int32_t ccount[kMaxClusters];
    __shared__ int changed;
    const int j = blockIdx.x, tid = threadIdx.x, t = first + j;
    const int32_t l = lo[t], m = hi[t] - l;
    const double* x = xs + 3 * (int64_t)l;
    // centroid
    double c[3];
    for (int d = 0; d < 3; ++d) {
        double s = 0.0;
        for (int32_t q = tid; q < m; q += kTreeThreads) s += x[3 * q + d];
        red[tid] = s;
        __syncthreads();
        halving_sum(red, tid);
        c[d] = m > 0 ? red[0] / (double)m : 0.0;
        __syncthreads();
    }
    double r2 = 0.0;
    for (int32_t q = tid; q < m; q += kTreeThreads)
        r2 = fmax(r2, tdist2(x[3 * q], x[3 * q + 1], x[3 * q + 2], c[0], c[1], c[2]));
    red[tid] = r2;
    __syncthreads();
    for (int w = kTreeThreads / 2; w > 0; w >>= 1) {
        if (tid < w) red[tid] = fmax(red[tid], red[tid + w]);
        __syncthreads();
    }
    if (tid == 0) {
        center[3 * t] = c[0];
        center[3 * t + 1] = c[1];
        center[3 * t + 2] = c[2];
        radius[t] = sqrt(red[0]);
        ncl[j] = 0;
    }
    if ((int64_t)m <= minvalues || m == 0) return;
    __syncthreads();
    // first seed: the point farthest from the center (lowest position on ties)
    {
        double bv = -1.0;
        int32_t bi = -1;
        for (int32_t q = tid; q < m; q += kTreeThreads) {
            const double d2 = tdist2(x[3 * q], x[3 * q + 1], x[3 * q + 2], c[0], c[1], c[2]);
            if (d2 > bv) { bv = d2; bi = q; }
        }
        red[tid] = bv;
        redi[tid] = bi;
        __syncthreads();
        argmax_reduce(red, redi, tid);
        const int32_t f = redi[0] < 0 ? 0 : redi[0];  // no comparison held (NaN): position 0, as the host
        if (tid < 3) cent[tid] = x[3 * f + tid];
        __syncthreads();
    }
    // next seeds: farthest from the seeds so far
    int kc = 1;
    for (; kc < k; ++kc) {
        double bv = 0.0;
        int32_t bi = -1;
        for (int32_t q = tid; q < m; q += kTreeThreads) {
            double md = std::numeric_limits<double>::max();
            for (int s = 0; s < kc; ++s)
                md = fmin(md, tdist2(x[3 * q], x[3 * q + 1], x[3 * q + 2], cent[3 * s], cent[3 * s + 1],
                                     cent[3 * s + 2]));
            if (md > bv) { bv = md; bi = q; }
        }
        red[tid] = bv;
        redi[tid] = bi;
        __syncthreads();
        argmax_reduce(red, redi, tid);
        const int32_t arg = redi[0];
        if (arg < 0) break;  // every point on a seed: uniform across the workgroup
        if (tid < 3) cent[3 * kc + tid] = x[3 * arg + tid];
        __syncthreads();
    }
    if (kc < 2) return;
    int8_t* a = asg + l;
    for (int32_t q = tid; q < m; q += kTreeThreads) a[q] = -1;
    __syncthreads();
    for (int it = 0; it < 100; ++it) {
        if (tid == 0) changed = 0;
        __syncthreads();
        int ch = 0;
        for (int32_t q = tid; q < m; q += kTreeThreads) {
            int best = 0;
            double bd = tdist2(x[3 * q], x[3 * q + 1], x[3 * q + 2], cent[0], cent[1], cent[2]);
            for (int s = 1; s < kc; ++s) {
                const double d2 = tdist2(x[3 * q], x[3 * q + 1], x[3 * q + 2], cent[3 * s], cent[3 * s + 1],
                                         cent[3 * s + 2]);
                if (d2 < bd) { bd = d2; best = s; }
            }
            if (best != a[q]) { a[q] = (int8_t)best; ch = 1; }
        }
        if (ch) changed = 1;  // benign: every writer stores 1
        __syncthreads();
        if (!changed) break;
        for (int s = 0; s < kc; ++s) {
            int32_t cn = 0;
            for (int32_t q = tid; q < m; q += kTreeThreads) cn += a[q] == s;
            redi[tid] = cn;
            __syncthreads();
            for (int w = kTreeThreads / 2; w > 0; w >>= 1) {
                if (tid < w) redi[tid] += redi[tid + w];
                __syncthreads();
            }
            const int32_t total = redi[0];
            for (int d = 0; d < 3; ++d) {
                double sum = 0.0;
                for (int32_t q = tid; q < m; q += kTreeThreads)
                    if (a[q] == s) sum += x[3 * q + d];
                __syncthreads();
                red[tid] = sum;
                __syncthreads();
                halving_sum(red, tid);
                if (tid == 0 && total > 0) cent[3 * s + d] = red[0] / (double)total;
                __syncthreads();
            }
        }
    }
    // final counts
    for (int s = 0; s < kc; ++s) {
        int32_t cn = 0;
        for (int32_t q = tid; q < m; q += kTreeThreads) cn += a[q] == s;
        redi[tid] = cn;
        __syncthreads();
        for (int w = kTreeThreads / 2; w > 0; w >>= 1) {
            if (tid < w) redi[tid] += redi[tid + w];
            __syncthreads();
        }
        if (tid == 0) ccount[s] = redi[0];
        __syncthreads();
    }
    if (tid == 0) {
        int nonempty = 0;
        for (int s = 0; s < kMaxClusters; ++s) {
            const int32_t cs = s < kc ? ccount[s] : 0;
            cntk[kMaxClusters * j + s] = cs;
            nonempty += cs > 0;
        }
        ncl[j] = nonempty < 2 ? 0 : nonempty;
    }
}

// One workgroup per split node: stable partition of its positions by cluster
// (wave ballots give the rank among equal clusters of a 256-position chunk)
// and the child records.
__global__ void __launch_bounds__(kTreeThreads) k_km_emit(
    int32_t first, int32_t first_new, const int32_t* __restrict__ off, const int32_t* __restrict__ ncl,
    const int32_t* __restrict__ cntk, const int8_t* __restrict__ asg, const int32_t* __restrict__ perm,
    int32_t* __restrict__ perm2, int32_t* __restrict__ parent, int32_t* __restrict__ level, int32_t* __restrict__ lo,
    int32_t* __restrict__ hi, int32_t* __restrict__ nch) {
    __shared__ int32_t wtot[kTreeThreads / 64][kMaxClusters];
    __shared__ int32_t start[kMaxClusters];
    const int j = blockIdx.x, tid = threadIdx.x, t = first + j;
    if (tid == 0) nch[t] = ncl[j];
    if (ncl[j] == 0) return;
    const int32_t l = lo[t], m = hi[t] - l;
    if (tid == 0) {
        int32_t s = l;
        int id = first_new + off[j];
        for (int c = 0; c < kMaxClusters; ++c) {
            const int32_t cn = cntk[kMaxClusters * j + c];
            start[c] = s;
            if (cn > 0) {
                parent[id] = t;
                level[id] = level[t] + 1;
                lo[id] = s;
                hi[id] = s + cn;
                ++id;
            }
            s += cn;
        }
    }
    __syncthreads();
    const int lane = tid & 63, wave = tid >> 6;
    const uint64_t below = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    for (int32_t base = 0; base < m; base += kTreeThreads) {
        const int32_t q = base + tid;
        const int c = q < m ? (int)asg[l + q] : -1;
        int32_t rank = 0;
        for (int s = 0; s < kMaxClusters; ++s) {
            const uint64_t mask = __ballot(c == s);
            if (lane == 0) wtot[wave][s] = __popcll(mask);
            if (c == s) rank = __popcll(mask & below);
        }
        __syncthreads();
        if (c >= 0) {
            int32_t pos = start[c] + rank;
            for (int w = 0; w < wave; ++w) pos += wtot[w][c];
            perm2[pos] = perm[l + q];
        }
        __syncthreads();
        if (tid < kMaxClusters) {
            int32_t add = 0;
            for (int w = 0; w < kTreeThreads / 64; ++w) add += wtot[w][tid];
            start[tid] += add;
        }
        __syncthreads();
    }
}

__global__ void k_tree_radius(const double* __restrict__ half, int64_t nn, double* __restrict__ radius) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < nn) radius[i] = half[i] * sqrt(3.0);
}

// Scratch of a build: hipcub temporary storage and scans.
struct TreeScan {
    DevBuf<char> tmp;
    DevBuf<int32_t> off;
    int64_t scan(const int32_t* in, int64_t cnt, hipStream_t st) {
        // in[cnt] must be 0: the scan runs over cnt + 1 values, off[cnt] is the total
        off.reserve((size_t)cnt + 1);
        size_t bytes = 0;
        NCA_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, bytes, in, off.p, (int)cnt + 1, st));
        if (tmp.n < bytes) tmp.alloc(bytes);
        NCA_HIP(hipcub::DeviceScan::ExclusiveSum(tmp.p, bytes, in, off.p, (int)cnt + 1, st));
        int32_t total = 0;
        NCA_HIP(hipMemcpyAsync(&total, off.p + cnt, sizeof(int32_t), hipMemcpyDeviceToHost, st));
        NCA_HIP(hipStreamSynchronize(st));
        return total;
    }
};

std::unique_ptr<nca_tree> download_tree(NodeArrays& nd, int64_t nn, const DevBuf<int32_t>& perm, int64_t n,
                                        const double* radius_dev, TreeScan& sc, hipStream_t st) {
    auto t = std::make_unique<nca_tree>();
    t->n = n;
    std::vector<int32_t> par(nn), lev(nn), lo(nn), hi(nn), pm(n), cp(nn + 1);
    // childptr: exclusive scan of the child counts (children are the ids 1 .. nn-1 in order)
    DevBuf<int32_t> nch1;
    nch1.alloc((size_t)nn + 1);
    NCA_HIP(hipMemcpyAsync(nch1.p, nd.nch.p, sizeof(int32_t) * nn, hipMemcpyDeviceToDevice, st));
    NCA_HIP(hipMemsetAsync(nch1.p + nn, 0, sizeof(int32_t), st));
    sc.scan(nch1.p, nn, st);
    t->center.resize(3 * nn);
    t->radius.resize(nn);
    NCA_HIP(hipMemcpyAsync(par.data(), nd.parent.p, sizeof(int32_t) * nn, hipMemcpyDeviceToHost, st));
    NCA_HIP(hipMemcpyAsync(lev.data(), nd.level.p, sizeof(int32_t) * nn, hipMemcpyDeviceToHost, st));
    NCA_HIP(hipMemcpyAsync(lo.data(), nd.lo.p, sizeof(int32_t) * nn, hipMemcpyDeviceToHost, st));
    NCA_HIP(hipMemcpyAsync(hi.data(), nd.hi.p, sizeof(int32_t) * nn, hipMemcpyDeviceToHost, st));
    NCA_HIP(hipMemcpyAsync(cp.data(), sc.off.p, sizeof(int32_t) * (nn + 1), hipMemcpyDeviceToHost, st));
    if (n) NCA_HIP(hipMemcpyAsync(pm.data(), perm.p, sizeof(int32_t) * n, hipMemcpyDeviceToHost, st));
    NCA_HIP(hipMemcpyAsync(t->center.data(), nd.center.p, sizeof(double) * 3 * nn, hipMemcpyDeviceToHost, st));
    NCA_HIP(hipMemcpyAsync(t->radius.data(), radius_dev, sizeof(double) * nn, hipMemcpyDeviceToHost, st));
    NCA_HIP(hipStreamSynchronize(st));
    t->parent.assign(par.begin(), par.end());
    t->level.assign(lev.begin(), lev.end());
    t->lo.assign(lo.begin(), lo.end());
    t->hi.assign(hi.begin(), hi.end());
    t->childptr.assign(cp.begin(), cp.end());
    t->perm.assign(pm.begin(), pm.end());
    return t;
}

std::unique_ptr<nca_tree> build_twon_device(const nca_points* pts, int64_t minvalues, double minhalfsize,
                                            hipStream_t st) {
    const int64_t n = pts->n;
    // bounding cube
    double mn[3] = {0, 0, 0}, mx[3] = {0, 0, 0};
    if (n > 0) {
        const int nb = (int)std::min<int64_t>(256, tblocks(n));
        DevBuf<double> part((size_t)6 * nb);
        hipLaunchKernelGGL(k_bbox_partial, dim3(nb), dim3(kTreeThreads), 0, st, pts->xyz.p, n, part.p);
        NCA_HIP(hipGetLastError());
        std::vector<double> hp((size_t)6 * nb);
        NCA_HIP(hipMemcpyAsync(hp.data(), part.p, sizeof(double) * 6 * nb, hipMemcpyDeviceToHost, st));
        NCA_HIP(hipStreamSynchronize(st));
        for (int d = 0; d < 3; ++d) {
            mn[d] = hp[d];
            mx[d] = hp[3 + d];
        }
        for (int b = 1; b < nb; ++b)
            for (int d = 0; d < 3; ++d) {
                mn[d] = std::min(mn[d], hp[6 * b + d]);
                mx[d] = std::max(mx[d], hp[6 * b + 3 + d]);
            }
    }
    double h = 0.0;
    for (int d = 0; d < 3; ++d) h = std::max(h, 0.5 * (mx[d] - mn[d]));
    if (h <= 0.0) h = 1e-300;
    NodeArrays nd;
    nd.reserve(0, std::max<int64_t>(64, n / 16), st);
    {
        const int32_t z = 0, nn32 = (int32_t)n;
        const double c[3] = {0.5 * (mn[0] + mx[0]), 0.5 * (mn[1] + mx[1]), 0.5 * (mn[2] + mx[2])};
        const int32_t minus1 = -1;
        NCA_HIP(hipMemcpyAsync(nd.parent.p, &minus1, sizeof(int32_t), hipMemcpyHostToDevice, st));
        NCA_HIP(hipMemcpyAsync(nd.level.p, &z, sizeof(int32_t), hipMemcpyHostToDevice, st));
        NCA_HIP(hipMemcpyAsync(nd.lo.p, &z, sizeof(int32_t), hipMemcpyHostToDevice, st));
        NCA_HIP(hipMemcpyAsync(nd.hi.p, &nn32, sizeof(int32_t), hipMemcpyHostToDevice, st));
        NCA_HIP(hipMemcpyAsync(nd.center.p, c, sizeof(double) * 3, hipMemcpyHostToDevice, st));
        NCA_HIP(hipMemcpyAsync(nd.half.p, &h, sizeof(double), hipMemcpyHostToDevice, st));
        NCA_HIP(hipStreamSynchronize(st));
    }
    DevBuf<int32_t> perm, pnode, perm2, pnode2, idx, idx2, cnt8, ncl, childmap;
    DevBuf<uint64_t> keys, keys2;
    const size_t np = (size_t)std::max<int64_t>(n, 1);
    perm.alloc(np);
    pnode.alloc(np);
    perm2.alloc(np);
    pnode2.alloc(np);
    idx.alloc(np);
    idx2.alloc(np);
    keys.alloc(np);
    keys2.alloc(np);
    if (n > 0) {
        std::vector<int32_t> id(n);
        for (int64_t i = 0; i < n; ++i) id[i] = (int32_t)i;
        perm.upload(id, st);
        NCA_HIP(hipMemsetAsync(pnode.p, 0, sizeof(int32_t) * n, st));
    }
    int endbit = 1;
    while ((1ull << endbit) < ((uint64_t)n << 3) + 8) ++endbit;
    TreeScan sc;
    int64_t first = 0, end = 1;
    while (end > first && n > 0) {
        const int32_t nl = (int32_t)(end - first);
        hipLaunchKernelGGL(k_twon_keys, dim3(tblocks(n)), dim3(256), 0, st, pts->xyz.p, perm.p, pnode.p, n,
                           (int32_t)first, nd.lo.p, nd.hi.p, nd.center.p, nd.half.p, minvalues, minhalfsize,
                           keys.p, idx.p);
        NCA_HIP(hipGetLastError());
        size_t bytes = 0;
        NCA_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, bytes, keys.p, keys2.p, idx.p, idx2.p, (int)n, 0,
                                                   endbit, st));
        if (sc.tmp.n < bytes) sc.tmp.alloc(bytes);
        NCA_HIP(hipcub::DeviceRadixSort::SortPairs(sc.tmp.p, bytes, keys.p, keys2.p, idx.p, idx2.p, (int)n, 0,
                                                   endbit, st));
        hipLaunchKernelGGL(k_tree_gather, dim3(tblocks(n)), dim3(256), 0, st, idx2.p, n, perm.p, pnode.p, perm2.p,
                           pnode2.p);
        NCA_HIP(hipGetLastError());
        std::swap(perm, perm2);
        std::swap(pnode, pnode2);
        cnt8.reserve((size_t)8 * nl);
        ncl.reserve((size_t)nl + 1);
        childmap.reserve((size_t)8 * nl);
        NCA_HIP(hipMemsetAsync(ncl.p + nl, 0, sizeof(int32_t), st));
        hipLaunchKernelGGL(k_twon_split, dim3(tblocks(nl)), dim3(256), 0, st, (int32_t)first, nl, keys2.p,
                           nd.lo.p, nd.hi.p, nd.half.p, minvalues, minhalfsize, cnt8.p, ncl.p);
        NCA_HIP(hipGetLastError());
        const int64_t added = sc.scan(ncl.p, nl, st);
        NCA_REQUIRE(end + added < (1LL << 31), NCA_ERR_LIMIT, "nca_tree_twon: more than 2^31 nodes");
        nd.reserve(end, end + added, st);
        hipLaunchKernelGGL(k_twon_emit, dim3(tblocks(nl)), dim3(256), 0, st, (int32_t)first, nl, (int32_t)end,
                           sc.off.p, ncl.p, cnt8.p, nd.parent.p, nd.level.p, nd.lo.p, nd.hi.p, nd.center.p,
                           nd.half.p, nd.nch.p, childmap.p);
        NCA_HIP(hipGetLastError());
        if (added > 0) {
            hipLaunchKernelGGL(k_twon_move, dim3(tblocks(n)), dim3(256), 0, st, keys2.p, n, (int32_t)first, ncl.p,
                               childmap.p, pnode.p);
            NCA_HIP(hipGetLastError());
        }
        first = end;
        end += added;
    }
    const int64_t nn = end;
    if (n == 0) NCA_HIP(hipMemsetAsync(nd.nch.p, 0, sizeof(int32_t), st));
    DevBuf<double> rad((size_t)nn);
    hipLaunchKernelGGL(k_tree_radius, dim3(tblocks(nn)), dim3(256), 0, st, nd.half.p, nn, rad.p);
    NCA_HIP(hipGetLastError());
    return download_tree(nd, nn, perm, n, rad.p, sc, st);
}

std::unique_ptr<nca_tree> build_kmeans_device(const nca_points* pts, int64_t nclusters, int64_t minvalues,
                                              hipStream_t st) {
    const int64_t n = pts->n;
    const int k = (int)std::max<int64_t>(2, nclusters);
    NodeArrays nd;
    nd.reserve(0, std::max<int64_t>(64, n / 16), st);
    {
        const int32_t z = 0, nn32 = (int32_t)n, minus1 = -1;
        NCA_HIP(hipMemcpyAsync(nd.parent.p, &minus1, sizeof(int32_t), hipMemcpyHostToDevice, st));
        NCA_HIP(hipMemcpyAsync(nd.level.p, &z, sizeof(int32_t), hipMemcpyHostToDevice, st));
        NCA_HIP(hipMemcpyAsync(nd.lo.p, &z, sizeof(int32_t), hipMemcpyHostToDevice, st));
        NCA_HIP(hipMemcpyAsync(nd.hi.p, &nn32, sizeof(int32_t), hipMemcpyHostToDevice, st));
        NCA_HIP(hipStreamSynchronize(st));
    }
    const size_t np = (size_t)std::max<int64_t>(n, 1);
    DevBuf<int32_t> perm(np), perm2(np), cntk, ncl;
    DevBuf<double> xs(3 * np);
    DevBuf<int8_t> asg(np);
    {
        std::vector<int32_t> id(np);
        for (size_t i = 0; i < np; ++i) id[i] = (int32_t)i;
        perm.upload(id, st);
    }
    TreeScan sc;
    int64_t first = 0, end = 1;
    while (end > first) {
        const int32_t nl = (int32_t)(end - first);
        if (n > 0) {
            hipLaunchKernelGGL(k_km_gather, dim3(tblocks(n)), dim3(256), 0, st, pts->xyz.p, perm.p, n, xs.p);
            NCA_HIP(hipGetLastError());
        }
        cntk.reserve((size_t)kMaxClusters * nl);
        ncl.reserve((size_t)nl + 1);
        NCA_HIP(hipMemsetAsync(ncl.p + nl, 0, sizeof(int32_t), st));
        hipLaunchKernelGGL(k_km_node, dim3(nl), dim3(kTreeThreads), 0, st, (int32_t)first, xs.p, nd.lo.p, nd.hi.p,
                           (int32_t)k, minvalues, nd.center.p, nd.half.p, asg.p, cntk.p, ncl.p);
        NCA_HIP(hipGetLastError());
        const int64_t added = sc.scan(ncl.p, nl, st);
        NCA_REQUIRE(end + added < (1LL << 31), NCA_ERR_LIMIT, "nca_tree_kmeans: more than 2^31 nodes");
        nd.reserve(end, end + added, st);
        if (added > 0) NCA_HIP(hipMemcpyAsync(perm2.p, perm.p, sizeof(int32_t) * n, hipMemcpyDeviceToDevice, st));
        hipLaunchKernelGGL(k_km_emit, dim3(nl), dim3(kTreeThreads), 0, st, (int32_t)first, (int32_t)end, sc.off.p,
                           ncl.p, cntk.p, asg.p, perm.p, perm2.p, nd.parent.p, nd.level.p, nd.lo.p, nd.hi.p,
                           nd.nch.p);
        NCA_HIP(hipGetLastError());
        if (added > 0) std::swap(perm, perm2);
        first = end;
        end += added;
    }
    return download_tree(nd, end, perm, n, nd.half.p, sc, st);
}

}  // namespace
}  // namespace nca

using namespace nca;

extern "C" {

int32_t nca_tree_twon(nca_ctx* ctx, const nca_points* pts, int64_t minvalues, double minhalfsize, int64_t* nnodes,
                      nca_tree** out) {
    return guarded([&] {
        NCA_REQUIRE(ctx && pts && out, NCA_ERR_ARGUMENT, "nca_tree_twon: invalid argument");
        NCA_REQUIRE(pts->n < (1LL << 31) - 8, NCA_ERR_LIMIT, "nca_tree_twon: more than 2^31 points");
        set_device(ctx);
        auto t = build_twon_device(pts, minvalues, minhalfsize, ctx->stream);
        if (nnodes) *nnodes = (int64_t)t->parent.size();
        *out = t.release();
    });
}

int32_t nca_tree_kmeans(nca_ctx* ctx, const nca_points* pts, int64_t nclusters, int64_t minvalues,
                        int64_t* nnodes, nca_tree** out) {
    return guarded([&] {
        NCA_REQUIRE(ctx && pts && out, NCA_ERR_ARGUMENT, "nca_tree_kmeans: invalid argument");
        NCA_REQUIRE(nclusters <= kMaxClusters, NCA_ERR_LIMIT,
                    "nca_tree_kmeans: more than " + std::to_string(kMaxClusters) + " clusters per node");
        NCA_REQUIRE(pts->n < (1LL << 31) - 8, NCA_ERR_LIMIT, "nca_tree_kmeans: more than 2^31 points");
        set_device(ctx);
        auto t = build_kmeans_device(pts, nclusters, minvalues, ctx->stream);
        if (nnodes) *nnodes = (int64_t)t->parent.size();
        *out = t.release();
    });
}

int32_t nca_tree_export(const nca_tree* t, int64_t* parent, int64_t* level, int64_t* lo, int64_t* hi,
                        int64_t* childptr, int64_t* children, int64_t* perm, double* center, double* radius) {
    return guarded([&] {
        NCA_REQUIRE(t, NCA_ERR_ARGUMENT, "nca_tree_export: null handle");
        const size_t nn = t->parent.size();
        for (size_t i = 0; i < nn; ++i) {
            if (parent) parent[i] = t->parent[i] + 1;
            if (level) level[i] = t->level[i] + 1;
            if (lo) lo[i] = t->lo[i] + 1;
            if (hi) hi[i] = t->hi[i];
            if (radius) radius[i] = t->radius[i];
        }
        if (childptr)
            for (size_t i = 0; i <= nn; ++i) childptr[i] = t->childptr[i] + 1;
        if (children)
            for (size_t i = 1; i < nn; ++i) children[i - 1] = (int64_t)i + 1;
        if (perm)
            for (int64_t i = 0; i < t->n; ++i) perm[i] = t->perm[i] + 1;
        if (center) std::copy(t->center.begin(), t->center.end(), center);
    });
}

int32_t nca_tree_destroy(nca_tree* t) {
    return guarded([&] { delete t; });
}

}  // extern "C"
