// Admissibility partition on the device: the dual traversal of a test and a
// trial cluster tree into near (both leaves, not admissible) and far
// (admissible) cluster pairs, as CSR by test node with the trial nodes of a
// row in increasing order. Level-synchronous: every frontier pair is
// classified, then far pairs, near pairs and the next frontier are written
// at scanned offsets. Same criteria and operations as the host traversal
// (host/nca_tree.cpp interactions), so the partitions are equal.
//
// Compiled with -ffp-contract=off: the distance is the host's expression.

#include <hipcub/hipcub.hpp>

#include <vector>

#include "nca_common.hpp"

namespace nca {

struct PartCrit {
    int wideband;
    double eta, k, etalf, etahf, gamma;
};

struct PartTree {
    const int32_t* childptr;  // nn + 1, 0-based
    const int32_t* children;
    const double* center;     // 3 per node
    const double* radius;
};

__device__ __forceinline__ bool near_pair(const PartTree& a, const PartTree& b, int t, int s, const PartCrit& c) {
    const double ra = a.radius[t], rb = b.radius[s];
    const double dx = a.center[3 * t] - b.center[3 * s], dy = a.center[3 * t + 1] - b.center[3 * s + 1],
                 dz = a.center[3 * t + 2] - b.center[3 * s + 2];
    const double d = sqrt(dx * dx + dy * dy + dz * dz) - (ra + rb);
    if (!c.wideband) return 2.0 * max(ra, rb) > c.eta * d;
    const double sep = max(d, 0.0);
    if (c.k * min(2.0 * ra, 2.0 * rb) <= c.gamma) return 2.0 * max(ra, rb) > c.etalf * sep;
    return 4.0 * c.k * max(ra * ra, rb * rb) > c.etahf * sep;
}

// per frontier pair: far (1, 0, 0), near (0, 1, 0) or the number of child pairs
__global__ void k_part_classify(const int32_t* __restrict__ pt, const int32_t* __restrict__ ps, int64_t n,
                                PartTree a, PartTree b, PartCrit c, int32_t* __restrict__ nfar,
                                int32_t* __restrict__ nnear, int32_t* __restrict__ nchild) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int t = pt[i], s = ps[i];
    int f = 0, ne = 0, ch = 0;
    if (!near_pair(a, b, t, s, c)) {
        f = 1;
    } else {
        const int ct = a.childptr[t + 1] - a.childptr[t], cs = b.childptr[s + 1] - b.childptr[s];
        if (ct == 0 && cs == 0) ne = 1;
        else ch = (ct == 0 ? 1 : ct) * (cs == 0 ? 1 : cs);
    }
    nfar[i] = f;
    nnear[i] = ne;
    nchild[i] = ch;
}

// writes the far and near pairs as 64-bit keys t * nb + s and the child pairs
// in the host's order (test child outer, trial child inner)
__global__ void k_part_emit(const int32_t* __restrict__ pt, const int32_t* __restrict__ ps, int64_t n, PartTree a,
                            PartTree b, const int32_t* __restrict__ nfar, const int32_t* __restrict__ nnear,
                            const int32_t* __restrict__ nchild, const int64_t* __restrict__ ofar,
                            const int64_t* __restrict__ onear, const int64_t* __restrict__ ochild, int64_t nb,
                            int64_t farbase, int64_t nearbase, int64_t* __restrict__ farkeys,
                            int64_t* __restrict__ nearkeys, int32_t* __restrict__ qt, int32_t* __restrict__ qs) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int t = pt[i], s = ps[i];
    if (nfar[i]) farkeys[farbase + ofar[i]] = (int64_t)t * nb + s;
    if (nnear[i]) nearkeys[nearbase + onear[i]] = (int64_t)t * nb + s;
    if (nchild[i]) {
        int64_t o = ochild[i];
        const int t0 = a.childptr[t], t1 = a.childptr[t + 1], s0 = b.childptr[s], s1 = b.childptr[s + 1];
        if (t1 == t0) {
            for (int k = s0; k < s1; ++k, ++o) { qt[o] = t; qs[o] = b.children[k]; }
        } else if (s1 == s0) {
            for (int k = t0; k < t1; ++k, ++o) { qt[o] = a.children[k]; qs[o] = s; }
        } else {
            for (int k = t0; k < t1; ++k)
                for (int m = s0; m < s1; ++m, ++o) { qt[o] = a.children[k]; qs[o] = b.children[m]; }
        }
    }
}

// CSR of sorted keys: rows by key / nb, nodes = key % nb
__global__ void k_part_csr(const int64_t* __restrict__ keys, int64_t n, int64_t nb, int64_t* __restrict__ ptr,
                           int64_t nrows, int64_t* __restrict__ nodes) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int64_t r = keys[i] / nb;
    nodes[i] = keys[i] % nb + 1;
    const int64_t rp = i == 0 ? -1 : keys[i - 1] / nb;
    for (int64_t q = rp + 1; q <= r; ++q) ptr[q] = i + 1;  // rows rp+1 .. r start at i
    if (i == n - 1)
        for (int64_t q = r + 1; q <= nrows; ++q) ptr[q] = n + 1;
}

__global__ void k_part_empty(int64_t* __restrict__ ptr, int64_t nrows) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i <= nrows) ptr[i] = 1;
}

inline unsigned pblocks(int64_t n) { return (unsigned)((n + 255) / 256); }

}  // namespace nca

struct nca_partition {
    std::vector<int64_t> nearptr, nearnodes, farptr, farnodes;  // 1-based
};

using namespace nca;

extern "C" {

int32_t nca_interactions(nca_ctx* ctx, int64_t nn_a, const int64_t* childptr_a, const int64_t* children_a,
                         const double* center_a, const double* radius_a, int64_t nn_b, const int64_t* childptr_b,
                         const int64_t* children_b, const double* center_b, const double* radius_b,
                         int32_t wideband, const double* params, int64_t* nnear, int64_t* nfar,
                         nca_partition** out) {
    return guarded([&] {
        NCA_REQUIRE(ctx && params && out && nn_a > 0 && nn_b > 0 && childptr_a && childptr_b && center_a &&
                        center_b && radius_a && radius_b,
                    NCA_ERR_ARGUMENT, "nca_interactions: invalid argument");
        NCA_REQUIRE(nn_a < (1LL << 31) && nn_b < (1LL << 31), NCA_ERR_LIMIT, "nca_interactions: too many nodes");
        set_device(ctx);
        hipStream_t st = ctx->stream;
        // trees (1-based child CSR in, 0-based on the device)
        DevBuf<int32_t> cpa, cha, cpb, chb;
        DevBuf<double> cea, raa, ceb, rab;
        auto tree_up = [&](int64_t nn, const int64_t* cp, const int64_t* ch, const double* ce, const double* ra,
                           DevBuf<int32_t>& dcp, DevBuf<int32_t>& dch, DevBuf<double>& dce, DevBuf<double>& dra) {
            std::vector<int32_t> p(nn + 1), c(std::max<int64_t>(cp[nn] - cp[0], 1));
            for (int64_t i = 0; i <= nn; ++i) p[i] = (int32_t)(cp[i] - cp[0]);
            for (int64_t k = 0; k < cp[nn] - cp[0]; ++k) {
                NCA_REQUIRE(ch[k] >= 1 && ch[k] <= nn, NCA_ERR_ARGUMENT, "nca_interactions: child out of range");
                c[k] = (int32_t)(ch[k] - 1);
            }
            dcp.upload(p, st);
            dch.upload(c, st);
            dce.upload(ce, (size_t)(3 * nn), st);
            dra.upload(ra, (size_t)nn, st);
        };
        tree_up(nn_a, childptr_a, children_a, center_a, radius_a, cpa, cha, cea, raa);
        tree_up(nn_b, childptr_b, children_b, center_b, radius_b, cpb, chb, ceb, rab);
        PartTree A{cpa.p, cha.p, cea.p, raa.p}, B{cpb.p, chb.p, ceb.p, rab.p};
        PartCrit crit{wideband, params[0], params[1], params[2], params[3], params[4]};
        // frontier
        DevBuf<int32_t> ft, fs, nt, ns_, kf, kn, kc;
        DevBuf<int64_t> of, on, oc, farkeys, nearkeys;
        DevBuf<char> tmp;
        std::vector<int32_t> zero(1, 0);
        ft.upload(zero, st);
        fs.upload(zero, st);
        int64_t n = 1, nfar_tot = 0, nnear_tot = 0;
        // scan helper: exclusive sum of n int32 into int64, returns the total
        auto scan = [&](const int32_t* in, DevBuf<int64_t>& outbuf, int64_t cnt) -> int64_t {
            outbuf.alloc((size_t)cnt + 1);
            size_t bytes = 0;
            NCA_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, bytes, in, outbuf.p, (int)cnt + 1, st));
            if (tmp.n < bytes) tmp.alloc(bytes);
            // the element past the end reads kc/kf/kn padding (0), so outbuf[cnt] is the total
            NCA_HIP(hipcub::DeviceScan::ExclusiveSum(tmp.p, bytes, in, outbuf.p, (int)cnt + 1, st));
            int64_t total = 0;
            NCA_HIP(hipMemcpyAsync(&total, outbuf.p + cnt, sizeof(int64_t), hipMemcpyDeviceToHost, st));
            NCA_HIP(hipStreamSynchronize(st));
            return total;
        };
        std::vector<int64_t> fkeys_host_sizes;
        while (n > 0) {
            NCA_REQUIRE(n < (1LL << 31) - 1, NCA_ERR_LIMIT, "nca_interactions: frontier above 2^31 pairs");
            kf.alloc((size_t)n + 1);
            kn.alloc((size_t)n + 1);
            kc.alloc((size_t)n + 1);
            hipLaunchKernelGGL(k_part_classify, dim3(pblocks(n)), dim3(256), 0, st, ft.p, fs.p, n, A, B, crit, kf.p,
                               kn.p, kc.p);
            NCA_HIP(hipMemsetAsync(kf.p + n, 0, sizeof(int32_t), st));
            NCA_HIP(hipMemsetAsync(kn.p + n, 0, sizeof(int32_t), st));
            NCA_HIP(hipMemsetAsync(kc.p + n, 0, sizeof(int32_t), st));
            const int64_t af = scan(kf.p, of, n), an = scan(kn.p, on, n), ac = scan(kc.p, oc, n);
            // grow the key arrays (copying what they hold)
            auto grow = [&](DevBuf<int64_t>& buf, int64_t have, int64_t need) {
                if ((int64_t)buf.n >= need) return;
                DevBuf<int64_t> nb_;
                nb_.alloc((size_t)std::max<int64_t>(need, 2 * (int64_t)buf.n));
                if (have > 0)
                    NCA_HIP(hipMemcpyAsync(nb_.p, buf.p, sizeof(int64_t) * have, hipMemcpyDeviceToDevice, st));
                buf = std::move(nb_);
            };
            grow(farkeys, nfar_tot, nfar_tot + af);
            grow(nearkeys, nnear_tot, nnear_tot + an);
            nt.alloc((size_t)std::max<int64_t>(ac, 1));
            ns_.alloc((size_t)std::max<int64_t>(ac, 1));
            hipLaunchKernelGGL(k_part_emit, dim3(pblocks(n)), dim3(256), 0, st, ft.p, fs.p, n, A, B, kf.p, kn.p,
                               kc.p, of.p, on.p, oc.p, nn_b, nfar_tot, nnear_tot, farkeys.p, nearkeys.p, nt.p, ns_.p);
            NCA_HIP(hipGetLastError());
            nfar_tot += af;
            nnear_tot += an;
            std::swap(ft, nt);
            std::swap(fs, ns_);
            n = ac;
        }
        // sort the keys and build both CSRs
        auto csr = [&](DevBuf<int64_t>& keys, int64_t cnt, std::vector<int64_t>& ptr, std::vector<int64_t>& nodes) {
            ptr.assign(nn_a + 1, 1);
            nodes.assign(cnt, 0);
            if (cnt == 0) return;
            DevBuf<int64_t> sorted, dptr, dnodes;
            sorted.alloc((size_t)cnt);
            size_t bytes = 0;
            NCA_HIP(hipcub::DeviceRadixSort::SortKeys(nullptr, bytes, keys.p, sorted.p, (int)cnt, 0, 64, st));
            if (tmp.n < bytes) tmp.alloc(bytes);
            NCA_HIP(hipcub::DeviceRadixSort::SortKeys(tmp.p, bytes, keys.p, sorted.p, (int)cnt, 0, 64, st));
            dptr.alloc((size_t)nn_a + 1);
            dnodes.alloc((size_t)cnt);
            hipLaunchKernelGGL(k_part_empty, dim3(pblocks(nn_a + 1)), dim3(256), 0, st, dptr.p, nn_a);
            hipLaunchKernelGGL(k_part_csr, dim3(pblocks(cnt)), dim3(256), 0, st, sorted.p, cnt, nn_b, dptr.p, nn_a,
                               dnodes.p);
            NCA_HIP(hipGetLastError());
            NCA_HIP(hipMemcpyAsync(ptr.data(), dptr.p, sizeof(int64_t) * (nn_a + 1), hipMemcpyDeviceToHost, st));
            NCA_HIP(hipMemcpyAsync(nodes.data(), dnodes.p, sizeof(int64_t) * cnt, hipMemcpyDeviceToHost, st));
            NCA_HIP(hipStreamSynchronize(st));
        };
        auto p = std::make_unique<nca_partition>();
        csr(farkeys, nfar_tot, p->farptr, p->farnodes);
        csr(nearkeys, nnear_tot, p->nearptr, p->nearnodes);
        if (nnear) *nnear = nnear_tot;
        if (nfar) *nfar = nfar_tot;
        *out = p.release();
    });
}

int32_t nca_partition_export(const nca_partition* p, int64_t* nearptr, int64_t* nearnodes, int64_t* farptr,
                             int64_t* farnodes) {
    return guarded([&] {
        NCA_REQUIRE(p, NCA_ERR_ARGUMENT, "nca_partition_export: null handle");
        if (nearptr) std::copy(p->nearptr.begin(), p->nearptr.end(), nearptr);
        if (nearnodes) std::copy(p->nearnodes.begin(), p->nearnodes.end(), nearnodes);
        if (farptr) std::copy(p->farptr.begin(), p->farptr.end(), farptr);
        if (farnodes) std::copy(p->farnodes.begin(), p->farnodes.end(), farnodes);
    });
}

int32_t nca_partition_destroy(nca_partition* p) {
    return guarded([&] { delete p; });
}

}  // extern "C"
