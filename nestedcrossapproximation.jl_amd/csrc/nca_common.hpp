// Internal runtime pieces shared by the translation units of libnca_hip.so.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "../../include/nca_hip.h"

namespace nca {

struct Error : std::runtime_error {
    int32_t code;
    Error(int32_t c, const std::string& m) : std::runtime_error(m), code(c) {}
};

void set_last_error(const std::string& msg);

#define NCA_HIP(call)                                                                       \
    do {                                                                                    \
        hipError_t e_ = (call);                                                             \
        if (e_ != hipSuccess)                                                               \
            throw ::nca::Error(NCA_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
    } while (0)

#define NCA_REQUIRE(cond, code, msg)                         \
    do {                                                     \
        if (!(cond)) throw ::nca::Error((code), (msg));      \
    } while (0)

template <class F>
int32_t guarded(F&& f) {
    try {
        f();
        return NCA_OK;
    } catch (const Error& e) {
        set_last_error(e.what());
        return e.code;
    } catch (const std::exception& e) {
        set_last_error(e.what());
        return NCA_ERR_HIP;
    } catch (...) {
        set_last_error("unknown error");
        return NCA_ERR_HIP;
    }
}

// Owning device allocation.
template <class T>
struct DevBuf {
    T* p = nullptr;
    size_t n = 0;
    DevBuf() = default;
    explicit DevBuf(size_t count) { alloc(count); }
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    DevBuf(DevBuf&& o) noexcept : p(o.p), n(o.n) { o.p = nullptr; o.n = 0; }
    DevBuf& operator=(DevBuf&& o) noexcept {
        if (this != &o) { release(); p = o.p; n = o.n; o.p = nullptr; o.n = 0; }
        return *this;
    }
    ~DevBuf() { release(); }
    void alloc(size_t count) {
        release();
        n = count;
        if (count) NCA_HIP(hipMalloc(&p, count * sizeof(T)));
    }
    // Grow only: keeps the allocation when it already holds count elements.
    void reserve(size_t count) {
        if (count > n || !p) alloc(count);
    }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
    // Set-up copies: complete before returning, so the host source may be a
    // temporary and later null-stream reads see the data.
    void upload(const T* h, size_t count, hipStream_t s) {
        if (count > n) alloc(count);
        if (count) {
            NCA_HIP(hipMemcpyAsync(p, h, count * sizeof(T), hipMemcpyHostToDevice, s));
            NCA_HIP(hipStreamSynchronize(s));
        }
    }
    void upload(const std::vector<T>& v, hipStream_t s) { upload(v.data(), v.size(), s); }
    size_t bytes() const { return n * sizeof(T); }
};

}  // namespace nca

struct nca_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
};

struct nca_points {
    nca_ctx* ctx = nullptr;
    int64_t n = 0;
    nca::DevBuf<double> xyz;  // 3 x n, point-major (x, y, z per point)
    std::vector<double> host; // kept for host-side bookkeeping
};

struct nca_stack {
    nca_ctx* ctx = nullptr;
    int64_t nb = 0, nrows = 0, maxrank = 0;
    std::vector<int64_t> rowoffsets;  // nb+1, 0-based cumulative
    std::vector<int64_t> ncols;       // nb
    nca::DevBuf<double2> buffer;      // nrows x maxrank, column-major
};

struct nca_coupling {
    nca_ctx* ctx = nullptr;
    int64_t ntest = 0, ntrial = 0, nblocks = 0, nentries = 0;
    std::vector<int64_t> ptr, nodes;  // 0-based CSR (test node -> trial nodes)
    std::vector<int64_t> offsets;     // nblocks+1, 0-based complex offsets
    std::vector<int32_t> testrank, trialrank;
    // the pivots (0-based point indices) of every node of both sides, 0-based
    // CSR: block (t, s) = A[tpiv[tptr[t] ..], spiv[sptr[s] ..]]
    std::vector<int64_t> tptr, sptr;
    std::vector<int32_t> tpiv, spiv;
    nca::DevBuf<double2> blocks;
};

// Skeleton factors X kept on the device, indexed by block id (a node, or a
// direction index): one device chunk per compressed level, each holding that
// level's factors in stack order.
struct nca_factors {
    nca_ctx* ctx = nullptr;
    int64_t nids = 0;
    std::vector<int32_t> chunk, m, rank;  // per id; chunk -1: no factor
    std::vector<int64_t> off;             // per id: complex offset in its chunk
    std::vector<nca::DevBuf<double2>> chunks;
    std::vector<int64_t> chunk_size;      // complex entries of the factors in a chunk
    // chunks that are not packed (the factors at bounded offsets): their ids
    // in download order; empty for packed chunks
    std::vector<std::vector<int32_t>> chunk_ids;
};

namespace nca {
inline void set_device(const nca_ctx* ctx) { NCA_HIP(hipSetDevice(ctx->device)); }

// One block copy on the device: element (a, k) of the rows x cols source
// (column-major, leading dimension ld) goes to dst[dst + a * sa + k * sk].
struct PackJob {
    const double2* src;
    int64_t dst;
    int32_t ld, rows, cols, sa, sk, pad;
};

// Runs the jobs into dst on stream st (the jobs are uploaded first).
void pack_blocks(const std::vector<PackJob>& jobs, double2* dst, hipStream_t st);

// Device address, rows and rank of the factor of block id (0-based).
struct FactorRef {
    const double2* p;
    int64_t m, r;
};
inline FactorRef factor_of(const nca_factors* f, int64_t id) {
    NCA_REQUIRE(id >= 0 && id < f->nids && f->chunk[id] >= 0, NCA_ERR_ARGUMENT,
                "resident factors: block " + std::to_string(id + 1) + " has no factor");
    return FactorRef{f->chunks[f->chunk[id]].p + f->off[id], f->m[id], f->rank[id]};
}
}  // namespace nca
