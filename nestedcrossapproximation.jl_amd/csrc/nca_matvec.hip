// The H^2 matrix-vector product on gfx950 (src/nca/petrovgalerkinnca.jl:431-589):
//   gather   xs = x[perm_src]                         (source points in tree order)
//   stage 1  xh_s = B_s xs_s          per source leaf  (aggregate.jl:1-19)
//   stage 2  xh_p += E_c xh_c         per level, deepest first (aggregate.jl:44-67)
//   stage 3  yh_t += S_ts xh_s        coupling (couple.jl:1-24; reverse :119-142)
//   stage 4  yh_c += E_c yh_p         per level, root first (disaggregate.jl:1-23)
//   stage 5  y[perm_dst] = alpha (B_t yh_t + near) + beta y   per destination leaf
//            (disaggregate.jl:62-80 fused with the near field, which is evaluated
//            on the fly from the point coordinates instead of being stored).
// op T/C swap the two sides and use every operator transposed (conjugated).
// Coefficient vectors hold the R right-hand sides interleaved: xh[k * R + q].

#include <algorithm>
#include <atomic>
#include <chrono>
#include <thread>
#include <cstdio>
#include <type_traits>
#include <cstdlib>
#include <memory>
#include <mutex>

#include "nca_common.hpp"
#include "nca_math.hpp"

namespace nca {

constexpr int kWave = 64;

struct SideDev {
    // nn cluster nodes; ne coefficient blocks: the nodes themselves, or the
    // direction indices (cluster, direction) of a directional operator
    int64_t npts = 0, nn = 0, ne = 0, K = 0;
    DevBuf<int32_t> perm, iperm, lo, hi, cptr;  // iperm: original index -> tree position; cptr: ne+1
    // per node the first coefficient of its blocks (nn+1); empty when ne == nn
    DevBuf<int32_t> ncptr;
    const int32_t* node_cptr() const { return ncptr.p ? ncptr.p : cptr.p; }
    DevBuf<double> pts;  // coordinates in tree order, 3 per position
    DevBuf<double> pts_h16;  // the same times t (the MIXED f16 near field; made with the c32 copies)
    // leaf bases
    int64_t nbasis = 0;
    DevBuf<int32_t> bnodes;
    DevBuf<int64_t> boff;
    DevBuf<double2> bblocks;
    // leaf projection tiles: (basis index, first row) pairs, 16 rows each
    int64_t nptile = 0;
    DevBuf<int32_t> ptile;
    // transfers, grouped by level: parents [levptr[l], levptr[l+1]) of parents array
    int64_t nlev = 0;
    std::vector<int64_t> levptr;
    DevBuf<int32_t> parents, chds;
    DevBuf<int64_t> chdptr, toff;
    DevBuf<double2> tblocks;
    // downward transfers by child when a child block has several parents
    // (directions): per level [dlevptr[l], dlevptr[l+1]) of dchild, each with
    // its parents dpar[dptr[i], dptr[i+1]) and their transfer blocks dtoff
    bool down_by_child = false;
    std::vector<int64_t> dlevptr;
    DevBuf<int32_t> dchild, dpar;
    DevBuf<int64_t> dptr, dtoff;
    // per block: 1 if it has coefficients and no downward transfer writes it
    // (roots, children of parents without coefficients)
    DevBuf<uint8_t> orph;
    // c32 copies of the bases and transfers for the MIXED path
    DevBuf<float2> bblocks32, tblocks32;
    // all leaves of the tree (destinations of stage 5) and their basis slot (-1: none)
    int64_t nleaves = 0;
    DevBuf<int32_t> leaves, leafbasis;
    // basis slots of nodes that are not leaves (directional bases above
    // low-frequency children), applied after the leaves
    int64_t nibasis = 0;
    DevBuf<int32_t> ibslot;
    // near lists by node of this side (as destination): ptr (nn+1), nodes of the other side
    DevBuf<int64_t> nearptr;
    DevBuf<int32_t> nearnodes;
    // the same near lists as merged contiguous point ranges of the other side;
    // bit 31 of nrhi: the range may hold a point at distance 0 from the node's
    DevBuf<int64_t> nrptr;
    DevBuf<int32_t> nrlo, nrhi;
    int64_t max_range = 0;  // points of the longest merged range
    // 16-row tiles of the leaves (leaf index, first row) for the MFMA expansion kernel
    int64_t ntiles = 0;
    DevBuf<int32_t> tile_leaf, tile_row0;
    // MIXED f16 forms: per node, 1 when every near-field tile of the leaf
    // (near_rel) or every coupling tile of the block (coup_rel) with this
    // side as destination meets the rel32 condition
    DevBuf<uint8_t> near_rel, coup_rel;
    // host copies used during creation
    std::vector<int32_t> h_lo, h_hi, h_cptr, h_perm, h_bnodes, h_leaves, h_leafslot, h_leaves_node_order;
    std::vector<int32_t> h_ncptr, h_enode;  // per node first coefficient (nn+1); node of every block
    // host copies of the merged near ranges (to reorder the leaves)
    std::vector<int64_t> h_nrptr;
    std::vector<int32_t> h_nrlo, h_nrhi;
    // per node: tight bounding box of its points (x, y, z minima, then maxima)
    std::vector<double> h_box;
    int64_t basis_entries = 0, transfer_entries = 0;
};

}  // namespace nca

struct nca_h2 {
    nca_ctx* ctx = nullptr;
    double k = 0.0;
    int sign = 1;
    // the largest phase |k| r / (2 pi) over the near pairs (from the point boxes)
    double near_turns = 0.0;
    // the largest distance of a near pair from the point boxes (bounds the
    // phase) and, with the c32 copies, the measured smallest (above the
    // zero-distance guard) and largest
    double near_rmax = 0.0, near_rmin = 0.0, near_rmax_pairs = 0.0;
    // MIXED near field on f16 MFMA: 1 when the near pairs allow it (distances
    // within 2^17 of each other), with the point scale h16_t = 2^-e
    int h16 = 0;
    double h16_t = 1.0;
    // MIXED coupling evaluated on the fly on the f16 MFMA (couple_k3 = 2): 1
    // when available (far pivot pairs spanning at most 2^17 in distance); the
    // pivot scale, and the measured smallest and largest far pivot distances
    int coup_h16 = 0;
    double coup_t = 1.0, coup_rmin = 0.0, coup_rmax = 0.0;
    // f16 near field / coupling with FP32 distances from tile-relative
    // coordinates (NearK::rel32): allowed by the measured tile statistics,
    // and used unless rel32_off (nca_h2_rel32)
    int rel32_off = 0;
    // the fraction of near-field / coupling tiles of both directions that
    // meet the rel32 condition
    double near_rel_frac = 0.0, coup_rel_frac = 0.0;
    // scaled pivot coordinates of both sides in coefficient order (3 per coefficient)
    nca::DevBuf<double> piv_test, piv_trial;
    // the far blocks' coefficients as merged ranges (bit 31 of the end: the
    // zero-distance guard): per test block (trial coefficients) and per trial
    // block (test coefficients), and the 16-row tiles of the trial blocks
    nca::DevBuf<int64_t> cf_ptr, cr_ptr;
    nca::DevBuf<int32_t> cf_lo, cf_hi, cr_lo, cr_hi;
    int64_t cr_ntiles = 0;
    nca::DevBuf<int32_t> cr_tile_node, cr_tile_row0;
    int precision = NCA_PREC_FP64;
    int64_t M = 0, N = 0;
    nca::SideDev test, trial;
    const nca_coupling* coupling = nullptr;
    // forward coupling: per test node entry base and column base, per column the source coefficient
    nca::DevBuf<int64_t> f_base, f_colbase;
    nca::DevBuf<int32_t> f_ncols, f_colcoef;
    // 16-row tiles of the forward coupling (test node, first row)
    int64_t f_ntiles = 0;
    nca::DevBuf<int32_t> f_tile_node, f_tile_row0;
    // reverse coupling: per trial node list of (test node, block offset)
    nca::DevBuf<int64_t> r_ptr, r_off;
    nca::DevBuf<int32_t> r_test;
    // transposed coupling for T/C products, made on the first one: per trial
    // block its r_s x C' row of the blocks S_ts^T (column-major, contiguous),
    // in the forward plan's format, so the forward kernels apply it
    bool rv_plan = false;
    int64_t rv_ntiles = 0;
    nca::DevBuf<int64_t> rv_base, rv_colbase;
    nca::DevBuf<int32_t> rv_ncols, rv_colcoef, rv_tile_node, rv_tile_row0;
    nca::DevBuf<double2> rv_blocks;   // FP64 (made for the first FP64 T/C product)
    nca::DevBuf<float2> rv_blocks32;  // c32 (made for the first MIXED T/C product)
    nca::DevBuf<int64_t> rv_dst;  // per block of the reverse plan (r_off order): offset in the transposed store
    nca::DevBuf<int32_t> rv_rs;   // per block of the reverse plan: its trial rank
    // stored near field (nca_h2_store_near): forward products read the near
    // entries from HBM instead of evaluating them. One "row" per test leaf (its
    // points in tree order) over the concatenated points of its merged near
    // ranges, in the forward coupling format, so the forward coupling kernels
    // apply it into a per-point buffer that the leaf kernels add.
    bool near_stored = false, ns_plan = false;
    int64_t ns_ntiles = 0, ns_entries = 0, ns_nleaf = 0;
    nca::DevBuf<int64_t> ns_base, ns_colbase;
    nca::DevBuf<int32_t> ns_ncols, ns_colcoef, ns_tile_node, ns_tile_row0, ns_cptr;
    nca::DevBuf<double2> ns_blocks;   // FP64 entries (made for the first FP64 product)
    nca::DevBuf<float2> ns_blocks32;  // c32 entries (made for the first MIXED product)
    nca::DevBuf<double2> ws_ynear;    // near-field result per point (tree order) x R
    // workspace
    int64_t ws_rhs = 0;
    // run a last chunk of 1..3 right-hand sides on the MFMA kernels too: -1
    // by precision (MIXED yes: its f16 near field costs the same at one column
    // and beats the VALU kernel; FP64 no: its MFMA near field costs twice the
    // VALU kernel's time at one column), 0 / 1 set by NCA_MFMA_SMALL
    int mfma_small = -1;
    // MIXED coupling with more than 16 right-hand sides reads a compact copy
    // of the chunk's source coefficients (worth it when the coupling gathers
    // every coefficient several times)
    int couple_x32 = 1;
    // ws_xh: source coefficients in the operator's element type (c32 in
    // MIXED, in the same allocation); ws_yh: destination coefficients, FP64;
    // ws_yc: the c32 coupling results of MIXED forward products
    nca::DevBuf<double2> ws_xs, ws_xh, ws_yh, ws_xin, ws_yio;
    nca::DevBuf<float2> ws_xs32, ws_xh32, ws_yc;
    // the f16 near field's split right-hand sides, the gather's per-workgroup
    // column maxima and the column scales (s_q, 1 / s_q)
    nca::DevBuf<uint4> ws_xs16;
    nca::DevBuf<float> ws_xpart, ws_xsc;
    // the on-the-fly coupling's split source coefficients (K x R), the column
    // maxima of a chunk and its scales (up to 32 columns)
    nca::DevBuf<uint4> ws_xh16;
    nca::DevBuf<float> ws_xhpart, ws_xhsc;
    // The device workspace the stage kernels of the product in flight use:
    // the operator's own buffers above, or carved from the static scratch
    // array of nca_h2_matvec_device_noalloc.
    struct WsView {
        double2* xs = nullptr;    // gathered right-hand sides, FP64
        float2* xs32 = nullptr;   // gathered right-hand sides, c32 (MIXED)
        double2* xh = nullptr;    // source coefficients (c32 in MIXED, same bytes)
        double2* yh = nullptr;    // destination coefficients, FP64
        float2* yc = nullptr;     // c32 coupling results (MIXED forward)
        float2* xh32 = nullptr;   // compact c32 copy of a 16-column chunk's source coefficients
        double2* ynear = nullptr; // stored near field applied, per point (tree order), FP64
        uint4* xs16 = nullptr;    // split right-hand sides of the f16 near field (MIXED)
        float* xpart = nullptr;   // per gather workgroup and column: largest |re|, |im|
        float* xsc = nullptr;     // per column: s_q, 1 / s_q
        uint4* xh16 = nullptr;    // split source coefficients of the on-the-fly coupling
        float* xhpart = nullptr;  // their per-workgroup column maxima (a chunk of up to 32 columns)
        float* xhsc = nullptr;    // the chunk's column scales
    } ws;
    // c32 copy of the coupling blocks for the MIXED path
    nca::DevBuf<float2> S32;
    bool mixed_ready = false;
    // (cos, sin)(n pi / 128), n = 0..255, for the FP64 near field
    nca::DevBuf<double2> sctab;
    nca_h2_stats stats{};
    // Launch choices, set by nca_h2_autotune (defaults measured on the
    // 1M-point sphere with 16 RHS; environment overrides at creation):
    // near_k3 / couple_k3 [precision]: complex products on MFMA with three
    //   real products (Karatsuba) instead of four (NCA_MFMA_K3, NCA_MIXED_K3);
    //   in FP32 three raise the MIXED error against FP64 from 1.0e-6 to 1.5e-6;
    //   near_k3 = 2 (MIXED, the default): the near field on f16 MFMA with
    //   split operands when the operator allows it (h16), else as 1;
    //   couple_k3 = 2 (MIXED, the default): the coupling blocks evaluated on
    //   the fly on the f16 path when the operator allows it (coup_h16), else as 1;
    // near_xcd / couple_xcd: work_block order (NCA_ORDER_NEAR, NCA_ORDER_COUPLE);
    // leaf_order: order_leaves mode (NCA_LEAF_ORDER);
    // couple_order: order_coupling_tiles mode (NCA_COUPLE_ORDER);
    // near_stagger: near-field walks start at a tile-dependent range (NCA_NEAR_STAGGER).
    int near_k3[2] = {1, 2}, couple_k3[2] = {1, 2};
    int near_xcd = 0, couple_xcd = 0, leaf_order = 2, couple_order = 3, near_stagger = 0;
    std::vector<int32_t> h_ncols;  // host copy: columns of every test node's coupling row
    // Every entry point that touches the workspace, the launch plans or the
    // precision holds this lock, so products of several host threads over one
    // operator run one after the other (each in full) on the context stream.
    mutable std::mutex mu;
    // Bumped whenever a device buffer a captured sweep points to is freed or
    // reallocated, or a launch plan changes; a graph captured at another
    // generation refuses to launch.
    uint64_t generation = 0;
};

// Rows of padding after the gathered right-hand sides and the tree-ordered
// points of each side (read, never used, by the last near-field step of a range).
constexpr int64_t kNearPad = 16;

// One captured device sweep (see nca_h2_graph_create).
struct nca_h2_graph {
    nca_h2* h2 = nullptr;
    int precision = NCA_PREC_FP64;
    uint64_t generation = 0;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
};

namespace nca {

// ---------------------------------------------------------------------------
// Storage types: operator blocks and the gathered right-hand sides are double2
// (FP64) or float2 (MIXED); coefficient vectors and outputs are always double2.
__device__ __forceinline__ double2 to_d2(double2 v) { return v; }
__device__ __forceinline__ double2 to_d2(float2 v) { return make_double2(v.x, v.y); }
__device__ __forceinline__ void store_as(double2* p, double2 v) { *p = v; }
__device__ __forceinline__ void store_as(float2* p, double2 v) { *p = make_float2((float)v.x, (float)v.y); }

template <bool CJ, class T2>
__device__ __forceinline__ double2 ldc(const T2* p, int64_t i) {
    double2 v = to_d2(p[i]);
    if (CJ) v.y = -v.y;
    return v;
}

__global__ void k_fill(double2* __restrict__ p, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = make_double2(0.0, 0.0);
}

__global__ void k_to_c32(const double2* __restrict__ in, float2* __restrict__ out, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = make_float2((float)in[i].x, (float)in[i].y);
}

// Transposed coupling blocks, one wave per block b of the reverse plan: the
// r_t x r_s block at S + src[b] (column-major) goes to ST + dst[b] as r_s x r_t
// (column-major); r_t from the test block's coefficient range, r_s given.
template <class T2>
__global__ void __launch_bounds__(256)
    k_transpose_blocks(const T2* __restrict__ S, T2* __restrict__ ST, const int64_t* __restrict__ src,
                       const int64_t* __restrict__ dst, const int32_t* __restrict__ rtest,
                       const int32_t* __restrict__ cptr_test, const int32_t* __restrict__ rs, int64_t nb) {
    const int64_t b = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    if (b >= nb) return;
    const int t = rtest[b];
    const int rt = cptr_test[t + 1] - cptr_test[t], c = rs[b];
    const T2* in = S + src[b];
    T2* out = ST + dst[b];
    for (int d = threadIdx.x & 63; d < rt * c; d += 64) {
        const int k = d % c, a = d / c;
        out[d] = in[a + (int64_t)k * rt];
    }
}

// out[k*16 + q] = c32(in[k*R + q0 + q]), q < 16: one 16-column chunk of a
// coefficient vector, compacted.
template <class IT>
__global__ void k_chunk_to_c32(const IT* __restrict__ in, int64_t R, int64_t q0, int64_t K, int nq,
                               float2* __restrict__ out) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= K * 16) return;
    const int q = (int)(i & 15);
    const double2 v = q < nq ? to_d2(in[(i >> 4) * R + q0 + q]) : make_double2(0.0, 0.0);
    out[i] = make_float2((float)v.x, (float)v.y);
}

// Stored near field, one workgroup per test leaf i: its nt x C row of
// G(x_row, y_col) (column-major at base[i]); rows are its points (tree
// positions cptr[i] ..), columns the source positions colcoef[colbase[i] ..];
// the entry of helmholtz() (zero at r = 0), rounded to T2.
template <class T2>
__global__ void __launch_bounds__(256)
    k_near_store(const double* __restrict__ dpts, const double* __restrict__ spts, const int32_t* __restrict__ cptr,
                 const int64_t* __restrict__ base, const int64_t* __restrict__ colbase,
                 const int32_t* __restrict__ ncols, const int32_t* __restrict__ colcoef, double k, int sign,
                 T2* __restrict__ S) {
    const int64_t i = blockIdx.x;
    const int r0 = cptr[i], nt = cptr[i + 1] - r0, C = ncols[i];
    const int64_t b = base[i], cb = colbase[i];
    const int64_t total = (int64_t)nt * C;
    for (int64_t e = threadIdx.x; e < total; e += blockDim.x) {
        const int col = (int)(e / nt), row = (int)(e - (int64_t)col * nt);
        const double* x = dpts + 3 * (int64_t)(r0 + row);
        const double* y = spts + 3 * (int64_t)colcoef[cb + col];
        store_as(S + b + e, helmholtz(k, sign, x[0], x[1], x[2], y[0], y[1], y[2]));
    }
}

// xs[iperm[i]*R + q] = x[i + q*ldx], moved through LDS in tiles of 64
// consecutive input points x 16 columns: reads run along the points of one
// column, writes along the 16 columns of one point. xpart (optional): per
// workgroup row blockIdx.x and column q < R the largest |re|, |im| of the
// tile's values rounded to FP32, at xpart[blockIdx.x R + q].
template <class XT>
__global__ void __launch_bounds__(256)
    k_gather_tiled(const double2* __restrict__ x, int64_t ldx, const int32_t* __restrict__ iperm, int64_t npts,
                   int64_t R, int64_t Rw, XT* __restrict__ xs, float* __restrict__ xpart) {
    // R live columns, Rw >= R columns written (the padding ones as zeros)
    __shared__ double2 tile[16][65];
    const int t = threadIdx.x;
    const int64_t p0 = (int64_t)blockIdx.x * 64, c0 = (int64_t)blockIdx.y * 16;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        const int ql = (t >> 6) + 4 * r;
        const int64_t p = p0 + (t & 63), q = c0 + ql;
        tile[ql][t & 63] = (p < npts && q < R) ? x[p + q * ldx] : make_double2(0.0, 0.0);
    }
    __syncthreads();
    if (xpart) {
        // column t / 16 of the tile, points t % 16 + 16 r, then over the 16 lanes
        const int ql = t >> 4;
        float m = 0.0f;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const double2 v = tile[ql][(t & 15) + 16 * r];
            m = fmaxf(m, fmaxf(fabsf((float)v.x), fabsf((float)v.y)));
        }
#pragma unroll
        for (int o = 8; o > 0; o >>= 1) m = fmaxf(m, __shfl_xor(m, o, 16));
        if ((t & 15) == 0 && c0 + ql < R) xpart[(int64_t)blockIdx.x * R + c0 + ql] = m;
    }
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        const int pl = (t >> 4) + 16 * r;
        const int64_t p = p0 + pl, q = c0 + (t & 15);
        if (p < npts && q < Rw) store_as(xs + (int64_t)iperm[p] * Rw + q, tile[t & 15][pl]);
    }
}

// Stage 1: one wave per source leaf. TR: the block is the destination-side
// basis used transposed (U_t, |t| x r), else V_s (r x |s|).
template <int NR, bool TR, bool CJ, class BT, class XT>
__global__ void __launch_bounds__(kWave)
    k_project(const int32_t* __restrict__ bnodes, const int64_t* __restrict__ boff, const BT* __restrict__ blocks,
              const int32_t* __restrict__ lo, const int32_t* __restrict__ hi, const int32_t* __restrict__ cptr,
              const XT* __restrict__ xs, int64_t R, int64_t q0, BT* __restrict__ xh) {
    const int i = blockIdx.x;
    const int s = bnodes[i];
    const int r = cptr[s + 1] - cptr[s];
    const int l0 = lo[s], ns = hi[s] - l0;
    const BT* B = blocks + boff[i];
    for (int a = threadIdx.x; a < r; a += kWave) {
        double2 acc[NR];
#pragma unroll
        for (int q = 0; q < NR; ++q) acc[q] = make_double2(0.0, 0.0);
        for (int j = 0; j < ns; ++j) {
            const double2 b = ldc<CJ>(B, TR ? (int64_t)j + (int64_t)a * ns : (int64_t)a + (int64_t)j * r);
            const XT* xv = xs + (int64_t)(l0 + j) * R + q0;
#pragma unroll
            for (int q = 0; q < NR; ++q) acc[q] = cfma(b, to_d2(xv[q]), acc[q]);
        }
        BT* out = xh + (int64_t)(cptr[s] + a) * R + q0;
#pragma unroll
        for (int q = 0; q < NR; ++q) store_as(out + q, cadd(to_d2(out[q]), acc[q]));
    }
}

// Stage 2 (TR = false: E (r_p x r_c), forward trial transfers; TR = true: E
// (r_c x r_p) used transposed) and stage 4 (DOWN: children accumulate from the
// parent; TR = false: E (r_c x r_p), TR = true: E (r_p x r_c) transposed).
// One wave per parent of the level; a parent's edges run on the same lanes.
template <int NR, bool DOWN, bool TR, bool CJ, class BT, class CT>
__global__ void __launch_bounds__(kWave)
    k_transfer(const int32_t* __restrict__ parents, const int64_t* __restrict__ chdptr,
               const int32_t* __restrict__ chds, const int64_t* __restrict__ toff, const BT* __restrict__ blocks,
               const int32_t* __restrict__ cptr, int64_t pbase, int64_t R, int64_t q0, CT* __restrict__ h) {
    const int64_t pi = pbase + blockIdx.x;
    const int p = parents[pi];
    const int rp = cptr[p + 1] - cptr[p];
    for (int64_t e = chdptr[pi]; e < chdptr[pi + 1]; ++e) {
        const int c = chds[e];
        const int rc = cptr[c + 1] - cptr[c];
        const BT* E = blocks + toff[e];
        const int nout = DOWN ? rc : rp, nin = DOWN ? rp : rc;
        const int outbase = DOWN ? cptr[c] : cptr[p], inbase = DOWN ? cptr[p] : cptr[c];
        for (int a = threadIdx.x; a < nout; a += kWave) {
            double2 acc[NR];
#pragma unroll
            for (int q = 0; q < NR; ++q) acc[q] = make_double2(0.0, 0.0);
            for (int j = 0; j < nin; ++j) {
                const double2 b = ldc<CJ>(E, TR ? (int64_t)j + (int64_t)a * nin : (int64_t)a + (int64_t)j * nout);
                const CT* hv = h + (int64_t)(inbase + j) * R + q0;
#pragma unroll
                for (int q = 0; q < NR; ++q) acc[q] = cfma(b, to_d2(hv[q]), acc[q]);
            }
            CT* out = h + (int64_t)(outbase + a) * R + q0;
#pragma unroll
            for (int q = 0; q < NR; ++q) store_as(out + q, cadd(to_d2(out[q]), acc[q]));
        }
    }
}

// Stage 3 forward: one wave per test cluster t. The coupling blocks of its row
// form one r_t x C column-major matrix; lane = (row, column slot).
template <int NR, bool CJ, class ST>
__global__ void __launch_bounds__(kWave)
    k_couple_fwd(const ST* __restrict__ S, const int64_t* __restrict__ base, const int64_t* __restrict__ colbase,
                 const int32_t* __restrict__ ncols, const int32_t* __restrict__ colcoef,
                 const int32_t* __restrict__ cptr_dst, const ST* __restrict__ xh, int64_t R, int64_t q0,
                 double2* __restrict__ yh) {
    __shared__ double2 red[kWave];
    const int t = blockIdx.x;
    const int rt = cptr_dst[t + 1] - cptr_dst[t];
    const int C = ncols[t];
    if (rt == 0 || C == 0) return;
    const ST* St = S + base[t];
    const int32_t* cc = colcoef + colbase[t];
    const int lane = threadIdx.x;
    for (int r0 = 0; r0 < rt; r0 += kWave) {
        const int rows = min(rt - r0, kWave);
        const int ncs = kWave / rows;
        const int row = r0 + lane % rows, slot = lane / rows;
        double2 acc[NR];
#pragma unroll
        for (int q = 0; q < NR; ++q) acc[q] = make_double2(0.0, 0.0);
        if (slot < ncs) {
            // four columns in flight: their index, coefficient and operand loads
            // are independent of each other
#pragma unroll 4
            for (int col = slot; col < C; col += ncs) {
                const double2 sv = ldc<CJ>(St, (int64_t)row + (int64_t)col * rt);
                const ST* xv = xh + (int64_t)cc[col] * R + q0;
#pragma unroll
                for (int q = 0; q < NR; ++q) acc[q] = cfma(sv, to_d2(xv[q]), acc[q]);
            }
        }
#pragma unroll
        for (int q = 0; q < NR; ++q) {
            __syncthreads();
            red[lane] = (slot < ncs) ? acc[q] : make_double2(0.0, 0.0);
            __syncthreads();
            if (slot == 0) {
                double2 sum = red[lane];
                for (int c = 1; c < ncs; ++c) sum = cadd(sum, red[lane + c * rows]);
                double2* out = yh + (int64_t)(cptr_dst[t] + row) * R + q0 + q;
                *out = cadd(*out, sum);
            }
        }
    }
}

// Stage 3 reverse (T/C): one wave per trial cluster s: yh_s += sum_t S_ts^T xh_t,
// for the nq <= NR live columns q0 .. q0 + nq - 1.
template <int NR, bool CJ, class ST>
__global__ void __launch_bounds__(kWave)
    k_couple_rev(const ST* __restrict__ S, const int64_t* __restrict__ rptr, const int32_t* __restrict__ rtest,
                 const int64_t* __restrict__ roff, const int32_t* __restrict__ cptr_src,
                 const int32_t* __restrict__ cptr_dst, const ST* __restrict__ xh, int64_t R, int64_t q0,
                 int nq, double2* __restrict__ yh) {
    const int s = blockIdx.x;
    const int rs = cptr_dst[s + 1] - cptr_dst[s];
    if (rs == 0) return;
    for (int a = threadIdx.x; a < rs; a += kWave) {
        double2 acc[NR];
#pragma unroll
        for (int q = 0; q < NR; ++q) acc[q] = make_double2(0.0, 0.0);
        for (int64_t e = rptr[s]; e < rptr[s + 1]; ++e) {
            const int t = rtest[e];
            const int rt = cptr_src[t + 1] - cptr_src[t];
            const ST* Sb = S + roff[e] + (int64_t)a * rt;
            const ST* xv0 = xh + (int64_t)cptr_src[t] * R + q0;
            for (int j = 0; j < rt; ++j) {
                const double2 sv = ldc<CJ>(Sb, j);
#pragma unroll
                for (int q = 0; q < NR; ++q)
                    if (q < nq) acc[q] = cfma(sv, to_d2(xv0[(int64_t)j * R + q]), acc[q]);
            }
        }
        double2* out = yh + (int64_t)(cptr_dst[s] + a) * R + q0;
#pragma unroll
        for (int q = 0; q < NR; ++q)
            if (q < nq) out[q] = cadd(out[q], acc[q]);
    }
}

constexpr double kInvFourPi = 0x1.45f306dc9c883p-4;

// Near-field constants of one product: k, k / (2 pi) and k / (4 pi); phase32:
// every near pair's phase k r stays within kPhase32Turns revolutions (from
// the point boxes), so MIXED forms r and the phase in FP32 (k2pif = k / (2 pi)).
struct NearK {
    double k;
    double k2pi;
    double k4pi;
    float k2pif;
    int phase32;
    // f16 path (points scaled by t): the zero-distance guard's 1e-30 and
    // 2e-30 times t^2, and t / (4 pi)
    double gadd, gmin;
    float tsc;
    // f16 path with phase32: rel32 = distances in FP32 from coordinates taken
    // relative to each tile's first row, for the tiles whose node is flagged
    // (rows within kRel32Span times the tile's smallest pair distance of its
    // first row, no pair strictly between 0 and the guard); the guard's FP32
    // form then is r^2 + gaddf > gminf, both far below every scaled nonzero
    // distance
    int rel32;
    float gaddf, gminf;
};
constexpr double kPhase32Turns = 1.0;
// turns: the largest phase k r / (2 pi) over the operator's near pairs
inline NearK near_consts(double k, double turns) {
    return NearK{k,     k * (0.5 / M_PI), k * (0.25 / M_PI), (float)(k * (0.5 / M_PI)), turns <= kPhase32Turns ? 1 : 0,
                 1e-30, 2e-30,            (float)(0.25 / M_PI),       0,
                 0x1p-126f, 0x1p-100f};
}
// The same for points scaled by t (a power of two): k / t, guards times t^2.
inline NearK near_consts_scaled(double k, double turns, double t) {
    NearK nk = near_consts(k / t, turns);
    nk.gadd = 1e-30 * t * t;
    nk.gmin = 2e-30 * t * t;
    nk.tsc = (float)(t * (0.25 / M_PI));
    return nk;
}

// Table of (cos, sin)(n pi / 128), n = 0..255, for the FP64 near field.
constexpr int kSinCosTab = 256;
constexpr double kTabScale = 0x1.45f306dc9c883p+5;   // 128 / pi
constexpr double kPi128Hi = 0x1.921fb54442d18p-6;    // pi / 128, leading part
constexpr double kPi128Lo = 0x1.1a62633145c07p-60;   // pi / 128 - kPi128Hi

// 4 pi G(x, y) = (cos, sin)(k r) / r for the FP64 near field (the 1/(4 pi) is
// applied to the accumulated sum; NEG: the conjugate, for a negative phase
// sign or the adjoint), branch free, r = 0 gives 0 (GUARD; without it the
// source range is known to stay more than 1e-14 away): 1/r from the hardware
// reciprocal square root (~2^-23) refined by one Newton step (~2^-46); the phase x = k r split
// as x = n pi / 128 + y, |y| <= pi / 256 (two-part constant), then sin/cos x =
// table(n) rotated by sin/cos y from short Taylor polynomials (truncation below
// 5e-15).
template <bool NEG, bool GUARD>
__device__ __forceinline__ double2 helmholtz_mv(const NearK& nk, const double2* tab, double x0, double x1,
                                                double x2, double y0, double y1, double y2) {
    const double dx = x0 - y0, dy = x1 - y1, dz = x2 - y2;
    const double r2 = fma(dz, dz, fma(dy, dy, dx * dx));
    const double r2s = GUARD ? fmax(r2, 1e-300) : r2;
    double ri = __builtin_amdgcn_rsq(r2s);
    const double h = 0.5 * r2s;
    ri = ri * fma(-h * ri, ri, 1.5);
    const double x = nk.k * (r2s * ri);
    const double q = rint(x * kTabScale);
    double y = fma(-q, kPi128Hi, x);
    y = fma(-q, kPi128Lo, y);
    const double2 cs = tab[(int)q & (kSinCosTab - 1)];
    const double z = y * y;
    const double sy = fma(y * z, fma(z, 0x1.1111111111111p-7, -0x1.5555555555555p-3), y);
    const double cy = fma(z, fma(z, 0x1.5555555555555p-5, -0.5), 1.0);
    const double c = fma(cs.x, cy, -(cs.y * sy));
    const double s = fma(cs.y, cy, cs.x * sy);
    const double sc = !GUARD || r2 > 0.0 ? ri : 0.0;
    return make_double2(c * sc, NEG ? -(s * sc) : s * sc);
}

// 4 pi G(x, y) for the MIXED near field (NEG as above), branch free:
// coordinate differences, squared distance and the phase k r / (2 pi) in
// FP64; sin/cos of the phase reduced to [0, 1) revolutions (its fraction,
// exact in FP64, then rounded to FP32: 6e-8 revolutions at most) and the 1/r
// scale in FP32 (hardware sin/cos of 2 pi f, hardware reciprocal square
// root). GUARD (a source range that may hold a point at distance 0): the
// squared distance carries 1e-30 (below half an ulp of any r^2 >= 1e-12), so
// r = 0 needs no clamp, and r^2 <= 2e-30 counts as r = 0 and gives 0; other
// ranges are more than 1e-14 away and skip both. The phase takes r0 = r^2 rsq(r^2) plus one Newton correction
// (r^2 - r0^2) rsq(r^2) / 2, relative error ~1e-14. PH32 (every near pair
// within kPhase32Turns revolutions): r = r^2 rsq(r^2) and the phase k r / (2 pi)
// in FP32 instead, relative error < 4e-7, i.e. < 2.5e-6 radians at one
// revolution (five VALU operations fewer per entry).
template <bool NEG, bool GUARD, bool PH32>
__device__ __forceinline__ float2 helmholtz_f32(const NearK& nk, double x0, double x1, double x2, double y0,
                                                double y1, double y2) {
    const double dx = x0 - y0, dy = x1 - y1, dz = x2 - y2;
    const double r2 = fma(dz, dz, fma(dy, dy, GUARD ? fma(dx, dx, 1e-30) : dx * dx));
    const float r2f = (float)r2;
    const float rf = __builtin_amdgcn_rsqf(r2f);
    float f;
    if constexpr (PH32) {
        f = __builtin_amdgcn_fractf((r2f * rf) * nk.k2pif);
    } else {
        const double ri = (double)rf;
        const double r0 = r2 * ri;
        const double e = fma(-r0, r0, r2);
        const double t = fma(e, ri * nk.k4pi, r0 * nk.k2pi);
        f = (float)__builtin_amdgcn_fract(t);
    }
    const float c = __builtin_amdgcn_cosf(f), s = __builtin_amdgcn_sinf(f);
    const float sc = !GUARD || r2 > 2e-30 ? rf : 0.0f;
    return make_float2(c * sc, NEG ? -(s * sc) : s * sc);
}

// ---------------------------------------------------------------------------
// MFMA kernels (16 right-hand sides): a 16 x 4 operand tile (rows x sources or
// columns) times a 4 x 16 tile (sources x right-hand sides). Lane l supplies
// A[l % 16][l / 16] and B[l / 16][l % 16]; the 16 x 16 result layout differs
// per instruction on gfx950 (checked on the device):
//   v_mfma_f64_16x16x4f64: lane l holds D[l / 16 + 4 v][l % 16], v = 0..3
//   v_mfma_f32_16x16x4f32: lane l holds D[4 (l / 16) + v][l % 16]
typedef double dbl4 __attribute__((ext_vector_type(4)));
typedef float flt4 __attribute__((ext_vector_type(4)));

struct PrecF64 {
    using T = double;
    using T2 = double2;
    using V4 = dbl4;
    static constexpr bool kF64 = true;
    __device__ static V4 mma(T a, T b, V4 c) { return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0); }
    __device__ static int row(int kk, int v) { return kk + 4 * v; }
    __device__ static T2 cvt(double2 v) { return v; }
    __device__ static T2 cvt(float2 v) { return make_double2(v.x, v.y); }
    static constexpr bool kTab = true;
    template <bool NEG, bool GUARD, bool PH32>
    __device__ static T2 eval(const NearK& nk, const double2* tab, double x0, double x1, double x2, double y0,
                              double y1, double y2) {
        return helmholtz_mv<NEG, GUARD>(nk, tab, x0, x1, x2, y0, y1, y2);
    }

};

struct PrecF32 {
    using T = float;
    using T2 = float2;
    using V4 = flt4;
    static constexpr bool kF64 = false;
    __device__ static V4 mma(T a, T b, V4 c) { return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0); }
    __device__ static int row(int kk, int v) { return 4 * kk + v; }
    __device__ static T2 cvt(double2 v) { return make_float2((float)v.x, (float)v.y); }
    __device__ static T2 cvt(float2 v) { return v; }
    static constexpr bool kTab = false;
    template <bool NEG, bool GUARD, bool PH32>
    __device__ static T2 eval(const NearK& nk, const double2*, double x0, double x1, double x2, double y0, double y1,
                              double y2) {
        return helmholtz_f32<NEG, GUARD, PH32>(nk, x0, x1, x2, y0, y1, y2);
    }
};

template <class P>
__device__ __forceinline__ typename P::T2 czero2() {
    typename P::T2 v;
    v.x = 0;
    v.y = 0;
    return v;
}

// Complex product accumulator. K3: three real products (Karatsuba),
// P3 = (Ar + Ai)(Br + Bi), Im = P3 - P1 - P2; else four, Im = P3 + P4.
template <class P, bool K3>
struct Cacc {
    typename P::V4 p1, p2, p3, p4;
    __device__ __forceinline__ void zero() {
        for (int v = 0; v < 4; ++v) p1[v] = p2[v] = p3[v] = p4[v] = 0;
    }
    __device__ __forceinline__ void mma(typename P::T2 a, typename P::T2 b) {
        p1 = P::mma(a.x, b.x, p1);
        p2 = P::mma(a.y, b.y, p2);
        if (K3) {
            p3 = P::mma(a.x + a.y, b.x + b.y, p3);
        } else {
            p3 = P::mma(a.x, b.y, p3);
            p4 = P::mma(a.y, b.x, p4);
        }
    }
    __device__ __forceinline__ void scale(typename P::T f) {
        for (int v = 0; v < 4; ++v) {
            p1[v] *= f;
            p2[v] *= f;
            p3[v] *= f;
            if (!K3) p4[v] *= f;
        }
    }
    __device__ __forceinline__ double2 get(int v) const {
        const double r1 = p1[v], r2 = p2[v], r3 = p3[v];
        return K3 ? make_double2(r1 - r2, r3 - r1 - r2) : make_double2(r1 - r2, r3 + (double)p4[v]);
    }
};

// Real embedding of a complex product for tiles of at most HE rows (HE = 8 or
// 4): MFMA rows [0, HE) carry the real part and [HE, 2 HE) the imaginary part
// of the same output rows; two real MFMAs per step, D += A1 Br + A2 Bi with
// A1 = (Ar; Ai), A2 = (-Ai; Ar). The caller gives every lane the complex A
// entry of row (lane % 16) % HE (lanes of MFMA rows >= 2 HE: any finite value).
template <class P, int HE>
struct Cemb {
    typename P::V4 d;
    __device__ __forceinline__ void zero() {
        for (int v = 0; v < 4; ++v) d[v] = 0;
    }
    __device__ __forceinline__ void mma(typename P::T2 a, typename P::T2 b) {
        const bool re = (threadIdx.x & 15) < HE;
        const typename P::T a1 = re ? a.x : a.y;
        const typename P::T a2 = re ? -a.y : a.x;
        d = P::mma(a1, b.x, d);
        d = P::mma(a2, b.y, d);
    }
    __device__ __forceinline__ void scale(typename P::T f) {
        for (int v = 0; v < 4; ++v) d[v] *= f;
    }
    // emit(row, value) for the output rows (< HE) this lane holds; every lane
    // of the wave must call it (cross-lane exchange in FP32).
    template <class F>
    __device__ __forceinline__ void store(F emit) const {
        const int kk = (threadIdx.x & 63) >> 4;
        if constexpr (P::kF64) {  // lane holds MFMA rows kk + 4 v
            if (HE == 8) {
                emit(kk, make_double2(d[0], d[2]));
                emit(kk + 4, make_double2(d[1], d[3]));
            } else {
                emit(kk, make_double2(d[0], d[1]));
            }
        } else {  // lane holds MFMA rows 4 kk + v
            float im[4];
            for (int v = 0; v < 4; ++v) im[v] = __shfl_xor(d[v], HE == 8 ? 32 : 16);
            if (kk < HE / 4)
                for (int v = 0; v < 4; ++v) emit(4 * kk + v, make_double2(d[v], im[v]));
        }
    }
};

// Workgroup -> work block, a bijection on [0, gridDim.x). Workgroups are dealt
// round robin over the 8 XCDs (each with its own L2). order 0: identity;
// order 1: each XCD gets one contiguous eighth; order G >= 2: rounds of 8 G
// workgroups in which each XCD gets G consecutive work blocks (the last
// partial round keeps the identity).
__device__ __forceinline__ int work_block(int order) {
    const int b = blockIdx.x;
    if (order == 0) return b;
    const int n = gridDim.x, x = b & 7, i = b >> 3;
    if (order == 1) return x * (n >> 3) + min(x, n & 7) + i;
    const int r8 = 8 * order;
    if (b >= (n / r8) * r8) return b;
    return (b / r8) * r8 + x * order + i % order;
}

template <bool CJ, class P>
__device__ __forceinline__ typename P::T2 conjif(typename P::T2 v) {
    if (CJ) v.y = -v.y;
    return v;
}

// Lane i of each 16-lane row receives the value of lane (i + S) mod 16 of the
// same row (DPP row rotate right by 16 - S).
template <int S>
__device__ __forceinline__ float rot16(float v) {
    if (S == 0) return v;
    return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), 0x120 + (16 - S), 0xf, 0xf, false));
}
template <int S>
__device__ __forceinline__ double rot16(double v) {
    if (S == 0) return v;
    const long long u = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_mov_dpp((int)u, 0x120 + (16 - S), 0xf, 0xf, false);
    const int hi = __builtin_amdgcn_mov_dpp((int)(u >> 32), 0x120 + (16 - S), 0xf, 0xf, false);
    return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}
template <int S, class T2>
__device__ __forceinline__ T2 rot16v(T2 v) {
    v.x = rot16<S>(v.x);
    v.y = rot16<S>(v.y);
    return v;
}

// Raw buffer loads: reads past the end of a descriptor's range return 0, so
// the near-field walk and the coupling pipeline below run without clamps or
// masks. The varying part of an address goes into the scalar offset, the
// lane's part into the vector offset (both within the range check).
constexpr int kBufDword3 = 0x00020000;  // gfx9 buffer descriptor, raw dword access

template <class T>
__device__ __forceinline__ T buf_load(__amdgpu_buffer_rsrc_t r, uint32_t voff, uint32_t soff);
template <>
__device__ __forceinline__ float2 buf_load<float2>(__amdgpu_buffer_rsrc_t r, uint32_t voff, uint32_t soff) {
    return __builtin_bit_cast(float2, __builtin_amdgcn_raw_buffer_load_b64(r, voff, soff, 0));
}
template <>
__device__ __forceinline__ double2 buf_load<double2>(__amdgpu_buffer_rsrc_t r, uint32_t voff, uint32_t soff) {
    return __builtin_bit_cast(double2, __builtin_amdgcn_raw_buffer_load_b128(r, voff, soff, 0));
}
template <>
__device__ __forceinline__ double buf_load<double>(__amdgpu_buffer_rsrc_t r, uint32_t voff, uint32_t soff) {
    return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, voff, soff, 0));
}
template <>
__device__ __forceinline__ int buf_load<int>(__amdgpu_buffer_rsrc_t r, uint32_t voff, uint32_t soff) {
    return (int)__builtin_amdgcn_raw_buffer_load_b32(r, voff, soff, 0);
}
template <>
__device__ __forceinline__ uint4 buf_load<uint4>(__amdgpu_buffer_rsrc_t r, uint32_t voff, uint32_t soff) {
    return __builtin_bit_cast(uint4, __builtin_amdgcn_raw_buffer_load_b128(r, voff, soff, 0));
}
template <>
__device__ __forceinline__ uint2 buf_load<uint2>(__amdgpu_buffer_rsrc_t r, uint32_t voff, uint32_t soff) {
    return __builtin_bit_cast(uint2, __builtin_amdgcn_raw_buffer_load_b64(r, voff, soff, 0));
}

__device__ __forceinline__ __amdgpu_buffer_rsrc_t buf_rsrc(const void* base, uint32_t bytes) {
    return __builtin_amdgcn_make_buffer_rsrc((void*)base, (short)0, (int)bytes, kBufDword3);
}

// Operands of one near-field step: B rows of the G groups for each of the NC
// 16-column chunks, and the lane's source point.
template <class P, int G, int NC>
struct NearOps {
    typename P::T2 b[NC][G];
    // coordinate j < 3 of the source point in lane j of each 16-lane row
    // (G = 1) or of each quad (G > 1)
    double pq;
};

// Coordinate L of a 16-lane row's point held in lane L of the row: one
// 64-bit DPP move (row_newbcast).
template <int L>
__device__ __forceinline__ double row_double(double v) {
    const long long u = __double_as_longlong(v);
    return __longlong_as_double(__builtin_amdgcn_mov_dpp(u, 0x150 + L, 0xf, 0xf, false));
}

// Coordinate L of a quad's point held in lane L of the quad (DPP quad_perm).
template <int L>
__device__ __forceinline__ double quad_double(double v) {
    const long long u = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_mov_dpp((int)u, 0x55 * L, 0xf, 0xf, false);
    const int hi = __builtin_amdgcn_mov_dpp((int)(u >> 32), 0x55 * L, 0xf, 0xf, false);
    return __hiloint2double(hi, lo);
}

// Loads of one step: xs rows (vx[c][m]: the lane's offset of group m in
// column chunk c) and the
// source points, at byte offsets sp / sx into the range's points (rp) and
// right-hand sides (rx); reads past the range end return 0. Every lane of a
// 16-lane row (G = 1) or of a quad (G > 1) needs the same source point, so
// each lane loads one coordinate of it (vp) and the computing step
// broadcasts them.
template <class P, int G, int NC>
__device__ __forceinline__ void near_load(NearOps<P, G, NC>& o, __amdgpu_buffer_rsrc_t rp,
                                          __amdgpu_buffer_rsrc_t rx, uint32_t sp, uint32_t sx, uint32_t vp,
                                          const uint32_t (&vx)[NC][G]) {
#pragma unroll
    for (int c = 0; c < NC; ++c)
#pragma unroll
        for (int m = 0; m < G; ++m) o.b[c][m] = buf_load<typename P::T2>(rx, vx[c][m], sx);
    o.pq = buf_load<double>(rp, vp, sp);
}

// The lane's byte offset of its source-point coordinate in a step: coordinate
// min(j, 2) of source kk (16-row tiles, j = lane % 16) or of the quad's source
// ysrc (smaller tiles, j = lane % 4).
template <int G>
__device__ __forceinline__ uint32_t near_point_offset(int lane, int kk, int ysrc) {
    return G == 1 ? (uint32_t)kk * 24u + 8u * (uint32_t)min(lane & 15, 2)
                  : (uint32_t)ysrc * 24u + 8u * (uint32_t)min(lane & 3, 2);
}

// One near-field step of a tile with H valid-row slots (H = 16, 8 or 4).
// Lane (i, kk) = (lane % 16, lane / 16) evaluates G for row i % H and source
// 4 (i / H) + kk of the step, so a step evaluates 64 entries over 4 * 16 / H
// sources; MFMA m takes A = the values of group m, rotated down to lanes
// 0..H-1 of each 16-lane row (the rows >= H of the product are not stored),
// and B = xs of sources 4 m .. 4 m + 3 of the step. Sources past the range end
// read points 0 and B = 0; groups that start there (left: sources of the range
// from the step on) skip their products. The evaluated entries serve the NC
// column chunks (one accumulator each).
template <class P, int H, class ACC, bool EMB, bool NEG, bool GUARD, bool PH32, int NC>
__device__ __forceinline__ void near_compute(ACC (&acc)[NC], const NearK& nk, const double2* tab, double x0,
                                             double x1, double x2, const NearOps<P, 16 / H, NC>& o, int left) {
    using T2 = typename P::T2;
    constexpr int G = 16 / H;
    T2 g;
    if constexpr (G == 1)
        g = P::template eval<NEG, GUARD, PH32>(nk, tab, x0, x1, x2, row_double<0>(o.pq), row_double<1>(o.pq),
                                               row_double<2>(o.pq));
    else
        g = P::template eval<NEG, GUARD, PH32>(nk, tab, x0, x1, x2, quad_double<0>(o.pq), quad_double<1>(o.pq),
                                               quad_double<2>(o.pq));
    const int i = threadIdx.x & 15;
    // the range's last step: sources past its end read the point (0, 0, 0),
    // at which an unguarded kernel value may be infinite, with zero
    // right-hand sides; their values are zeroed (a guarded evaluation gives 0
    // there)
    if constexpr (!GUARD)
        if (left < 4 * G && 4 * (i / H) + (int)((threadIdx.x & 63) >> 4) >= left) g = czero2<P>();
#pragma unroll
    for (int m = 0; m < G; ++m) {
        if (m > 0 && 4 * m >= left) continue;
        T2 a;
        if constexpr (!EMB) {
            if (m == 0) a = g;
            else if (m == 1) a = rot16v<H>(g);
            else if (m == 2) a = rot16v<(2 * H) & 15>(g);
            else a = rot16v<(3 * H) & 15>(g);
        } else if constexpr (H == 8) {
            // embedding: lanes i and i + 8 both need row i % 8 of group m
            const T2 r8 = rot16v<8>(g);
            a = ((i < 8) == (m == 0)) ? g : r8;
        } else {
            // embedding, 4 rows: lanes 0..3 need lane 4 m + i, lanes 4..7 lane 4 m + i - 4
            const T2 ra = m == 0 ? g : (m == 1 ? rot16v<4>(g) : (m == 2 ? rot16v<8>(g) : rot16v<12>(g)));
            const T2 rb = m == 0 ? rot16v<12>(g) : (m == 1 ? g : (m == 2 ? rot16v<4>(g) : rot16v<8>(g)));
            a = i < 4 ? ra : rb;
        }
#pragma unroll
        for (int c = 0; c < NC; ++c) acc[c].mma(a, o.b[c][m]);
    }
}

// Calls steps(guard, ph32) with compile-time flags: FP64 always guarded, one
// phase form; MIXED guarded only for ranges that need it, phase in FP32 when
// the operator's near pairs allow it (both uniform over the wave).
template <class P, class F>
__device__ __forceinline__ void near_dispatch(F& steps, bool guard, bool ph32) {
    if constexpr (P::kF64) {
        steps(std::true_type(), std::false_type());
    } else if (ph32) {
        if (guard) steps(std::true_type(), std::true_type());
        else steps(std::false_type(), std::true_type());
    } else {
        if (guard) steps(std::true_type(), std::false_type());
        else steps(std::false_type(), std::false_type());
    }
}

// Steps per iteration of the MIXED near-field walk for tiles of more than 8
// rows (their loads are all issued before the first of them computes).
#ifndef NCA_NEAR_UNROLL
#define NCA_NEAR_UNROLL 2
#endif

// Near-field walk of one tile over its merged source ranges: one buffer
// descriptor pair per range (points, and the right-hand sides from column q0
// on). MIXED tiles of more than 8 rows take NCA_NEAR_UNROLL steps per
// iteration while the range has sources for all of them (the last may read
// past its end: zeros), all their loads issued before the first computes;
// the rest of a range, FP64 tiles and smaller ones take one step per
// iteration (in FP64 a second register set would cost occupancy: 76 VGPRs,
// 6 waves). Control flow is uniform over the wave; byte offsets inside
// a range are 32-bit (range length x right-hand sides x element size < 2^32).
// rot: the walk starts at range rot mod nr (tiles that share sources then read
// them at different times).
template <class P, int H, class ACC, bool EMB, bool NEG, int NC>
__device__ __forceinline__ void near_walk(ACC (&acc)[NC], const NearK& nk, const double2* tab, int lane,
                                          double x0, double x1, double x2, int64_t e0, int64_t e1, int rot,
                                          const int32_t* __restrict__ nrlo, const int32_t* __restrict__ nrhi,
                                          const double* __restrict__ spts, const typename P::T2* __restrict__ xsq,
                                          uint32_t R, const int (&jl)[NC]) {
    using T2 = typename P::T2;
    constexpr int S = 4 * (16 / H), G = 16 / H;
    constexpr int U = (H == 16 && !P::kF64 && NC == 1) ? NCA_NEAR_UNROLL : 1;
    constexpr uint32_t kT = sizeof(T2);
    const int kk = lane >> 4, ysrc = 4 * ((lane & 15) / H) + kk;
    const uint32_t vp = near_point_offset<G>(lane, kk, ysrc);
    uint32_t vx[NC][G];
#pragma unroll
    for (int c = 0; c < NC; ++c)
#pragma unroll
        for (int m = 0; m < G; ++m) vx[c][m] = ((uint32_t)(4 * m + kk) * R + (uint32_t)jl[c]) * kT;
    const uint32_t rowb = R * kT;
    const int nr = (int)(e1 - e0);
    for (int c = 0; c < nr; ++c) {
        int ec = c + (nr ? rot % nr : 0);
        if (ec >= nr) ec -= nr;
        const int64_t e = e0 + ec;
        const int lo = nrlo[e], hr = nrhi[e];
        const int hi = hr & 0x7fffffff;  // bit 31: the range keeps the zero-distance guard
        const uint32_t n = (uint32_t)(hi - lo);
        const __amdgpu_buffer_rsrc_t rp = buf_rsrc(spts + 3 * (int64_t)lo, n * 24u);
        const __amdgpu_buffer_rsrc_t rx = buf_rsrc(xsq + (int64_t)lo * R, n * rowb);
        auto steps = [&](auto guard, auto ph32) {
            constexpr bool GU = decltype(guard)::value, PH = decltype(ph32)::value;
            int left = (int)n;
            uint32_t sp = 0, sx = 0;
            if constexpr (U > 1) {
                for (; left > (U - 1) * S; left -= U * S, sp += U * S * 24u, sx += U * S * rowb) {
                    NearOps<P, G, NC> o[U];
#pragma unroll
                    for (int u = 0; u < U; ++u)
                        near_load<P, G, NC>(o[u], rp, rx, sp + u * S * 24u, sx + u * S * rowb, vp, vx);
#pragma unroll
                    for (int u = 0; u < U; ++u)
                        near_compute<P, H, ACC, EMB, NEG, GU, PH, NC>(acc, nk, tab, x0, x1, x2, o[u],
                                                                      left - u * S);
                }
            }
            for (; left > 0; left -= S, sp += S * 24u, sx += S * rowb) {
                NearOps<P, G, NC> o;
                near_load<P, G, NC>(o, rp, rx, sp, sx, vp, vx);
                near_compute<P, H, ACC, EMB, NEG, GU, PH, NC>(acc, nk, tab, x0, x1, x2, o, left);
            }
        };
        // FP64: one guarded step loop (a second copy costs the kernel occupancy)
        near_dispatch<P>(steps, hr < 0, nk.phase32 != 0);
    }
}

// Stage 5 + near field for 16 NC right-hand sides (NC = 1 or 2 chunks of 16
// columns: each evaluated entry serves all of them): one wave per 16-row tile
// of a destination leaf. The near-field sources over each merged contiguous
// point range, with 16, 8 or 4 row slots by the rows of the tile, then the
// basis rows (A = U_t or V_t^T, B = yh_t). Rows past the tile end repeat its
// last row and are not stored.
template <class P, bool FWD, bool CJ, bool K3, bool NEG, int NC>
__global__ void __launch_bounds__(256, 4)
    k_leaf_out_mfma(const int32_t* __restrict__ tile_leaf, const int32_t* __restrict__ tile_row0, int64_t ntiles,
                    const int32_t* __restrict__ leaves, const int32_t* __restrict__ leafbasis,
                    const int64_t* __restrict__ boff, const typename P::T2* __restrict__ bblocks,
                    const int32_t* __restrict__ dlo, const int32_t* __restrict__ dhi,
                    const int32_t* __restrict__ dperm, const int32_t* __restrict__ dcptr,
                    const double* __restrict__ dpts, const int64_t* __restrict__ nrptr,
                    const int32_t* __restrict__ nrlo, const int32_t* __restrict__ nrhi,
                    const double* __restrict__ spts, const double2* __restrict__ yh,
                    const typename P::T2* __restrict__ xs, int64_t R, int64_t q0, NearK nk,
                    const double2* __restrict__ sctab, double2 alpha, double2 beta, double2* __restrict__ y,
                    int64_t ldy, int order, int stagger, int nq, const double2* __restrict__ ynear) {
    using T2 = typename P::T2;
    __shared__ double2 tab[P::kTab ? kSinCosTab : 1];
    if (P::kTab) {
        static_assert(!P::kTab || kSinCosTab == 256, "one table entry per thread");
        tab[threadIdx.x] = sctab[threadIdx.x];
        __syncthreads();
    }
    const int lane = threadIdx.x & 63;
    const int w = __builtin_amdgcn_readfirstlane(work_block(order) * 4 + (int)(threadIdx.x >> 6));
    if (w >= ntiles) return;
    const int li = tile_leaf[w], row0 = tile_row0[w];
    const int t = leaves[li], bi = leafbasis[li];
    const int l0 = dlo[t], nt = dhi[t] - l0;
    const int nrows = min(16, nt - row0);
    const int i = lane & 15, kk = lane >> 4, j = lane & 15;
    const int64_t e0 = nrptr[t], e1 = nrptr[t + 1];
    const int rot = stagger ? w * 7 : 0;
    // columns past the live ones read the last live one (not stored)
    int jl[NC];
#pragma unroll
    for (int c = 0; c < NC; ++c) jl[c] = min(16 * c + j, nq - 1);
    const int rt = dcptr[t + 1] - dcptr[t];
    const T2* B = bi >= 0 ? bblocks + boff[bi] : nullptr;
    // near field first (4 pi G), scaled by 1 / (4 pi), then the basis rows
    auto tile = [&](auto& acc, auto hval, auto emb, int rsel) {
        constexpr int H = decltype(hval)::value;
        constexpr bool EMB = decltype(emb)::value;
        using ACC = std::remove_reference_t<decltype(acc[0])>;
#pragma unroll
        for (int c = 0; c < NC; ++c) acc[c].zero();
        // the near field: evaluated here, or stored and applied before (ynear)
        if (!ynear) {
            const double* xp = dpts + 3 * (int64_t)(l0 + row0 + min(rsel, nrows - 1));
            near_walk<P, H, ACC, EMB, NEG, NC>(acc, nk, tab, lane, xp[0], xp[1], xp[2], e0, e1, rot, nrlo, nrhi,
                                               spts, xs + q0, (uint32_t)R, jl);
#pragma unroll
            for (int c = 0; c < NC; ++c) acc[c].scale((typename P::T)kInvFourPi);
        }
        if (B) {
            const int rowc = row0 + min(rsel, nrows - 1);
            const double2* yt = yh + (int64_t)dcptr[t] * R + q0;
            for (int k0 = 0; k0 < rt; k0 += 4) {
                const int k = k0 + kk;
                const int kc = min(k, rt - 1);
                const T2 a =
                    conjif<CJ, P>(B[FWD ? (int64_t)rowc + (int64_t)kc * nt : (int64_t)kc + (int64_t)rowc * rt]);
#pragma unroll
                for (int c = 0; c < NC; ++c) {
                    T2 b = P::cvt(yt[(int64_t)kc * R + jl[c]]);
                    if (k >= rt) b = czero2<P>();
                    acc[c].mma(a, b);
                }
            }
        }
    };
    auto emitter = [&](int c) {
        return [&, c](int row, double2 v) {
            const int col = 16 * c + j;
            if (row < nrows && col < nq) {
                double2* out = y + (int64_t)dperm[l0 + row0 + row] + (q0 + col) * ldy;
                if (ynear) v = cadd(v, ynear[(int64_t)(l0 + row0 + row) * R + q0 + col]);
                double2 val = cmul(alpha, v);
                if (beta.x != 0.0 || beta.y != 0.0) val = cadd(val, cmul(beta, *out));
                *out = val;
            }
        };
    };
    if (nrows > 8) {
        Cacc<P, K3> acc[NC];
        tile(acc, std::integral_constant<int, 16>(), std::false_type(), i);
#pragma unroll
        for (int c = 0; c < NC; ++c) {
            const auto emit = emitter(c);
#pragma unroll
            for (int v = 0; v < 4; ++v) emit(P::row(kk, v), acc[c].get(v));
        }
    } else if (nrows > 4) {
        Cemb<P, 8> acc[NC];
        tile(acc, std::integral_constant<int, 8>(), std::true_type(), i & 7);
#pragma unroll
        for (int c = 0; c < NC; ++c) acc[c].store(emitter(c));
    } else {
        Cemb<P, 4> acc[NC];
        tile(acc, std::integral_constant<int, 4>(), std::true_type(), i & 3);
#pragma unroll
        for (int c = 0; c < NC; ++c) acc[c].store(emitter(c));
    }
}

// ---------------------------------------------------------------------------
// MIXED near field on f16 MFMA (v_mfma_f32_16x16x32_f16: eight times the K
// depth of the f32 16x16x4 product at about 0.6 of its issue time). Every
// FP32 operand a is split into f16 parts ah = f16(a), al = f16(a - ah)
// (|a - ah - al| <= 2^-22 |a| while both parts stay normal), and every real
// product is formed as ah bh + ah bl + al bh with FP32 accumulation (the
// dropped al bl is below 2^-22 |a b|). The parts stay in f16's range because
// the kernel is evaluated on copies of the points scaled by t = 2^-e, chosen
// so that 1/r' = 1/(t r) lies in (2^14, 2^15] at the closest near pair (the
// operator takes this path only when its near pairs span at most 2^17 in
// distance, so 1/r' >= 2^-3 at the farthest), and the gathered right-hand
// sides are scaled per column by s_q = 2^m (largest |re|, |im| in [2^14,
// 2^15)); the tile's sums are scaled back by t / (4 pi s_q).
//
// Gathered right-hand sides for this path (xs16): per tree position p and
// column one 16-byte entry holding the f16 parts of p and p + 4 side by side
// (re hi, im hi, re lo, im lo, each as a pair), in the order of the c32
// copy: every tile form below puts sources s and s + 4 of a block into one
// K-slot pair, so one 16-byte load gives a lane the four B registers of a
// slot pair without permutes. A tile's near field is walked per merged source
// range in blocks of sub-steps whose kernel values and right-hand sides fill
// the 32 K-slots of one f16 product: 16-row tiles take 8 sub-steps of 4
// sources (lane (i, kk) evaluates row i x source 4 m + kk, K-slot 8 kk + m),
// 8-row tiles 4 sub-steps of 8 sources and 4-row tiles 2 sub-steps of 16
// (row i % H, source S m + 4 (i / H) + kk; the other sources of the lane's
// row come from the lanes that evaluated them, by DPP), in the real
// embedding of the f32 kernel (MFMA rows [0, H) the real parts, [H, 2 H) the
// imaginary parts). The last block of a range is padded with zero slots;
// sources past a range end read zero right-hand sides (buffer range check).
// Each block's 32 source points go to a 1 KiB LDS buffer of the wave with
// one direct-to-LDS buffer load (16 bytes a lane), and every lane reads the
// coordinates of its sources from there (two LDS reads a sub-step, broadcast
// within the lanes that share a source): one texture instruction a block
// instead of one a sub-step, and no cross-lane moves on the VALU.

typedef _Float16 h16x8 __attribute__((ext_vector_type(8)));
typedef _Float16 h16x2 __attribute__((ext_vector_type(2)));
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ flt4 mma_h16(u32x4 a, u32x4 b, flt4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_f16(__builtin_bit_cast(h16x8, a), __builtin_bit_cast(h16x8, b), c, 0,
                                                  0, 0);
}
// c + a b for split operands: ah bh + ah bl + al bh
__device__ __forceinline__ flt4 mma_h16x3(u32x4 ah, u32x4 al, u32x4 bh, u32x4 bl, flt4 c) {
    c = mma_h16(ah, bh, c);
    c = mma_h16(ah, bl, c);
    return mma_h16(al, bh, c);
}
// The products accumulate in the running FP32 sums directly. The f16
// product rounds its result with few guard bits below the exponent of the
// larger of the accumulator and the products (measured: up to 13 ulp of a
// large accumulator, against 1.9 for the f32 16x16x4 product); summing each
// block from zero and adding it on VALU halves the MIXED error of the 1M
// sphere (6.8e-7 -> 3.3e-7 against FP64) at 2.5 % of the sweep, which the
// 1e-5 MIXED bar does not need.
// f16 parts of two FP32 values: hi = (f16(a), f16(b)), lo = (f16(a - hi.x),
// f16(b - hi.y)), in three instructions: v_cvt_pk_f16_f32 for hi, then
// v_fma_mix (a x 1 - hi, exact, rounded once to f16) into each half of lo.
// Written out so that hi is the rounding of the given FP32 value itself (the
// compiler may otherwise round a product feeding the conversion to f16
// directly for one use and through FP32 for another; the two can differ by
// one f16 ulp, and lo would then not complete hi).
__device__ __forceinline__ void split_h16(float a, float b, uint32_t& hi, uint32_t& lo) {
    asm("v_cvt_pk_f16_f32 %0, %1, %2" : "=v"(hi) : "v"(a), "v"(b));
    asm("v_fma_mixlo_f16 %0, %1, 1.0, -%2 op_sel_hi:[0,0,1]" : "=v"(lo) : "v"(a), "v"(hi));
    asm("v_fma_mixhi_f16 %0, %1, 1.0, -%2 op_sel:[0,0,1] op_sel_hi:[0,0,1]" : "+v"(lo) : "v"(b), "v"(hi));
}
// (low half of u, low half of v) and (high half of u, high half of v)
__device__ __forceinline__ uint32_t pack_lo(uint32_t u, uint32_t v) { return __builtin_amdgcn_perm(v, u, 0x05040100u); }
__device__ __forceinline__ uint32_t pack_hi(uint32_t u, uint32_t v) { return __builtin_amdgcn_perm(v, u, 0x07060302u); }
// DPP row rotation of a 32-bit value (lane i of a 16-lane row gets lane (i + S) mod 16)
template <int S>
__device__ __forceinline__ uint32_t rot16u(uint32_t v) {
    return (uint32_t)__builtin_amdgcn_mov_dpp((int)v, 0x120 + (16 - S), 0xf, 0xf, false);
}

// 4 pi G on the scaled points (NEG, GUARD and PH32 as for helmholtz_f32, the
// guard thresholds from nk); PH32: the phase is below one revolution plus
// rounding, inside the hardware sin/cos domain without reduction.
template <bool NEG, bool GUARD, bool PH32>
__device__ __forceinline__ float2 helmholtz_h16(const NearK& nk, double x0, double x1, double x2, double y0,
                                                double y1, double y2) {
    const double dx = x0 - y0, dy = x1 - y1, dz = x2 - y2;
    const double r2 = fma(dz, dz, fma(dy, dy, GUARD ? fma(dx, dx, nk.gadd) : dx * dx));
    const float r2f = (float)r2;
    const float rf = __builtin_amdgcn_rsqf(r2f);
    float f;
    if constexpr (PH32) {
        f = (r2f * rf) * nk.k2pif;
    } else {
        const double ri = (double)rf;
        const double r0 = r2 * ri;
        const double e = fma(-r0, r0, r2);
        const double t = fma(e, ri * nk.k4pi, r0 * nk.k2pi);
        f = (float)__builtin_amdgcn_fract(t);
    }
    const float c = __builtin_amdgcn_cosf(f), s = __builtin_amdgcn_sinf(f);
    const float sc = !GUARD || r2 > nk.gmin ? rf : 0.0f;
    return make_float2(c * sc, NEG ? -(s * sc) : s * sc);
}

// The same from FP32 coordinates relative to the tile's first row (rel32):
// the difference of two such coordinates errs by a few units of the row's
// distance from the first row plus the pair's distance, which the rel32
// condition keeps within a small multiple of the pair's distance.
template <bool NEG, bool GUARD>
__device__ __forceinline__ float2 helmholtz_h16r(const NearK& nk, float x0, float x1, float x2, float y0, float y1,
                                                 float y2) {
    const float dx = x0 - y0, dy = x1 - y1, dz = x2 - y2;
    const float r2 = fmaf(dz, dz, fmaf(dy, dy, GUARD ? fmaf(dx, dx, nk.gaddf) : dx * dx));
    const float rf = __builtin_amdgcn_rsqf(r2);
    const float f = (r2 * rf) * nk.k2pif;
    const float c = __builtin_amdgcn_cosf(f), s = __builtin_amdgcn_sinf(f);
    const float sc = !GUARD || r2 > nk.gminf ? rf : 0.0f;
    return make_float2(c * sc, NEG ? -(s * sc) : s * sc);
}

typedef __attribute__((address_space(3))) double lds_double;
typedef __attribute__((address_space(3))) float lds_float;
// LDS element and row coordinate types of the f16 walk: FP64 points, or FP32
// coordinates relative to the tile's first row (RL)
template <bool RL>
struct H16Pt {
    typedef lds_double L;
    typedef double X;
};
template <>
struct H16Pt<true> {
    typedef lds_float L;
    typedef float X;
};
template <bool NEG, bool GU, bool PH, bool RL, class X, class L>
__device__ __forceinline__ float2 h16_eval(const NearK& nk, X x0, X x1, X x2, const L* p) {
    if constexpr (RL) return helmholtz_h16r<NEG, GU>(nk, x0, x1, x2, p[0], p[1], p[2]);
    else return helmholtz_h16<NEG, GU, PH>(nk, x0, x1, x2, p[0], p[1], p[2]);
}
// Copies the 1,024 bytes of a range's points from byte offset sp on (32
// points and the start of the next block; bytes past the range end read 0)
// into the wave's LDS buffer: lane l's 16 bytes to bytes 16 l (direct-to-LDS
// buffer load). The wave's earlier reads of the buffer complete first.
__device__ __forceinline__ void h16_stage_points(lds_double* lds, __amdgpu_buffer_rsrc_t rp, uint32_t sp, int lane) {
    asm volatile("s_waitcnt lgkmcnt(0)" ::: "memory");
    __builtin_amdgcn_raw_ptr_buffer_load_lds(rp, (__attribute__((address_space(3))) void*)lds, 16,
                                             (uint32_t)lane * 16u + sp, 0, 0, 0);
    asm volatile("" ::: "memory");
}
// Waits until at most N vector memory loads are in flight (loads complete in
// issue order): with the N loads issued after a staging copy, the copy is in
// LDS. The compiler does not order LDS reads after direct-to-LDS loads.
template <int N>
__device__ __forceinline__ void h16_wait_vm() {
    asm volatile("s_waitcnt vmcnt(%0)" ::"n"(N) : "memory");
}

// B operands of one block of the f16 walk, from the paired split copy
// (per source s and column: the f16 parts of sources s and s + 4, see
// k_split16): slot pair v (v < 4) of each column chunk is the entry of
// source 8 v + kk of the block (vx[c]: the lane's column and kk, sx: the
// block's offset, rowb: bytes per source); entries whose first source lies
// at or past nsrc (the block's evaluated sources) are zero.
template <int NC, bool FULL>
__device__ __forceinline__ void h16_load_b(uint4 (&xb)[NC][4], __amdgpu_buffer_rsrc_t rx, uint32_t sx,
                                           const uint32_t (&vx)[NC], uint32_t rowb, int nsrc) {
#pragma unroll
    for (int v = 0; v < 4; ++v) {
        if (FULL || 8 * v < nsrc) {
#pragma unroll
            for (int c = 0; c < NC; ++c) xb[c][v] = buf_load<uint4>(rx, vx[c], sx + (uint32_t)(8 * v) * rowb);
        } else {
#pragma unroll
            for (int c = 0; c < NC; ++c) xb[c][v] = make_uint4(0u, 0u, 0u, 0u);
        }
    }
}
// Re and im, high and low f16 parts of the four slot pairs.
__device__ __forceinline__ void h16_b_parts(const uint4 (&xb)[4], u32x4& brh, u32x4& bih, u32x4& brl, u32x4& bil) {
#pragma unroll
    for (int v = 0; v < 4; ++v) {
        brh[v] = xb[v].x;
        bih[v] = xb[v].y;
        brl[v] = xb[v].z;
        bil[v] = xb[v].w;
    }
}

// One block of a 16-row tile's walk: nsub (1..8) sub-steps of 4 sources,
// their points staged in LDS (lp: this lane's first source there, the
// staging load the last vector load issued), their right-hand sides at byte
// offset sx of the range; P, Q, I accumulate Gr xr, Gi xi and Gr xi + Gi xr.
template <int NC, bool NEG, bool GU, bool PH, bool FULL, bool RL, class LandF>
__device__ __forceinline__ void h16_block16(flt4 (&P)[NC], flt4 (&Q)[NC], flt4 (&I)[NC], const NearK& nk,
                                            typename H16Pt<RL>::X x0, typename H16Pt<RL>::X x1,
                                            typename H16Pt<RL>::X x2, const typename H16Pt<RL>::L* lp,
                                            __amdgpu_buffer_rsrc_t rx, uint32_t sx,
                                            const uint32_t (&vx)[NC], uint32_t rowb, int nsub, int nval,
                                            LandF land) {
    uint4 xb[NC][4];
    h16_load_b<NC, FULL>(xb, rx, sx, vx, rowb, 4 * nsub);
    land(std::bool_constant<FULL>());
    const int kk = (threadIdx.x & 63) >> 4;
    u32x4 arh, aih, arl, ail;
#pragma unroll
    for (int v = 0; v < 4; ++v) {
        float2 g[2];
#pragma unroll
        for (int e = 0; e < 2; ++e) {
            const int m = 2 * v + e;
            if (FULL || m < nsub) {
                g[e] = h16_eval<NEG, GU, PH, RL>(nk, x0, x1, x2, lp + 12 * m);
                // the range's last sub-step: sources past its end read the point
                // (0, 0, 0), at which the kernel may be infinite, with zero
                // right-hand sides; their values are zeroed
                if (!FULL && 4 * m + 4 > nval && 4 * m + kk >= nval) g[e] = make_float2(0.0f, 0.0f);
            } else {
                g[e] = make_float2(0.0f, 0.0f);
            }
        }
        uint32_t h, l;
        split_h16(g[0].x, g[1].x, h, l);
        arh[v] = h;
        arl[v] = l;
        split_h16(g[0].y, g[1].y, h, l);
        aih[v] = h;
        ail[v] = l;
    }
#pragma unroll
    for (int c = 0; c < NC; ++c) {
        u32x4 brh, bih, brl, bil;
        h16_b_parts(xb[c], brh, bih, brl, bil);
        P[c] = mma_h16x3(arh, arl, brh, brl, P[c]);
        Q[c] = mma_h16x3(aih, ail, bih, bil, Q[c]);
        I[c] = mma_h16x3(aih, ail, brh, brl, mma_h16x3(arh, arl, bih, bil, I[c]));
    }
}

// One block of an 8-row (H = 8) or 4-row (H = 4) tile's walk in the real
// embedding: nsub (1 .. 32 / S) sub-steps of S = 64 / H sources (points in
// LDS as for h16_block16). Lane (i, kk)
// evaluates row i % H x source S m + 4 (i / H) + kk and receives the other
// sources of its row's sub-step from lanes i +- H (rotations); sel1 / sel2
// pick, by v_perm, the parts of two slots for A1 (real rows: Gr, imaginary
// rows: Gi) and A2 (real rows: Gi, imaginary rows: Gr) in slot order (imrow:
// the lane's MFMA row holds an imaginary part). D1 += A1 xr, D2 += A2 xi; the
// real rows are D1 - D2, the imaginary rows D1 + D2.
template <int H, int NC, bool NEG, bool GU, bool PH, bool FULL, bool RL, class LandF>
__device__ __forceinline__ void h16_block_emb(flt4 (&D1)[NC], flt4 (&D2)[NC], const NearK& nk,
                                              typename H16Pt<RL>::X x0, typename H16Pt<RL>::X x1,
                                              typename H16Pt<RL>::X x2, const typename H16Pt<RL>::L* lp,
                                              __amdgpu_buffer_rsrc_t rx, uint32_t sx,
                                              const uint32_t (&vx)[NC], uint32_t rowb, int nsub, int nval,
                                              bool imrow, uint32_t sel1, uint32_t sel2, LandF land) {
    constexpr int S = 64 / H, NS = 32 / S;
    uint4 xb[NC][4];
    h16_load_b<NC, FULL>(xb, rx, sx, vx, rowb, S * nsub);
    land(std::bool_constant<FULL>());
    // this lane's source in a sub-step, after the sub-step's first
    const int own = 4 * ((int)(threadIdx.x & 15) / H) + (int)((threadIdx.x & 63) >> 4);
    u32x4 a1h, a2h, a1l, a2l;
#pragma unroll
    for (int m = 0; m < NS; ++m) {
        float2 g = make_float2(0.0f, 0.0f);
        if (FULL || m < nsub) {
            g = h16_eval<NEG, GU, PH, RL>(nk, x0, x1, x2, lp + 3 * S * m);
            // sources past the range end (the point (0, 0, 0)): zero values
            if (!FULL && S * m + S > nval && S * m + own >= nval) g = make_float2(0.0f, 0.0f);
        }
        uint32_t oh, ol;
        split_h16(g.x, g.y, oh, ol);
        if constexpr (H == 8) {
            // slots 2 m (lanes i < 8 own) and 2 m + 1 (lanes i >= 8 own); the order is in the selectors
            const uint32_t ph = rot16u<8>(oh), pl = rot16u<8>(ol);
            a1h[m] = __builtin_amdgcn_perm(ph, oh, sel1);
            a2h[m] = __builtin_amdgcn_perm(ph, oh, sel2);
            a1l[m] = __builtin_amdgcn_perm(pl, ol, sel1);
            a2l[m] = __builtin_amdgcn_perm(pl, ol, sel2);
        } else {
            // slots 4 m + q from the lane 4 q + i % 4 of the row: rotation r_d gives
            // group (i / 4 + d) mod 4; real rows (i / 4 = 0) pair (r0, r1), (r2, r3),
            // imaginary rows (i / 4 = 1) (r3, r0), (r1, r2)
            const uint32_t r1h = rot16u<4>(oh), r2h = rot16u<8>(oh), r3h = rot16u<12>(oh);
            const uint32_t r1l = rot16u<4>(ol), r2l = rot16u<8>(ol), r3l = rot16u<12>(ol);
            const uint32_t xh = imrow ? r3h : r1h, yhh = imrow ? r1h : r3h;
            const uint32_t xl = imrow ? r3l : r1l, yl = imrow ? r1l : r3l;
            a1h[2 * m] = __builtin_amdgcn_perm(xh, oh, sel1);
            a2h[2 * m] = __builtin_amdgcn_perm(xh, oh, sel2);
            a1l[2 * m] = __builtin_amdgcn_perm(xl, ol, sel1);
            a2l[2 * m] = __builtin_amdgcn_perm(xl, ol, sel2);
            a1h[2 * m + 1] = __builtin_amdgcn_perm(yhh, r2h, sel1);
            a2h[2 * m + 1] = __builtin_amdgcn_perm(yhh, r2h, sel2);
            a1l[2 * m + 1] = __builtin_amdgcn_perm(yl, r2l, sel1);
            a2l[2 * m + 1] = __builtin_amdgcn_perm(yl, r2l, sel2);
        }
    }
#pragma unroll
    for (int c = 0; c < NC; ++c) {
        u32x4 brh, bih, brl, bil;
        h16_b_parts(xb[c], brh, bih, brl, bil);
        D1[c] = mma_h16x3(a1h, a1l, brh, brl, D1[c]);
        D2[c] = mma_h16x3(a2h, a2l, bih, bil, D2[c]);
    }
}

// The f16 walk of one tile: H row slots (16, 8 or 4), this lane's row at
// xp (3 doubles), merged source ranges e0 .. e0 + nr - 1 (from e0 + rot on,
// cyclically; bit 31 of the end: the range keeps the zero-distance guard),
// source points spts (3 doubles each, staged per block in the wave's 1 KiB
// LDS buffer lds), split right-hand sides xsq (R per source, 16 bytes each:
// the paired parts of k_split16, from the first column on), columns jl.
// Accumulates
// Gr xr, Gi xi and Gr xi + Gi xr into A1, A2, A3 (H = 16), or A1 xr and
// A2 xi into A1, A2 (the real embedding, H = 8 or 4). RL: the points are
// staged as FP32 coordinates relative to cp (the tile's first row; lane l
// converts doubles 2 l and 2 l + 1 of a block), else copied as they are by
// one direct-to-LDS load a block.
template <int H, int NC, bool NEG, bool PH, bool RL>
__device__ __forceinline__ void h16_walk(flt4 (&A1)[NC], flt4 (&A2)[NC], flt4 (&A3)[NC], const NearK& nk,
                                         const double* xp, const double* cp, int lane, int64_t e0, int nr, int rot,
                                         const int32_t* __restrict__ nrlo, const int32_t* __restrict__ nrhi,
                                         const double* __restrict__ spts, const uint4* __restrict__ xsq, int64_t R,
                                         const int (&jl)[NC], bool imrow, uint32_t sel1, uint32_t sel2,
                                         lds_double* lds) {
    constexpr int S = 64 / H, NB = 32 / S;
    using X = typename H16Pt<RL>::X;
    const int kk = lane >> 4;
    X x0, x1, x2;
    double ca = 0.0, cb = 0.0;
    if constexpr (RL) {
        const double c[3] = {cp[0], cp[1], cp[2]};
        x0 = (float)(xp[0] - c[0]);
        x1 = (float)(xp[1] - c[1]);
        x2 = (float)(xp[2] - c[2]);
        ca = c[(2 * lane) % 3];
        cb = c[(2 * lane + 1) % 3];
    } else {
        x0 = xp[0];
        x1 = xp[1];
        x2 = xp[2];
    }
    const int ysrc = 4 * ((lane & 15) / H) + kk;
    const typename H16Pt<RL>::L* lp = (const typename H16Pt<RL>::L*)lds + 3 * ysrc;
    const uint32_t rowb = (uint32_t)R * 16u;
    uint32_t vx[NC];
#pragma unroll
    for (int c = 0; c < NC; ++c) vx[c] = ((uint32_t)kk * (uint32_t)R + (uint32_t)jl[c]) * 16u;
    for (int cc = 0; cc < nr; ++cc) {
        int ec = cc + rot;
        if (ec >= nr) ec -= nr;
        const int64_t e = e0 + ec;
        const int lo = nrlo[e], hr = nrhi[e];
        const int hi = hr & 0x7fffffff;
        const uint32_t n = (uint32_t)(hi - lo);
        const __amdgpu_buffer_rsrc_t rp = buf_rsrc(spts + 3 * (int64_t)lo, n * 24u);
        const __amdgpu_buffer_rsrc_t rx = buf_rsrc(xsq + (int64_t)lo * R, n * rowb);
        // full blocks of 32 sources, then the tail (rem sources: fewer
        // sub-steps, and the sources past the range end masked)
        const int nfull = (int)(n / 32u), rem = (int)(n % 32u);
        auto blocks = [&](auto guard) {
            constexpr bool GU = decltype(guard)::value;
            uint32_t sp = 0, sx = 0;
            // one block: stage its points, then hand the block function the
            // step that makes them readable once its right-hand side loads
            // are issued
            auto block = [&](auto full, int nsub, int nval) {
                constexpr bool FU = decltype(full)::value;
                double2 pv;
                if constexpr (RL) pv = buf_load<double2>(rp, (uint32_t)lane * 16u, sp);
                else h16_stage_points(lds, rp, sp, lane);
                auto land = [&](auto) {
                    if constexpr (RL) {
                        // earlier LDS reads of this wave precede the write in
                        // LDS order; the compiler keeps the reads below after it
                        asm volatile("" ::: "memory");
                        ((lds_float*)lds)[2 * lane] = (float)(pv.x - ca);
                        ((lds_float*)lds)[2 * lane + 1] = (float)(pv.y - cb);
                        asm volatile("" ::: "memory");
                    } else {
                        if constexpr (FU) h16_wait_vm<4 * NC>();
                        else h16_wait_vm<0>();
                    }
                };
                if constexpr (H == 16)
                    h16_block16<NC, NEG, GU, PH, FU, RL>(A1, A2, A3, nk, x0, x1, x2, lp, rx, sx, vx, rowb, nsub,
                                                        nval, land);
                else
                    h16_block_emb<H, NC, NEG, GU, PH, FU, RL>(A1, A2, nk, x0, x1, x2, lp, rx, sx, vx, rowb, nsub,
                                                             nval, imrow, sel1, sel2, land);
            };
            for (int b = 0; b < nfull; ++b, sp += 768u, sx += 32u * rowb) block(std::true_type(), NB, 32);
            if (rem > 0) block(std::false_type(), (rem + S - 1) / S, rem);
        };
        if (hr < 0) blocks(std::true_type());
        else blocks(std::false_type());
    }
}

// One tile on the f16 path with nrows valid rows: the 16-row form above 8
// rows, else the real embedding with 8 or 4 row slots; rowptr(rsel) gives
// the coordinates of row slot rsel (rows past the tile repeat its last); rel:
// the tile takes the relative FP32 distance form (with phase32). The
// sums, scaled by osc (per column chunk), go to fin(acc, rsel) as a
// Cacc<PrecF32, false> or Cemb<PrecF32, H> accumulator per chunk; lds: the
// wave's 1 KiB LDS buffer for the source points.
template <int NC, bool NEG, class RowF, class FinF>
__device__ __forceinline__ void h16_tile(const NearK& nk, bool rel, int nrows, int lane, RowF rowptr, int64_t e0,
                                         int nr, int rot, const int32_t* __restrict__ nrlo,
                                         const int32_t* __restrict__ nrhi,
                                         const double* __restrict__ spts, const uint4* __restrict__ xsq, int64_t R,
                                         const int (&jl)[NC], const float (&osc)[NC], FinF fin, lds_double* lds) {
    using P = PrecF32;
    const int i = lane & 15, kk = lane >> 4;
    auto run = [&](auto ph, auto rl) {
        constexpr bool PH = decltype(ph)::value, RL = decltype(rl)::value;
        if (nrows > 8) {
            flt4 Pa[NC], Qa[NC], Ia[NC];
#pragma unroll
            for (int c = 0; c < NC; ++c)
                for (int v = 0; v < 4; ++v) Pa[c][v] = Qa[c][v] = Ia[c][v] = 0.0f;
            h16_walk<16, NC, NEG, PH, RL>(Pa, Qa, Ia, nk, rowptr(i), rowptr(0), lane, e0, nr, rot, nrlo, nrhi, spts,
                                          xsq, R, jl, false, 0u, 0u, lds);
            Cacc<P, false> acc[NC];
#pragma unroll
            for (int c = 0; c < NC; ++c) {
                acc[c].p1 = Pa[c] * osc[c];
                acc[c].p2 = Qa[c] * osc[c];
                acc[c].p3 = Ia[c] * osc[c];
                for (int v = 0; v < 4; ++v) acc[c].p4[v] = 0.0f;
            }
            fin(acc, i);
        } else {
            auto emb = [&](auto hval) {
                constexpr int H = decltype(hval)::value;
                // v_perm(S0, S1): S1 the lower slot's source on real rows, the
                // higher one on imaginary rows (there the lane's own slot is the
                // higher); real rows take re halves for A1, imaginary rows im halves
                const bool imrow = (i / H) == 1;
                const uint32_t sel1 = imrow ? 0x03020706u : 0x05040100u;
                const uint32_t sel2 = imrow ? 0x01000504u : 0x07060302u;
                flt4 D1[NC], D2[NC];
#pragma unroll
                for (int c = 0; c < NC; ++c)
                    for (int v = 0; v < 4; ++v) D1[c][v] = D2[c][v] = 0.0f;
                h16_walk<H, NC, NEG, PH, RL>(D1, D2, D1, nk, rowptr(i % H), rowptr(0), lane, e0, nr, rot, nrlo, nrhi,
                                             spts, xsq, R, jl, imrow, sel1, sel2, lds);
                // this lane holds MFMA rows 4 kk + v: real parts when kk < H / 4
                const float sg = kk < H / 4 ? -1.0f : 1.0f;
                Cemb<P, H> acc[NC];
#pragma unroll
                for (int c = 0; c < NC; ++c)
                    for (int v = 0; v < 4; ++v) acc[c].d[v] = (D1[c][v] + sg * D2[c][v]) * osc[c];
                fin(acc, i % H);
            };
            if (nrows > 4) emb(std::integral_constant<int, 8>());
            else emb(std::integral_constant<int, 4>());
        }
    };
    if (nk.phase32 && rel) run(std::true_type(), std::true_type());
    else if (nk.phase32) run(std::true_type(), std::false_type());
    else run(std::false_type(), std::false_type());
}

// Hands every output entry of a tile's accumulators (16-row form or real
// embedding) to emitter(c)(row, value).
template <int NC, class EF>
__device__ __forceinline__ void emit_tile(Cacc<PrecF32, false> (&acc)[NC], EF emitter) {
    const int kk = (threadIdx.x & 63) >> 4;
#pragma unroll
    for (int c = 0; c < NC; ++c) {
        const auto emit = emitter(c);
#pragma unroll
        for (int v = 0; v < 4; ++v) emit(PrecF32::row(kk, v), acc[c].get(v));
    }
}
template <int HE, int NC, class EF>
__device__ __forceinline__ void emit_tile(Cemb<PrecF32, HE> (&acc)[NC], EF emitter) {
#pragma unroll
    for (int c = 0; c < NC; ++c) acc[c].store(emitter(c));
}

// Stage 5 + near field of MIXED products on the f16 path for 16 NC
// right-hand sides (the tiles, tile order and output of k_leaf_out_mfma;
// dpts / spts: the scaled point copies; xs16: the split right-hand sides of
// all R columns; xsc: per column s_q and 1 / s_q). The near sums are scaled
// back (t / (4 pi s_q)), then the basis rows are added on the f32 MFMA.
template <bool FWD, bool CJ, bool NEG, int NC>
__global__ void __launch_bounds__(256, NC == 1 ? 5 : 3)
    k_leaf_out_h16(const int32_t* __restrict__ tile_leaf, const int32_t* __restrict__ tile_row0, int64_t ntiles,
                   const int32_t* __restrict__ leaves, const int32_t* __restrict__ leafbasis,
                   const int64_t* __restrict__ boff, const float2* __restrict__ bblocks,
                   const int32_t* __restrict__ dlo, const int32_t* __restrict__ dhi,
                   const int32_t* __restrict__ dperm, const int32_t* __restrict__ dcptr,
                   const double* __restrict__ dpts, const int64_t* __restrict__ nrptr,
                   const int32_t* __restrict__ nrlo, const int32_t* __restrict__ nrhi,
                   const double* __restrict__ spts, const double2* __restrict__ yh, const uint4* __restrict__ xs16,
                   const float* __restrict__ xsc, int64_t R, int64_t q0, NearK nk, const uint8_t* __restrict__ nrel,
                   double2 alpha, double2 beta, double2* __restrict__ y, int64_t ldy, int order, int stagger,
                   int nq) {
    using P = PrecF32;
    const int lane = threadIdx.x & 63;
    const int w = __builtin_amdgcn_readfirstlane(work_block(order) * 4 + (int)(threadIdx.x >> 6));
    if (w >= ntiles) return;
    const int li = tile_leaf[w], row0 = tile_row0[w];
    const int t = leaves[li], bi = leafbasis[li];
    const int l0 = dlo[t], nt = dhi[t] - l0;
    const int nrows = min(16, nt - row0);
    const int kk = lane >> 4, j = lane & 15;
    const int64_t e0 = nrptr[t];
    const int nr = (int)(nrptr[t + 1] - e0);
    const int rot = stagger && nr ? (w * 7) % nr : 0;
    int jl[NC];
    float osc[NC];
#pragma unroll
    for (int c = 0; c < NC; ++c) {
        jl[c] = min(16 * c + j, nq - 1);
        osc[c] = nk.tsc * xsc[2 * (q0 + jl[c]) + 1];
    }
    const int rt = dcptr[t + 1] - dcptr[t];
    const float2* B = bi >= 0 ? bblocks + boff[bi] : nullptr;
    auto emitter = [&](int c) {
        return [&, c](int row, double2 v) {
            const int col = 16 * c + j;
            if (row < nrows && col < nq) {
                double2* out = y + (int64_t)dperm[l0 + row0 + row] + (q0 + col) * ldy;
                double2 val = cmul(alpha, v);
                if (beta.x != 0.0 || beta.y != 0.0) val = cadd(val, cmul(beta, *out));
                *out = val;
            }
        };
    };
    // then the basis rows: acc += A B with A = U_t (or V_t^T) rows, B = yh_t
    auto fin = [&](auto& acc, int rsel) {
        if (B) {
            const int rowc = row0 + min(rsel, nrows - 1);
            const double2* yt = yh + (int64_t)dcptr[t] * R + q0;
            for (int k0 = 0; k0 < rt; k0 += 4) {
                const int k = k0 + kk;
                const int kc = min(k, rt - 1);
                const float2 a =
                    conjif<CJ, P>(B[FWD ? (int64_t)rowc + (int64_t)kc * nt : (int64_t)kc + (int64_t)rowc * rt]);
#pragma unroll
                for (int c = 0; c < NC; ++c) {
                    float2 b = P::cvt(yt[(int64_t)kc * R + jl[c]]);
                    if (k >= rt) b = czero2<P>();
                    acc[c].mma(a, b);
                }
            }
        }
        emit_tile(acc, emitter);
    };
    auto rowptr = [&](int rsel) { return dpts + 3 * (int64_t)(l0 + row0 + min(rsel, nrows - 1)); };
    __shared__ double h16pts[4][128];
    const int wv = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    const bool rel = nk.rel32 && nrel[t];
    h16_tile<NC, NEG>(nk, rel, nrows, lane, rowptr, e0, nr, rot, nrlo, nrhi, spts, xs16 + q0, R, jl, osc, fin,
                      (lds_double*)h16pts[wv]);
}

// Stage 3 of MIXED products on the f16 path, the coupling blocks evaluated
// on the fly (block (t, s) = G(test pivots of t, trial pivots of s), the
// store's entries): one wave per tile of at most 16 rows of a destination
// block, its rows the scaled coordinates of the block's pivots (dpiv, in
// coefficient order), its sources the pivots of its far blocks as merged
// coefficient ranges (crptr / crlo / crhi, bit 31 of the end: the guard) of
// spiv, B = the split source coefficients xs16 (R per coefficient) with the
// chunk's column scales xsc (relative to q0). The tiles and the output are
// those of k_couple_fwd_mfma (yh, or yfin in FP64 for the blocks fin marks).
template <bool NEG, int NC, class YT>
__global__ void __launch_bounds__(256, NC == 1 ? 5 : 3)
    k_couple_h16(const int32_t* __restrict__ tile_node, const int32_t* __restrict__ tile_row0, int64_t ntiles,
                 const int32_t* __restrict__ cptr_dst, const double* __restrict__ dpiv,
                 const int64_t* __restrict__ crptr, const int32_t* __restrict__ crlo,
                 const int32_t* __restrict__ crhi, const double* __restrict__ spiv, const uint4* __restrict__ xs16,
                 const float* __restrict__ xsc, int64_t R, int64_t q0, int nq, NearK nk,
                 const uint8_t* __restrict__ crel, YT* __restrict__ yh, double2* __restrict__ yfin,
                 const uint8_t* __restrict__ fin, int order) {
    const int lane = threadIdx.x & 63;
    const int w = __builtin_amdgcn_readfirstlane(work_block(order) * 4 + (int)(threadIdx.x >> 6));
    if (w >= ntiles) return;
    const int t = tile_node[w], row0 = tile_row0[w];
    if (t < 0) return;  // padding tile
    const int r0 = cptr_dst[t], rt = cptr_dst[t + 1] - r0;
    const int nrows = min(16, rt - row0);
    const int j = lane & 15;
    const int64_t e0 = crptr[t];
    const int nr = (int)(crptr[t + 1] - e0);
    int jl[NC];
    float osc[NC];
#pragma unroll
    for (int c = 0; c < NC; ++c) {
        jl[c] = min(16 * c + j, nq - 1);
        osc[c] = nk.tsc * xsc[2 * jl[c] + 1];
    }
    const int64_t o0 = (int64_t)(r0 + row0) * R + q0 + j;
    const bool tofin = fin && fin[t];
    auto emitter = [&](int c) {
        return [&, c](int row, double2 v) {
            if (row < nrows && 16 * c + j < nq) {
                const int64_t o = o0 + 16 * c + (int64_t)row * R;
                if (tofin) yfin[o] = v;
                else store_as(yh + o, v);
            }
        };
    };
    auto fn = [&](auto& acc, int) { emit_tile(acc, emitter); };
    auto rowptr = [&](int rsel) { return dpiv + 3 * (int64_t)(r0 + row0 + min(rsel, nrows - 1)); };
    __shared__ double h16pts[4][128];
    const int wv = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    const bool rel = nk.rel32 && crel[t];
    h16_tile<NC, NEG>(nk, rel, nrows, lane, rowptr, e0, nr, 0, crlo, crhi, spiv, xs16 + q0, R, jl, osc, fn,
                      (lds_double*)h16pts[wv]);
}

// Squared-distance statistics of one tile (one wave; lanes over the sources
// of its merged ranges e0 .. e1 - 1 of spts, the rows xrow + 3 r, r < nrows,
// in a loop) to out[0 .. 3]: the smallest pair above the zero-distance guard
// (2e-30; +inf without one), the largest pair (0 without one), the smallest
// nonzero pair (+inf without one) and the largest squared distance of a row
// from the first row.
__device__ __forceinline__ void tile_pair_stats(const double* __restrict__ xrow, int nrows, int64_t e0, int64_t e1,
                                                const int32_t* __restrict__ rlo, const int32_t* __restrict__ rhi,
                                                const double* __restrict__ spts, int lane, double* __restrict__ out) {
    double m = HUGE_VAL, mx = 0.0, mp = HUGE_VAL;
    for (int64_t e = e0; e < e1; ++e) {
        const int lo = rlo[e], hi = rhi[e] & 0x7fffffff;
        for (int s = lo + lane; s < hi; s += 64) {
            const double y0 = spts[3 * (int64_t)s], y1 = spts[3 * (int64_t)s + 1], y2 = spts[3 * (int64_t)s + 2];
            for (int r = 0; r < nrows; ++r) {
                const double* xp = xrow + 3 * (int64_t)r;
                const double dx = xp[0] - y0, dy = xp[1] - y1, dz = xp[2] - y2;
                const double r2 = fma(dz, dz, fma(dy, dy, dx * dx));
                if (r2 > 2e-30) m = fmin(m, r2);
                if (r2 > 0.0) mp = fmin(mp, r2);
                mx = fmax(mx, r2);
            }
        }
    }
    double ext = 0.0;
    for (int r = lane; r < nrows; r += 64) {
        const double dx = xrow[3 * (int64_t)r] - xrow[0], dy = xrow[3 * (int64_t)r + 1] - xrow[1],
                     dz = xrow[3 * (int64_t)r + 2] - xrow[2];
        ext = fmax(ext, fma(dz, dz, fma(dy, dy, dx * dx)));
    }
    for (int o = 32; o > 0; o >>= 1) {
        m = fmin(m, __shfl_xor(m, o));
        mx = fmax(mx, __shfl_xor(mx, o));
        mp = fmin(mp, __shfl_xor(mp, o));
        ext = fmax(ext, __shfl_xor(ext, o));
    }
    if (lane == 0) {
        out[0] = m;
        out[1] = mx;
        out[2] = mp;
        out[3] = ext;
    }
}

// tile_pair_stats of each coupling tile (rows: the tile's pivots, dpiv +
// 3 (cptr[t] + row), sources: spiv over its merged ranges; one wave per
// tile): out[4 w .. 4 w + 3] (+inf, 0, +inf, 0 for a padding tile).
__global__ void __launch_bounds__(256)
    k_pair_rminmax(const int32_t* __restrict__ tile_node, const int32_t* __restrict__ tile_row0, int64_t ntiles,
                   const int32_t* __restrict__ cptr_dst, const double* __restrict__ dpiv,
                   const int64_t* __restrict__ crptr, const int32_t* __restrict__ crlo,
                   const int32_t* __restrict__ crhi, const double* __restrict__ spiv, double* __restrict__ out) {
    const int lane = threadIdx.x & 63;
    const int64_t w = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    if (w >= ntiles) return;
    const int t = tile_node[w];
    if (t < 0) {
        if (lane == 0) {
            out[4 * w] = HUGE_VAL;
            out[4 * w + 1] = 0.0;
            out[4 * w + 2] = HUGE_VAL;
            out[4 * w + 3] = 0.0;
        }
        return;
    }
    const int r0 = cptr_dst[t] + tile_row0[w], nrows = min(16, cptr_dst[t + 1] - r0);
    tile_pair_stats(dpiv + 3 * (int64_t)r0, nrows, crptr[t], crptr[t + 1], crlo, crhi, spiv, lane, out + 4 * w);
}

// Per-workgroup column maxima of |re|, |im| of a c32 [n][R] array over
// columns q0 .. q0 + nc - 1 (64 rows x 16 columns per workgroup; lanes along
// the columns): part[blockIdx.x nc + c].
__global__ void __launch_bounds__(256)
    k_colmax_c32(const float2* __restrict__ in, int64_t n, int64_t R, int64_t q0, int nc, float* __restrict__ part) {
    __shared__ float red[16][17];
    const int t = threadIdx.x, cl = t & 15, rg = t >> 4;
    const int c = (int)blockIdx.y * 16 + cl;
    float m = 0.0f;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        const int64_t row = (int64_t)blockIdx.x * 64 + rg + 16 * r;
        if (row < n && c < nc) {
            const float2 v = in[row * R + q0 + c];
            m = fmaxf(m, fmaxf(fabsf(v.x), fabsf(v.y)));
        }
    }
    red[rg][cl] = m;
    __syncthreads();
    if (t < 16) {
        float a = red[0][t];
        for (int q = 1; q < 16; ++q) a = fmaxf(a, red[q][t]);
        const int cc = (int)blockIdx.y * 16 + t;
        if (cc < nc) part[(int64_t)blockIdx.x * nc + cc] = a;
    }
}

// tile_pair_stats of each near-field destination tile (one wave per tile):
// out[4 w .. 4 w + 3].
__global__ void __launch_bounds__(256)
    k_near_rmin(const int32_t* __restrict__ tile_leaf, const int32_t* __restrict__ tile_row0, int64_t ntiles,
                const int32_t* __restrict__ leaves, const int32_t* __restrict__ dlo, const int32_t* __restrict__ dhi,
                const double* __restrict__ dpts, const int64_t* __restrict__ nrptr,
                const int32_t* __restrict__ nrlo, const int32_t* __restrict__ nrhi, const double* __restrict__ spts,
                double* __restrict__ out) {
    const int lane = threadIdx.x & 63;
    const int64_t w = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    if (w >= ntiles) return;
    const int t = leaves[tile_leaf[w]], row0 = tile_row0[w];
    const int l0 = dlo[t], nrows = min(16, dhi[t] - l0 - row0);
    tile_pair_stats(dpts + 3 * (int64_t)(l0 + row0), nrows, nrptr[t], nrptr[t + 1], nrlo, nrhi, spts, lane,
                    out + 4 * w);
}

// The forward coupling's column map: for every test node t with columns,
// out[colbase[t] ..] = the source coefficients of its blocks in block order
// (block i: trial node s = nodes[i], coefficients scptr[s] .. scptr[s] +
// srank[s] - 1). One wave per test node; a wave scan places 64 blocks at a
// time.
__global__ void __launch_bounds__(256)
    k_colcoef(int64_t ntest, const int64_t* __restrict__ ptr, const int32_t* __restrict__ nodes,
              const int32_t* __restrict__ srank, const int32_t* __restrict__ scptr,
              const int64_t* __restrict__ colbase, const int32_t* __restrict__ ncols, int32_t* __restrict__ out) {
    const int lane = threadIdx.x & 63;
    const int64_t t = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    if (t >= ntest || ncols[t] == 0) return;
    int64_t c = colbase[t];
    const int64_t i1 = ptr[t + 1];
    for (int64_t i0 = ptr[t]; i0 < i1; i0 += 64) {
        const int64_t i = i0 + lane;
        int r = 0, q = 0;
        if (i < i1) {
            const int sn = nodes[i];
            r = srank[sn];
            q = scptr[sn];
        }
        int incl = r;
        for (int o = 1; o < 64; o <<= 1) {
            const int v = __shfl_up(incl, o);
            if (lane >= o) incl += v;
        }
        for (int j = 0; j < r; ++j) out[c + (incl - r) + j] = q + j;
        c += __shfl(incl, 63);
    }
}

// out[3 k + d] = pts[3 pos[k] + d] (k < K), zeros for k in [K, K + pad).
__global__ void k_gather_pivots(const double* __restrict__ pts, const int32_t* __restrict__ pos, int64_t K,
                                double* __restrict__ out) {
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= K + kNearPad) return;
    for (int d = 0; d < 3; ++d) out[3 * k + d] = k < K ? pts[3 * (int64_t)pos[k] + d] : 0.0;
}

// out = in * t (points, a power of two: exact)
__global__ void k_scale_pts(const double* __restrict__ in, double* __restrict__ out, int64_t n, double t) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = in[i] * t;
}

// Per column q of the gathered right-hand sides: the largest |re|, |im| over
// the per-workgroup maxima of the gather (nparts per column), then s_q = 2^m
// with that maximum times s_q in [2^14, 2^15) (1 for a zero or non-finite
// maximum): xsc[2 q] = s_q, xsc[2 q + 1] = 1 / s_q. One workgroup per column.
__global__ void __launch_bounds__(256)
    k_colscale(const float* __restrict__ part, int64_t nparts, int64_t R, float* __restrict__ xsc) {
    __shared__ float red[256];
    const int64_t q = blockIdx.x;
    float m = 0.0f;
    for (int64_t b = threadIdx.x; b < nparts; b += 256) m = fmaxf(m, part[b * R + q]);
    red[threadIdx.x] = m;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) red[threadIdx.x] = fmaxf(red[threadIdx.x], red[threadIdx.x + s]);
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        const float mx = red[0];
        int ex = 0;
        if (mx > 0.0f && mx <= 3.4e38f) {
            (void)frexpf(mx, &ex);  // mx = f 2^ex, f in [0.5, 1)
            ex = 15 - ex;
            ex = ex > 100 ? 100 : (ex < -100 ? -100 : ex);
        }
        xsc[2 * q] = ldexpf(1.0f, ex);
        xsc[2 * q + 1] = ldexpf(1.0f, -ex);
    }
}

// out[p R + q0 + c] = the f16 parts of s_c in[p R + q0 + c] and of s_c
// in[(p + 4) R + q0 + c] (zero for p + 4 >= n) as the f16 walk's B slot
// pairs: (re high of p, of p + 4), (im high ...), (re low ...), (im low ...)
// (p < n, c < nc; s_c = xsc[2 c]).
__global__ void k_split16(const float2* __restrict__ in, const float* __restrict__ xsc, int64_t n, int64_t R,
                          int64_t q0, int nc, uint4* __restrict__ out) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n * nc) return;  // n R < 2^31 (checked per product)
    const uint32_t p = (uint32_t)i / (uint32_t)nc, c = (uint32_t)i % (uint32_t)nc;
    const int64_t o = (int64_t)p * R + q0 + c;
    const float s = xsc[2 * c];
    const float2 v = in[o];
    uint32_t h0, l0, h1 = 0u, l1 = 0u;
    split_h16(v.x * s, v.y * s, h0, l0);
    if ((int64_t)p + 4 < n) {
        const float2 w = in[o + 4 * R];
        split_h16(w.x * s, w.y * s, h1, l1);
    }
    out[o] = make_uint4(pack_lo(h0, h1), pack_hi(h0, h1), pack_lo(l0, l1), pack_hi(l0, l1));
}

// Stage 5 + near field for NQ = 1..3 right-hand sides on VALU (an MFMA
// product would pad them to 16 columns): the tiles, merged source ranges and
// point broadcasts of k_leaf_out_mfma. Lane (i, kk) = (lane % 16, lane / 16)
// evaluates G for row i % H and source s = 4 (i / H) + kk of each step of
// S = 64 / H sources and accumulates G x_s for its own source (the
// right-hand sides of sources past the range end read 0 and are masked);
// the basis rows are spread over the same lanes (row i % H, basis columns
// s, s + S, ...); the lanes of one row are summed by shuffles at the end.
// Accumulation in the precision's type (FP32 for MIXED, as in its MFMA tiles).
template <class P, int H, bool NEG, int NQ>
__device__ __forceinline__ void small_walk(typename P::T2 (&acc)[NQ], const NearK& nk, const double2* tab, int lane,
                                           double x0, double x1, double x2, int64_t e0, int64_t e1, int rot,
                                           const int32_t* __restrict__ nrlo, const int32_t* __restrict__ nrhi,
                                           const double* __restrict__ spts, const typename P::T2* __restrict__ xsq,
                                           uint32_t R) {
    using T2 = typename P::T2;
    using T = typename P::T;
    constexpr int S = 4 * (16 / H), G = 16 / H;
    constexpr uint32_t kT = sizeof(T2);
    const int kk = lane >> 4, ysrc = 4 * ((lane & 15) / H) + kk;
    const uint32_t vp = near_point_offset<G>(lane, kk, ysrc);
    const uint32_t rowb = R * kT, vx = (uint32_t)ysrc * rowb;
    const int nr = (int)(e1 - e0);
    for (int c = 0; c < nr; ++c) {
        int ec = c + (nr ? rot % nr : 0);
        if (ec >= nr) ec -= nr;
        const int64_t e = e0 + ec;
        const int lo = nrlo[e], hr = nrhi[e];
        const int hi = hr & 0x7fffffff;
        const uint32_t n = (uint32_t)(hi - lo);
        const __amdgpu_buffer_rsrc_t rp = buf_rsrc(spts + 3 * (int64_t)lo, n * 24u);
        const __amdgpu_buffer_rsrc_t rx = buf_rsrc(xsq + (int64_t)lo * R, n * rowb);
        auto steps = [&](auto guard, auto ph32) {
            constexpr bool GU = decltype(guard)::value, PH = decltype(ph32)::value;
            int left = (int)n;
            uint32_t sp = 0, sx = 0;
            for (; left > 0; left -= S, sp += S * 24u, sx += S * rowb) {
                T2 b[NQ];
#pragma unroll
                for (int q = 0; q < NQ; ++q) b[q] = buf_load<T2>(rx, vx + (uint32_t)q * kT, sx);
                T2 g;
                const double pq = buf_load<double>(rp, vp, sp);
                if constexpr (G == 1)
                    g = P::template eval<NEG, GU, PH>(nk, tab, x0, x1, x2, row_double<0>(pq), row_double<1>(pq),
                                                      row_double<2>(pq));
                else
                    g = P::template eval<NEG, GU, PH>(nk, tab, x0, x1, x2, quad_double<0>(pq), quad_double<1>(pq),
                                                      quad_double<2>(pq));
                if (ysrc >= left) g = czero2<P>();
#pragma unroll
                for (int q = 0; q < NQ; ++q) {
                    acc[q].x = fma(g.x, b[q].x, fma(-g.y, b[q].y, acc[q].x));
                    acc[q].y = fma(g.x, b[q].y, fma(g.y, b[q].x, acc[q].y));
                }
            }
        };
        near_dispatch<P>(steps, hr < 0, nk.phase32 != 0);
    }
    (void)sizeof(T);
}

template <class P, bool FWD, bool CJ, bool NEG, int NQ>
__global__ void __launch_bounds__(256)
    k_leaf_out_small(const int32_t* __restrict__ tile_leaf, const int32_t* __restrict__ tile_row0, int64_t ntiles,
                     const int32_t* __restrict__ leaves, const int32_t* __restrict__ leafbasis,
                     const int64_t* __restrict__ boff, const typename P::T2* __restrict__ bblocks,
                     const int32_t* __restrict__ dlo, const int32_t* __restrict__ dhi,
                     const int32_t* __restrict__ dperm, const int32_t* __restrict__ dcptr,
                     const double* __restrict__ dpts, const int64_t* __restrict__ nrptr,
                     const int32_t* __restrict__ nrlo, const int32_t* __restrict__ nrhi,
                     const double* __restrict__ spts, const double2* __restrict__ yh,
                     const typename P::T2* __restrict__ xs, int64_t R, int64_t q0, NearK nk,
                     const double2* __restrict__ sctab, double2 alpha, double2 beta, double2* __restrict__ y,
                     int64_t ldy, int order, int stagger, const double2* __restrict__ ynear) {
    using T2 = typename P::T2;
    using T = typename P::T;
    __shared__ double2 tab[P::kTab ? kSinCosTab : 1];
    if (P::kTab) {
        tab[threadIdx.x] = sctab[threadIdx.x];
        __syncthreads();
    }
    const int lane = threadIdx.x & 63;
    const int w = __builtin_amdgcn_readfirstlane(work_block(order) * 4 + (int)(threadIdx.x >> 6));
    if (w >= ntiles) return;
    const int li = tile_leaf[w], row0 = tile_row0[w];
    const int t = leaves[li], bi = leafbasis[li];
    const int l0 = dlo[t], nt = dhi[t] - l0;
    const int nrows = min(16, nt - row0);
    const int64_t e0 = nrptr[t], e1 = nrptr[t + 1];
    const int rot = stagger ? w * 7 : 0;
    const int rt = dcptr[t + 1] - dcptr[t];
    const T2* B = bi >= 0 ? bblocks + boff[bi] : nullptr;
    auto tile = [&](auto hval) {
        constexpr int H = decltype(hval)::value;
        constexpr int S = 4 * (16 / H);
        const int rr = (lane & 15) % H, slot = 4 * ((lane & 15) / H) + (lane >> 4);
        const int rowc = row0 + min(rr, nrows - 1);
        T2 acc[NQ];
#pragma unroll
        for (int q = 0; q < NQ; ++q) acc[q] = czero2<P>();
        if (!ynear) {
            const double* xp = dpts + 3 * (int64_t)(l0 + rowc);
            small_walk<P, H, NEG, NQ>(acc, nk, tab, lane, xp[0], xp[1], xp[2], e0, e1, rot, nrlo, nrhi, spts,
                                      xs + q0, (uint32_t)R);
#pragma unroll
            for (int q = 0; q < NQ; ++q) {
                acc[q].x *= (T)kInvFourPi;
                acc[q].y *= (T)kInvFourPi;
            }
        }
        if (B) {
            const double2* yt = yh + (int64_t)dcptr[t] * R + q0;
            for (int k = slot; k < rt; k += S) {
                const T2 a = conjif<CJ, P>(B[FWD ? (int64_t)rowc + (int64_t)k * nt : (int64_t)k + (int64_t)rowc * rt]);
#pragma unroll
                for (int q = 0; q < NQ; ++q) {
                    const T2 v = P::cvt(yt[(int64_t)k * R + q]);
                    acc[q].x = fma(a.x, v.x, fma(-a.y, v.y, acc[q].x));
                    acc[q].y = fma(a.x, v.y, fma(a.y, v.x, acc[q].y));
                }
            }
        }
        // sum the lanes of each row: over kk (lanes +16, +32) and over the
        // source groups of a 16-lane row (lanes +H, +2H, ...)
#pragma unroll
        for (int q = 0; q < NQ; ++q) {
#pragma unroll
            for (int m = H; m < 64; m *= 2) {
                acc[q].x += __shfl_xor(acc[q].x, m);
                acc[q].y += __shfl_xor(acc[q].y, m);
            }
        }
        if (lane < H && rr < nrows) {
            double2* out = y + (int64_t)dperm[l0 + row0 + rr] + q0 * ldy;
#pragma unroll
            for (int q = 0; q < NQ; ++q) {
                double2 v = make_double2((double)acc[q].x, (double)acc[q].y);
                if (ynear) v = cadd(v, ynear[(int64_t)(l0 + row0 + rr) * R + q0 + q]);
                double2 val = cmul(alpha, v);
                if (beta.x != 0.0 || beta.y != 0.0) val = cadd(val, cmul(beta, out[q * ldy]));
                out[q * ldy] = val;
            }
        }
    };
    if (nrows > 8) tile(std::integral_constant<int, 16>());
    else if (nrows > 4) tile(std::integral_constant<int, 8>());
    else tile(std::integral_constant<int, 4>());
}

// Stage 3 forward, one tile: acc[c] += (coupling row of the tile) (source
// coefficients of column chunk c), four columns per step. Software pipeline
// over groups of kU steps, two register sets in turn: the operands of group
// n + 1 and the coefficient indices of group n + 2 are in flight while the
// MFMAs of group n run. Sb: the tile's r_t x C block (column-major, ur = r_t rows), rowsel:
// the lane's row. Columns past C read A = 0 (and source index 0), so the
// pipeline needs neither clamps nor masks.
template <class P, bool CJ, int NC, class ACC, class XT>
__device__ __forceinline__ void couple_run(ACC (&acc)[NC], const typename P::T2* __restrict__ Sb, uint32_t ur,
                                           uint32_t rowsel, const int32_t* __restrict__ cc, int C,
                                           const XT* const (&xq)[NC], uint32_t xstride, int kk) {
    using T2 = typename P::T2;
    // steps per group: two chunks take half the depth (a quarter in FP64), for registers
    constexpr int kU = NC == 1 ? 4 : (P::kF64 ? 1 : 2);
    constexpr uint32_t kA = sizeof(T2);
    ur = __builtin_amdgcn_readfirstlane(ur);
    C = __builtin_amdgcn_readfirstlane(C);
    const __amdgpu_buffer_rsrc_t ra =
        __builtin_amdgcn_make_buffer_rsrc((void*)Sb, (short)0, (int)((uint32_t)C * ur * kA), kBufDword3);
    const __amdgpu_buffer_rsrc_t rc = __builtin_amdgcn_make_buffer_rsrc((void*)cc, (short)0, C * 4, kBufDword3);
    const uint32_t va = ((uint32_t)kk * ur + rowsel) * kA, vc = (uint32_t)kk * 4u;
    T2 a0[kU], a1[kU];
    XT b0[NC][kU], b1[NC][kU];
    int ci0[kU], ci1[kU];
    auto load_idx = [&](int g0, int (&ci)[kU]) {
#pragma unroll
        for (int u = 0; u < kU; ++u) ci[u] = buf_load<int>(rc, vc, (uint32_t)(g0 + 4 * u) * 4u);
    };
    auto load_ops = [&](int g0, const int (&ci)[kU], T2 (&a)[kU], XT (&b)[NC][kU]) {
#pragma unroll
        for (int u = 0; u < kU; ++u) {
            a[u] = buf_load<T2>(ra, va, (uint32_t)(g0 + 4 * u) * ur * kA);
#pragma unroll
            for (int c = 0; c < NC; ++c) b[c][u] = xq[c][(uint32_t)ci[u] * xstride];
        }
    };
    auto mma = [&](const T2 (&a)[kU], const XT (&b)[NC][kU]) {
#pragma unroll
        for (int u = 0; u < kU; ++u)
#pragma unroll
            for (int c = 0; c < NC; ++c) acc[c].mma(conjif<CJ, P>(a[u]), P::cvt(b[c][u]));
    };
    load_idx(0, ci0);
    load_ops(0, ci0, a0, b0);
    load_idx(4 * kU, ci1);
    for (int c0 = 0; c0 < C; c0 += 8 * kU) {
        load_ops(c0 + 4 * kU, ci1, a1, b1);
        load_idx(c0 + 8 * kU, ci0);
        __builtin_amdgcn_sched_barrier(0);
        mma(a0, b0);
        if (c0 + 4 * kU >= C) break;
        load_ops(c0 + 8 * kU, ci0, a0, b0);
        load_idx(c0 + 12 * kU, ci1);
        __builtin_amdgcn_sched_barrier(0);
        mma(a1, b1);
    }
}

// Stage 3 forward for 16 NC right-hand sides (NC chunks of 16 columns: each
// coupling entry loaded once serves all of them): one wave per tile of at
// most 16 rows of a test cluster; A = the r_t x C coupling row, B = the source
// coefficients of its columns. Tiles of more than 8 rows use the complex
// (three- or four-product) accumulator, smaller ones the real embedding.
// Control flow is uniform over the wave; offsets inside a block and into xh
// are 32-bit (coefficients x right-hand sides < 2^31). XT: type of the source
// coefficients (the operator's, or a compact c32 copy in MIXED).
template <class P, bool CJ, bool K3, class XT, class YT, int NC>
__global__ void __launch_bounds__(256, 4)
    k_couple_fwd_mfma(const typename P::T2* __restrict__ S, const int32_t* __restrict__ tile_node,
                      const int32_t* __restrict__ tile_row0, int64_t ntiles, const int64_t* __restrict__ base,
                      const int64_t* __restrict__ colbase, const int32_t* __restrict__ ncols,
                      const int32_t* __restrict__ colcoef, const int32_t* __restrict__ cptr_dst,
                      const XT* __restrict__ xq0, uint32_t xstride, int64_t R, int64_t q0, int nq,
                      YT* __restrict__ yh, double2* __restrict__ yfin, const uint8_t* __restrict__ fin,
                      int order) {
    const int lane = threadIdx.x & 63;
    const int w = __builtin_amdgcn_readfirstlane(work_block(order) * 4 + (int)(threadIdx.x >> 6));
    if (w >= ntiles) return;
    const int t = tile_node[w], row0 = tile_row0[w];
    if (t < 0) return;  // padding tile
    const int rt = cptr_dst[t + 1] - cptr_dst[t];
    const int C = ncols[t];
    const int nrows = min(16, rt - row0);
    const int i = lane & 15, kk = lane >> 4, j = lane & 15;
    const int32_t* cc = colcoef + colbase[t];
    // columns past the live ones read the last live one (not stored)
    const XT* xq[NC];
#pragma unroll
    for (int c = 0; c < NC; ++c) xq[c] = xq0 + min(16 * c + j, nq - 1);
    // every row of every test block with coefficients has one tile (rows
    // without far columns store zeros), so the output needs no clearing
    // fin[t]: a block no downward transfer writes takes its final value here, in FP64
    const int64_t o0 = (int64_t)(cptr_dst[t] + row0) * R + q0 + j;
    const bool tofin = fin && fin[t];
    auto emitter = [&](int c) {
        return [&, c](int row, double2 v) {
            if (row < nrows && 16 * c + j < nq) {
                const int64_t o = o0 + 16 * c + (int64_t)row * R;
                if (tofin) yfin[o] = v;
                else store_as(yh + o, v);
            }
        };
    };
    auto run = [&](auto& acc, int rowsel) {
#pragma unroll
        for (int c = 0; c < NC; ++c) acc[c].zero();
        if (C > 0)
            couple_run<P, CJ, NC>(acc, S + base[t], (uint32_t)rt, (uint32_t)(row0 + min(rowsel, nrows - 1)), cc, C,
                                  xq, xstride, kk);
    };
    if (nrows > 8) {
        Cacc<P, K3> acc[NC];
        run(acc, i);
#pragma unroll
        for (int c = 0; c < NC; ++c) {
            const auto emit = emitter(c);
#pragma unroll
            for (int v = 0; v < 4; ++v) emit(P::row(kk, v), acc[c].get(v));
        }
    } else if (nrows > 4) {
        Cemb<P, 8> acc[NC];
        run(acc, i & 7);
#pragma unroll
        for (int c = 0; c < NC; ++c) acc[c].store(emitter(c));
    } else {
        Cemb<P, 4> acc[NC];
        run(acc, i & 3);
#pragma unroll
        for (int c = 0; c < NC; ++c) acc[c].store(emitter(c));
    }
}

// Stages 1, 2 and 4 for 16 right-hand sides as small MFMA GEMMs, one wave per
// node (4 per workgroup): out (m x 16) += A (m x K) B (K x 16) over 16-row
// tiles of m and 4-wide steps of K. A rows past m repeat row m - 1 (not
// stored); B rows past K are zero. Control flow is uniform over the wave.
template <class P, bool K3, class AF, class BF>
__device__ __forceinline__ void gemm16_tile(Cacc<P, K3>& acc, int lane, int a0, int m, int K, AF afun, BF bfun) {
    using T2 = typename P::T2;
    const int i = lane & 15, kk = lane >> 4;
    const int ar = a0 + min(i, m - 1 - a0);
    for (int k0 = 0; k0 < K; k0 += 4) {
        const int k = k0 + kk, kc = min(k, K - 1);
        const T2 a = afun(ar, kc);
        T2 b = bfun(kc, i);
        if (k >= K) b = czero2<P>();
        acc.mma(a, b);
    }
}

// out (+)= the tile: add into out, or store (the caller writes every element once)
template <class P, bool K3, class CT>
__device__ __forceinline__ void gemm16_store(const Cacc<P, K3>& acc, int lane, int a0, int m, CT* out,
                                             int64_t R, int nq, bool add) {
    const int kk = lane >> 4, j = lane & 15;
#pragma unroll
    for (int v = 0; v < 4; ++v) {
        const int row = a0 + P::row(kk, v);
        if (row < m && j < nq) {
            CT* o = out + (int64_t)row * R + j;
            store_as(o, add ? cadd(to_d2(*o), acc.get(v)) : acc.get(v));
        }
    }
}

// out = base + the tile (FP64 sum of an FP32 tile result and a c32 base of
// the same layout), for the rows < m and columns < nq.
template <class P, bool K3>
__device__ __forceinline__ void gemm16_store_base(const Cacc<P, K3>& acc, int lane, int a0, int m, double2* out,
                                                  const float2* base, int64_t R, int nq) {
    const int kk = lane >> 4, j = lane & 15;
#pragma unroll
    for (int v = 0; v < 4; ++v) {
        const int row = a0 + P::row(kk, v);
        if (row < m && j < nq) {
            const int64_t o = (int64_t)row * R + j;
            out[o] = cadd(to_d2(base[o]), acc.get(v));
        }
    }
}

// Stage 1: xh_s = V_s xs_s (TR: U_s^T). One wave per 16-row tile of a source
// basis (ptile: basis index, first row), the tiles of a basis side by side in
// one workgroup where they fit, so its source rows are read from the same L1;
// the tiles cover every coefficient of the node, so xh needs no clearing.
template <class P, bool TR, bool CJ, bool K3>
__global__ void __launch_bounds__(256)
    k_project_mfma(const int32_t* __restrict__ ptile, int64_t ntiles, const int32_t* __restrict__ bnodes,
                   const int64_t* __restrict__ boff, const typename P::T2* __restrict__ blocks,
                   const int32_t* __restrict__ lo, const int32_t* __restrict__ hi, const int32_t* __restrict__ cptr,
                   const typename P::T2* __restrict__ xs, int64_t R, int64_t q0, int nq,
                   typename P::T2* __restrict__ xh) {
    using T2 = typename P::T2;
    const int lane = threadIdx.x & 63;
    const int w = __builtin_amdgcn_readfirstlane((int)(blockIdx.x * 4 + (threadIdx.x >> 6)));
    if (w >= ntiles) return;
    const int b = ptile[2 * w], a0 = ptile[2 * w + 1];
    const int s = bnodes[b];
    const int r = cptr[s + 1] - cptr[s];
    const int l0 = lo[s], ns = hi[s] - l0;
    const T2* B = blocks + boff[b];
    const T2* xv = xs + (int64_t)l0 * R + q0;
    Cacc<P, K3> acc;
    acc.zero();
    gemm16_tile<P, K3>(
        acc, lane, a0, r, ns,
        [&](int a, int j) { return conjif<CJ, P>(B[TR ? (int64_t)j + (int64_t)a * ns : (int64_t)a + (int64_t)j * r]); },
        [&](int j, int q) { return xv[(int64_t)j * R + min(q, nq - 1)]; });
    gemm16_store<P, K3>(acc, lane, a0, r, xh + (int64_t)cptr[s] * R + q0, R, nq, false);
}

// Stage 2 (up): xh_p = sum_c E_c xh_c, E_c r_p x r_c (TR: r_c x r_p used
// transposed; accumulate: += instead, for the child-centric downward pass);
// stage 4 (DOWN): yh_c += E_c yh_p, E_c r_c x r_p (TR: r_p x r_c transposed).
// One wave per parent of the level. base (MIXED forward downward pass): the
// c32 coupling result of the written blocks; with it, out = base + the tile
// (instead of +=), summed and stored in FP64.
template <class P, bool DOWN, bool TR, bool CJ, bool K3, class CT>
__global__ void __launch_bounds__(256)
    k_transfer_mfma(const int32_t* __restrict__ parents, const int64_t* __restrict__ chdptr,
                    const int32_t* __restrict__ chds, const int64_t* __restrict__ toff,
                    const typename P::T2* __restrict__ blocks, const int32_t* __restrict__ cptr, int64_t pbase,
                    int64_t np, int64_t R, int64_t q0, int nq, int accumulate, CT* __restrict__ h,
                    const float2* __restrict__ base) {
    using T2 = typename P::T2;
    const int lane = threadIdx.x & 63;
    const int w = __builtin_amdgcn_readfirstlane((int)(blockIdx.x * 4 + (threadIdx.x >> 6)));
    if (w >= np) return;
    const int64_t pi = pbase + w;
    const int p = parents[pi];
    const int rp = cptr[p + 1] - cptr[p];
    if (rp == 0) return;
    const int64_t e0 = chdptr[pi], e1 = chdptr[pi + 1];
    if (DOWN) {
        const CT* hp = h + (int64_t)cptr[p] * R + q0;
        for (int64_t e = e0; e < e1; ++e) {
            const int c = chds[e];
            const int rc = cptr[c + 1] - cptr[c];
            const T2* E = blocks + toff[e];
            for (int a0 = 0; a0 < rc; a0 += 16) {
                Cacc<P, K3> acc;
                acc.zero();
                gemm16_tile<P, K3>(
                    acc, lane, a0, rc, rp,
                    [&](int a, int b) { return conjif<CJ, P>(E[TR ? (int64_t)b + (int64_t)a * rp : (int64_t)a + (int64_t)b * rc]); },
                    [&](int b, int q) { return P::cvt(hp[(int64_t)b * R + min(q, nq - 1)]); });
                if constexpr (std::is_same_v<CT, double2>) {
                    if (base) {
                        gemm16_store_base<P, K3>(acc, lane, a0, rc, h + (int64_t)cptr[c] * R + q0,
                                                 base + (int64_t)cptr[c] * R + q0, R, nq);
                        continue;
                    }
                }
                gemm16_store<P, K3>(acc, lane, a0, rc, h + (int64_t)cptr[c] * R + q0, R, nq, true);
            }
        }
    } else {
        for (int a0 = 0; a0 < rp; a0 += 16) {
            Cacc<P, K3> acc;
            acc.zero();
            for (int64_t e = e0; e < e1; ++e) {
                const int c = chds[e];
                const int rc = cptr[c + 1] - cptr[c];
                if (rc == 0) continue;
                const T2* E = blocks + toff[e];
                const CT* hc = h + (int64_t)cptr[c] * R + q0;
                gemm16_tile<P, K3>(
                    acc, lane, a0, rp, rc,
                    [&](int a, int b) { return conjif<CJ, P>(E[TR ? (int64_t)b + (int64_t)a * rc : (int64_t)a + (int64_t)b * rp]); },
                    [&](int b, int q) { return P::cvt(hc[(int64_t)b * R + min(q, nq - 1)]); });
            }
            if constexpr (std::is_same_v<CT, double2>) {
                if (base) {
                    gemm16_store_base<P, K3>(acc, lane, a0, rp, h + (int64_t)cptr[p] * R + q0,
                                             base + (int64_t)cptr[p] * R + q0, R, nq);
                    continue;
                }
            }
            gemm16_store<P, K3>(acc, lane, a0, rp, h + (int64_t)cptr[p] * R + q0, R, nq, accumulate != 0);
        }
    }
}

// Stage 5 for a basis above the leaves: y[rows of t] += alpha U_t yh_t
// (!FWD: V_t^T, CJ: conjugated) for nq right-hand sides, after the leaf
// kernel; one workgroup per such basis, one thread per (row, rhs).
template <bool FWD, bool CJ, class T2>
__global__ void __launch_bounds__(256)
    k_basis_out(const int32_t* __restrict__ slots, const int32_t* __restrict__ bnodes,
                const int64_t* __restrict__ boff, const T2* __restrict__ blocks, const int32_t* __restrict__ lo,
                const int32_t* __restrict__ hi, const int32_t* __restrict__ perm, const int32_t* __restrict__ cptr,
                const double2* __restrict__ yh, int64_t R, int64_t q0, int64_t nq, double2 alpha,
                double2* __restrict__ y, int64_t ldy) {
    const int w = slots[blockIdx.x];
    const int t = bnodes[w];
    const int l0 = lo[t], nt = hi[t] - l0;
    const int c0 = cptr[t], r = cptr[t + 1] - c0;
    const T2* B = blocks + boff[w];
    for (int64_t idx = threadIdx.x; idx < (int64_t)nt * nq; idx += blockDim.x) {
        const int row = (int)(idx / nq), q = (int)(idx % nq);
        double2 acc = make_double2(0.0, 0.0);
        for (int k = 0; k < r; ++k) {
            const T2 b = B[FWD ? (int64_t)row + (int64_t)k * nt : (int64_t)k + (int64_t)row * r];
            double2 bv = make_double2((double)b.x, CJ ? -(double)b.y : (double)b.y);
            const double2 v = yh[(int64_t)(c0 + k) * R + q0 + q];
            acc = cadd(acc, cmul(bv, v));
        }
        double2* o = y + (int64_t)perm[l0 + row] + (q0 + q) * ldy;
        *o = cadd(*o, cmul(alpha, acc));
    }
}

inline unsigned blocks_for(int64_t n, int b) { return (unsigned)((n + b - 1) / b); }

// ---------------------------------------------------------------------------
// Optional stage timing: events recorded on the stream between the stages.
struct StageEvents {
    hipEvent_t ev[7];
    bool on = false;
    void mark(int i, hipStream_t st) {
        if (on) NCA_HIP(hipEventRecord(ev[i], st));
    }
};

// Operator blocks of one precision: the FP64 originals or their c32 copies.
template <bool MIX>
struct Ops;
template <>
struct Ops<false> {
    using T2 = double2;
    using P = PrecF64;
    static const T2* basis(const SideDev& s) { return s.bblocks.p; }
    static const T2* transfer(const SideDev& s) { return s.tblocks.p; }
    static const T2* coupling(const nca_h2* h) { return h->coupling->blocks.p; }
    static T2* xs(nca_h2* h) { return h->ws.xs; }
};
template <>
struct Ops<true> {
    using T2 = float2;
    using P = PrecF32;
    static const T2* basis(const SideDev& s) { return s.bblocks32.p; }
    static const T2* transfer(const SideDev& s) { return s.tblocks32.p; }
    static const T2* coupling(const nca_h2* h) { return h->S32.p; }
    static T2* xs(nca_h2* h) { return h->ws.xs32; }
};

// The stored near entries of the current precision when forward products
// use them (made and with a workspace), else null.
template <bool MIX>
const typename Ops<MIX>::T2* stored_near(const nca_h2* h) {
    if (!h->near_stored || !h->ws.ynear) return nullptr;
    if constexpr (MIX) return h->ns_blocks32.p;
    else return h->ns_blocks.p;
}

// The operator's MIXED products evaluate the coupling on the fly (no stored
// c32 or transposed copy needed).
inline bool couple_otf(const nca_h2* h) {
    return h->precision == NCA_PREC_MIXED && h->coup_h16 == 1 && h->couple_k3[1] == 2;
}

// This MIXED product's coupling is evaluated on the fly on the f16 path.
template <bool MIX>
bool couple_h16(const nca_h2* h) {
    return MIX && h->coup_h16 == 1 && h->couple_k3[1] == 2 && h->ws.xh16 && h->ws.xhpart && h->ws.xhsc;
}

// The near field of this MIXED product's MFMA chunks runs on the f16 path:
// the operator allows it, the launch choice asks for it, the workspace holds
// its buffers, and it is evaluated (not stored).
template <bool MIX>
bool near_h16(const nca_h2* h, bool fwd) {
    return MIX && h->h16 == 1 && h->near_k3[1] == 2 && h->ws.xs16 && h->ws.xpart && h->ws.xsc &&
           !(fwd && stored_near<MIX>(h));
}

// Stage 3 of a MIXED product evaluated on the fly for columns q0 .. q0 + nq -
// 1 (nq <= 32): the source coefficients of those columns split into f16
// parts with their column scales, then one k_couple_h16 launch; forward the
// c32 result goes to yc (yfin for the blocks fin marks) or, with yc null,
// to yh; transpose and adjoint write yh.
template <int NC, int OP>
void run_couple_h16(nca_h2* h, int64_t R, int nq, int64_t q0, float2* yc, const uint8_t* fin) {
    constexpr bool FWD = OP == NCA_OP_N;
    constexpr bool CJ = OP == NCA_OP_C;
    hipStream_t st = h->ctx->stream;
    const SideDev& src = FWD ? h->trial : h->test;
    const SideDev& dst = FWD ? h->test : h->trial;
    const float2* xh = reinterpret_cast<const float2*>(h->ws.xh);
    if (src.K > 0) {
        hipLaunchKernelGGL(k_colmax_c32, dim3(blocks_for(src.K, 64), blocks_for(nq, 16)), dim3(256), 0, st, xh,
                           src.K, R, q0, nq, h->ws.xhpart);
        hipLaunchKernelGGL(k_colscale, dim3((unsigned)nq), dim3(256), 0, st, h->ws.xhpart,
                           (int64_t)blocks_for(src.K, 64), (int64_t)nq, h->ws.xhsc);
        hipLaunchKernelGGL(k_split16, dim3(blocks_for(src.K * nq, 256)), dim3(256), 0, st, xh, h->ws.xhsc, src.K, R,
                           q0, nq, h->ws.xh16);
    }
    const int64_t ntiles = FWD ? h->f_ntiles : h->cr_ntiles;
    if (ntiles > 0) {
        const bool neg = (h->sign < 0) != CJ;
        NearK nk = near_consts_scaled(h->k, std::fabs(h->k) * 0.5 / M_PI * h->coup_rmax, h->coup_t);
        nk.rel32 = !h->rel32_off && dst.coup_rel.p != nullptr;
        const int32_t* tn = FWD ? h->f_tile_node.p : h->cr_tile_node.p;
        const int32_t* tr = FWD ? h->f_tile_row0.p : h->cr_tile_row0.p;
        const double* dpiv = FWD ? h->piv_test.p : h->piv_trial.p;
        const double* spiv = FWD ? h->piv_trial.p : h->piv_test.p;
        const int64_t* crp = FWD ? h->cf_ptr.p : h->cr_ptr.p;
        const int32_t* crl = FWD ? h->cf_lo.p : h->cr_lo.p;
        const int32_t* crh = FWD ? h->cf_hi.p : h->cr_hi.p;
        if (yc) {
            auto kern = neg ? k_couple_h16<true, NC, float2> : k_couple_h16<false, NC, float2>;
            hipLaunchKernelGGL(kern, dim3(blocks_for(ntiles, 4)), dim3(256), 0, st, tn, tr, ntiles, dst.cptr.p, dpiv,
                               crp, crl, crh, spiv, h->ws.xh16, h->ws.xhsc, R, q0, nq, nk, dst.coup_rel.p, yc,
                               h->ws.yh, fin, h->couple_xcd);
        } else {
            auto kern = neg ? k_couple_h16<true, NC, double2> : k_couple_h16<false, NC, double2>;
            hipLaunchKernelGGL(kern, dim3(blocks_for(ntiles, 4)), dim3(256), 0, st, tn, tr, ntiles, dst.cptr.p, dpiv,
                               crp, crl, crh, spiv, h->ws.xh16, h->ws.xhsc, R, q0, nq, nk, dst.coup_rel.p,
                               h->ws.yh, h->ws.yh, nullptr, h->couple_xcd);
        }
    }
    NCA_HIP(hipGetLastError());
}

// Stage 5 and the near field on the MFMA kernel for columns q0 .. q0 + nq - 1
// (nq <= 16 NC): one launch for NC chunks of 16 columns, each kernel
// evaluation serving all of them; then the bases above the leaves. With the
// near field stored, the forward coupling kernel applies it first into
// ws.ynear (each stored entry read once for the NC chunks) and the leaf
// kernel adds that instead of evaluating.
template <int NC, int OP, bool MIX>
void run_leaf(nca_h2* h, int64_t R, int nq, int64_t q0, double2* y, int64_t ldy, double2 alpha, double2 beta) {
    constexpr bool FWD = OP == NCA_OP_N;
    constexpr bool CJ = OP == NCA_OP_C;
    using O = Ops<MIX>;
    using T2 = typename O::T2;
    using P = typename O::P;
    hipStream_t st = h->ctx->stream;
    SideDev& src = FWD ? h->trial : h->test;
    SideDev& dst = FWD ? h->test : h->trial;
    const double2* yh = h->ws.yh;
    const T2* NS = FWD ? stored_near<MIX>(h) : nullptr;
    const double2* ynear = nullptr;
    if (NS && h->ns_ntiles > 0) {
        const bool k3 = h->couple_k3[MIX ? 1 : 0] != 0;
        auto kern = k3 ? k_couple_fwd_mfma<P, false, true, T2, double2, NC>
                       : k_couple_fwd_mfma<P, false, false, T2, double2, NC>;
        hipLaunchKernelGGL(kern, dim3(blocks_for(h->ns_ntiles, 4)), dim3(256), 0, st, NS, h->ns_tile_node.p,
                           h->ns_tile_row0.p, h->ns_ntiles, h->ns_base.p, h->ns_colbase.p, h->ns_ncols.p,
                           h->ns_colcoef.p, h->ns_cptr.p, O::xs(h) + q0, (uint32_t)R, R, q0, nq, h->ws.ynear,
                           h->ws.ynear, nullptr, 0);
        ynear = h->ws.ynear;
    }
    if constexpr (MIX) {
        if (dst.ntiles > 0 && !NS && near_h16<MIX>(h, FWD)) {
            const bool neg = (h->sign < 0) != CJ;
            NearK nk = near_consts_scaled(h->k, h->near_turns, h->h16_t);
            nk.rel32 = !h->rel32_off && dst.near_rel.p != nullptr;
            auto kern = neg ? k_leaf_out_h16<FWD, CJ, true, NC> : k_leaf_out_h16<FWD, CJ, false, NC>;
            hipLaunchKernelGGL(kern, dim3(blocks_for(dst.ntiles, 4)), dim3(256), 0, st, dst.tile_leaf.p,
                               dst.tile_row0.p, dst.ntiles, dst.leaves.p, dst.leafbasis.p, dst.boff.p, O::basis(dst),
                               dst.lo.p, dst.hi.p, dst.perm.p, dst.node_cptr(), dst.pts_h16.p, dst.nrptr.p,
                               dst.nrlo.p, dst.nrhi.p, src.pts_h16.p, yh, h->ws.xs16, h->ws.xsc, R, q0, nk,
                               dst.near_rel.p, alpha, beta, y, ldy, h->near_xcd, h->near_stagger, nq);
            if (dst.nibasis > 0)
                hipLaunchKernelGGL((k_basis_out<FWD, CJ, T2>), dim3((unsigned)dst.nibasis), dim3(256), 0, st,
                                   dst.ibslot.p, dst.bnodes.p, dst.boff.p, O::basis(dst), dst.lo.p, dst.hi.p,
                                   dst.perm.p, dst.node_cptr(), yh, R, q0, (int64_t)nq, alpha, y, ldy);
            NCA_HIP(hipGetLastError());
            return;
        }
    }
    if (dst.ntiles > 0) {
        // conj(G) for a negative phase sign, again conjugated for the adjoint
        const bool neg = (h->sign < 0) != CJ;
        const NearK nk = near_consts(h->k, h->near_turns);
        const bool k3 = h->near_k3[MIX ? 1 : 0] != 0;
        auto kern = k3 ? (neg ? k_leaf_out_mfma<P, FWD, CJ, true, true, NC> : k_leaf_out_mfma<P, FWD, CJ, true, false, NC>)
                       : (neg ? k_leaf_out_mfma<P, FWD, CJ, false, true, NC>
                              : k_leaf_out_mfma<P, FWD, CJ, false, false, NC>);
        hipLaunchKernelGGL(kern, dim3(blocks_for(dst.ntiles, 4)), dim3(256), 0, st, dst.tile_leaf.p,
                           dst.tile_row0.p, dst.ntiles, dst.leaves.p, dst.leafbasis.p, dst.boff.p, O::basis(dst),
                           dst.lo.p, dst.hi.p, dst.perm.p, dst.node_cptr(), dst.pts.p, dst.nrptr.p, dst.nrlo.p,
                           dst.nrhi.p, src.pts.p, yh, O::xs(h), R, q0, nk, h->sctab.p, alpha, beta, y, ldy,
                           h->near_xcd, h->near_stagger, nq, ynear);
    }
    if (dst.nibasis > 0)
        hipLaunchKernelGGL((k_basis_out<FWD, CJ, T2>), dim3((unsigned)dst.nibasis), dim3(256), 0, st, dst.ibslot.p,
                           dst.bnodes.p, dst.boff.p, O::basis(dst), dst.lo.p, dst.hi.p, dst.perm.p, dst.node_cptr(),
                           yh, R, q0, (int64_t)nq, alpha, y, ldy);
    NCA_HIP(hipGetLastError());
}

// Stage 3 forward on the MFMA kernel for columns q0 .. q0 + nq - 1 (nq <=
// 16 NC): one launch for NC chunks of 16 columns, each coupling entry read
// once for all of them.
template <int NC, int OP, bool MIX>
void run_couple(nca_h2* h, int64_t R, int nq, int64_t q0) {
    static_assert(OP == NCA_OP_N, "forward coupling");
    constexpr bool CJ = false;
    using O = Ops<MIX>;
    using T2 = typename O::T2;
    using P = typename O::P;
    hipStream_t st = h->ctx->stream;
    SideDev& src = h->trial;
    SideDev& dst = h->test;
    if (h->f_ntiles == 0) return;
    if constexpr (MIX) {
        if (couple_h16<MIX>(h)) {
            run_couple_h16<NC, OP>(h, R, nq, q0, h->ws.yc, dst.orph.p);
            return;
        }
    }
    const T2* xh = reinterpret_cast<const T2*>(h->ws.xh);
    double2* yh = h->ws.yh;
    const T2* S = O::coupling(h);
    // MIXED forward products: the coupling result (one FP32 tile per block)
    // goes to the c32 buffer yc; the downward pass sums it with the parents'
    // transfers in FP64 and stores yh once per block; the blocks no downward
    // transfer writes take the coupling result in yh directly
    using YT = std::conditional_t<MIX, float2, double2>;
    YT* ycoup = MIX ? reinterpret_cast<YT*>(h->ws.yc) : reinterpret_cast<YT*>(yh);
    const uint8_t* fin = MIX ? dst.orph.p : nullptr;
    const bool k3 = h->couple_k3[MIX ? 1 : 0] != 0;
    auto kern = k3 ? k_couple_fwd_mfma<P, CJ, true, T2, YT, NC> : k_couple_fwd_mfma<P, CJ, false, T2, YT, NC>;
    if (NC == 1 && MIX && h->couple_x32 && R > 16) {
        // one 16-column chunk of more: its source coefficients as a compact
        // copy when the coupling gathers them several times over (couple_x32)
        hipLaunchKernelGGL(k_chunk_to_c32<T2>, dim3(blocks_for(src.K * 16, 256)), dim3(256), 0, st, xh, R, q0,
                           src.K, nq, h->ws.xh32);
        hipLaunchKernelGGL(kern, dim3(blocks_for(h->f_ntiles, 4)), dim3(256), 0, st, S, h->f_tile_node.p,
                           h->f_tile_row0.p, h->f_ntiles, h->f_base.p, h->f_colbase.p, h->f_ncols.p, h->f_colcoef.p,
                           dst.cptr.p, reinterpret_cast<const T2*>(h->ws.xh32), 16u, R, q0, nq, ycoup, yh, fin,
                           h->couple_xcd);
    } else {
        hipLaunchKernelGGL(kern, dim3(blocks_for(h->f_ntiles, 4)), dim3(256), 0, st, S, h->f_tile_node.p,
                           h->f_tile_row0.p, h->f_ntiles, h->f_base.p, h->f_colbase.p, h->f_ncols.p, h->f_colcoef.p,
                           dst.cptr.p, xh + q0, (uint32_t)R, R, q0, nq, ycoup, yh, fin, h->couple_xcd);
    }
    NCA_HIP(hipGetLastError());
}

// Phases of one column chunk.
enum : int { kPhaseUp = 1, kPhaseCouple = 2, kPhaseDown = 4, kPhaseLeaf = 8, kPhaseAll = 15 };

// Columns q0 .. q0 + nq - 1 of R; NR = 16 runs the MFMA kernels for any
// nq <= 16 (lanes of the columns past nq compute, and are not stored).
// phases: the stages to run (the caller runs the others for this chunk
// together with the next one).
template <int NR, int OP, bool MIX>
void run_chunk(nca_h2* h, int64_t R, int nq, int64_t q0, double2* y, int64_t ldy, double2 alpha, double2 beta,
               StageEvents* evs, int phases = kPhaseAll) {
    constexpr bool FWD = OP == NCA_OP_N;
    constexpr bool CJ = OP == NCA_OP_C;
    using O = Ops<MIX>;
    using T2 = typename O::T2;
    using P = typename O::P;
    hipStream_t st = h->ctx->stream;
    SideDev& src = FWD ? h->trial : h->test;
    SideDev& dst = FWD ? h->test : h->trial;
    const T2* xs = O::xs(h);
    // source coefficients in the operator's precision (each is one tile's
    // product; in MIXED an FP32 one, and c32 storage halves their traffic);
    // destination coefficients in FP64, since they sum the coupling and the
    // downward contributions of every level
    T2* xh = reinterpret_cast<T2*>(h->ws.xh);
    double2* yh = h->ws.yh;
    constexpr bool kCombine = MIX && FWD && NR == 16;
    float2* yc = kCombine ? h->ws.yc : nullptr;
    if (phases & kPhaseUp) {
    if (evs) evs->mark(1, st);
    if (src.nbasis > 0) {
        if (NR == 16) {
            if (src.nptile > 0)
                hipLaunchKernelGGL((k_project_mfma<P, !FWD, CJ, true>), dim3(blocks_for(src.nptile, 4)), dim3(256), 0,
                                   st, src.ptile.p, src.nptile, src.bnodes.p, src.boff.p, O::basis(src), src.lo.p,
                                   src.hi.p, src.node_cptr(), xs, R, q0, nq, xh);
        } else
            hipLaunchKernelGGL((k_project<NR, !FWD, CJ, T2, T2>), dim3((unsigned)src.nbasis), dim3(kWave), 0, st,
                               src.bnodes.p, src.boff.p, O::basis(src), src.lo.p, src.hi.p, src.node_cptr(), xs, R,
                               q0, xh);
    }
    if (evs) evs->mark(2, st);
    // upward: deepest level first; source transfers (trial E forward, test E transposed)
    for (int64_t l = src.nlev - 1; l >= 0; --l) {
        const int64_t np = src.levptr[l + 1] - src.levptr[l];
        if (np == 0) continue;
        if (NR == 16)
            hipLaunchKernelGGL((k_transfer_mfma<P, false, !FWD, CJ, true, T2>), dim3(blocks_for(np, 4)), dim3(256), 0,
                               st, src.parents.p, src.chdptr.p, src.chds.p, src.toff.p, O::transfer(src), src.cptr.p,
                               src.levptr[l], np, R, q0, nq, 0, xh, nullptr);
        else
            hipLaunchKernelGGL((k_transfer<NR, false, !FWD, CJ, T2, T2>), dim3((unsigned)np), dim3(kWave), 0, st,
                               src.parents.p, src.chdptr.p, src.chds.p, src.toff.p, O::transfer(src), src.cptr.p,
                               src.levptr[l], R, q0, xh);
    }
    if (evs) evs->mark(3, st);
    }
    if (phases & kPhaseCouple) {
    const T2* S = O::coupling(h);
    if (h->coupling->nentries > 0 && couple_h16<MIX>(h) && (!FWD || NR < 16)) {
        // on the fly: transpose / adjoint, and forward chunks of 1 to 3
        // columns (into the cleared yh, which their VALU downward pass adds to)
        run_couple_h16<1, OP>(h, R, nq, q0, nullptr, nullptr);
    } else if constexpr (FWD) {
        if constexpr (NR == 16) {
            run_couple<1, OP, MIX>(h, R, nq, q0);
        } else if (h->coupling->nentries > 0) {
            hipLaunchKernelGGL((k_couple_fwd<NR, CJ, T2>), dim3((unsigned)dst.ne), dim3(kWave), 0, st, S, h->f_base.p,
                               h->f_colbase.p, h->f_ncols.p, h->f_colcoef.p, dst.cptr.p, xh, R, q0, yh);
        }
    } else if (h->coupling->nentries > 0) {
        // the transposed store in the forward format when it is made (the
        // forward kernels then apply S^T), else the reverse CSR walk
        const T2* ST = MIX ? reinterpret_cast<const T2*>(h->rv_blocks32.p) : reinterpret_cast<const T2*>(h->rv_blocks.p);
        if (h->rv_plan && ST) {
            if constexpr (NR == 16) {
                if (h->rv_ntiles > 0) {
                    const bool k3 = h->couple_k3[MIX ? 1 : 0] != 0;
                    auto kern = k3 ? k_couple_fwd_mfma<P, CJ, true, T2, double2, 1>
                                   : k_couple_fwd_mfma<P, CJ, false, T2, double2, 1>;
                    hipLaunchKernelGGL(kern, dim3(blocks_for(h->rv_ntiles, 4)), dim3(256), 0, st, ST,
                                       h->rv_tile_node.p, h->rv_tile_row0.p, h->rv_ntiles, h->rv_base.p,
                                       h->rv_colbase.p, h->rv_ncols.p, h->rv_colcoef.p, dst.cptr.p, xh + q0,
                                       (uint32_t)R, R, q0, nq, yh, yh, nullptr, h->couple_xcd);
                }
            } else {
                hipLaunchKernelGGL((k_couple_fwd<NR, CJ, T2>), dim3((unsigned)dst.ne), dim3(kWave), 0, st, ST,
                                   h->rv_base.p, h->rv_colbase.p, h->rv_ncols.p, h->rv_colcoef.p, dst.cptr.p, xh, R,
                                   q0, yh);
            }
        } else {
            hipLaunchKernelGGL((k_couple_rev<NR, CJ, T2>), dim3((unsigned)dst.ne), dim3(kWave), 0, st, S, h->r_ptr.p,
                               h->r_test.p, h->r_off.p, src.cptr.p, dst.cptr.p, xh, R, q0, nq, yh);
        }
    }
    if (evs) evs->mark(4, st);
    }
    if (phases & kPhaseDown) {
    // downward: root level first; destination transfers (test E forward, trial E transposed)
    if (dst.down_by_child) {
        // child blocks with several parents: one wave per child sums over its
        // parents, the upward kernel's form (out = child, in = parents)
        for (size_t l = 0; l + 1 < dst.dlevptr.size(); ++l) {
            const int64_t nc = dst.dlevptr[l + 1] - dst.dlevptr[l];
            if (nc == 0) continue;
            if (NR == 16)
                hipLaunchKernelGGL((k_transfer_mfma<P, false, !FWD, CJ, true, double2>), dim3(blocks_for(nc, 4)),
                                   dim3(256), 0, st, dst.dchild.p, dst.dptr.p, dst.dpar.p, dst.dtoff.p,
                                   O::transfer(dst), dst.cptr.p, dst.dlevptr[l], nc, R, q0, nq, 1, yh, yc);
            else
                hipLaunchKernelGGL((k_transfer<NR, false, !FWD, CJ, T2, double2>), dim3((unsigned)nc), dim3(kWave), 0,
                                   st, dst.dchild.p, dst.dptr.p, dst.dpar.p, dst.dtoff.p, O::transfer(dst), dst.cptr.p,
                                   dst.dlevptr[l], R, q0, yh);
        }
    }
    for (int64_t l = 0; l < (dst.down_by_child ? 0 : dst.nlev); ++l) {
        const int64_t np = dst.levptr[l + 1] - dst.levptr[l];
        if (np == 0) continue;
        if (NR == 16)
            hipLaunchKernelGGL((k_transfer_mfma<P, true, !FWD, CJ, true, double2>), dim3(blocks_for(np, 4)), dim3(256),
                               0, st, dst.parents.p, dst.chdptr.p, dst.chds.p, dst.toff.p, O::transfer(dst),
                               dst.cptr.p, dst.levptr[l], np, R, q0, nq, 1, yh, yc);
        else
            hipLaunchKernelGGL((k_transfer<NR, true, !FWD, CJ, T2, double2>), dim3((unsigned)np), dim3(kWave), 0, st,
                               dst.parents.p, dst.chdptr.p, dst.chds.p, dst.toff.p, O::transfer(dst), dst.cptr.p,
                               dst.levptr[l], R, q0, yh);
    }
    if (evs) evs->mark(5, st);
    }
    if (!(phases & kPhaseLeaf)) return;
    if constexpr (NR == 16) {
        run_leaf<1, OP, MIX>(h, R, nq, q0, y, ldy, alpha, beta);
    } else {
        // 1..3 columns: the near field on VALU over the leaf tiles (or the
        // stored one, applied by the VALU forward coupling kernel)
        const T2* NS = FWD ? stored_near<MIX>(h) : nullptr;
        const double2* ynear = nullptr;
        if (NS && h->ns_nleaf > 0) {
            const int64_t n16 = dst.npts * R;
            hipLaunchKernelGGL(k_fill, dim3(blocks_for(n16, 256)), dim3(256), 0, st, h->ws.ynear, n16);
            hipLaunchKernelGGL((k_couple_fwd<NR, false, T2>), dim3((unsigned)h->ns_nleaf), dim3(kWave), 0, st, NS,
                               h->ns_base.p, h->ns_colbase.p, h->ns_ncols.p, h->ns_colcoef.p, h->ns_cptr.p, xs, R,
                               q0, h->ws.ynear);
            ynear = h->ws.ynear;
        }
        if (dst.ntiles > 0) {
            const bool neg = (h->sign < 0) != CJ;
            const NearK nk = near_consts(h->k, h->near_turns);
            auto kern = neg ? k_leaf_out_small<P, FWD, CJ, true, NR> : k_leaf_out_small<P, FWD, CJ, false, NR>;
            hipLaunchKernelGGL(kern, dim3(blocks_for(dst.ntiles, 4)), dim3(256), 0, st, dst.tile_leaf.p,
                               dst.tile_row0.p, dst.ntiles, dst.leaves.p, dst.leafbasis.p, dst.boff.p, O::basis(dst),
                               dst.lo.p, dst.hi.p, dst.perm.p, dst.node_cptr(), dst.pts.p, dst.nrptr.p, dst.nrlo.p,
                               dst.nrhi.p, src.pts.p, yh, xs, R, q0, nk, h->sctab.p, alpha, beta, y, ldy,
                               h->near_xcd, h->near_stagger, ynear);
        }
        if (dst.nibasis > 0)
            hipLaunchKernelGGL((k_basis_out<FWD, CJ, T2>), dim3((unsigned)dst.nibasis), dim3(256), 0, st,
                               dst.ibslot.p, dst.bnodes.p, dst.boff.p, O::basis(dst), dst.lo.p, dst.hi.p, dst.perm.p,
                               dst.node_cptr(), yh, R, q0, (int64_t)nq, alpha, y, ldy);
    }
    if (evs) evs->mark(6, st);
    NCA_HIP(hipGetLastError());
}

template <int OP, bool MIX>
void run_op(nca_h2* h, const double2* x, int64_t ldx, double2* y, int64_t ldy, int64_t R, double2 alpha,
            double2 beta, StageEvents* evs) {
    constexpr bool FWD = OP == NCA_OP_N;
    hipStream_t st = h->ctx->stream;
    SideDev& src = FWD ? h->trial : h->test;
    SideDev& dst = FWD ? h->test : h->trial;
    if (evs) evs->mark(0, st);
    // the f16 near field of the MFMA chunks (if any) needs the split copy of
    // the gathered right-hand sides and their column scales
    const bool msmall = h->mfma_small >= 0 ? h->mfma_small != 0 : MIX;
    const bool h16 = near_h16<MIX>(h, FWD) && (R >= 4 || msmall) && src.npts * R > 0;
    if (src.npts * R > 0)
        hipLaunchKernelGGL((k_gather_tiled<typename Ops<MIX>::T2>),
                           dim3(blocks_for(src.npts, 64), blocks_for(R, 16)), dim3(256), 0, st, x, ldx, src.iperm.p,
                           src.npts, R, R, Ops<MIX>::xs(h), h16 ? h->ws.xpart : nullptr);
    if constexpr (MIX) {
        if (h16) {
            hipLaunchKernelGGL(k_colscale, dim3((unsigned)R), dim3(256), 0, st, h->ws.xpart,
                               (int64_t)blocks_for(src.npts, 64), R, h->ws.xsc);
            hipLaunchKernelGGL(k_split16, dim3(blocks_for(src.npts * R, 256)), dim3(256), 0, st, h->ws.xs32,
                               h->ws.xsc, src.npts, R, (int64_t)0, (int)R, h->ws.xs16);
        }
    }
    // The MFMA chunks write every coefficient of xh (projection, upward pass)
    // and, forward, of yh (coupling; in MIXED the coupling and the downward
    // pass) exactly once; the VALU chunks and the reverse coupling add, so they
    // need cleared vectors.
    const bool valu = !msmall && (R % 16) > 0 && (R % 16) < 4;
    // source coefficients: K x R elements of the operator's type, destination
    // coefficients FP64; cleared in 16-byte units
    using CT = typename Ops<MIX>::T2;
    const int64_t nx16 = (src.K * R * (int64_t)sizeof(CT) + 15) / 16, ny16 = dst.K * R;
    if (nx16 > 0 && valu)
        hipLaunchKernelGGL(k_fill, dim3(blocks_for(nx16, 256)), dim3(256), 0, st, h->ws.xh, nx16);
    if (ny16 > 0 && (valu || !FWD))
        hipLaunchKernelGGL(k_fill, dim3(blocks_for(ny16, 256)), dim3(256), 0, st, h->ws.yh, ny16);
    // chunks of up to 16 columns on the MFMA kernels; a last chunk of 1 to 3
    // columns on the VALU kernels unless msmall. Two MFMA chunks in
    // a row share one near-field launch (each kernel evaluation serves 32
    // columns).
    int64_t q0 = 0;
    while (q0 < R) {
        const int64_t left = R - q0;
        if (left >= 4 || msmall) {
            const int nq = (int)std::min<int64_t>(16, left);
            const int nq2 = (int)std::min<int64_t>(16, left - nq);
            if (nq == 16 && nq2 > 0 && (nq2 >= 4 || msmall)) {
                // two chunks: stages 1-2 of each, one forward MIXED coupling
                // launch (each coupling entry read once) or one per chunk,
                // stage 4 of each, one near-field launch
                run_chunk<16, OP, MIX>(h, R, nq, q0, y, ldy, alpha, beta, evs, kPhaseUp);
                run_chunk<16, OP, MIX>(h, R, nq2, q0 + 16, y, ldy, alpha, beta, evs, kPhaseUp);
                if constexpr (OP == NCA_OP_N) {
                    run_couple<2, OP, MIX>(h, R, nq + nq2, q0);
                } else {
                    run_chunk<16, OP, MIX>(h, R, nq, q0, y, ldy, alpha, beta, evs, kPhaseCouple);
                    run_chunk<16, OP, MIX>(h, R, nq2, q0 + 16, y, ldy, alpha, beta, evs, kPhaseCouple);
                }
                run_chunk<16, OP, MIX>(h, R, nq, q0, y, ldy, alpha, beta, evs, kPhaseDown);
                run_chunk<16, OP, MIX>(h, R, nq2, q0 + 16, y, ldy, alpha, beta, evs, kPhaseDown);
                run_leaf<2, OP, MIX>(h, R, nq + nq2, q0, y, ldy, alpha, beta);
                q0 += nq + nq2;
                continue;
            }
            run_chunk<16, OP, MIX>(h, R, nq, q0, y, ldy, alpha, beta, evs);
            q0 += nq;
        } else if (left == 3) { run_chunk<3, OP, MIX>(h, R, 3, q0, y, ldy, alpha, beta, evs); q0 += 3; }
        else if (left == 2) { run_chunk<2, OP, MIX>(h, R, 2, q0, y, ldy, alpha, beta, evs); q0 += 2; }
        else { run_chunk<1, OP, MIX>(h, R, 1, q0, y, ldy, alpha, beta, evs); q0 += 1; }
    }
}

// The set-up of the on-the-fly coupling (defined with the plans below).
void setup_couple_h16(nca_h2* h);

// The c32 copy of the coupling blocks (MIXED products that stream the stored
// coupling: couple_k3 0 or 1, or the on-the-fly form unavailable). Made
// outside any capture: by the MIXED set-up, or when such a choice is applied.
void ensure_s32(nca_h2* h) {
    if (h->S32.p) return;
    hipStream_t st = h->ctx->stream;
    const int64_t n = h->coupling->nentries;
    h->S32.alloc((size_t)std::max<int64_t>(n, 1));
    if (n > 0) hipLaunchKernelGGL(k_to_c32, dim3(blocks_for(n, 256)), dim3(256), 0, st, h->coupling->blocks.p, h->S32.p, n);
    NCA_HIP(hipGetLastError());
    NCA_HIP(hipStreamSynchronize(st));
}

// The rel32 condition over tile statistics (tile_pair_stats, 4 per tile):
// no pair strictly between 0 and the zero-distance guard, and the largest
// distance of a row from the tile's first row c within kRel32Span of the
// tile's smallest pair distance. For a pair (x, y) at distance r, |x - c| <=
// kRel32Span r and |y - c| <= |x - c| + r; rounding x - c and y - c to FP32
// and subtracting errs by at most 2^-24 (2 |x - c| + 2 r), i.e. by at most
// 2^-23 (1 + kRel32Span) of r.
// rel32_flags: per node of nn, 1 when every tile of the node (tnode[w])
// meets the condition; returns the fraction of tiles that do.
constexpr double kRel32Span = 32.0;
double rel32_flags(const std::vector<double>& hm, const std::vector<int32_t>& tnode, int64_t nn,
                   std::vector<uint8_t>& flag) {
    flag.assign((size_t)std::max<int64_t>(nn, 1), 1);
    int64_t ok = 0, nt = 0;
    for (size_t w = 0; w < tnode.size(); ++w) {
        const int32_t t = tnode[w];
        if (t < 0) continue;  // padding tile
        const double* q = hm.data() + 4 * w;
        const bool good = q[2] > 2e-30 && (!std::isfinite(q[0]) || q[3] <= kRel32Span * kRel32Span * q[0]);
        ++nt;
        if (good) ++ok;
        else flag[(size_t)t] = 0;
    }
    return nt ? (double)ok / (double)nt : 0.0;
}
// Node of each near-field tile of a side in the current leaf order (the
// tiles run leaf by leaf, 16 rows each).
std::vector<int32_t> near_tile_nodes(const SideDev& sd) {
    std::vector<int32_t> out;
    out.reserve((size_t)sd.ntiles);
    for (int32_t t : sd.h_leaves)
        for (int r0 = 0; r0 < sd.h_hi[t] - sd.h_lo[t]; r0 += 16) out.push_back(t);
    return out;
}

// c32 copies of the operator for the MIXED path (made once: at creation in
// MIXED, by nca_h2_set_precision, or on first use). It allocates once and
// frees nothing a captured sweep uses, so the workspace generation stays.
void ensure_mixed(nca_h2* h) {
    if (h->mixed_ready) return;
    hipStream_t st = h->ctx->stream;
    auto conv = [&](const DevBuf<double2>& in, DevBuf<float2>& out, int64_t n) {
        out.alloc((size_t)std::max<int64_t>(n, 1));
        if (n > 0)
            hipLaunchKernelGGL(k_to_c32, dim3(blocks_for(n, 256)), dim3(256), 0, st, in.p, out.p, n);
    };
    for (SideDev* sd : {&h->test, &h->trial}) {
        conv(sd->bblocks, sd->bblocks32, sd->basis_entries);
        conv(sd->tblocks, sd->tblocks32, sd->transfer_entries);
    }
    // f16 near field: the smallest near distance above the zero-distance
    // guard (the forward tiles cover every near pair); the path is taken when
    // the near distances span at most 2^17, with t = 2^-14-e for r_min = f 2^e
    // (f in [0.5, 1)), so that 1 / (t r) lies in [2^-3, 2^15]
    h->h16 = 0;
    h->near_rmin = h->near_rmax_pairs = 0.0;
    if (h->test.ntiles > 0) {
        DevBuf<double> mins((size_t)(4 * h->test.ntiles));
        hipLaunchKernelGGL(k_near_rmin, dim3(blocks_for(h->test.ntiles, 4)), dim3(256), 0, st, h->test.tile_leaf.p,
                           h->test.tile_row0.p, h->test.ntiles, h->test.leaves.p, h->test.lo.p, h->test.hi.p,
                           h->test.pts.p, h->test.nrptr.p, h->test.nrlo.p, h->test.nrhi.p, h->trial.pts.p, mins.p);
        NCA_HIP(hipGetLastError());
        std::vector<double> hm((size_t)(4 * h->test.ntiles));
        NCA_HIP(hipMemcpyAsync(hm.data(), mins.p, hm.size() * sizeof(double), hipMemcpyDeviceToHost, st));
        NCA_HIP(hipStreamSynchronize(st));
        double r2 = HUGE_VAL, r2x = 0.0;
        for (size_t q = 0; q < hm.size(); q += 4) {
            r2 = std::min(r2, hm[q]);
            r2x = std::max(r2x, hm[q + 1]);
        }
        const double rmin = std::sqrt(r2), rmax = std::sqrt(r2x);
        h->near_rmin = std::isfinite(rmin) ? rmin : 0.0;
        h->near_rmax_pairs = rmax;
        std::vector<uint8_t> flag;
        const double ff = rel32_flags(hm, near_tile_nodes(h->test), (int64_t)h->test.h_lo.size(), flag);
        h->test.near_rel.upload(flag, st);
        h->near_rel_frac = ff;
        // the reverse tiles (transpose and adjoint products) as well
        if (h->trial.ntiles > 0) {
            DevBuf<double> rv((size_t)(4 * h->trial.ntiles));
            hipLaunchKernelGGL(k_near_rmin, dim3(blocks_for(h->trial.ntiles, 4)), dim3(256), 0, st,
                               h->trial.tile_leaf.p, h->trial.tile_row0.p, h->trial.ntiles, h->trial.leaves.p,
                               h->trial.lo.p, h->trial.hi.p, h->trial.pts.p, h->trial.nrptr.p, h->trial.nrlo.p,
                               h->trial.nrhi.p, h->test.pts.p, rv.p);
            NCA_HIP(hipGetLastError());
            std::vector<double> hv((size_t)(4 * h->trial.ntiles));
            NCA_HIP(hipMemcpyAsync(hv.data(), rv.p, hv.size() * sizeof(double), hipMemcpyDeviceToHost, st));
            NCA_HIP(hipStreamSynchronize(st));
            const double fr = rel32_flags(hv, near_tile_nodes(h->trial), (int64_t)h->trial.h_lo.size(), flag);
            h->trial.near_rel.upload(flag, st);
            h->near_rel_frac = (ff * (double)h->test.ntiles + fr * (double)h->trial.ntiles) /
                               (double)(h->test.ntiles + h->trial.ntiles);
        }
        if (std::isfinite(rmin) && rmin > 0.0 && rmax <= 131072.0 * rmin) {
            int e = 0;
            (void)std::frexp(rmin, &e);
            h->h16_t = std::ldexp(1.0, -14 - e);
            for (SideDev* sd : {&h->test, &h->trial}) {
                const int64_t n = 3 * (sd->npts + kNearPad);
                sd->pts_h16.alloc((size_t)n);
                hipLaunchKernelGGL(k_scale_pts, dim3(blocks_for(n, 256)), dim3(256), 0, st, sd->pts.p,
                                   sd->pts_h16.p, n, h->h16_t);
            }
            h->h16 = 1;
        }
    }
    NCA_HIP(hipGetLastError());
    NCA_HIP(hipStreamSynchronize(st));
    setup_couple_h16(h);
    if (h->coup_h16 != 1 || h->couple_k3[1] != 2) ensure_s32(h);
    h->mixed_ready = true;
}

// The transposed coupling of T/C products and the stored near field of
// forward products in the current precision (defined with the plans below).
void ensure_reverse(nca_h2* h);
void ensure_near_store(nca_h2* h);

void ensure_workspace(nca_h2* h, int64_t R) {
    const int64_t np = std::max(h->M, h->N), K = std::max(h->test.K, h->trial.K);
    if (R > h->ws_rhs) {
        ++h->generation;
        h->ws_xh.alloc((size_t)std::max<int64_t>(K * R, 1));
        h->ws_yh.alloc((size_t)std::max<int64_t>(K * R, 1));
        h->ws_xs.release();
        h->ws_xs32.release();
        h->ws_yc.release();
        h->ws_xs16.release();
        h->ws_xpart.release();
        h->ws_xsc.release();
        h->ws_xh16.release();
        h->ws_rhs = R;
    }
    if (h->precision == NCA_PREC_MIXED) {
        if (!h->ws_xs32.p) h->ws_xs32.alloc((size_t)std::max<int64_t>((np + kNearPad) * h->ws_rhs, 1));
        if (!h->ws_xh32.p) h->ws_xh32.alloc((size_t)std::max<int64_t>(K * 16, 1));
        if (!h->ws_yc.p) h->ws_yc.alloc((size_t)std::max<int64_t>(K * h->ws_rhs, 1));
        if (h->h16 == 1) {
            if (!h->ws_xs16.p) h->ws_xs16.alloc((size_t)std::max<int64_t>(np * h->ws_rhs, 1));
            if (!h->ws_xpart.p) h->ws_xpart.alloc((size_t)std::max<int64_t>(blocks_for(np, 64) * h->ws_rhs, 1));
            if (!h->ws_xsc.p) h->ws_xsc.alloc((size_t)std::max<int64_t>(2 * h->ws_rhs, 1));
        }
        if (h->coup_h16 == 1) {
            if (!h->ws_xh16.p) h->ws_xh16.alloc((size_t)std::max<int64_t>(K * h->ws_rhs, 1));
            if (!h->ws_xhpart.p) h->ws_xhpart.alloc((size_t)std::max<int64_t>(blocks_for(K, 64) * 32, 1));
            if (!h->ws_xhsc.p) h->ws_xhsc.alloc(64);
        }
    } else {
        if (!h->ws_xs.p) h->ws_xs.alloc((size_t)std::max<int64_t>((np + kNearPad) * h->ws_rhs, 1));
    }
    if (h->near_stored && h->ws_ynear.n < (size_t)(np * h->ws_rhs)) {
        ++h->generation;
        h->ws_ynear.alloc((size_t)std::max<int64_t>(np * h->ws_rhs, 1));
    }
}

// Argument checks of a device product; false when there is nothing to do.
bool check_product(const nca_h2* h, int32_t op, int64_t R, int64_t ldx, int64_t ldy) {
    NCA_REQUIRE(op >= NCA_OP_N && op <= NCA_OP_C, NCA_ERR_ARGUMENT, "nca_h2_matvec: op must be 0 (N), 1 (T) or 2 (C)");
    NCA_REQUIRE(R >= 0, NCA_ERR_ARGUMENT, "nca_h2_matvec: negative nrhs");
    const int64_t nx = op == NCA_OP_N ? h->N : h->M, ny = op == NCA_OP_N ? h->M : h->N;
    NCA_REQUIRE(ldx >= nx && ldy >= ny, NCA_ERR_DIMENSION, "DimensionMismatch: leading dimension too small");
    NCA_REQUIRE(std::max({h->M, h->N, h->test.K, h->trial.K}) * R < ((int64_t)1 << 31), NCA_ERR_LIMIT,
                "nca_h2_matvec: points (or coefficients) x nrhs must stay below 2^31 (split the right-hand sides)");
    NCA_REQUIRE(std::max(h->test.max_range, h->trial.max_range) * R * (int64_t)sizeof(double2) < ((int64_t)1 << 32),
                NCA_ERR_LIMIT,
                "nca_h2_matvec: a near-field source range x nrhs x 16 bytes must stay below 2^32 (split the "
                "right-hand sides)");
    return R > 0 && ny > 0;
}

// The sweep itself on the workspace h->ws (set by the caller).
void run_sweep(nca_h2* h, int32_t op, int64_t R, const double2* x, int64_t ldx, double2* y, int64_t ldy,
               const double* alpha, const double* beta, StageEvents* evs) {
    const bool mix = h->precision == NCA_PREC_MIXED;
    const double2 al = make_double2(alpha ? alpha[0] : 1.0, alpha ? alpha[1] : 0.0);
    const double2 be = make_double2(beta ? beta[0] : 0.0, beta ? beta[1] : 0.0);
    if (mix) {
        if (op == NCA_OP_N) run_op<NCA_OP_N, true>(h, x, ldx, y, ldy, R, al, be, evs);
        else if (op == NCA_OP_T) run_op<NCA_OP_T, true>(h, x, ldx, y, ldy, R, al, be, evs);
        else run_op<NCA_OP_C, true>(h, x, ldx, y, ldy, R, al, be, evs);
    } else {
        if (op == NCA_OP_N) run_op<NCA_OP_N, false>(h, x, ldx, y, ldy, R, al, be, evs);
        else if (op == NCA_OP_T) run_op<NCA_OP_T, false>(h, x, ldx, y, ldy, R, al, be, evs);
        else run_op<NCA_OP_C, false>(h, x, ldx, y, ldy, R, al, be, evs);
    }
}

void matvec_device(nca_h2* h, int32_t op, int64_t R, const double2* x, int64_t ldx, double2* y, int64_t ldy,
                   const double* alpha, const double* beta, StageEvents* evs = nullptr) {
    if (!check_product(h, op, R, ldx, ldy)) return;
    if (h->precision == NCA_PREC_MIXED) ensure_mixed(h);
    if (op != NCA_OP_N && !couple_otf(h)) ensure_reverse(h);
    if (op == NCA_OP_N && h->near_stored) ensure_near_store(h);
    ensure_workspace(h, R);
    h->ws = nca_h2::WsView{h->ws_xs.p,   h->ws_xs32.p,  h->ws_xh.p,   h->ws_yh.p,
                           h->ws_yc.p,   h->ws_xh32.p,  h->near_stored ? h->ws_ynear.p : nullptr,
                           h->ws_xs16.p, h->ws_xpart.p, h->ws_xsc.p, h->ws_xh16.p, h->ws_xhpart.p, h->ws_xhsc.p};
    run_sweep(h, op, R, x, ldx, y, ldy, alpha, beta, evs);
}

// ---------------------------------------------------------------------------
// Allocation-free products (nca_h2_matvec_device_noalloc). Their workspace is
// this static device array: kScratchBytes on every GPU the library runs on,
// held until the process ends, shared by every such call on that GPU (calls
// whose device work overlaps on one GPU race). Host code never takes the
// array's name as an address: scratch_address() asks the runtime for the
// device address once per device and every later use goes through it.
constexpr size_t kScratchBytes = (size_t)64 << 20;
static __device__ __align__(256) double2 scratch[kScratchBytes / sizeof(double2)];

constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static void* g_scratch_ptr[kMaxDevices] = {};

// Device address of `scratch` on `device` (the current device).
void* scratch_address(int device) {
    NCA_REQUIRE(device >= 0 && device < kMaxDevices, NCA_ERR_ARGUMENT,
                "nca_h2_matvec_device_noalloc: device id out of range");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[device]) {
        void* ptr = nullptr;
        NCA_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
        NCA_REQUIRE(ptr != nullptr, NCA_ERR_HIP, "nca_h2_matvec_device_noalloc: no device address for the scratch");
        g_scratch_ptr[device] = ptr;
    }
    return g_scratch_ptr[device];
}

// The workspace of one product of R right-hand sides carved from `base`
// (256-byte aligned parts); returns the bytes it needs.
size_t carve_workspace(const nca_h2* h, int64_t R, char* base, nca_h2::WsView& v) {
    const bool mix = h->precision == NCA_PREC_MIXED;
    const size_t np = (size_t)std::max(h->M, h->N), K = (size_t)std::max(h->test.K, h->trial.K);
    const size_t r = (size_t)R;
    size_t off = 0;
    auto take = [&](size_t bytes) -> char* {
        char* p = base ? base + off : nullptr;
        off += (std::max<size_t>(bytes, 16) + 255) & ~(size_t)255;
        return p;
    };
    v = nca_h2::WsView{};
    if (mix) v.xs32 = reinterpret_cast<float2*>(take((np + kNearPad) * r * sizeof(float2)));
    else v.xs = reinterpret_cast<double2*>(take((np + kNearPad) * r * sizeof(double2)));
    v.xh = reinterpret_cast<double2*>(take(K * r * (mix ? sizeof(float2) : sizeof(double2))));
    v.yh = reinterpret_cast<double2*>(take(K * r * sizeof(double2)));
    if (mix) {
        v.yc = reinterpret_cast<float2*>(take(K * r * sizeof(float2)));
        if (h->couple_x32 && R > 16) v.xh32 = reinterpret_cast<float2*>(take(K * 16 * sizeof(float2)));
        if (h->h16 == 1 && h->near_k3[1] == 2) {
            v.xs16 = reinterpret_cast<uint4*>(take(np * r * sizeof(uint4)));
            v.xpart = reinterpret_cast<float*>(take((size_t)blocks_for((int64_t)np, 64) * r * sizeof(float)));
            v.xsc = reinterpret_cast<float*>(take(2 * r * sizeof(float)));
        }
        if (h->coup_h16 == 1 && h->couple_k3[1] == 2) {
            v.xh16 = reinterpret_cast<uint4*>(take(K * r * sizeof(uint4)));
            v.xhpart = reinterpret_cast<float*>(take((size_t)blocks_for((int64_t)K, 64) * 32 * sizeof(float)));
            v.xhsc = reinterpret_cast<float*>(take(64 * sizeof(float)));
        }
    }
    // the stored near field is used when its entries of this precision exist
    if (h->near_stored && h->ns_plan && (mix ? h->ns_blocks32.p != nullptr : h->ns_blocks.p != nullptr))
        v.ynear = reinterpret_cast<double2*>(take(np * r * sizeof(double2)));
    return off;
}

// ---------------------------------------------------------------------------
// Creation: 1-based Julia-shaped plans -> device arrays.
// ne coefficient blocks; dirptr (nn+1, 1-based) gives the blocks of every
// node (null: one block per node). Bases belong to nodes and cover all
// coefficients of their node; transfers connect blocks and, with multiparent,
// a child block may have several parents.
// Host loop over [0, n) on up to 16 threads (dynamic chunks); f must not throw.
template <class F>
void parallel_for(int64_t n, F&& f) {
    const int64_t nt = std::min<int64_t>(std::max<unsigned>(1u, std::thread::hardware_concurrency()), 16);
    if (n < 8192 || nt == 1) {
        for (int64_t i = 0; i < n; ++i) f(i);
        return;
    }
    std::atomic<int64_t> next{0};
    constexpr int64_t kChunk = 512;
    std::vector<std::thread> pool;
    for (int64_t k = 0; k < nt; ++k)
        pool.emplace_back([&] {
            for (;;) {
                const int64_t b = next.fetch_add(kChunk);
                if (b >= n) break;
                const int64_t e = std::min(n, b + kChunk);
                for (int64_t i = b; i < e; ++i) f(i);
            }
        });
    for (auto& t : pool) t.join();
}

// Transpose of a CSR (nrows rows, entries col(i) < ncols) on up to 16
// threads: tptr (ncols + 1) and, for every entry i of row r in column c, its
// place slot(c, r) in column order, rows increasing within a column (as a
// serial counting sort gives). emit(i, r, pos) is called once per entry.
template <class Col, class Emit>
void csr_transpose(int64_t nrows, const int64_t* ptr, int64_t ncols, Col col, std::vector<int64_t>& tptr, Emit emit) {
    const int64_t nt = std::min<int64_t>(std::max<unsigned>(1u, std::thread::hardware_concurrency()), 16);
    const int64_t nnz = ptr[nrows] - ptr[0];
    const int64_t nth = nnz < (1 << 16) ? 1 : nt;
    // contiguous row ranges of about nnz / nth entries
    std::vector<int64_t> rb(nth + 1, nrows);
    rb[0] = 0;
    for (int64_t k = 1, r = 0; k < nth; ++k) {
        const int64_t want = ptr[0] + nnz * k / nth;
        while (r < nrows && ptr[r] < want) ++r;
        rb[k] = r;
    }
    std::vector<std::vector<int64_t>> cnt(nth, std::vector<int64_t>(ncols, 0));
    auto run = [&](auto&& body) {
        std::vector<std::thread> pool;
        for (int64_t k = 0; k < nth; ++k) pool.emplace_back([&, k] { body(k); });
        for (auto& t : pool) t.join();
    };
    run([&](int64_t k) {
        for (int64_t r = rb[k]; r < rb[k + 1]; ++r)
            for (int64_t i = ptr[r]; i < ptr[r + 1]; ++i) cnt[k][col(i)]++;
    });
    tptr.assign(ncols + 1, 0);
    for (int64_t c = 0; c < ncols; ++c) {
        int64_t run_ = tptr[c];
        for (int64_t k = 0; k < nth; ++k) {
            const int64_t v = cnt[k][c];
            cnt[k][c] = run_;
            run_ += v;
        }
        tptr[c + 1] = run_;
    }
    run([&](int64_t k) {
        for (int64_t r = rb[k]; r < rb[k + 1]; ++r)
            for (int64_t i = ptr[r]; i < ptr[r + 1]; ++i) emit(i, r, cnt[k][col(i)]++);
    });
}

void build_side(SideDev& sd, hipStream_t st, int64_t npts, int64_t nn, int64_t ne, const int64_t* dirptr,
                bool multiparent, const int64_t* perm, const int64_t* lo, const int64_t* hi, const int64_t* cptr,
                const nca_points* pts, int64_t nbasis, const int64_t* bnodes, const int64_t* boff,
                const double* bblocks, bool testside, int64_t nlev, const int64_t* levptr, const int64_t* levidcs,
                const int64_t* chdptr, const int64_t* chds, const int64_t* toff, const double* tblocks,
                const std::vector<PackJob>* bjobs = nullptr, const std::vector<PackJob>* tjobs = nullptr) {
    sd.npts = npts;
    sd.nn = nn;
    sd.ne = ne;
    NCA_REQUIRE(pts->n == npts, NCA_ERR_DIMENSION, "nca_h2_create: point count differs from the matrix size");
    sd.h_perm.resize(npts);
    std::vector<char> seen(npts, 0);
    for (int64_t i = 0; i < npts; ++i) {
        const int64_t p = perm[i] - 1;
        NCA_REQUIRE(p >= 0 && p < npts && !seen[p], NCA_ERR_ARGUMENT, "nca_h2_create: perm is not a permutation");
        seen[p] = 1;
        sd.h_perm[i] = (int32_t)p;
    }
    sd.h_lo.resize(nn);
    sd.h_hi.resize(nn);
    sd.h_cptr.resize(ne + 1);
    for (int64_t t = 0; t < nn; ++t) {
        sd.h_lo[t] = (int32_t)(lo[t] - 1);
        sd.h_hi[t] = (int32_t)hi[t];
        NCA_REQUIRE(sd.h_lo[t] >= 0 && sd.h_hi[t] >= sd.h_lo[t] && sd.h_hi[t] <= npts, NCA_ERR_ARGUMENT,
                    "nca_h2_create: node range out of bounds");
    }
    for (int64_t e = 0; e <= ne; ++e) sd.h_cptr[e] = (int32_t)(cptr[e] - 1);
    NCA_REQUIRE(sd.h_cptr[0] == 0, NCA_ERR_ARGUMENT, "nca_h2_create: the coefficient plan must start at 1");
    for (int64_t e = 0; e < ne; ++e)
        NCA_REQUIRE(sd.h_cptr[e + 1] >= sd.h_cptr[e], NCA_ERR_ARGUMENT, "nca_h2_create: decreasing coefficient plan");
    sd.K = sd.h_cptr[ne];
    sd.h_enode.resize(ne);
    sd.h_ncptr.resize(nn + 1);
    if (dirptr) {
        NCA_REQUIRE(dirptr[0] == 1 && dirptr[nn] - 1 == ne, NCA_ERR_ARGUMENT,
                    "nca_h2_create: direction pointer does not cover the direction indices");
        for (int64_t t = 0; t < nn; ++t) {
            NCA_REQUIRE(dirptr[t + 1] >= dirptr[t], NCA_ERR_ARGUMENT, "nca_h2_create: decreasing direction pointer");
            for (int64_t e = dirptr[t] - 1; e < dirptr[t + 1] - 1; ++e) sd.h_enode[e] = (int32_t)t;
        }
        for (int64_t t = 0; t <= nn; ++t) sd.h_ncptr[t] = sd.h_cptr[dirptr[t] - 1];
    } else {
        NCA_REQUIRE(ne == nn, NCA_ERR_ARGUMENT, "nca_h2_create: one coefficient block per node expected");
        for (int64_t e = 0; e < ne; ++e) sd.h_enode[e] = (int32_t)e;
        sd.h_ncptr = sd.h_cptr;
    }
    // kNearPad zero points after the last one (slack: the near-field loads are
    // range-checked and read zeros past a range end)
    std::vector<double> pp(3 * (npts + kNearPad), 0.0);
    for (int64_t i = 0; i < npts; ++i)
        for (int d = 0; d < 3; ++d) pp[3 * i + d] = pts->host[3 * sd.h_perm[i] + d];
    sd.h_box.assign(6 * (size_t)nn, 0.0);
    parallel_for(nn, [&](int64_t t) {
        double* b = &sd.h_box[6 * t];
        for (int d = 0; d < 3; ++d) {
            b[d] = HUGE_VAL;
            b[3 + d] = -HUGE_VAL;
        }
        for (int64_t i = sd.h_lo[t]; i < sd.h_hi[t]; ++i)
            for (int d = 0; d < 3; ++d) {
                b[d] = std::min(b[d], pp[3 * i + d]);
                b[3 + d] = std::max(b[3 + d], pp[3 * i + d]);
            }
    });
    sd.perm.upload(sd.h_perm, st);
    {
        std::vector<int32_t> ip(npts);
        for (int64_t i = 0; i < npts; ++i) ip[sd.h_perm[i]] = (int32_t)i;
        sd.iperm.upload(ip, st);
    }
    sd.lo.upload(sd.h_lo, st);
    sd.hi.upload(sd.h_hi, st);
    sd.cptr.upload(sd.h_cptr, st);
    if (dirptr) sd.ncptr.upload(sd.h_ncptr, st);
    sd.pts.upload(pp, st);
    // bases
    sd.nbasis = nbasis;
    std::vector<int32_t> bn(nbasis);
    std::vector<int64_t> bo(nbasis + 1);
    for (int64_t i = 0; i < nbasis; ++i) {
        const int64_t t = bnodes[i] - 1;
        NCA_REQUIRE(t >= 0 && t < nn, NCA_ERR_ARGUMENT, "nca_h2_create: basis node out of range");
        bn[i] = (int32_t)t;
        const int64_t expect = (int64_t)(sd.h_hi[t] - sd.h_lo[t]) * (sd.h_ncptr[t + 1] - sd.h_ncptr[t]);
        NCA_REQUIRE(boff[i + 1] - boff[i] == expect, NCA_ERR_DIMENSION,
                    std::string("DimensionMismatch: ") + (testside ? "test" : "trial") + " basis of node " +
                        std::to_string(t + 1) + " has " + std::to_string(boff[i + 1] - boff[i]) + " entries, " +
                        std::to_string(expect) + " expected");
    }
    for (int64_t i = 0; i <= nbasis; ++i) bo[i] = nbasis ? boff[i] - boff[0] : 0;
    sd.basis_entries = nbasis ? bo[nbasis] : 0;
    sd.h_bnodes = bn;
    sd.bnodes.upload(bn, st);
    sd.boff.upload(bo, st);
    {
        std::vector<int32_t> pt;
        for (int64_t i = 0; i < nbasis; ++i) {
            const int32_t t = bn[i];
            const int64_t r = sd.h_ncptr[t + 1] - sd.h_ncptr[t];
            if (sd.h_hi[t] == sd.h_lo[t]) continue;
            for (int64_t a0 = 0; a0 < r; a0 += 16) {
                pt.push_back((int32_t)i);
                pt.push_back((int32_t)a0);
            }
        }
        sd.nptile = (int64_t)pt.size() / 2;
        sd.ptile.upload(pt, st);
    }
    if (nbasis && bjobs) {
        // packed on the device from resident factors
        sd.bblocks.alloc((size_t)bo[nbasis]);
        pack_blocks(*bjobs, sd.bblocks.p, st);
    } else if (nbasis) {
        sd.bblocks.upload(reinterpret_cast<const double2*>(bblocks) + boff[0], (size_t)bo[nbasis], st);
    }
    // transfers
    sd.nlev = nlev;
    sd.levptr.assign(nlev + 1, 0);
    const int64_t npar = nlev ? levptr[nlev] - levptr[0] : 0;
    for (int64_t l = 0; l <= nlev; ++l) sd.levptr[l] = nlev ? levptr[l] - levptr[0] : 0;
    std::vector<int32_t> par(npar), ch;
    std::vector<int64_t> cptr0(npar + 1, 0), to;
    const int64_t nedge = npar ? chdptr[npar] - chdptr[0] : 0;
    ch.resize(nedge);
    to.resize(nedge + 1);
    std::vector<char> ischild(ne, 0);
    for (int64_t i = 0; i < npar; ++i) {
        const int64_t p = levidcs[i] - 1;
        NCA_REQUIRE(p >= 0 && p < ne, NCA_ERR_ARGUMENT, "nca_h2_create: transfer parent out of range");
        par[i] = (int32_t)p;
        cptr0[i + 1] = chdptr[i + 1] - chdptr[0];
        for (int64_t e = chdptr[i] - chdptr[0]; e < chdptr[i + 1] - chdptr[0]; ++e) {
            const int64_t c = chds[e] - 1;
            NCA_REQUIRE(c >= 0 && c < ne && (multiparent || !ischild[c]), NCA_ERR_ARGUMENT,
                        "nca_h2_create: transfer child out of range or with two parents");
            ischild[c] = 1;
            ch[e] = (int32_t)c;
            const int64_t rp = sd.h_cptr[p + 1] - sd.h_cptr[p], rc = sd.h_cptr[c + 1] - sd.h_cptr[c];
            NCA_REQUIRE(toff[e + 1] - toff[e] == rp * rc, NCA_ERR_DIMENSION,
                        "DimensionMismatch: transfer block of child " + std::to_string(c + 1) + " has " +
                            std::to_string(toff[e + 1] - toff[e]) + " entries, " + std::to_string(rp * rc) +
                            " expected");
        }
    }
    for (int64_t e = 0; e <= nedge; ++e) to[e] = nedge ? toff[e] - toff[0] : 0;
    sd.transfer_entries = nedge ? to[nedge] : 0;
    sd.parents.upload(par, st);
    sd.chdptr.upload(cptr0, st);
    sd.chds.upload(ch, st);
    sd.toff.upload(to, st);
    if (nedge && to[nedge] && tjobs) {
        sd.tblocks.alloc((size_t)to[nedge]);
        pack_blocks(*tjobs, sd.tblocks.p, st);
    } else if (nedge && to[nedge]) {
        sd.tblocks.upload(reinterpret_cast<const double2*>(tblocks) + toff[0], (size_t)to[nedge], st);
    }
    sd.down_by_child = multiparent;
    if (multiparent) {
        // per level of parents: the children (first appearance order), each
        // with its incoming (parent, block) edges in plan order
        std::vector<int32_t> dch, dpa;
        std::vector<int64_t> dpt(1, 0), dto;
        sd.dlevptr.assign(1, 0);
        std::vector<int64_t> slot(ne, -1);
        for (int64_t l = 0; l < nlev; ++l) {
            std::vector<int32_t> kids;
            std::vector<std::vector<std::pair<int32_t, int64_t>>> in;
            for (int64_t i = sd.levptr[l]; i < sd.levptr[l + 1]; ++i)
                for (int64_t e = cptr0[i]; e < cptr0[i + 1]; ++e) {
                    const int32_t c = ch[e];
                    if (slot[c] < 0) {
                        slot[c] = (int64_t)kids.size();
                        kids.push_back(c);
                        in.emplace_back();
                    }
                    in[slot[c]].emplace_back(par[i], to[e]);
                }
            for (size_t q = 0; q < kids.size(); ++q) {
                dch.push_back(kids[q]);
                for (auto& pe : in[q]) {
                    dpa.push_back(pe.first);
                    dto.push_back(pe.second);
                }
                dpt.push_back((int64_t)dpa.size());
                slot[kids[q]] = -1;
            }
            sd.dlevptr.push_back((int64_t)dch.size());
        }
        sd.dchild.upload(dch, st);
        sd.dpar.upload(dpa, st);
        sd.dptr.upload(dpt, st);
        sd.dtoff.upload(dto, st);
    }
    // the blocks the downward pass writes: every child with coefficients (by
    // child), or every child with coefficients of a parent with coefficients
    auto rank = [&](int64_t e) { return sd.h_cptr[e + 1] - sd.h_cptr[e]; };
    std::vector<char> written(ne, 0);
    for (int64_t i = 0; i < npar; ++i)
        for (int64_t e = cptr0[i]; e < cptr0[i + 1]; ++e)
            if (rank(ch[e]) > 0 && (multiparent || rank(par[i]) > 0)) written[ch[e]] = 1;
    std::vector<uint8_t> orph(std::max<int64_t>(ne, 1), 0);
    for (int64_t e = 0; e < ne; ++e) orph[e] = rank(e) > 0 && !written[e];
    sd.orph.upload(orph, st);
}

// Leaves of the side = nodes whose range is not split by any other node: a
// node is a leaf iff it is nobody's transfer parent and no node's range is a
// strict subset starting inside it. With ranges only, a leaf is a node that no
// other node lies strictly inside; computed by sorting ranges.
void build_leaves(SideDev& sd, hipStream_t st) {
    const int64_t nn = sd.nn;
    std::vector<int64_t> order(nn);
    for (int64_t i = 0; i < nn; ++i) order[i] = i;
    // sort by (lo asc, hi desc): a node is a leaf iff the next node in this
    // order does not start inside its range (nested ranges follow their parent)
    std::sort(order.begin(), order.end(), [&](int64_t a, int64_t b) {
        if (sd.h_lo[a] != sd.h_lo[b]) return sd.h_lo[a] < sd.h_lo[b];
        if (sd.h_hi[a] != sd.h_hi[b]) return sd.h_hi[a] > sd.h_hi[b];
        return a < b;
    });
    std::vector<int32_t> leaves;
    std::vector<char> isleaf(nn, 0);
    for (int64_t k = 0; k < nn; ++k) {
        const int64_t a = order[k];
        if (sd.h_hi[a] == sd.h_lo[a]) continue;
        bool leaf = true;
        if (k + 1 < nn) {
            const int64_t b = order[k + 1];
            if (sd.h_lo[b] < sd.h_hi[a]) leaf = false;
        }
        if (leaf) { isleaf[a] = 1; }
    }
    std::vector<int64_t> covered(sd.npts + 1, 0);
    for (int64_t a = 0; a < nn; ++a)
        if (isleaf[a]) {
            leaves.push_back((int32_t)a);
            covered[sd.h_lo[a]] += 1;
            covered[sd.h_hi[a]] -= 1;
        }
    int64_t run = 0;
    for (int64_t i = 0; i < sd.npts; ++i) {
        run += covered[i];
        NCA_REQUIRE(run == 1, NCA_ERR_ARGUMENT, "nca_h2_create: the leaves do not partition the points");
    }
    sd.nleaves = (int64_t)leaves.size();
    sd.h_leaves = leaves;
    sd.h_leaves_node_order = leaves;
}

// Order of the leaves (and their 16-row tiles) in the near-field launch, from
// the merged near ranges: 0 node id, 1 point order, 2 decreasing near-field
// work (longest first). Uploads leaves, leafbasis and the tiles.
void order_leaves(SideDev& sd, int mode, hipStream_t st) {
    const std::vector<int32_t>& leafslot = sd.h_leafslot;
    const std::vector<int64_t>& mptr = sd.h_nrptr;
    const std::vector<int32_t>& mlo = sd.h_nrlo;
    const std::vector<int32_t>& mhi = sd.h_nrhi;
    std::vector<int64_t> idx(sd.nleaves);
    for (int64_t i = 0; i < sd.nleaves; ++i) idx[i] = i;
    const std::vector<int32_t>& lv = sd.h_leaves_node_order;
    if (mode == 1)
        std::sort(idx.begin(), idx.end(), [&](int64_t a, int64_t b) { return sd.h_lo[lv[a]] < sd.h_lo[lv[b]]; });
    if (mode == 2) {
        std::vector<int64_t> work(sd.nleaves, 0);
        for (int64_t i = 0; i < sd.nleaves; ++i)
            for (int64_t e = mptr[lv[i]]; e < mptr[lv[i] + 1]; ++e) work[i] += mhi[e] - mlo[e];
        std::stable_sort(idx.begin(), idx.end(), [&](int64_t a, int64_t b) { return work[a] > work[b]; });
    }
    std::vector<int32_t> leaves(sd.nleaves), lb(sd.nleaves), tl, tr;
    for (int64_t i = 0; i < sd.nleaves; ++i) {
        leaves[i] = lv[idx[i]];
        lb[i] = leafslot[idx[i]];
    }
    sd.h_leaves = leaves;
    sd.leaves.upload(leaves, st);
    sd.leafbasis.upload(lb, st);
    for (int64_t li = 0; li < sd.nleaves; ++li)
        for (int r0 = 0; r0 < sd.h_hi[leaves[li]] - sd.h_lo[leaves[li]]; r0 += 16) {
            tl.push_back((int32_t)li);
            tr.push_back(r0);
        }
    sd.ntiles = (int64_t)tl.size();
    sd.tile_leaf.upload(tl, st);
    sd.tile_row0.upload(tr, st);
}


// Host phase timing of nca_h2_create (environment NCA_CREATE_TIMING): each
// mark prints the time since the previous one to stderr.
struct PhaseTimer {
    bool on;
    std::chrono::steady_clock::time_point t0 = std::chrono::steady_clock::now();
    explicit PhaseTimer(bool o) : on(o) {}
    void mark(const char* what) {
        if (!on) return;
        const auto t = std::chrono::steady_clock::now();
        fprintf(stderr, "[nca_h2_create] %-16s %8.1f ms\n", what,
                std::chrono::duration<double, std::milli>(t - t0).count());
        t0 = t;
    }
};

// Coupling tiles (16 rows of a destination block) level by level (block
// order inside a level), each level's tiles cut into 8 contiguous runs dealt
// so that workgroup b (4 tiles, placed on XCD b mod 8) takes its XCD's run;
// runs are padded with empty tiles (block -1). rank: rows of every block of
// the side sd.
void level_xcd_tiles(const SideDev& sd, const std::vector<int32_t>& rank, std::vector<int32_t>& tnode,
                     std::vector<int32_t>& trow) {
    const int64_t nn = sd.nn, ne = (int64_t)rank.size();
    const std::vector<int32_t>& enode = sd.h_enode;
    std::vector<int64_t> dfs(nn);
    for (int64_t t = 0; t < nn; ++t) dfs[t] = t;
    std::sort(dfs.begin(), dfs.end(), [&](int64_t a, int64_t b) {
        if (sd.h_lo[a] != sd.h_lo[b]) return sd.h_lo[a] < sd.h_lo[b];
        if (sd.h_hi[a] != sd.h_hi[b]) return sd.h_hi[a] > sd.h_hi[b];
        return a < b;
    });
    std::vector<int> ndepth(nn, 0);
    std::vector<int64_t> stack;
    for (const int64_t t : dfs) {
        while (!stack.empty() && !(sd.h_lo[t] >= sd.h_lo[stack.back()] && sd.h_hi[t] <= sd.h_hi[stack.back()]))
            stack.pop_back();
        ndepth[t] = (int)stack.size();
        stack.push_back(t);
    }
    std::vector<int> depth(ne, 0);
    for (int64_t t = 0; t < ne; ++t) depth[t] = ndepth[enode[t]];
    const int maxd = ne ? *std::max_element(depth.begin(), depth.end()) : 0;
    for (int d = 0; d <= maxd; ++d) {
        std::vector<int32_t> ln, lr;
        for (int64_t t = 0; t < ne; ++t)
            if (depth[t] == d && rank[t] > 0)
                for (int r0 = 0; r0 < rank[t]; r0 += 16) {
                    ln.push_back((int32_t)t);
                    lr.push_back(r0);
                }
        const int64_t nblk = ((int64_t)ln.size() + 3) / 4, per = (nblk + 7) / 8;
        for (int64_t i = 0; i < per; ++i)
            for (int64_t x = 0; x < 8; ++x) {
                const int64_t blk = x * per + i;
                for (int u = 0; u < 4; ++u) {
                    const int64_t k = blk * 4 + u;
                    const bool ok = blk < nblk && k < (int64_t)ln.size() && i < per;
                    tnode.push_back(ok ? ln[k] : -1);
                    trow.push_back(ok ? lr[k] : 0);
                }
            }
    }
}

// The transposed coupling of T/C products in the current precision (made
// once per precision; allocates and frees nothing a captured sweep uses).
void ensure_reverse(nca_h2* h) {
    const nca_coupling* cp = h->coupling;
    const bool mix = h->precision == NCA_PREC_MIXED;
    if (cp->nentries == 0 || (h->rv_plan && (mix ? h->rv_blocks32.p != nullptr : h->rv_blocks.p != nullptr)))
        return;  // nothing to make (also the case inside a graph capture)
    hipStream_t st = h->ctx->stream;
    if (!h->rv_plan) {
        const int64_t ntest = h->test.ne, ntrial = h->trial.ne, nb = cp->nblocks;
        std::vector<int64_t> rptr;
        std::vector<int32_t> rtest(nb);
        // the reverse plan's order (the one r_test and r_off hold)
        csr_transpose(
            ntest, cp->ptr.data(), ntrial, [&](int64_t i) { return cp->nodes[i]; }, rptr,
            [&](int64_t, int64_t t, int64_t pos) { rtest[pos] = (int32_t)t; });
        std::vector<int64_t> base(ntrial), colbase(ntrial + 1, 0), rdst(nb);
        std::vector<int32_t> ncols(ntrial), rs(nb);
        for (int64_t s = 0; s < ntrial; ++s) {
            int64_t C = 0;
            const int32_t r = cp->trialrank[s];
            for (int64_t e = rptr[s]; e < rptr[s + 1]; ++e) {
                rs[e] = r;
                rdst[e] = C;  // column of the block in its row, made absolute below
                if (r > 0) C += cp->testrank[rtest[e]];
            }
            ncols[s] = (int32_t)C;
            colbase[s + 1] = colbase[s] + C;
        }
        int64_t off = 0;
        for (int64_t s = 0; s < ntrial; ++s) {
            base[s] = off;
            for (int64_t e = rptr[s]; e < rptr[s + 1]; ++e) rdst[e] = off + rdst[e] * cp->trialrank[s];
            off += (int64_t)cp->trialrank[s] * ncols[s];
        }
        NCA_REQUIRE(off == cp->nentries, NCA_ERR_ARGUMENT, "nca_h2: coupling store and ranks disagree");
        std::vector<int32_t> colcoef((size_t)colbase[ntrial]);
        parallel_for(ntrial, [&](int64_t s) {
            if (ncols[s] == 0) return;
            int64_t c = colbase[s];
            for (int64_t e = rptr[s]; e < rptr[s + 1]; ++e)
                for (int j = 0; j < cp->testrank[rtest[e]]; ++j) colcoef[c++] = h->test.h_cptr[rtest[e]] + j;
        });
        colbase.pop_back();
        std::vector<int32_t> tnode, trow;
        level_xcd_tiles(h->trial, cp->trialrank, tnode, trow);
        h->rv_ntiles = (int64_t)tnode.size();
        h->rv_tile_node.upload(tnode, st);
        h->rv_tile_row0.upload(trow, st);
        h->rv_base.upload(base, st);
        h->rv_colbase.upload(colbase, st);
        h->rv_ncols.upload(ncols, st);
        h->rv_colcoef.upload(colcoef, st);
        h->rv_dst.upload(rdst, st);
        h->rv_rs.upload(rs, st);
        h->rv_plan = true;
    }
    const int64_t nb = cp->nblocks;
    if (mix && !h->rv_blocks32.p) {
        ensure_mixed(h);
        ensure_s32(h);
        h->rv_blocks32.alloc((size_t)cp->nentries);
        hipLaunchKernelGGL(k_transpose_blocks<float2>, dim3(blocks_for(nb, 4)), dim3(256), 0, st, h->S32.p,
                           h->rv_blocks32.p, h->r_off.p, h->rv_dst.p, h->r_test.p, h->test.cptr.p, h->rv_rs.p, nb);
    } else if (!mix && !h->rv_blocks.p) {
        h->rv_blocks.alloc((size_t)cp->nentries);
        hipLaunchKernelGGL(k_transpose_blocks<double2>, dim3(blocks_for(nb, 4)), dim3(256), 0, st,
                           cp->blocks.p, h->rv_blocks.p, h->r_off.p, h->rv_dst.p, h->r_test.p, h->test.cptr.p,
                           h->rv_rs.p, nb);
    }
    NCA_HIP(hipGetLastError());
    NCA_HIP(hipStreamSynchronize(st));
}

// On-the-fly MIXED coupling: the pivot coordinates of both sides in
// coefficient order (blocks: nodes, or a directional operator's direction
// indices, each with its own pivots; gathered on the
// device from the tree-ordered points), the far blocks as merged coefficient
// ranges per test block and per trial block (a range keeps the zero-distance
// guard when the two clusters' point boxes come within 1e-14: radius-0
// clusters can be admissible with themselves), the trial blocks' 16-row
// tiles, and the smallest and largest far pivot distances (measured on the
// device); available when those span at most 2^17, with the pivot coordinates
// scaled by t = 2^-14-e for r_min = f 2^e, f in [0.5, 1).
void setup_couple_h16(nca_h2* h) {
    h->coup_h16 = 0;
    const nca_coupling* cp = h->coupling;
    if (cp->nentries == 0 || h->f_ntiles == 0 || (int64_t)cp->tptr.size() != h->test.ne + 1 ||
        (int64_t)cp->sptr.size() != h->trial.ne + 1)
        return;
    hipStream_t st = h->ctx->stream;
    // pivot coordinates (unscaled first): coefficient cptr[t] + i is pivot i of block t
    auto pivots = [&](const SideDev& sd, const std::vector<int64_t>& pptr, const std::vector<int32_t>& piv,
                      DevBuf<double>& out) {
        std::vector<int32_t> inv(sd.npts);
        for (int64_t i = 0; i < sd.npts; ++i) inv[sd.h_perm[i]] = (int32_t)i;
        std::vector<int32_t> pos((size_t)sd.K);
        for (int64_t t = 0; t < sd.ne; ++t) {
            const int64_t r = sd.h_cptr[t + 1] - sd.h_cptr[t];
            NCA_REQUIRE(r == pptr[t + 1] - pptr[t], NCA_ERR_ARGUMENT, "nca_h2: coupling pivots and ranks disagree");
            for (int64_t i = 0; i < r; ++i) pos[sd.h_cptr[t] + i] = inv[piv[pptr[t] + i]];
        }
        DevBuf<int32_t> dpos;
        dpos.upload(pos, st);
        out.alloc((size_t)(3 * (sd.K + kNearPad)));
        hipLaunchKernelGGL(k_gather_pivots, dim3(blocks_for(sd.K + kNearPad, 256)), dim3(256), 0, st, sd.pts.p,
                           dpos.p, sd.K, out.p);
        NCA_HIP(hipGetLastError());
        NCA_HIP(hipStreamSynchronize(st));
    };
    pivots(h->test, cp->tptr, cp->tpiv, h->piv_test);
    pivots(h->trial, cp->sptr, cp->spiv, h->piv_trial);
    auto box_d2 = [](const double* a, const double* b) {
        double d2 = 0.0;
        for (int d = 0; d < 3; ++d) {
            const double g = std::max({0.0, a[d] - b[3 + d], b[d] - a[3 + d]});
            d2 += g * g;
        }
        return d2;
    };
    // merged coefficient ranges of the far blocks of each destination block
    auto ranges = [&](const SideDev& dsd, const SideDev& ssd, const std::vector<int64_t>& ptr,
                      const std::vector<int32_t>& nodes, const std::vector<int32_t>& srank, DevBuf<int64_t>& dptr,
                      DevBuf<int32_t>& dlo, DevBuf<int32_t>& dhi) {
        const int64_t nn = dsd.ne;
        struct Rg { int32_t lo, hi; bool guard; };
        // block t's merged ranges in a per-thread scratch vector (run twice: to
        // count, then to write, instead of one heap vector per block)
        auto ranges_of = [&](int64_t t, std::vector<Rg>& rg) {
            rg.clear();
            for (int64_t i = ptr[t]; i < ptr[t + 1]; ++i) {
                const int32_t q = nodes[i];
                // blocks are nodes, or direction indices of their node h_enode
                if (srank[q] > 0)
                    rg.push_back({ssd.h_cptr[q], ssd.h_cptr[q + 1],
                                  !(box_d2(&dsd.h_box[6 * (size_t)dsd.h_enode[t]],
                                           &ssd.h_box[6 * (size_t)ssd.h_enode[q]]) > 1e-28)});
            }
            std::sort(rg.begin(), rg.end(), [](const Rg& a, const Rg& b) { return a.lo < b.lo; });
            size_t k = 0;
            for (size_t q = 0; q < rg.size(); ++q) {
                if (k > 0 && rg[k - 1].hi == rg[q].lo && rg[k - 1].guard == rg[q].guard) rg[k - 1].hi = rg[q].hi;
                else rg[k++] = rg[q];
            }
            rg.resize(k);
        };
        std::vector<int64_t> cnt(nn, 0);
        parallel_for(nn, [&](int64_t t) {
            thread_local std::vector<Rg> rg;
            ranges_of(t, rg);
            cnt[t] = (int64_t)rg.size();
        });
        std::vector<int64_t> mp(nn + 1, 0);
        for (int64_t t = 0; t < nn; ++t) mp[t + 1] = mp[t] + cnt[t];
        std::vector<int32_t> lo((size_t)mp[nn]), hi((size_t)mp[nn]);
        parallel_for(nn, [&](int64_t t) {
            thread_local std::vector<Rg> rg;
            ranges_of(t, rg);
            int64_t o = mp[t];
            for (auto& r : rg) {
                lo[o] = r.lo;
                hi[o] = (int32_t)((uint32_t)r.hi | (r.guard ? 0x80000000u : 0u));
                ++o;
            }
        });
        dptr.upload(mp, st);
        dlo.upload(lo, st);
        dhi.upload(hi, st);
    };
    std::vector<int32_t> fnodes(cp->nodes.begin(), cp->nodes.end());
    ranges(h->test, h->trial, cp->ptr, fnodes, cp->trialrank, h->cf_ptr, h->cf_lo, h->cf_hi);
    std::vector<int64_t> rptr;
    std::vector<int32_t> rnodes((size_t)cp->nblocks);
    csr_transpose(
        h->test.ne, cp->ptr.data(), h->trial.ne, [&](int64_t i) { return cp->nodes[i]; }, rptr,
        [&](int64_t, int64_t t, int64_t pos) { rnodes[pos] = (int32_t)t; });
    ranges(h->trial, h->test, rptr, rnodes, cp->testrank, h->cr_ptr, h->cr_lo, h->cr_hi);
    std::vector<int32_t> tnode, trow;
    level_xcd_tiles(h->trial, cp->trialrank, tnode, trow);
    h->cr_ntiles = (int64_t)tnode.size();
    h->cr_tile_node.upload(tnode, st);
    h->cr_tile_row0.upload(trow, st);
    // far pivot distances over the forward tiles (every far pair once), and
    // the reverse tiles' statistics for the rel32 condition
    auto stats = [&](int64_t nt, const int32_t* tn, const int32_t* tr, const int32_t* cpd, const double* dp,
                     const int64_t* rp, const int32_t* rl, const int32_t* rh, const double* sp) {
        std::vector<double> out((size_t)(4 * nt));
        if (nt == 0) return out;
        DevBuf<double> mm((size_t)(4 * nt));
        hipLaunchKernelGGL(k_pair_rminmax, dim3(blocks_for(nt, 4)), dim3(256), 0, st, tn, tr, nt, cpd, dp, rp, rl,
                           rh, sp, mm.p);
        NCA_HIP(hipGetLastError());
        NCA_HIP(hipMemcpyAsync(out.data(), mm.p, out.size() * sizeof(double), hipMemcpyDeviceToHost, st));
        NCA_HIP(hipStreamSynchronize(st));
        return out;
    };
    const std::vector<double> hm = stats(h->f_ntiles, h->f_tile_node.p, h->f_tile_row0.p, h->test.cptr.p,
                                         h->piv_test.p, h->cf_ptr.p, h->cf_lo.p, h->cf_hi.p, h->piv_trial.p);
    const std::vector<double> hr = stats(h->cr_ntiles, h->cr_tile_node.p, h->cr_tile_row0.p, h->trial.cptr.p,
                                         h->piv_trial.p, h->cr_ptr.p, h->cr_lo.p, h->cr_hi.p, h->piv_test.p);
    {
        std::vector<int32_t> fnode((size_t)h->f_ntiles);
        if (h->f_ntiles > 0) {
            NCA_HIP(hipMemcpyAsync(fnode.data(), h->f_tile_node.p, fnode.size() * sizeof(int32_t),
                                   hipMemcpyDeviceToHost, st));
            NCA_HIP(hipStreamSynchronize(st));
        }
        std::vector<uint8_t> flag;
        const double ff = rel32_flags(hm, fnode, h->test.ne, flag);
        h->test.coup_rel.upload(flag, st);
        const double fr = rel32_flags(hr, tnode, h->trial.ne, flag);
        h->trial.coup_rel.upload(flag, st);
        h->coup_rel_frac = (ff * (double)h->f_ntiles + fr * (double)h->cr_ntiles) /
                           (double)std::max<int64_t>(h->f_ntiles + h->cr_ntiles, 1);
    }
    double r2 = HUGE_VAL, r2x = 0.0;
    for (size_t q = 0; q < hm.size(); q += 4) {
        r2 = std::min(r2, hm[q]);
        r2x = std::max(r2x, hm[q + 1]);
    }
    const double rmin = std::sqrt(r2), rmax = std::sqrt(r2x);
    h->coup_rmin = std::isfinite(rmin) ? rmin : 0.0;
    h->coup_rmax = rmax;
    if (!(std::isfinite(rmin) && rmin > 0.0 && rmax <= 131072.0 * rmin)) {
        h->piv_test.release();
        h->piv_trial.release();
        return;
    }
    int e = 0;
    (void)std::frexp(rmin, &e);
    h->coup_t = std::ldexp(1.0, -14 - e);
    for (DevBuf<double>* b : {&h->piv_test, &h->piv_trial}) {
        const int64_t n = (int64_t)b->n;
        hipLaunchKernelGGL(k_scale_pts, dim3(blocks_for(n, 256)), dim3(256), 0, st, b->p, b->p, n, h->coup_t);
    }
    NCA_HIP(hipGetLastError());
    NCA_HIP(hipStreamSynchronize(st));
    h->coup_h16 = 1;
}

// The stored near field of forward products (the reference CPU path keeps
// the near field as dense block rows, assemblenears
// src/nca/abstractnca.jl:134-178): per test leaf, in point order, one row of
// its points over the points of its merged near ranges, entries evaluated
// once on the device in the current precision (made once per precision;
// allocates and frees nothing a captured sweep uses).
void ensure_near_store(nca_h2* h) {
    const bool mix = h->precision == NCA_PREC_MIXED;
    if (h->ns_plan && (mix ? h->ns_blocks32.p != nullptr : h->ns_blocks.p != nullptr)) return;
    hipStream_t st = h->ctx->stream;
    SideDev& dst = h->test;
    const SideDev& src = h->trial;
    if (!h->ns_plan) {
        std::vector<int32_t> lv = dst.h_leaves_node_order;
        std::sort(lv.begin(), lv.end(), [&](int32_t a, int32_t b) { return dst.h_lo[a] < dst.h_lo[b]; });
        const int64_t nleaf = (int64_t)lv.size();
        std::vector<int32_t> cptr(nleaf + 1, 0), ncols(nleaf);
        std::vector<int64_t> base(nleaf), colbase(nleaf + 1, 0);
        for (int64_t i = 0; i < nleaf; ++i) {
            cptr[i] = dst.h_lo[lv[i]];
            NCA_REQUIRE(i == 0 ? cptr[0] == 0 : cptr[i] == dst.h_hi[lv[i - 1]], NCA_ERR_ARGUMENT,
                        "nca_h2_store_near: the test leaves do not tile the points");
            int64_t C = 0;
            for (int64_t e = dst.h_nrptr[lv[i]]; e < dst.h_nrptr[lv[i] + 1]; ++e)
                C += (dst.h_nrhi[e] & 0x7fffffff) - dst.h_nrlo[e];
            NCA_REQUIRE(C < ((int64_t)1 << 31), NCA_ERR_LIMIT, "nca_h2_store_near: a near row is too long");
            ncols[i] = (int32_t)C;
            colbase[i + 1] = colbase[i] + C;
        }
        cptr[nleaf] = nleaf ? dst.h_hi[lv[nleaf - 1]] : 0;
        NCA_REQUIRE(cptr[nleaf] == dst.npts, NCA_ERR_ARGUMENT, "nca_h2_store_near: the test leaves do not tile the points");
        int64_t off = 0;
        for (int64_t i = 0; i < nleaf; ++i) {
            base[i] = off;
            off += (int64_t)(cptr[i + 1] - cptr[i]) * ncols[i];
        }
        std::vector<int32_t> colcoef((size_t)colbase[nleaf]);
        parallel_for(nleaf, [&](int64_t i) {
            int64_t c = colbase[i];
            for (int64_t e = dst.h_nrptr[lv[i]]; e < dst.h_nrptr[lv[i] + 1]; ++e)
                for (int32_t q = dst.h_nrlo[e]; q < (dst.h_nrhi[e] & 0x7fffffff); ++q) colcoef[c++] = q;
        });
        colbase.pop_back();
        // 16-row tiles, the longest rows first (balance)
        std::vector<int64_t> order(nleaf);
        for (int64_t i = 0; i < nleaf; ++i) order[i] = i;
        std::stable_sort(order.begin(), order.end(), [&](int64_t a, int64_t b) { return ncols[a] > ncols[b]; });
        std::vector<int32_t> tnode, trow;
        for (const int64_t i : order)
            for (int r0 = 0; r0 < cptr[i + 1] - cptr[i]; r0 += 16) {
                tnode.push_back((int32_t)i);
                trow.push_back(r0);
            }
        h->ns_nleaf = nleaf;
        h->ns_entries = off;
        h->ns_ntiles = (int64_t)tnode.size();
        h->ns_cptr.upload(cptr, st);
        h->ns_base.upload(base, st);
        h->ns_colbase.upload(colbase, st);
        h->ns_ncols.upload(ncols, st);
        h->ns_colcoef.upload(colcoef, st);
        h->ns_tile_node.upload(tnode, st);
        h->ns_tile_row0.upload(trow, st);
        h->ns_plan = true;
    }
    if (h->ns_nleaf > 0) {
        if (mix) {
            h->ns_blocks32.alloc((size_t)std::max<int64_t>(h->ns_entries, 1));
            hipLaunchKernelGGL(k_near_store<float2>, dim3((unsigned)h->ns_nleaf), dim3(256), 0, st, dst.pts.p,
                               src.pts.p, h->ns_cptr.p, h->ns_base.p, h->ns_colbase.p, h->ns_ncols.p, h->ns_colcoef.p,
                               h->k, h->sign, h->ns_blocks32.p);
        } else {
            h->ns_blocks.alloc((size_t)std::max<int64_t>(h->ns_entries, 1));
            hipLaunchKernelGGL(k_near_store<double2>, dim3((unsigned)h->ns_nleaf), dim3(256), 0, st, dst.pts.p,
                               src.pts.p, h->ns_cptr.p, h->ns_base.p, h->ns_colbase.p, h->ns_ncols.p, h->ns_colcoef.p,
                               h->k, h->sign, h->ns_blocks.p);
        }
        NCA_HIP(hipGetLastError());
    }
    NCA_HIP(hipStreamSynchronize(st));
}

// Launch order of the forward-coupling tiles (16 rows of one test cluster;
// every cluster with coefficients has tiles, also one without far columns):
// 0 node id, 1 depth-first point order (nested ranges after their parent),
// 2 decreasing row work (columns), 3 per level, one contiguous run per XCD.
void order_coupling_tiles(nca_h2* h, int mode, hipStream_t st) {
    // rows are coefficient blocks (test nodes, or their directions); the tree
    // orders use the node of each block
    const int64_t ntest = h->test.ne;
    const std::vector<int32_t>& ncols = h->h_ncols;
    const std::vector<int32_t>& enode = h->test.h_enode;
    std::vector<int64_t> torder(ntest);
    for (int64_t t = 0; t < ntest; ++t) torder[t] = t;
    if (mode == 1)
        std::sort(torder.begin(), torder.end(), [&](int64_t a, int64_t b) {
            const SideDev& sd = h->test;
            const int32_t na = enode[a], nb = enode[b];
            if (sd.h_lo[na] != sd.h_lo[nb]) return sd.h_lo[na] < sd.h_lo[nb];
            if (sd.h_hi[na] != sd.h_hi[nb]) return sd.h_hi[na] > sd.h_hi[nb];
            return a < b;
        });
    if (mode == 2)
        std::stable_sort(torder.begin(), torder.end(), [&](int64_t a, int64_t b) { return ncols[a] > ncols[b]; });
    std::vector<int32_t> tnode, trow;
    if (mode != 3) {
        for (const int64_t t : torder)
            if (h->coupling->testrank[t] > 0)
                for (int r0 = 0; r0 < h->coupling->testrank[t]; r0 += 16) {
                    tnode.push_back((int32_t)t);
                    trow.push_back(r0);
                }
    } else {
        level_xcd_tiles(h->test, h->coupling->testrank, tnode, trow);
    }
    h->f_ntiles = (int64_t)tnode.size();
    h->f_tile_node.upload(tnode, st);
    h->f_tile_row0.upload(trow, st);
}

// Stage times (ms) of nreps profiled FWD sweeps over 16 RHS of zeros: the
// median of the coupling and of the leaf expansion + near field stage.
void time_stages(nca_h2* h, const double2* x, double2* y, int nreps, double& couple_ms, double& near_ms) {
    std::vector<double> c, n;
    StageEvents evs;
    for (auto& e : evs.ev) NCA_HIP(hipEventCreate(&e));
    evs.on = true;
    try {
        for (int r = 0; r < nreps; ++r) {
            matvec_device(h, NCA_OP_N, 16, x, h->N, y, h->M, nullptr, nullptr, &evs);
            NCA_HIP(hipEventSynchronize(evs.ev[6]));
            float a = 0.f, b = 0.f;
            NCA_HIP(hipEventElapsedTime(&a, evs.ev[3], evs.ev[4]));
            NCA_HIP(hipEventElapsedTime(&b, evs.ev[5], evs.ev[6]));
            c.push_back(a);
            n.push_back(b);
        }
    } catch (...) {
        for (auto& e : evs.ev) (void)hipEventDestroy(e);
        throw;
    }
    for (auto& e : evs.ev) (void)hipEventDestroy(e);
    std::sort(c.begin(), c.end());
    std::sort(n.begin(), n.end());
    couple_ms = c[c.size() / 2];
    near_ms = n[n.size() / 2];
}

// Coefficient blocks of a directional operator: per side the number of
// direction indices and their pointer per node (1-based, nn+1). The plain
// operator has one block per node (dirptr null).
// Block copies that pack the stores from resident factors (test/trial bases
// and transfers).
struct ResidentJobs {
    std::vector<PackJob> tb, tt, sb, st;
};

struct BlockLayout {
    int64_t ne_test = 0, ne_trial = 0;
    const int64_t* test_dirptr = nullptr;
    const int64_t* trial_dirptr = nullptr;
    const ResidentJobs* jobs = nullptr;  // stores packed from resident factors
};

// Basis store from resident factors: basis i of node (or direction index)
// bnodes[i] is X (m x r, test) or X^T (r x m, trial), at boff[i].
void basis_jobs(const nca_factors* f, bool testside, int64_t nbasis, const int64_t* bnodes, const int64_t* boff,
                const int64_t* cptr, std::vector<PackJob>& out) {
    out.clear();
    out.reserve(nbasis);
    for (int64_t i = 0; i < nbasis; ++i) {
        const int64_t id = bnodes[i] - 1;
        const FactorRef F = factor_of(f, id);
        NCA_REQUIRE(F.r == cptr[id + 1] - cptr[id] && boff[i + 1] - boff[i] == F.m * F.r, NCA_ERR_DIMENSION,
                    "resident factors: the factor of basis block " + std::to_string(id + 1) +
                        " differs from the plan's rank or size");
        const int64_t dst = boff[i] - boff[0];
        if (testside) out.push_back({F.p, dst, (int32_t)F.m, (int32_t)F.m, (int32_t)F.r, 1, (int32_t)F.m, 0});
        else out.push_back({F.p, dst, (int32_t)F.m, (int32_t)F.m, (int32_t)F.r, (int32_t)F.r, 1, 0});
    }
}

// Transfer store from resident factors: the rows of a parent's factor are
// its children's pivots, in the order the plan lists the children, so the
// block of child c is X_p[oc : oc + r_c, :] (test, r_c x r_p) or its transpose
// (trial), oc = the ranks of the children before c.
void transfer_jobs(const nca_factors* f, bool testside, int64_t nlev, const int64_t* levptr, const int64_t* levidcs,
                   const int64_t* chdptr, const int64_t* chds, const int64_t* toff, const int64_t* cptr,
                   std::vector<PackJob>& out) {
    out.clear();
    const int64_t npar = nlev ? levptr[nlev] - levptr[0] : 0;
    for (int64_t i = 0; i < npar; ++i) {
        const int64_t p = levidcs[i] - 1;
        const FactorRef P = factor_of(f, p);
        NCA_REQUIRE(P.r == cptr[p + 1] - cptr[p], NCA_ERR_DIMENSION,
                    "resident factors: the factor of transfer parent " + std::to_string(p + 1) +
                        " differs from the plan's rank");
        int64_t oc = 0;
        for (int64_t e = chdptr[i] - chdptr[0]; e < chdptr[i + 1] - chdptr[0]; ++e) {
            const int64_t c = chds[e] - 1;
            const int64_t rc = cptr[c + 1] - cptr[c];
            NCA_REQUIRE(oc + rc <= P.m && toff[e + 1] - toff[e] == rc * P.r, NCA_ERR_DIMENSION,
                        "resident factors: transfer block of child " + std::to_string(c + 1) +
                            " does not fit its parent's factor");
            const int64_t dst = toff[e] - toff[0];
            if (rc > 0) {
                if (testside)
                    out.push_back({P.p + oc, dst, (int32_t)P.m, (int32_t)rc, (int32_t)P.r, 1, (int32_t)rc, 0});
                else
                    out.push_back({P.p + oc, dst, (int32_t)P.m, (int32_t)rc, (int32_t)P.r, (int32_t)P.r, 1, 0});
            }
            oc += rc;
        }
        NCA_REQUIRE(oc == P.m, NCA_ERR_DIMENSION,
                    "resident factors: the children listed for transfer parent " + std::to_string(p + 1) +
                        " do not make up its factor's rows");
    }
}

// nca_h2_create with d's coefficient plans, transfers and coupling plan over
// the coefficient blocks of lay (trees, bases and near field over nodes).
void create_core(nca_ctx* ctx, const nca_helmholtz* kernel, const nca_points* test, const nca_points* trial,
                 const nca_h2_desc* d, const BlockLayout& lay, const nca_coupling* coupling, int32_t precision,
                 nca_h2** out) {
    {
        const bool directional = lay.test_dirptr != nullptr;
        NCA_REQUIRE(ctx && kernel && test && trial && d && coupling && out, NCA_ERR_ARGUMENT,
                    "nca_h2_create: invalid argument");
        NCA_REQUIRE(precision == NCA_PREC_FP64 || precision == NCA_PREC_MIXED, NCA_ERR_ARGUMENT,
                    "nca_h2_create: unknown precision");
        set_device(ctx);
        hipStream_t st = ctx->stream;
        auto h = std::make_unique<nca_h2>();
        PhaseTimer timer(getenv("NCA_CREATE_TIMING") != nullptr);
        h->ctx = ctx;
        h->k = kernel->wavenumber;
        h->sign = kernel->sign;
        h->precision = precision;
        h->M = d->M;
        h->N = d->N;
        h->coupling = coupling;
        {
            std::vector<double2> tab(kSinCosTab);
            for (int n = 0; n < kSinCosTab; ++n) {
                const long double a = (long double)n * 3.14159265358979323846264338327950288L / 128.0L;
                tab[n] = make_double2((double)cosl(a), (double)sinl(a));
            }
            h->sctab.upload(tab, st);
        }
        // directional operators: many small coupling blocks over short ranges,
        // where streaming the stored c32 blocks beats evaluating them (262k
        // wideband sphere: 0.6 ms against 1.3 ms); the evaluated form stays a
        // launch choice (couple_k3 = 2)
        if (directional) h->couple_k3[1] = 1;
        if (const char* v = getenv("NCA_MFMA_K3")) h->near_k3[0] = h->couple_k3[0] = atoi(v) != 0;
        if (const char* v = getenv("NCA_MIXED_K3")) h->near_k3[1] = h->couple_k3[1] = std::min(std::max(atoi(v), 0), 2);
        if (const char* v = getenv("NCA_ORDER_NEAR")) h->near_xcd = atoi(v);
        if (const char* v = getenv("NCA_ORDER_COUPLE")) h->couple_xcd = atoi(v);
        if (const char* v = getenv("NCA_LEAF_ORDER")) h->leaf_order = atoi(v);
        if (const char* v = getenv("NCA_COUPLE_ORDER")) h->couple_order = atoi(v);
        if (const char* v = getenv("NCA_NEAR_STAGGER")) h->near_stagger = atoi(v) != 0;
        if (const char* v = getenv("NCA_MFMA_SMALL")) h->mfma_small = atoi(v) != 0;
        const char* cx32 = getenv("NCA_COUPLE_X32");
        timer.mark("sides");
        build_side(h->test, st, d->M, d->nn_test, lay.ne_test, lay.test_dirptr, directional, d->test_perm,
                   d->test_lo, d->test_hi, d->test_cptr, test, d->nbasis_test, d->basis_test_nodes,
                   d->basis_test_off, d->basis_test, true, d->nlev_test, d->tt_levptr, d->tt_levidcs, d->tt_chdptr,
                   d->tt_chds, d->tt_off, d->tt_blocks, lay.jobs ? &lay.jobs->tb : nullptr,
                   lay.jobs ? &lay.jobs->tt : nullptr);
        build_side(h->trial, st, d->N, d->nn_trial, lay.ne_trial, lay.trial_dirptr, directional, d->trial_perm,
                   d->trial_lo, d->trial_hi, d->trial_cptr, trial, d->nbasis_trial, d->basis_trial_nodes,
                   d->basis_trial_off, d->basis_trial, false, d->nlev_trial, d->st_levptr, d->st_levidcs,
                   d->st_chdptr, d->st_chds, d->st_off, d->st_blocks, lay.jobs ? &lay.jobs->sb : nullptr,
                   lay.jobs ? &lay.jobs->st : nullptr);
        timer.mark("build_side");
        build_leaves(h->test, st);
        build_leaves(h->trial, st);
        // basis slot of every leaf
        for (SideDev* sd : {&h->test, &h->trial}) {
            std::vector<int32_t> slot(sd->nn, -1), lb;
            const std::vector<int32_t>& bn = sd->h_bnodes;
            for (int64_t i = 0; i < sd->nbasis; ++i) slot[bn[i]] = (int32_t)i;
            const std::vector<int32_t>& lv = sd->h_leaves;
            std::vector<char> isleaf(sd->nn, 0);
            for (int32_t l : lv) {
                lb.push_back(slot[l]);
                isleaf[l] = 1;
            }
            // directional bases may sit above the leaves; they are applied
            // after the leaves (their nodes do not overlap: the nodes below
            // one of them are low frequency and carry no such basis)
            std::vector<int32_t> ib;
            std::vector<char> cover(sd->npts + 1, 0);
            for (int64_t i = 0; i < sd->nbasis; ++i) {
                if (isleaf[bn[i]]) continue;
                NCA_REQUIRE(directional, NCA_ERR_ARGUMENT,
                            "nca_h2_create: a basis belongs to a node that is not a leaf");
                for (int32_t q = sd->h_lo[bn[i]]; q < sd->h_hi[bn[i]]; ++q) {
                    NCA_REQUIRE(!cover[q], NCA_ERR_ARGUMENT, "nca_h2_create: nested bases above the leaves");
                    cover[q] = 1;
                }
                ib.push_back((int32_t)i);
            }
            sd->nibasis = (int64_t)ib.size();
            sd->ibslot.upload(ib, st);
            sd->h_leafslot = lb;
        }
        timer.mark("leaves");
        // coupling plan must match the store (over coefficient blocks)
        const int64_t ntest = lay.ne_test;
        const int64_t nn_test = d->nn_test;
        NCA_REQUIRE(coupling->ntest == ntest && coupling->ntrial == lay.ne_trial, NCA_ERR_DIMENSION,
                    "nca_h2_create: coupling store of another tree pair");
        for (int64_t t = 0; t <= ntest; ++t)
            NCA_REQUIRE(d->cp_ptr[t] - d->cp_ptr[0] == coupling->ptr[t], NCA_ERR_ARGUMENT,
                        "nca_h2_create: coupling plan differs from the coupling store");
        for (int64_t i = 0; i < coupling->nblocks; ++i)
            NCA_REQUIRE(d->cp_nodes[i] - 1 == coupling->nodes[i], NCA_ERR_ARGUMENT,
                        "nca_h2_create: coupling plan differs from the coupling store");
        for (int64_t t = 0; t < ntest; ++t)
            NCA_REQUIRE(coupling->testrank[t] == h->test.h_cptr[t + 1] - h->test.h_cptr[t], NCA_ERR_DIMENSION,
                        "nca_h2_create: coupling ranks differ from the coefficient plan");
        for (int64_t s = 0; s < lay.ne_trial; ++s)
            NCA_REQUIRE(coupling->trialrank[s] == h->trial.h_cptr[s + 1] - h->trial.h_cptr[s], NCA_ERR_DIMENSION,
                        "nca_h2_create: coupling ranks differ from the coefficient plan");
        timer.mark("coupling checks");
        // forward coupling: per test node base/colbase/ncols; per column its source coefficient
        {
            std::vector<int64_t> base(ntest), colbase(ntest + 1, 0);
            std::vector<int32_t> ncols(ntest);
            parallel_for(ntest, [&](int64_t t) {
                base[t] = coupling->offsets[coupling->ptr[t]];
                int64_t C = 0;
                if (coupling->testrank[t] > 0)
                    for (int64_t i = coupling->ptr[t]; i < coupling->ptr[t + 1]; ++i)
                        C += coupling->trialrank[coupling->nodes[i]];
                ncols[t] = (int32_t)C;
            });
            for (int64_t t = 0; t < ntest; ++t) colbase[t + 1] = colbase[t] + ncols[t];
            const int64_t ncolumns = colbase[ntest];
            colbase.pop_back();
            h->h_ncols = ncols;
            order_coupling_tiles(h.get(), h->couple_order, st);
            h->f_base.upload(base, st);
            h->f_colbase.upload(colbase, st);
            h->f_ncols.upload(ncols, st);
            // the column map itself is built on the device from the block lists
            h->f_colcoef.alloc((size_t)std::max<int64_t>(ncolumns, 1));
            if (ncolumns > 0) {
                std::vector<int32_t> nodes32((size_t)coupling->nblocks);
                parallel_for(coupling->nblocks, [&](int64_t i) { nodes32[i] = (int32_t)coupling->nodes[i]; });
                DevBuf<int64_t> dptr;
                DevBuf<int32_t> dnodes, drank, dcptr;
                dptr.upload(coupling->ptr, st);
                dnodes.upload(nodes32, st);
                drank.upload(coupling->trialrank, st);
                dcptr.upload(h->trial.h_cptr, st);
                hipLaunchKernelGGL(k_colcoef, dim3(blocks_for(ntest, 4)), dim3(256), 0, st, ntest, dptr.p, dnodes.p,
                                   drank.p, dcptr.p, h->f_colbase.p, h->f_ncols.p, h->f_colcoef.p);
                NCA_HIP(hipGetLastError());
                NCA_HIP(hipStreamSynchronize(st));
            }
            h->stats.plan_bytes += ncolumns * 4 + ntest * 20;
            // the c32 copy costs a pass over all source coefficients: keep it
            // when the columns read each coefficient more than twice on average
            h->couple_x32 = cx32 ? atoi(cx32) != 0 : ncolumns > 2 * h->trial.K;
        }
        timer.mark("forward plan");
        // reverse coupling: blocks by trial block
        {
            const int64_t ntrial = lay.ne_trial;
            std::vector<int64_t> rptr, roff(coupling->nblocks);
            std::vector<int32_t> rtest(coupling->nblocks);
            csr_transpose(
                ntest, coupling->ptr.data(), ntrial, [&](int64_t i) { return coupling->nodes[i]; }, rptr,
                [&](int64_t i, int64_t t, int64_t pos) {
                    rtest[pos] = (int32_t)t;
                    roff[pos] = coupling->offsets[i];
                });
            h->r_ptr.upload(rptr, st);
            h->r_off.upload(roff, st);
            h->r_test.upload(rtest, st);
        }
        timer.mark("reverse plan");
        // near lists (over nodes), both orientations
        {
            const int64_t ntest = nn_test;
            const int64_t ntrial = d->nn_trial;
            const int64_t nnz = d->nearptr[ntest] - d->nearptr[0];
            std::vector<int64_t> np(ntest + 1), tp;
            std::vector<int32_t> nn32(nnz), tn(nnz);
            int64_t near_entries = 0;
            for (int64_t t = 0; t <= ntest; ++t) np[t] = d->nearptr[t] - d->nearptr[0];
            std::atomic<bool> bad{false};
            parallel_for(nnz, [&](int64_t i) {
                const int64_t s = d->nearnodes[i] - 1;
                if (s < 0 || s >= ntrial) bad = true;
                else nn32[i] = (int32_t)s;
            });
            NCA_REQUIRE(!bad, NCA_ERR_ARGUMENT, "nca_h2_create: near node out of range");
            std::vector<int64_t> rowent(ntest, 0);
            parallel_for(ntest, [&](int64_t t) {
                int64_t e = 0;
                for (int64_t i = np[t]; i < np[t + 1]; ++i) e += h->trial.h_hi[nn32[i]] - h->trial.h_lo[nn32[i]];
                rowent[t] = e * (h->test.h_hi[t] - h->test.h_lo[t]);
            });
            for (int64_t t = 0; t < ntest; ++t) near_entries += rowent[t];
            // the largest distance of a near pair: farthest corners of the point boxes
            std::vector<double> dmax2(ntest, 0.0);
            parallel_for(ntest, [&](int64_t t) {
                if (h->test.h_hi[t] <= h->test.h_lo[t]) return;
                const double* a = &h->test.h_box[6 * t];
                double m = 0.0;
                for (int64_t i = np[t]; i < np[t + 1]; ++i) {
                    const int32_t q = nn32[i];
                    if (h->trial.h_hi[q] <= h->trial.h_lo[q]) continue;
                    const double* b = &h->trial.h_box[6 * (size_t)q];
                    double d2 = 0.0;
                    for (int d = 0; d < 3; ++d) {
                        const double g = std::max(a[3 + d] - b[d], b[3 + d] - a[d]);
                        d2 += g * g;
                    }
                    m = std::max(m, d2);
                }
                dmax2[t] = m;
            });
            double dm2 = 0.0;
            for (int64_t t = 0; t < ntest; ++t) dm2 = std::max(dm2, dmax2[t]);
            h->near_turns = std::fabs(h->k) * 0.5 / M_PI * std::sqrt(dm2);
            h->near_rmax = std::sqrt(dm2);
            csr_transpose(
                ntest, np.data(), ntrial, [&](int64_t i) { return nn32[i]; }, tp,
                [&](int64_t, int64_t t, int64_t pos) { tn[pos] = (int32_t)t; });
            h->test.nearptr.upload(np, st);
            h->test.nearnodes.upload(nn32, st);
            h->trial.nearptr.upload(tp, st);
            h->trial.nearnodes.upload(tn, st);
            // merged contiguous source ranges per destination node, both orientations
            auto merge = [](SideDev& dsd, const SideDev& ssd, const std::vector<int64_t>& ptr,
                            const std::vector<int32_t>& nodes, hipStream_t s) {
                // per node: its source ranges sorted and merged where contiguous
                // and alike: a range keeps the zero-distance guard unless the
                // point boxes are more than 1e-14 apart (then every pair has
                // r^2 > 1e-28, far above the guard's 2e-30)
                const int64_t nn = dsd.nn;
                std::vector<int64_t> cnt(nn, 0);
                struct Rg { int32_t lo, hi; bool guard; };
                auto box_d2 = [](const double* a, const double* b) {
                    double d2 = 0.0;
                    for (int d = 0; d < 3; ++d) {
                        const double g = std::max({0.0, a[d] - b[3 + d], b[d] - a[3 + d]});
                        d2 += g * g;
                    }
                    return d2;
                };
                // node t's merged ranges in a per-thread scratch vector (run twice:
                // to count, then to write, instead of one heap vector per node)
                auto ranges_of = [&](int64_t t, std::vector<Rg>& rg) {
                    rg.clear();
                    for (int64_t i = ptr[t]; i < ptr[t + 1]; ++i) {
                        const int32_t q = nodes[i];
                        if (ssd.h_hi[q] > ssd.h_lo[q])
                            rg.push_back({ssd.h_lo[q], ssd.h_hi[q],
                                          !(box_d2(&dsd.h_box[6 * t], &ssd.h_box[6 * (size_t)q]) > 1e-28)});
                    }
                    std::sort(rg.begin(), rg.end(), [](const Rg& a, const Rg& b) { return a.lo < b.lo; });
                    size_t k = 0;
                    for (size_t q = 0; q < rg.size(); ++q) {
                        if (k > 0 && rg[k - 1].hi == rg[q].lo && rg[k - 1].guard == rg[q].guard) rg[k - 1].hi = rg[q].hi;
                        else rg[k++] = rg[q];
                    }
                    rg.resize(k);
                };
                parallel_for(nn, [&](int64_t t) {
                    thread_local std::vector<Rg> rg;
                    ranges_of(t, rg);
                    cnt[t] = (int64_t)rg.size();
                });
                std::vector<int64_t> mptr(nn + 1, 0);
                for (int64_t t = 0; t < nn; ++t) mptr[t + 1] = mptr[t] + cnt[t];
                std::vector<int32_t> mlo((size_t)mptr[nn]), mhi((size_t)mptr[nn]), mhid((size_t)mptr[nn]);
                parallel_for(nn, [&](int64_t t) {
                    thread_local std::vector<Rg> rg;
                    ranges_of(t, rg);
                    int64_t o = mptr[t];
                    for (auto& r : rg) {
                        mlo[o] = r.lo;
                        mhi[o] = r.hi;
                        mhid[o] = (int32_t)((uint32_t)r.hi | (r.guard ? 0x80000000u : 0u));
                        ++o;
                    }
                });
                dsd.max_range = 0;
                for (size_t q = 0; q < mlo.size(); ++q) dsd.max_range = std::max<int64_t>(dsd.max_range, mhi[q] - mlo[q]);
                dsd.nrptr.upload(mptr, s);
                dsd.nrlo.upload(mlo, s);
                dsd.nrhi.upload(mhid, s);
                dsd.h_nrptr = std::move(mptr);
                dsd.h_nrlo = std::move(mlo);
                dsd.h_nrhi = std::move(mhi);
            };
            merge(h->test, h->trial, np, nn32, st);
            merge(h->trial, h->test, tp, tn, st);
            order_leaves(h->test, h->leaf_order, st);
            order_leaves(h->trial, h->leaf_order, st);
            h->stats.near_pairs = nnz;
            h->stats.near_entries = near_entries;
            h->stats.plan_bytes += nnz * 12 + (ntest + ntrial) * 8;
        }
        timer.mark("near plans");
        h->stats.coupling_blocks = coupling->nblocks;
        h->stats.coupling_entries = coupling->nentries;
        h->stats.basis_entries_test = h->test.basis_entries;
        h->stats.basis_entries_trial = h->trial.basis_entries;
        h->stats.transfer_entries_test = h->test.transfer_entries;
        h->stats.transfer_entries_trial = h->trial.transfer_entries;
        h->stats.coeffs_test = h->test.K;
        h->stats.coeffs_trial = h->trial.K;
        if (h->precision == NCA_PREC_MIXED) ensure_mixed(h.get());
        NCA_HIP(hipStreamSynchronize(st));
        timer.mark("stats + sync");
        *out = h.release();
    }
}

}  // namespace nca

using namespace nca;

extern "C" {

int32_t nca_h2_create(nca_ctx* ctx, const nca_helmholtz* kernel, const nca_points* test, const nca_points* trial,
                      const nca_h2_desc* d, const nca_coupling* coupling, int32_t precision, nca_h2** out) {
    return guarded([&] {
        NCA_REQUIRE(d, NCA_ERR_ARGUMENT, "nca_h2_create: invalid argument");
        BlockLayout lay;
        lay.ne_test = d->nn_test;
        lay.ne_trial = d->nn_trial;
        create_core(ctx, kernel, test, trial, d, lay, coupling, precision, out);
    });
}

int32_t nca_h2_create_resident(nca_ctx* ctx, const nca_helmholtz* kernel, const nca_points* test,
                               const nca_points* trial, const nca_h2_desc* d, const nca_coupling* coupling,
                               int32_t precision, const nca_factors* testf, const nca_factors* trialf,
                               nca_h2** out) {
    return guarded([&] {
        NCA_REQUIRE(d && testf && trialf, NCA_ERR_ARGUMENT, "nca_h2_create_resident: invalid argument");
        NCA_REQUIRE(testf->nids == d->nn_test && trialf->nids == d->nn_trial, NCA_ERR_DIMENSION,
                    "nca_h2_create_resident: factors of another tree");
        NCA_REQUIRE(testf->ctx == ctx && trialf->ctx == ctx, NCA_ERR_ARGUMENT,
                    "nca_h2_create_resident: factors of another context");
        set_device(ctx);
        ResidentJobs rj;
        basis_jobs(testf, true, d->nbasis_test, d->basis_test_nodes, d->basis_test_off, d->test_cptr, rj.tb);
        basis_jobs(trialf, false, d->nbasis_trial, d->basis_trial_nodes, d->basis_trial_off, d->trial_cptr, rj.sb);
        transfer_jobs(testf, true, d->nlev_test, d->tt_levptr, d->tt_levidcs, d->tt_chdptr, d->tt_chds, d->tt_off,
                      d->test_cptr, rj.tt);
        transfer_jobs(trialf, false, d->nlev_trial, d->st_levptr, d->st_levidcs, d->st_chdptr, d->st_chds, d->st_off,
                      d->trial_cptr, rj.st);
        BlockLayout lay;
        lay.ne_test = d->nn_test;
        lay.ne_trial = d->nn_trial;
        lay.jobs = &rj;
        create_core(ctx, kernel, test, trial, d, lay, coupling, precision, out);
    });
}

// The directional plans lowered onto the block form of create_core: the
// bases of a node's directions become one basis of the node over all its
// coefficients (test [U_1 ... U_k], trial [V_1; ...; V_k]); the coupling plan
// becomes a CSR over test direction indices.
}  // extern "C"

namespace nca {
// nca_dh2_create, with the stores packed on the device from resident factors
// when testf / trialf are given.
void dh2_create_impl(nca_ctx* ctx, const nca_helmholtz* kernel, const nca_points* test, const nca_points* trial,
                     const nca_dh2_desc* d, const nca_coupling* coupling, int32_t precision,
                     const nca_factors* testf, const nca_factors* trialf, nca_h2** out) {
    {
        NCA_REQUIRE(d && coupling, NCA_ERR_ARGUMENT, "nca_dh2_create: invalid argument");
        NCA_REQUIRE(d->test_dirptr && d->trial_dirptr, NCA_ERR_ARGUMENT, "nca_dh2_create: missing direction pointers");
        struct Lowered {
            std::vector<int64_t> nodes, off;
            std::vector<double2> blocks;
        };
        auto lower = [](bool testside, int64_t nn, const int64_t* lo, const int64_t* hi, const int64_t* dirptr,
                        const int64_t* cptr, int64_t nb, const int64_t* bnodes, const int64_t* bdirptr,
                        const int64_t* bdiridx, const int64_t* boff, const double* blocks, Lowered& L,
                        const nca_factors* f, std::vector<PackJob>* jobs) {
            const double2* B = reinterpret_cast<const double2*>(blocks);
            L.off.assign(1, 0);
            const char* side = testside ? "test" : "trial";
            // checks and sizes in order, then the packing in parallel
            for (int64_t i = 0; i < nb; ++i) {
                const int64_t t = bnodes[i] - 1;
                NCA_REQUIRE(t >= 0 && t < nn, NCA_ERR_ARGUMENT, "nca_dh2_create: basis node out of range");
                const int64_t m = hi[t] - lo[t] + 1;
                const int64_t e0 = dirptr[t] - 1, e1 = dirptr[t + 1] - 1;
                const int64_t c0 = cptr[e0] - 1, R = cptr[e1] - cptr[e0];
                const int64_t j0 = bdirptr[i] - 1, j1 = bdirptr[i + 1] - 1;
                // listed directions: increasing, of this node; the others hold no coefficients
                int64_t prev = e0 - 1, listed = 0;
                for (int64_t j = j0; j < j1; ++j) {
                    const int64_t e = bdiridx[j] - 1;
                    NCA_REQUIRE(e > prev && e < e1, NCA_ERR_ARGUMENT,
                                std::string("nca_dh2_create: ") + side + " basis directions of node " +
                                    std::to_string(t + 1) + " are not increasing direction indices of the node");
                    for (int64_t q = prev + 1; q < e; ++q)
                        NCA_REQUIRE(cptr[q + 1] == cptr[q], NCA_ERR_ARGUMENT,
                                    "nca_dh2_create: a direction with coefficients has no basis");
                    const int64_t r = cptr[e + 1] - cptr[e];
                    NCA_REQUIRE(boff[j + 1] - boff[j] == m * r, NCA_ERR_DIMENSION,
                                std::string("DimensionMismatch: ") + side + " basis of direction index " +
                                    std::to_string(e + 1) + " has " + std::to_string(boff[j + 1] - boff[j]) +
                                    " entries, " + std::to_string(m * r) + " expected");
                    listed += r;
                    prev = e;
                }
                for (int64_t q = prev + 1; q < e1; ++q)
                    NCA_REQUIRE(cptr[q + 1] == cptr[q], NCA_ERR_ARGUMENT,
                                "nca_dh2_create: a direction with coefficients has no basis");
                NCA_REQUIRE(listed == R, NCA_ERR_ARGUMENT, "nca_dh2_create: basis ranks differ from the plan");
                L.nodes.push_back(t + 1);
                L.off.push_back(L.off.back() + m * R);
            }
            if (f) {
                // the node blocks packed on the device from the direction factors
                jobs->clear();
                for (int64_t i = 0; i < nb; ++i) {
                    const int64_t t = bnodes[i] - 1;
                    const int64_t m = hi[t] - lo[t] + 1;
                    const int64_t e0 = dirptr[t] - 1, e1 = dirptr[t + 1] - 1;
                    const int64_t c0 = cptr[e0] - 1, R = cptr[e1] - cptr[e0];
                    for (int64_t j = bdirptr[i] - 1; j < bdirptr[i + 1] - 1; ++j) {
                        const int64_t e = bdiridx[j] - 1, r = cptr[e + 1] - cptr[e];
                        const FactorRef F = factor_of(f, e);
                        NCA_REQUIRE(F.m == m && F.r == r, NCA_ERR_DIMENSION,
                                    "resident factors: the factor of direction index " + std::to_string(e + 1) +
                                        " differs from the plan's size");
                        if (r == 0) continue;
                        const int64_t ro = cptr[e] - 1 - c0;
                        if (testside)  // U_t(:, ro + k) = X_e(:, k)
                            jobs->push_back({F.p, L.off[i] + ro * m, (int32_t)m, (int32_t)m, (int32_t)r, 1,
                                             (int32_t)m, 0});
                        else  // V_t(ro + k, :) = X_e(:, k)^T
                            jobs->push_back({F.p, L.off[i] + ro, (int32_t)m, (int32_t)m, (int32_t)r, (int32_t)R, 1,
                                             0});
                    }
                }
                return;
            }
            L.blocks.resize((size_t)std::max<int64_t>(L.off.back(), 1));
            parallel_for(nb, [&](int64_t i) {
                const int64_t t = bnodes[i] - 1;
                const int64_t m = hi[t] - lo[t] + 1;
                const int64_t e0 = dirptr[t] - 1, e1 = dirptr[t + 1] - 1;
                const int64_t c0 = cptr[e0] - 1, R = cptr[e1] - cptr[e0];
                const int64_t base = L.off[i];
                for (int64_t j = bdirptr[i] - 1; j < bdirptr[i + 1] - 1; ++j) {
                    const int64_t e = bdiridx[j] - 1, r = cptr[e + 1] - cptr[e];
                    const double2* Be = B + boff[j];
                    const int64_t ro = cptr[e] - 1 - c0;  // first coefficient of e within the node
                    if (testside) {
                        // U_t(:, ro + k) = U_e(:, k)
                        for (int64_t k = 0; k < r; ++k)
                            for (int64_t a = 0; a < m; ++a) L.blocks[base + a + (ro + k) * m] = Be[a + k * m];
                    } else {
                        // V_t(ro + k, :) = V_e(k, :)
                        for (int64_t a = 0; a < m; ++a)
                            for (int64_t k = 0; k < r; ++k) L.blocks[base + (ro + k) + a * R] = Be[k + a * r];
                    }
                }
            });
        };
        Lowered lt, ls;
        ResidentJobs rj;
        if (testf || trialf) {
            NCA_REQUIRE(testf && trialf, NCA_ERR_ARGUMENT, "nca_dh2_create_resident: factors of both sides needed");
            NCA_REQUIRE(testf->nids == d->ne_test && trialf->nids == d->ne_trial, NCA_ERR_DIMENSION,
                        "nca_dh2_create_resident: factors of another directional tree");
            NCA_REQUIRE(testf->ctx == ctx && trialf->ctx == ctx, NCA_ERR_ARGUMENT,
                        "nca_dh2_create_resident: factors of another context");
            set_device(ctx);
        }
        lower(true, d->nn_test, d->test_lo, d->test_hi, d->test_dirptr, d->test_cptr, d->nbnodes_test, d->bt_nodes,
              d->bt_dirptr, d->bt_diridx, d->bt_off, d->bt_blocks, lt, testf, &rj.tb);
        lower(false, d->nn_trial, d->trial_lo, d->trial_hi, d->trial_dirptr, d->trial_cptr, d->nbnodes_trial,
              d->bs_nodes, d->bs_dirptr, d->bs_diridx, d->bs_off, d->bs_blocks, ls, trialf, &rj.sb);
        if (testf) {
            transfer_jobs(testf, true, d->nlev_test, d->tt_levptr, d->tt_levidcs, d->tt_chdptr, d->tt_chds, d->tt_off,
                          d->test_cptr, rj.tt);
            transfer_jobs(trialf, false, d->nlev_trial, d->st_levptr, d->st_levidcs, d->st_chdptr, d->st_chds,
                          d->st_off, d->trial_cptr, rj.st);
        }
        // coupling plan: blocks of every test node, grouped by increasing test direction index
        const int64_t nblk = d->cp_nodeptr[d->nn_test] - d->cp_nodeptr[0];
        NCA_REQUIRE(nblk == coupling->nblocks, NCA_ERR_ARGUMENT, "nca_dh2_create: coupling plan differs from the store");
        std::vector<int64_t> cp_ptr(d->ne_test + 1, 0);
        int64_t last = -1;
        for (int64_t t = 0; t < d->nn_test; ++t)
            for (int64_t i = d->cp_nodeptr[t] - d->cp_nodeptr[0]; i < d->cp_nodeptr[t + 1] - d->cp_nodeptr[0]; ++i) {
                const int64_t e = d->cp_tidcs[i] - 1;
                NCA_REQUIRE(e >= d->test_dirptr[t] - 1 && e < d->test_dirptr[t + 1] - 1 && e >= last,
                            NCA_ERR_ARGUMENT,
                            "nca_dh2_create: coupling blocks not grouped by the test directions of their node");
                last = e;
                cp_ptr[e + 1]++;
            }
        for (int64_t e = 0; e < d->ne_test; ++e) cp_ptr[e + 1] += cp_ptr[e];
        for (auto& v : cp_ptr) v += 1;
        nca_h2_desc h{};
        h.M = d->M;
        h.N = d->N;
        h.nn_test = d->nn_test;
        h.nn_trial = d->nn_trial;
        h.test_perm = d->test_perm;
        h.test_lo = d->test_lo;
        h.test_hi = d->test_hi;
        h.trial_perm = d->trial_perm;
        h.trial_lo = d->trial_lo;
        h.trial_hi = d->trial_hi;
        h.test_cptr = d->test_cptr;
        h.trial_cptr = d->trial_cptr;
        h.nbasis_test = (int64_t)lt.nodes.size();
        h.nbasis_trial = (int64_t)ls.nodes.size();
        h.basis_test_nodes = lt.nodes.data();
        h.basis_test_off = lt.off.data();
        h.basis_trial_nodes = ls.nodes.data();
        h.basis_trial_off = ls.off.data();
        h.basis_test = reinterpret_cast<const double*>(lt.blocks.data());
        h.basis_trial = reinterpret_cast<const double*>(ls.blocks.data());
        h.nlev_test = d->nlev_test;
        h.nlev_trial = d->nlev_trial;
        h.tt_levptr = d->tt_levptr;
        h.tt_levidcs = d->tt_levidcs;
        h.tt_chdptr = d->tt_chdptr;
        h.tt_chds = d->tt_chds;
        h.tt_off = d->tt_off;
        h.st_levptr = d->st_levptr;
        h.st_levidcs = d->st_levidcs;
        h.st_chdptr = d->st_chdptr;
        h.st_chds = d->st_chds;
        h.st_off = d->st_off;
        h.tt_blocks = d->tt_blocks;
        h.st_blocks = d->st_blocks;
        h.cp_ptr = cp_ptr.data();
        h.cp_nodes = d->cp_sidcs + (d->cp_nodeptr[0] - 1);
        h.nearptr = d->nearptr;
        h.nearnodes = d->nearnodes;
        BlockLayout lay;
        lay.ne_test = d->ne_test;
        lay.ne_trial = d->ne_trial;
        lay.test_dirptr = d->test_dirptr;
        lay.trial_dirptr = d->trial_dirptr;
        if (testf) lay.jobs = &rj;
        create_core(ctx, kernel, test, trial, &h, lay, coupling, precision, out);
    }
}
}  // namespace nca

extern "C" {

int32_t nca_dh2_create(nca_ctx* ctx, const nca_helmholtz* kernel, const nca_points* test, const nca_points* trial,
                       const nca_dh2_desc* d, const nca_coupling* coupling, int32_t precision, nca_h2** out) {
    return guarded([&] { dh2_create_impl(ctx, kernel, test, trial, d, coupling, precision, nullptr, nullptr, out); });
}

int32_t nca_dh2_create_resident(nca_ctx* ctx, const nca_helmholtz* kernel, const nca_points* test,
                                const nca_points* trial, const nca_dh2_desc* d, const nca_coupling* coupling,
                                int32_t precision, const nca_factors* testf, const nca_factors* trialf,
                                nca_h2** out) {
    return guarded([&] {
        NCA_REQUIRE(testf && trialf, NCA_ERR_ARGUMENT, "nca_dh2_create_resident: invalid argument");
        dh2_create_impl(ctx, kernel, test, trial, d, coupling, precision, testf, trialf, out);
    });
}

int32_t nca_h2_destroy(nca_h2* h2) {
    return guarded([&] {
        if (!h2) return;
        set_device(h2->ctx);
        (void)hipStreamSynchronize(h2->ctx->stream);
        delete h2;
    });
}

int32_t nca_h2_set_precision(nca_h2* h2, int32_t precision) {
    return guarded([&] {
        NCA_REQUIRE(h2, NCA_ERR_ARGUMENT, "nca_h2_set_precision: null operator");
        NCA_REQUIRE(precision == NCA_PREC_FP64 || precision == NCA_PREC_MIXED, NCA_ERR_ARGUMENT,
                    "nca_h2_set_precision: unknown precision");
        std::lock_guard<std::mutex> lock(h2->mu);
        h2->precision = precision;
        // the c32 operator copies are made here, so that later products
        // (nca_h2_matvec_device_noalloc among them) allocate nothing for them
        if (precision == NCA_PREC_MIXED) {
            set_device(h2->ctx);
            ensure_mixed(h2);
        }
    });
}

int32_t nca_h2_matvec_device(nca_h2* h2, int32_t op, int64_t nrhs, const void* x, int64_t ldx, void* y,
                             int64_t ldy, const double* alpha, const double* beta) {
    return guarded([&] {
        NCA_REQUIRE(h2 && (nrhs == 0 || (x && y)), NCA_ERR_ARGUMENT, "nca_h2_matvec_device: invalid argument");
        std::lock_guard<std::mutex> lock(h2->mu);
        set_device(h2->ctx);
        matvec_device(h2, op, nrhs, static_cast<const double2*>(x), ldx, static_cast<double2*>(y), ldy, alpha, beta);
    });
}

int32_t nca_h2_store_near(nca_h2* h2, int32_t on) {
    return guarded([&] {
        NCA_REQUIRE(h2, NCA_ERR_ARGUMENT, "nca_h2_store_near: null operator");
        std::lock_guard<std::mutex> lock(h2->mu);
        set_device(h2->ctx);
        if (on) {
            h2->near_stored = true;
            ensure_near_store(h2);
        } else if (h2->near_stored || h2->ns_plan) {
            ++h2->generation;  // a captured sweep may read the entries or ws_ynear
            NCA_HIP(hipStreamSynchronize(h2->ctx->stream));
            h2->near_stored = h2->ns_plan = false;
            h2->ns_blocks.release();
            h2->ns_blocks32.release();
            h2->ws_ynear.release();
            h2->ns_colcoef.release();
            h2->ns_nleaf = h2->ns_ntiles = h2->ns_entries = 0;
        }
    });
}

int32_t nca_h2_matvec_device_noalloc(nca_h2* h2, int32_t op, int64_t nrhs, const void* x, int64_t ldx, void* y,
                                     int64_t ldy, const double* alpha, const double* beta) {
    return guarded([&] {
        NCA_REQUIRE(h2 && (nrhs == 0 || (x && y)), NCA_ERR_ARGUMENT,
                    "nca_h2_matvec_device_noalloc: invalid argument");
        std::lock_guard<std::mutex> lock(h2->mu);
        set_device(h2->ctx);
        void* base = scratch_address(h2->ctx->device);
        if (!check_product(h2, op, nrhs, ldx, ldy)) return;
        NCA_REQUIRE(h2->precision != NCA_PREC_MIXED || h2->mixed_ready, NCA_ERR_ARGUMENT,
                    "nca_h2_matvec_device_noalloc: the MIXED operator copies are missing (set the precision with "
                    "nca_h2_set_precision first)");
        nca_h2::WsView v;
        const size_t need = carve_workspace(h2, nrhs, nullptr, v);
        NCA_REQUIRE(need <= kScratchBytes, NCA_ERR_LIMIT,
                    "nca_h2_matvec_device_noalloc: the product needs " + std::to_string(need) +
                        " bytes of workspace, above the " + std::to_string(kScratchBytes) +
                        "-byte scratch (fewer right-hand sides per call, or nca_h2_matvec_device)");
        carve_workspace(h2, nrhs, static_cast<char*>(base), v);
        h2->ws = v;
        run_sweep(h2, op, nrhs, static_cast<const double2*>(x), ldx, static_cast<double2*>(y), ldy, alpha, beta,
                  nullptr);
    });
}

int32_t nca_h2_graph_create(nca_h2* h2, int32_t op, int64_t nrhs, const void* x, int64_t ldx, void* y,
                            int64_t ldy, const double* alpha, const double* beta, nca_h2_graph** out) {
    return guarded([&] {
        NCA_REQUIRE(h2 && out && nrhs > 0 && x && y, NCA_ERR_ARGUMENT, "nca_h2_graph_create: invalid argument");
        std::lock_guard<std::mutex> lock(h2->mu);
        set_device(h2->ctx);
        hipStream_t st = h2->ctx->stream;
        // everything that allocates or converts runs before the capture
        if (h2->precision == NCA_PREC_MIXED) ensure_mixed(h2);
        if (op != NCA_OP_N && !couple_otf(h2)) ensure_reverse(h2);
        if (op == NCA_OP_N && h2->near_stored) ensure_near_store(h2);
        ensure_workspace(h2, nrhs);
        NCA_HIP(hipStreamSynchronize(st));
        auto g = std::make_unique<nca_h2_graph>();
        g->h2 = h2;
        g->precision = h2->precision;
        g->generation = h2->generation;
        NCA_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
        try {
            matvec_device(h2, op, nrhs, static_cast<const double2*>(x), ldx, static_cast<double2*>(y), ldy, alpha,
                          beta);
        } catch (...) {
            hipGraph_t broken = nullptr;
            (void)hipStreamEndCapture(st, &broken);
            if (broken) (void)hipGraphDestroy(broken);
            throw;
        }
        NCA_HIP(hipStreamEndCapture(st, &g->graph));
        NCA_HIP(hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0));
        *out = g.release();
    });
}

int32_t nca_h2_graph_launch(nca_h2_graph* g) {
    return guarded([&] {
        NCA_REQUIRE(g && g->exec, NCA_ERR_ARGUMENT, "nca_h2_graph_launch: invalid argument");
        std::lock_guard<std::mutex> lock(g->h2->mu);
        NCA_REQUIRE(g->h2->precision == g->precision, NCA_ERR_ARGUMENT,
                    "nca_h2_graph_launch: the operator's precision changed since the capture");
        NCA_REQUIRE(g->h2->generation == g->generation, NCA_ERR_ARGUMENT,
                    "nca_h2_graph_launch: the operator's workspace or launch plans were reallocated since the "
                    "capture (a product with more right-hand sides, or nca_h2_autotune); capture the sweep again");
        set_device(g->h2->ctx);
        NCA_HIP(hipGraphLaunch(g->exec, g->h2->ctx->stream));
    });
}

int32_t nca_h2_graph_destroy(nca_h2_graph* g) {
    return guarded([&] {
        if (!g) return;
        set_device(g->h2->ctx);
        if (g->exec) (void)hipGraphExecDestroy(g->exec);
        if (g->graph) (void)hipGraphDestroy(g->graph);
        delete g;
    });
}

int32_t nca_h2_autotune(nca_h2* h2, int32_t reps, nca_h2_tuning* out) {
    return guarded([&] {
        NCA_REQUIRE(h2 && reps > 0, NCA_ERR_ARGUMENT, "nca_h2_autotune: invalid argument");
        std::lock_guard<std::mutex> lock(h2->mu);
        if (h2->M == 0 || h2->N == 0) {
            if (out) *out = nca_h2_tuning{};
            return;
        }
        set_device(h2->ctx);
        hipStream_t st = h2->ctx->stream;
        // the tuner reorders the launch plans and may reallocate them
        ++h2->generation;
        const int prec = h2->precision == NCA_PREC_MIXED ? 1 : 0;
        DevBuf<double2> x((size_t)std::max<int64_t>(h2->N * 16, 1)), y((size_t)std::max<int64_t>(h2->M * 16, 1));
        hipLaunchKernelGGL(k_fill, dim3(blocks_for((int64_t)x.n, 256)), dim3(256), 0, st, x.p, (int64_t)x.n);
        hipLaunchKernelGGL(k_fill, dim3(blocks_for((int64_t)y.n, 256)), dim3(256), 0, st, y.p, (int64_t)y.n);
        double c0 = 0, n0 = 0;
        time_stages(h2, x.p, y.p, reps, c0, n0);  // warm-up and the starting configuration
        time_stages(h2, x.p, y.p, reps, c0, n0);
        nca_h2_tuning r{};
        r.couple_ms_start = c0;
        r.near_ms_start = n0;
        int count = 0;
        // near field: products x leaf order x workgroup order
        int bk3 = h2->near_k3[prec], blo = h2->leaf_order, bx = h2->near_xcd, bst = h2->near_stagger;
        double best = n0;
        // MIXED: also the f16 near field (near_k3 = 2) when the operator allows it
        const int nk3 = prec == 1 && h2->h16 == 1 ? 3 : 2;
        for (int k3 = 0; k3 < nk3; ++k3)
            for (int lo = 0; lo < 3; ++lo)
                for (int xo = 0; xo < 4; ++xo) {
                    h2->near_k3[prec] = k3;
                    h2->near_xcd = xo & 1;
                    h2->near_stagger = xo >> 1;
                    if (h2->leaf_order != lo) {
                        order_leaves(h2->test, lo, st);
                        order_leaves(h2->trial, lo, st);
                        h2->leaf_order = lo;
                    }
                    double c, n;
                    time_stages(h2, x.p, y.p, reps, c, n);
                    ++count;
                    if (n < best) { best = n; bk3 = k3; blo = lo; bx = xo & 1; bst = xo >> 1; }
                }
        h2->near_k3[prec] = bk3;
        h2->near_xcd = bx;
        h2->near_stagger = bst;
        if (h2->leaf_order != blo) {
            order_leaves(h2->test, blo, st);
            order_leaves(h2->trial, blo, st);
            h2->leaf_order = blo;
        }
        r.near_ms = best;
        // coupling: products x tile order x workgroup order
        int ck3 = h2->couple_k3[prec], co = h2->couple_order, cx = h2->couple_xcd;
        double cbest = c0;
        // MIXED: also the coupling evaluated on the fly (couple_k3 = 2); the
        // stored forms need the c32 copy
        const int ck = prec == 1 && h2->coup_h16 == 1 ? 3 : 2;
        if (prec == 1) ensure_s32(h2);
        if (h2->f_ntiles > 0)
            for (int k3 = 0; k3 < ck; ++k3)
                for (int o = 0; o < 4; ++o)
                    for (int xo = 0; xo < 2; ++xo) {
                        h2->couple_k3[prec] = k3;
                        h2->couple_xcd = xo;
                        if (h2->couple_order != o) {
                            order_coupling_tiles(h2, o, st);
                            h2->couple_order = o;
                        }
                        double c, n;
                        time_stages(h2, x.p, y.p, reps, c, n);
                        ++count;
                        if (c < cbest) { cbest = c; ck3 = k3; co = o; cx = xo; }
                    }
        h2->couple_k3[prec] = ck3;
        h2->couple_xcd = cx;
        if (h2->couple_order != co) {
            order_coupling_tiles(h2, co, st);
            h2->couple_order = co;
        }
        r.couple_ms = cbest;
        r.near_k3 = bk3;
        r.leaf_order = blo;
        r.near_xcd_order = bx;
        r.near_stagger = bst;
        r.couple_k3 = ck3;
        r.couple_tile_order = co;
        r.couple_xcd_order = cx;
        r.candidates = count;
        NCA_HIP(hipStreamSynchronize(st));
        if (out) *out = r;
    });
}

int32_t nca_h2_tuning_apply(nca_h2* h2, const nca_h2_tuning* t) {
    return guarded([&] {
        NCA_REQUIRE(h2 && t, NCA_ERR_ARGUMENT, "nca_h2_tuning_apply: invalid argument");
        NCA_REQUIRE(t->near_k3 >= 0 && t->near_k3 <= (h2->precision == NCA_PREC_MIXED ? 2 : 1) &&
                        t->couple_k3 >= 0 && t->couple_k3 <= (h2->precision == NCA_PREC_MIXED ? 2 : 1) &&
                        t->leaf_order >= 0 && t->leaf_order <= 2 && t->near_xcd_order >= 0 &&
                        t->near_xcd_order <= 1 && t->near_stagger >= 0 && t->near_stagger <= 1 &&
                        t->couple_tile_order >= 0 && t->couple_tile_order <= 3 && t->couple_xcd_order >= 0 &&
                        t->couple_xcd_order <= 1,
                    NCA_ERR_ARGUMENT, "nca_h2_tuning_apply: a launch choice out of range");
        std::lock_guard<std::mutex> lock(h2->mu);
        set_device(h2->ctx);
        hipStream_t st = h2->ctx->stream;
        ++h2->generation;  // the launch plans may be reordered
        const int prec = h2->precision == NCA_PREC_MIXED ? 1 : 0;
        h2->near_k3[prec] = t->near_k3;
        h2->near_xcd = t->near_xcd_order;
        h2->near_stagger = t->near_stagger;
        if (h2->leaf_order != t->leaf_order) {
            order_leaves(h2->test, t->leaf_order, st);
            order_leaves(h2->trial, t->leaf_order, st);
            h2->leaf_order = t->leaf_order;
        }
        h2->couple_k3[prec] = t->couple_k3;
        // a stored MIXED coupling form needs the c32 copy
        if (prec == 1 && h2->mixed_ready && !couple_otf(h2)) ensure_s32(h2);
        h2->couple_xcd = t->couple_xcd_order;
        if (h2->f_ntiles > 0 && h2->couple_order != t->couple_tile_order) {
            order_coupling_tiles(h2, t->couple_tile_order, st);
            h2->couple_order = t->couple_tile_order;
        }
        NCA_HIP(hipStreamSynchronize(st));
    });
}

int32_t nca_h2_matvec_profile(nca_h2* h2, int32_t op, int64_t nrhs, const void* x, int64_t ldx, void* y,
                              int64_t ldy, const double* alpha, const double* beta, double* stage_ms) {
    return guarded([&] {
        NCA_REQUIRE(h2 && stage_ms && (nrhs == 0 || (x && y)), NCA_ERR_ARGUMENT,
                    "nca_h2_matvec_profile: invalid argument");
        NCA_REQUIRE(nrhs <= 16, NCA_ERR_ARGUMENT, "nca_h2_matvec_profile: at most 16 right-hand sides");
        std::lock_guard<std::mutex> lock(h2->mu);
        set_device(h2->ctx);
        StageEvents evs;
        for (auto& e : evs.ev) NCA_HIP(hipEventCreate(&e));
        evs.on = true;
        try {
            matvec_device(h2, op, nrhs, static_cast<const double2*>(x), ldx, static_cast<double2*>(y), ldy, alpha,
                          beta, &evs);
            NCA_HIP(hipEventSynchronize(evs.ev[6]));
            for (int i = 0; i < 6; ++i) {
                float ms = 0.f;
                NCA_HIP(hipEventElapsedTime(&ms, evs.ev[i], evs.ev[i + 1]));
                stage_ms[i] = ms;
            }
        } catch (...) {
            for (auto& e : evs.ev) (void)hipEventDestroy(e);
            throw;
        }
        for (auto& e : evs.ev) (void)hipEventDestroy(e);
    });
}

int32_t nca_h2_matvec(nca_h2* h2, int32_t op, int64_t nrhs, const double* x, int64_t ldx, double* y, int64_t ldy,
                      const double* alpha, const double* beta) {
    return guarded([&] {
        NCA_REQUIRE(h2 && (nrhs == 0 || (x && y)), NCA_ERR_ARGUMENT, "nca_h2_matvec: invalid argument");
        NCA_REQUIRE(op >= NCA_OP_N && op <= NCA_OP_C, NCA_ERR_ARGUMENT, "nca_h2_matvec: op must be 0 (N), 1 (T) or 2 (C)");
        std::lock_guard<std::mutex> lock(h2->mu);
        set_device(h2->ctx);
        const int64_t nx = op == NCA_OP_N ? h2->N : h2->M, ny = op == NCA_OP_N ? h2->M : h2->N;
        NCA_REQUIRE(ldx >= nx && ldy >= ny, NCA_ERR_DIMENSION, "DimensionMismatch: leading dimension too small");
        if (nrhs == 0 || ny == 0) return;
        hipStream_t st = h2->ctx->stream;
        h2->ws_xin.reserve((size_t)std::max<int64_t>(nx * nrhs, 1));
        h2->ws_yio.reserve((size_t)(ny * nrhs));
        if (nx > 0)
            NCA_HIP(hipMemcpy2DAsync(h2->ws_xin.p, nx * sizeof(double2), x, ldx * sizeof(double2), nx * sizeof(double2),
                                     nrhs, hipMemcpyHostToDevice, st));
        const bool usey = beta && (beta[0] != 0.0 || beta[1] != 0.0);
        if (usey)
            NCA_HIP(hipMemcpy2DAsync(h2->ws_yio.p, ny * sizeof(double2), y, ldy * sizeof(double2), ny * sizeof(double2),
                                     nrhs, hipMemcpyHostToDevice, st));
        matvec_device(h2, op, nrhs, h2->ws_xin.p, nx, h2->ws_yio.p, ny, alpha, beta);
        NCA_HIP(hipMemcpy2DAsync(y, ldy * sizeof(double2), h2->ws_yio.p, ny * sizeof(double2), ny * sizeof(double2),
                                 nrhs, hipMemcpyDeviceToHost, st));
        NCA_HIP(hipStreamSynchronize(st));
    });
}

int32_t nca_h2_get_stats(const nca_h2* h2, nca_h2_stats* out) {
    return guarded([&] {
        NCA_REQUIRE(h2 && out, NCA_ERR_ARGUMENT, "nca_h2_get_stats: invalid argument");
        std::lock_guard<std::mutex> lock(h2->mu);
        *out = h2->stats;
    });
}

int32_t nca_h2_couple_form(nca_h2* h2, int32_t* form, double* rmin, double* rmax) {
    return guarded([&] {
        NCA_REQUIRE(h2 && form, NCA_ERR_ARGUMENT, "nca_h2_couple_form: invalid argument");
        std::lock_guard<std::mutex> lock(h2->mu);
        set_device(h2->ctx);
        int32_t f = 0;
        if (h2->precision == NCA_PREC_MIXED) {
            ensure_mixed(h2);
            f = couple_otf(h2) ? 2 : 1;
        }
        *form = f;
        if (rmin) *rmin = h2->coup_rmin;
        if (rmax) *rmax = h2->coup_rmax;
    });
}

int32_t nca_h2_rel32(nca_h2* h2, int32_t mode, int32_t* near_on, int32_t* couple_on, double* near_frac,
                     double* couple_frac) {
    return guarded([&] {
        NCA_REQUIRE(h2 && mode >= -1 && mode <= 1, NCA_ERR_ARGUMENT, "nca_h2_rel32: invalid argument");
        std::lock_guard<std::mutex> lock(h2->mu);
        set_device(h2->ctx);
        if (mode >= 0 && h2->rel32_off != (mode == 0)) {
            ++h2->generation;  // captured sweeps hold the previous form
            h2->rel32_off = mode == 0 ? 1 : 0;
        }
        const bool mix = h2->precision == NCA_PREC_MIXED;
        if (mix) ensure_mixed(h2);
        const bool on = mix && !h2->rel32_off;
        // in use: the f16 forms run (forms 2 of nca_h2_near_form / _couple_form)
        // with every phase below one revolution
        const bool nf = on && !h2->near_stored && h2->near_k3[1] == 2 && h2->h16 == 1 &&
                        h2->near_turns <= kPhase32Turns;
        const bool cf = on && couple_otf(h2) && std::fabs(h2->k) * 0.5 / M_PI * h2->coup_rmax <= kPhase32Turns;
        if (near_on) *near_on = nf && h2->near_rel_frac > 0.0 ? 1 : 0;
        if (couple_on) *couple_on = cf && h2->coup_rel_frac > 0.0 ? 1 : 0;
        if (near_frac) *near_frac = mix ? h2->near_rel_frac : 0.0;
        if (couple_frac) *couple_frac = mix ? h2->coup_rel_frac : 0.0;
    });
}

int32_t nca_h2_near_form(nca_h2* h2, int32_t* form, double* rmin, double* rmax) {
    return guarded([&] {
        NCA_REQUIRE(h2 && form, NCA_ERR_ARGUMENT, "nca_h2_near_form: invalid argument");
        std::lock_guard<std::mutex> lock(h2->mu);
        set_device(h2->ctx);
        int32_t f = 0;
        if (h2->precision == NCA_PREC_MIXED) {
            ensure_mixed(h2);
            f = h2->near_stored ? 1 : (h2->near_k3[1] == 2 && h2->h16 == 1 ? 2 : 1);
        }
        *form = f;
        if (rmin) *rmin = h2->near_rmin;
        if (rmax) *rmax = h2->near_rmax_pairs > 0.0 ? h2->near_rmax_pairs : h2->near_rmax;
    });
}

}  // extern "C"
