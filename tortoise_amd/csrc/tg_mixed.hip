// tg_mixed.hip -- batched mixed-precision collision operator on MFMA.
//
// For B members at once the operator needs U = K Z and V = K^T Z (Z = n x B,
// z = f or 1 - f per panel and member).  Both come from one GEMM with the
// stacked matrix A = [K; K^T] (2n x n):
//
//   C = A Z,  A = 2^-sa (A_hi + A_lo),  Z = 2^-sf (Z_hi + Z_lo)   (fp16 parts)
//   C ~ 2^-(sa+sf) (A_hi Z_hi + A_hi Z_lo + A_lo Z_hi)
//
// three v_mfma_f32_32x32x16_f16 products per tile; the FP32 sums of at most
// 32*FOLD products are folded into FP64 registers, the epilogue works in FP64
// (per-member bath occupation, Pauli blocking, number-defect removal).  The
// row and column sums K 1, K^T 1 used by the epilogue are exact FP64 sums of
// the *split* matrix, so the split error cancels in K 1 - K g.
//
// Node order: K is stored in energy order (see tg_assemble.hip), which makes
// the emission kernel a band: every 128-row panel of A has a contiguous
// column range outside which it is exactly zero, and the GEMM only streams
// that range.  Inputs and outputs stay in the caller's node order; the gather
// kernel copies f into energy order and the epilogue scatters back.
//
// Packed layouts (16-byte units of eight fp16 values along the reduction k):
//   A band: per panel its band only, iteration-major, each iteration one
//     16 KB block [kg in 0..3][hi, lo][row in 0..127] (kg = k/8 within the
//     iteration); the panel's blocks are contiguous from band_off[panel]
//   Z_{hi,lo}[it][col][kgl]   it = kg/4, kgl = kg%4, col = member in [0, nt):
//     a member's four k-groups of one iteration are 64 contiguous bytes (the
//     pack kernel writes whole 64-byte pieces), and a tile of members of one
//     iteration is one contiguous block
// Every MFMA fragment and LDS tile is a contiguous 16-byte read per lane in
// these layouts (lanes 0-31: k-group 2s, lanes 32-63: k-group 2s+1), and a
// panel streams its band as consecutive 16 KB blocks.  The band is packed
// straight from K's entries (tg_kentry.hpp): no dense copy of A exists.
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <numeric>
#include <vector>

#include "tg_dp5_kernels.hpp"
#include "tg_internal.hpp"

namespace tg {

typedef _Float16 half8 __attribute__((ext_vector_type(8)));
typedef float floatx16 __attribute__((ext_vector_type(16)));
typedef float floatx4 __attribute__((ext_vector_type(4)));
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));

constexpr int kMT = 128;        // rows of A per workgroup
constexpr int kKS = 32;         // reduction depth per iteration (two MFMA k-steps)
constexpr int kKG = kKS / 8;    // 16-byte k-groups per iteration
constexpr int kMixThreads = 256;
constexpr int kSF = 12;         // F scale exponent: f in [-8, 8] stays inside fp16
// a persistent work unit is (panel << 1) | chunk: a panel's band is one
// chunk or two (partial-sum slots 0 and 1)
constexpr int kMaxChunks = 2;

struct MixedData {
    int64_t rows = 0;          // 2n
    int64_t kt = 0;            // first row of the K^T half (n rounded up to kMT)
    int64_t rows_pad = 0;      // 2 kt
    int64_t kpad = 0;          // n rounded up to kKS
    int nt = 256;              // members per GEMM tile (padded batch)
    int sa = 0;                // A scale exponent
    int npanel = 0;
    uint4* A_band = nullptr;   // band-major packed A (hi and lo)
    int64_t* band_off = nullptr;    // per panel: first 16-byte unit of its band
    int64_t band_units = 0;
    uint4* Z_all = nullptr;    // F_hi, F_lo, G_hi, G_lo back to back (one allocation)
    uint4* F_hi = nullptr;
    uint4* F_lo = nullptr;
    uint4* G_hi = nullptr;     // operand g = 1 - f
    uint4* G_lo = nullptr;
    uint8_t* useg = nullptr;   // npanel x nt operand choice
    int2* krange = nullptr;    // per panel [kit_begin, kit_end) in iterations
    std::vector<int2> h_krange;
    const int32_t* perm = nullptr;  // sorted position -> node (owned by the context)
    double* rk_s = nullptr;    // split-matrix K 1 (sorted order)
    double* ck_s = nullptr;    // split-matrix K^T 1 (sorted order)
    double* rk_c = nullptr;    // the same in the caller's order (the hat form's defect weights)
    double* ck_c = nullptr;
    const double* V_s = nullptr;    // node volumes (sorted order, owned by the context)
    const double* eps_s = nullptr;  // node energies (sorted order, owned by the context)
    uint8_t* P = nullptr;      // slots x nt x rows_pad partial sums (FP64, FP32 if fold == 0)
    int nsplit = 1;            // K-splits of the tiled kernel (its slots)
    // the persistent kernel's work units: a band longer than cut_min
    // iterations is cut in two chunks (partial-sum slots 0 and 1)
    int cut_min = 0;
    std::vector<int32_t> h_cut;      // per panel: iterations of chunk 0 (the band length when not cut)
    int32_t* cut = nullptr;
    int nunit = 0;                   // units: (panel << 1) | chunk
    int pslots = 1;                  // slots of the persistent kernel (2 when some band is cut)
    uint8_t* allmask = nullptr;      // per panel: the slots a schedule of every unit writes
    int2* srange = nullptr;          // per unit: [first iteration, count] of its whole chunk
    int2* durange = nullptr;         // per unit: the trimmed range the last schedule gave it
    int nw = 128;              // members per workgroup tile
    int fold = 2;              // iterations per FP64 fold (0: FP32 partials)
    int waves = 4;             // waves per workgroup
    int depth = 2;             // iterations of register prefetch
    bool persist = true;       // persistent balanced kernel (else the tiled grid kernel)
    int ncu = 0;               // compute units of the device
    int da = 2;                // persistent kernel: iterations of register prefetch
    int al = 0;                // persistent kernel: A staged through LDS
    int sched_nw = 0;          // tile width the schedule below was built for (0: none)
    int bins_per_xcd = 0;
    int32_t* sched_off = nullptr;    // per bin: first entry in sched_panel (nbins + 1)
    int32_t* sched_panel = nullptr;  // work units in bin order
    // zero operands: which iterations hold a nonzero f (g) operand of a live
    // member, which panels some member runs with f (g); from them the units
    // with work, their iteration ranges and their schedule, rebuilt on the
    // device per evaluation
    bool skip = true;
    int32_t nit = 0;                 // iterations of 32 k over the padded k range
    uint8_t* nzF = nullptr;          // nit
    uint8_t* nzG = nullptr;          // nit
    uint8_t* useF = nullptr;         // npanel
    uint8_t* useG = nullptr;         // npanel
    uint8_t* actp = nullptr;         // npanel: the slots (chunks) the last schedule computed
    int32_t* lpt_order = nullptr;    // units by decreasing length
    int32_t* dsched_off = nullptr;   // nbins + 1
    int32_t* dsched_panel = nullptr; // nunit
    int64_t* dstat = nullptr;        // the last schedule's iterations and nonzeros of [K; K^T]
    int64_t* nnz_pref = nullptr;     // exclusive prefix of K's nonzeros over the band's blocks
    int32_t* blk_first = nullptr;    // per panel: its first block of the band
    int64_t band_nnz = 0;           // nonzeros of [K; K^T] in the band
    int32_t* dslot = nullptr;        // nt: a fused DP5 attempt's column of each member (-1: not running)
    int32_t* dncol = nullptr;        // 1: the columns running members take (leading ones)
    float tuned_ms = 0.0f;     // GEMM + epilogue time of the configuration the autotuner picked (0: not tuned)
    size_t P_cap = 0;          // bytes
    double* bath = nullptr;    // per-member bath occupation (device)
    std::vector<double> bath_set;   // host copy of what `bath` holds (empty: unknown)
    double* xb = nullptr;      // hat-averaged occupation factors of the packed states (nt x n,
                               // member-major; null with nodal factors)
    bool rowcache = true;      // member rows fit in LDS
    int64_t work_items = 0;
};

static MixedData* md_of(tg_ctx* c) { return (MixedData*)c->mixed; }

// partial-sum slots of the configured kernel, and the per-panel slot masks
// the epilogue reads (the last schedule's, the static ones of a schedule of
// every unit, or null: every panel has all slots of the tiled kernel)
static int slots(const MixedData* m) { return m->persist ? m->pslots : m->nsplit; }
static const uint8_t* slot_masks(const MixedData* m, bool dyn) {
    return m->persist ? (dyn ? m->actp : m->allmask) : nullptr;
}
// dynamic LDS of active_schedule_kernel
static size_t sched_lds(const MixedData* m) {
    return sizeof(int32_t) * (2 * ((size_t)m->nit + 1) + m->npanel + m->nunit);
}

// ------------------------------------------------------------------ packing
// Band-major packed A from K's entries (tg_kentry.hpp): block (panel,
// iteration) holds the k-groups 4 it .. 4 it + 3 of the panel's band, hi and
// lo, 128 rows each.  A[r][q] = K[r][q] for r < n; the K^T half starts at
// the panel-aligned row kt: A[kt + r][q] = K[q][r].  Thread (row, k-group)
// evaluates its 8 entries and writes their hi and lo units.
constexpr int kBandT = 2 * kKG * kMT;
__global__ void __launch_bounds__(kKG * kMT)
band_pack_kernel(KSrc k, const int32_t* __restrict__ perm, int64_t n, int64_t kt, double scale,
                 const int2* __restrict__ krange, const int64_t* __restrict__ band_off,
                 uint4* __restrict__ band, int32_t* __restrict__ blk_nnz) {
    __shared__ int wsum[kBandT / 2 / 64];
    const int panel = blockIdx.x;
    const int itl = blockIdx.y;
    const int2 kr = krange[panel];
    if (itl >= kr.y - kr.x) return;
    const int r = threadIdx.x % kMT, kg = threadIdx.x / kMT;
    const int64_t row = (int64_t)panel * kMT + r;
    const int64_t col0 = ((int64_t)(kr.x + itl) * kKG + kg) * 8;
    union { half8 v; uint4 u; } h, l;
    int nz = 0;
    for (int q = 0; q < 8; ++q) {
        const int64_t col = col0 + q;
        double x = 0.0;
        if (col < n) {
            if (row < n) x = k_entry(k, perm[row], perm[col]);
            else if (row >= kt && row < kt + n) x = k_entry(k, perm[col], perm[row - kt]);
        }
        nz += x != 0.0;
        _Float16 a, b;
        split_fp16(x, scale, a, b);
        h.v[q] = a;
        l.v[q] = b;
    }
    uint4* dst = band + band_off[panel] + (int64_t)itl * kBandT;
    dst[(kg * 2 + 0) * kMT + r] = h.u;
    dst[(kg * 2 + 1) * kMT + r] = l.u;
    // K's (structural) nonzeros in this block
    for (int o = 32; o > 0; o >>= 1) nz += __shfl_down(nz, o, 64);
    if ((threadIdx.x & 63) == 0) wsum[threadIdx.x >> 6] = nz;
    __syncthreads();
    if (threadIdx.x == 0) {
        int t = 0;
        for (int w = 0; w < kBandT / 2 / 64; ++w) t += wsum[w];
        blk_nnz[band_off[panel] / kBandT + itl] = t;
    }
}

// exact FP64 row sums of the split matrix over each row's band (ascending
// columns): rows [0, n) give K 1, rows [kt, kt + n) give K^T 1 (sorted order)
__global__ void band_rowsum_kernel(const uint4* __restrict__ band, const int64_t* __restrict__ band_off,
                                   const int2* __restrict__ krange, int64_t rows_pad, int64_t n, int64_t kt,
                                   double inv_scale, double* __restrict__ rk, double* __restrict__ ck) {
    const int64_t row = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (row >= rows_pad || (row >= n && row < kt) || row >= kt + n) return;
    const int panel = (int)(row / kMT), r = (int)(row % kMT);
    const int2 kr = krange[panel];
    const uint4* b0 = band + band_off[panel];
    double s = 0.0;
    for (int it = 0; it < kr.y - kr.x; ++it)
        for (int kg = 0; kg < kKG; ++kg) {
            union { half8 v; uint4 u; } h, l;
            h.u = b0[(int64_t)it * kBandT + (kg * 2 + 0) * kMT + r];
            l.u = b0[(int64_t)it * kBandT + (kg * 2 + 1) * kMT + r];
            for (int q = 0; q < 8; ++q) s += (double)h.v[q] + (double)l.v[q];
        }
    s *= inv_scale;
    if (row < n) rk[row] = s;
    else ck[row - kt] = s;
}

// Member-per-workgroup kernels.  Workgroup w serves member
// b = (w % 8) * (nt / 8) + w / 8, so the members one XCD serves are
// consecutive and the 16-byte units of eight members that share a 128-byte
// line of the packed operands meet in one L2.  With ROW the member's row of
// f (caller order) is staged in LDS by coalesced reads and the energy-order
// gathers read LDS; without it (rows larger than the LDS) they read global
// memory.
constexpr int kMemT = 1024;

__device__ inline int member_of_block(int w, int nt) { return (w % 8) * (nt / 8) + w / 8; }

// A member row moves between global memory and LDS in 16-byte pieces; a row
// that starts 8 bytes past a 16-byte boundary moves its first element alone,
// and an odd remainder its last.
__device__ inline int64_t row_head(const double* p) { return ((uintptr_t)p & 15) ? 1 : 0; }

// Pieces per thread of the staging: every row that fits the LDS
// (n <= 20352) is staged by one round of 1024 x 10 16-byte loads.
constexpr int kStageUnr = 10;
// k-groups of 8 positions per thread in the staged pack
constexpr int kPackKG = 3;

// stage row b of f (caller order) in LDS, n <= 2 kMemT kStageUnr: all of a
// thread's loads are in flight before any is stored (indices clamp to the
// last piece so the loads issue unconditionally; only the stores are
// guarded).  There is no loop, so loads the caller issued before stay in
// flight; the caller issues the barrier (lds_barrier keeps them in flight
// across it too).
__device__ inline void stage_member_row(const double* __restrict__ fb, int64_t n, double* row) {
    const int64_t h = row_head(fb);
    const int64_t np = (n - h) >> 1;
    const double2* __restrict__ src = reinterpret_cast<const double2*>(fb + h);
    if (np > 0) {
        double2 v[kStageUnr];
#pragma unroll
        for (int j = 0; j < kStageUnr; ++j) v[j] = src[min((int64_t)threadIdx.x + j * kMemT, np - 1)];
#pragma unroll
        for (int j = 0; j < kStageUnr; ++j) {
            const int64_t q = (int64_t)threadIdx.x + j * kMemT;
            if (q < np) {
                row[h + 2 * q] = v[j].x;
                row[h + 2 * q + 1] = v[j].y;
            }
        }
    }
    if (threadIdx.x == 0 && h) row[0] = fb[0];
    if (threadIdx.x == kMemT - 1 && h + 2 * np < n) row[n - 1] = fb[n - 1];
}

// Workgroup barrier for LDS data only: waits for this wave's LDS stores, not
// for its global loads, so loads issued before it stay in flight (the
// compiler inserts the waits for them where their values are used).
__device__ inline void lds_barrier() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// Operand of (panel, member): g = 1 - f when the panel's middle node is in
// the filled part of the member's distribution (f > 1/2), else f.  The
// products the epilogue needs come back as K z and K 1 - K z, so the
// Pauli-blocked out-scattering K^T(1 - f) of filled nodes is a sum of small
// terms instead of the difference of two large ones.
struct PackOut {
    u32x4* Fhi;
    u32x4* Flo;
    u32x4* Ghi;
    u32x4* Glo;
    uint8_t* useg;          // npanel x nt operand choice
    uint8_t* nzF;           // null, or per iteration: holds a nonzero f (g) operand
    uint8_t* nzG;
    uint8_t* useF;          // per panel: some live member runs it with f (g)
    uint8_t* useG;
    // hat-averaged occupation factors of the packed state, member-major n
    // values per member, from the stencil hs (null: nodal factors, not written)
    const uint32_t* hs;
    double* xb;
    int dim;
};

// member b's occupation factors x = S f from its state in LDS `row` (caller
// order, complete) into o.xb
__device__ __forceinline__ void hat_from_row(const double* row, int64_t n, const PackOut& o, int b) {
    double* __restrict__ xb = o.xb + (int64_t)b * n;
    for (int64_t i = threadIdx.x; i < n; i += kMemT)
        xb[i] = hat_value_dim(o.dim, o.hs, i, [&](int64_t j) { return row[j]; });
}

// the caller-order node at the middle of panel pn (K rows and K^T rows index
// the same sorted nodes)
__device__ inline int32_t panel_node(int pn, int64_t kt, int64_t n, const int32_t* __restrict__ perm) {
    const int64_t r = (int64_t)pn * kMT + kMT / 2;
    const int64_t p = r < kt ? r : r - kt;
    return perm[p > n - 1 ? n - 1 : p];
}

// k-group kg of member b: the values xv of its 8 energy-order positions
// (padding and idle members pack as zero) -> fp16 hi/lo of f and g = 1 - f
__device__ __forceinline__ void pack_group(int64_t kg, const double (&xv)[8], bool live, int64_t n, int b, int nt,
                                           const PackOut& o) {
    const double scale = (double)(1 << kSF);
    half8 fh, fl, gh, gl;
#pragma unroll
    for (int q = 0; q < 8; ++q) {
        const bool valid = live && kg * 8 + q < n;
        const double x = valid ? xv[q] : 0.0;
        const double y = valid ? 1.0 - xv[q] : 0.0;
        _Float16 hi, lo;
        split_fp16(x, scale, hi, lo);
        fh[q] = hi;
        fl[q] = lo;
        split_fp16(y, scale, hi, lo);
        gh[q] = hi;
        gl[q] = lo;
    }
    // four neighbouring threads (the k-groups of one iteration) write 64
    // contiguous bytes of each operand
    const int64_t u = ((kg / kKG) * nt + b) * kKG + kg % kKG;
    const u32x4 fh4 = __builtin_bit_cast(u32x4, fh), fl4 = __builtin_bit_cast(u32x4, fl);
    const u32x4 gh4 = __builtin_bit_cast(u32x4, gh), gl4 = __builtin_bit_cast(u32x4, gl);
    o.Fhi[u] = fh4;
    o.Flo[u] = fl4;
    o.Ghi[u] = gh4;
    o.Glo[u] = gl4;
    // iterations holding a nonzero operand (every writer stores the same 1)
    if (o.nzF) {
        const u32x4 fo = fh4 | fl4, go = gh4 | gl4;
        if ((fo.x | fo.y | fo.z | fo.w) != 0u) o.nzF[kg / kKG] = 1;
        if ((go.x | go.y | go.z | go.w) != 0u) o.nzG[kg / kKG] = 1;
    }
}

// Node indices of this thread's k-groups and of its panel for the packing
// from an LDS row (a row that fits the LDS has kgroups <= kPackKG * kMemT and
// npanel <= kMemT, checked when the rows are sized); loaded ahead of the row
// so their latency overlaps whatever the row needs.
struct PackNodes {
    int32_t node[kPackKG][8];
    int32_t pnode;
};

__device__ __forceinline__ void load_pack_nodes(PackNodes& pk, int64_t n, int64_t kgroups, int64_t kt, int npanel,
                                                const int32_t* __restrict__ perm) {
#pragma unroll
    for (int j = 0; j < kPackKG; ++j) {
        const int64_t kg = min((int64_t)threadIdx.x + (int64_t)j * kMemT, kgroups - 1);
#pragma unroll
        for (int q = 0; q < 8; ++q) pk.node[j][q] = perm[min(kg * 8 + q, n - 1)];
    }
    const int pn = threadIdx.x;
    pk.pnode = panel_node(pn < npanel ? pn : 0, kt, n, perm);
}

// member b's operands and operand choices from its f in LDS `row` (caller
// order, complete: the caller has passed a barrier since writing it)
__device__ __forceinline__ void pack_from_row(const double* row, const PackNodes& pk, bool live, int col, int64_t n,
                                              int64_t kgroups, int npanel, int nt, const PackOut& o) {
    const int pn = threadIdx.x;
    if (pn < npanel) {
        const uint8_t v = live && row[pk.pnode] > 0.5 ? 1 : 0;
        o.useg[(int64_t)pn * nt + col] = v;
        if (o.useF && live) (v ? o.useG : o.useF)[pn] = 1;
    }
#pragma unroll
    for (int j = 0; j < kPackKG; ++j) {
        const int64_t kg = (int64_t)threadIdx.x + (int64_t)j * kMemT;
        if (kg < kgroups) {
            double xv[8];
#pragma unroll
            for (int q = 0; q < 8; ++q) xv[q] = row[pk.node[j][q]];
            pack_group(kg, xv, live, n, col, nt, o);
        }
    }
}

// f (caller order) -> packed fp16 hi/lo of both GEMM operands z = f and
// z = g = 1 - f in energy order (the GEMM streams one of them per (panel,
// member)), and the operand choice of every panel.  Members >= B and nodes
// >= n pack as zero; with a gate (the batched DP5's per-member step sizes)
// members whose step is 0 are skipped.
template <bool ROW>
__global__ void __launch_bounds__(kMemT)
pack_member_kernel(const double* __restrict__ f, int64_t n, int B, const int32_t* __restrict__ perm,
                   int64_t kgroups, int64_t kt, int npanel, int nt, PackOut o, const double* __restrict__ gate) {
    extern __shared__ double row[];
    const int b = member_of_block(blockIdx.x, nt);
    if (b >= nt) return;
    // a member whose step is 0 this attempt (finished) keeps its old columns:
    // its products are not read
    if (gate && b < B && gate[b] == 0.0) return;
    const bool live = b < B;
    const double* fb = f + (int64_t)(live ? b : 0) * n;
    if constexpr (ROW) {
        PackNodes pk;
        load_pack_nodes(pk, n, kgroups, kt, npanel, perm);
        if (live) {
            stage_member_row(fb, n, row);
            lds_barrier();
        }
        pack_from_row(row, pk, live, b, n, kgroups, npanel, nt, o);
        if (o.xb && live) hat_from_row(row, n, o, b);
    } else {
        if (o.xb && live) {
            double* __restrict__ xb = o.xb + (int64_t)b * n;
            for (int64_t i = threadIdx.x; i < n; i += kMemT)
                xb[i] = hat_value_dim(o.dim, o.hs, i, [&](int64_t j) { return fb[j]; });
        }
        for (int pn = threadIdx.x; pn < npanel; pn += kMemT) {
            const uint8_t v = live && fb[panel_node(pn, kt, n, perm)] > 0.5 ? 1 : 0;
            o.useg[(int64_t)pn * nt + b] = v;
            if (o.useF && live) (v ? o.useG : o.useF)[pn] = 1;
        }
        for (int64_t kg = threadIdx.x; kg < kgroups; kg += kMemT) {
            double xv[8];
#pragma unroll
            for (int q = 0; q < 8; ++q) xv[q] = fb[perm[min(kg * 8 + q, n - 1)]];
            pack_group(kg, xv, live, n, b, nt, o);
        }
    }
}

// --------------------------------------------------------------------- GEMM
// Workgroup: 128 rows x NW members (one member tile of the padded batch) and
// WAVES waves in a 2 x WAVES/2 grid; a wave owns 64 rows x 2*NW/WAVES
// members = 2 x TN tiles of 32 x 32.  One iteration stages A (4 k-groups x
// 128 rows, hi and lo) and Z (4 k-groups x NW, hi and lo) in LDS,
// prefetching the next iteration's tile into registers while the MFMAs run.
//
// Accuracy: with FOLD > 0 the FP32 accumulators are added into FP64
// registers every FOLD iterations (a chain of at most 32*FOLD products per
// FP32 sum), so a workgroup can run a whole panel's band with no K-split and
// the partial sums leave the kernel in FP64.  FOLD == 0 keeps one FP32 chain
// per K-split and writes FP32 partials; it needs many splits for the same
// accuracy.
//
// Block order: the dispatcher hands workgroup w to XCD w % 8.  The logical
// block L = (w % 8) * per_xcd + w / 8 makes consecutive L run on the same
// XCD, and L enumerates (panel, split, member tile) with the member tile
// fastest, so the member tiles of one panel read the same A tile through
// one L2 and neighbouring panels share most of their Z window.
template <int FOLD>
struct Partial { typedef double type; };
template <>
struct Partial<0> { typedef float type; };

template <int NTH, int N>
__device__ __forceinline__ void stage_a(u32x4 (&ra)[N], const u32x4* __restrict__ Ab, int64_t itl, int tid) {
#pragma unroll
    for (int j = 0; j < N; ++j) {
        const int q = tid + j * NTH;
        const int which = q / (kKG * kMT);
        const int kg = (q / kMT) % kKG;
        const int r = q % kMT;
        ra[j] = Ab[((itl * kKG + kg) * 2 + which) * kMT + r];
    }
}

// Unit q of a Z tile (NW members, the four k-groups of one iteration, hi and
// lo): consecutive q are consecutive 16-byte units in memory (member-major,
// four k-groups per member); in LDS the tile is k-group-major with rows of
// NW + 1 units, so the MFMA fragments are contiguous reads and the
// transposing stores spread over the banks.
template <int NW>
__device__ __forceinline__ void z_unit(int q, int& which, int& m, int& kgl) {
    which = q / (kKG * NW);
    const int rem = q % (kKG * NW);
    m = rem / kKG;
    kgl = rem % kKG;
}

template <int NTH, int NW, int N>
__device__ __forceinline__ void stage_z(u32x4 (&rf)[N], const u32x4* const (&zhi)[N],
                                        const u32x4* const (&zlo)[N], int nt, int col0, int64_t it,
                                        int tid) {
#pragma unroll
    for (int j = 0; j < N; ++j) {
        int which, m, kgl;
        z_unit<NW>(tid + j * NTH, which, m, kgl);
        rf[j] = (which ? zlo[j] : zhi[j])[(it * nt + col0 + m) * kKG + kgl];
    }
}

template <int NW, int WAVES, int FOLD, int DEPTH>
__global__ void __launch_bounds__(64 * WAVES)
mixed_gemm_kernel(const u32x4* __restrict__ Aband, const int64_t* __restrict__ band_off,
                  const u32x4* __restrict__ Fhi, const u32x4* __restrict__ Flo,
                  const u32x4* __restrict__ Ghi, const u32x4* __restrict__ Glo,
                  const uint8_t* __restrict__ useg, int64_t rows_pad, int nt,
                  const int2* __restrict__ krange, int nmt, int nsplit, int64_t nblocks,
                  int64_t per_xcd, const int32_t* __restrict__ ncol,
                  typename Partial<FOLD>::type* __restrict__ P) {
    constexpr int NTH = 64 * WAVES;
    constexpr int WC = WAVES / 2;                // waves along the members
    constexpr int WN = NW / WC;                  // members per wave
    constexpr int TN = WN / 32;                  // 32-wide member tiles per wave
    static_assert(TN >= 1 && WN % 32 == 0, "member tile too narrow for the wave grid");
    constexpr int A_PER_T = 2 * kKG * kMT / NTH;
    constexpr int F_PER_T = 2 * kKG * NW / NTH;
    __shared__ u32x4 sA[2][kKG][kMT];
    __shared__ u32x4 sF[2][kKG][NW + 1];

    const int64_t wg = blockIdx.x;
    const int64_t L = (wg % 8) * per_xcd + wg / 8;
    if (L >= nblocks) return;
    const int mt = (int)(L % nmt);
    if (ncol && mt * NW >= *ncol) return;   // a member tile with no running member
    const int64_t ps = L / nmt;
    const int split = (int)(ps % nsplit);
    const int panel = (int)(ps / nsplit);

    const int tid = threadIdx.x;
    const int lane = tid & 63, wave = tid >> 6;
    const int wr = wave / WC, wc = wave % WC;
    const int64_t row0 = (int64_t)panel * kMT;
    const int col0 = mt * NW;

    const int2 kr = krange[panel];
    const u32x4* Ab = Aband + band_off[panel];
    const int nit = kr.y - kr.x;
    const int it0 = kr.x + (int)((int64_t)nit * split / nsplit);
    const int it1 = kr.x + (int)((int64_t)nit * (split + 1) / nsplit);

    floatx16 acc[2][TN];
#pragma unroll
    for (int t = 0; t < 2; ++t)
#pragma unroll
        for (int u = 0; u < TN; ++u) acc[t][u] = floatx16{0};
    double sum[2][TN][16];
    if constexpr (FOLD > 0) {
#pragma unroll
        for (int t = 0; t < 2; ++t)
#pragma unroll
            for (int u = 0; u < TN; ++u)
#pragma unroll
                for (int e = 0; e < 16; ++e) sum[t][u][e] = 0.0;
    }

    // operand of the member of every Z unit this thread stages
    const u32x4* Zhi[F_PER_T];
    const u32x4* Zlo[F_PER_T];
#pragma unroll
    for (int j = 0; j < F_PER_T; ++j) {
        int which, m, kgl;
        z_unit<NW>(tid + j * NTH, which, m, kgl);
        const bool g = useg[(int64_t)panel * nt + col0 + m] != 0;
        Zhi[j] = g ? Ghi : Fhi;
        Zlo[j] = g ? Glo : Flo;
    }
    // DEPTH register buffers: iteration it is staged from buffer (it - it0) %
    // DEPTH, whose loads were issued DEPTH iterations earlier.  Every
    // iteration of the padded count issues its loads and LDS stores
    // unconditionally (beyond it1 the index is clamped and the tile reloaded),
    // so the wait before staging a buffer counts exactly the DEPTH-1 newer
    // buffers still in flight; only the MFMAs are skipped past it1.
    u32x4 ra[DEPTH][A_PER_T], rf[DEPTH][F_PER_T];
    if (it0 < it1) {
        const int last = it1 - 1;
        const int nit_pad = (it1 - it0 + DEPTH - 1) / DEPTH * DEPTH;
#pragma unroll
        for (int d = 0; d < DEPTH; ++d) {
            const int il = min(it0 + d, last);
            stage_a<NTH>(ra[d], Ab, il - kr.x, tid);
            stage_z<NTH, NW>(rf[d], Zhi, Zlo, nt, col0, il, tid);
        }
        int since = 0;
        for (int itb = it0; itb < it0 + nit_pad; itb += DEPTH) {
#pragma unroll
            for (int d = 0; d < DEPTH; ++d) {
                const int it = itb + d;
#pragma unroll
                for (int j = 0; j < A_PER_T; ++j) {
                    const int q = tid + j * NTH;
                    sA[q / (kKG * kMT)][(q / kMT) % kKG][q % kMT] = ra[d][j];
                }
#pragma unroll
                for (int j = 0; j < F_PER_T; ++j) {
                    const int q = tid + j * NTH;
                    int which, m, kgl;
                    z_unit<NW>(q, which, m, kgl);
                    sF[which][kgl][m] = rf[d][j];
                }
                __syncthreads();
                {
                    const int il = min(it + DEPTH, last);
                    stage_a<NTH>(ra[d], Ab, il - kr.x, tid);
                    stage_z<NTH, NW>(rf[d], Zhi, Zlo, nt, col0, il, tid);
                }
                if (it < it1) {
#pragma unroll
                    for (int ks = 0; ks < 2; ++ks) {
                        const int kg = 2 * ks + (lane >> 5);
                        half8 ah[2], al[2], fh[TN], fl[TN];
#pragma unroll
                        for (int t = 0; t < 2; ++t) {
                            const int r = wr * 64 + t * 32 + (lane & 31);
                            ah[t] = __builtin_bit_cast(half8, sA[0][kg][r]);
                            al[t] = __builtin_bit_cast(half8, sA[1][kg][r]);
                        }
#pragma unroll
                        for (int u = 0; u < TN; ++u) {
                            const int col = wc * WN + u * 32 + (lane & 31);
                            fh[u] = __builtin_bit_cast(half8, sF[0][kg][col]);
                            fl[u] = __builtin_bit_cast(half8, sF[1][kg][col]);
                        }
#pragma unroll
                        for (int t = 0; t < 2; ++t)
#pragma unroll
                            for (int u = 0; u < TN; ++u) {
                                acc[t][u] = __builtin_amdgcn_mfma_f32_32x32x16_f16(ah[t], fh[u], acc[t][u], 0, 0, 0);
                                acc[t][u] = __builtin_amdgcn_mfma_f32_32x32x16_f16(ah[t], fl[u], acc[t][u], 0, 0, 0);
                                acc[t][u] = __builtin_amdgcn_mfma_f32_32x32x16_f16(al[t], fh[u], acc[t][u], 0, 0, 0);
                            }
                    }
                }
                __syncthreads();
                if constexpr (FOLD > 0) {
                    if (it < it1 && ++since == FOLD) {
                        since = 0;
#pragma unroll
                        for (int t = 0; t < 2; ++t)
#pragma unroll
                            for (int u = 0; u < TN; ++u) {
#pragma unroll
                                for (int e = 0; e < 16; ++e) sum[t][u][e] += (double)acc[t][u][e];
                                acc[t][u] = floatx16{0};
                            }
                    }
                }
            }
        }
    }
    // C tile -> P[split][member][row]; lane holds member (lane & 31) of each
    // 32-wide tile and rows 8*(e/4) + 4*(lane/32) + (e%4) of each 32-row tile
#pragma unroll
    for (int t = 0; t < 2; ++t)
#pragma unroll
        for (int u = 0; u < TN; ++u) {
            const int col = col0 + wc * WN + u * 32 + (lane & 31);
            typename Partial<FOLD>::type* dst =
                P + ((int64_t)split * nt + col) * rows_pad + row0 + wr * 64 + t * 32 + 4 * (lane >> 5);
#pragma unroll
            for (int e = 0; e < 16; ++e) {
                if constexpr (FOLD > 0)
                    dst[8 * (e / 4) + (e % 4)] = sum[t][u][e] + (double)acc[t][u][e];
                else
                    dst[8 * (e / 4) + (e % 4)] = acc[t][u][e];
            }
        }
}

// Persistent variant: one workgroup of 4 MG waves per CU (MG waves per
// SIMD).  Wave w owns 32 rows of the panel (row group w % 4) and NW / MG
// members of the tile (member group w / 4), i.e. TN = NW / (32 MG) MFMA
// tiles of 32 x 32.  A fragments go from global memory straight into the
// wave's registers (no LDS), DA iterations ahead; the Z tile is the operand
// all row groups share and goes through double-buffered LDS with one barrier
// per iteration, also DA iterations ahead.  With MG = 2 a SIMD runs two
// waves, so one wave's FP64 folds and memory waits overlap the other's
// MFMAs.
//
// The work units are chunks of the panels' bands: a band is one chunk, or
// (optionally) two cut at a fixed iteration set once per context, each
// summed on its own and written to partial-sum slot 0 or 1; the epilogue
// adds the slots in slot order.  A unit computes the iterations urange gives
// it: its whole chunk, or (schedules that leave out zero operands) the part
// from the first to the last iteration holding a nonzero operand of a kind
// its panel uses, the start moved back to the chunk's fold grid.  The
// iterations left out add products of zero operands, exactly 0 to the FP32
// sums, and the FP64 folds fall on the same iterations, so the result is
// bit for bit that of the whole chunk.
// The host balances the units over the workgroups (longest first onto the
// least loaded bin, a fixed order, so every call runs the same schedule).
// A bin is nmt workgroups on one XCD that run the same unit list for the
// nmt member tiles side by side, so a panel's A band is fetched into that
// XCD's L2 once for all member tiles.  Bin b runs on XCD b % 8: consecutive
// bins of a schedule (the longest units first) spread over the eight L2s,
// also when only a few bins have work.  With whole panels and the same FOLD
// the summation order is that of mixed_gemm_kernel with one split.
// The member tiles a persistent launch computes: all nmt, or with ncol (the
// columns running members fill, from the fused DP5's compaction) the tiles
// those columns reach.  Each XCD's workgroups then regroup into bins of that
// many, so fewer running tiles give more bins and shorter panel lists.
__device__ inline int active_tiles(const int32_t* ncol, int nw, int nmt) {
    if (!ncol) return nmt;
    const int t = (*ncol + nw - 1) / nw;
    return t < nmt ? t : nmt;
}

template <int NW, int FOLD, int DA, int MG, bool AL>
__global__ void __launch_bounds__(256 * MG, 1)
mixed_gemm_persist_kernel(const u32x4* __restrict__ Aband, const int64_t* __restrict__ band_off,
                          const u32x4* __restrict__ Fhi, const u32x4* __restrict__ Flo,
                          const u32x4* __restrict__ Ghi, const u32x4* __restrict__ Glo,
                          const uint8_t* __restrict__ useg, int64_t rows_pad, int nt,
                          const int2* __restrict__ krange, const int32_t* __restrict__ sched_off,
                          const int32_t* __restrict__ sched_panel, int nmt, int bins_per_xcd,
                          const int32_t* __restrict__ ncol, const int2* __restrict__ urange,
                          int64_t slot_stride, double* __restrict__ P) {
    static_assert(FOLD > 0, "the persistent kernel folds into FP64");
    constexpr int NTH = 256 * MG;
    constexpr int WN = NW / MG;                    // members per wave
    constexpr int TN = WN / 32;
    constexpr int Z_PER_T = 2 * kKG * NW / NTH;
    static_assert(TN >= 1 && WN % 32 == 0, "member group narrower than one MFMA tile");
    static_assert(Z_PER_T >= 1 && Z_PER_T * NTH == 2 * kKG * NW, "Z tile must split evenly over the threads");
    // unroll U = lcm(2, DA, FOLD): register slot, LDS buffer and fold fixed per step
    constexpr int U = std::lcm(std::lcm(2, DA), FOLD);
    static_assert(U % DA == 0 && U % 2 == 0 && U % FOLD == 0, "unroll must hold whole prefetch, buffer and fold cycles");
    constexpr int A_PER_T = 2 * kKG * kMT / NTH;
    __shared__ u32x4 sZ[2][2][kKG][NW + 1];
    __shared__ u32x4 sA[AL ? 2 : 1][kKG][2][AL ? kMT : 1];

    // workgroup w runs on XCD w % 8; the T workgroups of a bin (one per
    // member tile) share an XCD, so a panel's A band is read into one L2
    const int T = active_tiles(ncol, NW, nmt);
    if (T == 0) return;
    const int w = blockIdx.x;
    const int wi = w / 8;
    const int mt = wi % T;
    const int bin = (wi / T) * 8 + w % 8;
    if (wi / T >= bins_per_xcd * nmt / T) return;
    const int tid = threadIdx.x;
    const int lane = tid & 63, wave = tid >> 6;
    const int rg = wave % 4, mg = wave / 4;
    const int col0 = mt * NW;

    for (int s = sched_off[bin]; s < sched_off[bin + 1]; ++s) {
        // work unit: chunk ch of the panel, urange[unit].y iterations from
        // iteration urange[unit].x of its band, summed into slot ch
        const int unit = sched_panel[s];
        const int panel = unit >> 1, ch = unit & 1;
        const int64_t row0 = (int64_t)panel * kMT;
        const int arow = rg * 32 + (lane & 31);
        const int2 kr = krange[panel];
        const int2 ur = urange[unit];
        const int k0 = ur.x, N = ur.y;
        const u32x4* Ab = Aband + band_off[panel] + (int64_t)k0 * (2 * kKG * kMT);
        // operand of the member of every Z unit this thread stages
        const u32x4* Zhi[Z_PER_T];
        const u32x4* Zlo[Z_PER_T];
#pragma unroll
        for (int j = 0; j < Z_PER_T; ++j) {
            int which, m, kgl;
            z_unit<NW>(tid + j * NTH, which, m, kgl);
            const bool g = useg[(int64_t)panel * nt + col0 + m] != 0;
            Zhi[j] = g ? Ghi : Fhi;
            Zlo[j] = g ? Glo : Flo;
        }

        floatx16 acc[TN];
        double sum[TN][16];
#pragma unroll
        for (int u = 0; u < TN; ++u) {
            acc[u] = floatx16{0};
#pragma unroll
            for (int e = 0; e < 16; ++e) sum[u][e] = 0.0;
        }
        if (N > 0) {
            u32x4 ra[AL ? 1 : DA][4], rz[DA][Z_PER_T];
            // AL: this thread's units of the panel's 16 KB A block, staged in LDS
            u32x4 rst[AL ? DA : 1][A_PER_T];
            auto load_ast = [&](u32x4 (&r)[A_PER_T], int k) {
                const u32x4* blk = Ab + (int64_t)min(k, N - 1) * (2 * kKG * kMT);
#pragma unroll
                for (int j = 0; j < A_PER_T; ++j) r[j] = blk[tid + j * NTH];
            };
            auto store_ast = [&](const u32x4 (&r)[A_PER_T], int buf) {
#pragma unroll
                for (int j = 0; j < A_PER_T; ++j) (&sA[buf][0][0][0])[tid + j * NTH] = r[j];
            };
            // A(k): lane's 16-byte units of k-steps 0 and 1 (hi, lo); Z(k): this
            // thread's units of the shared tile; iteration indices clamp at the end
            auto load_a = [&](u32x4 (&r)[4], int k) {
                const int64_t kg = (int64_t)min(k, N - 1) * kKG + (lane >> 5);
#pragma unroll
                for (int ks = 0; ks < 2; ++ks) {
                    r[2 * ks] = Ab[((kg + 2 * ks) * 2) * kMT + arow];
                    r[2 * ks + 1] = Ab[((kg + 2 * ks) * 2 + 1) * kMT + arow];
                }
            };
            auto load_z = [&](u32x4 (&r)[Z_PER_T], int k) {
                stage_z<NTH, NW>(r, Zhi, Zlo, nt, col0, (int64_t)(kr.x + k0 + min(k, N - 1)), tid);
            };
            auto store_z = [&](const u32x4 (&r)[Z_PER_T], int buf) {
#pragma unroll
                for (int j = 0; j < Z_PER_T; ++j) {
                    int which, m, kgl;
                    z_unit<NW>(tid + j * NTH, which, m, kgl);
                    sZ[buf][which][kgl][m] = r[j];
                }
            };
            // one iteration: MFMAs on A(k) (registers, slot k % DA) and Z(k) (LDS
            // buffer k & 1); then A(k+DA) into the freed slot, Z(k+1) (loaded DA
            // iterations ago) into the other LDS buffer and Z(k+1+DA) into its
            // slot; the FP64 fold after every FOLD iterations.  uu = k mod U is a
            // compile-time constant at every call, so slots, buffers and the fold
            // position are fixed in the code.
            auto step = [&](int k, int uu) {
                const int da = uu % DA, buf = uu & 1;
#pragma unroll
                for (int ks = 0; ks < 2; ++ks) {
                    const int kg = 2 * ks + (lane >> 5);
                    const half8 ah = __builtin_bit_cast(half8, AL ? sA[buf][kg][0][arow] : ra[AL ? 0 : da][2 * ks]);
                    const half8 al = __builtin_bit_cast(half8, AL ? sA[buf][kg][1][arow] : ra[AL ? 0 : da][2 * ks + 1]);
#pragma unroll
                    for (int u = 0; u < TN; ++u) {
                        const int col = mg * WN + u * 32 + (lane & 31);
                        const half8 zh = __builtin_bit_cast(half8, sZ[buf][0][kg][col]);
                        const half8 zl = __builtin_bit_cast(half8, sZ[buf][1][kg][col]);
                        acc[u] = __builtin_amdgcn_mfma_f32_32x32x16_f16(ah, zh, acc[u], 0, 0, 0);
                        acc[u] = __builtin_amdgcn_mfma_f32_32x32x16_f16(ah, zl, acc[u], 0, 0, 0);
                        acc[u] = __builtin_amdgcn_mfma_f32_32x32x16_f16(al, zh, acc[u], 0, 0, 0);
                    }
                }
                const int dz = (uu + 1) % DA;
                if constexpr (AL) {
                    store_ast(rst[dz], (uu + 1) & 1);
                    load_ast(rst[dz], k + 1 + DA);
                } else {
                    load_a(ra[da], k + DA);
                }
                store_z(rz[dz], (uu + 1) & 1);
                load_z(rz[dz], k + 1 + DA);
                if ((uu + 1) % FOLD == 0) {
#pragma unroll
                    for (int u = 0; u < TN; ++u) {
#pragma unroll
                        for (int e = 0; e < 16; ++e) sum[u][e] += (double)acc[u][e];
                        acc[u] = floatx16{0};
                    }
                }
                __syncthreads();
            };
            // prologue: Z(0) to LDS buffer 0, then the loads of "iterations" -DA..-1
            // in the order the loop issues them (A(j+DA), Z(j+1+DA)), so the wait
            // counts at the loop head are the same from the prologue and from the
            // back edge
            if constexpr (AL) {
                load_ast(rst[0], 0);
                store_ast(rst[0], 0);
            }
            load_z(rz[0], 0);
            store_z(rz[0], 0);
#pragma unroll
            for (int d = 0; d < DA; ++d) {
                if constexpr (AL) load_ast(rst[(d + 1) % DA], d + 1);
                else load_a(ra[d], d);
                load_z(rz[(d + 1) % DA], d + 1);
            }
            __syncthreads();
            const int nfull = N / U * U;
            for (int k0 = 0; k0 < nfull; k0 += U) {
#pragma unroll
                for (int uu = 0; uu < U; ++uu) step(k0 + uu, uu);
            }
#pragma unroll
            for (int uu = 0; uu + 1 < U; ++uu)
                if (nfull + uu < N) step(nfull + uu, uu);
#pragma unroll
            for (int u = 0; u < TN; ++u) {
#pragma unroll
                for (int e = 0; e < 16; ++e) sum[u][e] += (double)acc[u][e];
            }
        }
        // C tile -> P[member][row]; lane holds member (lane & 31) of each
        // 32-wide tile and rows 8*(e/4) + 4*(lane/32) + (e%4) of the wave's 32
#pragma unroll
        for (int u = 0; u < TN; ++u) {
            const int col = col0 + mg * WN + u * 32 + (lane & 31);
            double* dst = P + ch * slot_stride + (int64_t)col * rows_pad + row0 + rg * 32 + 4 * (lane >> 5);
#pragma unroll
            for (int e = 0; e < 16; ++e) dst[8 * (e / 4) + (e % 4)] = sum[u][e];
        }
    }
}

// ------------------------------------------------------------------ epilogue
// One workgroup per member: FP64 sum of the split partials, Pauli-blocked
// assembly with the member's bath occupation, FP64 block sums of sum V I and
// sum V |f(1-f)| for the number-defect removal (and of sum V eps I and
// sum V eps |f(1-f)| for the member's energy flux, flux[b] = the e-ph
// electron energy flux of the corrected I), result in the caller's node
// order.  With ROW the member's f is staged in LDS and every position
// replaces the f of its node by I in place (the energy order is a
// permutation, so each LDS slot is read and written by one thread only); the
// row then goes out by coalesced stores together with the correction.
// Without it the gather reads and the scatter writes global memory and the
// correction is a second pass.  Gated members (step 0) are skipped.
// The energy-order part of the epilogue for member b: the products of the
// computed panels (P), the member's f (staged in LDS `row` with ROW, else
// read from fb) and the per-position constants give I at every position;
// I replaces f in `row` (ROW) or goes to Ib.  Returns the FP64 partial sums
// of V I and V |f(1-f)| (and, with want_e, of eps V I and eps V |f(1-f)|).
template <bool ROW, typename PT>
__device__ __forceinline__ void finish_energy(const PT* __restrict__ P, int nslot, int nt, int64_t rows_pad,
                                              int64_t n, int64_t kt, int b, const double* __restrict__ fb,
                                              const int32_t* __restrict__ perm, const double* __restrict__ rk,
                                              const double* __restrict__ ck, const double* __restrict__ V,
                                              double nb, const uint8_t* __restrict__ useg,
                                              const uint8_t* __restrict__ pmask, int npanel, uint8_t* sflag,
                                              double inv_scale, const double* __restrict__ eps_s, bool want_e,
                                              bool hatw, double* row, double* __restrict__ Ib, double& dI,
                                              double& dw, double& eI, double& ew) {
    // no contraction into FMAs: every kernel that inlines this (LDS-row and
    // global-memory epilogues, the fused DP5 passes) rounds the same way
#pragma clang fp contract(off)
    const double np1 = nb + 1.0;
    // UNR positions per thread per round with all their loads issued before
    // use; the trip count is the same for every thread (the barrier after the
    // first round's loads needs all of them)
#ifndef TG_FIN_UNR
#define TG_FIN_UNR 4
#endif
    constexpr int UNR = TG_FIN_UNR;
    double u[UNR], v[UNR], u1[UNR], v1[UNR], rkp[UNR], ckp[UNR], Vp[UNR], Ep[UNR];
    int32_t node[UNR], gu[UNR], gv[UNR];   // gu, gv: the panels' operand bytes
    // the partial sums (streamed from HBM), read only from the slots the GEMM
    // wrote: bit 1 (2) of the panel byte for slot 0 (1); a panel left out
    // has neither, its product is 0 and its bytes need not move.  Slot 1
    // goes to registers of its own, so every load of a round is issued before
    // the first sum.  Slots 2 and up (tiled kernel with more K-splits) add
    // into slot 1's value ...
    const int64_t slot = (int64_t)nt * rows_pad;
    auto load_products = [&](int64_t p0) {
#pragma unroll
        for (int j = 0; j < UNR; ++j) {
            const int64_t p = min(p0 + (int64_t)j * kMemT, n - 1);
            const PT* P0 = P + (int64_t)b * rows_pad;
            u[j] = 0.0;
            v[j] = 0.0;
            u1[j] = 0.0;
            v1[j] = 0.0;
            if (gu[j] & 2) u[j] = (double)P0[p];
            if (gu[j] & 4) u1[j] = (double)P0[slot + p];
            if (gv[j] & 2) v[j] = (double)P0[kt + p];
            if (gv[j] & 4) v1[j] = (double)P0[slot + kt + p];
            for (int s = 2; s < nslot; ++s) {
                if (gu[j] & 2) u1[j] += (double)P0[s * slot + p];
                if (gv[j] & 2) v1[j] += (double)P0[s * slot + kt + p];
            }
        }
    };
    // ... after the operand bytes and the panel slot masks (pmask: bit s for
    // slot s; null when every panel has all nslot slots), which say which
    // panels those are.  With ROW one byte per panel (the member's operand
    // choice in bit 0, the slots in bits 1 and 2) is staged in LDS (sflag)
    // once, so a round waits for one memory trip, not two ...
    auto panel_byte = [&](int64_t pn) {
        int g = useg[pn * nt + b];
        g |= pmask ? (pmask[pn] & 3) << 1 : (nslot > 1 ? 6 : 2);
        return g;
    };
    auto load_bytes = [&](int64_t p0) {
#pragma unroll
        for (int j = 0; j < UNR; ++j) {
            const int64_t p = min(p0 + (int64_t)j * kMemT, n - 1);
            const int64_t pu = p / kMT, pv = (kt + p) / kMT;
            if constexpr (ROW) {
                gu[j] = sflag[pu];
                gv[j] = sflag[pv];
            } else {
                gu[j] = panel_byte(pu);
                gv[j] = panel_byte(pv);
            }
        }
    };
    // ... and the per-position constants every member shares (L2-resident)
    auto load_constants = [&](int64_t p0) {
#pragma unroll
        for (int j = 0; j < UNR; ++j) {
            const int64_t p = min(p0 + (int64_t)j * kMemT, n - 1);
            rkp[j] = rk[p];
            ckp[j] = ck[p];
            Vp[j] = V[p];
            node[j] = perm[p];
            if (want_e) Ep[j] = eps_s[p];
        }
    };
    auto finish_round = [&](int64_t p0) {
        // every load of the round is used here, so the scheduler keeps them
        // issued back to back instead of sinking each to its first use (one
        // memory round trip per round, not one per position)
#pragma unroll
        for (int j = 0; j < UNR; ++j)
            asm volatile("" ::"v"(u[j]), "v"(v[j]), "v"(u1[j]), "v"(v1[j]), "v"(rkp[j]), "v"(ckp[j]), "v"(Vp[j]),
                         "v"(node[j]), "v"(gu[j]), "v"(gv[j]));
#pragma unroll
        for (int j = 0; j < UNR; ++j) {
            if (p0 + (int64_t)j * kMemT >= n) break;
            // bit 0 = g (z = 1 - f) rather than f; the slots in slot order
            const double uu = (u[j] + u1[j]) * inv_scale;
            const double vv = (v[j] + v1[j]) * inv_scale;
            // u = K z, v = K^T z' with z, z' = f or g = 1 - f per panel
            const bool gfu = gu[j] & 1, gfv = gv[j] & 1;
            const double Kf = gfu ? rkp[j] - uu : uu, Kg = gfu ? uu : rkp[j] - uu;
            const double KTf = gfv ? ckp[j] - vv : vv, KTg = gfv ? vv : ckp[j] - vv;
            const double x = ROW ? row[node[j]] : fb[node[j]];
            // (1-f)(W f) - f (W^T g),  W = (n+1) K + n K^T
            const double Ii = ((1.0 - x) * (np1 * Kf + nb * KTf) - x * (np1 * KTg + nb * Kg)) / Vp[j];
            // number-defect weights: |x(1-x)|, or the hat form's
            const double wj = hatw ? hat_defect_weight(x, np1, nb, rkp[j], ckp[j], Vp[j]) : fabs(x * (1.0 - x));
            const double vi = Vp[j] * Ii, vw = Vp[j] * wj;
            dI += vi;
            dw += vw;
            if (want_e) {
                eI += Ep[j] * vi;
                ew += Ep[j] * vw;
            }
            if constexpr (ROW)
                row[node[j]] = Ii;
            else
                Ib[node[j]] = Ii;
        }
    };
    const int64_t step = (int64_t)kMemT * UNR;
    if constexpr (ROW) {
        // the panel bytes, the first round's constants and the row staging
        // are in flight together; the partial sums follow the barrier
        int g = 0;
        if (threadIdx.x < npanel) g = panel_byte(threadIdx.x);
        load_constants(threadIdx.x);
        stage_member_row(fb, n, row);
        if (threadIdx.x < npanel) sflag[threadIdx.x] = (uint8_t)g;
        lds_barrier();
        load_bytes(threadIdx.x);
        load_products(threadIdx.x);
    } else {
        load_bytes(threadIdx.x);
        load_products(threadIdx.x);
        load_constants(threadIdx.x);
    }
    finish_round(threadIdx.x);
    for (int64_t base = step; base < n; base += step) {
        load_bytes(base + threadIdx.x);
        load_products(base + threadIdx.x);
        load_constants(base + threadIdx.x);
        finish_round(base + threadIdx.x);
    }
}

// Workgroup sums of the four epilogue partials (every thread calls it; the
// result is valid in thread 0).  red: 4 x (kMemT / 64) LDS doubles.
__device__ inline void block_sum4(double& a, double& b, double& c, double& d, double (*red)[kMemT / 64]) {
    for (int o = 32; o > 0; o >>= 1) {
        a += __shfl_down(a, o, 64);
        b += __shfl_down(b, o, 64);
        c += __shfl_down(c, o, 64);
        d += __shfl_down(d, o, 64);
    }
    if ((threadIdx.x & 63) == 0) {
        red[0][threadIdx.x >> 6] = a;
        red[1][threadIdx.x >> 6] = b;
        red[2][threadIdx.x >> 6] = c;
        red[3][threadIdx.x >> 6] = d;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        a = b = c = d = 0.0;
        for (int w = 0; w < kMemT / 64; ++w) {
            a += red[0][w];
            b += red[1][w];
            c += red[2][w];
            d += red[3][w];
        }
    }
}

template <bool ROW, typename PT>
__global__ void __launch_bounds__(kMemT)
finish_member_kernel(const PT* __restrict__ P, int nslot, int nt, int64_t rows_pad, int64_t n,
                     int64_t kt, int B, const double* __restrict__ f,
                     const int32_t* __restrict__ perm, const double* __restrict__ rk,
                     const double* __restrict__ ck, const double* __restrict__ V,
                     const double* __restrict__ bath, const uint8_t* __restrict__ useg,
                     const uint8_t* __restrict__ pmask, int npanel, double inv_scale, int conserve,
                     const double* __restrict__ gate,
                     double* __restrict__ I, const double* __restrict__ eps_s,
                     double* __restrict__ flux, const double* __restrict__ rkc,
                     const double* __restrict__ ckc, const double* __restrict__ Vc) {
#pragma clang fp contract(off)
    extern __shared__ double row[];
    __shared__ double red[4][kMemT / 64];
    __shared__ double s_b;
    __shared__ uint8_t sflag[ROW ? kMemT : 1];
    const int b = member_of_block(blockIdx.x, nt);
    if (b >= B || (gate && gate[b] == 0.0)) return;
    const double nb = bath[b];
    const double* fb = f + (int64_t)b * n;
    double* Ib = I + (int64_t)b * n;
    double dI = 0.0, dw = 0.0, eI = 0.0, ew = 0.0;
    // f holds the hat form's occupation factors when rkc is given (the hat
    // form's defect weights need the caller-order K 1, K^T 1 and V)
    finish_energy<ROW, PT>(P, nslot, nt, rows_pad, n, kt, b, fb, perm, rk, ck, V, nb, useg, pmask, npanel, sflag,
                           inv_scale, eps_s, flux != nullptr, rkc != nullptr, row, Ib, dI, dw, eI, ew);
    double sc = 0.0;
    if (conserve || flux) {
        block_sum4(dI, dw, eI, ew, red);
        if (threadIdx.x == 0) {
            s_b = (conserve && dw > 0.0) ? dI / dw : 0.0;
            if (flux) flux[b] = eI - s_b * ew;
        }
    }
    __syncthreads();
    if (conserve) sc = s_b;
    // the row (LDS, or the scattered I) goes out with 16-byte stores, minus
    // the number correction sc |f (1 - f)|
    if (ROW || conserve) {
        auto out = [&](int64_t i, double y) {
            if (!conserve) return y;
            const double x = fb[i];
            const double w = rkc ? hat_defect_weight(x, nb + 1.0, nb, rkc[i], ckc[i], Vc[i]) : fabs(x * (1.0 - x));
            return y - sc * w;
        };
        constexpr int WUNR = 5;
        const int64_t h = row_head(Ib);
        const int64_t np = (n - h) >> 1;
        double2* __restrict__ dst = reinterpret_cast<double2*>(Ib + h);
        for (int64_t j0 = threadIdx.x; j0 < np; j0 += (int64_t)kMemT * WUNR) {
            double2 y[WUNR];
#pragma unroll
            for (int j = 0; j < WUNR; ++j) {
                const int64_t i = h + 2 * min(j0 + (int64_t)j * kMemT, np - 1);
                y[j].x = out(i, ROW ? row[i] : Ib[i]);
                y[j].y = out(i + 1, ROW ? row[i + 1] : Ib[i + 1]);
            }
#pragma unroll
            for (int j = 0; j < WUNR; ++j)
                if (j0 + (int64_t)j * kMemT < np) dst[j0 + (int64_t)j * kMemT] = y[j];
        }
        if (threadIdx.x == 0 && h) Ib[0] = out(0, ROW ? row[0] : Ib[0]);
        if (threadIdx.x == kMemT - 1 && h + 2 * np < n) Ib[n - 1] = out(n - 1, ROW ? row[n - 1] : Ib[n - 1]);
    }
}

// a[perm[p]] = as[p], b[perm[p]] = bs[p] (energy order to the caller's)
__global__ void scatter_pair_kernel(int64_t n, const int32_t* __restrict__ perm, const double* __restrict__ as,
                                    const double* __restrict__ bs, double* __restrict__ a, double* __restrict__ b) {
    const int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (p >= n) return;
    a[perm[p]] = as[p];
    b[perm[p]] = bs[p];
}

// ------------------------------------------------- fused passes of the DP5
// What a fused pass needs of the packed operator (by value).
template <typename PT>
struct MixedView {
    const PT* P;
    int nslot, nt, B, npanel;
    int64_t rows_pad, n, kt, kgroups;
    const int32_t* perm;
    const double* rk;
    const double* ck;
    const double* V;
    const double* eps_s;
    const uint8_t* useg;
    const uint8_t* pmask;   // per-panel slot masks (null: every panel has nslot slots)
    const int32_t* slot;   // null: member b's column is b
    double inv_scale;
    PackOut o;
    // the hat form's defect weights (with o.xb): K 1, K^T 1 and V in the caller's order
    const double* rkc;
    const double* ckc;
    const double* Vc;
};


// One member per workgroup (member_of_block), member rows in LDS; the passes
// are described with Dp5PassArgs.  The GEMM that follows a kDp5Stage or
// kDp5Mid pass reads the operands it packs; a kDp5Mid or kDp5Last pass reads
// the partial sums of the GEMM before it.  A kDp5Mid pass may write ys over
// fin: each position is read (for the number-defect removal) before the same
// thread writes it, and the LDS copy of fin is complete before any write.
template <int MODE, typename PT>
__global__ void __launch_bounds__(kMemT)
dp5_member_kernel(Dp5PassArgs a, MixedView<PT> v) {
#pragma clang fp contract(off)
    extern __shared__ double row[];
    __shared__ double red[4][kMemT / 64];
    __shared__ double s_b;
    __shared__ uint8_t sflag[kMemT];
    const int b = member_of_block(blockIdx.x, v.nt);
    if (b >= v.nt) return;
    const bool live = b < v.B;
    if (MODE == kDp5Last && !live) return;
    const double hb = live ? a.h[b] : 0.0;
    if (live && hb == 0.0) return;   // finished member: nothing of this attempt is read
    // with slots the running members take the leading columns of the
    // operands and partial sums (padding members have none)
    if (v.slot && !live) return;
    const int col = v.slot ? v.slot[b] : b;
    const int64_t n = v.n;
    const int64_t base = (int64_t)b * n;
    // the bath of the finished k (thread 0 rewrites it for the next stage
    // behind a barrier below)
    const double nbk = live ? a.bath[b] : 0.0;
    if (MODE != kDp5Stage && live) {
        // k_s = I(fin): the epilogue over the energy order, I in place of fin in LDS
        double dI = 0.0, dw = 0.0, eI = 0.0, ew = 0.0;
        // the occupation factors of fin: the hat averages the previous pass
        // wrote, or fin itself
        finish_energy<true, PT>(v.P, v.nslot, v.nt, v.rows_pad, n, v.kt, col, (v.o.xb ? v.o.xb : a.fin) + base,
                                v.perm, v.rk, v.ck,
                                v.V, nbk, v.useg, v.pmask, v.npanel, sflag, v.inv_scale, v.eps_s,
                                a.flux != nullptr, v.o.xb != nullptr, row, nullptr, dI, dw, eI, ew);
        block_sum4(dI, dw, eI, ew, red);
        if (threadIdx.x == 0) {
            s_b = (a.conserve && dw > 0.0) ? dI / dw : 0.0;
            if (a.flux) a.flux[b] = eI - s_b * ew;
        }
        __syncthreads();
    }
    // the next stage's bath (every thread has read this one's above)
    if (MODE != kDp5Last && live && a.bs.tab && threadIdx.x == 0) {
        const double ts = a.md[b].t + a.cs * hb;
        a.bath[b] = sched_value(a.bs.tab + (int64_t)b * a.bs.ns, a.bs.ns, a.bs.t_start, a.bs.dt, ts);
    }
    if (live) {
        const double sc = MODE != kDp5Stage ? s_b : 0.0;
        double c[7], cb[7], cf[6];
#pragma unroll
        for (int j = 0; j < 7; ++j) {
            c[j] = a.ca[j] * hb;
            cb[j] = a.cb[j] * hb;
        }
#pragma unroll
        for (int j = 0; j < 6; ++j) cf[j] = a.cfin[j] * hb;
        const double c7 = hb * a.c7;
        const int ng = a.ng;
        const bool want_y4 = MODE == kDp5Mid && a.y4 != nullptr;
        // fin for the number-defect weights |x(1-x)| (nodal factors; the hat
        // factors come from o.xb) or for y5 in the last pass
        const bool need_fin = MODE == kDp5Last || (MODE == kDp5Mid && a.conserve && !v.o.xb && a.nfin == 0);
        // one position per thread per sweep, the next sweep's loads issued
        // before this sweep's arithmetic and stores (a store never writes a
        // position a later load reads: ys may overwrite fin only at the
        // position just read)
        struct Sweep {
            double y, f, x, k[6], y4;
        };
        auto load = [&](Sweep& w, int64_t p) {
            const int64_t i = base + p;
            if (MODE != kDp5Last) {
                w.y = a.y[i];
#pragma unroll
                for (int j = 0; j < 6; ++j)
                    if (j < ng) w.k[j] = a.kg[j][i];
            }
            if (need_fin) w.f = a.fin[i];
            if (MODE != kDp5Stage && a.conserve && v.o.xb) w.x = v.o.xb[i];
            if (MODE == kDp5Last) w.y4 = a.y4[i];
        };
        auto step = [&](Sweep& w, int64_t p) {
            const int64_t i = base + p;
            double kn = 0.0;
            if (MODE != kDp5Stage) {
                kn = row[p];
                if (a.conserve) {
                    if (!v.o.xb) {
                        if (MODE == kDp5Mid && a.nfin > 0) {
                            // fin recomputed bit for bit: the previous pass's sum
                            double x = w.y;
#pragma unroll
                            for (int j = 0; j < 6; ++j)
                                if (j < a.nfin) x = x + cf[j] * w.k[j];
                            w.x = x;
                        } else {
                            w.x = w.f;
                        }
                    }
                    const double x = w.x;
                    const double wt = v.o.xb ? hat_defect_weight(x, nbk + 1.0, nbk, v.rkc[p], v.ckc[p], v.Vc[p])
                                             : fabs(x * (1.0 - x));
                    kn = kn - sc * wt;
                }
                a.kout[i] = kn;
            }
            if (MODE == kDp5Last) {
                const double y5 = w.f;
                const double err = y5 - (w.y4 + c7 * kn);
                const double ay = fabs(y5), a1 = fabs(1.0 - y5);
                const double scale = a.rtol * (ay < a1 ? ay : a1) + a.atol;
                row[p] = err / scale;
            } else {
                double sum = w.y;
#pragma unroll
                for (int j = 0; j < 6; ++j)
                    if (j < ng) sum = sum + c[j] * w.k[j];
                if (MODE == kDp5Mid) sum = sum + c[ng] * kn;
                a.ys[i] = sum;
                row[p] = sum;
                if (want_y4) {
                    double y4 = w.y;
#pragma unroll
                    for (int j = 0; j < 6; ++j)
                        if (j < ng) y4 = y4 + cb[j] * w.k[j];
                    a.y4[i] = y4 + cb[ng] * kn;
                }
            }
        };
        Sweep cur, nxt;
        int64_t p = threadIdx.x;
        if (p < n) load(cur, p);
        for (; p < n; p += kMemT) {
            const bool more = p + kMemT < n;
            if (more) load(nxt, p + kMemT);
            step(cur, p);
            if (more) cur = nxt;
        }
    }
    PackNodes pk;
    if (MODE != kDp5Last) load_pack_nodes(pk, n, v.kgroups, v.kt, v.npanel, v.perm);
    __syncthreads();
    if constexpr (MODE != kDp5Last) {
        pack_from_row(row, pk, live, col, n, v.kgroups, v.npanel, v.nt, v.o);
        // the next pass's finish reads these (every read of this member's
        // factors in this pass is behind the barrier above)
        if (v.o.xb && live) hat_from_row(row, n, v.o, b);
    } else {
        if (!a.norms) return;
        // sum over elements of sum_a r_a^2 + (sum_a r_a)^2 (the P1 integral of
        // r^2 up to the factor the controller divides by)
        const int nv = a.dim + 1;
        double s = 0.0, z1 = 0.0, z2 = 0.0, z3 = 0.0;
        for (int64_t e = threadIdx.x; e < a.nel; e += kMemT) {
            double sr = 0.0, sr2 = 0.0;
            for (int q = 0; q < nv; ++q) {
                const double r = row[a.conn[e * nv + q]];
                sr += r;
                sr2 += r * r;
            }
            s += sr2 + sr * sr;
        }
        block_sum4(s, z1, z2, z3, red);
        if (threadIdx.x == 0) a.norms[b] = s;
    }
}

// ------------------------------------------------------- panels with work
// A panel whose k range holds, for every member, only zero fp16 operand
// units (f and 1 - f below the split's range, e.g. the empty top and the
// filled bottom of a cold distribution) has C = A Z = 0 exactly: it is left
// out of the GEMM and its products read as 0; so is either chunk of a cut
// band on its own.  One workgroup decides which work units have work (some
// iteration of the unit's k range holds a nonzero operand of a kind some
// member runs the panel with: a difference of prefix counts of the
// iteration flags), writes each panel's slot mask (bit c: chunk c computed)
// for the epilogue, deals the units in decreasing length to the bins in
// snake order (the k-th unit of pass s to bin k on even passes,
// nbins - 1 - k on odd ones: every complete pass gives each bin one unit, so
// a bin's slots and offsets have closed forms), and clears the flags for the
// next evaluation.  Dynamic LDS: 2 (nit + 1) + npanel + nunit ints.
constexpr int kSchedT = 1024;

// exclusive prefix of one int per thread over the workgroup; *total gets the
// sum (wtot: kSchedT / 64 ints of LDS)
__device__ inline int block_exclusive_scan(int v, int* wtot, int* total) {
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    int inc = v;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const int y = __shfl_up(inc, o, 64);
        if (lane >= o) inc += y;
    }
    if (lane == 63) wtot[wv] = inc;
    __syncthreads();
    int wbase = 0, all = 0;
    for (int w = 0; w < kSchedT / 64; ++w) {
        if (w < wv) wbase += wtot[w];
        all += wtot[w];
    }
    __syncthreads();   // wtot may be reused after this
    *total = all;
    return wbase + inc - v;
}

// The columns of a fused DP5 attempt: member b < B that runs (step h[b]
// != 0) takes column slot[b], the running members in member order from 0;
// ncol[0] = their count.  Member tiles past it are not computed, so the
// batched GEMM's work shrinks as members finish.  Columns are computed
// independently: a member's results do not depend on its column.
__global__ void __launch_bounds__(kSchedT)
compact_kernel(const double* __restrict__ h, int B, int nt, int32_t* __restrict__ slot, int32_t* __restrict__ ncol) {
    __shared__ int wtot[kSchedT / 64];
    int base = 0;
    for (int b0 = 0; b0 < nt; b0 += kSchedT) {
        const int b = b0 + (int)threadIdx.x;
        const int run = (b < B && h[b] != 0.0) ? 1 : 0;
        int tot;
        const int r = block_exclusive_scan(run, wtot, &tot);
        if (b < nt) slot[b] = run ? base + r : -1;
        base += tot;
    }
    if (threadIdx.x == 0) ncol[0] = base;
}

__global__ void __launch_bounds__(kSchedT)
active_schedule_kernel(int npanel, int nunit, int wg_per_xcd, const int32_t* __restrict__ ncol, int nw, int nmt,
                       int nit, int fold, const int2* __restrict__ krange, const int32_t* __restrict__ cut,
                       const int32_t* __restrict__ blk_first, const int64_t* __restrict__ nnz_pref,
                       int2* __restrict__ urange, const int32_t* __restrict__ order, uint8_t* __restrict__ nzF,
                       uint8_t* __restrict__ nzG, uint8_t* __restrict__ useF, uint8_t* __restrict__ useG,
                       uint8_t* __restrict__ pmask, int32_t* __restrict__ sched_off,
                       int32_t* __restrict__ sched_panel, int64_t* __restrict__ stat) {
    extern __shared__ int32_t sh[];
    // bins of the GEMM launch that reads this schedule (active_tiles)
    const int T = active_tiles(ncol, nw, nmt);
    const int nbins = 8 * (wg_per_xcd / (T > 0 ? T : 1));
    int32_t* cF = sh;                 // nit + 1: iterations before it holding a nonzero f operand
    int32_t* cG = cF + nit + 1;       // nit + 1: the same for g
    int32_t* msk = cG + nit + 1;      // npanel: the panel's slot mask
    int32_t* flag = msk + npanel;     // nunit: 1 if the unit at that position of `order` has work
    __shared__ int wtot[kSchedT / 64];
    __shared__ unsigned long long iters, nnzs;
    const int tid = threadIdx.x;
    if (tid == 0) {
        iters = 0;
        nnzs = 0;
    }
    // prefix counts of the iteration flags: thread t counts its contiguous chunk
    {
        const int chunk = (nit + kSchedT - 1) / kSchedT;
        const int i0 = min(tid * chunk, nit), i1 = min(i0 + chunk, nit);
        int nf = 0, ng = 0;
        for (int i = i0; i < i1; ++i) {
            nf += nzF[i] != 0;
            ng += nzG[i] != 0;
        }
        int tf, tgv;
        int pf = block_exclusive_scan(nf, wtot, &tf);
        int pg = block_exclusive_scan(ng, wtot, &tgv);
        for (int i = i0; i < i1; ++i) {
            cF[i] = pf;
            cG[i] = pg;
            pf += nzF[i] != 0;
            pg += nzG[i] != 0;
        }
        if (tid == 0) {
            cF[nit] = tf;
            cG[nit] = tgv;
        }
    }
    __syncthreads();
    for (int pn = tid; pn < npanel; pn += kSchedT) {
        const int2 kr = krange[pn];
        const int k1 = kr.x + cut[pn];
        const bool uf = useF[pn] != 0, ug = useG[pn] != 0;
        // iterations [a, e) with a nonzero operand of a kind the panel uses:
        // the first (last) one is where the prefix count first (last) steps
        auto first_nz = [&](const int32_t* cnt, int a, int e) {
            int lo = a, hi = e - 1;   // cnt[lo + 1] > cnt[a] holds at hi
            while (lo < hi) {
                const int mid = (lo + hi) >> 1;
                if (cnt[mid + 1] > cnt[a]) hi = mid;
                else lo = mid + 1;
            }
            return lo;
        };
        auto last_nz = [&](const int32_t* cnt, int a, int e) {
            int lo = a, hi = e - 1;   // cnt[hi] < cnt[e] holds at lo
            while (lo < hi) {
                const int mid = (lo + hi + 1) >> 1;
                if (cnt[mid] < cnt[e]) lo = mid;
                else hi = mid - 1;
            }
            return lo;
        };
        int m = 0;
        unsigned long long it = 0, nz = 0;
        auto unit = [&](int c, int a, int e) {
            const bool wf = uf && a < e && cF[e] > cF[a], wg = ug && a < e && cG[e] > cG[a];
            if (!wf && !wg) return;
            int f0 = e, l0 = a;
            if (wf) {
                f0 = min(f0, first_nz(cF, a, e));
                l0 = max(l0, last_nz(cF, a, e));
            }
            if (wg) {
                f0 = min(f0, first_nz(cG, a, e));
                l0 = max(l0, last_nz(cG, a, e));
            }
            const int s0 = a + (f0 - a) / fold * fold;
            urange[2 * pn + c] = make_int2(s0 - kr.x, l0 + 1 - s0);
            m |= 1 << c;
            it += l0 + 1 - s0;
            const int64_t g = blk_first[pn] + (s0 - kr.x);
            nz += nnz_pref[g + l0 + 1 - s0] - nnz_pref[g];
        };
        unit(0, kr.x, k1);
        unit(1, k1, kr.y);
        msk[pn] = m;
        pmask[pn] = (uint8_t)m;
        if (it) {
            atomicAdd(&iters, it);
            atomicAdd(&nnzs, nz);
        }
    }
    __syncthreads();
    for (int sidx = tid; sidx < nunit; sidx += kSchedT) {
        const int u = order[sidx];
        flag[sidx] = (msk[u >> 1] >> (u & 1)) & 1;
    }
    __syncthreads();
    // ranks over `order`: thread t counts its contiguous chunk
    const int chunk = (nunit + kSchedT - 1) / kSchedT;
    const int s0 = min(tid * chunk, nunit), s1 = min(s0 + chunk, nunit);
    int cnt = 0;
    for (int sidx = s0; sidx < s1; ++sidx) cnt += flag[sidx];
    int A;
    int r = block_exclusive_scan(cnt, wtot, &A);
    const int full = A / nbins, rem = A % nbins;
    auto bin_off = [&](int b) {
        int pb = 0;
        if (rem) pb = (full % 2 == 0) ? min(b, rem) : max(0, b - (nbins - rem));
        return full * b + pb;
    };
    for (int sidx = s0; sidx < s1; ++sidx) {
        if (!flag[sidx]) continue;
        const int pass = r / nbins, k = r % nbins;
        const int b = (pass % 2 == 0) ? k : nbins - 1 - k;
        sched_panel[bin_off(b) + pass] = order[sidx];
        ++r;
    }
    for (int b = tid; b <= nbins; b += kSchedT) sched_off[b] = bin_off(b);
    if (tid == 0) {
        stat[0] = (int64_t)iters;
        stat[1] = (int64_t)nnzs;
    }
    // the flags were read above (before the first barrier): clear them
    for (int i = tid; i < nit; i += kSchedT) {
        nzF[i] = 0;
        nzG[i] = 0;
    }
    for (int i = tid; i < npanel; i += kSchedT) {
        useF[i] = 0;
        useG[i] = 0;
    }
}

// ------------------------------------------------------------------- host
template <int NW, int WAVES, int FOLD, int DEPTH>
static void launch_gemm_t(tg_ctx* c, MixedData* m, const int32_t* ncol) {
    const int nmt = m->nt / NW;
    const int64_t nblocks = (int64_t)m->npanel * m->nsplit * nmt;
    const int64_t per = (nblocks + 7) / 8;
    hipLaunchKernelGGL((mixed_gemm_kernel<NW, WAVES, FOLD, DEPTH>), dim3((unsigned)(8 * per)), dim3(64 * WAVES),
                       0, c->stream, (const u32x4*)m->A_band, m->band_off,
                       (const u32x4*)m->F_hi, (const u32x4*)m->F_lo, (const u32x4*)m->G_hi,
                       (const u32x4*)m->G_lo, m->useg, m->rows_pad, m->nt, m->krange, nmt, m->nsplit,
                       nblocks, per, ncol, (typename Partial<FOLD>::type*)m->P);
}

// the tile configurations the library carries: (members, waves, fold, prefetch depth)
#define TG_MIXED_CONFIGS(X)                                                                     \
    X(64, 4, 0, 1) X(64, 4, 1, 1) X(64, 4, 1, 2) X(64, 4, 1, 3) X(64, 4, 2, 2) X(128, 4, 0, 1)  \
    X(128, 4, 1, 1) X(128, 8, 1, 1) X(128, 8, 1, 2) X(128, 8, 1, 3) X(128, 8, 2, 2)             \
    X(256, 4, 0, 1) X(256, 4, 0, 2) X(256, 8, 0, 2)

// the persistent configurations: (members, fold, prefetch depth, waves per
// SIMD, A path: 0 straight into registers, 1 staged through LDS)
#define TG_PERSIST_CONFIGS(X)                                                                       \
    X(64, 2, 2, 1, 0) X(64, 2, 3, 1, 0) X(128, 1, 2, 1, 0) X(128, 2, 2, 1, 0)                           \
    X(128, 4, 2, 1, 0) X(128, 2, 2, 2, 0) X(128, 2, 3, 2, 0) X(128, 4, 2, 2, 0) X(128, 4, 3, 2, 0)       \
    X(64, 2, 2, 2, 0) X(64, 2, 3, 2, 0) X(128, 2, 2, 2, 1) X(128, 2, 3, 2, 1) X(128, 4, 2, 2, 1)         \
    X(64, 2, 2, 2, 1) X(128, 2, 1, 2, 1)

static bool config_known(bool persist, int nw, int waves, int fold, int depth, int al) {
    if (persist) {
#define TG_X(a, c_, d_, g_, z_) if (nw == a && fold == c_ && depth == d_ && waves == 4 * g_ && al == z_) return true;
        TG_PERSIST_CONFIGS(TG_X)
#undef TG_X
        return false;
    }
#define TG_X(a, b, c_, d_) if (nw == a && waves == b && fold == c_ && depth == d_) return true;
    TG_MIXED_CONFIGS(TG_X)
#undef TG_X
    return false;
}

// The persistent kernel's work units (a panel's band, or its two chunks
// when it is longer than cut_min iterations: the first half, rounded down,
// and the rest), by decreasing length (ties by unit index), with the slot
// mask every panel has when all units are computed.
static void unit_list(const MixedData* m, std::vector<int32_t>& units, std::vector<int64_t>& cost,
                      std::vector<uint8_t>& mask) {
    units.clear();
    cost.assign(2 * (size_t)m->npanel, 0);
    mask.assign(m->npanel, 0);
    for (int pn = 0; pn < m->npanel; ++pn) {
        const int N = m->h_krange[pn].y - m->h_krange[pn].x, c = m->h_cut[pn];
        if (c > 0) {
            units.push_back(2 * pn);
            cost[2 * pn] = c + 1;
            mask[pn] |= 1;
        }
        if (N - c > 0) {
            units.push_back(2 * pn + 1);
            cost[2 * pn + 1] = N - c + 1;
            mask[pn] |= 2;
        }
    }
    std::stable_sort(units.begin(), units.end(), [&](int a, int b) { return cost[a] > cost[b]; });
}

// Longest-processing-time schedule of the units over the bins of the
// persistent kernel: units by decreasing length, each onto the bin with the
// least work so far (ties to the lowest bin); the cost of a unit is its
// iteration count plus one for its write-out.  Built on the host, copied
// once; evaluations only launch.
static tg_status build_schedule(tg_ctx* c, MixedData* m) {
    if (m->sched_nw == m->nw) return TG_OK;
    const int nmt = m->nt / m->nw;
    const int per_xcd = std::max(1, m->ncu / 8);
    m->bins_per_xcd = std::max(1, per_xcd / nmt);
    const int nbins = 8 * m->bins_per_xcd;
    std::vector<int32_t> order;
    std::vector<int64_t> cost;
    std::vector<uint8_t> mask;
    unit_list(m, order, cost, mask);
    std::vector<int64_t> load(nbins, 0);
    std::vector<std::vector<int32_t>> lists(nbins);
    for (int u : order) {
        int best = 0;
        for (int b = 1; b < nbins; ++b)
            if (load[b] < load[best]) best = b;
        load[best] += cost[u];
        lists[best].push_back(u);
    }
    std::vector<int32_t> off(nbins + 1, 0), units;
    units.reserve(order.size());
    for (int b = 0; b < nbins; ++b) {
        units.insert(units.end(), lists[b].begin(), lists[b].end());
        off[b + 1] = (int32_t)units.size();
    }
    dev_free(&m->sched_off);
    dev_free(&m->sched_panel);
    TG_TRY(dev_alloc(c, &m->sched_off, (size_t)nbins + 1));
    TG_TRY(dev_alloc(c, &m->sched_panel, std::max<size_t>(units.size(), 1)));
    TG_HIP(c, hipMemcpyAsync(m->sched_off, off.data(), sizeof(int32_t) * (nbins + 1), hipMemcpyHostToDevice,
                             c->stream));
    if (!units.empty())
        TG_HIP(c, hipMemcpyAsync(m->sched_panel, units.data(), sizeof(int32_t) * units.size(),
                                 hipMemcpyHostToDevice, c->stream));
    TG_HIP(c, hipStreamSynchronize(c->stream));
    m->sched_nw = m->nw;
    return TG_OK;
}

template <int NW, int FOLD, int DA, int MG, int AL>
static void launch_persist_t(tg_ctx* c, MixedData* m, bool dyn, const int32_t* ncol) {
    static_assert(AL == 0 || AL == 1, "A path");
    const int nmt = m->nt / NW;
    hipLaunchKernelGGL((mixed_gemm_persist_kernel<NW, FOLD, DA, MG, AL == 1>), dim3((unsigned)(8 * m->bins_per_xcd * nmt)),
                       dim3(256 * MG), 0, c->stream, (const u32x4*)m->A_band, m->band_off,
                       (const u32x4*)m->F_hi, (const u32x4*)m->F_lo, (const u32x4*)m->G_hi,
                       (const u32x4*)m->G_lo, m->useg, m->rows_pad, m->nt, m->krange,
                       dyn ? m->dsched_off : m->sched_off, dyn ? m->dsched_panel : m->sched_panel, nmt,
                       m->bins_per_xcd, ncol, dyn ? m->durange : m->srange, (int64_t)m->nt * m->rows_pad,
                       (double*)m->P);
}

// ncol (device, optional): member tiles from *ncol on are not computed
static tg_status launch_gemm(tg_ctx* c, MixedData* m, bool dyn = false, const int32_t* ncol = nullptr) {
    if (m->nt % m->nw != 0) return set_error(c, TG_ERR_ARG, "tile width %d does not divide %d", m->nw, m->nt);
    if (m->persist) {
        if (m->sched_nw != m->nw || m->nsplit != 1)
            return set_error(c, TG_ERR_STATE, "persistent GEMM schedule not built for tile %d", m->nw);
#define TG_X(a, c_, d_, g_, z_)                                                                   \
    if (m->nw == a && m->fold == c_ && m->da == d_ && m->waves == 4 * g_ && m->al == z_)                  \
        launch_persist_t<a, c_, d_, g_, z_>(c, m, dyn, ncol); else
        TG_PERSIST_CONFIGS(TG_X)
#undef TG_X
        return set_error(c, TG_ERR_ARG, "no persistent kernel for tile %d members, fold %d, depth %d, %d waves, A in LDS %d",
                         m->nw, m->fold, m->da, m->waves, m->al);
        TG_CHECK_LAUNCH(c);
        return TG_OK;
    }
#define TG_X(a, b, c_, d_)                                                        \
    if (m->nw == a && m->waves == b && m->fold == c_ && m->depth == d_) launch_gemm_t<a, b, c_, d_>(c, m, ncol); else
    TG_MIXED_CONFIGS(TG_X)
#undef TG_X
    return set_error(c, TG_ERR_ARG, "no kernel for tile %d members, %d waves, fold %d, depth %d", m->nw,
                     m->waves, m->fold, m->depth);
    TG_CHECK_LAUNCH(c);
    return TG_OK;
}

static PackOut pack_out(MixedData* m, bool dyn, const uint32_t* hat_stencil, int hat_dim) {
    PackOut o;
    o.Fhi = (u32x4*)m->F_hi;
    o.Flo = (u32x4*)m->F_lo;
    o.Ghi = (u32x4*)m->G_hi;
    o.Glo = (u32x4*)m->G_lo;
    o.useg = m->useg;
    o.nzF = dyn ? m->nzF : nullptr;
    o.nzG = dyn ? m->nzG : nullptr;
    o.useF = dyn ? m->useF : nullptr;
    o.useG = dyn ? m->useG : nullptr;
    o.hs = m->xb ? hat_stencil : nullptr;
    o.xb = m->xb;
    o.dim = hat_dim;
    return o;
}

static tg_status launch_pack(tg_ctx* c, MixedData* m, const double* d_f, int B, const double* gate,
                             bool dyn = false) {
    const int64_t n = c->n;
    const int64_t kgroups = m->kpad / 8;
    const PackOut o = pack_out(m, dyn, c->d_hat, c->dim);
#define TG_PACK(ROW)                                                                                \
    hipLaunchKernelGGL(pack_member_kernel<ROW>, dim3(m->nt), dim3(kMemT), ROW ? sizeof(double) * n : 0, \
                       c->stream, d_f, n, B, m->perm, kgroups, m->kt, m->npanel, m->nt, o, gate)
    if (m->rowcache) TG_PACK(true);
    else TG_PACK(false);
#undef TG_PACK
    TG_CHECK_LAUNCH(c);
    return TG_OK;
}

static tg_status launch_finish(tg_ctx* c, MixedData* m, const double* d_f, double* d_I, int B,
                               int conserve, const double* gate = nullptr, double* flux = nullptr,
                               bool dyn = false) {
    const int64_t n = c->n;
    const size_t rowbytes = m->rowcache ? sizeof(double) * (size_t)n : 0;
    const double inv_scale = std::ldexp(1.0, -(m->sa + kSF));
    // the occupation factors of the rows' nodes: the hat averages the pack
    // wrote, or the nodal state
    const double* xf = m->xb ? m->xb : d_f;
    const uint8_t* pmask = slot_masks(m, dyn);
#define TG_FINISH(ROW, PT)                                                                           \
    hipLaunchKernelGGL((finish_member_kernel<ROW, PT>), dim3(m->nt), dim3(kMemT), ROW ? rowbytes : 0, \
                       c->stream, (const PT*)m->P, slots(m), m->nt, m->rows_pad, n, m->kt, B, xf,    \
                       m->perm, m->rk_s, m->ck_s, m->V_s, m->bath, m->useg, pmask,                  \
                       m->npanel, inv_scale, conserve,                                               \
                       gate, d_I, m->eps_s, flux, m->xb ? m->rk_c : nullptr, m->ck_c, c->d_V)
    if (m->fold) {
        if (m->rowcache) TG_FINISH(true, double);
        else TG_FINISH(false, double);
    } else {
        if (m->rowcache) TG_FINISH(true, float);
        else TG_FINISH(false, float);
    }
#undef TG_FINISH
    TG_CHECK_LAUNCH(c);
    return TG_OK;
}

static tg_status ensure_P(tg_ctx* c, MixedData* m) {
    const size_t psize = m->fold ? sizeof(double) : sizeof(float);
    const size_t need = (size_t)slots(m) * m->nt * m->rows_pad * psize;
    if (m->P_cap < need) {
        dev_free(&m->P);
        m->P_cap = 0;
        TG_TRY(dev_alloc(c, &m->P, need));
        m->P_cap = need;
    }
    return TG_OK;
}

// Autotuner: times every accuracy-safe configuration (FP64 fold, so no
// K-split is needed for accuracy) -- the persistent kernels, and the tiled
// grid kernels with one and two K-splits -- on this band, the GEMM and the
// epilogue, whose partial-sum reads grow with the split count, and keeps the
// fastest.  The operands are zeroed first; the time does not depend on their
// values.
constexpr int kTuneReps = 5;

struct TuneCand {
    bool persist;
    int nw, waves, fold, depth, nsplit;
    int al;
};

static void apply_cand(MixedData* m, const TuneCand& cd) {
    m->persist = cd.persist;
    m->nw = cd.nw;
    m->fold = cd.fold;
    m->nsplit = cd.nsplit;
    m->waves = cd.waves;
    m->al = cd.al;
    if (cd.persist) m->da = cd.depth;
    else m->depth = cd.depth;
}

static TuneCand current_cand(const MixedData* m) {
    return TuneCand{m->persist, m->nw, m->waves, m->fold, m->persist ? m->da : m->depth, m->nsplit, m->al};
}

static tg_status autotune(tg_ctx* c, MixedData* m) {
    std::vector<TuneCand> cands;
#define TG_X(a, c_, d_, g_, z_) if (a <= m->nt) cands.push_back(TuneCand{true, a, 4 * g_, c_, d_, 1, z_});
    TG_PERSIST_CONFIGS(TG_X)
#undef TG_X
#define TG_X(a, b, c_, d_)                                                        \
    if (c_ > 0 && a <= m->nt)                                                     \
        for (int ns = 1; ns <= 2; ++ns) cands.push_back(TuneCand{false, a, b, c_, d_, ns, 0});
    TG_MIXED_CONFIGS(TG_X)
#undef TG_X
    if (cands.empty()) return TG_OK;
    hipStream_t st = c->stream;
    const int64_t kgroups = m->kpad / 8;
    TG_HIP(c, hipMemsetAsync(m->F_hi, 0, sizeof(uint4) * kgroups * m->nt, st));
    TG_HIP(c, hipMemsetAsync(m->F_lo, 0, sizeof(uint4) * kgroups * m->nt, st));
    TG_HIP(c, hipMemsetAsync(m->G_hi, 0, sizeof(uint4) * kgroups * m->nt, st));
    TG_HIP(c, hipMemsetAsync(m->G_lo, 0, sizeof(uint4) * kgroups * m->nt, st));
    TG_HIP(c, hipMemsetAsync(m->useg, 0, (size_t)m->npanel * m->nt, st));
    TG_HIP(c, hipMemsetAsync(m->bath, 0, sizeof(double) * c->max_batch, st));
    m->bath_set.clear();
    double *d_zero = nullptr, *d_out = nullptr;
    TG_TRY(dev_alloc(c, &d_zero, (size_t)c->n * m->nt));
    TG_HIP(c, hipMemsetAsync(d_zero, 0, sizeof(double) * c->n * m->nt, st));
    TG_TRY(dev_alloc(c, &d_out, (size_t)c->n * m->nt));
    const int Bt = std::min(m->nt, (int)c->max_batch);
    hipEvent_t e0 = nullptr, e1 = nullptr;
    TG_HIP(c, hipEventCreate(&e0));
    TG_HIP(c, hipEventCreate(&e1));
    // the default keeps its place unless another candidate is >3% faster, so
    // near-equal candidates do not flip from one context to the next
    const TuneCand dflt = current_cand(m);
    float best = 0.0f, dflt_ms = 0.0f;
    TuneCand pick = cands[0];
    tg_status status = TG_OK;
    for (const TuneCand& cd : cands) {
        apply_cand(m, cd);
        if (cd.persist) status = build_schedule(c, m);
        if (status == TG_OK) status = ensure_P(c, m);
        if (status == TG_OK) status = launch_gemm(c, m);     // warm-up
        if (status != TG_OK && cd.persist) {
            // a persistent configuration the device cannot launch (LDS or
            // register limits) is skipped
            (void)hipGetLastError();
            status = TG_OK;
            continue;
        }
        if (status == TG_OK) status = launch_finish(c, m, d_zero, d_out, Bt, 0);
        if (status != TG_OK) break;
        if (hipEventRecord(e0, st) != hipSuccess) { status = set_error(c, TG_ERR_HIP, "event record"); break; }
        for (int r = 0; r < kTuneReps && status == TG_OK; ++r) {
            status = launch_gemm(c, m);
            if (status == TG_OK) status = launch_finish(c, m, d_zero, d_out, Bt, 0);
        }
        if (status != TG_OK) break;
        if (hipEventRecord(e1, st) != hipSuccess || hipEventSynchronize(e1) != hipSuccess) {
            status = set_error(c, TG_ERR_HIP, "event timing");
            break;
        }
        float ms = 0.0f;
        (void)hipEventElapsedTime(&ms, e0, e1);
        ms /= kTuneReps;
        if (cd.persist == dflt.persist && cd.nw == dflt.nw && cd.fold == dflt.fold && cd.depth == dflt.depth &&
            cd.nsplit == dflt.nsplit && cd.waves == dflt.waves && cd.al == dflt.al)
            dflt_ms = ms;
        if (best == 0.0f || ms < best) {
            best = ms;
            pick = cd;
        }
    }
    if (status == TG_OK && dflt_ms > 0.0f && best > 0.97f * dflt_ms) {
        best = dflt_ms;
        pick = dflt;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipStreamSynchronize(st);
    dev_free(&d_zero);
    dev_free(&d_out);
    TG_TRY(status);
    apply_cand(m, pick);
    if (m->persist) TG_TRY(build_schedule(c, m));
    TG_TRY(ensure_P(c, m));
    m->tuned_ms = best;
    return TG_OK;
}

static int nt_for(int B) {
    if (B <= 64) return 64;
    if (B <= 128) return 128;
    return 256;
}

tg_status mixed_prepare(tg_ctx* c) {
    if (c->mixed) return TG_OK;
    if (!c->assembled) return set_error(c, TG_ERR_STATE, "kernel not assembled");
    if (c->max_batch > 256) return set_error(c, TG_ERR_ARG, "mixed path supports max_batch <= 256");
    MixedData* m = new MixedData();
    c->mixed = m;
    const int64_t n = c->n;
    m->rows = 2 * n;
    m->kt = (n + kMT - 1) / kMT * kMT;
    m->rows_pad = 2 * m->kt;
    m->kpad = (n + kKS - 1) / kKS * kKS;
    m->npanel = (int)(m->rows_pad / kMT);
    m->nt = nt_for(c->max_batch);
    const int64_t kgroups = m->kpad / 8;
    hipStream_t st = c->stream;

    // node order shared with the plain path (energy order, built at context creation)
    m->perm = c->d_perm;

    // scale: largest |K| maps to 2^14 (fp16 max 65504)
    double kmax = 0.0;
    {
        // K 1 bounds every entry of a nonnegative K; use the exact max instead
        std::vector<double> rk(n);
        TG_HIP(c, hipMemcpyAsync(rk.data(), c->d_rk, sizeof(double) * n, hipMemcpyDeviceToHost, st));
        TG_HIP(c, hipStreamSynchronize(st));
        for (double v : rk) kmax = std::max(kmax, std::fabs(v));
    }
    m->sa = kmax > 0.0 ? (int)std::floor(std::log2(16384.0 / kmax)) : 0;
    const double scale = std::ldexp(1.0, m->sa);

    // band extent of every panel: the exact extents of K's rows (A rows < n)
    // and columns (the K^T half) where the fp16 split is nonzero, in k-groups
    std::vector<int32_t> rf, rl, cf, cl;
    TG_TRY(k_extents(c, scale, rf, rl, cf, cl, nullptr));
    m->h_krange.resize(m->npanel);
    int64_t items = 0;
    for (int pn = 0; pn < m->npanel; ++pn) {
        int32_t lo = INT32_MAX, hi = -1;
        for (int r = 0; r < kMT; ++r) {
            const int64_t row = (int64_t)pn * kMT + r;
            int32_t f = -1, l = -1;
            if (row < n) {
                f = rf[row];
                l = rl[row];
            } else if (row >= m->kt && row < m->kt + n) {
                f = cf[row - m->kt];
                l = cl[row - m->kt];
            }
            if (l >= 0) {
                lo = std::min(lo, f / 8);      // columns -> k-groups
                hi = std::max(hi, l / 8);
            }
        }
        int2 kr;
        if (hi < 0) {
            kr.x = kr.y = 0;
        } else {
            kr.x = lo / kKG;             // k-groups -> iterations
            kr.y = hi / kKG + 1;
        }
        m->h_krange[pn] = kr;
        items += kr.y - kr.x;
    }
    m->work_items = items;
    TG_TRY(dev_alloc(c, &m->krange, (size_t)m->npanel));
    TG_HIP(c, hipMemcpyAsync(m->krange, m->h_krange.data(), sizeof(int2) * m->npanel,
                             hipMemcpyHostToDevice, st));
    // the persistent kernel's chunks: with TG_MIXED_CUT = L > 0 bands
    // longer than L iterations are cut in two, so the units spread over more
    // bins when only some panels have work.  Off by default: on the 128 x
    // 128-section ensemble (B = 256) cuts at 24 iterations gave 9% less GEMM
    // time with a quarter of the band computed and 20% more with all of it
    // (the second slot's writes and reads, A bands fetched into two L2s).
    if (const char* e = std::getenv("TG_MIXED_CUT")) m->cut_min = std::max(0, std::atoi(e));
    m->h_cut.resize(m->npanel);
    m->pslots = 1;
    for (int pn = 0; pn < m->npanel; ++pn) {
        const int N = m->h_krange[pn].y - m->h_krange[pn].x;
        const bool two = m->cut_min > 0 && N > m->cut_min;
        m->h_cut[pn] = two ? N / 2 : N;
        if (two) m->pslots = 2;
    }
    TG_TRY(dev_alloc(c, &m->cut, (size_t)m->npanel));
    TG_HIP(c, hipMemcpyAsync(m->cut, m->h_cut.data(), sizeof(int32_t) * m->npanel, hipMemcpyHostToDevice, st));
    // band-major packed A from K's entries, then the split-matrix row sums
    // over it; sorted node volumes and energies
    {
        std::vector<int64_t> off(m->npanel);
        int64_t units = 0;
        int maxit = 0;
        for (int pn = 0; pn < m->npanel; ++pn) {
            off[pn] = units;
            const int nit = m->h_krange[pn].y - m->h_krange[pn].x;
            units += (int64_t)nit * kBandT;
            maxit = std::max(maxit, nit);
        }
        m->band_units = units;
        TG_TRY(dev_alloc(c, &m->band_off, (size_t)m->npanel));
        TG_HIP(c, hipMemcpyAsync(m->band_off, off.data(), sizeof(int64_t) * m->npanel, hipMemcpyHostToDevice, st));
        TG_TRY(dev_alloc(c, &m->A_band, (size_t)std::max<int64_t>(units, 1)));
        const int64_t nblk = units / kBandT;
        int32_t* d_cnt = nullptr;
        TG_TRY(dev_alloc(c, &d_cnt, (size_t)std::max<int64_t>(nblk, 1)));
        if (maxit > 0) {
            hipLaunchKernelGGL(band_pack_kernel, dim3(m->npanel, maxit), dim3(kKG * kMT), 0, st,
                               ksrc_of(c, c->d_Q), c->d_perm, n, m->kt, scale, m->krange, m->band_off,
                               m->A_band, d_cnt);
            TG_CHECK_LAUNCH(c);
        }
        TG_TRY(dev_alloc(c, &m->rk_s, (size_t)n));
        TG_TRY(dev_alloc(c, &m->ck_s, (size_t)n));
        TG_TRY(dev_alloc(c, &m->rk_c, (size_t)n));
        TG_TRY(dev_alloc(c, &m->ck_c, (size_t)n));
        m->V_s = c->d_V_s;
        m->eps_s = c->d_eps_s;
        hipLaunchKernelGGL(band_rowsum_kernel, dim3(ceil_div(m->rows_pad, 256)), dim3(256), 0, st, m->A_band,
                           m->band_off, m->krange, m->rows_pad, n, m->kt, 1.0 / scale, m->rk_s, m->ck_s);
        TG_CHECK_LAUNCH(c);
        hipLaunchKernelGGL(scatter_pair_kernel, dim3(ceil_div(n, 256)), dim3(256), 0, st, n, m->perm, m->rk_s,
                           m->ck_s, m->rk_c, m->ck_c);
        TG_CHECK_LAUNCH(c);
        // K's nonzeros per block (counted by the pack) and their prefix over
        // the band: the schedule sums them over the iterations a launch computes
        std::vector<int32_t> cnt(std::max<int64_t>(nblk, 1), 0);
        {
            hipError_t e = hipSuccess;
            if (nblk > 0) e = hipMemcpyAsync(cnt.data(), d_cnt, sizeof(int32_t) * nblk, hipMemcpyDeviceToHost, st);
            if (e == hipSuccess) e = hipStreamSynchronize(st);
            dev_free(&d_cnt);
            TG_HIP(c, e);
        }
        std::vector<int64_t> pref((size_t)nblk + 1, 0);
        for (int64_t g = 0; g < nblk; ++g) pref[g + 1] = pref[g] + cnt[g];
        m->band_nnz = pref[nblk];
        std::vector<int32_t> first(m->npanel);
        for (int pn = 0; pn < m->npanel; ++pn) first[pn] = (int32_t)(off[pn] / kBandT);
        TG_TRY(dev_alloc(c, &m->nnz_pref, pref.size()));
        TG_TRY(dev_alloc(c, &m->blk_first, (size_t)m->npanel));
        TG_HIP(c, hipMemcpyAsync(m->nnz_pref, pref.data(), sizeof(int64_t) * pref.size(), hipMemcpyHostToDevice, st));
        TG_HIP(c, hipMemcpyAsync(m->blk_first, first.data(), sizeof(int32_t) * m->npanel, hipMemcpyHostToDevice, st));
        TG_HIP(c, hipStreamSynchronize(st));
    }
    // configuration.  Default (kept by the autotuner unless beaten by >3%):
    // the persistent kernel, 128 members per tile (64 for smaller batches),
    // 8 waves, A staged through LDS, FP64 fold every 2 iterations, 2
    // iterations of prefetch.  The tiled grid kernel: 128 members (8 waves;
    // 64 members: 4 waves), fold 2, prefetch 2, one K-split or two when the
    // panels alone leave CUs idle.
    {
        hipDeviceProp_t prop;
        TG_HIP(c, hipGetDeviceProperties(&prop, c->device));
        m->ncu = prop.multiProcessorCount;
    }
    m->nw = std::min(128, m->nt);
    m->waves = 8;
    m->fold = 2;
    m->depth = 2;
    m->da = 2;
    m->persist = true;
    m->al = 1;
    m->nsplit = 1;
    // explicit overrides (measurements of the tile configuration) replace the autotuner
    bool fixed = false;
    for (const char* v : {"TG_MIXED_NW", "TG_MIXED_FOLD", "TG_MIXED_WAVES", "TG_MIXED_DEPTH", "TG_MIXED_SPLITS",
                          "TG_MIXED_PERSIST", "TG_MIXED_AL"})
        if (std::getenv(v)) fixed = true;
    if (const char* e = std::getenv("TG_MIXED_AUTOTUNE"))
        if (std::atoi(e) == 0) fixed = true;
    // the waves and the K-split belong to the tiled kernel: setting either
    // selects it unless TG_MIXED_PERSIST says otherwise
    if (std::getenv("TG_MIXED_WAVES") || std::getenv("TG_MIXED_SPLITS")) m->persist = false;
    if (const char* e = std::getenv("TG_MIXED_PERSIST")) m->persist = std::atoi(e) != 0;
    if (!m->persist && m->nw == 64) m->waves = 4;
    if (const char* e = std::getenv("TG_MIXED_NW")) {
        const int v = std::atoi(e);
        if ((v == 64 || v == 128 || v == 256) && v <= m->nt) m->nw = v;
    }
    if (const char* e = std::getenv("TG_MIXED_FOLD")) m->fold = std::atoi(e);
    if (const char* e = std::getenv("TG_MIXED_WAVES")) m->waves = std::atoi(e);
    if (const char* e = std::getenv("TG_MIXED_DEPTH")) m->depth = m->da = std::atoi(e);
    if (const char* e = std::getenv("TG_MIXED_AL")) m->al = std::atoi(e) != 0 ? 1 : 0;
    if (!m->persist) m->al = 0;
    if (!config_known(m->persist, m->nw, m->waves, m->fold, m->persist ? m->da : m->depth, m->al)) {
        m->persist = true;
        m->nw = std::min(128, m->nt);
        m->waves = 8;
        m->fold = 2;
        m->da = 2;
        m->al = 1;
    }
    if (!m->persist) {
        const int nmt = m->nt / m->nw;
        const int ns = (int)std::ceil(512.0 / ((double)m->npanel * nmt));
        m->nsplit = std::max(1, std::min(ns, 2));
        if (const char* e = std::getenv("TG_MIXED_SPLITS")) {
            const int v = std::atoi(e);
            if (v >= 1 && v <= 64) m->nsplit = v;
        }
    }
    // member rows in LDS when they fit (with the epilogue's reduction scratch)
    {
        hipDeviceProp_t prop;
        TG_HIP(c, hipGetDeviceProperties(&prop, c->device));
        const size_t rowbytes = sizeof(double) * (size_t)n;
        // the staged pack holds every k-group index and the panel index of a
        // thread in registers
        m->rowcache = rowbytes + 2048 <= (size_t)prop.sharedMemPerBlock &&
                      n <= 2 * (int64_t)kMemT * kStageUnr &&
                      m->kpad / 8 <= (int64_t)kPackKG * kMemT && m->npanel <= kMemT;
        if (const char* e = std::getenv("TG_MIXED_ROWCACHE"))
            if (std::atoi(e) == 0) m->rowcache = false;
        if (m->rowcache) {
            TG_HIP(c, hipFuncSetAttribute((const void*)pack_member_kernel<true>,
                                          hipFuncAttributeMaxDynamicSharedMemorySize, (int)rowbytes));
            TG_HIP(c, hipFuncSetAttribute((const void*)finish_member_kernel<true, double>,
                                          hipFuncAttributeMaxDynamicSharedMemorySize, (int)rowbytes));
            TG_HIP(c, hipFuncSetAttribute((const void*)finish_member_kernel<true, float>,
                                          hipFuncAttributeMaxDynamicSharedMemorySize, (int)rowbytes));
            const void* passes[] = {(const void*)dp5_member_kernel<kDp5Stage, double>,
                                    (const void*)dp5_member_kernel<kDp5Mid, double>,
                                    (const void*)dp5_member_kernel<kDp5Last, double>,
                                    (const void*)dp5_member_kernel<kDp5Stage, float>,
                                    (const void*)dp5_member_kernel<kDp5Mid, float>,
                                    (const void*)dp5_member_kernel<kDp5Last, float>};
            for (const void* k : passes)
                TG_HIP(c, hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)rowbytes));
        }
    }
    // the four operand arrays in one allocation, each kgroups x nt units
    TG_TRY(dev_alloc(c, &m->Z_all, 4 * (size_t)kgroups * m->nt));
    m->F_hi = m->Z_all;
    m->F_lo = m->Z_all + (size_t)kgroups * m->nt;
    m->G_hi = m->Z_all + 2 * (size_t)kgroups * m->nt;
    m->G_lo = m->Z_all + 3 * (size_t)kgroups * m->nt;
    TG_TRY(dev_alloc(c, &m->useg, (size_t)m->npanel * m->nt));
    TG_TRY(dev_alloc(c, &m->bath, (size_t)c->max_batch));
    if (c->occ == TG_OCC_HAT) {
        TG_TRY(dev_alloc(c, &m->xb, (size_t)n * m->nt));
        // defined values for the autotuner's epilogue runs on zero operands
        TG_HIP(c, hipMemsetAsync(m->xb, 0, sizeof(double) * (size_t)n * m->nt, c->stream));
    }
    if (m->persist) TG_TRY(build_schedule(c, m));
    // panels with work: flags cleared here once, then by the schedule kernel
    // after each use; the units by decreasing length for the dealing
    if (const char* e = std::getenv("TG_MIXED_SKIP"))
        if (std::atoi(e) == 0) m->skip = false;
    {
        m->nit = (int32_t)(m->kpad / kKS);
        TG_TRY(dev_alloc(c, &m->nzF, (size_t)m->nit));
        TG_TRY(dev_alloc(c, &m->nzG, (size_t)m->nit));
        TG_TRY(dev_alloc(c, &m->useF, (size_t)m->npanel));
        TG_TRY(dev_alloc(c, &m->useG, (size_t)m->npanel));
        std::vector<int32_t> order;
        std::vector<int64_t> cost;
        std::vector<uint8_t> mask;
        unit_list(m, order, cost, mask);
        m->nunit = (int)order.size();
        TG_TRY(dev_alloc(c, &m->actp, (size_t)m->npanel));
        TG_TRY(dev_alloc(c, &m->allmask, (size_t)m->npanel));
        TG_HIP(c, hipMemcpyAsync(m->allmask, mask.data(), (size_t)m->npanel, hipMemcpyHostToDevice, st));
        TG_HIP(c, hipMemcpyAsync(m->actp, mask.data(), (size_t)m->npanel, hipMemcpyHostToDevice, st));
        TG_TRY(dev_alloc(c, &m->lpt_order, std::max<size_t>(order.size(), 1)));
        {
            std::vector<int2> sr(2 * (size_t)m->npanel, make_int2(0, 0));
            for (int pn = 0; pn < m->npanel; ++pn) {
                const int N = m->h_krange[pn].y - m->h_krange[pn].x, cc = m->h_cut[pn];
                sr[2 * pn] = make_int2(0, cc);
                sr[2 * pn + 1] = make_int2(cc, N - cc);
            }
            TG_TRY(dev_alloc(c, &m->srange, sr.size()));
            TG_TRY(dev_alloc(c, &m->durange, sr.size()));
            TG_HIP(c, hipMemcpyAsync(m->srange, sr.data(), sizeof(int2) * sr.size(), hipMemcpyHostToDevice, st));
            TG_HIP(c, hipMemcpyAsync(m->durange, sr.data(), sizeof(int2) * sr.size(), hipMemcpyHostToDevice, st));
        }
        if (!order.empty())
            TG_HIP(c, hipMemcpyAsync(m->lpt_order, order.data(), sizeof(int32_t) * order.size(),
                                     hipMemcpyHostToDevice, st));
        // one entry per workgroup of any persistent launch (every workgroup
        // is a bin when one member tile runs) and one
        TG_TRY(dev_alloc(c, &m->dsched_off, (size_t)std::max(m->ncu, 8) + 8 * (size_t)m->nt + 1));
        TG_TRY(dev_alloc(c, &m->dsched_panel, std::max<size_t>(order.size(), 1)));
        // the schedule's LDS (64 KB): without it, every unit is computed
        if (sched_lds(m) > 65536) m->skip = false;
        TG_TRY(dev_alloc(c, &m->dstat, 2));
        TG_HIP(c, hipMemsetAsync(m->dstat, 0xff, 2 * sizeof(int64_t), st));   // -1: no evaluation yet
        TG_TRY(dev_alloc(c, &m->dslot, (size_t)m->nt));
        TG_TRY(dev_alloc(c, &m->dncol, 1));
        TG_HIP(c, hipMemsetAsync(m->nzF, 0, (size_t)m->nit, st));
        TG_HIP(c, hipMemsetAsync(m->nzG, 0, (size_t)m->nit, st));
        TG_HIP(c, hipMemsetAsync(m->useF, 0, (size_t)m->npanel, st));
        TG_HIP(c, hipMemsetAsync(m->useG, 0, (size_t)m->npanel, st));
        TG_HIP(c, hipStreamSynchronize(st));
    }
    if (!fixed) TG_TRY(autotune(c, m));
    TG_HIP(c, hipStreamSynchronize(st));
    return TG_OK;
}

void mixed_stats(tg_ctx* c, tg_kernel_stats* out) {
    MixedData* m = md_of(c);
    if (!m) return;
    out->mixed_band = m->work_items * (int64_t)kMT * kKS;
    out->mixed_splits = m->nsplit;
    out->mixed_width = m->nt;
    out->mixed_scale = m->sa;
    out->mixed_tile = m->nw;
    out->mixed_fold = m->fold;
    out->mixed_waves = m->waves;
    out->mixed_depth = m->persist ? m->da : m->depth;
    out->mixed_persist = m->persist ? 1 + m->al : 0;
    out->mixed_tuned_ms = m->tuned_ms;
    out->mixed_skip = m->persist && m->skip ? 1 : 0;
    out->mixed_last_band = out->mixed_band;
    out->mixed_nnz = m->band_nnz;
    out->mixed_last_nnz = m->band_nnz;
    out->mixed_cuts = 0;
    if (m->persist)
        for (int pn = 0; pn < m->npanel; ++pn)
            out->mixed_cuts += m->h_cut[pn] < m->h_krange[pn].y - m->h_krange[pn].x ? 1 : 0;
    if (out->mixed_skip && m->dstat) {
        int64_t st[2] = {-1, -1};
        if (hipMemcpyAsync(st, m->dstat, sizeof(st), hipMemcpyDeviceToHost, c->stream) == hipSuccess &&
            hipStreamSynchronize(c->stream) == hipSuccess && st[0] >= 0) {
            out->mixed_last_band = st[0] * (int64_t)kMT * kKS;
            out->mixed_last_nnz = st[1];
        }
    }
}

void mixed_destroy(tg_ctx* c) {
    MixedData* m = md_of(c);
    if (!m) return;
    dev_free(&m->A_band);
    dev_free(&m->band_off);
    dev_free(&m->Z_all);
    m->F_hi = m->F_lo = m->G_hi = m->G_lo = nullptr;
    dev_free(&m->useg);
    dev_free(&m->xb);
    dev_free(&m->krange);
    dev_free(&m->sched_off);
    dev_free(&m->sched_panel);
    dev_free(&m->nzF);
    dev_free(&m->nzG);
    dev_free(&m->useF);
    dev_free(&m->useG);
    dev_free(&m->lpt_order);
    dev_free(&m->actp);
    dev_free(&m->allmask);
    dev_free(&m->cut);
    dev_free(&m->srange);
    dev_free(&m->durange);
    dev_free(&m->dsched_off);
    dev_free(&m->dsched_panel);
    dev_free(&m->dstat);
    dev_free(&m->nnz_pref);
    dev_free(&m->blk_first);
    dev_free(&m->dslot);
    dev_free(&m->dncol);
    dev_free(&m->rk_s);
    dev_free(&m->ck_s);
    dev_free(&m->rk_c);
    dev_free(&m->ck_c);
    dev_free(&m->P);
    dev_free(&m->bath);
    delete m;
    c->mixed = nullptr;
}

// per-member bath occupations for the following evaluations
tg_status mixed_set_bath(tg_ctx* c, const double* h_bath, int32_t B) {
    TG_TRY(mixed_prepare(c));
    MixedData* m = md_of(c);
    if (B < 1 || B > m->nt) return set_error(c, TG_ERR_ARG, "batch %d exceeds the packed width %d", B, m->nt);
    std::vector<double> bath(B);
    for (int b = 0; b < B; ++b) bath[b] = h_bath ? h_bath[b] : c->bath_n;
    // the device copy already holds these values (bit for bit): no copy and no
    // stream synchronisation, so repeated evaluations keep the queue full
    if (bath.size() <= m->bath_set.size() &&
        std::memcmp(bath.data(), m->bath_set.data(), sizeof(double) * B) == 0)
        return TG_OK;
    TG_HIP(c, hipMemcpyAsync(m->bath, bath.data(), sizeof(double) * B, hipMemcpyHostToDevice,
                             c->stream));
    TG_HIP(c, hipStreamSynchronize(c->stream));
    m->bath_set = bath;
    return TG_OK;
}

// the units with work, their iteration ranges and their schedule from the
// flags the pack set
// ncol (device, optional): the columns running members fill (the GEMM that
// reads the schedule gets the same)
static tg_status launch_schedule(tg_ctx* c, MixedData* m, const int32_t* ncol = nullptr) {
    hipLaunchKernelGGL(active_schedule_kernel, dim3(1), dim3(kSchedT), sched_lds(m), c->stream, m->npanel,
                       m->nunit, m->bins_per_xcd * (m->nt / m->nw), ncol, m->nw, m->nt / m->nw,
                       m->nit, m->fold, m->krange, m->cut, m->blk_first, m->nnz_pref, m->durange, m->lpt_order,
                       m->nzF, m->nzG, m->useF,
                       m->useG, m->actp,
                       m->dsched_off, m->dsched_panel, m->dstat);
    TG_CHECK_LAUNCH(c);
    return TG_OK;
}

template <int MODE, typename PT>
static void launch_pass_t(tg_ctx* c, MixedData* m, const Dp5PassArgs& a, int B, bool dyn) {
    MixedView<PT> v;
    v.P = (const PT*)m->P;
    v.nslot = slots(m);
    v.nt = m->nt;
    v.B = B;
    v.npanel = m->npanel;
    v.rows_pad = m->rows_pad;
    v.n = c->n;
    v.kt = m->kt;
    v.kgroups = m->kpad / 8;
    v.perm = m->perm;
    v.rk = m->rk_s;
    v.ck = m->ck_s;
    v.V = m->V_s;
    v.eps_s = m->eps_s;
    v.useg = m->useg;
    v.pmask = slot_masks(m, dyn);
    v.slot = m->dslot;
    v.inv_scale = std::ldexp(1.0, -(m->sa + kSF));
    v.o = pack_out(m, dyn, c->d_hat, c->dim);
    v.rkc = m->rk_c;
    v.ckc = m->ck_c;
    v.Vc = c->d_V;
    hipLaunchKernelGGL((dp5_member_kernel<MODE, PT>), dim3(m->nt), dim3(kMemT), sizeof(double) * (size_t)c->n,
                       c->stream, a, v);
}

bool mixed_dp5_fusable(tg_ctx* c) {
    MixedData* m = md_of(c);
    return m && m->rowcache && c->ee == nullptr;
}

// one fused member pass of the batched DP5 (Dp5PassArgs); kernel launches only
tg_status mixed_dp5_pass(tg_ctx* c, int mode, const Dp5PassArgs& a, int32_t B) {
    MixedData* m = md_of(c);
    if (!mixed_dp5_fusable(c)) return set_error(c, TG_ERR_STATE, "fused DP5 passes need member rows in LDS and no e-e channel");
    if (B < 1 || B > m->nt) return set_error(c, TG_ERR_ARG, "batch %d exceeds the packed width %d", B, m->nt);
    if (a.ng < 0 || a.ng > 6) return set_error(c, TG_ERR_ARG, "fused DP5 pass with %d stage terms", a.ng);
    TG_TRY(ensure_P(c, m));
    const bool dyn = m->persist && m->skip;
#define TG_PASS(MODE)                                                              \
    if (m->fold) launch_pass_t<MODE, double>(c, m, a, B, dyn);                     \
    else launch_pass_t<MODE, float>(c, m, a, B, dyn);
    switch (mode) {
    case kDp5Stage: TG_PASS(kDp5Stage) break;
    case kDp5Mid: TG_PASS(kDp5Mid) break;
    case kDp5Last: TG_PASS(kDp5Last) break;
    default: return set_error(c, TG_ERR_ARG, "fused DP5 pass mode %d", mode);
    }
#undef TG_PASS
    TG_CHECK_LAUNCH(c);
    return TG_OK;
}

// the GEMM of the operands the last kDp5Stage / kDp5Mid pass packed (with its
// panel schedule); kernel launches only
tg_status mixed_dp5_gemm(tg_ctx* c) {
    MixedData* m = md_of(c);
    if (!m) return set_error(c, TG_ERR_STATE, "batched path not prepared");
    const bool dyn = m->persist && m->skip;
    // the persistent GEMM regroups its bins for the running tiles only with
    // the schedule built for them (dyn); the tiled kernel skips empty tiles
    const int32_t* ncol = (dyn || !m->persist) ? m->dncol : nullptr;
    if (dyn) TG_TRY(launch_schedule(c, m, ncol));
    return launch_gemm(c, m, dyn, ncol);
}

// the columns of this attempt's running members (h: the per-member step
// sizes after the attempt's prep); kernel launch only
tg_status mixed_dp5_compact(tg_ctx* c, const double* h, int32_t B) {
    MixedData* m = md_of(c);
    if (!m || !m->dslot) return set_error(c, TG_ERR_STATE, "batched path not prepared");
    if (B < 1 || B > m->nt) return set_error(c, TG_ERR_ARG, "batch %d exceeds the packed width %d", B, m->nt);
    hipLaunchKernelGGL(compact_kernel, dim3(1), dim3(kSchedT), 0, c->stream, h, B, m->nt, m->dslot, m->dncol);
    TG_CHECK_LAUNCH(c);
    return TG_OK;
}

// one batched evaluation with the bath set by mixed_set_bath: kernel launches
// only, so a caller may capture it in a graph once it has run once (buffers
// are allocated on first use)
tg_status mixed_eval_device(tg_ctx* c, const double* d_f, double* d_I, int32_t B, int conserve,
                            const double* gate, double* flux) {
    MixedData* m = md_of(c);
    if (!m) return set_error(c, TG_ERR_STATE, "batched path not prepared");
    if (B < 1 || B > m->nt) return set_error(c, TG_ERR_ARG, "batch %d exceeds the packed width %d", B, m->nt);
    TG_TRY(ensure_P(c, m));
    // panels with all-zero operands leave the persistent GEMM
    const bool dyn = m->persist && m->skip;
    TG_TRY(launch_pack(c, m, d_f, B, gate, dyn));
    if (dyn) TG_TRY(launch_schedule(c, m));
    time_begin(c);
    TG_TRY(launch_gemm(c, m, dyn));
    time_end(c);
    TG_TRY(launch_finish(c, m, d_f, d_I, B, conserve, gate, flux, dyn));
    return TG_OK;
}

double* mixed_bath_device(tg_ctx* c) {
    MixedData* m = md_of(c);
    return m ? m->bath : nullptr;
}

void mixed_bath_forget(tg_ctx* c) {
    MixedData* m = md_of(c);
    if (m) m->bath_set.clear();
}

tg_status mixed_apply_device(tg_ctx* c, const double* d_f, double* d_I, int32_t B,
                             const double* h_bath, int conserve) {
    TG_TRY(mixed_set_bath(c, h_bath, B));
    return mixed_eval_device(c, d_f, d_I, B, conserve, nullptr);
}

}  // namespace tg
