// Pair-binning kernels for gfx950 (MI355X).
//
// For every forest pair (f1, f2) of the neighbour list and every pixel pair
// (i in f1, j in f2) the reference (cuda_kernels.cpp:90-162) evaluates
//     c  = sqrt(0.5 (1 + cos12)),  s = sqrt(0.5 (1 - cos12)),
//     rp = |r1 - r2| c,  rt = (r1 + r2) s,
//     binp = floor(rp * NP / rpmax),  bint = floor(rt * NT / rtmax)
// and, when both bins are in range, adds w1 w2 and dw1 dw2 to the bin.
//
// Execution model (one persistent grid):
//   * work = tiles (runs of items of one output slot, taken from an atomic
//     queue); item = (forest1, a range of its neighbours); task = (64*PPL-pixel
//     slab of forest1, one neighbour).  A wave runs one task at a time.
//   * lanes hold forest-1 pixels in registers; forest-2 pixels are walked in a
//     wave-uniform loop and read with scalar loads (constant address space),
//     so every lane sees the same j and no LDS traffic is spent on inputs.
//   * per-task j range is pruned to pixels that can fall inside rmax for some
//     lane (forests are monotone in dc); pruned pairs are provably out of range.
//   * histograms are privatised in LDS (one copy per workgroup, optionally
//     replicated) and merged into per-slot FP64 global sums with atomics at the
//     end of each tile.
//   * binnings larger than the LDS histogram are cut into bands (blocks of
//     r_perp columns, and of r_parallel rows when one column per row no longer
//     fits).  A tile is walked once per band; each task's j range is pruned to
//     the pixels that can land in the band (for sorted forests r_perp grows with
//     j, so a forest pair mostly touches one or two column bands), and the band's
//     sums are flushed before the next band starts.
//
// Precision modes:
//   LYA_MODE_FP64   the reference formula in FP64 (no FMA contraction),
//                   FP64 LDS sums.
//   LYA_MODE_MIXED  FP32 geometry on per-forest offsets (dc - base), binner
//                   folded into the half-angles; any pixel pair whose scaled
//                   rp or rt lies within the rounding bound tau of a bin edge
//                   is recomputed with the FP64 formula, so the bin indices
//                   equal the FP64 ones exactly.  Inputs w, dw are FP32; their
//                   products are formed exactly in FP64 and summed in FP64 LDS.
//   LYA_MODE_MIXED_F32  the same geometry with FP32 products and FP32 LDS
//                   partials, flushed every work item into FP64 LDS sums
//                   (fewer LDS bytes, ~1e-6 relative error on the sums).
#include <hip/hip_runtime.h>
#include <algorithm>
#include <climits>
#include <cmath>
#include <cstdint>
#include <type_traits>

#include "../../include/lya2pcf.h"
#include "errors.h"

#pragma clang fp contract(off)

#define CAS __attribute__((address_space(4)))

// forest-2 pixels per scalar-load block
#ifndef LYA_UNROLL_FP64
#define LYA_UNROLL_FP64 4
#endif
#ifndef LYA_UNROLL_MIXED
#define LYA_UNROLL_MIXED 8
#endif

namespace {

constexpr int HS = LYA_MAX_BINS;  // LDS histogram entries per replica

template <class T>
__device__ __forceinline__ const CAS T* cptr(const T* p) {
    return (const CAS T*)(p);
}

__device__ __forceinline__ int uni(int x) { return __builtin_amdgcn_readfirstlane(x); }
__device__ __forceinline__ float unif(float x) {
    return __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(x)));
}
__device__ __forceinline__ double unid(double x) {
    const uint64_t b = (uint64_t)__double_as_longlong(x);
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)b);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)(b >> 32));
    return __longlong_as_double((long long)(((uint64_t)hi << 32) | lo));
}

// Half-angle factors of a forest pair, the reference expression evaluated in
// FP64 without contraction.  cos12 is clamped to [-1, 1] (the reference would
// produce NaN for cos12 > 1, which only happens for coincident lines of sight).
__device__ __forceinline__ void halfangles(double x1, double y1, double z1, double x2, double y2, double z2,
                                           double& c, double& s) {
    double cos12 = x1 * x2 + y1 * y2;
    cos12 = cos12 + z1 * z2;
    cos12 = fmin(1.0, fmax(-1.0, cos12));
    c = sqrt(0.5 * (1.0 + cos12));
    s = sqrt(0.5 * (1.0 - cos12));
}

// Keeps x as one VGPR value so that a + x*k folds into a single v_mad_u32_u24
// instead of being reassociated into mul + shift + add3 (no instruction emitted).
__device__ __forceinline__ uint32_t opaque(uint32_t x) {
    asm volatile("" : "+v"(x));
    return x;
}

template <class T>
__device__ __forceinline__ void lds_add(unsigned char* lds, uint32_t byte_off, T v) {
    __hip_atomic_fetch_add((T*)(lds + byte_off), v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

// LDS layout (bytes).  MIXED_F32: per replica [cnt u32 | w f32 | dw f32] x HS,
// then the FP64 accumulators [w f64 | dw f64] x HS.  FP64 and MIXED: per
// replica [cnt u32 | w f64 | dw f64] x HS, without the count array when pairs
// are not counted (less LDS to clear and flush per tile).
template <int MODE, bool CNT>
struct Layout {
    static constexpr bool F32 = (MODE == LYA_MODE_MIXED_F32);
    static constexpr bool HAS_CNT = F32 || CNT;
    static constexpr uint32_t REPL = F32 ? 12u * HS : (CNT ? 20u : 16u) * HS;
    static constexpr uint32_t W = HAS_CNT ? 4u * HS : 0u;      // w offset in a replica
    static constexpr uint32_t DW = F32 ? 8u * HS : W + 8u * HS;  // dw offset
    static constexpr __host__ __device__ uint32_t bytes(int R) { return REPL * (uint32_t)R + (F32 ? 16u * HS : 0u); }
    static constexpr __host__ __device__ uint32_t acc(int R) { return REPL * (uint32_t)R; }
};

struct TaskCtx {
    int f1;
    int64_t p1;
    int n1;
    int slot;
};

// Bins held by the LDS histogram: rows [bp0, bp0 + rows) x columns [bt0, bt0 + cols)
// of the binning (the whole binning when it fits).
struct Band {
    uint32_t bp0, bt0, rows, cols;
};

// Band grid of a binning too large for the LDS histogram: br x bc bins per band,
// nbp x nbt bands (a kernel argument; unused when the binning fits).
struct BandShape {
    int32_t br, bc, nbp, nbt;
};

// Geometry of a forest pair shared by every slab of it: the half-angle factors and
// the pruning thresholds.  Dp bounds |r1 - r2| and Rt bounds r1 + r2 for pairs in
// rows below np_hi and columns below nt_hi; Rlo = r1 + r2 below which r_perp is under
// the band's first column (banded kernels).  The margins (1e-3 bin) dominate the FP64
// rounding of the thresholds, so pruned pairs are out of range exactly.
struct PairGeo {
    double c, s, Dp, Rt, Rlo;
};
template <bool BANDED = false>
__device__ __forceinline__ PairGeo pair_geo(const lya_xi_problem& p, int f1, int f2, double bp64, double bt64,
                                            const Band& bd = Band{0u, 0u, 0u, 0u}) {
    PairGeo g;
    halfangles(cptr(p.ux)[f1], cptr(p.uy)[f1], cptr(p.uz)[f1], cptr(p.ux)[f2], cptr(p.uy)[f2], cptr(p.uz)[f2],
               g.c, g.s);
    g.c = unid(g.c);
    g.s = unid(g.s);
    const double np_hi = BANDED ? (double)(bd.bp0 + bd.rows) : (double)p.np_rp;
    const double nt_hi = BANDED ? (double)(bd.bt0 + bd.cols) : (double)p.np_rt;
    g.Dp = (np_hi + 1e-3) / (g.c * bp64);
    g.Rt = (nt_hi + 1e-3) / (g.s * bt64);
    g.Rlo = BANDED && bd.bt0 > 0u ? ((double)bd.bt0 - 1e-3) / (g.s * bt64) : -INFINITY;
    return g;
}

// Forest-2 j range, within [j0, j1), that can hold an in-range pair for some
// forest-1 pixel of the slab [i0, ilast].  Needs both forests sorted by dc.
// BANDED: pairs that land in the band only: r_perp = (r1 + r2) s grows with j, so the
// band's columns give a lower bound on j as well; the row bound uses the band's last
// row (rows from 0 upwards, a superset).
template <bool BANDED = false>
__device__ __forceinline__ void prune(const lya_xi_problem& p, int f1, int64_t p1, int i0, int ilast, int f2,
                                      int64_t p2, int j0, int j1, const PairGeo& g, int lane, int& jlo, int& jhi) {
    jlo = j0;
    jhi = j1;
    if (!(cptr(p.sorted)[f1] & cptr(p.sorted)[f2])) return;
    const double r1min = cptr(p.dc)[p1 + i0];
    const double r1max = cptr(p.dc)[p1 + ilast];
    const double lo = r1min - g.Dp;
    const double hi = fmin(r1max + g.Dp, g.Rt - r1min);
    // band: r2 below this bound puts r_perp under the band's first column for every lane
    const double lot = g.Rlo - r1max;
    // the whole range below lo (or lot) or from hi on: empty, without the counting pass
    if (j1 <= j0 || cptr(p.dc)[p2 + j1 - 1] <= lo || cptr(p.dc)[p2 + j0] >= hi ||
        (BANDED && cptr(p.dc)[p2 + j1 - 1] < lot)) {
        jlo = jhi = j0;
        return;
    }
    // counting pass, 256 pixels per round: the four loads of a round are issued together
    // (one load latency per round instead of one per 64 pixels)
    int clo = 0, chi = 0;
    for (int jb = j0; jb < j1; jb += 256) {
        double r2[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int jj = jb + 64 * u + lane;
            r2[u] = (jj < j1) ? p.dc[p2 + jj] : INFINITY;
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            if constexpr (BANDED)
                clo += __popcll(__ballot(r2[u] <= lo || r2[u] < lot));
            else
                clo += __popcll(__ballot(r2[u] <= lo));
            chi += __popcll(__ballot(r2[u] < hi));
        }
    }
    jlo = j0 + uni(clo);
    jhi = j0 + uni(chi > clo ? chi : clo);
}

// Work units of a forest pair.  Only the forest-1 pixels [ia, ib) that can pair with
// the neighbour (one counting pass against the neighbour's whole range) take part: m
// slabs of SL = 64*PPL pixels from ia; with split > 0 the rest goes to nv slabs of 64
// pixels and a tail of at most `split` (< 64) pixels [ib - tail, ib), walked
// wave-uniformly against the neighbour's pixels in lanes: the neighbour's pixels
// [ja, jb) that can pair with the tail, in 64-pixel slabs from ja.  A short tail in lanes
// would leave most lanes idle over a window of ~2 rmax; forest 2 in the lanes keeps them
// busy.  split = 0: slabs of SL only, the last one partial.  Trimming the range first
// skips the slabs that cannot pair at all (about half of all slabs at 100k forests)
// with one pass per pair instead of one test per slab.
struct Units {
    int ia, ib, m, nv, tail, upq;
};
__device__ __forceinline__ Units units_of(int ia, int ib, int SL, int split) {
    Units u;
    u.ia = ia;
    u.ib = ib;
    const int n = ib - ia;
    if (split > 0) {
        u.m = n / SL;
        const int r = n - u.m * SL;
        u.nv = r >> 6;
        u.tail = r & 63;
        if (u.tail > split) {  // a longer tail takes one more 64-pixel slab
            ++u.nv;
            u.tail = 0;
        }
    } else {
        u.m = (n + SL - 1) / SL;
        u.nv = 0;
        u.tail = 0;
    }
    u.upq = u.m + u.nv + (u.tail > 0 ? 1 : 0);
    return u;
}

// Arguments of unit `sub` of the pair (forest 1 = t, neighbour f2 at p2 with n2 pixels):
// the lanes' forest tl (lanes past tl.n1 hold no pixel) from pixel i0 against the
// pixels [u0, u1) of forest fu.  A tail unit has forest 2 in the lanes and steps i0
// through its 64-pixel slabs from ja to jb.
struct UnitArgs {
    TaskCtx tl;
    int fu, u0, u1, i0;
    bool tail, wide;
};
__device__ __forceinline__ UnitArgs unit_args(const TaskCtx& t, const Units& u, int SL, int sub, int f2, int64_t p2,
                                              int n2, int ja, int jb) {
    UnitArgs a;
    a.tl = t;
    a.tl.n1 = u.ib;
    a.fu = f2;
    a.u0 = 0;
    a.u1 = n2;
    a.wide = sub < u.m;
    a.i0 = u.ia + (a.wide ? sub * SL : u.m * SL + (sub - u.m) * 64);
    a.tail = sub >= u.m + u.nv;
    if (a.tail) {
        a.tl.f1 = f2;
        a.tl.p1 = p2;
        a.tl.n1 = jb;
        a.fu = t.f1;
        a.u0 = u.ib - u.tail;
        a.u1 = u.ib;
        a.i0 = ja;
    }
    return a;
}

// The units of the pair (t, f2): the forest-1 range [ia, ib) that can pair with the
// neighbour, and for a tail the neighbour's range [ja, jb) that can pair with the tail.
// Returns false when no pixel pair of the two forests can be in range.
template <bool BANDED>
__device__ __forceinline__ bool pair_units(const lya_xi_problem& p, const TaskCtx& t, int f2, int64_t p2, int n2,
                                           const PairGeo& g, int SL, int split, int lane, Units& un, int& ja,
                                           int& jb) {
    if (n2 <= 0 || t.n1 <= 0) return false;
    int ia, ib;
    prune<BANDED>(p, f2, p2, 0, n2 - 1, t.f1, t.p1, 0, t.n1, g, lane, ia, ib);
    if (ia >= ib) return false;
    un = units_of(ia, ib, SL, split);
    ja = jb = 0;
    if (un.tail > 0) prune<BANDED>(p, t.f1, t.p1, ib - un.tail, ib - 1, f2, p2, 0, n2, g, lane, ja, jb);
    return true;
}

// The last jhi - j < U forest-2 pixels of a task: blocks of 4, 2 and 1 pixels, one
// scalar-load round each, instead of one round per pixel (fp64 6 %, mixed 2 % faster
// at 100k forests).
template <int U, class B>
__device__ __forceinline__ void rest_blocks(B&& block, int j, int jhi) {
    if constexpr (U > 4) {
        if (j + 4 <= jhi) {
            block(j, std::integral_constant<int, 4>{});
            j += 4;
        }
    }
    if constexpr (U > 2) {
        if (j + 2 <= jhi) {
            block(j, std::integral_constant<int, 2>{});
            j += 2;
        }
    }
    if (j < jhi) block(j, std::integral_constant<int, 1>{});
}

// ---------------------------------------------------------------------------
// One task: a 64*PPL-pixel slab of the lanes' forest (t, from pixel i0) against the
// pixels [u0, u1) of forest f2, walked wave-uniformly.  The bins and products are
// symmetric in the two forests, so either forest of a pair can take the lanes.
//   CNT   also count pixel pairs per bin (one more LDS atomic per pair)
//   POW2  FP64 mode with power-of-two binners: the binner is folded into the
//         half-angle factors, fl(|d| (c b)) = fl(|d| c) b exactly when b = 2^k,
//         so the bin equals the reference's floor(fl(fl(|d| c) b)).
//   BANDED  the LDS histogram holds the band bd of the binning: bins are taken
//         relative to (bd.bp0, bd.bt0) and pairs outside the band are out of range.
// ---------------------------------------------------------------------------
template <int MODE, int PPL, bool CNT, bool POW2, bool MERGE, bool BANDED>
__device__ __forceinline__ void run_task(const lya_xi_problem& p, unsigned char* lds, uint32_t hbase,
                                         uint32_t ntpad4, const TaskCtx& t, int i0, int f2, int u0, int u1,
                                         const PairGeo& g, int lane, int perm, double bp64, double bt64,
                                         const Band& bd) {
    static_assert(!(BANDED && MODE == LYA_MODE_MIXED_F32), "mixed32 has no banded form");
    using L = Layout<MODE, CNT>;
    // rows and columns of the LDS histogram
    const int NP = BANDED ? (int)bd.rows : p.np_rp, NT = BANDED ? (int)bd.cols : p.np_rt;
    const int64_t p2 = cptr(p.pix_off)[f2];
    const double c = g.c, s = g.s;
    const int ilast = min(t.n1, i0 + 64 * PPL) - 1;
    int jlo, jhi;
    prune<BANDED>(p, t.f1, t.p1, i0, ilast, f2, p2, u0, u1, g, lane, jlo, jhi);
    if (jlo >= jhi) return;
    // the lanes' pixels are loaded after the pruning pass: many tasks are empty, and
    // loading ahead of the pass measured 7 % slower
    double r1[PPL], w1[PPL], d1[PPL];
    int64_t pix1[PPL];
    bool valid[PPL];
#pragma unroll
    for (int k = 0; k < PPL; ++k) {
        const int i = i0 + k * 64 + perm;
        valid[k] = i < t.n1;
        pix1[k] = t.p1 + i;
        r1[k] = valid[k] ? p.dc[pix1[k]] : 1.0e9;
        if constexpr (MODE != LYA_MODE_MIXED_F32) {
            // FP64 sums: the pixel weights are read as FP64
            w1[k] = valid[k] ? p.we[pix1[k]] : 0.0;
            d1[k] = valid[k] ? p.dw[pix1[k]] : 0.0;
        }
    }
    // LDS addresses: count at a4 = hbase + 4 (bp ntpad + bt), sums at
    // a8 = hbase + W + 8 (bp ntpad + bt) = 2 a4 + (W - hbase).
    const uint32_t woff = L::W - hbase;
    const uint32_t ntpad8 = 2u * ntpad4;
    const uint32_t hbw = hbase + L::W;

    if constexpr (MODE != LYA_MODE_MIXED_F32) {
        // forest-2 pixels are wave-uniform: scalar loads through the constant cache
        const CAS double* r2p = cptr(p.dc) + p2;
        const CAS double* w2p = cptr(p.we) + p2;
        const CAS double* d2p = cptr(p.dw) + p2;
        // Merged kernels sum the products of lane pixel k over runs of equal bins
        // (consecutive forest-2 pixels of a lane often share a bin) and make one set of
        // LDS atomics per run.  The open run is carried from block to block of forest-2
        // pixels and flushed when its bin changes and once at the end of the task.
        // LDS index of bin (bp, bt) for an in-range pair.  Out of range: NONE without
        // merged runs; with them a trash entry in the histogram's padding column (row
        // lane % NP, column NT; the host gives merged kernels a padded row stride),
        // so that every run can be flushed without testing it, and the tile flush never
        // reads that column.  Lanes own their trash entry (two lanes share one at 50 rows);
        // a lane whose pairs are all out of range writes it once per task.
        constexpr uint32_t NONE = 0xFFFFFFFFu;
        const uint32_t trash = MERGE ? (uint32_t)(lane % NP) * (ntpad4 >> 2) + (uint32_t)NT : NONE;
        uint32_t cpk[PPL], can[PPL];
        double caw[PPL], cad[PPL];
#pragma unroll
        for (int k = 0; k < PPL; ++k) {
            cpk[k] = trash;
            can[k] = 0u;
            caw[k] = 0.0;
            cad[k] = 0.0;
        }
        auto flush_run = [&](uint32_t pk, double aw, double ad, uint32_t an) {
            const uint32_t a4 = (pk << 2) + hbase;
            if constexpr (CNT) lds_add<uint32_t>(lds, a4, an);
            const uint32_t a8 = (a4 << 1) + woff;
            lds_add<double>(lds, a8, aw);
            lds_add<double>(lds, a8 + (L::DW - L::W), ad);
        };
        auto run_sums_carry = [&](auto UC, const uint32_t* keys, double w1k, const double* w2, double d1k,
                                  const double* d2, uint32_t& pk, double& aw, double& ad, uint32_t& an) {
            constexpr int UU = decltype(UC)::value;
#pragma unroll
            for (int u = 0; u < UU; ++u) {
                const uint32_t key = keys[u];
                if (key != pk) {
                    flush_run(pk, aw, ad, an);
                    aw = 0.0;
                    ad = 0.0;
                    an = 0u;
                }
                pk = key;
                aw = fma(w1k, w2[u], aw);
                ad = fma(d1k, d2[u], ad);
                ++an;
            }
        };
        auto bin_key = [&](uint32_t bp, uint32_t bt) -> uint32_t {
            if constexpr (BANDED) {
                bp -= bd.bp0;   // bins below the band wrap to large values: out of range
                bt -= bd.bt0;
            }
            return (bp < (uint32_t)NP && bt < (uint32_t)NT) ? __umul24(bp, ntpad4 >> 2) + bt : trash;
        };
        auto accumulate = [&](uint32_t bp, uint32_t bt, double pw, double pd) {
            if constexpr (BANDED) {
                bp -= bd.bp0;
                bt -= bd.bt0;
            }
            if (bp < (uint32_t)NP && bt < (uint32_t)NT) {
                if constexpr (CNT) {
                    const uint32_t a4 = __umul24(bp, ntpad4) + opaque((bt << 2) + hbase);
                    lds_add<uint32_t>(lds, a4, 1u);
                    const uint32_t a8 = (a4 << 1) + woff;
                    lds_add<double>(lds, a8, pw);
                    lds_add<double>(lds, a8 + (L::DW - L::W), pd);
                } else {
                    const uint32_t a8 = __umul24(bp, ntpad8) + opaque((bt << 3) + hbw);
                    lds_add<double>(lds, a8, pw);
                    lds_add<double>(lds, a8 + (L::DW - L::W), pd);
                }
            }
        };
        auto accumulate_key = [&](uint32_t key, double pw, double pd) {
            if (key != NONE) {
                const uint32_t a4 = (key << 2) + hbase;
                if constexpr (CNT) lds_add<uint32_t>(lds, a4, 1u);
                const uint32_t a8 = (a4 << 1) + woff;
                lds_add<double>(lds, a8, pw);
                lds_add<double>(lds, a8 + (L::DW - L::W), pd);
            }
        };
        if constexpr (MODE == LYA_MODE_FP64) {
            const double cq = POW2 ? c * bp64 : c, sq = POW2 ? s * bt64 : s;
            // Merged one-pixel-per-lane kernels read the forest-2 distances through the
            // vector memory path (every lane the same address, into VGPRs) and the weights
            // through scalar loads: the first wait of a block then no longer waits for the
            // scalar loads and the LDS atomics in flight, which share one counter (4 %
            // faster at 100k forests; 13 more VGPRs, which the wider shapes do not have).
            // The index goes through an empty asm so that the compiler keeps it in a VGPR.
            constexpr bool VDC = MERGE && PPL == 1;
            const double* r2v = p.dc + p2;
            auto block = [&](int j, auto UC) {
                constexpr int UU = decltype(UC)::value;
                double r2[UU], w2[UU], d2[UU];
                const int jv = j;
                int jvv = jv;
                if constexpr (VDC) asm volatile("" : "+v"(jvv));
#pragma unroll
                for (int u = 0; u < UU; ++u) {
                    r2[u] = VDC ? r2v[jvv + u] : r2p[jv + u];
                    w2[u] = w2p[jv + u];
                    d2[u] = d2p[jv + u];
                }
                if constexpr (MERGE) {
#pragma unroll
                for (int k = 0; k < PPL; ++k) {
                    uint32_t keys[UU];
#pragma unroll
                    for (int u = 0; u < UU; ++u) {
                        double qp = fabs(r1[k] - r2[u]) * cq;
                        double qt = (r1[k] + r2[u]) * sq;
                        if constexpr (!POW2) {
                            qp = qp * bp64;
                            qt = qt * bt64;
                        }
                        keys[u] = bin_key((uint32_t)__double2int_rz(qp), (uint32_t)__double2int_rz(qt));
                    }
                    run_sums_carry(UC, keys, w1[k], w2, d1[k], d2, cpk[k], caw[k], cad[k], can[k]);
                }
                } else {
#pragma unroll
                for (int u = 0; u < UU; ++u) {
#pragma unroll
                    for (int k = 0; k < PPL; ++k) {
                        double qp = fabs(r1[k] - r2[u]) * cq;
                        double qt = (r1[k] + r2[u]) * sq;
                        if constexpr (!POW2) {
                            qp = qp * bp64;
                            qt = qt * bt64;
                        }
                        accumulate((uint32_t)__double2int_rz(qp), (uint32_t)__double2int_rz(qt), w1[k] * w2[u],
                                   d1[k] * d2[u]);
                    }
                }
                }
            };
            // merged runs grow with the block, so the merging kernel takes longer blocks
            constexpr int U = MERGE ? 2 * LYA_UNROLL_FP64 : LYA_UNROLL_FP64;
            int j = jlo;
            for (; j + U <= jhi; j += U) block(j, std::integral_constant<int, U>{});
            rest_blocks<U>(block, j, jhi);
            if constexpr (MERGE) {
#pragma unroll
                for (int k = 0; k < PPL; ++k) flush_run(cpk[k], caw[k], cad[k], can[k]);
            }
        } else {
            // FP32 geometry, FP64 recheck near edges, exact FP64 sums
            const double base2 = cptr(p.base)[f2];
            const double cb = c * bp64, sb = s * bt64;
            const float c4 = unif((float)cb);
            const float nc4 = -c4;
            const float s4 = unif((float)sb);
            const float tau = p.tau;
            const float tau2 = 2.0f * tau;
            // Lanes past the end of forest 1 get the coordinates 2^20 + 1/2 (out of range)
            // and zero slopes, so that they stay exactly there: their fractions are 1/2
            // and never call the recheck.  (With the slopes applied they would fall on
            // the 1/8 grid of FP32 at 2^20 and hit a zero fraction every eighth pair.)
            float P1[PPL], T1[PPL], ncl[PPL], scl[PPL];
#pragma unroll
            for (int k = 0; k < PPL; ++k) {
                P1[k] = valid[k] ? (float)((r1[k] - base2) * cb) : 1048576.5f;
                T1[k] = valid[k] ? (float)((r1[k] + base2) * sb + (double)tau) : 1048576.5f;
                ncl[k] = valid[k] ? nc4 : 0.0f;
                scl[k] = valid[k] ? s4 : 0.0f;
            }
            const CAS float* o2p = cptr(p.off32) + p2;
            // one LDS key per pair; the FP32 edge test is recomputed in the (rare)
            // recheck branch instead of keeping per-pair fractions and bins in registers
            auto block = [&](int j, auto UC) {
                constexpr int UU = decltype(UC)::value;
                uint32_t key[UU][PPL];
                float o2v[UU];
                double w2[UU], d2[UU];
                float frmin = 1.0f;
                const int jv = j;
#pragma unroll
                for (int u = 0; u < UU; ++u) {
                    o2v[u] = o2p[jv + u];
                    w2[u] = w2p[jv + u];
                    d2[u] = d2p[jv + u];
                }
                // r_par below half a bin is bin 0 in both precisions: lifting it to 1/2
                // keeps its zero edge out of the recheck (pixels on a common wavelength
                // grid meet at r_par = 0 exactly, once per lane and neighbour)
                auto qpt = [&](int u, int k, float& qp, float& qt) {
                    qp = fmaxf(__builtin_fabsf(__builtin_fmaf(o2v[u], ncl[k], P1[k])) + tau, 0.5f);
                    qt = __builtin_fmaf(o2v[u], scl[k], T1[k]);
                };
#pragma unroll
                for (int u = 0; u < UU; ++u) {
#pragma unroll
                    for (int k = 0; k < PPL; ++k) {
                        float qp, qt;
                        qpt(u, k, qp, qt);
                        key[u][k] = bin_key(__float2uint_rz(qp), __float2uint_rz(qt));
                        frmin = fminf(fminf(frmin, __builtin_amdgcn_fractf(qp)), __builtin_amdgcn_fractf(qt));
                    }
                }
                if (frmin < tau2) {
#pragma unroll
                    for (int u = 0; u < UU; ++u) {
                        // an opaque copy of the forest-2 offset: the recheck recomputes the
                        // fractions, so the main loop keeps none per pair and folds them into
                        // frmin with one v_min3 per pair
                        float ou = o2v[u];
                        asm volatile("" : "+v"(ou));
#pragma unroll
                        for (int k = 0; k < PPL; ++k) {
                            const float qp = fmaxf(__builtin_fabsf(__builtin_fmaf(ou, ncl[k], P1[k])) + tau, 0.5f);
                            const float qt = __builtin_fmaf(ou, scl[k], T1[k]);
                            if (fminf(__builtin_amdgcn_fractf(qp), __builtin_amdgcn_fractf(qt)) < tau2) {
                                const double r2 = r2p[jv + u];
                                key[u][k] = bin_key((uint32_t)__double2int_rz((fabs(r1[k] - r2) * c) * bp64),
                                                    (uint32_t)__double2int_rz(((r1[k] + r2) * s) * bt64));
                            }
                        }
                    }
                }
                if constexpr (MERGE) {
#pragma unroll
                    for (int k = 0; k < PPL; ++k) {
                        uint32_t keys[UU];
#pragma unroll
                        for (int u = 0; u < UU; ++u) keys[u] = key[u][k];
                        run_sums_carry(UC, keys, w1[k], w2, d1[k], d2, cpk[k], caw[k], cad[k], can[k]);
                    }
                } else {
#pragma unroll
                    for (int u = 0; u < UU; ++u) {
#pragma unroll
                        for (int k = 0; k < PPL; ++k) accumulate_key(key[u][k], w1[k] * w2[u], d1[k] * d2[u]);
                    }
                }
            };
            constexpr int U = LYA_UNROLL_MIXED;
            int j = jlo;
            for (; j + U <= jhi; j += U) block(j, std::integral_constant<int, U>{});
            rest_blocks<U>(block, j, jhi);
            if constexpr (MERGE) {
#pragma unroll
                for (int k = 0; k < PPL; ++k) flush_run(cpk[k], caw[k], cad[k], can[k]);
            }
        }
    } else {
        // FP32 geometry, FP32 products and FP32 LDS partials
        const double base2 = cptr(p.base)[f2];
        const double cb = c * bp64, sb = s * bt64;
        const float c4 = unif((float)cb);
        const float nc4 = -c4;
        const float s4 = unif((float)sb);
        const float tau = p.tau;
        const float tau2 = 2.0f * tau;
        // lanes past the end of forest 1: fixed out-of-range coordinates 2^20 + 1/2
        // (zero slopes keep their fractions at 1/2, clear of the recheck)
        float P1[PPL], T1[PPL], w1[PPL], d1[PPL], ncl[PPL], scl[PPL];
#pragma unroll
        for (int k = 0; k < PPL; ++k) {
            if (valid[k]) {
                P1[k] = (float)((r1[k] - base2) * cb);
                T1[k] = (float)((r1[k] + base2) * sb + (double)tau);
                w1[k] = p.we32[pix1[k]];
                d1[k] = p.dw32[pix1[k]];
                ncl[k] = nc4;
                scl[k] = s4;
            } else {
                P1[k] = 1048576.5f;
                T1[k] = 1048576.5f;
                w1[k] = 0.f;
                d1[k] = 0.f;
                ncl[k] = 0.f;
                scl[k] = 0.f;
            }
        }
        const CAS float* o2p = cptr(p.off32) + p2;
        const CAS float* w2p = cptr(p.we32) + p2;
        const CAS float* d2p = cptr(p.dw32) + p2;
        const CAS double* r2p = cptr(p.dc) + p2;
        constexpr int U = 8;
        auto block = [&](int j, auto UC) {
            constexpr int UU = decltype(UC)::value;
            uint32_t bp[UU][PPL], bt[UU][PPL];
            float fr[UU][PPL];
            float o2v[UU], w2v[UU], d2v[UU];
            float frmin = 1.0f;
            const int jv = j;
#pragma unroll
            for (int u = 0; u < UU; ++u) {
                o2v[u] = o2p[jv + u];
                w2v[u] = w2p[jv + u];
                d2v[u] = d2p[jv + u];
            }
#pragma unroll
            for (int u = 0; u < UU; ++u) {
#pragma unroll
                for (int k = 0; k < PPL; ++k) {
                    // r_par below half a bin: bin 0 either way, kept out of the recheck
                    const float qp = fmaxf(__builtin_fabsf(__builtin_fmaf(o2v[u], ncl[k], P1[k])) + tau, 0.5f);
                    const float qt = __builtin_fmaf(o2v[u], scl[k], T1[k]);
                    bp[u][k] = __float2uint_rz(qp);
                    bt[u][k] = __float2uint_rz(qt);
                    fr[u][k] = fminf(__builtin_amdgcn_fractf(qp), __builtin_amdgcn_fractf(qt));
                    frmin = fminf(frmin, fr[u][k]);
                }
            }
            if (frmin < tau2) {
#pragma unroll
                for (int u = 0; u < UU; ++u) {
#pragma unroll
                    for (int k = 0; k < PPL; ++k) {
                        if (fr[u][k] < tau2) {
                            const double r2 = r2p[jv + u];
                            bp[u][k] = (uint32_t)__double2int_rz((fabs(r1[k] - r2) * c) * bp64);
                            bt[u][k] = (uint32_t)__double2int_rz(((r1[k] + r2) * s) * bt64);
                        }
                    }
                }
            }
#pragma unroll
            for (int u = 0; u < UU; ++u) {
#pragma unroll
                for (int k = 0; k < PPL; ++k) {
                    const float pw = w1[k] * w2v[u], pd = d1[k] * d2v[u];
                    if (bp[u][k] < (uint32_t)NP && bt[u][k] < (uint32_t)NT) {
                        const uint32_t a = __umul24(bp[u][k], ntpad4) + opaque((bt[u][k] << 2) + hbase);
                        if constexpr (CNT) lds_add<uint32_t>(lds, a, 1u);
                        lds_add<float>(lds, a + L::W, pw);
                        lds_add<float>(lds, a + L::DW, pd);
                    }
                }
            }
        };
        int j = jlo;
        for (; j + U <= jhi; j += U) block(j, std::integral_constant<int, U>{});
        for (; j < jhi; ++j) block(j, std::integral_constant<int, 1>{});
    }
}

// Mixed mode: move the FP32 partial sums into the FP64 accumulators.
__device__ void flush_item(unsigned char* lds, int R, int nb_lds, int tid, int nthr) {
    using L = Layout<LYA_MODE_MIXED_F32, true>;
    double* aw = (double*)(lds + L::acc(R));
    double* ad = aw + HS;
    for (int b = tid; b < nb_lds; b += nthr) {
        double sw = 0.0, sd = 0.0;
        for (int r = 0; r < R; ++r) {
            float* hw = (float*)(lds + r * L::REPL + L::W);
            float* hd = (float*)(lds + r * L::REPL + L::DW);
            sw += (double)hw[b];
            sd += (double)hd[b];
            hw[b] = 0.f;
            hd[b] = 0.f;
        }
        aw[b] += sw;
        ad[b] += sd;
    }
}

// End of tile: merge the workgroup's sums into the slot's FP64 global sums.
// Next tile for this workgroup, or -1 when none is left (one thread calls it).  With
// tile groups (neighbouring slots, one group per XCD: consecutive workgroups sit on
// different XCDs), workgroup b serves group b % G first and then the others, so that an
// XCD's L2 mostly holds the forests of one part of the sky.  At most G atomics; the
// counters of exhausted groups only grow.
__device__ __forceinline__ int next_tile(const lya_xi_problem& p) {
    const int G = (p.tile_groups && p.n_tile_groups > 1) ? p.n_tile_groups : 1;
    if (G == 1) {
        const int t = (int)atomicAdd(p.queue, 1u);
        return t < p.n_tiles ? t : -1;
    }
    const int g0 = (int)(blockIdx.x % (unsigned)G);
    for (int k = 0; k < G; ++k) {
        const int g = g0 + k < G ? g0 + k : g0 + k - G;
        const int b = p.tile_groups[g], e = p.tile_groups[g + 1];
        if (b >= e) continue;
        const int t = b + (int)atomicAdd(p.queue + g, 1u);
        if (t < e) return t;
    }
    return -1;
}

template <int MODE, bool CNT>
__device__ void flush_tile(const lya_xi_problem& p, unsigned char* lds, int R, int ntpad, int slot, int tid,
                           int nthr, const Band& bd) {
    using L = Layout<MODE, CNT>;
    const int nb_lds = (int)bd.rows * ntpad;
    // global bin of LDS entry (bp, bt): (bd.bp0 + bp) np_rt + bd.bt0 + bt
    const int64_t gofs = (int64_t)slot * p.np_rp * p.np_rt + (int64_t)bd.bp0 * p.np_rt + bd.bt0;
    for (int b = tid; b < nb_lds; b += nthr) {
        uint32_t cnt = 0;
        double sw = 0.0, sd = 0.0;
        for (int r = 0; r < R; ++r) {
            if constexpr (L::HAS_CNT) {
                uint32_t* hc = (uint32_t*)(lds + r * L::REPL);
                cnt += hc[b];
                hc[b] = 0u;
            }
            if constexpr (!L::F32) {
                double* hw = (double*)(lds + r * L::REPL + L::W);
                double* hd = (double*)(lds + r * L::REPL + L::DW);
                sw += hw[b];
                sd += hd[b];
                hw[b] = 0.0;
                hd[b] = 0.0;
            }
        }
        if constexpr (L::F32) {
            double* aw = (double*)(lds + L::acc(R));
            double* ad = aw + HS;
            sw = aw[b];
            sd = ad[b];
            aw[b] = 0.0;
            ad[b] = 0.0;
        }
        const int bp = b / ntpad, bt = b - bp * ntpad;
        // columns from bd.cols on (padding, trash entries) never hold a pair
        if (bt < (int)bd.cols && (CNT ? (cnt != 0u) : (sw != 0.0 || sd != 0.0))) {
            const int64_t g = gofs + (int64_t)bp * p.np_rt + bt;
            if (CNT) atomicAdd((unsigned long long*)(p.out_cnt + g), (unsigned long long)cnt);
            unsafeAtomicAdd(p.out_w + g, sw);
            unsafeAtomicAdd(p.out_dw + g, sd);
        }
    }
}

// Shapes that can hold 8 waves per SIMD (32 per CU) ask the compiler for that
// occupancy: their VGPRs stay within 64 (all FP64 shapes, one pixel per lane in the
// FP32-geometry modes) and the LDS of the 32/WAVES workgroups it takes fits the CU's
// 160 KiB.  The compiler then keeps the SGPRs within that occupancy, spilling the rest
// to VGPR lanes; otherwise it uses all 106, which allows 7 waves per SIMD, i.e. three
// workgroups of 8 waves.
template <int MODE, int PPL, bool CNT, int R, int WAVES>
constexpr int tile_waves_per_eu() {
    const bool vgpr_fit = MODE == LYA_MODE_FP64 || PPL == 1;
    const bool lds_fit = WAVES >= 8 && (Layout<MODE, CNT>::bytes(R) + 4u) * (32u / WAVES) <= 160u * 1024u;
    return vgpr_fit && lds_fit ? 8 : 1;
}

template <int MODE, int PPL, int WAVES, bool CNT, bool POW2, int R, bool MERGE, bool BANDED>
__global__ void __launch_bounds__(WAVES * 64)
    __attribute__((amdgpu_waves_per_eu(tile_waves_per_eu<MODE, PPL, CNT, R, WAVES>()))) xi_kernel(const lya_xi_problem p, int ntpad,
                                                                                              const BandShape bs, int split) {
    using L = Layout<MODE, CNT>;
    // statically sized, so histogram addresses fold into the ds offset field
    __shared__ __align__(16) unsigned char lds[L::bytes(R)];
    __shared__ int s_tile;
    constexpr int NTHR = WAVES * 64;
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = uni(tid >> 6);
    // lane -> forest-1 pixel within the 64-pixel row: 0 consecutive, 1 interleaved by 2
    // (half-waves hold even / odd pixels), 2 interleaved by 4 (16-lane groups)
    const int lmap = p.lane_map;
    const int perm = lmap == 0 ? lane : (lmap == 1 ? (((lane & 31) << 1) | (lane >> 5))
                                                   : (((lane & 15) << 2) | (lane >> 4)));
    const double bp64 = (double)p.np_rp / p.rpmax;
    const double bt64 = (double)p.np_rt / p.rtmax;
    const uint32_t ntpad4 = (uint32_t)ntpad * 4u;
    const uint32_t hbase = (uint32_t)(wave % R) * L::REPL;
    const int nb_lds = p.np_rp * ntpad;

    const uint32_t nbytes = L::bytes(R);
    for (uint32_t o = (uint32_t)tid * 16u; o < nbytes; o += NTHR * 16u)
        *(uint4*)(lds + o) = make_uint4(0u, 0u, 0u, 0u);
    __syncthreads();

    for (;;) {
        if (tid == 0) s_tile = next_tile(p);
        __syncthreads();
        const int tile = uni(s_tile);
        if (tile < 0) break;
        const int ib = cptr(p.tiles)[2 * tile], ie = cptr(p.tiles)[2 * tile + 1];
        const int slot = cptr(p.items)[4 * ib + 3];
        // the tile is walked once per band of the binning (once when it fits the LDS)
        const int nbands = BANDED ? bs.nbp * bs.nbt : 1;
        for (int band = 0; band < nbands; ++band) {
            Band bd{0u, 0u, (uint32_t)p.np_rp, (uint32_t)p.np_rt};
            if constexpr (BANDED) {
                const int bpi = band / bs.nbt, bti = band - bpi * bs.nbt;
                bd.bp0 = (uint32_t)(bpi * bs.br);
                bd.bt0 = (uint32_t)(bti * bs.bc);
                bd.rows = (uint32_t)min(bs.br, p.np_rp - (int)bd.bp0);
                bd.cols = (uint32_t)min(bs.bc, p.np_rt - (int)bd.bt0);
            }
            for (int it = ib; it < ie; ++it) {
                TaskCtx t;
                t.f1 = cptr(p.items)[4 * it + 0];
                const int nb0 = cptr(p.items)[4 * it + 1];
                const int nb1 = cptr(p.items)[4 * it + 2];
                t.slot = slot;
                t.p1 = cptr(p.pix_off)[t.f1];
                t.n1 = (int)(cptr(p.pix_off)[t.f1 + 1] - t.p1);
                // one neighbour per wave: the pair's geometry and units once, then the units
                constexpr int SL = 64 * PPL;
                for (int q = nb0 + wave; q < nb1; q += WAVES) {
                    const int f2 = cptr(p.nb_idx)[q];
                    const int64_t p2 = cptr(p.pix_off)[f2];
                    const int n2 = (int)(cptr(p.pix_off)[f2 + 1] - p2);
                    const PairGeo g = pair_geo<BANDED>(p, t.f1, f2, bp64, bt64, bd);
                    Units un;
                    int ja, jb;
                    if (!pair_units<BANDED>(p, t, f2, p2, n2, g, SL, split, lane, un, ja, jb)) continue;
                    for (int sub = 0; sub < un.upq; ++sub) {
                        UnitArgs a = unit_args(t, un, SL, sub, f2, p2, n2, ja, jb);
                        // one call site per slab width (the inlined task body is large)
                        for (; a.i0 < a.tl.n1; a.i0 += 64) {
                            if (PPL > 1 && a.wide) {
                                run_task<MODE, PPL, CNT, POW2, MERGE, BANDED>(p, lds, hbase, ntpad4, a.tl, a.i0, a.fu,
                                                                              a.u0, a.u1, g, lane, perm, bp64, bt64,
                                                                              bd);
                            } else {
                                run_task<MODE, 1, CNT, POW2, MERGE, BANDED>(p, lds, hbase, ntpad4, a.tl, a.i0, a.fu,
                                                                            a.u0, a.u1, g, lane, perm, bp64, bt64, bd);
                            }
                            if (!a.tail) break;
                        }
                    }
                }
                if constexpr (MODE == LYA_MODE_MIXED_F32) {
                    // every wave done with the item; then the partials move to FP64 and are
                    // zeroed before any wave starts the next item
                    __syncthreads();
                    flush_item(lds, R, nb_lds, tid, NTHR);
                    __syncthreads();
                }
            }
            __syncthreads();
            flush_tile<MODE, CNT>(p, lds, R, ntpad, slot, tid, NTHR, bd);
            __syncthreads();
        }
    }
}

// ---------------------------------------------------------------------------
// Diagonal walk (FP64).  The in-range pixel pairs (i, j) of a forest pair lie in
// a band around the diagonals d = j - i.  Along one diagonal r1 and r2 grow
// together: |r1 - r2| (the r_parallel bin) stays nearly constant and r1 + r2
// (the r_perp bin) grows slowly, so consecutive pairs mostly share a bin (runs
// of ~30 pairs at 4 Mpc/h bins, ~100 in-band diagonals per forest pair).  Lanes
// walk diagonals with a wave-uniform forest-1 index i, keep the (w, dw, count)
// sums of the current bin in registers and add them to the LDS histogram only
// when the bin changes -- one LDS atomic per run instead of one per pair.
//
// Per item (forest 1 and a range of its neighbours), in batches of neighbours:
//   1. one wave per neighbour: half-angles and the diagonal range [dlo, dhi)
//      that can hold an in-range pair (per-pixel binary searches with the
//      1e-3-bin margins of prune(), so skipped diagonals are out of range);
//   2. sub-groups of 16 consecutive diagonals of one pair (their forest-2
//      loads are coalesced); the span of a sub-group is the union of its
//      diagonals' overlaps, each cut at the first step whose r1 + r2 exceeds
//      the r_perp range (r1 + r2 grows along a diagonal of sorted forests);
//   3. sub-groups sorted by (first step, last step) in LDS (bitonic);
//   4. a wave walks four consecutive sorted sub-groups over the union of their
//      spans; a lane outside its own diagonal's overlap reads forest-2 pixel 0
//      and bins nothing.
// ---------------------------------------------------------------------------
#ifndef LYA_WALK_UNROLL
#define LYA_WALK_UNROLL 8
#endif
constexpr int WALK_NB = 64;        // neighbours per batch
constexpr int WALK_HS = 2688;      // LDS bins: (np_rp + 1) x (np_rt + 1), last row and column = trash
constexpr int WALK_SG = 2048;      // sub-groups per batch
constexpr int WALK_G = 16;         // diagonals per sub-group
constexpr int WALK_MAX_LEN = 16000;  // longest forest: (2*len)/16 + 1 sub-groups fit a batch
constexpr uint32_t WALK_NONE = 0xFFFFFFFFu;

__device__ __forceinline__ int wave_min(int x) {
    for (int o = 32; o > 0; o >>= 1) x = min(x, __shfl_xor(x, o, 64));
    return x;
}
__device__ __forceinline__ int wave_max(int x) {
    for (int o = 32; o > 0; o >>= 1) x = max(x, __shfl_xor(x, o, 64));
    return x;
}

// Walk histogram in LDS: [cnt u32 | w f64 | dw f64] x WALK_HS, rows of ntpad =
// np_rt + 1 entries; row np_rp and column np_rt take the runs of out-of-range
// pairs and are never merged into the outputs.
struct WalkLayout {
    static constexpr uint32_t W = 4u * WALK_HS;
    static constexpr uint32_t DW = 12u * WALK_HS;
    static constexpr uint32_t BYTES = 20u * WALK_HS;
};

template <bool CNT>
__device__ void flush_walk(const lya_xi_problem& p, unsigned char* lds, int ntpad, int slot, int tid, int nthr) {
    using L = WalkLayout;
    const int NT = p.np_rt;
    const int nb_lds = p.np_rp * ntpad;
    const int64_t gofs = (int64_t)slot * p.np_rp * NT;
    for (int b = tid; b < nb_lds; b += nthr) {
        uint32_t* hc = (uint32_t*)lds;
        double* hw = (double*)(lds + L::W);
        double* hd = (double*)(lds + L::DW);
        const uint32_t cnt = hc[b];
        const double sw = hw[b], sd = hd[b];
        hc[b] = 0u;
        hw[b] = 0.0;
        hd[b] = 0.0;
        const int bp = b / ntpad, bt = b - bp * ntpad;
        if (bt < NT && (CNT ? (cnt != 0u) : (sw != 0.0 || sd != 0.0))) {
            const int64_t g = gofs + (int64_t)bp * NT + bt;
            if (CNT) atomicAdd((unsigned long long*)(p.out_cnt + g), (unsigned long long)cnt);
            unsafeAtomicAdd(p.out_w + g, sw);
            unsafeAtomicAdd(p.out_dw + g, sd);
        }
    }
}

template <int WAVES, bool CNT, bool POW2>
__global__ void __launch_bounds__(WAVES * 64) xi_walk_kernel(const lya_xi_problem p, int ntpad) {
    using L = WalkLayout;
    __shared__ __align__(16) unsigned char lds[L::BYTES];
    __shared__ uint64_t s_sg[WALK_SG];
    __shared__ int64_t s_p2[WALK_NB];
    __shared__ double s_cq[WALK_NB], s_sq[WALK_NB], s_rt[WALK_NB], s_dp[WALK_NB];
    __shared__ int s_n2[WALK_NB], s_dlo[WALK_NB], s_dhi[WALK_NB], s_sgoff[WALK_NB], s_srt[WALK_NB];
    __shared__ int s_tile, s_nq, s_total, s_next;
    constexpr int NTHR = WAVES * 64;
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = uni(tid >> 6);
    const int NP = p.np_rp, NT = p.np_rt;
    const double bp64 = (double)NP / p.rpmax;
    const double bt64 = (double)NT / p.rtmax;

    const uint32_t TRASH = (uint32_t)NP * (uint32_t)ntpad + (uint32_t)NT;

    for (uint32_t o = (uint32_t)tid * 16u; o < L::BYTES; o += NTHR * 16u)
        *(uint4*)(lds + o) = make_uint4(0u, 0u, 0u, 0u);
    __syncthreads();

    for (;;) {
        if (tid == 0) s_tile = next_tile(p);
        __syncthreads();
        const int tile = uni(s_tile);
        if (tile < 0) break;
        const int ib = cptr(p.tiles)[2 * tile], ie = cptr(p.tiles)[2 * tile + 1];
        const int slot = cptr(p.items)[4 * ib + 3];
        for (int it = ib; it < ie; ++it) {
            const int f1 = cptr(p.items)[4 * it + 0];
            const int nb0 = cptr(p.items)[4 * it + 1];
            const int nb1 = cptr(p.items)[4 * it + 2];
            const int64_t p1 = cptr(p.pix_off)[f1];
            const int n1 = (int)(cptr(p.pix_off)[f1 + 1] - p1);
            const bool sorted1 = cptr(p.sorted)[f1] != 0;
            for (int nbd = nb0; nbd < nb1;) {
                const int batch = min(WALK_NB, nb1 - nbd);
                // 1. half-angles of each neighbour (one thread each), then the diagonal
                //    range [dlo, dhi) from 8-pixel slabs of forest 1 (all threads).  For
                //    pixels i in a slab [i0, i1] every j with an in-range pair satisfies
                //    jl(i0) <= j < min(jhp(i1), jrt(i0)), where jl counts r2 <= r1 - Dp,
                //    jhp counts r2 < r1 + Dp and jrt counts r2 < Rt - r1 (jl and jhp grow
                //    with i, jrt shrinks), so d = j - i lies in
                //    [jl(i0) - i1, min(jhp(i1), jrt(i0)) - i0).  Dp and Rt carry the
                //    1e-3-bin margins of prune(): skipped diagonals are out of range.
                for (int q = tid; q < batch; q += NTHR) {
                    const int f2 = p.nb_idx[nbd + q];
                    const int64_t p2 = p.pix_off[f2];
                    const int n2 = (int)(p.pix_off[f2 + 1] - p2);
                    double c, s;
                    halfangles(p.ux[f1], p.uy[f1], p.uz[f1], p.ux[f2], p.uy[f2], p.uz[f2], c, s);
                    const bool both_sorted = sorted1 && p.sorted[f2] != 0;
                    s_p2[q] = p2;
                    s_n2[q] = n2;
                    s_cq[q] = POW2 ? c * bp64 : c;
                    s_sq[q] = POW2 ? s * bt64 : s;
                    s_dp[q] = ((double)NP + 1e-3) / (c * bp64);
                    s_rt[q] = ((double)NT + 1e-3) / (s * bt64);
                    s_srt[q] = both_sorted;
                    s_dlo[q] = both_sorted ? INT_MAX : -(n1 - 1);
                    s_dhi[q] = both_sorted ? INT_MIN : n2;
                }
                __syncthreads();
                {
                    const int nsl = (n1 + 7) >> 3;
                    for (int u = tid; u < batch * nsl; u += NTHR) {
                        const int q = u / nsl, sl = u - q * nsl;
                        if (!s_srt[q]) continue;
                        const int i0 = sl << 3, i1 = min(i0 + 7, n1 - 1);
                        const int n2 = s_n2[q];
                        const double* r2p = p.dc + s_p2[q];
                        const double ra = p.dc[p1 + i0], rb = p.dc[p1 + i1];
                        const double lo = ra - s_dp[q], hp = rb + s_dp[q], hr = s_rt[q] - ra;
                        int a0 = 0, b0 = n2, a1 = 0, b1 = n2, a2 = 0, b2 = n2;
                        while (a0 < b0 || a1 < b1 || a2 < b2) {
                            if (a0 < b0) {
                                const int m = (a0 + b0) >> 1;
                                if (r2p[m] <= lo) a0 = m + 1; else b0 = m;
                            }
                            if (a1 < b1) {
                                const int m = (a1 + b1) >> 1;
                                if (r2p[m] < hp) a1 = m + 1; else b1 = m;
                            }
                            if (a2 < b2) {
                                const int m = (a2 + b2) >> 1;
                                if (r2p[m] < hr) a2 = m + 1; else b2 = m;
                            }
                        }
                        const int jl = a0, jh = min(a1, a2);
                        if (jh > jl) {
                            atomicMin(&s_dlo[q], jl - i1);
                            atomicMax(&s_dhi[q], jh - i0);
                        }
                    }
                }
                __syncthreads();
                // neighbours whose sub-groups fit the batch (prefix), and their offsets
                if (wave == 0) {
                    const int cnt = (lane < batch && s_dhi[lane] > s_dlo[lane])
                                        ? (s_dhi[lane] - s_dlo[lane] + WALK_G - 1) / WALK_G : 0;
                    int inc = cnt;
                    for (int o = 1; o < 64; o <<= 1) {
                        const int v = __shfl_up(inc, o, 64);
                        if (lane >= o) inc += v;
                    }
                    const bool fits = lane < batch && inc <= WALK_SG;
                    const int nq = max(1, __popcll(__ballot(fits)));
                    if (lane < batch) s_sgoff[lane] = inc - cnt;
                    const int total = __shfl(inc, nq - 1, 64);
                    if (lane == 0) {
                        s_nq = nq;
                        s_total = min(total, WALK_SG);
                    }
                }
                __syncthreads();
                const int nq = s_nq, total = s_total;
                // 2. sub-group spans: four sub-groups per wave, one lane per diagonal
                for (int g0 = wave * 4; g0 < total; g0 += WAVES * 4) {
                    const int g = g0 + (lane >> 4), k = lane & 15;
                    int q = 0;
                    for (int b = 32; b > 0; b >>= 1)  // last neighbour with sgoff <= g
                        if (q + b < nq && s_sgoff[q + b] <= g) q += b;
                    const int d0 = s_dlo[q] + (g - s_sgoff[q]) * WALK_G;
                    int tlo = INT_MAX, thi = 0;
                    if (g < total) {
                        const int d = d0 + k;
                        const int n2 = s_n2[q];
                        if (d < s_dhi[q]) {
                            const int a = max(0, -d);
                            int b = min(n1, n2 - d);
                            if (b > a && s_srt[q]) {
                                const double Rt = s_rt[q];
                                const int64_t p2d = s_p2[q] + d;
                                int lo = a, hi = b;  // first step with r1 + r2 >= Rt
                                while (lo < hi) {
                                    const int m = (lo + hi) >> 1;
                                    if (p.dc[p1 + m] + p.dc[p2d + m] >= Rt) hi = m; else lo = m + 1;
                                }
                                b = lo;
                            }
                            if (b > a) {
                                tlo = a;
                                thi = b;
                            }
                        }
                    }
                    for (int o = 8; o > 0; o >>= 1) {
                        tlo = min(tlo, __shfl_xor(tlo, o, 16));
                        thi = max(thi, __shfl_xor(thi, o, 16));
                    }
                    if (k == 0 && g < total) {
                        const uint64_t t0 = thi > tlo ? (uint64_t)tlo : 0ull, t1 = thi > tlo ? (uint64_t)thi : 0ull;
                        s_sg[g] = (t0 << 48) | (t1 << 32) | ((uint64_t)q << 16) | (uint64_t)(uint32_t)(d0 + 32768);
                    }
                }
                __syncthreads();
                // 3. sort the sub-groups by (first step, last step)
                int pw = 1;
                while (pw < total) pw <<= 1;
                for (int t = total + tid; t < pw; t += NTHR) s_sg[t] = ~0ull;
                __syncthreads();
                for (int kk = 2; kk <= pw; kk <<= 1)
                    for (int jj = kk >> 1; jj > 0; jj >>= 1) {
                        for (int t = tid; t < pw; t += NTHR) {
                            const int u = t ^ jj;
                            if (u > t) {
                                const uint64_t x = s_sg[t], y = s_sg[u];
                                if ((x > y) == ((t & kk) == 0)) {
                                    s_sg[t] = y;
                                    s_sg[u] = x;
                                }
                            }
                        }
                        __syncthreads();
                    }
                if (tid == 0) s_next = 0;
                __syncthreads();
                // 4. walk
                for (;;) {
                    int gw = 0;
                    if (lane == 0) gw = atomicAdd(&s_next, 1);
                    const int gb = 4 * uni(gw);
                    if (gb >= total) break;
                    const int h = lane >> 4, k = lane & 15;
                    const uint64_t e = (gb + h < total) ? s_sg[gb + h] : 0ull;
                    const int tmin = (int)(e >> 48), tmax = (int)((e >> 32) & 0xFFFFu);
                    const int q = (int)((e >> 16) & 0xFFu), d = (int)(e & 0xFFFFu) - 32768 + k;
                    const bool ne = tmax > tmin;
                    const int x0 = ne ? tmin : INT_MAX, x1 = ne ? tmax : 0;
                    const int T0 = min(min(__builtin_amdgcn_readlane(x0, 0), __builtin_amdgcn_readlane(x0, 16)),
                                       min(__builtin_amdgcn_readlane(x0, 32), __builtin_amdgcn_readlane(x0, 48)));
                    const int T1 = max(max(__builtin_amdgcn_readlane(x1, 0), __builtin_amdgcn_readlane(x1, 16)),
                                       max(__builtin_amdgcn_readlane(x1, 32), __builtin_amdgcn_readlane(x1, 48)));
                    if (T1 <= T0) continue;
                    const bool valid = ne && d < s_dhi[q];
                    const int n2 = s_n2[q];
                    const int a = valid ? max(0, -d) : 0;
                    const int b = valid ? min(n1, n2 - d) : 0;
                    // forest-2 pixel of step t: jb + t.  Every step of the wave lies in
                    // [0, n1) and -n1 < d < n2, so jb + t stays within [-pad, n_pix + pad):
                    // the byte offset is fixed per lane and the step moves the uniform base.
                    // Lanes without a diagonal read the padding, whose dc sentinel puts
                    // every pair in a trash bin.
                    const int64_t jb = valid ? s_p2[q] + d : -(int64_t)p.pad;
                    const int64_t voff = (jb + p.pad) * 8;
                    const double cq = s_cq[q], sq = s_sq[q];
                    // steps where every lane with a diagonal is inside its overlap
                    const int A = wave_max(valid ? a : INT_MIN), B = wave_min(valid ? b : INT_MAX);
                    const int M0 = max(T0, min(A, T1)), M1 = max(M0, min(T1, B));
                    uint32_t prev = TRASH, an = 0u;
                    double aw = 0.0, ad = 0.0;
                    const char* dcb = (const char*)(p.dc - p.pad);
                    const char* web = (const char*)(p.we - p.pad);
                    const char* dwb = (const char*)(p.dw - p.pad);
                    auto step = [&](auto CHK, int t, double r1, double w1, double d1, double r2, double w2,
                                    double d2) {
                        double qp = fabs(r1 - r2) * cq;
                        double qt = (r1 + r2) * sq;
                        if constexpr (!POW2) {
                            qp = qp * bp64;
                            qt = qt * bt64;
                        }
                        // saturating conversions: out-of-range rows and columns clamp to the trash
                        const uint32_t kp = min((uint32_t)__double2int_rz(qp), (uint32_t)NP);
                        const uint32_t kt = min((uint32_t)__double2int_rz(qt), (uint32_t)NT);
                        uint32_t key = __umul24(kp, (uint32_t)ntpad) + kt;
                        if constexpr (decltype(CHK)::value) key = ((t >= a) & (t < b)) ? key : TRASH;
                        if (key != prev) {  // the run of `prev` ended
                            if constexpr (CNT) lds_add<uint32_t>(lds, 4u * prev, an);
                            lds_add<double>(lds, L::W + 8u * prev, aw);
                            lds_add<double>(lds, L::DW + 8u * prev, ad);
                            aw = 0.0;
                            ad = 0.0;
                            an = 0u;
                        }
                        aw = fma(w1, w2, aw);   // run sums, fused multiply-add
                        ad = fma(d1, d2, ad);
                        if constexpr (CNT) ++an;
                        prev = key;
                    };
                    const CAS double* r1p = cptr(p.dc) + p1;
                    const CAS double* w1p = cptr(p.we) + p1;
                    const CAS double* d1p = cptr(p.dw) + p1;
                    // a block's loads are all issued before its first step, so one memory
                    // latency is exposed per block rather than per step
                    auto block = [&](auto CHK, int t, auto UC) {
                        constexpr int UU = decltype(UC)::value;
                        double r2[UU], w2[UU], d2[UU], r1[UU], w1[UU], d1[UU];
                        const int64_t tb = (int64_t)t * 8 + voff;
#pragma unroll
                        for (int u = 0; u < UU; ++u) {
                            r2[u] = *(const double*)(dcb + tb + 8 * u);
                            w2[u] = *(const double*)(web + tb + 8 * u);
                            d2[u] = *(const double*)(dwb + tb + 8 * u);
                        }
#pragma unroll
                        for (int u = 0; u < UU; ++u) {
                            r1[u] = r1p[t + u];
                            w1[u] = w1p[t + u];
                            d1[u] = d1p[t + u];
                        }
#pragma unroll
                        for (int u = 0; u < UU; ++u) step(CHK, t + u, r1[u], w1[u], d1[u], r2[u], w2[u], d2[u]);
                    };
                    constexpr int U = LYA_WALK_UNROLL;
                    using Blk = std::integral_constant<int, U>;
                    using One = std::integral_constant<int, 1>;
                    int t = T0;
                    for (; t + U <= T1; t += U) {
                        // only blocks reaching outside [M0, M1) need the per-lane overlap test
                        if (t < M0 || t + U > M1)
                            block(std::true_type{}, t, Blk{});
                        else
                            block(std::false_type{}, t, Blk{});
                    }
                    for (; t < T1; ++t) block(std::true_type{}, t, One{});
                    if constexpr (CNT) lds_add<uint32_t>(lds, 4u * prev, an);
                    lds_add<double>(lds, L::W + 8u * prev, aw);
                    lds_add<double>(lds, L::DW + 8u * prev, ad);
                }
                __syncthreads();  // tables and sub-groups are rewritten by the next batch
                nbd += nq;
            }
        }
        __syncthreads();
        flush_walk<CNT>(p, lds, ntpad, slot, tid, NTHR);
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// Adjoint: g_dw[i] += alpha[A] dw[j], g_w[i] += beta[A] w[j] over the pairs of
// the list, lanes = forest-1 pixels (no atomics in the pair loop).
//   LDS holds (alpha, beta) interleaved, one 16-byte read per pair, plus a zero
//   entry at index nb that out-of-range pairs read (adding 0 * w leaves the sums
//   unchanged, so no select of the operands is needed).
//   POW2: FP64 mode with power-of-two binners folded into c and s (exact, as in
//   the forward kernel).
//   At 56-62 VGPRs and (nb + 1) * 16 bytes of LDS it holds 8 waves per SIMD (four
//   8-wave workgroups per CU at 50 x 50 bins); the compiler keeps the SGPRs within
//   that occupancy.
//   GTAB: binnings whose table exceeds a workgroup's LDS read alpha and beta from
//   global memory (L1/L2-cached), with out-of-range pairs selecting zero.
// ---------------------------------------------------------------------------
template <int MODE, int WAVES, int PPL, bool POW2, bool GTAB>
__global__ void __launch_bounds__(WAVES * 64) __attribute__((amdgpu_waves_per_eu(8))) xi_adjoint_kernel(const lya_xi_problem p, const double* alpha,
                                                                const double* beta, double* g_dw, double* g_w) {
    extern __shared__ __align__(16) unsigned char lds[];
    __shared__ int s_tile;
    constexpr int NTHR = WAVES * 64;
    constexpr int SL = 64 * PPL;   // forest-1 pixels of a task: lane l holds l, l + 64, ...
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = uni(tid >> 6);
    const int NP = p.np_rp, NT = p.np_rt;
    const int nb = NP * NT;
    double2* tab = (double2*)lds;   // [nb + 1] (alpha, beta); entry nb is zero
    if constexpr (!GTAB) {
        for (int b = tid; b <= nb; b += NTHR)
            tab[b] = b < nb ? make_double2(alpha[b], beta[b]) : make_double2(0.0, 0.0);
        __syncthreads();
    }
    const double bp64 = (double)NP / p.rpmax;
    const double bt64 = (double)NT / p.rtmax;

    for (;;) {
        if (tid == 0) s_tile = next_tile(p);
        __syncthreads();
        const int tile = uni(s_tile);
        __syncthreads();
        if (tile < 0) break;
        const int ib = cptr(p.tiles)[2 * tile], ie = cptr(p.tiles)[2 * tile + 1];
        for (int it = ib; it < ie; ++it) {
            const int f1 = cptr(p.items)[4 * it + 0];
            const int nb0 = cptr(p.items)[4 * it + 1];
            const int nb1 = cptr(p.items)[4 * it + 2];
            const int64_t p1 = cptr(p.pix_off)[f1];
            const int n1 = (int)(cptr(p.pix_off)[f1 + 1] - p1);
            // one neighbour per wave: the pair's geometry once, forest 1 trimmed to the
            // pixels [ia, ib) that can pair with it (as in the forward kernel), then its slabs
            for (int q = nb0 + wave; q < nb1; q += WAVES) {
                const int f2 = cptr(p.nb_idx)[q];
                const int64_t p2 = cptr(p.pix_off)[f2];
                const int n2 = (int)(cptr(p.pix_off)[f2 + 1] - p2);
                if (n1 <= 0 || n2 <= 0) continue;
                const PairGeo g = pair_geo(p, f1, f2, bp64, bt64);
                const double c = g.c, s = g.s;
                int ia, ib;
                prune(p, f2, p2, 0, n2 - 1, f1, p1, 0, n1, g, lane, ia, ib);
                for (int i0 = ia; i0 < ib; i0 += SL) {
                    const int ilast = min(ib, i0 + SL) - 1;
                    int jlo, jhi;
                    prune(p, f1, p1, i0, ilast, f2, p2, 0, n2, g, lane, jlo, jhi);
                    if (jlo >= jhi) continue;
                    int i[PPL];
                    bool valid[PPL];
                    double r1[PPL], gd[PPL], gw[PPL];
#pragma unroll
                    for (int m = 0; m < PPL; ++m) {
                        i[m] = i0 + m * 64 + lane;
                        valid[m] = i[m] < ib;
                        r1[m] = valid[m] ? p.dc[p1 + i[m]] : 1.0e9;   // far away: no pair in range
                        gd[m] = 0.0;
                        gw[m] = 0.0;
                    }
                    // in-range pairs gather alpha, beta of their bin; others read bin 0 times zero
                    auto gather = [&](int m, uint32_t bp, uint32_t bt, double dwj, double wj) {
                        const bool in = bp < (uint32_t)NP && bt < (uint32_t)NT;
                        double2 ab;
                        if constexpr (GTAB) {
                            const uint32_t A = in ? __umul24(bp, (uint32_t)NT) + bt : 0u;
                            ab = make_double2(in ? alpha[A] : 0.0, in ? beta[A] : 0.0);
                        } else {
                            ab = tab[in ? __umul24(bp, (uint32_t)NT) + bt : (uint32_t)nb];
                        }
                        gd[m] = fma(ab.x, dwj, gd[m]);
                        gw[m] = fma(ab.y, wj, gw[m]);
                    };
                    // forest-2 pixels in blocks of U wave-uniform scalar loads, as in the forward walk
                    constexpr int U = 8;
                    const CAS double* d2p = cptr(p.dw) + p2;
                    const CAS double* w2p = cptr(p.we) + p2;
                    if constexpr (MODE == LYA_MODE_FP64) {
                        const CAS double* r2p = cptr(p.dc) + p2;
                        const double cq = POW2 ? c * bp64 : c, sq = POW2 ? s * bt64 : s;
                        auto block = [&](int j, auto UC) {
                            constexpr int UU = decltype(UC)::value;
                            double r2[UU], d2[UU], w2[UU];
#pragma unroll
                            for (int u = 0; u < UU; ++u) {
                                r2[u] = r2p[j + u];
                                d2[u] = d2p[j + u];
                                w2[u] = w2p[j + u];
                            }
#pragma unroll
                            for (int u = 0; u < UU; ++u) {
#pragma unroll
                                for (int m = 0; m < PPL; ++m) {
                                    const double qp = POW2 ? fabs(r1[m] - r2[u]) * cq : (fabs(r1[m] - r2[u]) * c) * bp64;
                                    const double qt = POW2 ? (r1[m] + r2[u]) * sq : ((r1[m] + r2[u]) * s) * bt64;
                                    gather(m, (uint32_t)__double2int_rz(qp), (uint32_t)__double2int_rz(qt), d2[u], w2[u]);
                                }
                            }
                        };
                        int j = jlo;
                        for (; j + U <= jhi; j += U) block(j, std::integral_constant<int, U>{});
                        rest_blocks<U>(block, j, jhi);
                    } else {
                        const double base2 = cptr(p.base)[f2];
                        const double cb = c * bp64, sb = s * bt64;
                        const float c4 = unif((float)cb), s4 = unif((float)sb);
                        const float tau = p.tau, tau2 = 2.0f * tau;
                        // lanes past the end of forest 1: 2^20 + 1/2 with zero slopes (fraction
                        // 1/2, out of range, never rechecked)
                        float P1[PPL], T1[PPL], ncl[PPL], scl[PPL];
#pragma unroll
                        for (int m = 0; m < PPL; ++m) {
                            P1[m] = valid[m] ? (float)((r1[m] - base2) * cb) : 1048576.5f;
                            T1[m] = valid[m] ? (float)((r1[m] + base2) * sb + (double)tau) : 1048576.5f;
                            ncl[m] = valid[m] ? -c4 : 0.0f;
                            scl[m] = valid[m] ? s4 : 0.0f;
                        }
                        const CAS float* o2p = cptr(p.off32) + p2;
                        const CAS double* r2p = cptr(p.dc) + p2;
                        auto block = [&](int j, auto UC) {
                            constexpr int UU = decltype(UC)::value;
                            float o2[UU];
                            double d2[UU], w2[UU];
                            uint32_t bp[UU][PPL], bt[UU][PPL];
                            float frmin = 1.0f;
#pragma unroll
                            for (int u = 0; u < UU; ++u) {
                                o2[u] = o2p[j + u];
                                d2[u] = d2p[j + u];
                                w2[u] = w2p[j + u];
                            }
                            auto qpt = [&](int u, int m, float& qp, float& qt) {
                                // r_par below half a bin: bin 0 either way, kept out of the recheck
                                qp = fmaxf(__builtin_fabsf(__builtin_fmaf(o2[u], ncl[m], P1[m])) + tau, 0.5f);
                                qt = __builtin_fmaf(o2[u], scl[m], T1[m]);
                            };
#pragma unroll
                            for (int u = 0; u < UU; ++u) {
#pragma unroll
                                for (int m = 0; m < PPL; ++m) {
                                    float qp, qt;
                                    qpt(u, m, qp, qt);
                                    bp[u][m] = __float2uint_rz(qp);
                                    bt[u][m] = __float2uint_rz(qt);
                                    frmin = fminf(frmin, fminf(__builtin_amdgcn_fractf(qp), __builtin_amdgcn_fractf(qt)));
                                }
                            }
                            // the edge test is recomputed here rather than kept per pair
                            if (frmin < tau2) {
#pragma unroll
                                for (int u = 0; u < UU; ++u) {
#pragma unroll
                                    for (int m = 0; m < PPL; ++m) {
                                        float qp, qt;
                                        qpt(u, m, qp, qt);
                                        if (fminf(__builtin_amdgcn_fractf(qp), __builtin_amdgcn_fractf(qt)) < tau2) {
                                            const double r2 = r2p[j + u];
                                            bp[u][m] = (uint32_t)__double2int_rz((fabs(r1[m] - r2) * c) * bp64);
                                            bt[u][m] = (uint32_t)__double2int_rz(((r1[m] + r2) * s) * bt64);
                                        }
                                    }
                                }
                            }
#pragma unroll
                            for (int u = 0; u < UU; ++u)
#pragma unroll
                                for (int m = 0; m < PPL; ++m) gather(m, bp[u][m], bt[u][m], d2[u], w2[u]);
                        };
                        int j = jlo;
                        for (; j + U <= jhi; j += U) block(j, std::integral_constant<int, U>{});
                        rest_blocks<U>(block, j, jhi);
                    }
#pragma unroll
                    for (int m = 0; m < PPL; ++m) {
                        if (valid[m]) {
                            unsafeAtomicAdd(g_dw + p1 + i[m], gd[m]);
                            unsafeAtomicAdd(g_w + p1 + i[m], gw[m]);
                        }
                    }
                }
            }
        }
    }
}

// Work the tile kernels do, per the decomposition and pruning above (one band): one
// wave per item.  out[0] lane slots stepped (lanes x pruned j range, idle lanes
// included), out[1] the same over lanes holding a pixel, out[2] tasks, out[3] tasks
// with a non-empty j range.  Each lane slot is one pixel pair evaluated.
template <int PPL>
__global__ void work_stats_kernel(const lya_xi_problem p, int split, unsigned long long* out) {
    constexpr int SL = 64 * PPL;
    const int lane = threadIdx.x & 63;
    const int nw = (int)(gridDim.x * blockDim.x) >> 6;
    const double bp64 = (double)p.np_rp / p.rpmax, bt64 = (double)p.np_rt / p.rtmax;
    unsigned long long slots = 0, valid = 0, tasks = 0, busy = 0;
    for (int it = uni((int)((blockIdx.x * blockDim.x + threadIdx.x) >> 6)); it < p.n_items; it += nw) {
        TaskCtx t;
        t.f1 = p.items[4 * it];
        const int nb0 = p.items[4 * it + 1], nb1 = p.items[4 * it + 2];
        t.p1 = p.pix_off[t.f1];
        t.n1 = (int)(p.pix_off[t.f1 + 1] - t.p1);
        t.slot = p.items[4 * it + 3];
        for (int q = nb0; q < nb1; ++q) {
            const int f2 = p.nb_idx[q];
            const int64_t p2 = p.pix_off[f2];
            const int n2 = (int)(p.pix_off[f2 + 1] - p2);
            const PairGeo g = pair_geo(p, t.f1, f2, bp64, bt64);
            Units un;
            int ja, jb;
            if (!pair_units<false>(p, t, f2, p2, n2, g, SL, split, lane, un, ja, jb)) continue;
            for (int sub = 0; sub < un.upq; ++sub) {
                UnitArgs a = unit_args(t, un, SL, sub, f2, p2, n2, ja, jb);
                for (; a.i0 < a.tl.n1; a.i0 += 64) {
                    const int w = (PPL > 1 && a.wide) ? SL : 64;
                    const int ilast = min(a.tl.n1, a.i0 + w) - 1;
                    int jlo, jhi;
                    prune(p, a.tl.f1, a.tl.p1, a.i0, ilast, a.fu, p.pix_off[a.fu], a.u0, a.u1, g, lane, jlo,
                          jhi);
                    const unsigned long long span = (unsigned long long)max(0, jhi - jlo);
                    slots += (unsigned long long)w * span;
                    valid += (unsigned long long)(ilast - a.i0 + 1) * span;
                    tasks += 1;
                    busy += span > 0;
                    if (!a.tail) break;
                }
            }
        }
    }
    if (lane == 0) {
        atomicAdd(out, slots);
        atomicAdd(out + 1, valid);
        atomicAdd(out + 2, tasks);
        atomicAdd(out + 3, busy);
    }
}

// out[f] = sum of x[off[f] .. off[f+1]), one wave per segment, fixed summation order.
__global__ void segment_sum_kernel(const double* __restrict__ x, const int64_t* __restrict__ off, int32_t nseg,
                                   double* __restrict__ out) {
    const int64_t seg = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const int lane = threadIdx.x & 63;
    if (seg >= nseg) return;
    const int64_t b = off[seg], e = off[seg + 1];
    double acc = 0.0;
    for (int64_t k = b + lane; k < e; k += 64) acc += x[k];
    for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o);
    if (lane == 0) out[seg] = acc;
}

__global__ void zero_outputs(uint64_t* cnt, double* w, double* dw, int64_t n, uint32_t* queue) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        if (cnt) cnt[i] = 0ull;
        if (w) w[i] = 0.0;
        if (dw) dw[i] = 0.0;
    }
    if (blockIdx.x == 0 && threadIdx.x < 16) queue[threadIdx.x] = 0u;
}

__global__ void halfangle_kernel(const double* cos12, double* c, double* s, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    // same expression as halfangles() with the dot product replaced by its value
    double x = fmin(1.0, fmax(-1.0, cos12[i]));
    c[i] = sqrt(0.5 * (1.0 + x));
    s[i] = sqrt(0.5 * (1.0 - x));
}

// LDS atomic-add rate: every lane adds to its own entry of a 4096-entry table, 64
// entries apart from one add to the next (indices stay in [0, 4096)).
template <class T>
__global__ void __launch_bounds__(512) lds_atomic_probe(T* out, int iters) {
    __shared__ T h[4096];
    for (int i = threadIdx.x; i < 4096; i += blockDim.x) h[i] = (T)0;
    __syncthreads();
    uint32_t a = threadIdx.x & 4095u;
    for (int k = 0; k < iters; ++k) {
        __hip_atomic_fetch_add(h + a, (T)1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        a = (a + 64u) & 4095u;
    }
    __syncthreads();
    if (threadIdx.x == 0) out[blockIdx.x] = h[threadIdx.x];
}

__global__ void copy_kernel(const uint4* __restrict__ src, uint4* __restrict__ dst, int64_t n16) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n16; i += stride) dst[i] = src[i];
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
using XiKernel = void (*)(const lya_xi_problem, int, const BandShape, int);

// Banded kernels (binnings larger than the LDS histogram) are built for one
// workgroup shape: 8 waves, one histogram replica, fp64 and mixed modes.
template <int MODE, int PPL, bool CNT, bool POW2, int R, bool MERGE>
XiKernel pick_w(int waves, bool banded) {
    if (banded) {
        if constexpr (R == 1 && MODE != LYA_MODE_MIXED_F32)
            if (waves == 8) return xi_kernel<MODE, PPL, 8, CNT, POW2, 1, MERGE, true>;
        return nullptr;
    }
    if constexpr (R == 1) {
        if (waves == 4) return xi_kernel<MODE, PPL, 4, CNT, POW2, 1, MERGE, false>;
        if (waves == 16) return xi_kernel<MODE, PPL, 16, CNT, POW2, 1, MERGE, false>;
    }
    if (waves == 8) return xi_kernel<MODE, PPL, 8, CNT, POW2, R, MERGE, false>;
    return nullptr;
}
template <int MODE, bool CNT, bool POW2, int R, bool MERGE>
XiKernel pick_p(int ppl, int waves, bool banded) {
    if (ppl == 1) return pick_w<MODE, 1, CNT, POW2, R, MERGE>(waves, banded);
    if (ppl == 2) return pick_w<MODE, 2, CNT, POW2, R, MERGE>(waves, banded);
    if (ppl == 4) return pick_w<MODE, 4, CNT, POW2, R, MERGE>(waves, banded);
    return nullptr;
}
template <int MODE, bool POW2, bool MERGE>
XiKernel pick_r(bool cnt, int R, int ppl, int waves, bool banded) {
    if (R == 1)
        return cnt ? pick_p<MODE, true, POW2, 1, MERGE>(ppl, waves, banded)
                   : pick_p<MODE, false, POW2, 1, MERGE>(ppl, waves, banded);
    if (R == 2)
        return cnt ? pick_p<MODE, true, POW2, 2, MERGE>(ppl, waves, banded)
                   : pick_p<MODE, false, POW2, 2, MERGE>(ppl, waves, banded);
    return nullptr;
}
template <int MODE, bool POW2>
XiKernel pick_c(bool cnt, int R, int ppl, int waves, bool merge, bool banded) {
    if constexpr (MODE == LYA_MODE_MIXED_F32) {
        (void)merge;
        return pick_r<MODE, POW2, false>(cnt, R, ppl, waves, banded);
    } else {
        return merge ? pick_r<MODE, POW2, true>(cnt, R, ppl, waves, banded)
                     : pick_r<MODE, POW2, false>(cnt, R, ppl, waves, banded);
    }
}

using WalkKernel = void (*)(const lya_xi_problem, int);
template <bool CNT, bool POW2>
WalkKernel pick_walk(int waves) {
    if (waves == 8) return xi_walk_kernel<8, CNT, POW2>;
    if (waves == 16) return xi_walk_kernel<16, CNT, POW2>;
    return nullptr;
}

// Band grid for a histogram of rows x (columns + pad) LDS entries within HS: whole
// rows in column bands while a row of one column fits, else blocks of at most 48
// columns; bands of equal size.  Returns false when the binning fits unbanded.
bool band_shape(int np_rp, int np_rt, int pad, BandShape& bs) {
    bs = BandShape{np_rp, np_rt, 1, 1};
    if ((int64_t)np_rp * (np_rt + pad) <= HS) return false;
    if ((int64_t)np_rp * (1 + pad) <= HS) {
        const int bcmax = HS / np_rp - pad;
        bs.nbt = (np_rt + bcmax - 1) / bcmax;
        bs.bc = (np_rt + bs.nbt - 1) / bs.nbt;
        bs.br = np_rp;
        bs.nbp = 1;
    } else {
        bs.bc = std::min(np_rt, 48);
        bs.nbt = (np_rt + bs.bc - 1) / bs.bc;
        bs.bc = (np_rt + bs.nbt - 1) / bs.nbt;
        const int brmax = HS / (bs.bc + pad);
        bs.nbp = (np_rp + brmax - 1) / brmax;
        bs.br = (np_rp + bs.nbp - 1) / bs.nbp;
    }
    return true;
}

bool is_pow2(double x) {
    int e = 0;
    return x > 0 && std::frexp(x, &e) == 0.5;
}

struct Resolved {
    XiKernel fn;
    WalkKernel walk_fn;
    int mode, ppl, waves, R, ntpad, groups;
    bool banded;
    BandShape bs;
    int split;
    uint32_t lds;
};

int resolve(const lya_xi_problem* p, const lya_xi_config* cfg, Resolved& r) {
    r.mode = cfg ? cfg->mode : LYA_MODE_MIXED;
    r.ppl = (cfg && cfg->px_per_lane) ? cfg->px_per_lane : 1;
    r.waves = (cfg && cfg->waves) ? cfg->waves : 8;
    r.R = (cfg && cfg->replicas) ? cfg->replicas : 1;
    const int pad = (cfg && cfg->row_pad > 0) ? cfg->row_pad : 0;
    if (r.mode != LYA_MODE_FP64 && r.mode != LYA_MODE_MIXED && r.mode != LYA_MODE_MIXED_F32)
        return lya_fail(LYA_ERR_ARG, "unknown mode %d", r.mode);
    if (r.R != 1 && r.R != 2) return lya_fail(LYA_ERR_ARG, "replicas must be 1 or 2");
    if (r.R > r.waves) return lya_fail(LYA_ERR_ARG, "replicas > waves");
    const bool cnt = !(cfg && cfg->no_counts);
    const bool pow2 = p && is_pow2((double)p->np_rp / p->rpmax) && is_pow2((double)p->np_rt / p->rtmax);
    const bool walk = cfg && cfg->walk;
    const bool merge = cfg && cfg->merge;
    // longest forest-1 tail walked against forest 2 in lanes (0: whole slabs only)
    r.split = !cfg || cfg->tail_max == 0 ? 48 : (cfg->tail_max < 0 ? 0 : std::min(cfg->tail_max, 63));
    // LDS columns past the last bin column of a row: merged runs park out-of-range
    // runs in the padding column, so they need at least one
    const int padc = walk ? 1 : (merge && r.mode != LYA_MODE_MIXED_F32 ? std::max(pad, 1) : pad);
    r.banded = false;
    r.bs = BandShape{0, 0, 1, 1};
    r.fn = nullptr;
    r.walk_fn = nullptr;
    if (p) {
        if (p->np_rp < 1 || p->np_rt < 1 || !(p->rpmax > 0) || !(p->rtmax > 0))
            return lya_fail(LYA_ERR_ARG, "bad binning");
        if (!walk) r.banded = band_shape(p->np_rp, p->np_rt, padc, r.bs);
        if (r.banded && (r.mode == LYA_MODE_MIXED_F32 || r.R != 1 || r.waves != 8))
            return lya_fail(LYA_ERR_UNSUPPORTED,
                            "np_rp*(np_rt+pad)=%lld exceeds the LDS histogram (%d entries): binnings that large "
                            "run in bands, with the fp64 or mixed mode, 8 waves and one replica",
                            (long long)p->np_rp * (p->np_rt + padc), HS);
    }
    if (walk) {
        if (r.mode != LYA_MODE_FP64) return lya_fail(LYA_ERR_UNSUPPORTED, "walk 1 runs the FP64 mode only");
        if (p && p->max_len > WALK_MAX_LEN)
            return lya_fail(LYA_ERR_UNSUPPORTED, "walk 1: forests of %d pixels (> %d)", p->max_len, WALK_MAX_LEN);
        if (p && (p->pad < p->max_len || p->max_len < 0))
            return lya_fail(LYA_ERR_ARG, "walk 1 needs pad (%d) >= max_len (%d)", p->pad, p->max_len);
        r.R = 1;
        r.walk_fn = pow2 ? (cnt ? pick_walk<true, true>(r.waves) : pick_walk<false, true>(r.waves))
                         : (cnt ? pick_walk<true, false>(r.waves) : pick_walk<false, false>(r.waves));
    } else if (r.mode == LYA_MODE_FP64)
        r.fn = pow2 ? pick_c<LYA_MODE_FP64, true>(cnt, r.R, r.ppl, r.waves, merge, r.banded)
                    : pick_c<LYA_MODE_FP64, false>(cnt, r.R, r.ppl, r.waves, merge, r.banded);
    else if (r.mode == LYA_MODE_MIXED)
        r.fn = pick_c<LYA_MODE_MIXED, false>(cnt, r.R, r.ppl, r.waves, merge, r.banded);
    else
        r.fn = pick_c<LYA_MODE_MIXED_F32, false>(cnt, r.R, r.ppl, r.waves, merge, r.banded);
    if (!r.fn && !r.walk_fn) return lya_fail(LYA_ERR_ARG, "unsupported shape px_per_lane=%d waves=%d", r.ppl, r.waves);
    if (p) {
        // LDS row stride: the band's columns (all of them when unbanded) plus padding
        r.ntpad = (r.banded ? r.bs.bc : p->np_rt) + padc;
        if (walk && (int64_t)(p->np_rp + 1) * r.ntpad > WALK_HS)
            return lya_fail(LYA_ERR_UNSUPPORTED, "walk 1: (np_rp+1)*(np_rt+1)=%d exceeds %d LDS bins (use walk 0)",
                            (p->np_rp + 1) * r.ntpad, WALK_HS);
        if ((int64_t)r.bs.br * r.ntpad > HS)
            return lya_fail(LYA_ERR_UNSUPPORTED, "band of %d x %d bins exceeds the LDS histogram (%d entries)",
                            r.bs.br, r.ntpad, HS);
    } else {
        r.ntpad = 0;
    }
    r.lds = 16 + ((r.mode == LYA_MODE_MIXED_F32) ? Layout<LYA_MODE_MIXED_F32, true>::bytes(r.R)
                  : cnt ? Layout<LYA_MODE_FP64, true>::bytes(r.R)
                        : Layout<LYA_MODE_FP64, false>::bytes(r.R));  // static LDS
    if (walk) r.lds = 16 + WalkLayout::BYTES + WALK_SG * 8 + WALK_NB * (8 * 5 + 4 * 5);
    int dev = 0;
    hipDeviceProp_t prop;
    if (hipGetDevice(&dev) != hipSuccess || hipGetDeviceProperties(&prop, dev) != hipSuccess)
        return lya_fail(LYA_ERR_HIP, "no HIP device");
    if (r.lds > prop.sharedMemPerBlock)
        return lya_fail(LYA_ERR_UNSUPPORTED, "LDS %u bytes > %zu per workgroup", r.lds, prop.sharedMemPerBlock);
    int per_cu = 0;
    const void* kfn = r.walk_fn ? (const void*)r.walk_fn : (const void*)r.fn;
    hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kfn, r.waves * 64, 0);
    if (e != hipSuccess || per_cu < 1) per_cu = 1;
    r.groups = (cfg && cfg->groups > 0) ? cfg->groups : per_cu * prop.multiProcessorCount;
    return LYA_OK;
}

int check_problem(const lya_xi_problem* p, int mode) {
    if (!p) return lya_fail(LYA_ERR_ARG, "null problem");
    if (!p->pix_off || !p->ux || !p->uy || !p->uz || !p->sorted || !p->dc || !p->nb_idx || !p->items ||
        !p->tiles || !p->queue)
        return lya_fail(LYA_ERR_ARG, "null device pointer in problem");
    if (mode != LYA_MODE_MIXED_F32 && (!p->we || !p->dw)) return lya_fail(LYA_ERR_ARG, "mode needs we, dw (FP64)");
    if (mode != LYA_MODE_FP64 && (!p->off32 || !p->base)) return lya_fail(LYA_ERR_ARG, "mixed modes need off32, base");
    if (mode == LYA_MODE_MIXED_F32 && (!p->we32 || !p->dw32)) return lya_fail(LYA_ERR_ARG, "mixed32 needs we32, dw32");
    if (mode != LYA_MODE_FP64 && !(p->tau > 0.0f && p->tau < 0.25f))
        return lya_fail(LYA_ERR_ARG, "mixed mode needs 0 < tau < 0.25 (got %g)", (double)p->tau);
    if (p->n_tiles < 0 || p->n_items < 0 || p->n_slots < 0) return lya_fail(LYA_ERR_ARG, "negative sizes");
    if (p->n_tile_groups < 0 || p->n_tile_groups > 16 || (p->n_tile_groups > 1 && !p->tile_groups))
        return lya_fail(LYA_ERR_ARG, "n_tile_groups must be 0..16, with tile_groups when above 1");
    return LYA_OK;
}

int launch_all(const lya_xi_problem* p, const Resolved& r, hipStream_t st) {
    const int64_t n = (int64_t)p->n_slots * p->np_rp * p->np_rt;
    int zb = (int)std::min<int64_t>((n + 255) / 256, 4096);
    if (zb < 1) zb = 1;
    hipLaunchKernelGGL(zero_outputs, dim3(zb), dim3(256), 0, st, p->out_cnt, p->out_w, p->out_dw, n, p->queue);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return lya_fail(LYA_ERR_HIP, "zero_outputs launch: %s", hipGetErrorString(e));
    if (r.walk_fn)
        hipLaunchKernelGGL(r.walk_fn, dim3(r.groups), dim3(r.waves * 64), 0, st, *p, r.ntpad);
    else
        hipLaunchKernelGGL(r.fn, dim3(r.groups), dim3(r.waves * 64), 0, st, *p, r.ntpad, r.bs, r.split);
    e = hipGetLastError();
    if (e != hipSuccess) return lya_fail(LYA_ERR_HIP, "xi kernel launch: %s", hipGetErrorString(e));
    return LYA_OK;
}

}  // namespace

struct lya_graph_s {
    hipGraph_t graph;
    hipGraphExec_t exec;
};

extern "C" int lya_xi_grid(const lya_xi_config* cfg, int32_t* groups) {
    Resolved r;
    int rc = resolve(nullptr, cfg, r);
    if (rc) return rc;
    *groups = r.groups;
    return LYA_OK;
}

extern "C" int lya_xi_run(const lya_xi_problem* p, const lya_xi_config* cfg, void* stream) {
    Resolved r;
    int rc = resolve(p, cfg, r);
    if (rc) return rc;
    if ((rc = check_problem(p, r.mode))) return rc;
    if (!p->out_cnt || !p->out_w || !p->out_dw) return lya_fail(LYA_ERR_ARG, "null output");
    return launch_all(p, r, (hipStream_t)stream);
}

extern "C" int lya_xi_graph_create(const lya_xi_problem* p, const lya_xi_config* cfg, void* stream,
                                   lya_graph_t* out) {
    (void)stream;
    Resolved r;
    int rc = resolve(p, cfg, r);
    if (rc) return rc;
    if ((rc = check_problem(p, r.mode))) return rc;
    if (!p->out_cnt || !p->out_w || !p->out_dw || !out) return lya_fail(LYA_ERR_ARG, "null output");
    hipStream_t cs;
    if (hipStreamCreateWithFlags(&cs, hipStreamNonBlocking) != hipSuccess)
        return lya_fail(LYA_ERR_HIP, "capture stream create failed");
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal);
    if (e != hipSuccess) {
        hipStreamDestroy(cs);
        return lya_fail(LYA_ERR_HIP, "begin capture: %s", hipGetErrorString(e));
    }
    rc = launch_all(p, r, cs);
    e = hipStreamEndCapture(cs, &g);
    hipStreamDestroy(cs);
    if (rc) {
        if (g) hipGraphDestroy(g);
        return rc;
    }
    if (e != hipSuccess) return lya_fail(LYA_ERR_HIP, "end capture: %s", hipGetErrorString(e));
    hipGraphExec_t x;
    e = hipGraphInstantiate(&x, g, nullptr, nullptr, 0);
    if (e != hipSuccess) {
        hipGraphDestroy(g);
        return lya_fail(LYA_ERR_HIP, "graph instantiate: %s", hipGetErrorString(e));
    }
    *out = new lya_graph_s{g, x};
    return LYA_OK;
}

extern "C" int lya_graph_launch(lya_graph_t g, void* stream) {
    if (!g) return lya_fail(LYA_ERR_ARG, "null graph");
    hipError_t e = hipGraphLaunch(g->exec, (hipStream_t)stream);
    if (e != hipSuccess) return lya_fail(LYA_ERR_HIP, "graph launch: %s", hipGetErrorString(e));
    return LYA_OK;
}

extern "C" int lya_graph_destroy(lya_graph_t g) {
    if (!g) return LYA_OK;
    hipGraphExecDestroy(g->exec);
    hipGraphDestroy(g->graph);
    delete g;
    return LYA_OK;
}

extern "C" int lya_xi_tau(const lya_xi_problem* p, double max_abs_off, double max_dc, double max_c_binner,
                          double max_s_binner, double* tau) {
    // Bound on |(FP32 scaled coordinate) - (FP64 scaled coordinate)| for pairs
    // whose coordinate is below np + 3 bins (larger ones are out of range in
    // both evaluations).  Terms, each a unit roundoff u32 = 2^-24 times a
    // magnitude bound, follow the mixed-mode arithmetic of run_task:
    //   rounding of P1 (or T1), of the fma result, of |v| + tau, the FP32
    //   representation of off32 and of c*binner (or s*binner), plus the FP64
    //   roundings of the reference formula and of r1 + base2 (at most 2 max_dc)
    //   in T1.  r1 - base2 and r1 - r2 are exact in FP64 (Sterbenz).  The kernel's band is exact for an error e <= tau: a pair
    //   whose FP32 bin differs from its FP64 bin has its shifted FP32 coordinate
    //   within tau + e above an edge and is flagged (fract < 2 tau).  The bound
    //   alone would do; a margin of 25 % is kept.
    if (!p || !tau || !(max_abs_off >= 0) || !(max_dc >= 0)) return lya_fail(LYA_ERR_ARG, "lya_xi_tau: bad arguments");
    const double u32 = 1.0 / 16777216.0, u64 = 1.0 / 9007199254740992.0;
    auto bound = [&](double nb, double kb) {
        const double q = nb + 3.0;        // largest relevant coordinate
        const double m = max_abs_off * kb;  // largest |off * k|
        return u32 * (2.0 * q + (q + m) + 2.0 * m) + 8.0 * u64 * (q + m + 2.0 * max_dc * kb);
    };
    const double t = 1.25 * std::max(bound(p->np_rp, max_c_binner), bound(p->np_rt, max_s_binner));
    *tau = t;
    return LYA_OK;
}

extern "C" int lya_segment_sum(const double* x, const int64_t* off, int32_t nseg, double* out, void* stream) {
    if (nseg < 0 || (nseg > 0 && (!x || !off || !out))) return lya_fail(LYA_ERR_ARG, "segment sum: bad arguments");
    if (nseg == 0) return LYA_OK;
    const unsigned blocks = (unsigned)(((int64_t)nseg * 64 + 255) / 256);
    hipLaunchKernelGGL(segment_sum_kernel, dim3(blocks), dim3(256), 0, (hipStream_t)stream, x, off, nseg, out);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return lya_fail(LYA_ERR_HIP, "segment sum launch: %s", hipGetErrorString(e));
    return LYA_OK;
}

extern "C" int lya_xi_work_stats(const lya_xi_problem* p, const lya_xi_config* cfg, uint64_t* out, void* stream) {
    Resolved r;
    int rc = resolve(p, cfg, r);
    if (rc) return rc;
    if ((rc = check_problem(p, r.mode))) return rc;
    if (!out) return lya_fail(LYA_ERR_ARG, "work stats: null output");
    if (r.walk_fn) return lya_fail(LYA_ERR_UNSUPPORTED, "work stats: tile kernels only (walk 0)");
    hipStream_t st = (hipStream_t)stream;
    hipLaunchKernelGGL(zero_outputs, dim3(1), dim3(64), 0, st, out, (double*)nullptr, (double*)nullptr, (int64_t)4,
                       p->queue);
    auto fn = r.ppl == 1 ? work_stats_kernel<1> : (r.ppl == 2 ? work_stats_kernel<2> : work_stats_kernel<4>);
    hipLaunchKernelGGL(fn, dim3(2048), dim3(256), 0, st, *p, r.split, (unsigned long long*)out);
    const hipError_t e = hipGetLastError();
    return e == hipSuccess ? LYA_OK : lya_fail(LYA_ERR_HIP, "work stats launch: %s", hipGetErrorString(e));
}

extern "C" int lya_xi_adjoint(const lya_xi_problem* p, const lya_xi_config* cfg, const double* alpha,
                              const double* beta, double* g_dw, double* g_w, void* stream) {
    const int mode = cfg ? cfg->mode : LYA_MODE_MIXED;
    int rc = check_problem(p, mode);
    if (rc) return rc;
    if (!alpha || !beta || !g_dw || !g_w) return lya_fail(LYA_ERR_ARG, "null adjoint buffer");
    if (p->np_rp < 1 || p->np_rt < 1 || !(p->rpmax > 0) || !(p->rtmax > 0))
        return lya_fail(LYA_ERR_ARG, "bad binning");
    const int64_t nb = (int64_t)p->np_rp * p->np_rt;
    int dev = 0;
    hipDeviceProp_t prop;
    if (hipGetDevice(&dev) != hipSuccess || hipGetDeviceProperties(&prop, dev) != hipSuccess)
        return lya_fail(LYA_ERR_HIP, "no HIP device");
    // the (alpha, beta) table in LDS when it fits a workgroup, else read from global memory
    const bool gtab = (size_t)(nb + 1) * 16 > prop.sharedMemPerBlock;
    const size_t lds = gtab ? 0 : (size_t)(nb + 1) * 16;
    // one forest-1 pixel per lane: two (128-pixel tasks) need ~160 VGPRs for the batched
    // bin gathers, 3 waves per SIMD, and measured 38 % slower
    const bool pow2 = is_pow2((double)p->np_rp / p->rpmax) && is_pow2((double)p->np_rt / p->rtmax);
    auto fn = (mode == LYA_MODE_FP64)
                  ? (pow2 ? (gtab ? xi_adjoint_kernel<LYA_MODE_FP64, 8, 1, true, true>
                                  : xi_adjoint_kernel<LYA_MODE_FP64, 8, 1, true, false>)
                          : (gtab ? xi_adjoint_kernel<LYA_MODE_FP64, 8, 1, false, true>
                                  : xi_adjoint_kernel<LYA_MODE_FP64, 8, 1, false, false>))
                  : (gtab ? xi_adjoint_kernel<LYA_MODE_MIXED, 8, 1, false, true>
                          : xi_adjoint_kernel<LYA_MODE_MIXED, 8, 1, false, false>);
    if (hipFuncSetAttribute((const void*)fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds) != hipSuccess)
        return lya_fail(LYA_ERR_HIP, "adjoint: cannot reserve %zu bytes of LDS", lds);
    int per_cu = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, (const void*)fn, 512, lds) != hipSuccess || per_cu < 1)
        per_cu = 1;
    const int groups = (cfg && cfg->groups > 0) ? cfg->groups : per_cu * prop.multiProcessorCount;
    hipStream_t st = (hipStream_t)stream;
    hipLaunchKernelGGL(zero_outputs, dim3(1), dim3(64), 0, st, (uint64_t*)nullptr, (double*)nullptr,
                       (double*)nullptr, (int64_t)0, p->queue);
    hipLaunchKernelGGL(fn, dim3(groups), dim3(512), lds, st, *p, alpha, beta, g_dw, g_w);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return lya_fail(LYA_ERR_HIP, "adjoint launch: %s", hipGetErrorString(e));
    return LYA_OK;
}

extern "C" int lya_debug_halfangles(const double* cos12, double* c, double* s, int32_t n, void* stream) {
    if (n < 0 || (n > 0 && (!cos12 || !c || !s))) return lya_fail(LYA_ERR_ARG, "bad halfangle arguments");
    if (n == 0) return LYA_OK;
    hipLaunchKernelGGL(halfangle_kernel, dim3((n + 255) / 256), dim3(256), 0, (hipStream_t)stream, cos12, c, s, n);
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? LYA_OK : lya_fail(LYA_ERR_HIP, "halfangle launch: %s", hipGetErrorString(e));
}

extern "C" int lya_debug_lds_atomic(int32_t type, int32_t blocks, int32_t iters, void* out, void* stream) {
    if (blocks < 1 || iters < 0 || !out || type < 0 || type > 2) return lya_fail(LYA_ERR_ARG, "lds probe: bad arguments");
    hipStream_t st = (hipStream_t)stream;
    if (type == 0) hipLaunchKernelGGL(lds_atomic_probe<uint32_t>, dim3(blocks), dim3(512), 0, st, (uint32_t*)out, iters);
    if (type == 1) hipLaunchKernelGGL(lds_atomic_probe<float>, dim3(blocks), dim3(512), 0, st, (float*)out, iters);
    if (type == 2) hipLaunchKernelGGL(lds_atomic_probe<double>, dim3(blocks), dim3(512), 0, st, (double*)out, iters);
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? LYA_OK : lya_fail(LYA_ERR_HIP, "lds probe launch: %s", hipGetErrorString(e));
}

extern "C" int lya_copy_bytes(const void* src, void* dst, int64_t n, void* stream) {
    if (n < 0 || (n % 16) || (n > 0 && (!src || !dst))) return lya_fail(LYA_ERR_ARG, "copy: n must be a multiple of 16");
    if (n == 0) return LYA_OK;
    const int64_t n16 = n / 16;
    const int blocks = (int)std::min<int64_t>((n16 + 255) / 256, 8192);
    hipLaunchKernelGGL(copy_kernel, dim3(blocks), dim3(256), 0, (hipStream_t)stream, (const uint4*)src, (uint4*)dst, n16);
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? LYA_OK : lya_fail(LYA_ERR_HIP, "copy launch: %s", hipGetErrorString(e));
}
