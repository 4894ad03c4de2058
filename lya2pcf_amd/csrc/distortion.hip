// Distortion matrix of the projected correlation function, FP64, gfx950.
//
// For a forest pair (f1, f2) with pixel-pair bins A(i, j) the reference
// (cuda_kernels.cpp:165-278) builds, per active bin B of the pair,
//   eta12[B][i] = sum_{j:(i,j) in B} w_j / O2          eta13[B][i] = sum w_j dl_j / L2
//   eta21[B][j] = sum_{i:(i,j) in B} w_i / O1          eta31[B][j] = sum w_i dl_i / L1
//   eta22[B] = sum_B w_i w_j / (O1 O2)    eta23[B] = sum_B w_i w_j dl_j / (O1 L2)
//   eta32[B] = sum_B w_i dl_i w_j / (L1 O2)  eta33[B] = sum_B w_i dl_i w_j dl_j / (L1 L2)
// (O = sum w, L = sum w dl^2 of the forest) and adds, for every in-range pixel
// pair (i, j) in bin A and every active B,
//   D[A][B] += w_i w_j (dl_i dl_j eta33 + dl_i eta32 + dl_j eta23 + eta22
//                        - dl_i eta31[j] - dl_j eta13[i] - eta21[j] - eta12[i])
//   D[A][A] += w_i w_j,  wA[A] += w_i w_j.
//
// Summing the eight terms over the pairs of A first gives, per forest pair,
//   D_pair[A][B] = sum_k Z[A][k] v[k] Z[B][k]  (+ S00[A] on the diagonal)
// with one k per table row:
//   k = j (n2 rows)  Z = eta21[.][j]  v = -w_j O1     k = j  Z = eta31[.][j]  v = -w_j L1
//   k = i (n1 rows)  Z = eta12[.][i]  v = -w_i O2     k = i  Z = eta13[.][i]  v = -w_i L2
//   4 rows  Z = S00, S01, S10, S11 (bin sums of w_i w_j, w_i w_j dl_j, w_i dl_i w_j,
//           w_i dl_i w_j dl_j)  v = 1/(O1 O2), 1/(O1 L2), 1/(L1 O2), 1/(L1 L2).
// D_pair is symmetric, a rank-K (K = 2 n1 + 2 n2 + 4) product over the nact active
// bins: nact^2 K / 2 multiply-adds instead of n1 n2 nact histogram updates.
//
// MI355X organisation: one workgroup of 8 waves per CU, persistent over an atomic
// queue of forest pairs.  Per pair:
//   1. dc, w and w dl of both forests are staged in LDS; the FP64 bin of every
//      pixel pair (the xi formula) flags the active bins, which a wave scan
//      numbers 0..nact-1 in bin order.  No per-pair bin table is kept: bins are
//      recomputed where needed, which leaves the LDS to the Z chunks.
//   2. Z is built in chunks of at most 128 rows in LDS, zs[kk][b]: one thread per
//      pixel pair of the chunk (two per step of the loop) recomputes its bin and adds
//      its weight (LDS atomics).
//      The four rank rows follow from the eta21, eta31 rows of the j chunks
//      (S00[A] = O1 sum_j w_j eta21[A][j], ...) without another pass over the pairs;
//      they are kept in a small per-workgroup global scratch.
//   3. The upper triangle of D_pair is cut into 16x16 blocks, each one
//      v_mfma_f64_16x16x4f64 accumulator (FP64 matrix cores, four Z rows per
//      instruction) updated by every chunk.  A pass over Z gives every wave the same
//      number NQ <= 10 of consecutive blocks (NQ a compile-time constant per pass, so
//      that the accumulators stay in registers and the operand loads of the next four
//      rows run ahead of the MFMAs); larger pairs take several passes.  Z rows sit
//      ldz = npad + 4 doubles apart, which staggers the four rows an MFMA operand load
//      reads across the LDS banks.
//   4. The blocks go into the upper triangle of D (FP64 global atomics, act is
//      ascending so row <= column); lya_distortion_mirror() fills the lower
//      triangle once accumulation is over.
// Forests whose L = 0 (constant loglam) get zero slope terms (the reference
// divides by zero there).
#include <hip/hip_runtime.h>
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <mutex>
#include <type_traits>

#include "../../include/lya2pcf.h"
#include "errors.h"

#pragma clang fp contract(off)

#define CAS __attribute__((address_space(4)))

namespace {

typedef double double4v __attribute__((ext_vector_type(4)));

constexpr int DWAVES = 8;
constexpr int DTHREADS = DWAVES * 64;
constexpr int BS = 16;       // D_pair block edge: one v_mfma_f64_16x16x4f64 accumulator
constexpr int BPW = 10;      // blocks per wave per pass over Z (at most; 80 per pass)
constexpr int KC_MAX = 128;  // Z rows per LDS chunk
constexpr int ZSKEW = 4;     // extra doubles per Z row in LDS: rows k and k+1 start 32 bytes apart in the banks
constexpr size_t LDS_STATIC = 256;         // the kernel's static LDS, rounded up

template <class T>
__device__ __forceinline__ const CAS T* cptr(const T* p) {
    return (const CAS T*)(p);
}
__device__ __forceinline__ int uni(int x) { return __builtin_amdgcn_readfirstlane(x); }

struct DArgs {
    int32_t n_forest;
    const int64_t* pix_off;
    const double* dc;
    const double* we;
    const double* dl;
    const double* ux;
    const double* uy;
    const double* uz;
    const double* omega;
    const double* odl2;
    const int32_t* pairs;
    int32_t n_pairs;
    int32_t max_len;
    int32_t np_rp, np_rt;
    double rpmax, rtmax;
    double* D;
    double* wA;
    int32_t zcap;        // doubles of the LDS chunk buffer
    int32_t npad_max;    // bins padded to the block edge
    double* tscr;        // [groups][4][npad_max]: rank rows of each workgroup's current pair
    uint32_t* queue;     // pair counter
    uint32_t* skip;      // [2] pairs skipped: forest longer than max_len, too many active bins
};

// LDS bytes before the chunk buffer: bin maps, chunk weights, staged pixels.
__host__ __device__ constexpr size_t lds_fixed(int nb, int max_len) {
    return ((4 * (size_t)nb + 15) & ~(size_t)15) + 8 * KC_MAX + 48 * (size_t)max_len;
}

__global__ void __launch_bounds__(DTHREADS) distortion_kernel(const DArgs a) {
    extern __shared__ __align__(16) unsigned char lds[];
    __shared__ int s_pair, s_nact, s_wcnt[DWAVES];
    const int NB = a.np_rp * a.np_rt;
    const int NP = a.np_rp, NT = a.np_rt;
    int16_t* cmap = (int16_t*)lds;                                     // [NB] bin -> compact, -1
    int16_t* act = cmap + NB;                                          // [NB] compact -> bin
    double* vk = (double*)(lds + ((4 * (size_t)NB + 15) & ~(size_t)15));  // [KC_MAX] chunk weights
    double* r1s = vk + KC_MAX;                                         // [6][max_len] staged pixels
    double* sw1 = r1s + a.max_len;                                     //   w of forest 1
    double* swd1 = sw1 + a.max_len;                                    //   w dl of forest 1
    double* r2s = swd1 + a.max_len;
    double* sw2 = r2s + a.max_len;
    double* swd2 = sw2 + a.max_len;
    double* zs = (double*)(lds + lds_fixed(NB, a.max_len));           // [zcap]: Z chunk [kk][b]
    const int tid = threadIdx.x, lane = tid & 63, wave = uni(tid >> 6);
    const double bp64 = (double)NP / a.rpmax, bt64 = (double)NT / a.rtmax;

    // visits the cells (r, c), r < nr, c < nc, in row-major order strided over the
    // workgroup, two per call: f(rA, cA, rB, cB, hasB) with cell B = cell A + DTHREADS
    // (B repeats A past the end), so that the bins of both are computed before either
    // is stored
    auto grid2 = [&](int nr, int nc, auto&& f) {
        if (nr <= 0 || nc <= 0) return;
        int ra = tid / nc, ca = tid - ra * nc;
        int rb = ra + DTHREADS / nc, cb = ca + (DTHREADS - (DTHREADS / nc) * nc);
        if (cb >= nc) {
            cb -= nc;
            ++rb;
        }
        const int dr = (2 * DTHREADS) / nc, dc = 2 * DTHREADS - dr * nc;
        while (ra < nr) {
            const bool hb = rb < nr;
            f(ra, ca, hb ? rb : ra, hb ? cb : ca, hb);
            ra += dr;
            ca += dc;
            if (ca >= nc) {
                ca -= nc;
                ++ra;
            }
            rb += dr;
            cb += dc;
            if (cb >= nc) {
                cb -= nc;
                ++rb;
            }
        }
    };

    // the queue index of the next pair is drawn while this one runs and published at its end
    if (tid == 0) s_pair = (int)atomicAdd(a.queue, 1u);
    for (;;) {
        __syncthreads();
        const int pr = uni(s_pair);
        if (pr >= a.n_pairs) break;
        int nxt = 0;
        if (tid == 0) nxt = (int)atomicAdd(a.queue, 1u);
        auto publish = [&]() {
            if (tid == 0) s_pair = nxt;
        };
        const int f1 = cptr(a.pairs)[2 * pr], f2 = cptr(a.pairs)[2 * pr + 1];
        const int64_t p1 = cptr(a.pix_off)[f1], p2 = cptr(a.pix_off)[f2];
        const int n1 = (int)(cptr(a.pix_off)[f1 + 1] - p1), n2 = (int)(cptr(a.pix_off)[f2 + 1] - p2);
        if (n1 > a.max_len || n2 > a.max_len) {
            // longer than the staging arrays: skipped, and counted in queue[1]; the barrier
            // keeps thread 0 from publishing the next pair before every wave has read this one
            if (tid == 0) atomicAdd(a.skip, 1u);
            __syncthreads();
            publish();
            continue;
        }
        const double O1 = cptr(a.omega)[f1], O2 = cptr(a.omega)[f2];
        const double L1 = cptr(a.odl2)[f1], L2 = cptr(a.odl2)[f2];
        const double iO1 = 1.0 / O1, iO2 = 1.0 / O2;
        const double iL1 = (L1 != 0.0) ? 1.0 / L1 : 0.0, iL2 = (L2 != 0.0) ? 1.0 / L2 : 0.0;
        double cos12 = cptr(a.ux)[f1] * cptr(a.ux)[f2] + cptr(a.uy)[f1] * cptr(a.uy)[f2];
        cos12 = cos12 + cptr(a.uz)[f1] * cptr(a.uz)[f2];
        cos12 = fmin(1.0, fmax(-1.0, cos12));
        const double c = sqrt(0.5 * (1.0 + cos12)), s = sqrt(0.5 * (1.0 - cos12));
        // global bin of pixel pair (i, j), the xi kernel's FP64 expression, or -1
        auto gbin = [&](int i, int j) -> int {
            const double r1 = r1s[i], r2 = r2s[j];
            const uint32_t bp = (uint32_t)__double2int_rz((fabs(r1 - r2) * c) * bp64);
            const uint32_t bt = (uint32_t)__double2int_rz(((r1 + r2) * s) * bt64);
            return (bp < (uint32_t)NP && bt < (uint32_t)NT) ? (int)(bp * NT + bt) : -1;
        };

        // ---- 1. staging, active bins ------------------------------------------------
        for (int i = tid; i < n1; i += DTHREADS) {
            const double w = a.we[p1 + i];
            r1s[i] = a.dc[p1 + i];
            sw1[i] = w;
            swd1[i] = w * a.dl[p1 + i];
        }
        for (int j = tid; j < n2; j += DTHREADS) {
            const double w = a.we[p2 + j];
            r2s[j] = a.dc[p2 + j];
            sw2[j] = w;
            swd2[j] = w * a.dl[p2 + j];
        }
        for (int b = tid; b < NB; b += DTHREADS) cmap[b] = -1;
        __syncthreads();
        grid2(n2, n1, [&](int ja, int ia, int jb, int ib, bool hb) {
            const int ga = gbin(ia, ja), gb = gbin(ib, jb);
            if (ga >= 0) cmap[ga] = 1;
            if (hb && gb >= 0) cmap[gb] = 1;
        });
        __syncthreads();
        // compaction in bin order: wave w numbers the bins of its eighth of the binning,
        // after the counts of the waves before it
        {
            const int perw = (NB + DWAVES - 1) / DWAVES;
            const int wb0 = min(NB, wave * perw), wb1 = min(NB, wb0 + perw);
            const int per = (perw + 63) / 64;
            const int b0 = min(wb1, wb0 + lane * per), b1 = min(wb1, b0 + per);
            int cnt = 0;
            for (int b = b0; b < b1; ++b) cnt += (cmap[b] >= 0);
            int incl = cnt;
            for (int o = 1; o < 64; o <<= 1) {
                const int v = __shfl_up(incl, o);
                if (lane >= o) incl += v;
            }
            if (lane == 63) s_wcnt[wave] = incl;
            __syncthreads();
            int base = 0;
            for (int w = 0; w < wave; ++w) base += s_wcnt[w];
            int idx = base + incl - cnt;
            for (int b = b0; b < b1; ++b) {
                if (cmap[b] >= 0) {
                    cmap[b] = (int16_t)idx;
                    act[idx] = (int16_t)b;
                    ++idx;
                }
            }
            if (tid == DTHREADS - 1) s_nact = base + incl;
        }
        __syncthreads();
        const int nact = uni(s_nact);
        if (nact == 0) {
            publish();
            continue;
        }
        // the LDS chunk buffer must hold two Z rows of the pair's active bins (padded to
        // the block edge, plus the row skew); a pair with more active bins is skipped and
        // counted in queue[2]
        if (a.zcap / (((nact + BS - 1) / BS) * BS + ZSKEW) < 2) {
            if (tid == 0) atomicAdd(a.skip + 1, 1u);
            publish();
            continue;
        }
        // compact bin of pixel pair (i, j), or -1
        auto cbin = [&](int i, int j) -> int {
            const int g = gbin(i, j);
            return g >= 0 ? (int)cmap[g] : -1;
        };

        // ---- 2. D_pair = Z diag(v) Z^T over the upper triangle ----------------------
        const int NBK = (nact + BS - 1) / BS;
        const int npad = NBK * BS;
        const int ldz = npad + ZSKEW;   // LDS row stride of the Z chunk
        const int nblk = NBK * (NBK + 1) / 2;
        // chunk rows in zs[0 .. kc_max ldz); the four rank rows T00, T01, T10, T11 (see the
        // rank chunks below) in the workgroup's global scratch, T[q npad + b]
        double* T = a.tscr + (int64_t)blockIdx.x * 4 * a.npad_max;
        const int kc_max = min(KC_MAX, (a.zcap / ldz) & ~1);
        const int nchunk_j = (2 * n2 + kc_max - 1) / kc_max, nchunk_i = (2 * n1 + kc_max - 1) / kc_max;
        const int nrank = kc_max >= 4 ? 1 : 2;   // rank chunks of 4 rows, or of 2
        const int nchunk = nchunk_j + nchunk_i + nrank;
        // clears rows 0 .. kc-1 of the Z chunk (ldz is even and zs 16-byte aligned)
        auto zero_rows = [&](int kc) {
            double2* z2 = (double2*)zs;
            for (int e = tid; e < (kc * ldz) / 2; e += DTHREADS) z2[e] = make_double2(0.0, 0.0);
        };

        // One pass over Z for the blocks t0 .. t0 + DWAVES*NQ - 1: every wave takes NQ
        // consecutive blocks (the same NQ for all waves, a compile-time count, so that the
        // accumulators stay in registers and the operand loads can run ahead of the MFMAs).
        auto pass = [&](auto NQC, const int t0) {
            constexpr int NQ = decltype(NQC)::value;
            // this wave's blocks t0 + wave NQ + q as (row block, column block >= row block);
            // the lane's A and B operand column; blocks past nblk are computed on row 0
            // of the chunk and never written
            uint32_t boa[NQ], bob[NQ];
            bool bval[NQ];
#pragma unroll
            for (int q = 0; q < NQ; ++q) {
                int t = t0 + wave * NQ + q;
                bval[q] = t < nblk;
                int ba = 0;
                while (bval[q] && t >= NBK - ba) { t -= NBK - ba; ++ba; }
                boa[q] = (uint32_t)((bval[q] ? ba * BS : 0) + (lane & 15));
                bob[q] = (uint32_t)((bval[q] ? (ba + t) * BS : 0) + (lane & 15));
            }
            double4v acc[NQ];
#pragma unroll
            for (int q = 0; q < NQ; ++q) acc[q] = (double4v){0.0, 0.0, 0.0, 0.0};

            for (int ch = 0; ch < nchunk; ++ch) {
                // build the chunk's rows of Z in zs[kk][b] and their weights in vk[kk]
                int kc;
                if (ch < nchunk_j) {
                    // rows (eta21, eta31) of forest-2 pixels j0 .. j0 + kc/2
                    const int j0 = ch * (kc_max / 2);
                    const int nj = min(kc_max / 2, n2 - j0);
                    kc = 2 * nj;
                    zero_rows(kc);
                    if (ch == 0 && t0 == 0)
                        for (int e = tid; e < 4 * npad; e += DTHREADS) T[e] = 0.0;
                    if (tid < kc) {
                        const double wj = sw2[j0 + (tid >> 1)];
                        vk[tid] = (tid & 1) ? -wj * L1 : -wj * O1;
                    }
                    __syncthreads();
                    // consecutive threads on consecutive rows (forest-2 pixels) of the chunk: runs
                    // of equal bins along forest 1 would put neighbouring lanes on one address
                    grid2(n1, nj, [&](int ia, int ja, int ib, int jb, bool hb) {
                        const int ba = cbin(ia, j0 + ja), bb = cbin(ib, j0 + jb);
                        if (ba >= 0) {
                            atomicAdd(zs + (2 * ja) * ldz + ba, sw1[ia] * iO1);
                            atomicAdd(zs + (2 * ja + 1) * ldz + ba, swd1[ia] * iL1);
                        }
                        if (hb && bb >= 0) {
                            atomicAdd(zs + (2 * jb) * ldz + bb, sw1[ib] * iO1);
                            atomicAdd(zs + (2 * jb + 1) * ldz + bb, swd1[ib] * iL1);
                        }
                    });
                } else if (ch < nchunk_j + nchunk_i) {
                    // rows (eta12, eta13) of forest-1 pixels i0 .. i0 + kc/2
                    const int i0 = (ch - nchunk_j) * (kc_max / 2);
                    const int ni = min(kc_max / 2, n1 - i0);
                    kc = 2 * ni;
                    zero_rows(kc);
                    if (tid < kc) {
                        const double wi = sw1[i0 + (tid >> 1)];
                        vk[tid] = (tid & 1) ? -wi * L2 : -wi * O2;
                    }
                    __syncthreads();
                    grid2(n2, ni, [&](int ja, int ia, int jb, int ib, bool hb) {
                        const int ba = cbin(i0 + ia, ja), bb = cbin(i0 + ib, jb);
                        if (ba >= 0) {
                            atomicAdd(zs + (2 * ia) * ldz + ba, sw2[ja] * iO2);
                            atomicAdd(zs + (2 * ia + 1) * ldz + ba, swd2[ja] * iL2);
                        }
                        if (hb && bb >= 0) {
                            atomicAdd(zs + (2 * ib) * ldz + bb, sw2[jb] * iO2);
                            atomicAdd(zs + (2 * ib + 1) * ldz + bb, swd2[jb] * iL2);
                        }
                    });
                } else {
                    // rank rows.  With S00, S01, S10, S11 the bin sums of w_i w_j, w_i w_j dl_j,
                    // w_i dl_i w_j, w_i dl_i w_j dl_j:  S00 = O1 T00, S01 = O1 T01, S10 = L1 T10,
                    // S11 = L1 T11, where T00[b] = sum_j w_j eta21[b][j], T01 = sum_j w_j dl_j eta21,
                    // T10 = sum_j w_j eta31, T11 = sum_j w_j dl_j eta31 were accumulated from the
                    // j chunks.  S_q[A] S_q[B] v_q becomes T_q[A] T_q[B] (v_q times the factors).
                    kc = nrank == 1 ? 4 : 2;
                    const int r0 = (ch - nchunk_j - nchunk_i) * 2;
                    for (int e = tid; e < kc * npad; e += DTHREADS) {
                        const int r = e / npad, b = e - r * npad;
                        zs[r * ldz + b] = T[(r0 + r) * npad + b];
                    }
                    if (tid < kc) {
                        const int q = r0 + tid;
                        vk[tid] = (q < 2 ? O1 : L1) * ((q & 1) ? iL2 : iO2);
                    }
                }
                __syncthreads();
                if (ch < nchunk_j && t0 == 0) {
                    // rank rows from this chunk's eta21, eta31 rows: the chunk's j range is split
                    // into parts so that all threads take a share; parts add into T atomically
                    // (T is this workgroup's own scratch)
                    const int j0 = ch * (kc_max / 2), nj = kc / 2;
                    const int nparts = max(1, min(nj, DTHREADS / nact));
                    const int per = (nj + nparts - 1) / nparts;
                    for (int e = tid; e < nparts * nact; e += DTHREADS) {
                        const int part = e / nact, b = e - part * nact;
                        const int jb = part * per, je = min(nj, jb + per);
                        double t00 = 0.0, t01 = 0.0, t10 = 0.0, t11 = 0.0;
                        for (int jj = jb; jj < je; ++jj) {
                            const double z21 = zs[(2 * jj) * ldz + b], z31 = zs[(2 * jj + 1) * ldz + b];
                            const double wj = sw2[j0 + jj], wdj = swd2[j0 + jj];
                            t00 = fma(wj, z21, t00);
                            t01 = fma(wdj, z21, t01);
                            t10 = fma(wj, z31, t10);
                            t11 = fma(wdj, z31, t11);
                        }
                        if (nparts == 1) {
                            T[b] += t00;
                            T[npad + b] += t01;
                            T[2 * npad + b] += t10;
                            T[3 * npad + b] += t11;
                        } else {
                            unsafeAtomicAdd(T + b, t00);
                            unsafeAtomicAdd(T + npad + b, t01);
                            unsafeAtomicAdd(T + 2 * npad + b, t10);
                            unsafeAtomicAdd(T + 3 * npad + b, t11);
                        }
                    }
                }
                if (ch == nchunk_j + nchunk_i && t0 == 0) {
                    // identity term and weights: D[A][A] += S00[A], wA[A] += S00[A]
                    for (int b = tid; b < nact; b += DTHREADS) {
                        const int A = act[b];
                        const double s00 = O1 * zs[b];
                        unsafeAtomicAdd(a.wA + A, s00);
                        unsafeAtomicAdd(a.D + (int64_t)A * NB + A, s00);
                    }
                }
                // rank-kc update of the wave's blocks: lane l supplies A[l%16][k] = Z[row][k] v[k]
                // and B[k][l%16] = Z[col][k] for k = kk0 + l/16 (A = 0 past the chunk); the
                // operands of the next four rows are loaded before the MFMAs of these four
                const unsigned char* zb = (const unsigned char*)zs;
                auto row_of = [&](int kk0, double& v) -> const unsigned char* {
                    const int k = kk0 + (lane >> 4);
                    const bool kv = k < kc;
                    v = kv ? vk[k] : 0.0;
                    return zb + (uint32_t)((kv ? k : 0) * ldz) * 8u;
                };
                double an[NQ], bn[NQ], vn;
                {
                    const unsigned char* zr = row_of(0, vn);
#pragma unroll
                    for (int q = 0; q < NQ; ++q) {
                        an[q] = *(const double*)(zr + boa[q] * 8u);
                        bn[q] = *(const double*)(zr + bob[q] * 8u);
                    }
                }
                for (int kk0 = 0; kk0 < kc; kk0 += 4) {
                    double av[NQ], bw[NQ];
                    const double v = vn;
#pragma unroll
                    for (int q = 0; q < NQ; ++q) {
                        av[q] = an[q];
                        bw[q] = bn[q];
                    }
                    const unsigned char* zr = row_of(kk0 + 4, vn);
#pragma unroll
                    for (int q = 0; q < NQ; ++q) {
                        an[q] = *(const double*)(zr + boa[q] * 8u);
                        bn[q] = *(const double*)(zr + bob[q] * 8u);
                    }
#pragma unroll
                    for (int q = 0; q < NQ; ++q)
                        acc[q] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[q] * v, bw[q], acc[q], 0, 0, 0);
                }
                __syncthreads();
            }
            // upper triangle into D (act is ascending, so act[row] <= act[col] when row <= col);
            // accumulator element r of lane l is D_pair[l/16 + 4 r][l%16] of the block
#pragma unroll
            for (int q = 0; q < NQ; ++q) {
                if (!bval[q]) continue;
                const int col = (int)bob[q];
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int row = (int)boa[q] - (lane & 15) + (lane >> 4) + 4 * r;
                    if (row <= col && col < nact)
                        unsafeAtomicAdd(a.D + (int64_t)act[row] * NB + act[col], acc[q][r]);
                }
            }
        };
        for (int t0 = 0; t0 < nblk;) {
            // blocks per wave in this pass: all waves take the same count
            const int nq = min(BPW, (nblk - t0 + DWAVES - 1) / DWAVES);
            switch (nq) {
#define LYA_DIST_PASS(n) \
    case n: pass(std::integral_constant<int, n>{}, t0); break;
                LYA_DIST_PASS(1) LYA_DIST_PASS(2) LYA_DIST_PASS(3) LYA_DIST_PASS(4)
                LYA_DIST_PASS(5) LYA_DIST_PASS(6) LYA_DIST_PASS(7) LYA_DIST_PASS(8)
                LYA_DIST_PASS(9) LYA_DIST_PASS(10)
#undef LYA_DIST_PASS
                default: break;
            }
            t0 += DWAVES * nq;
        }
        publish();
    }
}

// The first n words of q (the pair counter, and for the allocation-free entry point
// the skip counters it keeps in its own scratch).
__global__ void zero_words(uint32_t* q, int n) {
    if ((int)threadIdx.x < n) q[threadIdx.x] = 0u;
}

// D[c][r] = D[r][c] for r < c (one thread per element of the lower triangle's rows).
__global__ void mirror_kernel(double* D, int nb) {
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= (int64_t)nb * nb) return;
    const int r = (int)(k / nb), c = (int)(k - (int64_t)r * nb);
    if (c < r) D[k] = D[(int64_t)c * nb + r];
}

}  // namespace

extern "C" int64_t lya_distortion_scratch_bytes(int32_t groups, int32_t max_len, int32_t nbins) {
    if (groups < 1 || max_len < 1 || nbins < 1) return -1;
    const int64_t npad = ((int64_t)nbins + BS - 1) / BS * BS;
    return 256 + (int64_t)groups * 4 * npad * 8;   // pair queue, then the rank rows per workgroup
}

namespace {

// Checks, LDS sizing and launches shared by lya_distortion_run and
// lya_distortion_run_noalloc: `work` holds the pair counter (first 256 bytes) and the
// rank rows (groups x 4 x npad doubles); the first `nzero` words of `work` are cleared
// before the kernel.
int distortion_launch(int32_t n_forest, const int64_t* pix_off, const double* dc, const double* we,
                      const double* dl, const double* ux, const double* uy, const double* uz, const double* omega,
                      const double* odl2, const int32_t* pairs, int32_t n_pairs, int32_t max_len, int32_t np_rp,
                      int32_t np_rt, double rpmax, double rtmax, double* D, double* wA, void* work,
                      int64_t work_bytes, uint32_t* skip, int nzero, int32_t groups, void* stream) {
    if (n_pairs < 0 || n_forest < 0) return lya_fail(LYA_ERR_ARG, "distortion: negative sizes");
    if (!pix_off || !dc || !we || !dl || !ux || !uy || !uz || !omega || !odl2 || !D || !wA || !work ||
        (n_pairs > 0 && !pairs))
        return lya_fail(LYA_ERR_ARG, "distortion: null pointer");
    if (np_rp < 1 || np_rt < 1 || !(rpmax > 0) || !(rtmax > 0)) return lya_fail(LYA_ERR_ARG, "distortion: bad binning");
    if ((int64_t)np_rp * np_rt > 16384) return lya_fail(LYA_ERR_UNSUPPORTED, "distortion: at most 16384 bins");
    const int NB = np_rp * np_rt;
    if (max_len < 1) return lya_fail(LYA_ERR_ARG, "distortion: max_len < 1");
    if ((int64_t)max_len > INT32_MAX / 48) return lya_fail(LYA_ERR_UNSUPPORTED, "distortion: max_len too large");
    if (groups < 1) return lya_fail(LYA_ERR_ARG, "distortion: groups < 1");
    const int64_t need = lya_distortion_scratch_bytes(groups, max_len, NB);
    if (work_bytes < need)
        return lya_fail(LYA_ERR_CAPACITY, "distortion: scratch %lld < %lld bytes", (long long)work_bytes,
                        (long long)need);
    int dev = 0;
    hipDeviceProp_t prop;
    if (hipGetDevice(&dev) != hipSuccess || hipGetDeviceProperties(&prop, dev) != hipSuccess)
        return lya_fail(LYA_ERR_HIP, "no HIP device");
    // the chunk buffer takes the rest of the workgroup's LDS.  It must hold two rows of
    // the active bins of a pair (padded to the block edge): at least one block here,
    // and the kernel skips (and counts) a pair with more active bins than it holds
    const size_t fixed = lds_fixed(NB, max_len) + LDS_STATIC;
    const int64_t npad_max = (int64_t)(NB + BS - 1) / BS * BS;
    const size_t budget = prop.sharedMemPerBlock;
    if (fixed + (size_t)16 * (BS + ZSKEW) > budget)
        return lya_fail(LYA_ERR_UNSUPPORTED, "distortion: %d bins and forests of %d pixels exceed the LDS", NB,
                        max_len);
    const int64_t zcap = (int64_t)((budget - fixed) / 8);
    const int32_t zc = (int32_t)std::min<int64_t>(zcap, (int64_t)(KC_MAX + 4) * (npad_max + ZSKEW));
    const size_t lds = lds_fixed(NB, max_len) + (size_t)8 * zc;
    if (hipFuncSetAttribute((const void*)distortion_kernel, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds) !=
        hipSuccess)
        return lya_fail(LYA_ERR_HIP, "distortion: cannot reserve %zu bytes of LDS", lds);
    DArgs a;
    a.n_forest = n_forest; a.pix_off = pix_off; a.dc = dc; a.we = we; a.dl = dl; a.ux = ux; a.uy = uy; a.uz = uz;
    a.omega = omega; a.odl2 = odl2; a.pairs = pairs; a.n_pairs = n_pairs; a.max_len = max_len;
    a.np_rp = np_rp; a.np_rt = np_rt; a.rpmax = rpmax; a.rtmax = rtmax; a.D = D; a.wA = wA;
    a.zcap = zc;
    a.npad_max = (int32_t)npad_max;
    a.queue = (uint32_t*)work;
    a.skip = skip;
    a.tscr = (double*)((unsigned char*)work + 256);
    hipStream_t st = (hipStream_t)stream;
    hipLaunchKernelGGL(zero_words, dim3(1), dim3(64), 0, st, a.queue, nzero);
    if (n_pairs > 0) hipLaunchKernelGGL(distortion_kernel, dim3(groups), dim3(DTHREADS), lds, st, a);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return lya_fail(LYA_ERR_HIP, "distortion launch: %s", hipGetErrorString(e));
    return LYA_OK;
}

}  // namespace

extern "C" int lya_distortion_run(int32_t n_forest, const int64_t* pix_off, const double* dc, const double* we,
                                  const double* dl, const double* ux, const double* uy, const double* uz,
                                  const double* omega, const double* odl2, const int32_t* pairs, int32_t n_pairs,
                                  int32_t max_len, int32_t np_rp, int32_t np_rt, double rpmax, double rtmax,
                                  double* D, double* wA, void* work, int64_t work_bytes, int32_t groups,
                                  void* stream) {
    // pair counter cleared per call; the skip counters (words 1 and 2) add up until the
    // caller clears them
    return distortion_launch(n_forest, pix_off, dc, we, dl, ux, uy, uz, omega, odl2, pairs, n_pairs, max_len, np_rp,
                             np_rt, rpmax, rtmax, D, wA, work, work_bytes,
                             work ? (uint32_t*)work + LYA_DIST_SKIP_LONG : nullptr, 1, groups, stream);
}

// ---------------------------------------------------------------------------
// Allocation-free variant.  The pair counter and rank rows live in `scratch`, a
// file-scope device array of LYA_DIST_NOALLOC_BYTES (each GPU the library runs on
// holds its own copy until the process ends).  Host code never passes the array's
// name to a kernel or a copy: its device address is looked up once per device with
// hipGetSymbolAddress and kept in g_scratch_ptr.
// ---------------------------------------------------------------------------
static __device__ __align__(256) unsigned char scratch[LYA_DIST_NOALLOC_BYTES];

namespace {
constexpr int MAX_DEVICES = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[MAX_DEVICES];

// Device address of this GPU's `scratch` (looked up on the first call per device).
int noalloc_scratch(void** out) {
    int dev = -1;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0) return lya_fail(LYA_ERR_HIP, "no HIP device");
    if (dev >= MAX_DEVICES) return lya_fail(LYA_ERR_UNSUPPORTED, "device %d: at most %d devices", dev, MAX_DEVICES);
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return lya_fail(LYA_ERR_HIP, "scratch symbol address on device %d: %s", dev, hipGetErrorString(e));
        g_scratch_ptr[dev] = p;
    }
    *out = g_scratch_ptr[dev];
    return LYA_OK;
}
}  // namespace

extern "C" int lya_distortion_run_noalloc(int32_t n_forest, const int64_t* pix_off, const double* dc,
                                          const double* we, const double* dl, const double* ux, const double* uy,
                                          const double* uz, const double* omega, const double* odl2,
                                          const int32_t* pairs, int32_t n_pairs, int32_t max_len, int32_t np_rp,
                                          int32_t np_rt, double rpmax, double rtmax, double* D, double* wA,
                                          uint32_t* skipped, int32_t groups, void* stream) {
    if (groups >= 1 && max_len >= 1 && np_rp >= 1 && np_rt >= 1 && (int64_t)np_rp * np_rt <= 16384) {
        const int64_t need = lya_distortion_scratch_bytes(groups, max_len, np_rp * np_rt);
        if (need > LYA_DIST_NOALLOC_BYTES)
            return lya_fail(LYA_ERR_CAPACITY,
                            "distortion (no allocation): %d workgroups x %d bins need %lld bytes of scratch, more "
                            "than the %lld of the built-in buffer (use fewer workgroups or lya_distortion_run)",
                            groups, np_rp * np_rt, (long long)need, (long long)LYA_DIST_NOALLOC_BYTES);
    }
    void* work = nullptr;
    int rc = noalloc_scratch(&work);
    if (rc) return rc;
    // the pair counter, and the skip counters when the caller keeps none, start at zero
    // on every call
    uint32_t* skip = skipped ? skipped : (uint32_t*)work + LYA_DIST_SKIP_LONG;
    return distortion_launch(n_forest, pix_off, dc, we, dl, ux, uy, uz, omega, odl2, pairs, n_pairs, max_len, np_rp,
                             np_rt, rpmax, rtmax, D, wA, work, LYA_DIST_NOALLOC_BYTES, skip, skipped ? 1 : 3, groups,
                             stream);
}

extern "C" int lya_distortion_mirror(double* D, int32_t nb, void* stream) {
    if (!D || nb < 1) return lya_fail(LYA_ERR_ARG, "distortion mirror: bad arguments");
    const int64_t n = (int64_t)nb * nb;
    hipLaunchKernelGGL(mirror_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, (hipStream_t)stream, D, nb);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return lya_fail(LYA_ERR_HIP, "distortion mirror launch: %s", hipGetErrorString(e));
    return LYA_OK;
}
