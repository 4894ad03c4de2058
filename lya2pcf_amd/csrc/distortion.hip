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
// MI355X organisation (one workgroup per forest pair, persistent over pairs):
//   1. FP64 bins of all pixel pairs (the xi formula) into LDS (u16, j-major);
//      active bins flagged and compacted (wave scan) to b = 0..nact-1.
//   2. eta12/13 (one lane per i walking j) and eta21/31 (one lane per j walking
//      i) accumulated in registers along runs of equal bin and written to a
//      per-workgroup global scratch [nact][n]; eta22..33 reduced per b.
//   3. For each active B: every pixel pair adds
//        w_i w_j [ (K_j + dl_i L_j) + (M_i + dl_j N_i) ]
//      with K_j = eta22 + dl_j eta23 - eta21[j], L_j = eta32 + dl_j eta33 - eta31[j],
//      M_i = -eta12[i], N_i = -eta13[i], into an LDS histogram over compact A;
//      the histogram is flushed once into column B of D (FP64 global atomics).
//   The D update volume per pair is nact^2 instead of n1 n2 nact.
// Forests whose L = 0 (constant loglam) get zero slope terms (the reference
// divides by zero there).
#include <hip/hip_runtime.h>
#include <algorithm>
#include <cmath>
#include <cstdint>

#include "../../include/lya2pcf.h"
#include "errors.h"

#pragma clang fp contract(off)

#define CAS __attribute__((address_space(4)))

namespace {

constexpr int DWAVES = 8;
constexpr int DTHREADS = DWAVES * 64;
constexpr int MAX_PAIRS_LDS = 44000;   // u16 bins of one forest pair held in LDS
constexpr uint16_t NOBIN = 0xFFFF;

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
    unsigned char* scratch;
    int64_t per_group;   // scratch bytes per workgroup
    uint32_t* queue;
};

__global__ void __launch_bounds__(DTHREADS) distortion_kernel(const DArgs a) {
    extern __shared__ __align__(16) unsigned char lds[];
    __shared__ int s_pair, s_nact;
    const int NB = a.np_rp * a.np_rt;
    uint16_t* binc = (uint16_t*)lds;                                   // [MAX_PAIRS_LDS]
    int16_t* cmap = (int16_t*)(lds + 2 * MAX_PAIRS_LDS);               // [NB] bin -> compact, -1
    int16_t* act = cmap + NB;                                          // [NB] compact -> bin
    double* hist = (double*)(lds + ((2 * MAX_PAIRS_LDS + 4 * NB + 15) & ~15));  // [NB]
    const int tid = threadIdx.x, lane = tid & 63, wave = uni(tid >> 6);
    const double bp64 = (double)a.np_rp / a.rpmax, bt64 = (double)a.np_rt / a.rtmax;

    unsigned char* scr = a.scratch + (int64_t)blockIdx.x * a.per_group;
    const int ML = a.max_len;
    double* e12 = (double*)scr;                  // [NB][ML]
    double* e13 = e12 + (int64_t)NB * ML;
    double* e21 = e13 + (int64_t)NB * ML;
    double* e31 = e21 + (int64_t)NB * ML;
    double* e2x = e31 + (int64_t)NB * ML;        // [NB][4]: eta22, eta23, eta32, eta33

    for (;;) {
        if (tid == 0) s_pair = (int)atomicAdd(a.queue, 1u);
        __syncthreads();
        const int pr = uni(s_pair);
        if (pr >= a.n_pairs) break;
        const int f1 = cptr(a.pairs)[2 * pr], f2 = cptr(a.pairs)[2 * pr + 1];
        const int64_t p1 = cptr(a.pix_off)[f1], p2 = cptr(a.pix_off)[f2];
        const int n1 = (int)(cptr(a.pix_off)[f1 + 1] - p1), n2 = (int)(cptr(a.pix_off)[f2 + 1] - p2);
        const double O1 = cptr(a.omega)[f1], O2 = cptr(a.omega)[f2];
        const double L1 = cptr(a.odl2)[f1], L2 = cptr(a.odl2)[f2];
        const double iO1 = 1.0 / O1, iO2 = 1.0 / O2;
        const double iL1 = (L1 != 0.0) ? 1.0 / L1 : 0.0, iL2 = (L2 != 0.0) ? 1.0 / L2 : 0.0;
        double cos12 = cptr(a.ux)[f1] * cptr(a.ux)[f2] + cptr(a.uy)[f1] * cptr(a.uy)[f2];
        cos12 = cos12 + cptr(a.uz)[f1] * cptr(a.uz)[f2];
        cos12 = fmin(1.0, fmax(-1.0, cos12));
        const double c = sqrt(0.5 * (1.0 + cos12)), s = sqrt(0.5 * (1.0 - cos12));
        const int npp = n1 * n2;

        // ---- 1. bins, active flags -------------------------------------------------
        for (int b = tid; b < NB; b += DTHREADS) cmap[b] = -1;
        __syncthreads();
        for (int k = tid; k < npp; k += DTHREADS) {
            const int j = k / n1, i = k - j * n1;
            const double r1 = a.dc[p1 + i], r2 = a.dc[p2 + j];
            const uint32_t bp = (uint32_t)__double2int_rz((fabs(r1 - r2) * c) * bp64);
            const uint32_t bt = (uint32_t)__double2int_rz(((r1 + r2) * s) * bt64);
            uint16_t A = NOBIN;
            if (bp < (uint32_t)a.np_rp && bt < (uint32_t)a.np_rt) {
                A = (uint16_t)(bp * a.np_rt + bt);
                cmap[A] = 1;
            }
            binc[k] = A;
        }
        __syncthreads();
        // compaction in bin order (wave 0)
        if (wave == 0) {
            const int per = (NB + 63) / 64;
            const int b0 = lane * per, b1 = min(NB, b0 + per);
            int cnt = 0;
            for (int b = b0; b < b1; ++b) cnt += (cmap[b] >= 0);
            int incl = cnt;
            for (int o = 1; o < 64; o <<= 1) {
                const int v = __shfl_up(incl, o);
                if (lane >= o) incl += v;
            }
            int idx = incl - cnt;
            for (int b = b0; b < b1; ++b) {
                if (cmap[b] >= 0) {
                    cmap[b] = (int16_t)idx;
                    act[idx] = (int16_t)b;
                    ++idx;
                }
            }
            if (lane == 63) s_nact = incl;
        }
        __syncthreads();
        const int nact = uni(s_nact);
        for (int k = tid; k < npp; k += DTHREADS) {
            const uint16_t A = binc[k];
            binc[k] = (A == NOBIN) ? NOBIN : (uint16_t)cmap[A];
        }
        // zero eta tables of this pair
        for (int64_t k = tid; k < (int64_t)nact * ML; k += DTHREADS) {
            e12[k] = 0.0; e13[k] = 0.0; e21[k] = 0.0; e31[k] = 0.0;
        }
        __syncthreads();

        // ---- 2. eta tables ---------------------------------------------------------
        if (tid < n1) {
            const int i = tid;
            int cur = -1;
            double s12 = 0.0, s13 = 0.0;
            for (int j = 0; j < n2; ++j) {
                const int b = binc[j * n1 + i];
                if (b == NOBIN) continue;
                if (b != cur) {
                    if (cur >= 0) { e12[(int64_t)cur * ML + i] += s12; e13[(int64_t)cur * ML + i] += s13; }
                    cur = b; s12 = 0.0; s13 = 0.0;
                }
                const double wj = cptr(a.we)[p2 + j];
                s12 += wj * iO2;
                s13 += wj * cptr(a.dl)[p2 + j] * iL2;
            }
            if (cur >= 0) { e12[(int64_t)cur * ML + i] += s12; e13[(int64_t)cur * ML + i] += s13; }
        } else if (tid >= 256 && tid - 256 < n2) {
            const int j = tid - 256;
            int cur = -1;
            double s21 = 0.0, s31 = 0.0;
            for (int i = 0; i < n1; ++i) {
                const int b = binc[j * n1 + i];
                if (b == NOBIN) continue;
                if (b != cur) {
                    if (cur >= 0) { e21[(int64_t)cur * ML + j] += s21; e31[(int64_t)cur * ML + j] += s31; }
                    cur = b; s21 = 0.0; s31 = 0.0;
                }
                const double wi = cptr(a.we)[p1 + i];
                s21 += wi * iO1;
                s31 += wi * cptr(a.dl)[p1 + i] * iL1;
            }
            if (cur >= 0) { e21[(int64_t)cur * ML + j] += s21; e31[(int64_t)cur * ML + j] += s31; }
        }
        __syncthreads();
        // eta22..33 per active bin; identity term and weights
        for (int b = tid; b < nact; b += DTHREADS) {
            double t22 = 0.0, t23 = 0.0, t32 = 0.0, t33 = 0.0;
            for (int i = 0; i < n1; ++i) {
                const double wi = cptr(a.we)[p1 + i];
                const double a1 = wi * iO1, b1 = wi * cptr(a.dl)[p1 + i] * iL1;
                const double x12 = e12[(int64_t)b * ML + i], x13 = e13[(int64_t)b * ML + i];
                t22 += a1 * x12; t23 += a1 * x13; t32 += b1 * x12; t33 += b1 * x13;
            }
            e2x[4 * b + 0] = t22; e2x[4 * b + 1] = t23; e2x[4 * b + 2] = t32; e2x[4 * b + 3] = t33;
        }
        __syncthreads();
        // identity D[A][A] and wA[A]: sum over pairs in A of w_i w_j
        for (int b = tid; b < nact; b += DTHREADS) hist[b] = 0.0;
        __syncthreads();
        for (int k = tid; k < npp; k += DTHREADS) {
            const uint16_t b = binc[k];
            if (b == NOBIN) continue;
            const int j = k / n1, i = k - j * n1;
            atomicAdd(&hist[b], cptr(a.we)[p1 + i] * a.we[p2 + j]);
        }
        __syncthreads();
        for (int b = tid; b < nact; b += DTHREADS) {
            const int A = act[b];
            unsafeAtomicAdd(a.wA + A, hist[b]);
            unsafeAtomicAdd(a.D + (int64_t)A * NB + A, hist[b]);
        }
        __syncthreads();

        // ---- 3. per active B: histogram over A, flushed into column B -------------
        const int nslab = (n1 + 63) / 64;
        const int slab = wave % nslab;
        const int kw = wave / nslab;
        const int mw = (DWAVES - slab + nslab - 1) / nslab;  // waves on this slab
        const bool haswork = kw < mw;
        const int i = slab * 64 + lane;
        const bool vi = haswork && i < n1;
        const double wi = vi ? a.we[p1 + i] : 0.0;
        const double dli = vi ? a.dl[p1 + i] : 0.0;
        for (int b = 0; b < nact; ++b) {
            for (int q = tid; q < nact; q += DTHREADS) hist[q] = 0.0;
            __syncthreads();
            const double h22 = e2x[4 * b + 0], h23 = e2x[4 * b + 1], h32 = e2x[4 * b + 2], h33 = e2x[4 * b + 3];
            if (haswork) {
                const double Mi = vi ? -e12[(int64_t)b * ML + i] : 0.0;
                const double Ni = vi ? -e13[(int64_t)b * ML + i] : 0.0;
                for (int j = kw; j < n2; j += mw) {
                    const double wj = cptr(a.we)[p2 + j], dlj = cptr(a.dl)[p2 + j];
                    const double Kj = (h22 + dlj * h23) - e21[(int64_t)b * ML + j];
                    const double Lj = (h32 + dlj * h33) - e31[(int64_t)b * ML + j];
                    if (vi) {
                        const uint16_t A = binc[j * n1 + i];
                        if (A != NOBIN) {
                            const double t = (Kj + dli * Lj) + (Mi + dlj * Ni);
                            atomicAdd(&hist[A], (wi * wj) * t);
                        }
                    }
                }
            }
            __syncthreads();
            const int B = act[b];
            for (int q = tid; q < nact; q += DTHREADS) {
                const double v = hist[q];
                if (v != 0.0) unsafeAtomicAdd(a.D + (int64_t)act[q] * NB + B, v);
            }
            __syncthreads();
        }
    }
}

__global__ void zero_queue(uint32_t* q) {
    if (threadIdx.x < 4) q[threadIdx.x] = 0u;
}

int64_t per_group_bytes(int32_t max_len, int32_t nbins) {
    return ((int64_t)4 * nbins * max_len + 4LL * nbins) * 8 + 256;
}

}  // namespace

extern "C" int64_t lya_distortion_scratch_bytes(int32_t groups, int32_t max_len, int32_t nbins) {
    if (groups < 1 || max_len < 1 || nbins < 1) return -1;
    return (int64_t)groups * per_group_bytes(max_len, nbins) + 64;
}

extern "C" int lya_distortion_run(int32_t n_forest, const int64_t* pix_off, const double* dc, const double* we,
                                  const double* dl, const double* ux, const double* uy, const double* uz,
                                  const double* omega, const double* odl2, const int32_t* pairs, int32_t n_pairs,
                                  int32_t max_len, int32_t np_rp, int32_t np_rt, double rpmax, double rtmax,
                                  double* D, double* wA, void* scratch, int64_t scratch_bytes, int32_t groups,
                                  void* stream) {
    if (n_pairs < 0 || n_forest < 0) return lya_fail(LYA_ERR_ARG, "distortion: negative sizes");
    if (!pix_off || !dc || !we || !dl || !ux || !uy || !uz || !omega || !odl2 || !D || !wA || !scratch ||
        (n_pairs > 0 && !pairs))
        return lya_fail(LYA_ERR_ARG, "distortion: null pointer");
    if (np_rp < 1 || np_rt < 1 || !(rpmax > 0) || !(rtmax > 0)) return lya_fail(LYA_ERR_ARG, "distortion: bad binning");
    const int NB = np_rp * np_rt;
    if (NB > 16384) return lya_fail(LYA_ERR_UNSUPPORTED, "distortion: at most 16384 bins");
    if (max_len < 1 || (int64_t)max_len * max_len > MAX_PAIRS_LDS)
        return lya_fail(LYA_ERR_UNSUPPORTED, "distortion: max_len^2 = %lld exceeds %d pixel pairs per forest pair",
                        (long long)max_len * max_len, MAX_PAIRS_LDS);
    if (max_len > 256) return lya_fail(LYA_ERR_UNSUPPORTED, "distortion: forests longer than 256 pixels");
    if (groups < 1) return lya_fail(LYA_ERR_ARG, "distortion: groups < 1");
    const int64_t need = lya_distortion_scratch_bytes(groups, max_len, NB);
    if (scratch_bytes < need)
        return lya_fail(LYA_ERR_CAPACITY, "distortion: scratch %lld < %lld bytes", (long long)scratch_bytes,
                        (long long)need);
    const size_t lds = ((2 * MAX_PAIRS_LDS + 4 * NB + 15) & ~15) + (size_t)8 * NB;
    int dev = 0;
    hipDeviceProp_t prop;
    if (hipGetDevice(&dev) != hipSuccess || hipGetDeviceProperties(&prop, dev) != hipSuccess)
        return lya_fail(LYA_ERR_HIP, "no HIP device");
    if (lds > prop.sharedMemPerBlock) return lya_fail(LYA_ERR_UNSUPPORTED, "distortion: LDS %zu too large", lds);
    (void)hipFuncSetAttribute((const void*)distortion_kernel, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    DArgs a;
    a.n_forest = n_forest; a.pix_off = pix_off; a.dc = dc; a.we = we; a.dl = dl; a.ux = ux; a.uy = uy; a.uz = uz;
    a.omega = omega; a.odl2 = odl2; a.pairs = pairs; a.n_pairs = n_pairs; a.max_len = max_len;
    a.np_rp = np_rp; a.np_rt = np_rt; a.rpmax = rpmax; a.rtmax = rtmax; a.D = D; a.wA = wA;
    a.per_group = per_group_bytes(max_len, NB);
    a.queue = (uint32_t*)((unsigned char*)scratch + (int64_t)groups * a.per_group);
    a.scratch = (unsigned char*)scratch;
    hipStream_t st = (hipStream_t)stream;
    hipLaunchKernelGGL(zero_queue, dim3(1), dim3(64), 0, st, a.queue);
    if (n_pairs > 0) hipLaunchKernelGGL(distortion_kernel, dim3(groups), dim3(DTHREADS), lds, st, a);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return lya_fail(LYA_ERR_HIP, "distortion launch: %s", hipGetErrorString(e));
    return LYA_OK;
}
