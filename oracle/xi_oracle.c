/*
 * CPU oracle for the lya2pcf pair-binning path.  TEST INFRASTRUCTURE ONLY:
 * only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load
 * this library, as the checker or as the timed CPU baseline; the product path
 * never calls it.
 *
 * FP64 restatements of the reference, compiled with -ffp-contract=off so every
 * expression is evaluated in IEEE double exactly as written:
 *
 *   oracle_pairs            forest_class.py:78-98 (neighbour rule of :96), brute force O(N^2)
 *   oracle_xi_gpu_form      cuda_kernels.cpp:112-156 (half-angles via sqrt, floor(rp*NP/rpmax))
 *   oracle_xi_cpu_form      correlation_procedures_cpu.py:19-84 (angle from ra/dec, arccos,
 *                           small-angle branch, theta < angmax gate, int(rp/rpmax*NP), per-pair
 *                           fresh histograms added in pair order)
 *   oracle_distortion_eta   cuda_kernels.cpp:165-278 eta algebra, FP64, dense tables
 *   oracle_distortion_direct  definition D_AB = sum_{(i,j) in A} w_i w_j sum_{(k,l) in B} eta_ik eta_jl
 *   oracle_xi_adjoint       gradient of chi^2 pieces: g_dw[i] += alpha[A] dw[j] (both members)
 *
 * Pair counts per bin are an addition of this build (the reference keeps only
 * sums of w and dw).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define EXPORT __attribute__((visibility("default")))

/* ---------------------------------------------------------------- pairs */
EXPORT int64_t oracle_pairs(int32_t nf, const double* ux, const double* uy, const double* uz, const double* ra,
                            const int64_t* name, double cos_max, int64_t* nb_off, int32_t* nb_idx) {
    int64_t k = 0;
    nb_off[0] = 0;
    for (int32_t i = 0; i < nf; ++i) {
        for (int32_t j = 0; j < nf; ++j) {
            double mu = ux[i] * ux[j] + uy[i] * uy[j];
            mu = mu + uz[i] * uz[j];
            if (mu > cos_max && name[i] != name[j] && ra[i] < ra[j]) {
                if (nb_idx) nb_idx[k] = j;
                ++k;
            }
        }
        nb_off[i + 1] = k;
    }
    return k;
}

/* ------------------------------------------------------- xi, GPU form */
static void halfangles(double x1, double y1, double z1, double x2, double y2, double z2, double* c, double* s) {
    double cos12 = x1 * x2 + y1 * y2;
    cos12 = cos12 + z1 * z2;
    if (cos12 > 1.0) cos12 = 1.0;
    if (cos12 < -1.0) cos12 = -1.0;
    *c = sqrt(0.5 * (1.0 + cos12));
    *s = sqrt(0.5 * (1.0 - cos12));
}

/*
 * f1_list[k] (k < n_f1) are processed as forest 1 into slot f1_slot[k];
 * outputs cnt/w/dw are [n_slots][np*nt], accumulated (caller zeroes).
 * Loop order: f1 list order, neighbour order, i major, j minor.
 */
EXPORT void oracle_xi_gpu_form(const int64_t* pix_off, const double* dc, const double* we, const double* dw,
                               const double* ux, const double* uy, const double* uz, const int64_t* nb_off,
                               const int32_t* nb_idx, const int32_t* f1_list, const int32_t* f1_slot, int32_t n_f1,
                               int32_t np, int32_t nt, double rpmax, double rtmax, uint64_t* cnt, double* w,
                               double* dwh) {
    const double bp = (double)np / rpmax, bt = (double)nt / rtmax;
    const int64_t nb = (int64_t)np * nt;
    for (int32_t k = 0; k < n_f1; ++k) {
        const int32_t f1 = f1_list[k];
        const int64_t o = (int64_t)f1_slot[k] * nb;
        for (int64_t q = nb_off[f1]; q < nb_off[f1 + 1]; ++q) {
            const int32_t f2 = nb_idx[q];
            double c, s;
            halfangles(ux[f1], uy[f1], uz[f1], ux[f2], uy[f2], uz[f2], &c, &s);
            for (int64_t i = pix_off[f1]; i < pix_off[f1 + 1]; ++i) {
                for (int64_t j = pix_off[f2]; j < pix_off[f2 + 1]; ++j) {
                    const double rp = fabs(dc[i] - dc[j]) * c;
                    const double rt = (dc[i] + dc[j]) * s;
                    const double qp = floor(rp * bp), qt = floor(rt * bt);
                    if (qp < np && qt < nt) {
                        const int64_t b = o + (int64_t)qp * nt + (int64_t)qt;
                        cnt[b] += 1;
                        w[b] += we[i] * we[j];
                        dwh[b] += dw[i] * dw[j];
                    }
                }
            }
        }
    }
}

/* ------------------------------------------------------- xi, CPU form */
EXPORT void oracle_xi_cpu_form(const int64_t* pix_off, const double* dc, const double* we, const double* dw,
                               const double* ra, const double* dec, const int64_t* nb_off, const int32_t* nb_idx,
                               const int32_t* f1_list, const int32_t* f1_slot, int32_t n_f1, int32_t np, int32_t nt,
                               double rpmax, double rtmax, double angmax, double chiquito, uint64_t* cnt, double* w,
                               double* dwh) {
    const int64_t nb = (int64_t)np * nt;
    double* tw = (double*)calloc((size_t)nb, sizeof(double));
    double* td = (double*)calloc((size_t)nb, sizeof(double));
    uint64_t* tc = (uint64_t*)calloc((size_t)nb, sizeof(uint64_t));
    unsigned char* touched = (unsigned char*)calloc((size_t)nb, 1);
    int64_t* list = (int64_t*)malloc(sizeof(int64_t) * (size_t)nb);
    for (int32_t k = 0; k < n_f1; ++k) {
        const int32_t f1 = f1_list[k];
        const int64_t o = (int64_t)f1_slot[k] * nb;
        for (int64_t q = nb_off[f1]; q < nb_off[f1 + 1]; ++q) {
            const int32_t f2 = nb_idx[q];
            double th;
            if (fabs(ra[f1] - ra[f2]) < chiquito && fabs(dec[f1] - dec[f2]) < chiquito) {
                const double a = (ra[f1] - ra[f2]) * cos(dec[f1]);
                const double b = dec[f1] - dec[f2];
                th = sqrt(a * a + b * b);
            } else {
                double x = sin(dec[f1]) * sin(dec[f2]) + cos(dec[f1]) * cos(dec[f2]) * cos(ra[f1] - ra[f2]);
                th = acos(x);
            }
            const double c = cos(0.5 * th), s = sin(0.5 * th);
            if (!(th < angmax)) continue;
            int64_t nl = 0;
            for (int64_t i = pix_off[f1]; i < pix_off[f1 + 1]; ++i) {
                for (int64_t j = pix_off[f2]; j < pix_off[f2 + 1]; ++j) {
                    const double rp = fabs(dc[i] - dc[j]) * c;
                    const double rt = (dc[i] + dc[j]) * s;
                    const double xp = rp / rpmax * (double)np;
                    const double xt = rt / rtmax * (double)nt;
                    /* int() truncates toward zero; both are >= 0 */
                    const double qp = trunc(xp), qt = trunc(xt);
                    if (qp < np && qt < nt) {
                        const int64_t b = (int64_t)qp * nt + (int64_t)qt;
                        if (!touched[b]) { touched[b] = 1; list[nl++] = b; }
                        tc[b] += 1;
                        tw[b] += we[i] * we[j];
                        td[b] += dw[i] * dw[j];
                    }
                }
            }
            for (int64_t m = 0; m < nl; ++m) {
                const int64_t b = list[m];
                cnt[o + b] += tc[b];
                w[o + b] += tw[b];
                dwh[o + b] += td[b];
                tc[b] = 0; tw[b] = 0.0; td[b] = 0.0; touched[b] = 0;
            }
        }
    }
    free(tw); free(td); free(tc); free(touched); free(list);
}

/* ------------------------------------------------------- distortion */
static int64_t bin_of(double r1, double r2, double c, double s, double bp, double bt, int32_t np, int32_t nt) {
    const double qp = floor((fabs(r1 - r2) * c) * bp), qt = floor(((r1 + r2) * s) * bt);
    if (qp < np && qt < nt) return (int64_t)qp * nt + (int64_t)qt;
    return -1;
}

/* eta algebra of cuda_kernels.cpp:165-278 for a list of forest pairs, FP64. */
EXPORT void oracle_distortion_eta(const int64_t* pix_off, const double* dc, const double* we, const double* dl,
                                  const double* ux, const double* uy, const double* uz, const double* omega,
                                  const double* odl2, const int32_t* pairs, int32_t n_pairs, int32_t np, int32_t nt,
                                  double rpmax, double rtmax, double* D, double* wA) {
    const double bp = (double)np / rpmax, bt = (double)nt / rtmax;
    const int64_t NB = (int64_t)np * nt;
    for (int32_t q = 0; q < n_pairs; ++q) {
        const int32_t f1 = pairs[2 * q], f2 = pairs[2 * q + 1];
        const int64_t p1 = pix_off[f1], p2 = pix_off[f2];
        const int64_t n1 = pix_off[f1 + 1] - p1, n2 = pix_off[f2 + 1] - p2;
        double c, s;
        halfangles(ux[f1], uy[f1], uz[f1], ux[f2], uy[f2], uz[f2], &c, &s);
        const double O1 = omega[f1], O2 = omega[f2], L1 = odl2[f1], L2 = odl2[f2];
        const double iL1 = L1 != 0.0 ? 1.0 / L1 : 0.0, iL2 = L2 != 0.0 ? 1.0 / L2 : 0.0;
        double* e12 = (double*)calloc((size_t)(NB * n1), sizeof(double));
        double* e13 = (double*)calloc((size_t)(NB * n1), sizeof(double));
        double* e21 = (double*)calloc((size_t)(NB * n2), sizeof(double));
        double* e31 = (double*)calloc((size_t)(NB * n2), sizeof(double));
        double* e22 = (double*)calloc((size_t)NB, sizeof(double));
        double* e23 = (double*)calloc((size_t)NB, sizeof(double));
        double* e32 = (double*)calloc((size_t)NB, sizeof(double));
        double* e33 = (double*)calloc((size_t)NB, sizeof(double));
        unsigned char* active = (unsigned char*)calloc((size_t)NB, 1);
        for (int64_t i = 0; i < n1; ++i)
            for (int64_t j = 0; j < n2; ++j) {
                const int64_t B = bin_of(dc[p1 + i], dc[p2 + j], c, s, bp, bt, np, nt);
                if (B < 0) continue;
                const double wi = we[p1 + i], wj = we[p2 + j], li = dl[p1 + i], lj = dl[p2 + j];
                active[B] = 1;
                e12[B * n1 + i] += wj / O2;
                e21[B * n2 + j] += wi / O1;
                e13[B * n1 + i] += wj * lj * iL2;
                e31[B * n2 + j] += wi * li * iL1;
                e22[B] += wi * wj / (O1 * O2);
                e23[B] += wi * wj * lj * iL2 / O1;
                e32[B] += wi * li * wj * iL1 / O2;
                e33[B] += wi * li * wj * lj * iL1 * iL2;
                wA[B] += wi * wj;
            }
        for (int64_t i = 0; i < n1; ++i)
            for (int64_t j = 0; j < n2; ++j) {
                const int64_t A = bin_of(dc[p1 + i], dc[p2 + j], c, s, bp, bt, np, nt);
                if (A < 0) continue;
                const double wij = we[p1 + i] * we[p2 + j], li = dl[p1 + i], lj = dl[p2 + j];
                D[A * NB + A] += wij;
                for (int64_t B = 0; B < NB; ++B) {
                    if (!active[B]) continue;
                    D[A * NB + B] += wij * (li * lj * e33[B] + li * e32[B] + lj * e23[B] + e22[B] -
                                            li * e31[B * n2 + j] - lj * e13[B * n1 + i] - e21[B * n2 + j] -
                                            e12[B * n1 + i]);
                }
            }
        free(e12); free(e13); free(e21); free(e31); free(e22); free(e23); free(e32); free(e33); free(active);
    }
}

/* Definition: eta_ik = delta_ik - w_k/O - dl_i dl_k w_k / L (projection of forest_class.py:54-66). */
EXPORT void oracle_distortion_direct(const int64_t* pix_off, const double* dc, const double* we, const double* dl,
                                     const double* ux, const double* uy, const double* uz, const double* omega,
                                     const double* odl2, const int32_t* pairs, int32_t n_pairs, int32_t np,
                                     int32_t nt, double rpmax, double rtmax, double* D, double* wA) {
    const double bp = (double)np / rpmax, bt = (double)nt / rtmax;
    const int64_t NB = (int64_t)np * nt;
    for (int32_t q = 0; q < n_pairs; ++q) {
        const int32_t f1 = pairs[2 * q], f2 = pairs[2 * q + 1];
        const int64_t p1 = pix_off[f1], p2 = pix_off[f2];
        const int64_t n1 = pix_off[f1 + 1] - p1, n2 = pix_off[f2 + 1] - p2;
        double c, s;
        halfangles(ux[f1], uy[f1], uz[f1], ux[f2], uy[f2], uz[f2], &c, &s);
        const double O1 = omega[f1], O2 = omega[f2];
        const double iL1 = odl2[f1] != 0.0 ? 1.0 / odl2[f1] : 0.0, iL2 = odl2[f2] != 0.0 ? 1.0 / odl2[f2] : 0.0;
        for (int64_t i = 0; i < n1; ++i)
            for (int64_t j = 0; j < n2; ++j) {
                const int64_t A = bin_of(dc[p1 + i], dc[p2 + j], c, s, bp, bt, np, nt);
                if (A < 0) continue;
                const double wij = we[p1 + i] * we[p2 + j];
                wA[A] += wij;
                for (int64_t k = 0; k < n1; ++k)
                    for (int64_t l = 0; l < n2; ++l) {
                        const int64_t B = bin_of(dc[p1 + k], dc[p2 + l], c, s, bp, bt, np, nt);
                        if (B < 0) continue;
                        const double eik = (i == k) - we[p1 + k] / O1 - dl[p1 + i] * dl[p1 + k] * we[p1 + k] * iL1;
                        const double ejl = (j == l) - we[p2 + l] / O2 - dl[p2 + j] * dl[p2 + l] * we[p2 + l] * iL2;
                        D[A * NB + B] += wij * eik * ejl;
                    }
            }
    }
}

/* ------------------------------------------------------- adjoint */
/* g_dw[i] += alpha[A] dw[j], g_w[i] += beta[A] w[j] for both members of every pair. */
EXPORT void oracle_xi_adjoint(const int64_t* pix_off, const double* dc, const double* we, const double* dw,
                              const double* ux, const double* uy, const double* uz, const int64_t* nb_off,
                              const int32_t* nb_idx, const int32_t* f1_list, int32_t n_f1, int32_t np, int32_t nt,
                              double rpmax, double rtmax, const double* alpha, const double* beta, double* g_dw,
                              double* g_w) {
    const double bp = (double)np / rpmax, bt = (double)nt / rtmax;
    for (int32_t k = 0; k < n_f1; ++k) {
        const int32_t f1 = f1_list[k];
        for (int64_t q = nb_off[f1]; q < nb_off[f1 + 1]; ++q) {
            const int32_t f2 = nb_idx[q];
            double c, s;
            halfangles(ux[f1], uy[f1], uz[f1], ux[f2], uy[f2], uz[f2], &c, &s);
            for (int64_t i = pix_off[f1]; i < pix_off[f1 + 1]; ++i)
                for (int64_t j = pix_off[f2]; j < pix_off[f2 + 1]; ++j) {
                    const int64_t A = bin_of(dc[i], dc[j], c, s, bp, bt, np, nt);
                    if (A < 0) continue;
                    g_dw[i] += alpha[A] * dw[j];
                    g_dw[j] += alpha[A] * dw[i];
                    g_w[i] += beta[A] * we[j];
                    g_w[j] += beta[A] * we[i];
                }
        }
    }
}

/* ------------------------------------------------ multi-core CPU baseline */
#ifdef _OPENMP
#include <omp.h>
#endif
/* The CPU form over the f1 list with `threads` OpenMP threads, each with
 * private histograms summed at the end (used only to time the CPU baseline;
 * per-bin sums differ from the sequential form only in summation order). */
EXPORT void oracle_xi_cpu_form_mt(const int64_t* pix_off, const double* dc, const double* we, const double* dw,
                                  const double* ra, const double* dec, const int64_t* nb_off, const int32_t* nb_idx,
                                  const int32_t* f1_list, const int32_t* f1_slot, int32_t n_f1, int32_t np, int32_t nt,
                                  double rpmax, double rtmax, double angmax, double chiquito, int32_t n_slots,
                                  int32_t threads, uint64_t* cnt, double* w, double* dwh) {
    const int64_t tot = (int64_t)n_slots * np * nt;
#ifdef _OPENMP
    if (threads < 1) threads = omp_get_max_threads();
#pragma omp parallel num_threads(threads)
    {
        const int t = omp_get_thread_num(), T = omp_get_num_threads();
#else
    {
        const int t = 0, T = 1;
        (void)threads;
#endif
        uint64_t* c = (uint64_t*)calloc((size_t)tot, sizeof(uint64_t));
        double* a = (double*)calloc((size_t)tot, sizeof(double));
        double* b = (double*)calloc((size_t)tot, sizeof(double));
        for (int32_t k = t; k < n_f1; k += T)
            oracle_xi_cpu_form(pix_off, dc, we, dw, ra, dec, nb_off, nb_idx, f1_list + k, f1_slot + k, 1, np, nt,
                               rpmax, rtmax, angmax, chiquito, c, a, b);
#ifdef _OPENMP
#pragma omp critical
#endif
        for (int64_t m = 0; m < tot; ++m) {
            cnt[m] += c[m];
            w[m] += a[m];
            dwh[m] += b[m];
        }
        free(c); free(a); free(b);
    }
}

/* ------------------------------------------------ multi-core GPU form */
/* oracle_xi_gpu_form over the f1 list with `threads` OpenMP threads (each forest 1
 * by one thread, into private histograms summed at the end).  Bins and counts equal
 * the sequential form; sums differ only in summation order.  Used to check the
 * full-size workloads, where the sequential form would take minutes. */
EXPORT void oracle_xi_gpu_form_mt(const int64_t* pix_off, const double* dc, const double* we, const double* dw,
                                  const double* ux, const double* uy, const double* uz, const int64_t* nb_off,
                                  const int32_t* nb_idx, const int32_t* f1_list, const int32_t* f1_slot, int32_t n_f1,
                                  int32_t np, int32_t nt, double rpmax, double rtmax, int32_t n_slots, int32_t threads,
                                  uint64_t* cnt, double* w, double* dwh) {
    const int64_t tot = (int64_t)n_slots * np * nt;
#ifdef _OPENMP
    if (threads < 1) threads = omp_get_max_threads();
#pragma omp parallel num_threads(threads)
    {
        const int t = omp_get_thread_num(), T = omp_get_num_threads();
#else
    {
        const int t = 0, T = 1;
        (void)threads;
#endif
        uint64_t* c = (uint64_t*)calloc((size_t)tot, sizeof(uint64_t));
        double* a = (double*)calloc((size_t)tot, sizeof(double));
        double* b = (double*)calloc((size_t)tot, sizeof(double));
        for (int32_t k = t; k < n_f1; k += T)
            oracle_xi_gpu_form(pix_off, dc, we, dw, ux, uy, uz, nb_off, nb_idx, f1_list + k, f1_slot + k, 1, np, nt,
                               rpmax, rtmax, c, a, b);
#ifdef _OPENMP
#pragma omp critical
#endif
        for (int64_t m = 0; m < tot; ++m) {
            cnt[m] += c[m];
            w[m] += a[m];
            dwh[m] += b[m];
        }
        free(c); free(a); free(b);
    }
}
