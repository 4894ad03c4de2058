/*
 * lya2pcf MI355X C ABI.
 *
 * Plain C, plain pointers and sizes.  Device buffers are caller-owned
 * (allocated by the caller, e.g. as PyTorch-ROCm tensors) and passed as
 * pointers; the library allocates no output.  Every function returns 0 on
 * success or a negative LYA_ERR_* code, with a message in lya_last_error().
 * No C++ exception crosses this boundary.  `stream` is a hipStream_t passed
 * as void* (NULL = the default stream).  Calls on one lya_xi_problem are not
 * thread-safe.
 *
 * Reference interfaces replaced (Rafael9104/lya2pcf):
 *   lya_build_pairs     <- quasar.neighborhood          forest_class.py:78-98
 *   lya_build_pairs_device  (the same, on the GPU)
 *   lya_build_work      <- per-forest1 launch loop      correlation_procedures_pycuda.py:141-166
 *   lya_xi_run          <- pair_correlation kernel      cuda_kernels.cpp:90-162, launched
 *                          once per forest1 with grid (n1,1,1), block (1,16,16)
 *                          (correlation_procedures_pycuda.py:157-163)
 *   lya_xi_graph_*      <- the same launch sequence, captured once
 *   lya_xi_adjoint      <- (no counterpart: gradient of the pair sums)
 *   lya_distortion_run  <- precompute_distances/compute_etas/order_active/compute_d
 *                          cuda_kernels.cpp:46-87,165-278 driven by
 *                          distortion_procedures_pycuda.py:195-279
 */
#ifndef LYA2PCF_H
#define LYA2PCF_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define LYA_OK 0
#define LYA_ERR_ARG (-1)
#define LYA_ERR_HIP (-2)
#define LYA_ERR_CAPACITY (-3)
#define LYA_ERR_UNSUPPORTED (-4)

#define LYA_MODE_FP64 0   /* plain FP64 pair loop, FP64 LDS histograms */
#define LYA_MODE_MIXED 1  /* FP32 geometry (FP32 dc offsets) + FP64 near-edge recheck; w, dw read as
                             FP64, products and LDS sums in FP64 */
#define LYA_MODE_MIXED_F32 2  /* FP32 geometry + recheck, FP32 w, dw, products and LDS partials
                                 -> FP64 merge */

/* Entries (numpix_rp * row stride) of the LDS histogram of one workgroup: 50 x 51 for the
   default binning.  At 16 bytes per entry four workgroups' histograms fit a CU's 160 KiB of
   LDS.  Larger binnings run in bands of bins that fit it (lya_xi_run). */
#define LYA_MAX_BINS 2550

const char* lya_last_error(void);
int lya_version(void);

/*
 * Forest-pair list (host, OpenMP).  Forest f2 is a neighbour of f1 when
 *   (x1*x2 + y1*y2) + z1*z2 > cos_max   (FP64, no fused multiply-add),
 *   name1 != name2  and  ra1 < ra2,
 * the rule of forest_class.py:96.  Output is CSR over f1: neighbours of f1
 * are nb_idx[nb_off[f1] .. nb_off[f1+1]), in ascending f2 order.
 * Call with nb_idx == NULL to get the counts (nb_off filled, *n_pairs set);
 * then call again with nb_idx of capacity >= *n_pairs.
 */
int lya_build_pairs(int32_t n_forest, const double* ux, const double* uy, const double* uz,
                    const double* ra, const int64_t* name, double cos_max,
                    int64_t* nb_off, int32_t* nb_idx, int64_t nb_capacity, int64_t* n_pairs);

/*
 * Work decomposition (host).  Forests f1 listed in f1_list (grouped by
 * output slot f1_slot) are cut into items {f1, nb_begin, nb_end, slot} of at
 * most item_cost pixel pairs (at least one neighbour each); consecutive items
 * of one slot are grouped into tiles {item_begin, item_end} of at most
 * tile_cost, and tiles are ordered by decreasing cost.
 * Call with items == NULL to get *n_items and *n_tiles; then with buffers of
 * 4*n_items and 2*n_tiles int32.  *total_cost = sum of n1*n2 over pairs.
 */
int lya_build_work(int32_t n_f1, const int32_t* f1_list, const int32_t* f1_slot,
                   const int64_t* pix_off, const int64_t* nb_off, const int32_t* nb_idx,
                   int64_t item_cost, int64_t tile_cost,
                   int32_t* items, int32_t* tiles, int32_t* n_items, int32_t* n_tiles,
                   double* total_cost);

/*
 * Tiles from an item list (host): the tile step of lya_build_work applied to
 * items {f1, nb_begin, nb_end, slot} with their costs (pixel pairs), e.g. the
 * output of lya_build_items_device.  Call with tiles == NULL for *n_tiles.
 * Replaces the same part of the per-forest1 launch loop as lya_build_work.
 */
int lya_build_tiles(int64_t n_items, const int32_t* items, const int64_t* item_cost,
                    int64_t tile_cost, int32_t* tiles, int32_t* n_tiles);

/*
 * Forest-pair list on the GPU: the rule and the output of lya_build_pairs
 * (identical arrays), for quasar.neighborhood (forest_class.py:78-98) called
 * for every forest.  ux, uy, uz, ra, name, nb_off (n_forest + 1) and nb_idx
 * are device pointers; scratch is a device buffer of at least
 * lya_pairs_device_scratch_bytes(n_forest, cos_max) bytes.
 * Call with nb_idx == NULL to fill nb_off and get *n_pairs (the call waits for
 * the stream to read it back); then call again with nb_idx of capacity >=
 * *n_pairs.  A forest with more than 16384 neighbours returns
 * LYA_ERR_UNSUPPORTED (use lya_build_pairs).
 */
int lya_pairs_device_scratch_bytes(int32_t n_forest, double cos_max, int64_t* bytes);
int lya_build_pairs_device(int32_t n_forest, const double* ux, const double* uy, const double* uz,
                           const double* ra, const int64_t* name, double cos_max,
                           int64_t* nb_off, int32_t* nb_idx, int64_t nb_capacity,
                           void* scratch, int64_t scratch_bytes, int64_t* n_pairs, void* stream);

/*
 * Work items on the GPU: the item cut of lya_build_work (same items, same
 * order) with the cost of each item.  f1_list, f1_slot, pix_off, nb_off,
 * nb_idx, items (4 * capacity int32) and costs (capacity int64) are device
 * pointers; scratch is at least lya_items_device_scratch_bytes(n_f1) bytes.
 * Call with items == NULL for *n_items and *total_cost (waits for the stream);
 * then with buffers.  Tiles follow from lya_build_tiles on the host copy.
 */
int lya_items_device_scratch_bytes(int32_t n_f1, int64_t* bytes);
int lya_build_items_device(int32_t n_f1, const int32_t* f1_list, const int32_t* f1_slot,
                           const int64_t* pix_off, const int64_t* nb_off, const int32_t* nb_idx,
                           int64_t item_cost, int32_t* items, int64_t* costs, int64_t items_capacity,
                           void* scratch, int64_t scratch_bytes, int64_t* n_items, double* total_cost,
                           void* stream);

/* One xi problem: all pointers are device pointers except where noted. */
typedef struct lya_xi_problem {
    int32_t n_forest;
    int32_t n_slots;
    const int64_t* pix_off;   /* [n_forest+1] pixel offsets of each forest            */
    const double* ux;         /* [n_forest] unit vector of the line of sight          */
    const double* uy;
    const double* uz;
    const double* base;       /* [n_forest] integer-valued origin of off32 (mixed)    */
    const uint8_t* sorted;    /* [n_forest] 1 if dc increases along the forest        */
    const double* dc;         /* [n_pix] comoving distance, FP64                      */
    const double* we;         /* [n_pix] weight, FP64 (FP64 and mixed modes)          */
    const double* dw;         /* [n_pix] weight*delta, FP64 (FP64 and mixed modes)    */
    const float* off32;       /* [n_pix] dc - base[forest], FP32 (mixed)              */
    const float* we32;        /* [n_pix] weight, FP32 (mixed32 mode)                  */
    const float* dw32;        /* [n_pix] weight*delta, FP32 (mixed32 mode)            */
    const int32_t* nb_idx;    /* [n_pairs] neighbour forest indices                   */
    const int32_t* items;     /* [n_items][4] f1, nb_begin, nb_end, slot              */
    const int32_t* tiles;     /* [n_tiles][2] item_begin, item_end                    */
    int32_t n_items;
    int32_t n_tiles;
    int32_t np_rp;            /* bins along r_parallel                                */
    int32_t np_rt;            /* bins along r_perp                                    */
    double rpmax;
    double rtmax;
    float tau;                /* mixed: half-width of the FP64 recheck band, bin units */
    int32_t lane_map;         /* lane -> pixel order: 0 consecutive, 1 by 2, 2 by 4   */
    int32_t max_len;          /* host: longest forest in pixels (walk 1: <= 16000)    */
    int32_t pad;              /* dc, we, dw readable at indices [-pad, 0) and         */
                              /* [n_pix, n_pix + pad); walk 1 needs pad >= max_len,  */
                              /* dc = 1e300 and we, dw finite there                  */
    uint64_t* out_cnt;        /* [n_slots][np_rp*np_rt] pair counts                   */
    double* out_w;            /* [n_slots][np_rp*np_rt] sum w1*w2                     */
    double* out_dw;           /* [n_slots][np_rp*np_rt] sum dw1*dw2                   */
    uint32_t* queue;          /* [>= 16] device scratch: work-queue counters          */
    const int32_t* tile_groups; /* [n_tile_groups+1] tile offsets of the tile groups;   */
                              /* NULL: one group.  Workgroup b takes tiles from group */
                              /* b % n_tile_groups first (consecutive workgroups sit  */
                              /* on different XCDs), then from the others             */
    int32_t n_tile_groups;    /* 0 or 1: one queue over all tiles; at most 16         */
} lya_xi_problem;

/* Kernel shape, chosen by the autotuner.  Zero fields take defaults. */
typedef struct lya_xi_config {
    int32_t mode;             /* LYA_MODE_FP64, LYA_MODE_MIXED or LYA_MODE_MIXED_F32  */
    int32_t px_per_lane;      /* 1, 2, 4: forest-1 tile of 64, 128, 256 pixels        */
    int32_t waves;            /* waves per workgroup: 4, 8 or 16 (replicas 2: 8 only) */
    int32_t replicas;         /* LDS histogram copies: 1 or 2                         */
    int32_t row_pad;          /* extra LDS columns per r_parallel row: 0 or 1         */
    int32_t groups;           /* workgroups in the persistent grid (0 = fill the GPU)  */
    int32_t no_counts;        /* 1: skip the per-bin pair counts (out_cnt left zero)  */
    int32_t walk;             /* 0: lanes hold forest-1 pixels (all modes);           */
                              /* 1: lanes walk diagonals (i, i+d) of each forest pair */
                              /*    and sum runs of equal bins in registers (FP64)    */
    int32_t merge;            /* walk 0, fp64 and mixed: 1 sums runs of equal bins    */
                              /* within each block of forest-2 pixels before the LDS  */
                              /* atomics (one atomic set per run)                     */
    int32_t tail_max;         /* walk 0.  Forest 1 is cut into slabs of 64*px_per_lane */
                              /* pixels, then 64-pixel slabs; a last piece of at most */
                              /* tail_max (< 64) pixels is walked against forest 2 in */
                              /* lanes (fewer idle lanes).  0: 48; -1: slabs of       */
                              /* 64*px_per_lane pixels only, the last one partial      */
} lya_xi_config;

/* Largest mixed-mode error bound for these data (host): returns tau in bin units.
 * max_abs_off = largest |dc - base| (the FP32 offsets), max_dc = largest dc. */
int lya_xi_tau(const lya_xi_problem* host_view_of_sizes, double max_abs_off, double max_dc, double max_c_binner,
               double max_s_binner, double* tau);

/* Zero outputs and queue, then launch the persistent kernel on `stream`.
 * Any np_rp x np_rt binning: when np_rp * (np_rt + pad) exceeds LYA_MAX_BINS the
 * tile kernels (fp64 and mixed modes, 8 waves, one replica) walk each tile once per
 * band of bins that fits the LDS histogram (column bands of whole rows, or blocks),
 * with each task pruned to the pixels that can land in the band.  mixed32, other
 * workgroup shapes and walk 1 (at most (np_rp+1)(np_rt+1) <= 2688) return
 * LYA_ERR_UNSUPPORTED for such binnings. */
int lya_xi_run(const lya_xi_problem* p, const lya_xi_config* cfg, void* stream);

/* The same sequence captured into a hipGraph (zeroing uses fill kernels). */
typedef struct lya_graph_s* lya_graph_t;
int lya_xi_graph_create(const lya_xi_problem* p, const lya_xi_config* cfg, void* stream,
                        lya_graph_t* out);
int lya_graph_launch(lya_graph_t g, void* stream);
int lya_graph_destroy(lya_graph_t g);

/* Workgroups the persistent grid uses for cfg (0 on error). */
int lya_xi_grid(const lya_xi_config* cfg, int32_t* groups);

/* Work of the tile kernels for (p, cfg) without binning (no counterpart in the
 * reference; a measurement helper).  out: device uint64[4] = pixel-pair slots
 * stepped (lanes x pruned forest-2 range, idle lanes included), the same over
 * lanes holding a pixel, tasks, tasks with a non-empty range.  One band. */
int lya_xi_work_stats(const lya_xi_problem* p, const lya_xi_config* cfg, uint64_t* out, void* stream);

/*
 * Adjoint of the pair sums.  For every pair (i in forest f1, j in neighbour
 * f2) in bin A:  g_dw[i] += alpha[A] * dw[j],  g_w[i] += beta[A] * w[j]
 * (alpha, beta: [np_rp*np_rt] FP64, global bins).  Gradients w.r.t. the
 * forest-2 side come from a second call with the transposed pair list.
 * Any binning: (alpha, beta) sit in LDS when (nb + 1) * 16 bytes fit a workgroup,
 * else they are read from global memory.
 * Uses the mixed geometry (exact bins) when cfg->mode is LYA_MODE_MIXED.
 * g_dw, g_w: [n_pix] FP64, accumulated (caller zeroes).
 */
int lya_xi_adjoint(const lya_xi_problem* p, const lya_xi_config* cfg, const double* alpha,
                   const double* beta, double* g_dw, double* g_w, void* stream);

/*
 * out[f] = sum of x[off[f] .. off[f+1]) for f < nseg (FP64, one wave per segment,
 * fixed summation order).  The per-forest weighted sums of the projection
 * (forest_class.py:54-66) in the gradient path, and the adjoint of broadcasting a
 * per-forest value to its pixels.
 */
int lya_segment_sum(const double* x, const int64_t* off, int32_t nseg, double* out, void* stream);

/*
 * Distortion matrix, FP64 accumulation.  Work items are forest pairs
 * {f1, f2} (already subsampled).  Per pair, with pixel bins A(i,j):
 *   D[A][B] += w_i w_j (eta terms of cuda_kernels.cpp:263-272) for every B
 *   active in the pair, D[A][A] += w_i w_j, wA[A] += w_i w_j.
 * dl = delta_lambda, omega = sum w, odl2 = sum w dl^2 per forest (FP64).
 * max_len must be at least the length of every forest in the pairs (it sizes the
 * LDS staging); pairs with a longer forest are skipped.
 * Any binning with nb = np_rp*np_rt <= 16384.  The LDS holds the Z rows of a pair's
 * active bins in chunks; a pair with more active bins than two rows of the LDS
 * chunk buffer hold is skipped (and counted).
 * Skipped pairs are counted in the uint32 words LYA_DIST_SKIP_LONG (forest longer
 * than max_len) and LYA_DIST_SKIP_BINS (too many active bins) of scratch: the
 * counts add up over calls; the caller zeroes scratch before the first call and
 * reads them back after the last.
 * D: [nb][nb] FP64 with nb = np_rp*np_rt, wA: [nb]; both accumulated.
 * Each pair's contribution is symmetric, and only the upper triangle of D
 * (row <= column) is accumulated: call lya_distortion_mirror() after the last
 * lya_distortion_run() to copy it into the lower triangle.
 * scratch: device buffer of at least lya_distortion_scratch_bytes() bytes.
 * groups: persistent workgroups (one per CU fills the GPU).
 * Replaces the kernels launched at distortion_procedures_pycuda.py:249-279.
 */
#define LYA_DIST_SKIP_LONG 1
#define LYA_DIST_SKIP_BINS 2
int64_t lya_distortion_scratch_bytes(int32_t groups, int32_t max_len, int32_t nbins);
int lya_distortion_run(int32_t n_forest, const int64_t* pix_off, const double* dc, const double* we,
                       const double* dl, const double* ux, const double* uy, const double* uz,
                       const double* omega, const double* odl2, const int32_t* pairs, int32_t n_pairs,
                       int32_t max_len, int32_t np_rp, int32_t np_rt, double rpmax, double rtmax,
                       double* D, double* wA, void* scratch, int64_t scratch_bytes, int32_t groups,
                       void* stream);
/*
 * lya_distortion_run without a scratch argument and without any allocation.
 * No counterpart in the reference: for callers that cannot allocate device
 * memory (code under HIP graph capture, serving loops, harnesses that forbid
 * hipMalloc).  The scratch is a built-in device buffer of LYA_DIST_NOALLOC_BYTES
 * per GPU, held until the process ends.  Calls needing more, i.e.
 * lya_distortion_scratch_bytes(groups, max_len, nb) > LYA_DIST_NOALLOC_BYTES
 * (above 417 workgroups at 50 x 50 bins, above 104 at 100 x 100), return
 * LYA_ERR_CAPACITY.  Every call on one GPU shares that buffer: calls that
 * overlap on one GPU (two streams, two threads) race and give wrong results.
 * skipped: NULL, or a caller-owned device uint32[2] to which the pairs skipped
 * (forest longer than max_len, more active bins than the LDS holds) are added.
 * Otherwise as lya_distortion_run.
 */
#define LYA_DIST_NOALLOC_BYTES (32ll << 20)
int lya_distortion_run_noalloc(int32_t n_forest, const int64_t* pix_off, const double* dc, const double* we,
                               const double* dl, const double* ux, const double* uy, const double* uz,
                               const double* omega, const double* odl2, const int32_t* pairs, int32_t n_pairs,
                               int32_t max_len, int32_t np_rp, int32_t np_rt, double rpmax, double rtmax,
                               double* D, double* wA, uint32_t* skipped, int32_t groups, void* stream);
/* D[c][r] = D[r][c] for r < c: completes the symmetric D accumulated by
 * lya_distortion_run (stream-ordered; idempotent). */
int lya_distortion_mirror(double* D, int32_t nb, void* stream);

/* Test helpers. */
/* c = sqrt(0.5*(1+cos12)), s = sqrt(0.5*(1-cos12)) evaluated by the device code path. */
int lya_debug_halfangles(const double* cos12, double* c, double* s, int32_t n, void* stream);
/* Plain device copy of n bytes (profiler calibration of HBM byte counters). */
int lya_copy_bytes(const void* src, void* dst, int64_t n, void* stream);
/* LDS atomic-add rate probe: `blocks` workgroups of 512 threads, each lane adding 1
 * `iters` times to its own LDS entry; type 0 uint32, 1 float, 2 double.  out holds
 * `blocks` elements of that type.  The caller times the launch. */
int lya_debug_lds_atomic(int32_t type, int32_t blocks, int32_t iters, void* out, void* stream);

#ifdef __cplusplus
}
#endif
#endif
