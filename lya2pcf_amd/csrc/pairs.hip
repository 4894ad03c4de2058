// Forest-pair list and work items built on the GPU.
//
// The neighbour rule is that of quasar.neighborhood (reference
// forest_class.py:78-98): forest j is a neighbour of forest i when
//     (x_i x_j + y_i y_j) + z_i z_j > cos(angmax)   (FP64, no contraction),
//     name_i != name_j   and   ra_i < ra_j.
// The reference takes candidates from a healpix disc query; any superset of
// the true neighbours gives the same list, so the candidates here come from a
// cubic grid over the unit-vector cube [-1,1]^3 whose cell edge h is at least
// the chord of the cut angle (with a margin): every neighbour of a forest sits
// in one of the 27 cells around it.  The grid is a dense table of at most
// MAX_M^3 cells; for tiny cut angles it is coarser than the chord, which only
// adds candidates.
//
// Steps (all on one stream, no host synchronisation except where a size is
// returned):
//   1. cell key per forest, cell populations (atomics), exclusive scan ->
//      cell start; scatter of the forests into cell order with a copy of the
//      unit vectors, ra and name in that order (coalesced candidate reads);
//   2. one wave per forest counts its neighbours (ballot + popcount);
//      exclusive scan -> CSR offsets nb_off;
//   3. one workgroup per forest collects its neighbours into LDS, sorts them
//      (bitonic network) and writes the row: ascending f2, the order of the
//      host list, so both builders give identical arrays.
// Work items (the cut of lya_build_work, one thread per forest 1) follow the
// same count / scan / fill pattern.
#include <hip/hip_runtime.h>
#include <algorithm>
#include <climits>
#include <cmath>
#include <cstdint>

#include "../../include/lya2pcf.h"
#include "errors.h"

#pragma clang fp contract(off)

namespace {

constexpr int MAX_M = 160;          // grid cells per axis at most (MAX_M^3 cells)
constexpr int SCAN_BLK = 1024;      // elements per scan block (256 threads x 4)
constexpr int MAX_ROW = 16384;      // neighbours per forest the LDS sort holds

struct DevGrid {
    double h;
    int m;
};

DevGrid grid_for(double cos_max) {
    const double cm = std::max(-1.0, std::min(1.0, cos_max));
    const double chord = std::sqrt(std::max(0.0, 2.0 - 2.0 * cm));
    DevGrid g;
    g.h = std::max(chord * 1.0001 + 1e-9, 2.0 / (MAX_M - 1));
    g.m = (int)std::ceil(2.0 / g.h) + 1;
    if (g.m > MAX_M) g.m = MAX_M;
    return g;
}

__device__ __forceinline__ int coord(double u, double h, int m) {
    const int c = (int)floor((u + 1.0) / h);
    return min(max(c, 0), m - 1);
}

inline size_t align256(size_t b) { return (b + 255) & ~size_t(255); }

// Scratch layout, shared by the size query and the builder.
struct PairScratch {
    int32_t* key;      // [n] cell of each forest
    int32_t* cnt;      // [MC] cell population, then scatter cursor
    int64_t* start;    // [MC+1] first slot of each cell in cell order
    int32_t* order;    // [n] forest index in cell order
    double* sx;        // [n] unit vectors, ra and name in cell order
    double* sy;
    double* sz;
    double* sra;
    int64_t* sname;
    int64_t* counts;   // [n] neighbours per forest
    int64_t* bsum;     // [nb] scan block sums
    int64_t* boff;     // [nb] scan block offsets
    int32_t* max_row;  // [1]
    size_t bytes;
};

int64_t scan_blocks(int64_t len) { return (len + SCAN_BLK - 1) / SCAN_BLK; }

PairScratch pair_layout(char* base, int64_t n, int64_t mc) {
    PairScratch s;
    size_t o = 0;
    auto take = [&](size_t b) {
        char* p = base ? base + o : nullptr;
        o += align256(b);
        return p;
    };
    const int64_t nb = std::max<int64_t>(1, std::max(scan_blocks(mc), scan_blocks(n)));
    s.key = (int32_t*)take(4 * n);
    s.cnt = (int32_t*)take(4 * mc);
    s.start = (int64_t*)take(8 * (mc + 1));
    s.order = (int32_t*)take(4 * n);
    s.sx = (double*)take(8 * n);
    s.sy = (double*)take(8 * n);
    s.sz = (double*)take(8 * n);
    s.sra = (double*)take(8 * n);
    s.sname = (int64_t*)take(8 * n);
    s.counts = (int64_t*)take(8 * n);
    s.bsum = (int64_t*)take(8 * nb);
    s.boff = (int64_t*)take(8 * nb);
    s.max_row = (int32_t*)take(4);
    s.bytes = o;
    return s;
}

// ---------------------------------------------------------------------------
// Exclusive scan: out[k] = sum_{t<k} in[t] for k < len, out[len] = total.
// ---------------------------------------------------------------------------
template <class T>
__global__ void __launch_bounds__(256) scan_reduce_kernel(const T* in, int64_t len, int64_t* bsum) {
    __shared__ int64_t part[256];
    const int64_t b0 = (int64_t)blockIdx.x * SCAN_BLK;
    int64_t v = 0;
    for (int e = 0; e < 4; ++e) {
        const int64_t k = b0 + threadIdx.x * 4 + e;
        if (k < len) v += (int64_t)in[k];
    }
    part[threadIdx.x] = v;
    __syncthreads();
    for (int st = 128; st > 0; st >>= 1) {
        if ((int)threadIdx.x < st) part[threadIdx.x] += part[threadIdx.x + st];
        __syncthreads();
    }
    if (threadIdx.x == 0) bsum[blockIdx.x] = part[0];
}

// inclusive scan of one value per thread over a 256-thread block
__device__ int64_t block_inclusive(int64_t v, int64_t* sh) {
    const int t = threadIdx.x;
    sh[t] = v;
    __syncthreads();
    for (int st = 1; st < 256; st <<= 1) {
        const int64_t add = t >= st ? sh[t - st] : 0;
        __syncthreads();
        sh[t] += add;
        __syncthreads();
    }
    const int64_t r = sh[t];
    __syncthreads();
    return r;
}

__global__ void __launch_bounds__(256) scan_blocksums_kernel(const int64_t* bsum, int64_t nb, int64_t* boff,
                                                             int64_t* total) {
    __shared__ int64_t sh[256];
    __shared__ int64_t chunk_total;
    int64_t carry = 0;
    for (int64_t c0 = 0; c0 < nb; c0 += 256) {
        const int64_t k = c0 + threadIdx.x;
        const int64_t v = k < nb ? bsum[k] : 0;
        const int64_t inc = block_inclusive(v, sh);
        if (k < nb) boff[k] = carry + inc - v;
        // the chunk total is the inclusive value of the last thread
        if (threadIdx.x == 255) chunk_total = inc;
        __syncthreads();
        carry += chunk_total;
        __syncthreads();
    }
    if (threadIdx.x == 0) *total = carry;
}

template <class T>
__global__ void __launch_bounds__(256) scan_apply_kernel(const T* in, int64_t len, const int64_t* boff,
                                                         int64_t* out) {
    __shared__ int64_t sh[256];
    const int64_t b0 = (int64_t)blockIdx.x * SCAN_BLK;
    int64_t v[4];
    int64_t tsum = 0;
    for (int e = 0; e < 4; ++e) {
        const int64_t k = b0 + threadIdx.x * 4 + e;
        v[e] = k < len ? (int64_t)in[k] : 0;
        tsum += v[e];
    }
    const int64_t inc = block_inclusive(tsum, sh);
    int64_t run = boff[blockIdx.x] + inc - tsum;
    for (int e = 0; e < 4; ++e) {
        const int64_t k = b0 + threadIdx.x * 4 + e;
        if (k < len) out[k] = run;
        run += v[e];
    }
}

template <class T>
int exclusive_scan(const T* in, int64_t len, int64_t* out, int64_t* bsum, int64_t* boff, hipStream_t st) {
    const int64_t nb = std::max<int64_t>(1, scan_blocks(len));
    hipLaunchKernelGGL(scan_reduce_kernel<T>, dim3((unsigned)nb), dim3(256), 0, st, in, len, bsum);
    hipLaunchKernelGGL(scan_blocksums_kernel, dim3(1), dim3(256), 0, st, bsum, nb, boff, out + len);
    hipLaunchKernelGGL(scan_apply_kernel<T>, dim3((unsigned)nb), dim3(256), 0, st, in, len, boff, out);
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? LYA_OK : lya_fail(LYA_ERR_HIP, "scan launch: %s", hipGetErrorString(e));
}

template <class T>
__global__ void fill_kernel(T* p, int64_t n, T v) {
    for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += (int64_t)gridDim.x * blockDim.x)
        p[k] = v;
}

template <class T>
void fill(T* p, int64_t n, T v, hipStream_t st) {
    const int blocks = (int)std::max<int64_t>(1, std::min<int64_t>((n + 255) / 256, 4096));
    hipLaunchKernelGGL(fill_kernel<T>, dim3(blocks), dim3(256), 0, st, p, n, v);
}

// ---------------------------------------------------------------------------
// Cell table
// ---------------------------------------------------------------------------
__global__ void cell_key_kernel(int n, const double* ux, const double* uy, const double* uz, double h, int m,
                                int32_t* key, int32_t* cnt) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int k = (coord(ux[i], h, m) * m + coord(uy[i], h, m)) * m + coord(uz[i], h, m);
    key[i] = k;
    atomicAdd(&cnt[k], 1);
}

__global__ void cell_scatter_kernel(int n, const int32_t* key, const int64_t* start, int32_t* cursor,
                                    const double* ux, const double* uy, const double* uz, const double* ra,
                                    const int64_t* name, int32_t* order, double* sx, double* sy, double* sz,
                                    double* sra, int64_t* sname) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int k = key[i];
    const int64_t pos = start[k] + atomicAdd(&cursor[k], 1);
    order[pos] = i;
    sx[pos] = ux[i];
    sy[pos] = uy[i];
    sz[pos] = uz[i];
    sra[pos] = ra[i];
    sname[pos] = name[i];
}

struct Query {
    double x, y, z, ra;
    int64_t name;
    int a, b, c;
};

__device__ __forceinline__ Query query(int i, const double* ux, const double* uy, const double* uz,
                                       const double* ra, const int64_t* name, double h, int m) {
    Query q;
    q.x = ux[i];
    q.y = uy[i];
    q.z = uz[i];
    q.ra = ra[i];
    q.name = name[i];
    q.a = coord(q.x, h, m);
    q.b = coord(q.y, h, m);
    q.c = coord(q.z, h, m);
    return q;
}

// the neighbour test of forest_class.py:96, operand order of the host builder
__device__ __forceinline__ bool is_neighbour(const Query& q, int64_t t, const double* sx, const double* sy,
                                             const double* sz, const double* sra, const int64_t* sname,
                                             double cos_max) {
    double mu = q.x * sx[t] + q.y * sy[t];
    mu = mu + q.z * sz[t];
    return mu > cos_max && q.name != sname[t] && q.ra < sra[t];
}

// one wave per forest: neighbour count
__global__ void __launch_bounds__(256) pair_count_kernel(int n, const double* ux, const double* uy,
                                                         const double* uz, const double* ra, const int64_t* name,
                                                         double cos_max, double h, int m, const int64_t* start,
                                                         const double* sx, const double* sy, const double* sz,
                                                         const double* sra, const int64_t* sname,
                                                         int64_t* counts, int32_t* max_row) {
    const int lane = threadIdx.x & 63;
    const int i = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (i >= n) return;  // wave-uniform
    const Query q = query(i, ux, uy, uz, ra, name, h, m);
    int64_t cnt = 0;
    for (int da = -1; da <= 1; ++da)
        for (int db = -1; db <= 1; ++db)
            for (int dc = -1; dc <= 1; ++dc) {
                const int A = q.a + da, B = q.b + db, Cc = q.c + dc;
                if (A < 0 || B < 0 || Cc < 0 || A >= m || B >= m || Cc >= m) continue;
                const int k = (A * m + B) * m + Cc;
                const int64_t t0 = start[k], t1 = start[k + 1];
                for (int64_t base = t0; base < t1; base += 64) {
                    const int64_t t = base + lane;
                    const bool ok = t < t1 && is_neighbour(q, t, sx, sy, sz, sra, sname, cos_max);
                    cnt += __popcll(__ballot(ok));
                }
            }
    if (lane == 0) {
        counts[i] = cnt;
        atomicMax(max_row, cnt < INT_MAX ? (int32_t)cnt : INT_MAX);
    }
}

// one workgroup (4 waves) per forest: collect, sort ascending, write the row
template <int CAP>
__global__ void __launch_bounds__(256) pair_fill_kernel(int n, const double* ux, const double* uy,
                                                        const double* uz, const double* ra, const int64_t* name,
                                                        double cos_max, double h, int m, const int64_t* start,
                                                        const int32_t* order, const double* sx, const double* sy,
                                                        const double* sz, const double* sra,
                                                        const int64_t* sname, const int64_t* nb_off,
                                                        int32_t* nb_idx) {
    __shared__ int32_t buf[CAP];
    __shared__ int32_t s_n;
    const int i = blockIdx.x;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int64_t row0 = nb_off[i];
    const int len = (int)(nb_off[i + 1] - row0);
    if (len == 0) return;  // block-uniform
    if (tid == 0) s_n = 0;
    __syncthreads();
    const Query q = query(i, ux, uy, uz, ra, name, h, m);
    for (int da = -1; da <= 1; ++da)
        for (int db = -1; db <= 1; ++db)
            for (int dc = -1; dc <= 1; ++dc) {
                const int A = q.a + da, B = q.b + db, Cc = q.c + dc;
                if (A < 0 || B < 0 || Cc < 0 || A >= m || B >= m || Cc >= m) continue;
                const int k = (A * m + B) * m + Cc;
                const int64_t t0 = start[k], t1 = start[k + 1];
                for (int64_t base = t0 + 64 * wave; base < t1; base += 256) {
                    const int64_t t = base + lane;
                    const bool ok = t < t1 && is_neighbour(q, t, sx, sy, sz, sra, sname, cos_max);
                    const uint64_t bal = __ballot(ok);
                    if (bal == 0) continue;  // wave-uniform
                    int wb = 0;
                    if (lane == 0) wb = atomicAdd(&s_n, __popcll(bal));
                    wb = __shfl(wb, 0, 64);
                    const int pos = wb + (int)__builtin_amdgcn_mbcnt_hi((uint32_t)(bal >> 32),
                                                                        __builtin_amdgcn_mbcnt_lo((uint32_t)bal, 0u));
                    if (ok && pos < CAP) buf[pos] = order[t];
                }
            }
    __syncthreads();
    int pw = 1;
    while (pw < len) pw <<= 1;
    for (int t = len + tid; t < pw; t += 256) buf[t] = INT_MAX;
    __syncthreads();
    for (int k = 2; k <= pw; k <<= 1)
        for (int j = k >> 1; j > 0; j >>= 1) {
            for (int t = tid; t < pw; t += 256) {
                const int u = t ^ j;
                if (u > t) {
                    const int a = buf[t], b = buf[u];
                    const bool up = (t & k) == 0;
                    if ((a > b) == up) {
                        buf[t] = b;
                        buf[u] = a;
                    }
                }
            }
            __syncthreads();
        }
    for (int t = tid; t < len; t += 256) nb_idx[row0 + t] = buf[t];
}

// ---------------------------------------------------------------------------
// Work items: the greedy cut of host lya_build_work, one thread per forest 1.
// ---------------------------------------------------------------------------
__global__ void item_count_kernel(int n_f1, const int32_t* f1_list, const int64_t* pix_off, const int64_t* nb_off,
                                  const int32_t* nb_idx, int64_t item_cost, int64_t* item_cnt,
                                  unsigned long long* total) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n_f1) return;
    const int f = f1_list[k];
    const int64_t n1 = pix_off[f + 1] - pix_off[f];
    const int64_t b = nb_off[f], e = nb_off[f + 1];
    int64_t s = b, acc = 0, ni = 0, tot = 0;
    for (int64_t q = b; q < e; ++q) {
        const int f2 = nb_idx[q];
        const int64_t c = n1 * (pix_off[f2 + 1] - pix_off[f2]);
        tot += c;
        if (q > s && acc + c > item_cost) {
            ++ni;
            s = q;
            acc = 0;
        }
        acc += c;
    }
    if (e > s) ++ni;
    item_cnt[k] = ni;
    atomicAdd(total, (unsigned long long)tot);
}

__global__ void item_fill_kernel(int n_f1, const int32_t* f1_list, const int32_t* f1_slot, const int64_t* pix_off,
                                 const int64_t* nb_off, const int32_t* nb_idx, int64_t item_cost,
                                 const int64_t* item_off, int32_t* items, int64_t* costs) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n_f1) return;
    const int f = f1_list[k];
    const int slot = f1_slot[k];
    const int64_t n1 = pix_off[f + 1] - pix_off[f];
    const int64_t b = nb_off[f], e = nb_off[f + 1];
    int64_t s = b, acc = 0, o = item_off[k];
    for (int64_t q = b; q < e; ++q) {
        const int f2 = nb_idx[q];
        const int64_t c = n1 * (pix_off[f2 + 1] - pix_off[f2]);
        if (q > s && acc + c > item_cost) {
            items[4 * o + 0] = f;
            items[4 * o + 1] = (int32_t)s;
            items[4 * o + 2] = (int32_t)q;
            items[4 * o + 3] = slot;
            costs[o] = acc;
            ++o;
            s = q;
            acc = 0;
        }
        acc += c;
    }
    if (e > s) {
        items[4 * o + 0] = f;
        items[4 * o + 1] = (int32_t)s;
        items[4 * o + 2] = (int32_t)e;
        items[4 * o + 3] = slot;
        costs[o] = acc;
    }
}

int launch_check(const char* what) {
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? LYA_OK : lya_fail(LYA_ERR_HIP, "%s launch: %s", what, hipGetErrorString(e));
}

int copy_to_host(void* dst, const void* src, size_t b, hipStream_t st) {
    hipError_t e = hipMemcpyAsync(dst, src, b, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    return e == hipSuccess ? LYA_OK : lya_fail(LYA_ERR_HIP, "device-to-host copy: %s", hipGetErrorString(e));
}

}  // namespace

extern "C" int lya_pairs_device_scratch_bytes(int32_t n_forest, double cos_max, int64_t* bytes) {
    if (n_forest < 0 || !bytes) return lya_fail(LYA_ERR_ARG, "lya_pairs_device_scratch_bytes: bad argument");
    const DevGrid g = grid_for(cos_max);
    const int64_t mc = (int64_t)g.m * g.m * g.m;
    *bytes = (int64_t)pair_layout(nullptr, n_forest, mc).bytes;
    return LYA_OK;
}

extern "C" int lya_build_pairs_device(int32_t n, const double* ux, const double* uy, const double* uz,
                                      const double* ra, const int64_t* name, double cos_max, int64_t* nb_off,
                                      int32_t* nb_idx, int64_t cap, void* scratch, int64_t scratch_bytes,
                                      int64_t* n_pairs, void* stream) {
    if (n < 0 || !nb_off || !n_pairs || (n > 0 && (!ux || !uy || !uz || !ra || !name || !scratch)))
        return lya_fail(LYA_ERR_ARG, "lya_build_pairs_device: null argument or n < 0");
    hipStream_t st = (hipStream_t)stream;
    const DevGrid g = grid_for(cos_max);
    const int64_t mc = (int64_t)g.m * g.m * g.m;
    const PairScratch s = pair_layout((char*)scratch, n, mc);
    if (scratch_bytes < (int64_t)s.bytes)
        return lya_fail(LYA_ERR_CAPACITY, "lya_build_pairs_device: scratch %lld < %zu bytes", (long long)scratch_bytes,
                        s.bytes);
    if (n == 0) {
        fill<int64_t>(nb_off, 1, 0, st);
        *n_pairs = 0;
        return launch_check("fill");
    }
    const unsigned nblk = (unsigned)((n + 255) / 256);
    int rc;
    // 1. cell table
    fill<int32_t>(s.cnt, mc, 0, st);
    fill<int32_t>(s.max_row, 1, 0, st);
    hipLaunchKernelGGL(cell_key_kernel, dim3(nblk), dim3(256), 0, st, n, ux, uy, uz, g.h, g.m, s.key, s.cnt);
    if ((rc = launch_check("cell key"))) return rc;
    if ((rc = exclusive_scan<int32_t>(s.cnt, mc, s.start, s.bsum, s.boff, st))) return rc;
    fill<int32_t>(s.cnt, mc, 0, st);
    hipLaunchKernelGGL(cell_scatter_kernel, dim3(nblk), dim3(256), 0, st, n, s.key, s.start, s.cnt, ux, uy, uz, ra,
                       name, s.order, s.sx, s.sy, s.sz, s.sra, s.sname);
    if ((rc = launch_check("cell scatter"))) return rc;
    // 2. counts and offsets
    hipLaunchKernelGGL(pair_count_kernel, dim3((unsigned)((n + 3) / 4)), dim3(256), 0, st, n, ux, uy, uz, ra, name,
                       cos_max, g.h, g.m, s.start, s.sx, s.sy, s.sz, s.sra, s.sname, s.counts, s.max_row);
    if ((rc = launch_check("pair count"))) return rc;
    if ((rc = exclusive_scan<int64_t>(s.counts, n, nb_off, s.bsum, s.boff, st))) return rc;
    int64_t total = 0;
    int32_t max_row = 0;
    if ((rc = copy_to_host(&total, nb_off + n, sizeof(total), st))) return rc;
    if ((rc = copy_to_host(&max_row, s.max_row, sizeof(max_row), st))) return rc;
    *n_pairs = total;
    if (!nb_idx) return LYA_OK;
    if (cap < total)
        return lya_fail(LYA_ERR_CAPACITY, "lya_build_pairs_device: capacity %lld < %lld pairs", (long long)cap,
                        (long long)total);
    if (total > INT32_MAX)
        return lya_fail(LYA_ERR_CAPACITY, "lya_build_pairs_device: %lld pairs exceed the int32 item range",
                        (long long)total);
    if (max_row > MAX_ROW)
        return lya_fail(LYA_ERR_UNSUPPORTED,
                        "lya_build_pairs_device: a forest has %d neighbours (> %d); use lya_build_pairs", max_row,
                        MAX_ROW);
    // 3. rows, sorted
    auto fillrows = [&](auto kern) {
        hipLaunchKernelGGL(kern, dim3((unsigned)n), dim3(256), 0, st, n, ux, uy, uz, ra, name, cos_max, g.h, g.m,
                           s.start, s.order, s.sx, s.sy, s.sz, s.sra, s.sname, nb_off, nb_idx);
    };
    if (max_row <= 1024)
        fillrows(pair_fill_kernel<1024>);
    else if (max_row <= 4096)
        fillrows(pair_fill_kernel<4096>);
    else
        fillrows(pair_fill_kernel<MAX_ROW>);
    return launch_check("pair fill");
}

extern "C" int lya_items_device_scratch_bytes(int32_t n_f1, int64_t* bytes) {
    if (n_f1 < 0 || !bytes) return lya_fail(LYA_ERR_ARG, "lya_items_device_scratch_bytes: bad argument");
    const int64_t nb = std::max<int64_t>(1, scan_blocks(n_f1));
    *bytes = (int64_t)(align256(8 * (size_t)n_f1) + align256(8 * ((size_t)n_f1 + 1)) + 2 * align256(8 * nb) +
                       align256(8));
    return LYA_OK;
}

extern "C" int lya_build_items_device(int32_t n_f1, const int32_t* f1_list, const int32_t* f1_slot,
                                      const int64_t* pix_off, const int64_t* nb_off, const int32_t* nb_idx,
                                      int64_t item_cost, int32_t* items, int64_t* costs, int64_t items_capacity,
                                      void* scratch, int64_t scratch_bytes, int64_t* n_items, double* total_cost,
                                      void* stream) {
    if (n_f1 < 0 || !n_items || (n_f1 > 0 && (!f1_list || !f1_slot || !pix_off || !nb_off || !nb_idx || !scratch)))
        return lya_fail(LYA_ERR_ARG, "lya_build_items_device: null argument or n_f1 < 0");
    if (item_cost < 1) return lya_fail(LYA_ERR_ARG, "lya_build_items_device: item_cost must be >= 1");
    int64_t need = 0;
    lya_items_device_scratch_bytes(n_f1, &need);
    if (scratch_bytes < need)
        return lya_fail(LYA_ERR_CAPACITY, "lya_build_items_device: scratch %lld < %lld bytes",
                        (long long)scratch_bytes, (long long)need);
    if (n_f1 == 0) {
        *n_items = 0;
        if (total_cost) *total_cost = 0.0;
        return LYA_OK;
    }
    hipStream_t st = (hipStream_t)stream;
    const int64_t nb = std::max<int64_t>(1, scan_blocks(n_f1));
    char* base = (char*)scratch;
    int64_t* cnt = (int64_t*)base;
    int64_t* off = (int64_t*)(base + align256(8 * (size_t)n_f1));
    int64_t* bsum = (int64_t*)((char*)off + align256(8 * ((size_t)n_f1 + 1)));
    int64_t* boff = (int64_t*)((char*)bsum + align256(8 * nb));
    unsigned long long* tot = (unsigned long long*)((char*)boff + align256(8 * nb));
    const unsigned nblk = (unsigned)((n_f1 + 255) / 256);
    int rc;
    fill<unsigned long long>(tot, 1, 0ull, st);
    hipLaunchKernelGGL(item_count_kernel, dim3(nblk), dim3(256), 0, st, n_f1, f1_list, pix_off, nb_off, nb_idx,
                       item_cost, cnt, tot);
    if ((rc = launch_check("item count"))) return rc;
    if ((rc = exclusive_scan<int64_t>(cnt, n_f1, off, bsum, boff, st))) return rc;
    int64_t ni = 0;
    unsigned long long t = 0;
    if ((rc = copy_to_host(&ni, off + n_f1, sizeof(ni), st))) return rc;
    if ((rc = copy_to_host(&t, tot, sizeof(t), st))) return rc;
    *n_items = ni;
    if (total_cost) *total_cost = (double)t;
    if (!items || !costs) return LYA_OK;
    if (items_capacity < ni)
        return lya_fail(LYA_ERR_CAPACITY, "lya_build_items_device: capacity %lld < %lld items",
                        (long long)items_capacity, (long long)ni);
    if (ni > INT32_MAX) return lya_fail(LYA_ERR_CAPACITY, "lya_build_items_device: too many items");
    hipLaunchKernelGGL(item_fill_kernel, dim3(nblk), dim3(256), 0, st, n_f1, f1_list, f1_slot, pix_off, nb_off,
                       nb_idx, item_cost, off, items, costs);
    return launch_check("item fill");
}
