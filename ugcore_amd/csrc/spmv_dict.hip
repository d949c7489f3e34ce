// Sliced ELL with a column dictionary per slice: the SpMV layout of
// GPUSparseMatrix when the rows can be grouped into slices of 64 whose entries
// touch at most 255 distinct columns (a 4x4x4 block of lattice points has
// 6x6x6 = 216 neighbours in a 27-point pattern, whatever the DoF numbering).
//
//   * slice t holds rows order[64t .. 64t+63] (-1: an empty lane);
//   * its distinct columns, sorted, are dict[dptr[t] .. dptr[t+1]);
//   * entry k of lane l is value sval[vptr[t] + 2*(64*(k/2) + l) + k%2]
//     (two consecutive entries of a row form one 16-byte load) and the byte
//     k%4 of word sidx[iptr[t] + 64*(k/4) + l], the entry's position in the
//     dictionary; a padded entry has value 0 and position 255, which the
//     product maps to 0 (so non-finite vector entries stay out, as they do in
//     the CSR product);
//   * the product stages x at the dictionary's columns in LDS once per slice
//     (216 gathers instead of 27 per row) and streams 9 bytes per entry.
// Each row is summed as the sliced-ELL kernel sums it (even and odd entries of
// the slice width's groups of four in two sums, the last w % 4 entries into
// the first, in order), so the two layouts agree bit for bit where a row's
// slice has the same width in both placements and to rounding otherwise; only
// the storage of the column indices differs from the CSR, which download_csr
// still returns.
#include <hip/hip_runtime.h>

#include <climits>
#include <cstdint>

#include "ugx_internal.hpp"

namespace ugamd {

namespace {

constexpr int kDictSlots = 255;       // dictionary entries per slice (position 255 = zero)
constexpr int kDictMaxWidth = 64;     // longest row of a slice that the layout takes
constexpr int kHashSlots = 512;       // LDS hash of a slice's columns (at most 255 + 64 occupied)
constexpr int kChunk = 8;             // entries per row moved through LDS at a time

typedef double dbl2 __attribute__((ext_vector_type(2)));

inline unsigned nblk(int64_t n, int t) { return (unsigned)((n + t - 1) / t > 0 ? (n + t - 1) / t : 1); }

__device__ __forceinline__ int hash_slot(int32_t c) { return (int)(((uint32_t)c * 2654435761u) >> 23); }

// slot of column c, which is in the table
__device__ __forceinline__ int hash_find(const int32_t* hk, int32_t c) {
    int h = hash_slot(c);
    while (hk[h] != c) h = (h + 1) & (kHashSlots - 1);
    return h;
}

// Kernels below run one 64-lane workgroup (one wavefront) per slice. The
// slice's rows: start and length of lane l's row in rbs[l], lens[l] (length 0
// for an empty lane)
__device__ __forceinline__ void slice_rows(int64_t nrows, const int32_t* __restrict__ rowptr,
                                           const int32_t* __restrict__ order, int64_t t, int32_t* rbs, int32_t* lens) {
    const int lane = threadIdx.x;
    const int32_t r = order[t * 64 + lane];
    int32_t rb = 0, len = 0;
    if (r >= 0 && r < nrows) {
        rb = rowptr[r];
        len = rowptr[r + 1] - rb;
    }
    rbs[lane] = rb;
    lens[lane] = len;
}

// inserts the slice's columns into the LDS hash table hk (all -1 on entry;
// *cnt 0): the entries are read row by row (lane e of a round reads entry
// e % w of row e / w), so a round reads contiguous CSR runs. Returns the
// number of distinct columns, or a number above 255 as soon as a round ends
// above it (then the table is incomplete). A round adds at most 64 keys, so
// at most 319 of the 512 slots are ever occupied and every probe ends.
__device__ int slice_hash(int w, const int32_t* rbs, const int32_t* lens, const int32_t* __restrict__ col,
                          int32_t* hk, int* cnt) {
    const int lane = threadIdx.x;
    const int total = 64 * w;
    for (int base = 0; base < total; base += 64) {
        const int e = base + lane;
        if (e < total) {
            const int j = e / w, k = e - j * w;
            if (k < lens[j]) {
                const int32_t c = col[rbs[j] + k];
                int h = hash_slot(c);
                for (;;) {
                    const int32_t old = atomicCAS(&hk[h], -1, c);
                    if (old == -1) {
                        atomicAdd(cnt, 1);
                        break;
                    }
                    if (old == c) break;
                    h = (h + 1) & (kHashSlots - 1);
                }
            }
        }
        __syncthreads();
        const int u = *cnt;
        __syncthreads();
        if (u > kDictSlots) return u;
    }
    return *cnt;
}

__device__ __forceinline__ void cmp_swap(int32_t& a, int32_t& b, bool up) {
    if ((a > b) == up) {
        const int32_t t = a;
        a = b;
        b = t;
    }
}

// bitonic sort of the 256 keys v[0..3] of the wave (key 4*lane + m in v[m]),
// ascending
__device__ __forceinline__ void wave_sort256(int32_t v[4]) {
    const int lane = threadIdx.x;
    for (int size = 2; size <= 256; size <<= 1) {
        for (int stride = size >> 1; stride > 0; stride >>= 1) {
            if (stride >= 4) {
#pragma unroll
                for (int m = 0; m < 4; ++m) {
                    const int i = 4 * lane + m;
                    const int32_t o = __shfl_xor(v[m], stride >> 2, 64);
                    const bool up = (i & size) == 0, lower = (i & stride) == 0;
                    v[m] = (lower == up) ? min(v[m], o) : max(v[m], o);
                }
            } else {
                const bool up = ((4 * lane) & size) == 0;   // size >= 4 here, or size 2 (both halves alike below)
                if (stride == 2) {
                    cmp_swap(v[0], v[2], up);
                    cmp_swap(v[1], v[3], up);
                } else if (size == 2) {
                    cmp_swap(v[0], v[1], true);
                    cmp_swap(v[2], v[3], false);
                } else {
                    cmp_swap(v[0], v[1], up);
                    cmp_swap(v[2], v[3], up);
                }
            }
        }
    }
}

// distinct columns per slice (above 255, or INT_MAX when the slice is wider
// than the layout takes: the slice does not fit)
__global__ __launch_bounds__(64) void k_dict_count(int64_t nrows, const int32_t* __restrict__ rowptr,
                                                   const int32_t* __restrict__ col, const int32_t* __restrict__ order,
                                                   const int32_t* __restrict__ width, int32_t* __restrict__ ucount) {
    __shared__ int32_t hk[kHashSlots];
    __shared__ int32_t rbs[64], lens[64];
    __shared__ int cnt;
    const int64_t t = blockIdx.x;
    const int lane = threadIdx.x;
    const int w = width[t];
    if (w > kDictMaxWidth) {
        if (lane == 0) ucount[t] = INT_MAX;
        return;
    }
    for (int h = lane; h < kHashSlots; h += 64) hk[h] = -1;
    if (lane == 0) cnt = 0;
    slice_rows(nrows, rowptr, order, t, rbs, lens);
    __syncthreads();
    const int u = slice_hash(w, rbs, lens, col, hk, &cnt);
    if (lane == 0) ucount[t] = u;
}

// moves the slice's values (and, with pb, the dictionary positions of the
// entries) through LDS in chunks of kChunk entries per row: rows are read
// contiguously, the value pairs and position words written contiguously.
// hk/hv: column -> position (used when pb is given)
__device__ void slice_values_out(int w, const int32_t* rbs, const int32_t* lens, const int32_t* __restrict__ col,
                                 const double* __restrict__ val, const int32_t* hk, const int32_t* hv, uint8_t* pb,
                                 double* vt, double* __restrict__ sv, uint32_t* __restrict__ si) {
    const int lane = threadIdx.x;
    const int npairs = (w + 1) >> 1, nwords = (w + 3) >> 2;
    for (int k0 = 0; k0 < w; k0 += kChunk) {
#pragma unroll
        for (int it = 0; it < kChunk; ++it) {
            const int e = it * 64 + lane;
            const int j = e / kChunk, kk = e % kChunk, k = k0 + kk;
            uint32_t pos = 255;
            double v = 0.0;
            if (k < lens[j]) {
                const int64_t a = (int64_t)rbs[j] + k;
                v = val[a];
                if (pb) pos = (uint32_t)hv[hash_find(hk, col[a])];
            }
            if (pb) pb[kk * 64 + j] = (uint8_t)pos;
            vt[(kk >> 1) * 128 + 2 * j + (kk & 1)] = v;
        }
        __syncthreads();
        const int npc = min(kChunk / 2, npairs - k0 / 2);
        for (int o = lane; o < 128 * npc; o += 64) sv[(int64_t)64 * k0 + o] = vt[o];
        if (pb) {
            const int nqc = min(kChunk / 4, nwords - k0 / 4);
            for (int qq = 0; qq < nqc; ++qq) {
                const uint8_t* b = pb + 4 * qq * 64 + lane;
                si[(int64_t)64 * (k0 / 4 + qq) + lane] =
                    (uint32_t)b[0] | ((uint32_t)b[64] << 8) | ((uint32_t)b[128] << 16) | ((uint32_t)b[192] << 24);
            }
        }
        __syncthreads();
    }
}

// the dictionary (sorted distinct columns), positions and values of every
// slice; the count pass has checked that every slice fits
__global__ __launch_bounds__(64) void k_dict_fill(int64_t nrows, const int32_t* __restrict__ rowptr,
                                                  const int32_t* __restrict__ col, const double* __restrict__ val,
                                                  const int32_t* __restrict__ order, const int32_t* __restrict__ width,
                                                  const int32_t* __restrict__ dptr, const int64_t* __restrict__ vptr,
                                                  const int64_t* __restrict__ iptr, int32_t* __restrict__ dict,
                                                  double* __restrict__ sval, uint32_t* __restrict__ sidx) {
    __shared__ int32_t hk[kHashSlots], hv[kHashSlots];
    __shared__ int32_t rbs[64], lens[64], list[256];
    __shared__ int cnt;
    __shared__ uint8_t pb[kChunk * 64];
    __shared__ double vt[kChunk * 64];
    const int64_t t = blockIdx.x;
    const int lane = threadIdx.x;
    const int w = width[t];
    for (int h = lane; h < kHashSlots; h += 64) hk[h] = -1;
    if (lane == 0) cnt = 0;
    slice_rows(nrows, rowptr, order, t, rbs, lens);
    __syncthreads();
    const int u = slice_hash(w, rbs, lens, col, hk, &cnt);
    if (u > kDictSlots) return;   // not a slice that fits (the count pass keeps such layouts from being built)
    // compact the table (lane l owns slots 8l .. 8l+7), then sort
    int n = 0;
    for (int i = 0; i < 8; ++i) n += hk[8 * lane + i] >= 0 ? 1 : 0;
    int off = n;
    for (int d = 1; d < 64; d <<= 1) {
        const int o = __shfl_up(off, d, 64);
        if (lane >= d) off += o;
    }
    off -= n;
    for (int i = 0; i < 8; ++i) {
        const int32_t c = hk[8 * lane + i];
        if (c >= 0) list[off++] = c;
    }
    for (int i = u + lane; i < 256; i += 64) list[i] = INT_MAX;
    __syncthreads();
    int32_t v[4] = {list[4 * lane], list[4 * lane + 1], list[4 * lane + 2], list[4 * lane + 3]};
    wave_sort256(v);
    const int32_t d0 = dptr[t];
#pragma unroll
    for (int m = 0; m < 4; ++m) {
        const int i = 4 * lane + m;
        if (i < u) {
            dict[d0 + i] = v[m];
            hv[hash_find(hk, v[m])] = i;
        }
    }
    __syncthreads();
    slice_values_out(w, rbs, lens, col, val, hk, hv, pb, vt, sval + vptr[t], sidx + iptr[t]);
}

// the values alone, for a matrix whose pattern (and so dictionary and
// positions) is unchanged since k_dict_fill: up to 32 columns of all 64 rows
// at a time through LDS, each row's run read once and in one piece (lane e of
// a round reads entry e % wc of row e / wc), the loads of 8 rounds issued
// before their stores; the value pairs then written out contiguously
constexpr int kValCols = 32;
__global__ __launch_bounds__(64) void k_dict_values(int64_t nrows, const int32_t* __restrict__ rowptr,
                                                    const double* __restrict__ val, const int32_t* __restrict__ order,
                                                    const int32_t* __restrict__ width, const int64_t* __restrict__ vptr,
                                                    double* __restrict__ sval) {
    __shared__ int32_t rbs[64], lens[64];
    __shared__ double vt[kValCols * 64];
    const int64_t t = blockIdx.x;
    slice_rows(nrows, rowptr, order, t, rbs, lens);
    __syncthreads();
    const int w = width[t], lane = threadIdx.x;
    double* sv = sval + vptr[t];
    for (int k0 = 0; k0 < w; k0 += kValCols) {
        const int wc = min(kValCols, w - k0), wp = wc + (wc & 1);   // a padded odd entry is written as 0
        const int total = 64 * wp;
        for (int e0 = lane; e0 < total; e0 += 8 * 64) {
            double v[8];
            int o[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                const int e = e0 + u * 64;
                const int j = e / wp, kk = e - j * wp, k = k0 + kk;
                o[u] = e < total ? (kk >> 1) * 128 + 2 * j + (kk & 1) : -1;
                v[u] = (e < total && k < lens[j]) ? val[(int64_t)rbs[j] + k] : 0.0;
            }
#pragma unroll
            for (int u = 0; u < 8; ++u)
                if (o[u] >= 0) vt[o[u]] = v[u];
        }
        __syncthreads();
        for (int q = lane; q < 64 * wp; q += 64) sv[(int64_t)64 * k0 + q] = vt[q];
        __syncthreads();
    }
}

// y = alpha*v1 + beta*A*x (mode 0), or y = beta*A*x on rows with entries
// (mode 1); one wave per slice, four slices per workgroup
__global__ __launch_bounds__(256) void k_dict_spmv(int64_t nrows, int64_t nslices, const int32_t* __restrict__ width,
                                                   const int32_t* __restrict__ dptr, const int32_t* __restrict__ dict,
                                                   const int64_t* __restrict__ vptr, const int64_t* __restrict__ iptr,
                                                   const double* __restrict__ sval, const uint32_t* __restrict__ sidx,
                                                   const int32_t* __restrict__ order, double alpha, const double* v1,
                                                   double beta, const double* __restrict__ x, double* y, int mode,
                                                   const int32_t* __restrict__ rowptr) {
    __shared__ double xl[4][256];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int64_t t = (int64_t)blockIdx.x * 4 + wave;
    double* xs = xl[wave];
    int w = 0;
    if (t < nslices) {
        const int32_t d0 = dptr[t], u = dptr[t + 1] - d0;
        for (int j = lane; j < u; j += 64) xs[j] = x[dict[d0 + j]];
        if (lane == 0) xs[255] = 0.0;
        w = width[t];
    }
    __syncthreads();
    if (t >= nslices) return;
    const dbl2* vp = reinterpret_cast<const dbl2*>(sval + vptr[t]) + lane;
    const uint32_t* ip = sidx + iptr[t] + lane;
    double s0 = 0.0, s1 = 0.0;
    int k = 0;
    for (; k + 4 <= w; k += 4) {
        const uint32_t word = __builtin_nontemporal_load(ip + 64 * (k >> 2));
        const dbl2 a = __builtin_nontemporal_load(vp + 64 * (k >> 1));
        const dbl2 b = __builtin_nontemporal_load(vp + 64 * ((k >> 1) + 1));
        s0 = fma(a.x, xs[word & 255u], s0);
        s1 = fma(a.y, xs[(word >> 8) & 255u], s1);
        s0 = fma(b.x, xs[(word >> 16) & 255u], s0);
        s1 = fma(b.y, xs[word >> 24], s1);
    }
    // the last w % 4 entries all go to s0, in order, as in the sliced-ELL
    // kernel, so the two layouts round identically
    if (k < w) {
        const uint32_t word = __builtin_nontemporal_load(ip + 64 * (k >> 2));
        const dbl2 a = __builtin_nontemporal_load(vp + 64 * (k >> 1));
        s0 = fma(a.x, xs[word & 255u], s0);
        if (k + 1 < w) s0 = fma(a.y, xs[(word >> 8) & 255u], s0);
        if (k + 2 < w) {
            const dbl2 b = __builtin_nontemporal_load(vp + 64 * ((k >> 1) + 1));
            s0 = fma(b.x, xs[(word >> 16) & 255u], s0);
        }
    }
    const int32_t r = order[t * 64 + lane];
    if (r < 0 || r >= nrows) return;
    const double sum = s0 + s1;
    if (mode == 1) {
        if (rowptr[r + 1] > rowptr[r]) y[r] = beta * sum;
    } else {
        const double base = (alpha == 0.0) ? 0.0 : alpha * v1[r];
        y[r] = base + beta * sum;
    }
}

// the rows of a lattice in blocks of 64 points: 4x4x4 (3D) or 8x8 (2D). Slot
// 64*b + l of block b (x fastest over the blocks) holds the DoF of the block's
// point l (x fastest inside the block), -1 where the block extends past the
// lattice
__global__ void k_block_order(int dim, int W, const int32_t* __restrict__ dof_of_lat, int32_t* __restrict__ order) {
    const int e = dim == 3 ? 4 : 8;
    const int nb = (W + e - 1) / e;
    const int64_t nblocks = dim == 3 ? (int64_t)nb * nb * nb : (int64_t)nb * nb;
    const int64_t s = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (s >= nblocks * 64) return;
    const int64_t b = s >> 6;
    const int l = (int)(s & 63);
    if (dim == 3) {
        const int bx = (int)(b % nb), by = (int)((b / nb) % nb), bz = (int)(b / ((int64_t)nb * nb));
        const int x = 4 * bx + (l & 3), y = 4 * by + ((l >> 2) & 3), z = 4 * bz + (l >> 4);
        order[s] = (x < W && y < W && z < W) ? dof_of_lat[((int64_t)z * W + y) * W + x] : -1;
    } else {
        const int bx = (int)(b % nb), by = (int)(b / nb);
        const int x = 8 * bx + (l & 7), y = 8 * by + (l >> 3);
        order[s] = (x < W && y < W) ? dof_of_lat[(int64_t)y * W + x] : -1;
    }
}

__global__ void k_dict_widths(int64_t nslices, int64_t nrows, const int32_t* __restrict__ rowptr,
                              const int32_t* __restrict__ order, int32_t* __restrict__ width) {
    const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (t >= nslices) return;
    int32_t w = 0;
    for (int l = 0; l < 64; ++l) {
        const int32_t r = order[t * 64 + l];
        if (r >= 0 && r < nrows) {
            const int32_t len = rowptr[r + 1] - rowptr[r];
            w = len > w ? len : w;
        }
    }
    width[t] = w;
}

}  // namespace

int dict_widths(int64_t nslices, int64_t nrows, const int32_t* rowptr, const int32_t* order, int32_t* width,
                hipStream_t s) {
    if (nslices <= 0) return 0;
    hipLaunchKernelGGL(k_dict_widths, dim3(nblk(nslices, 256)), dim3(256), 0, s, nslices, nrows, rowptr, order, width);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

int dict_count(int64_t nslices, int64_t nrows, const int32_t* rowptr, const int32_t* col, const int32_t* order,
               const int32_t* width, int32_t* ucount, hipStream_t s) {
    if (nslices <= 0) return 0;
    hipLaunchKernelGGL(k_dict_count, dim3((unsigned)nslices), dim3(64), 0, s, nrows, rowptr, col, order, width,
                       ucount);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

int dict_fill(int64_t nslices, int64_t nrows, const int32_t* rowptr, const int32_t* col, const double* val,
              const int32_t* order, const int32_t* width, const int32_t* dptr, const int64_t* vptr,
              const int64_t* iptr, int32_t* dict, double* sval, uint32_t* sidx, hipStream_t s) {
    if (nslices <= 0) return 0;
    hipLaunchKernelGGL(k_dict_fill, dim3((unsigned)nslices), dim3(64), 0, s, nrows, rowptr, col, val, order, width,
                       dptr, vptr, iptr, dict, sval, sidx);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

int dict_values(int64_t nslices, int64_t nrows, const int32_t* rowptr, const double* val, const int32_t* order,
                const int32_t* width, const int64_t* vptr, double* sval, hipStream_t s) {
    if (nslices <= 0) return 0;
    hipLaunchKernelGGL(k_dict_values, dim3((unsigned)nslices), dim3(64), 0, s, nrows, rowptr, val, order, width, vptr,
                       sval);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

int dict_spmv(int64_t nrows, int64_t nslices, const int32_t* width, const int32_t* dptr, const int32_t* dict,
              const int64_t* vptr, const int64_t* iptr, const double* sval, const uint32_t* sidx,
              const int32_t* order, double alpha, const double* v1, double beta, const double* x, double* y, int mode,
              const int32_t* rowptr, hipStream_t s) {
    if (nslices <= 0) return 0;
    hipLaunchKernelGGL(k_dict_spmv, dim3(nblk(nslices, 4)), dim3(256), 0, s, nrows, nslices, width, dptr, dict, vptr,
                       iptr, sval, sidx, order, alpha, v1, beta, x, y, mode, rowptr);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

int64_t lattice_block_slots(int dim, int W) {
    const int e = dim == 3 ? 4 : 8;
    const int64_t nb = (W + e - 1) / e;
    return (dim == 3 ? nb * nb * nb : nb * nb) * 64;
}

int lattice_block_order(int dim, int W, const int32_t* dof_of_lat, int32_t* order, hipStream_t s) {
    const int64_t n = lattice_block_slots(dim, W);
    hipLaunchKernelGGL(k_block_order, dim3(nblk(n, 256)), dim3(256), 0, s, dim, W, dof_of_lat, order);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

}  // namespace ugamd
