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
constexpr int kSortCap = 64 * kDictMaxWidth;

typedef double dbl2 __attribute__((ext_vector_type(2)));

inline unsigned nblk(int64_t n, int t) { return (unsigned)((n + t - 1) / t > 0 ? (n + t - 1) / t : 1); }

// the columns of slice t's entries sorted in LDS (pads INT_MAX); returns the
// number of sorted keys (a power of two)
__device__ int slice_sorted_columns(int64_t nrows, const int32_t* __restrict__ rowptr, const int32_t* __restrict__ col,
                                    const int32_t* __restrict__ order, int64_t t, int w, int32_t* keys) {
    int n = 64 * w, P = 1;
    while (P < n) P <<= 1;
    for (int e = threadIdx.x; e < P; e += blockDim.x) {
        int32_t v = INT_MAX;
        if (e < n) {
            const int k = e >> 6, lane = e & 63;
            const int32_t r = order[t * 64 + lane];
            if (r >= 0 && r < nrows) {
                const int32_t rb = rowptr[r], len = rowptr[r + 1] - rb;
                if (k < len) v = col[rb + k];
            }
        }
        keys[e] = v;
    }
    __syncthreads();
    // bitonic sort, ascending
    for (int size = 2; size <= P; size <<= 1)
        for (int stride = size >> 1; stride > 0; stride >>= 1) {
            for (int i = threadIdx.x; i < P; i += blockDim.x) {
                const int j = i ^ stride;
                if (j > i) {
                    const bool up = (i & size) == 0;
                    const int32_t a = keys[i], b = keys[j];
                    if ((a > b) == up) {
                        keys[i] = b;
                        keys[j] = a;
                    }
                }
            }
            __syncthreads();
        }
    return P;
}

// distinct columns per slice (INT_MAX if the slice's width exceeds the layout)
__global__ __launch_bounds__(256) void k_dict_count(int64_t nrows, const int32_t* __restrict__ rowptr,
                                                    const int32_t* __restrict__ col, const int32_t* __restrict__ order,
                                                    const int32_t* __restrict__ width, int32_t* __restrict__ ucount) {
    __shared__ int32_t keys[kSortCap];
    __shared__ int32_t red[4];
    const int64_t t = blockIdx.x;
    const int w = width[t];
    if (w > kDictMaxWidth) {
        if (threadIdx.x == 0) ucount[t] = INT_MAX;
        return;
    }
    const int P = slice_sorted_columns(nrows, rowptr, col, order, t, w, keys);
    int c = 0;
    for (int i = threadIdx.x; i < P; i += blockDim.x)
        c += (keys[i] != INT_MAX && (i == 0 || keys[i] != keys[i - 1])) ? 1 : 0;
    for (int o = 32; o > 0; o >>= 1) c += __shfl_down(c, o, 64);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = c;
    __syncthreads();
    if (threadIdx.x == 0) ucount[t] = red[0] + red[1] + red[2] + red[3];
}

// the dictionary, positions and values of every slice
__global__ __launch_bounds__(256) void k_dict_fill(int64_t nrows, const int32_t* __restrict__ rowptr,
                                                   const int32_t* __restrict__ col, const double* __restrict__ val,
                                                   const int32_t* __restrict__ order, const int32_t* __restrict__ width,
                                                   const int32_t* __restrict__ dptr, const int64_t* __restrict__ vptr,
                                                   const int64_t* __restrict__ iptr, int32_t* __restrict__ dict,
                                                   double* __restrict__ sval, uint32_t* __restrict__ sidx) {
    __shared__ int32_t keys[kSortCap];
    __shared__ int32_t uniq[kDictSlots];
    __shared__ int nuniq;
    const int64_t t = blockIdx.x;
    const int w = width[t];
    const int P = slice_sorted_columns(nrows, rowptr, col, order, t, w, keys);
    if (threadIdx.x == 0) {
        // sequential compaction of at most 255 distinct keys (the count pass
        // checked the bound)
        int u = 0;
        for (int i = 0; i < P && keys[i] != INT_MAX; ++i)
            if (i == 0 || keys[i] != keys[i - 1]) uniq[u++] = keys[i];
        nuniq = u;
    }
    __syncthreads();
    const int u = nuniq;
    for (int j = threadIdx.x; j < u; j += blockDim.x) dict[dptr[t] + j] = uniq[j];
    // one thread per (word, lane): four entries k = 4q .. 4q+3 of lane l
    const int nq = (w + 3) >> 2;
    for (int e = threadIdx.x; e < nq * 64; e += blockDim.x) {
        const int q = e >> 6, lane = e & 63;
        const int32_t r = order[t * 64 + lane];
        int32_t rb = 0, len = 0;
        if (r >= 0 && r < nrows) {
            rb = rowptr[r];
            len = rowptr[r + 1] - rb;
        }
        uint32_t word = 0;
        for (int b = 0; b < 4; ++b) {
            const int k = 4 * q + b;
            uint32_t pos = 255;
            double v = 0.0;
            if (k < len) {
                const int32_t c = col[rb + k];
                int lo = 0, hi = u - 1;
                while (lo < hi) {
                    const int mid = (lo + hi) >> 1;
                    if (uniq[mid] < c) lo = mid + 1;
                    else hi = mid;
                }
                pos = (uint32_t)lo;
                v = val[rb + k];
            }
            word |= pos << (8 * b);
            if (k < ((w + 1) & ~1)) sval[vptr[t] + 2 * (64 * (k >> 1) + lane) + (k & 1)] = v;
        }
        sidx[iptr[t] + 64 * q + lane] = word;
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
    hipLaunchKernelGGL(k_dict_count, dim3((unsigned)nslices), dim3(256), 0, s, nrows, rowptr, col, order, width,
                       ucount);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

int dict_fill(int64_t nslices, int64_t nrows, const int32_t* rowptr, const int32_t* col, const double* val,
              const int32_t* order, const int32_t* width, const int32_t* dptr, const int64_t* vptr,
              const int64_t* iptr, int32_t* dict, double* sval, uint32_t* sidx, hipStream_t s) {
    if (nslices <= 0) return 0;
    hipLaunchKernelGGL(k_dict_fill, dim3((unsigned)nslices), dim3(256), 0, s, nrows, rowptr, col, val, order, width,
                       dptr, vptr, iptr, dict, sval, sidx);
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
