// Host implementation of the GPUAlgebra classes (include/ug_gpu/gpu_algebra.hpp).
#include <atomic>
#include "ug_gpu/gpu_algebra.hpp"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <numeric>
#include <type_traits>

#include "hierarchy.hpp"
#include "ugx_internal.hpp"

namespace ug {
namespace gpu {

using namespace ugamd;

#define UGX_HIP(cmd)                                                                                      \
    do {                                                                                                  \
        hipError_t e_ = (cmd);                                                                            \
        if (e_ != hipSuccess) throw UGError(std::string("HIP error: ") + hipGetErrorString(e_) + " at " + \
                                            __FILE__ + ":" + std::to_string(__LINE__));                   \
    } while (0)
#define UGX_LAUNCH(cmd)                                                                          \
    do {                                                                                         \
        if ((cmd) != 0) throw UGError(std::string("kernel launch failed: ") + #cmd + " at line " + \
                                      std::to_string(__LINE__));                                 \
    } while (0)

static double wall() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

// ---------------------------------------------------------------------------
// context / buffers
// ---------------------------------------------------------------------------
static DeviceContext g_ctx;
static bool g_ctx_ready = false;

DeviceContext* context_if_ready() { return g_ctx_ready ? &g_ctx : nullptr; }

DeviceContext& context() {
    DeviceContext& ctx = g_ctx;
    bool& init = g_ctx_ready;
    if (!init) {
        int n = 0;
        UGX_HIP(hipGetDeviceCount(&n));
        if (n < 1) throw UGError("no HIP device available");
        UGX_HIP(hipGetDevice(&ctx.device));
        UGX_HIP(hipStreamCreateWithFlags(&ctx.stream, hipStreamNonBlocking));
        UGX_HIP(hipMalloc(&ctx.red_partial, 65536 * sizeof(double)));
        UGX_HIP(hipMalloc(&ctx.red_out, 8 * sizeof(double)));
        UGX_HIP(hipHostMalloc(&ctx.red_host, 8 * sizeof(double), hipHostMallocDefault));
        UGX_HIP(hipMalloc(&ctx.flag, sizeof(int)));
        UGX_HIP(hipMemset(ctx.flag, 0, sizeof(int)));
        init = true;
    }
    return ctx;
}

void sync() { UGX_HIP(hipStreamSynchronize(context().stream)); }

template <typename T>
void DeviceBuffer<T>::resize(int64_t n) {
    if (n == n_) return;
    release();
    if (n > 0) {
        context();
        UGX_HIP(hipMalloc(&p_, sizeof(T) * (size_t)n));
    }
    n_ = n;
}
template <typename T>
void DeviceBuffer<T>::release() {
    if (p_) (void)hipFree(p_);
    p_ = nullptr;
    n_ = 0;
}
template <typename T>
void DeviceBuffer<T>::upload(const T* h, int64_t n, int64_t offset) {
    if (n <= 0) return;
    if (offset + n > n_) throw UGError("DeviceBuffer::upload out of range");
    UGX_HIP(hipMemcpyAsync(p_ + offset, h, sizeof(T) * (size_t)n, hipMemcpyHostToDevice, context().stream));
    sync();
}
template <typename T>
void DeviceBuffer<T>::download(T* h, int64_t n, int64_t offset) const {
    if (n <= 0) return;
    if (offset + n > n_) throw UGError("DeviceBuffer::download out of range");
    UGX_HIP(hipMemcpyAsync(h, p_ + offset, sizeof(T) * (size_t)n, hipMemcpyDeviceToHost, context().stream));
    sync();
}
template <typename T>
void DeviceBuffer<T>::zero() {
    if (n_ > 0) UGX_HIP(hipMemsetAsync(p_, 0, sizeof(T) * (size_t)n_, context().stream));
}
template class DeviceBuffer<double>;
template class DeviceBuffer<float>;
template class DeviceBuffer<int32_t>;
template class DeviceBuffer<uint8_t>;
template class DeviceBuffer<uint16_t>;
template class DeviceBuffer<uint32_t>;
template class DeviceBuffer<int64_t>;

// ---------------------------------------------------------------------------
// GPUVector
// ---------------------------------------------------------------------------
GPUVector::GPUVector(const GPUVector& o) {
    resize(o.n_);
    if (n_) UGX_HIP(hipMemcpyAsync(buf_.get(), o.buf_.get(), sizeof(double) * n_, hipMemcpyDeviceToDevice, context().stream));
}
GPUVector& GPUVector::operator=(const GPUVector& o) {
    if (this == &o) return *this;
    resize(o.n_);
    if (n_) UGX_HIP(hipMemcpyAsync(buf_.get(), o.buf_.get(), sizeof(double) * n_, hipMemcpyDeviceToDevice, context().stream));
    return *this;
}
void GPUVector::resize(int64_t n) {
    buf_.resize(n);
    n_ = n;
}
void GPUVector::set(double v) { UGX_LAUNCH(vec_set<double>(n_, buf_.get(), v, context().stream)); }

static double reduce_to_host(int slot) {
    auto& c = context();
    UGX_HIP(hipMemcpyAsync(c.red_host + slot, c.red_out + slot, sizeof(double), hipMemcpyDeviceToHost, c.stream));
    sync();
    return c.red_host[slot];
}

double GPUVector::norm() const {
    auto& c = context();
    UGX_LAUNCH(dot<double>(n_, buf_.get(), buf_.get(), c.red_partial, c.red_out, true, c.stream));
    return reduce_to_host(0);
}
double GPUVector::dotprod(const GPUVector& b) const {
    if (b.n_ != n_) throw UGError("GPUVector::dotprod size mismatch");
    auto& c = context();
    UGX_LAUNCH(dot<double>(n_, buf_.get(), b.buf_.get(), c.red_partial, c.red_out, false, c.stream));
    return reduce_to_host(0);
}
GPUVector& GPUVector::operator+=(const GPUVector& b) {
    UGX_LAUNCH(vec_scale_add<double>(n_, buf_.get(), 1.0, buf_.get(), 1.0, b.buf_.get(), 0.0, nullptr, context().stream));
    return *this;
}
GPUVector& GPUVector::operator-=(const GPUVector& b) {
    UGX_LAUNCH(vec_scale_add<double>(n_, buf_.get(), 1.0, buf_.get(), -1.0, b.buf_.get(), 0.0, nullptr, context().stream));
    return *this;
}
GPUVector& GPUVector::operator*=(double a) {
    UGX_LAUNCH(vec_scale_add<double>(n_, buf_.get(), a, buf_.get(), 0.0, buf_.get(), 0.0, nullptr, context().stream));
    return *this;
}

void VecScaleAdd(GPUVector& dest, double a1, const GPUVector& v1, double a2, const GPUVector& v2) {
    UGX_LAUNCH(vec_scale_add<double>(dest.size(), dest.data(), a1, v1.data(), a2, v2.data(), 0.0, nullptr,
                                     context().stream));
}
void VecScaleAdd(GPUVector& dest, double a1, const GPUVector& v1, double a2, const GPUVector& v2, double a3,
                 const GPUVector& v3) {
    UGX_LAUNCH(vec_scale_add<double>(dest.size(), dest.data(), a1, v1.data(), a2, v2.data(), a3, v3.data(),
                                     context().stream));
}

// ---------------------------------------------------------------------------
// GPUSparseMatrix
// ---------------------------------------------------------------------------
static int lanes_for(int64_t rows, int64_t nnz) {
    const double avg = rows > 0 ? (double)nnz / (double)rows : 0.0;
    if (avg <= 6) return 4;
    if (avg <= 14) return 8;
    if (avg <= 40) return 16;
    return 32;
}

void GPUSparseMatrix::set_slice_order_device(const int32_t* order, int64_t nslots, bool by_construction) {
    if (!order) {
        slice_order_.resize(0);
        build_spmv_layout();
        return;
    }
    if (nslots <= 0 || nslots % 64 != 0) throw UGError("set_slice_order_device: the slots must be a multiple of 64");
    if (by_construction) {
        slice_order_.resize(nslots);
        UGX_HIP(hipMemcpyAsync(slice_order_.get(), order, sizeof(int32_t) * (size_t)nslots, hipMemcpyDeviceToDevice,
                               context().stream));
        build_spmv_layout();
        return;
    }
    // every row exactly once
    std::vector<int32_t> h((size_t)nslots);
    UGX_HIP(hipMemcpy(h.data(), order, sizeof(int32_t) * (size_t)nslots, hipMemcpyDeviceToHost));
    std::vector<uint8_t> seen((size_t)rows_, 0);
    int64_t cnt = 0;
    for (int32_t r : h) {
        if (r < 0) continue;
        if (r >= rows_ || seen[(size_t)r]) throw UGError("set_slice_order_device: not a placement of the rows");
        seen[(size_t)r] = 1;
        ++cnt;
    }
    if (cnt != rows_) throw UGError("set_slice_order_device: rows are missing from the slices");
    slice_order_.resize(nslots);
    slice_order_.upload(h.data(), nslots);
    build_spmv_layout();
}

bool GPUSparseMatrix::build_dictionary() {
    auto& c = context();
    const int64_t ns = slice_order_.size() / 64;
    d_width_.resize(ns);
    UGX_LAUNCH(dict_widths(ns, rows_, rowptr_.get(), slice_order_.get(), d_width_.get(), c.stream));
    DeviceBuffer<int32_t> cnt(ns);
    UGX_LAUNCH(dict_count(ns, rows_, rowptr_.get(), col_.get(), slice_order_.get(), d_width_.get(), cnt.get(),
                          c.stream));
    std::vector<int32_t> w((size_t)ns), u((size_t)ns);
    d_width_.download(w.data(), ns);
    cnt.download(u.data(), ns);
    for (int64_t t = 0; t < ns; ++t)
        if (u[(size_t)t] > 255) return false;   // too many columns (or too wide) for one-byte positions
    std::vector<int32_t> dptr((size_t)ns + 1, 0);
    std::vector<int64_t> vptr((size_t)ns + 1, 0), iptr((size_t)ns + 1, 0);
    for (int64_t t = 0; t < ns; ++t) {
        dptr[(size_t)t + 1] = dptr[(size_t)t] + u[(size_t)t];
        vptr[(size_t)t + 1] = vptr[(size_t)t] + (int64_t)128 * ((w[(size_t)t] + 1) / 2);
        iptr[(size_t)t + 1] = iptr[(size_t)t] + (int64_t)64 * ((w[(size_t)t] + 3) / 4);
    }
    d_dptr_.resize(ns + 1);
    d_dptr_.upload(dptr.data(), ns + 1);
    d_vptr_.resize(ns + 1);
    d_vptr_.upload(vptr.data(), ns + 1);
    d_iptr_.resize(ns + 1);
    d_iptr_.upload(iptr.data(), ns + 1);
    d_dict_.resize(std::max<int64_t>(dptr[(size_t)ns], 1));
    d_val_.resize(std::max<int64_t>(vptr[(size_t)ns], 2));
    d_idx_.resize(std::max<int64_t>(iptr[(size_t)ns], 1));
    UGX_LAUNCH(dict_fill(ns, rows_, rowptr_.get(), col_.get(), val_.get(), slice_order_.get(), d_width_.get(),
                         d_dptr_.get(), d_vptr_.get(), d_iptr_.get(), d_dict_.get(), d_val_.get(), d_idx_.get(),
                         c.stream));
    sync();
    return true;
}

uint64_t GPUSparseMatrix::fresh_pattern_id() {
    static std::atomic<uint64_t> next{1};
    return next.fetch_add(1);
}

void GPUSparseMatrix::refresh_layout_values() {
    if (rows_ == 0 || building_) return;
    if (layout_pattern_id_ != pattern_id_ || (!dict_valid_ && !sell_valid_)) {
        build_spmv_layout();
        return;
    }
    auto& c = context();
    if (dict_valid_) {
        UGX_LAUNCH(dict_values(slice_order_.size() / 64, rows_, rowptr_.get(), val_.get(), slice_order_.get(),
                               d_width_.get(), d_vptr_.get(), d_val_.get(), c.stream));
    } else {
        const int32_t* order = has_row_order() ? row_order_.get() : nullptr;
        UGX_LAUNCH(sell_fill<double>(rows_, cols_, rowptr_.get(), col_.get(), val_.get(), order, sell_ptr_.get(),
                                     sell_width_.get(), sell_col_.get(), sell_val_.get(), c.stream, true));
    }
    sync();
}

void GPUSparseMatrix::build_spmv_layout() {
    sell_valid_ = false;
    dict_valid_ = false;
    layout_pattern_id_ = pattern_id_;
    if (rows_ == 0 || building_) return;
    if (dict_enabled_ && slice_order_.size() > 0 && build_dictionary()) {
        dict_valid_ = true;
        // the sliced-ELL copy is not needed next to the dictionary layout
        sell_col_.release();
        sell_val_.release();
        return;
    }
    d_val_.release();
    d_idx_.release();
    if (!sell_enabled_) return;
    auto& c = context();
    const int64_t nsl = (rows_ + 63) / 64;
    sell_width_.resize(nsl);
    const int32_t* order = has_row_order() ? row_order_.get() : nullptr;
    UGX_LAUNCH(sell_widths(rows_, rowptr_.get(), order, sell_width_.get(), c.stream));
    std::vector<int32_t> w(nsl);
    sell_width_.download(w.data(), nsl);
    std::vector<int64_t> ptr(nsl + 1, 0);
    for (int64_t t = 0; t < nsl; ++t) ptr[t + 1] = ptr[t] + (int64_t)64 * w[t];
    sell_ptr_.resize(nsl + 1);
    sell_ptr_.upload(ptr.data(), nsl + 1);
    sell_col_.resize(ptr[nsl]);
    sell_val_.resize(ptr[nsl]);
    UGX_LAUNCH(sell_fill<double>(rows_, cols_, rowptr_.get(), col_.get(), val_.get(), order, sell_ptr_.get(),
                                 sell_width_.get(), sell_col_.get(), sell_val_.get(), c.stream));
    sync();
    sell_valid_ = true;
}

void GPUSparseMatrix::allocate(int64_t rows, int64_t cols, int64_t nnz) {
    building_ = false;
    pattern_id_ = fresh_pattern_id();
    std::vector<std::vector<std::pair<int32_t, double>>>().swap(hrows_);
    sell_valid_ = false;
    dict_valid_ = false;
    slice_order_.resize(0);
    row_order_.resize(0);
    if (nnz >= ((int64_t)1 << 31)) throw UGError("GPUSparseMatrix: nnz exceeds int32 range");
    rows_ = rows;
    cols_ = cols;
    nnz_ = nnz;
    rowptr_.resize(rows + 1);
    col_.resize(nnz);
    val_.resize(nnz);
    lanes_ = lanes_for(rows, nnz);
}

void GPUSparseMatrix::set_row_order(const int32_t* order, int64_t n) {
    if (!order) {
        row_order_.resize(0);
        build_spmv_layout();
        return;
    }
    if (n != rows_) throw UGError("set_row_order: the order must list every row once");
    std::vector<uint8_t> seen(rows_, 0);
    for (int64_t i = 0; i < n; ++i) {
        const int32_t r = order[i];
        if (r < 0 || r >= rows_ || seen[r]) throw UGError("set_row_order: not a permutation of the rows");
        seen[r] = 1;
    }
    row_order_.resize(n);
    row_order_.upload(order, n);
    build_spmv_layout();
}

void GPUSparseMatrix::set_row_order_device(const int32_t* order, int64_t n) {
    if (n != rows_) throw UGError("set_row_order_device: the order must list every row once");
    row_order_.resize(n);
    UGX_HIP(hipMemcpyAsync(row_order_.get(), order, sizeof(int32_t) * (size_t)n, hipMemcpyDeviceToDevice,
                           context().stream));
    sync();
}

void GPUSparseMatrix::set_from_host_csr(int64_t rows, int64_t cols, const int32_t* rowptr, const int32_t* col,
                                        const double* val) {
    allocate(rows, cols, rowptr[rows]);
    set_origin(nullptr);
    rowptr_.upload(rowptr, rows + 1);
    col_.upload(col, nnz_);
    val_.upload(val, nnz_);
    build_spmv_layout();
}

void GPUSparseMatrix::require_device(const char* what) const {
    if (building_)
        throw UGError(std::string("GPUSparseMatrix::") + what +
                      ": the matrix is being built on the host (call finalize() first)");
}

void GPUSparseMatrix::values_changed_on_device() {
    set_origin(nullptr);
    refresh_layout_values();
}

void GPUSparseMatrix::resize_and_clear(int64_t rows, int64_t cols) {
    if (rows < 0 || cols < 0 || rows >= ((int64_t)1 << 31) || cols >= ((int64_t)1 << 31))
        throw UGError("GPUSparseMatrix::resize_and_clear: sizes must be in [0, 2^31)");
    // release the device CSR and its derived layouts
    rowptr_.release();
    col_.release();
    val_.release();
    sell_width_.release();
    sell_col_.release();
    sell_val_.release();
    sell_ptr_.release();
    row_order_.release();
    sell_valid_ = false;
    slice_order_.release();
    d_dict_.release();
    d_val_.release();
    d_idx_.release();
    dict_valid_ = false;
    pattern_id_ = fresh_pattern_id();
    t_rowptr_.release();
    t_col_.release();
    t_val_.release();
    set_origin(nullptr);
    rows_ = rows;
    cols_ = cols;
    nnz_ = 0;
    hrows_.assign((size_t)rows, {});
    building_ = true;
}

void GPUSparseMatrix::resize_and_keep_values(int64_t rows, int64_t cols) {
    if (rows < 0 || cols < 0 || rows >= ((int64_t)1 << 31) || cols >= ((int64_t)1 << 31))
        throw UGError("GPUSparseMatrix::resize_and_keep_values: sizes must be in [0, 2^31)");
    if (!building_) {
        // move the device CSR into a host build
        std::vector<int32_t> rp(rows_ + 1), cl(nnz_);
        std::vector<double> vl(nnz_);
        if (rows_ > 0) download_csr(rp.data(), cl.data(), vl.data());
        const int64_t old_rows = rows_;
        resize_and_clear(old_rows, cols_);
        for (int64_t r = 0; r < old_rows; ++r) {
            auto& row = hrows_[(size_t)r];
            row.reserve((size_t)(rp[r + 1] - rp[r]));
            for (int32_t k = rp[r]; k < rp[r + 1]; ++k) row.emplace_back(cl[k], vl[k]);
        }
        nnz_ = (int64_t)cl.size();
    }
    hrows_.resize((size_t)rows);
    nnz_ = 0;
    for (auto& row : hrows_) {
        if (cols < cols_) {
            auto it = std::lower_bound(row.begin(), row.end(), std::make_pair((int32_t)cols, -HUGE_VAL));
            row.erase(it, row.end());
        }
        nnz_ += (int64_t)row.size();
    }
    rows_ = rows;
    cols_ = cols;
}

void GPUSparseMatrix::clear_retain_structure() {
    if (building_) {
        for (auto& row : hrows_)
            for (auto& e : row) e.second = 0.0;
        return;
    }
    // written, not scaled: 0 * NaN would stay NaN
    if (nnz_ > 0) UGX_LAUNCH(vec_set<double>(nnz_, val_.get(), 0.0, context().stream));
    values_changed_on_device();
}

double& GPUSparseMatrix::operator()(int64_t r, int64_t c) {
    if (!building_)
        throw UGError("GPUSparseMatrix::operator(): the pattern is finalised on the device; entries are created "
                      "only between resize_and_clear and finalize");
    if (r < 0 || r >= rows_ || c < 0 || c >= cols_) throw UGError("GPUSparseMatrix::operator(): index out of range");
    auto& row = hrows_[(size_t)r];
    auto it = std::lower_bound(row.begin(), row.end(), (int32_t)c,
                               [](const std::pair<int32_t, double>& e, int32_t col) { return e.first < col; });
    if (it == row.end() || it->first != (int32_t)c) {
        it = row.insert(it, std::make_pair((int32_t)c, 0.0));
        ++nnz_;
    }
    return it->second;
}

int64_t GPUSparseMatrix::num_connections(int64_t r) const {
    if (r < 0 || r >= rows_) throw UGError("GPUSparseMatrix::num_connections: row out of range");
    if (building_) return (int64_t)hrows_[(size_t)r].size();
    int32_t b[2];
    rowptr_.download(b, 2, r);
    return b[1] - b[0];
}

void GPUSparseMatrix::row(int64_t r, int32_t* cols, double* vals) const {
    if (r < 0 || r >= rows_) throw UGError("GPUSparseMatrix::row: row out of range");
    if (building_) {
        const auto& row = hrows_[(size_t)r];
        for (size_t k = 0; k < row.size(); ++k) {
            if (cols) cols[k] = row[k].first;
            if (vals) vals[k] = row[k].second;
        }
        return;
    }
    int32_t b[2];
    rowptr_.download(b, 2, r);
    if (b[1] > b[0]) {
        if (cols) col_.download(cols, b[1] - b[0], b[0]);
        if (vals) val_.download(vals, b[1] - b[0], b[0]);
    }
}

double GPUSparseMatrix::get(int64_t r, int64_t c) const {
    if (r < 0 || r >= rows_ || c < 0 || c >= cols_) throw UGError("GPUSparseMatrix::get: index out of range");
    const int64_t n = num_connections(r);
    std::vector<int32_t> cl((size_t)n);
    std::vector<double> vl((size_t)n);
    row(r, cl.data(), vl.data());
    auto it = std::lower_bound(cl.begin(), cl.end(), (int32_t)c);
    return (it != cl.end() && *it == (int32_t)c) ? vl[(size_t)(it - cl.begin())] : 0.0;
}

bool GPUSparseMatrix::has_connection(int64_t r, int64_t c) const {
    if (r < 0 || r >= rows_ || c < 0 || c >= cols_) return false;
    const int64_t n = num_connections(r);
    std::vector<int32_t> cl((size_t)n);
    row(r, cl.data(), nullptr);
    return std::binary_search(cl.begin(), cl.end(), (int32_t)c);
}

void GPUSparseMatrix::finalize() {
    if (building_) {
        std::vector<int32_t> rp((size_t)rows_ + 1, 0), cl;
        std::vector<double> vl;
        int64_t nnz = 0;
        for (const auto& row : hrows_) nnz += (int64_t)row.size();
        if (nnz >= ((int64_t)1 << 31)) throw UGError("GPUSparseMatrix::finalize: nnz exceeds int32 range");
        cl.reserve((size_t)nnz);
        vl.reserve((size_t)nnz);
        for (int64_t r = 0; r < rows_; ++r) {
            for (const auto& e : hrows_[(size_t)r]) {
                cl.push_back(e.first);
                vl.push_back(e.second);
            }
            rp[(size_t)r + 1] = (int32_t)cl.size();
        }
        building_ = false;
        std::vector<std::vector<std::pair<int32_t, double>>>().swap(hrows_);
        set_from_host_csr(rows_, cols_, rp.data(), cl.data(), vl.data());
        return;
    }
    set_origin(nullptr);
    build_spmv_layout();
}

void GPUSparseMatrix::transpose_from(const GPUSparseMatrix& B, double scale, bool drop_zeros) {
    B.require_device("set_as_transpose_of");
    if (&B == this) throw UGError("GPUSparseMatrix::set_as_transpose_of: source and destination are the same matrix");
    auto& c = context();
    const int64_t nnz = B.nnz_;
    DeviceBuffer<uint32_t> key(nnz), key2(nnz);
    DeviceBuffer<int32_t> idx(nnz), idx2(nnz), row_of(nnz);
    size_t tb = 0;
    if (transpose_sort_bytes(nnz, (int32_t)B.cols_, &tb) != 0) throw UGError("set_as_transpose_of: sort set-up failed");
    DeviceBuffer<uint8_t> temp((int64_t)tb);
    DeviceBuffer<int32_t> rpt(B.cols_ + 1);
    UGX_LAUNCH(transpose_csr(B.rows_, (int32_t)B.cols_, nnz, B.rowptr_.get(), B.col_.get(), B.val_.get(),
                             drop_zeros ? 1 : 0, key.get(), key2.get(), idx.get(), idx2.get(), row_of.get(),
                             temp.get(), tb, rpt.get(), c.stream));
    int32_t kept = 0;
    rpt.download(&kept, 1, B.cols_);
    allocate(B.cols_, B.rows_, kept);
    set_origin(nullptr);
    UGX_HIP(hipMemcpyAsync(rowptr_.get(), rpt.get(), sizeof(int32_t) * (size_t)(rows_ + 1), hipMemcpyDeviceToDevice,
                           c.stream));
    UGX_LAUNCH(transpose_gather(kept, idx2.get(), row_of.get(), B.val_.get(), scale, col_.get(), val_.get(),
                                c.stream));
    sync();
    build_spmv_layout();
}

void GPUSparseMatrix::set_as_copy_of(const GPUSparseMatrix& B, double scale) {
    B.require_device("set_as_copy_of");
    if (&B == this) {
        if (scale != 1.0) this->scale(scale);
        return;
    }
    auto& c = context();
    allocate(B.rows_, B.cols_, B.nnz_);
    set_origin(nullptr);
    UGX_HIP(hipMemcpyAsync(rowptr_.get(), B.rowptr_.get(), sizeof(int32_t) * (size_t)(rows_ + 1),
                           hipMemcpyDeviceToDevice, c.stream));
    if (nnz_ > 0) {
        UGX_HIP(hipMemcpyAsync(col_.get(), B.col_.get(), sizeof(int32_t) * (size_t)nnz_, hipMemcpyDeviceToDevice,
                               c.stream));
        UGX_LAUNCH(scale_values(nnz_, B.val_.get(), scale, val_.get(), c.stream));
    }
    if (B.has_row_order()) {
        row_order_.resize(rows_);
        UGX_HIP(hipMemcpyAsync(row_order_.get(), B.row_order_.get(), sizeof(int32_t) * (size_t)rows_,
                               hipMemcpyDeviceToDevice, c.stream));
    }
    sync();
    build_spmv_layout();
}

void GPUSparseMatrix::scale(double a) {
    if (building_) {
        for (auto& row : hrows_)
            for (auto& e : row) e.second *= a;
        return;
    }
    if (nnz_ > 0) UGX_LAUNCH(scale_values(nnz_, val_.get(), a, val_.get(), context().stream));
    values_changed_on_device();
}

void GPUSparseMatrix::axpy_transposed(GPUVector& dest, double a1, const GPUVector& v1, double b1,
                                      const GPUVector& w1) const {
    require_device("axpy_transposed");
    if (dest.size() != cols_ || v1.size() != cols_ || w1.size() != rows_)
        throw UGError("GPUSparseMatrix::axpy_transposed size mismatch");
    if (&dest == &w1) throw UGError("GPUSparseMatrix::axpy_transposed: dest must not alias the multiplied vector");
    auto& c = context();
    if (t_version_ != version_ || t_rowptr_.size() != cols_ + 1) {
        // the transpose without zero entries (the product skips them)
        const int64_t nnz = nnz_;
        DeviceBuffer<uint32_t> key(nnz), key2(nnz);
        DeviceBuffer<int32_t> idx(nnz), idx2(nnz), row_of(nnz);
        size_t tb = 0;
        if (transpose_sort_bytes(nnz, (int32_t)cols_, &tb) != 0) throw UGError("axpy_transposed: sort set-up failed");
        DeviceBuffer<uint8_t> temp((int64_t)tb);
        t_rowptr_.resize(cols_ + 1);
        UGX_LAUNCH(transpose_csr(rows_, (int32_t)cols_, nnz, rowptr_.get(), col_.get(), val_.get(), 1, key.get(),
                                 key2.get(), idx.get(), idx2.get(), row_of.get(), temp.get(), tb, t_rowptr_.get(),
                                 c.stream));
        int32_t kept = 0;
        t_rowptr_.download(&kept, 1, cols_);
        t_col_.resize(kept);
        t_val_.resize(kept);
        UGX_LAUNCH(transpose_gather(kept, idx2.get(), row_of.get(), val_.get(), 1.0, t_col_.get(), t_val_.get(),
                                    c.stream));
        sync();
        t_nnz_ = kept;
        t_version_ = version_;
    }
    UGX_LAUNCH(ugamd::axpy_transposed(cols_, t_rowptr_.get(), t_col_.get(), t_val_.get(), a1, v1.data(),
                                      &dest == &v1, b1, w1.data(), dest.data(), c.stream));
}

void GPUSparseMatrix::set_dirichlet_rows(const int32_t* rows_host, int64_t n) {
    if (n <= 0) return;
    for (int64_t i = 0; i < n; ++i)
        if (rows_host[i] < 0 || rows_host[i] >= rows_ || rows_host[i] >= cols_)
            throw UGError("GPUSparseMatrix::set_dirichlet_rows: row out of range");
    if (building_) {
        for (int64_t i = 0; i < n; ++i) {
            for (auto& e : hrows_[(size_t)rows_host[i]]) e.second = 0.0;
            (*this)(rows_host[i], rows_host[i]) = 1.0;
        }
        return;
    }
    auto& c = context();
    DeviceBuffer<int32_t> rows(n);
    rows.upload(rows_host, n);
    UGX_HIP(hipMemsetAsync(c.flag, 0, sizeof(int), c.stream));
    UGX_LAUNCH(missing_diagonals(n, rows.get(), rowptr_.get(), col_.get(), c.flag, c.stream));
    int missing = 0;
    UGX_HIP(hipMemcpyAsync(&missing, c.flag, sizeof(int), hipMemcpyDeviceToHost, c.stream));
    sync();
    if (missing)
        throw UGError("GPUSparseMatrix::set_dirichlet_rows: a listed row stores no diagonal entry (the device pattern "
                      "is fixed; add the diagonal while building the pattern)");
    UGX_LAUNCH(dirichlet_rows(n, rows.get(), rowptr_.get(), col_.get(), val_.get(), c.stream));
    sync();
    values_changed_on_device();
}

void GPUSparseMatrix::download_csr(int32_t* rowptr, int32_t* col, double* val) const {
    require_device("download_csr");
    if (rowptr) rowptr_.download(rowptr, rows_ + 1);
    if (col) col_.download(col, nnz_);
    if (val) val_.download(val, nnz_);
}

void GPUSparseMatrix::axpy(GPUVector& dest, double a1, const GPUVector& v1, double b1, const GPUVector& w1) const {
    require_device("axpy");
    if (dest.size() != rows_ || w1.size() != cols_) throw UGError("GPUSparseMatrix::axpy size mismatch");
    if (&dest == &w1) throw UGError("GPUSparseMatrix::axpy: dest must not alias the multiplied vector");
    if (dict_valid_)
        UGX_LAUNCH(dict_spmv(rows_, slice_order_.size() / 64, d_width_.get(), d_dptr_.get(), d_dict_.get(),
                             d_vptr_.get(), d_iptr_.get(), d_val_.get(), d_idx_.get(), slice_order_.get(), a1,
                             v1.data(), b1, w1.data(), dest.data(), 0, rowptr_.get(), context().stream));
    else if (sell_valid_)
        UGX_LAUNCH((sell_spmv<double, double>(rows_, sell_ptr_.get(), sell_width_.get(), sell_col_.get(),
                                              sell_val_.get(), a1, v1.data(), b1, w1.data(), dest.data(), 0,
                                              rowptr_.get(), has_row_order() ? row_order_.get() : nullptr,
                                              context().stream)));
    else
        UGX_LAUNCH((spmv_csr<double, double>(rows_, rowptr_.get(), col_.get(), val_.get(), a1, v1.data(), b1,
                                             w1.data(), dest.data(), 0, lanes_, context().stream)));
}

void GPUSparseMatrix::apply_ignore_zero_rows(GPUVector& dest, double beta, const GPUVector& w) const {
    require_device("apply_ignore_zero_rows");
    if (dict_valid_)
        UGX_LAUNCH(dict_spmv(rows_, slice_order_.size() / 64, d_width_.get(), d_dptr_.get(), d_dict_.get(),
                             d_vptr_.get(), d_iptr_.get(), d_val_.get(), d_idx_.get(), slice_order_.get(), 0.0,
                             dest.data(), beta, w.data(), dest.data(), 1, rowptr_.get(), context().stream));
    else if (sell_valid_)
        UGX_LAUNCH((sell_spmv<double, double>(rows_, sell_ptr_.get(), sell_width_.get(), sell_col_.get(),
                                              sell_val_.get(), 0.0, dest.data(), beta, w.data(), dest.data(), 1,
                                              rowptr_.get(), has_row_order() ? row_order_.get() : nullptr,
                                              context().stream)));
    else
        UGX_LAUNCH((spmv_csr<double, double>(rows_, rowptr_.get(), col_.get(), val_.get(), 0.0, dest.data(), beta,
                                             w.data(), dest.data(), 1, lanes_, context().stream)));
}

// ---------------------------------------------------------------------------
// iterators / solvers
// ---------------------------------------------------------------------------
bool ILinearIterator::apply_update_defect(GPUVector& c, GPUVector& d) {
    if (!apply(c, d)) return false;
    op_->apply_sub(d, c);
    return true;
}

void StdConvCheck::start_defect(double d) {
    defects_.assign(1, d);
    initial_ = current_ = d;
    step_ = 0;
    if (verbose_) std::printf("%% %s\n  Iter      Defect         Rate\n%4d:    %e      -------\n", name_.c_str(), 0, d);
}
void StdConvCheck::update_defect(double d) {
    last_ = current_;
    current_ = d;
    ++step_;
    defects_.push_back(d);
    if (verbose_) std::printf("%4d:    %e    %e\n", step_, d, d / last_);
}
bool StdConvCheck::iteration_ended() const {
    if (!std::isfinite(current_)) return true;
    if (step_ >= max_steps_) return true;
    if (current_ < min_defect_) return true;
    if (reduction() < rel_reduction_) return true;
    return false;
}
bool StdConvCheck::post() const { return current_ < min_defect_ || reduction() < rel_reduction_; }

bool Jacobi::init(std::shared_ptr<MatrixOperator> J) {
    op_ = J;
    auto& A = J->get_matrix();
    dinv_.resize(A.num_rows());
    auto& c = context();
    UGX_HIP(hipMemsetAsync(c.flag, 0, sizeof(int), c.stream));
    UGX_LAUNCH((jacobi_diag_csr<double, double>(A.num_rows(), A.rowptr(), A.col(), A.val(), omega_, dinv_.get(),
                                                c.flag, c.stream)));
    int bad = 0;
    UGX_HIP(hipMemcpyAsync(&bad, c.flag, sizeof(int), hipMemcpyDeviceToHost, c.stream));
    sync();
    return bad == 0;
}

bool Jacobi::apply(GPUVector& c, const GPUVector& d) {
    UGX_LAUNCH((diag_mult<double, double>(d.size(), c.data(), dinv_.get(), d.data(), context().stream)));
    return true;
}

bool ILinearOperatorInverse::init(std::shared_ptr<MatrixOperator> J) {
    op_ = J;
    if (precond_ && !precond_->init(J)) return false;
    return true;
}

bool ILinearOperatorInverse::apply(GPUVector& x, const GPUVector& b) {
    GPUVector r(b);
    return apply_return_defect(x, r);
}

static double dev_dot(int64_t n, const double* a, const double* b, bool take_sqrt) {
    auto& c = context();
    UGX_LAUNCH(dot<double>(n, a, b, c.red_partial, c.red_out, take_sqrt, c.stream));
    return reduce_to_host(0);
}

// number of nonzero entries of a[0, n)
static double dev_count_nonzero(int64_t n, const double* a) {
    auto& c = context();
    UGX_LAUNCH(count_nonzero(n, a, c.red_partial, c.red_out, c.stream));
    return reduce_to_host(0);
}

// How the lattice path can start for x and b (one host synchronisation for
// three nonzero counts): 1 = x is zero and b vanishes on the Dirichlet DoFs (r = b);
// 2 = x and b both vanish there (r = b - A x with the Dirichlet-eliminated
// lattice operator, which equals A x for such x); 0 = neither.
int CG::lattice_start_kind(const GPUVector& x, const GPUVector& b, int64_t ni) {
    auto& c = context();
    const int64_t N = x.size(), nt = N - ni;
    if (b.size() != N || nt < 0) return 0;
    // nonzero counts (x == 0 exactly, not x.x == 0, which underflows)
    UGX_LAUNCH(count_nonzero(N, x.data(), c.red_partial, c.red_out + 0, c.stream));
    if (nt > 0) {
        UGX_LAUNCH(count_nonzero(nt, x.data() + ni, c.red_partial, c.red_out + 1, c.stream));
        UGX_LAUNCH(count_nonzero(nt, b.data() + ni, c.red_partial, c.red_out + 2, c.stream));
    } else {
        UGX_LAUNCH(vec_set<double>(2, c.red_out + 1, 0.0, c.stream));
    }
    UGX_HIP(hipMemcpyAsync(c.red_host, c.red_out, 3 * sizeof(double), hipMemcpyDeviceToHost, c.stream));
    sync();
    if (c.red_host[2] != 0.0) return 0;
    if (c.red_host[0] == 0.0) return 1;
    return c.red_host[1] == 0.0 ? 2 : 0;
}

bool CG::apply(GPUVector& x, const GPUVector& b) {
    // the lattice path straight from b: no defect copy, and the final defect
    // (which apply does not return) stays in lattice order
    used_fast_ = false;
    if (fast_ && op_) {
        auto* gmg = dynamic_cast<AssembledMultiGridCycle*>(precond_.get());
        if (gmg && gmg->lattice_ready(op_->get_matrix())) {
            const int kind = lattice_start_kind(x, b, gmg->num_inner_dofs());
            if (kind) {
                used_fast_ = true;
                return apply_lattice(x, b, *gmg, kind, nullptr);
            }
        }
    }
    return ILinearOperatorInverse::apply(x, b);
}

bool CG::apply_lattice(GPUVector& x, const GPUVector& src, AssembledMultiGridCycle& gmg, int kind,
                       GPUVector* defect_out) {
    // The iteration of cg.h:103-242 on lattice-ordered copies of x and r with
    // the surface operator and the preconditioner applied in lattice order.
    // Valid because p and z vanish on Dirichlet DoFs (r does on entry).
    // The scalars rho, lambda, alpha, beta stay in device memory; the host
    // reads |r| once per step (the convergence check) together with the
    // breakdown flag. (q, p) is summed inside the operator kernel, (z, r)
    // inside the cycle's final conversion, and x, r, |r|^2 are updated by one
    // kernel.
    auto& c = context();
    const int64_t N = gmg.lattice_size(), H = gmg.lattice_halo();
    if (ws_n_ != N || ws_h_ != H) {
        for (auto* b : {&ws_x_, &ws_r_, &ws_p_, &ws_q_, &ws_z_}) {
            b->resize(N + 2 * H);
            b->zero();   // the halos; the kernels write only [0, N)
        }
        ws_partial_.resize(AssembledMultiGridCycle::max_partials());
        ws_sc_.resize(8);
        ws_n_ = N;
        ws_h_ = H;
    }
    double *xl = ws_x_.get() + H, *rl = ws_r_.get() + H, *pl = ws_p_.get() + H, *ql = ws_q_.get() + H,
           *zl = ws_z_.get() + H;
    double* sc = ws_sc_.get();
    double* part = ws_partial_.get();
    ws_sc_.zero();
    // an FP32 cycle starts from an FP32 copy of r that cg_update writes along
    float* rfl = nullptr;
    if (gmg.lattice_uses_f32_defect()) {
        if (ws_rf_.size() != N + 2 * H) {
            ws_rf_.resize(N + 2 * H);
            ws_rf_.zero();   // the halos
        }
        rfl = ws_rf_.get() + H;
    }
    // the defect in lattice order: src is the defect (kind 0) or b (kinds 1, 2)
    if (kind == 1) {
        UGX_LAUNCH(vec_set<double>(N, xl, 0.0, c.stream));
    } else {
        UGX_LAUNCH((permute_to<double, double>(N, gmg.dof_of_lattice(), x.data(), xl, c.stream)));
    }
    // the FP32 copy of the defect comes with the permutation unless the
    // defect is formed after it (kind 2)
    const float* rf_first = nullptr;
    if (rfl && kind != 2) {
        UGX_LAUNCH(permute_to_dual(N, gmg.dof_of_lattice(), src.data(), rl, rfl, c.stream));
        rf_first = rfl;
    } else {
        UGX_LAUNCH((permute_to<double, double>(N, gmg.dof_of_lattice(), src.data(), rl, c.stream)));
    }
    if (kind == 2) gmg.surface_axpy_lattice(rl, 1.0, rl, -1.0, xl);   // r = b - A x
    // p = z = B r, rho = (z, r): the first z is written where p lives (the
    // loop's z buffer is written afresh before it is read)
    gmg.apply_lattice(pl, rl, sc + 0, part, rf_first);
    conv_->set_name(name());
    conv_->start_defect(dev_dot(N, rl, rl, true));
    // second direction buffer: the fused update writes p_new beside p_old
    if (ws_p2_.size() != N + 2 * H) {
        ws_p2_.resize(N + 2 * H);
        ws_p2_.zero();   // the halos
    }
    double* pl2 = ws_p2_.get() + H;
    bool q_ready = false;   // q = A p and lambda already formed with the direction update
    bool ok = true;
    while (!conv_->iteration_ended()) {
        // q = A p, lambda = (q, p)
        if (!q_ready && !gmg.surface_axpy_lattice(ql, 0.0, ql, 1.0, pl, sc + 1, part)) {
            gmg.surface_axpy_lattice(ql, 0.0, ql, 1.0, pl);
            UGX_LAUNCH(dot<double>(N, ql, pl, c.red_partial, sc + 1, false, c.stream));
        }
        q_ready = false;
        UGX_LAUNCH(cg_update(N, sc, pl, ql, xl, rl, part, c.stream, rfl));
        UGX_HIP(hipMemcpyAsync(c.red_host, sc + 5, 2 * sizeof(double), hipMemcpyDeviceToHost, c.stream));
        sync();
        if (c.red_host[1] != 0.0) {   // lambda == 0: breakdown (x, r unchanged)
            ok = false;
            break;
        }
        conv_->update_defect(std::sqrt(c.red_host[0]));
        if (conv_->iteration_ended()) break;
        gmg.apply_lattice(zl, rl, sc + 4, part, rfl);   // z = B r, rho_new = (z, r)
        UGX_LAUNCH(cg_beta(sc, c.stream));          // beta = rho_new / rho, rho = rho_new
        // p = z + beta p, fused with the next q = A p when the operator allows
        if (gmg.surface_direction_axpy_lattice(ql, pl, zl, sc + 3, pl2, sc + 1, part)) {
            std::swap(pl, pl2);
            q_ready = true;
        } else {
            UGX_LAUNCH(cg_direction(N, sc, pl, zl, c.stream));
        }
    }
    UGX_LAUNCH((permute_scatter<double, double>(N, gmg.dof_of_lattice(), xl, x.data(), c.stream)));
    if (defect_out)
        UGX_LAUNCH((permute_scatter<double, double>(N, gmg.dof_of_lattice(), rl, defect_out->data(), c.stream)));
    sync();
    return ok && conv_->post();
}

bool CG::apply_return_defect(GPUVector& x, GPUVector& b) {
    // cg.h:103-242
    used_fast_ = false;
    auto* gmg = fast_ ? dynamic_cast<AssembledMultiGridCycle*>(precond_.get()) : nullptr;
    if (gmg && !gmg->lattice_ready(op_->get_matrix())) gmg = nullptr;
    if (gmg) {
        // the initial defect formed in lattice order where x and b allow it
        const int kind = lattice_start_kind(x, b, gmg->num_inner_dofs());
        if (kind) {
            used_fast_ = true;
            return apply_lattice(x, b, *gmg, kind, &b);
        }
    }
    GPUVector& r = b;
    op_->apply_sub(r, x);
    if (gmg) {
        const int64_t ni = gmg->num_inner_dofs();
        const double tail = r.size() > ni ? dev_count_nonzero(r.size() - ni, r.data() + ni) : 0.0;
        if (tail == 0.0) {
            used_fast_ = true;
            return apply_lattice(x, r, *gmg, 0, &r);
        }
    }
    GPUVector q(r.size()), z(x.size()), p(x.size());
    if (precond_) {
        if (!precond_->apply(z, r)) return false;
    } else {
        z = r;
    }
    conv_->set_name(name());
    conv_->start(r);
    p = z;
    double rho_old = z.dotprod(r), rho;
    while (!conv_->iteration_ended()) {
        op_->apply(q, p);
        double lambda = q.dotprod(p);
        if (lambda == 0.0) {
            if (p.size()) return false;
            lambda = 1.0;
        }
        const double alpha = rho_old / lambda;
        VecScaleAdd(x, 1.0, x, alpha, p);
        VecScaleAdd(r, 1.0, r, -alpha, q);
        conv_->update(r);
        if (conv_->iteration_ended()) break;
        if (precond_) {
            if (!precond_->apply(z, r)) return false;
        } else {
            z = r;
        }
        rho = z.dotprod(r);
        const double beta = rho / rho_old;
        VecScaleAdd(p, beta, p, 1.0, z);
        rho_old = rho;
    }
    return conv_->post();
}

bool BiCGStab::apply_return_defect(GPUVector& x, GPUVector& b) {
    // bicgstab.h:112-385 (restart every num_restarts_ steps and when the
    // orthogonality of r to r0 drops to min_ortho_, bicgstab.h:157-215)
    op_->apply_sub(b, x);
    GPUVector& r = b;
    const int64_t n = r.size();
    GPUVector r0(n), p(n), v(n), t(n), s(n), q(n);
    conv_->set_name(name());
    conv_->start(r);
    double rho = 1, alpha = 1, omega = 1, norm_r0 = 0.0;
    bool restart = true;
    while (!conv_->iteration_ended()) {
        if (num_restarts_ > 0 && conv_->step() % num_restarts_ == 0) restart = true;
        if (restart) {
            r0 = r;
            p.set(0.0);
            alpha = 0.0;
            v.set(0.0);
            omega = 1.0;
            rho = 1.0;
            norm_r0 = conv_->defect();
            restart = false;
        }
        const double rho_old = rho;
        rho = r0.dotprod(r);
        const double norm_r = conv_->defect();
        if (std::fabs(rho) / (norm_r * norm_r0) <= min_ortho_) restart = true;
        if (rho_old == 0.0) return false;
        const double beta = (rho / rho_old) * (alpha / omega);
        VecScaleAdd(p, 1.0, r, beta, p, -beta * omega, v);
        if (precond_) {
            if (!precond_->apply(q, p)) return false;
        } else {
            q = p;
        }
        op_->apply(v, q);
        alpha = v.dotprod(r0);
        if (alpha == 0.0) return false;
        alpha = rho / alpha;
        VecScaleAdd(x, 1.0, x, alpha, q);
        VecScaleAdd(s, 1.0, r, -alpha, v);
        conv_->update(s);
        if (conv_->iteration_ended()) {
            r = s;
            break;
        }
        if (precond_) {
            if (!precond_->apply(q, s)) return false;
        } else {
            q = s;
        }
        op_->apply(t, q);
        const double tt = t.dotprod(t);
        omega = s.dotprod(t);
        if (tt == 0.0) return false;
        omega = omega / tt;
        VecScaleAdd(x, 1.0, x, omega, q);
        VecScaleAdd(r, 1.0, s, -omega, t);
        conv_->update(r);
        if (omega == 0.0) return false;
    }
    return conv_->post();
}

// ---------------------------------------------------------------------------
// DomainDiscretization (structured)
// ---------------------------------------------------------------------------
DomainDiscretization::DomainDiscretization(int kind, int num_refs, int ordering, int setup)
    : kind_(kind), refs_(num_refs) {
    if (kind < 0 || kind > 2) throw UGError("DomainDiscretization: unknown domain kind");
    if (num_refs < 0 || num_refs > (kind == 0 ? 9 : kind == 1 ? 13 : 6))
        throw UGError("DomainDiscretization: refinement out of range");
    if (ordering != 0 && ordering != 1) throw UGError("DomainDiscretization: unknown ordering");
    if (setup == 0)
        build_device_levels(kind, num_refs, ordering, levels_);
    else if (setup == 1)
        build_host_levels(kind, num_refs, ordering);
    else
        throw UGError("DomainDiscretization: unknown set-up (0 device, 1 host)");
    auto& c = context();
    for (auto& L : levels_) {
        L->lat_of_dof.resize(L->N);
        UGX_LAUNCH(invert_map(L->N, L->dof_of_lat.get(), L->lat_of_dof.get(), c.stream));
        L->dir.resize(L->N);
        UGX_LAUNCH(boundary_flags(L->dim, L->W, L->N, L->dir.get(), c.stream));
    }
    sync();
}

// the hierarchy and numbering built on the host (hierarchy.cpp), then uploaded
void DomainDiscretization::build_host_levels(int kind, int num_refs, int ordering) {
    auto nums = build_level_numbering(kind, num_refs);
    if (ordering == 1) {
        // lexicographic lattice order, inner DoFs first then Dirichlet DoFs
        for (auto& ln : nums) {
            const int64_t W = ln.n + 1, N = (int64_t)ln.dof_of_lattice.size();
            int32_t k = 0;
            for (int pass = 0; pass < 2; ++pass)
                for (int64_t l = 0; l < N; ++l) {
                    const int64_t x = l % W, y = (l / W) % W, z = ln.dim == 3 ? l / (W * W) : 0;
                    const bool bnd = x == 0 || y == 0 || x == W - 1 || y == W - 1 ||
                                     (ln.dim == 3 && (z == 0 || z == W - 1));
                    if (bnd == (pass == 1)) ln.dof_of_lattice[l] = k++;
                }
        }
    }
    for (auto& ln : nums) {
        auto L = std::make_unique<LevelGrid>();
        L->dim = ln.dim;
        L->n = ln.n;
        L->W = ln.n + 1;
        L->N = (int64_t)ln.dof_of_lattice.size();
        L->n_inner = ln.n_inner;
        if (ln.n_dof != L->N) throw UGError("DomainDiscretization: numbering does not cover the lattice");
        {
            // the numbering must be a permutation of the lattice points
            std::vector<uint8_t> seen(L->N, 0);
            for (int32_t d : ln.dof_of_lattice) {
                if (d < 0 || d >= L->N || seen[d]) throw UGError("DomainDiscretization: numbering is not a permutation");
                seen[d] = 1;
            }
        }
        L->dof_of_lat.resize(L->N);
        L->dof_of_lat.upload(ln.dof_of_lattice.data(), L->N);
        if (kind == 2) {
            const int64_t cells = (int64_t)L->n * L->n * L->n;
            if ((int64_t)ln.cell_tets.size() != cells * 6) throw UGError("DomainDiscretization: tetrahedra missing");
            L->cell_tets.resize(cells * 6);
            L->cell_tets.upload(ln.cell_tets.data(), cells * 6);
            if (!ln.parent_edge.empty()) {
                // every parent edge must join two coarse lattice points: the
                // odd coordinates of the child are exactly the non-zero
                // components of the direction
                const int64_t W = L->W;
                for (int64_t l = 0; l < L->N; ++l) {
                    const int code = ln.parent_edge[l];
                    const int64_t x = l % W, y = (l / W) % W, z = l / (W * W);
                    if (code == 0) {
                        if ((x | y | z) & 1) throw UGError("DomainDiscretization: inconsistent parent code");
                        continue;
                    }
                    if (code > 13) throw UGError("DomainDiscretization: inconsistent parent code");
                    const int d = code + 13;
                    const int dx = d % 3 - 1, dy = (d / 3) % 3 - 1, dz = d / 9 - 1;
                    if ((x & 1) != (dx != 0) || (y & 1) != (dy != 0) || (z & 1) != (dz != 0))
                        throw UGError("DomainDiscretization: inconsistent parent code");
                }
                L->pcode.resize(L->N);
                L->pcode.upload(ln.parent_edge.data(), L->N);
            }
        }
        levels_.push_back(std::move(L));
    }
}

void DomainDiscretization::set_velocity(double vx, double vy, double vz) {
    vel_[0] = vx;
    vel_[1] = vy;
    vel_[2] = vz;
}

void DomainDiscretization::set_jitter(uint64_t seed, double amp) {
    if (amp < 0.0 || amp >= 0.5) throw UGError("set_jitter: amplitude must lie in [0, 0.5)");
    jseed_ = seed;
    jamp_ = amp;
}

namespace {
// kernel parameters of level `lev` of a tetrahedral discretisation
ugamd::TetDisc tet_disc(const DomainDiscretization& D, const LevelGrid& L, int lev, const double* theta) {
    ugamd::TetDisc P;
    P.n = L.n;
    P.shift = D.top_level() - lev;
    P.n_fine = D.fine_cells();
    P.scale = 1.0 / (double)P.n_fine;
    P.theta = theta;
    P.jseed = D.jitter_seed();
    P.jamp = D.jitter_amp();
    P.jitter = (lev == D.top_level() && D.jitter_amp() != 0.0) ? 1 : 0;
    for (int d = 0; d < 3; ++d) P.vel[d] = D.velocity()[d];
    P.f = D.source_constant();
    return P;
}
}  // namespace

void DomainDiscretization::dof_positions(int lev, double* host_out) {
    if (lev < 0 || lev > top_level()) throw UGError("dof_positions: level out of range");
    LevelGrid& L = *levels_[lev];
    const int dim = kind_ == 1 ? 2 : 3;
    DeviceBuffer<double> out(L.N * dim);
    UGX_LAUNCH(ugamd::dof_positions(tet_disc(*this, L, lev, nullptr), kind_, dim, L.W, L.N, L.lat_of_dof.get(),
                                    out.get(), context().stream));
    out.download(host_out, L.N * dim);
}

void DomainDiscretization::set_theta(const double* theta_host) {
    const int64_t nf = fine_cells();
    const int64_t n = kind_ == 1 ? nf * nf : nf * nf * nf;
    theta_.resize(n);
    theta_.upload(theta_host, n);
    has_theta_ = true;
    kel_valid_ = false;
}

void DomainDiscretization::set_theta_device(const double* theta_dev) {
    const int64_t nf = fine_cells();
    const int64_t n = kind_ == 1 ? nf * nf : nf * nf * nf;
    theta_.resize(n);
    UGX_HIP(hipMemcpyAsync(theta_.get(), theta_dev, sizeof(double) * n, hipMemcpyDeviceToDevice, context().stream));
    has_theta_ = true;
    kel_valid_ = false;
}

const double* DomainDiscretization::exp_theta_device() {
    if (!has_theta_) return nullptr;
    if (!kel_valid_) {
        kel_.resize(theta_.size());
        UGX_LAUNCH(exp_values(theta_.size(), theta_.get(), kel_.get(), context().stream));
        kel_valid_ = true;
    }
    return kel_.get();
}

const double* DomainDiscretization::q1_tables_device() {
    if (q1_tab_.size() != kQ1Tab) {
        double tab[kQ1Tab];
        q1_lattice_tables(tab);
        q1_tab_.resize(kQ1Tab);
        q1_tab_.upload(tab, kQ1Tab);
    }
    return q1_tab_.get();
}

void DomainDiscretization::adjust_defect(GPUVector& d) {
    LevelGrid& L = *levels_.back();
    if (d.size() != L.N) throw UGError("adjust_defect: size mismatch");
    UGX_LAUNCH(zero_range(L.n_inner, L.N, d.data(), context().stream));
}

void DomainDiscretization::adjoint_gradient(const GPUVector& u, const GPUVector& lambda, double* grad_dev) {
    if (kind_ == 1) throw UGError("adjoint_gradient: not implemented for the triangle discretisation");
    LevelGrid& L = *levels_.back();
    if (u.size() != L.N || lambda.size() != L.N) throw UGError("adjoint_gradient: size mismatch");
    if (!has_theta_) throw UGError("adjoint_gradient: theta not set");
    if (kind_ == 2) {
        UGX_LAUNCH(adjoint_grad_fv1_tet(tet_disc(*this, L, top_level(), theta_.get()), L.cell_tets.get(),
                                        L.dof_of_lat.get(), L.n_inner, u.data(), lambda.data(), grad_dev,
                                        context().stream));
        return;
    }
    UGX_LAUNCH(adjoint_grad_q1_hex(L.n, 1.0 / L.n, theta_.get(), L.dof_of_lat.get(), L.n_inner, u.data(),
                                   lambda.data(), grad_dev, context().stream));
}

// ---------------------------------------------------------------------------
// ElementDiscretization: element loop over caller-supplied connectivity
// ---------------------------------------------------------------------------
ElementDiscretization::ElementDiscretization(int elem_type, int64_t num_elems, const int32_t* elem_dofs,
                                             int64_t num_dofs, const double* dof_coords, bool on_device)
    : type_(elem_type), ne_(num_elems), ndofs_(num_dofs) {
    if (elem_type < 0 || elem_type > 3)
        throw UGError("ElementDiscretization: element type 0 (triangle, FV1), 1 (tetrahedron, FV1), 2 "
                      "(hexahedron, Q1 FE) or 3 (hexahedron, Q2 FE)");
    nv_ = vertices_of(elem_type);
    dim_ = dim_of(elem_type);
    if (num_elems < 0 || num_dofs < 0 || num_dofs >= ((int64_t)1 << 31))
        throw UGError("ElementDiscretization: element and DoF counts must be >= 0 and below 2^31");
    if (num_elems * nv_ * nv_ >= ((int64_t)1 << 31))
        throw UGError("ElementDiscretization: num_elems * vertices^2 must stay below 2^31");
    if (num_elems > 0 && (!elem_dofs || !dof_coords)) throw UGError("ElementDiscretization: null connectivity");
    if (!on_device)
        for (int64_t i = 0; i < ne_ * nv_; ++i)
            if (elem_dofs[i] < 0 || elem_dofs[i] >= ndofs_)
                throw UGError("ElementDiscretization: DoF index out of range in the element table");
    auto& c = context();
    dofs_.resize(ne_ * nv_);
    coords_.resize(ndofs_ * dim_);
    if (on_device) {
        if (ne_ > 0)
            UGX_HIP(hipMemcpyAsync(dofs_.get(), elem_dofs, sizeof(int32_t) * ne_ * nv_, hipMemcpyDeviceToDevice,
                                   c.stream));
        if (ndofs_ > 0)
            UGX_HIP(hipMemcpyAsync(coords_.get(), dof_coords, sizeof(double) * ndofs_ * dim_,
                                   hipMemcpyDeviceToDevice, c.stream));
    } else {
        dofs_.upload(elem_dofs, ne_ * nv_);
        coords_.upload(dof_coords, ndofs_ * dim_);
    }
    build();
}

void ElementDiscretization::build() {
    auto& c = context();
    const int64_t n = ne_ * nv_ * nv_;
    // the pattern and the slot of every local entry
    size_t sort_b = 0, scan_b = 0;
    if (elem_pattern_temp_bytes(ne_, nv_, ndofs_, &sort_b, &scan_b) != 0)
        throw UGError("ElementDiscretization: sort set-up failed");
    DeviceBuffer<uint64_t> key(n), key2(n);
    DeviceBuffer<int32_t> idx(n), idx2(n), head(n), run(n), nnz_dev(1), slot_local(n);
    DeviceBuffer<uint8_t> temp((int64_t)std::max(sort_b, scan_b));
    rowptr_.resize(ndofs_ + 1);
    UGX_LAUNCH(elem_pattern(ne_, nv_, ndofs_, dofs_.get(), key.get(), key2.get(), idx.get(), idx2.get(), head.get(),
                            run.get(), temp.get(), sort_b, scan_b, rowptr_.get(), nnz_dev.get(), c.stream));
    int32_t nnz = 0;
    nnz_dev.download(&nnz, 1);
    nnz_ = nnz;
    col_.resize(nnz_);
    UGX_LAUNCH(elem_pattern_fill(ne_, nv_, key2.get(), idx2.get(), head.get(), run.get(), col_.get(),
                                 slot_local.get(), c.stream));
    key.release();
    key2.release();
    head.release();
    run.release();
    idx.release();
    idx2.release();
    // DoF -> element incidence, then the colouring rounds
    const int64_t ni = ne_ * nv_;
    size_t inc_b = 0;
    if (elem_incidence_temp_bytes(ni, ndofs_, &inc_b) != 0 || elem_incidence_temp_bytes(ne_, 64, &sort_b) != 0)
        throw UGError("ElementDiscretization: sort set-up failed");
    DeviceBuffer<uint32_t> ikey(std::max(ni, ne_)), ikey2(std::max(ni, ne_));
    DeviceBuffer<int32_t> iidx(std::max(ni, ne_)), inc_t(ni), inc_ptr(ndofs_ + 1), colour(ne_), flags(2);
    DeviceBuffer<uint8_t> itemp((int64_t)std::max(inc_b, sort_b));
    UGX_LAUNCH(elem_incidence(ne_, nv_, ndofs_, dofs_.get(), ikey.get(), ikey2.get(), iidx.get(), inc_t.get(),
                              itemp.get(), inc_b, inc_ptr.get(), c.stream));
    UGX_HIP(hipMemsetAsync(colour.get(), 0xff, sizeof(int32_t) * (size_t)std::max<int64_t>(ne_, 1), c.stream));
    UGX_HIP(hipMemsetAsync(flags.get(), 0, sizeof(int32_t) * 2, c.stream));
    int32_t done[2] = {0, 0};
    int rounds = 0;
    while (done[0] < ne_) {
        UGX_LAUNCH(elem_colour_round(ne_, nv_, dofs_.get(), inc_ptr.get(), inc_t.get(), colour.get(), flags.get(),
                                     flags.get() + 1, c.stream));
        flags.download(done, 2);
        if (done[1]) throw UGError("ElementDiscretization: an element has more than 63 neighbouring colours");
        if (++rounds > 100000) throw UGError("ElementDiscretization: the colouring does not terminate");
    }
    // the number of colours: 1 + the largest colour
    int ncol = 0;
    if (ne_ > 0) {
        std::vector<int32_t> col_h(ne_);
        colour.download(col_h.data(), ne_);
        for (int32_t x : col_h) ncol = std::max(ncol, x + 1);
    }
    order_.resize(ne_);
    DeviceBuffer<int32_t> colour_ptr(ncol + 1);
    UGX_LAUNCH(elem_colour_order(ne_, ncol, colour.get(), ikey.get(), ikey2.get(), iidx.get(), order_.get(),
                                 itemp.get(), sort_b, colour_ptr.get(), c.stream));
    std::vector<int32_t> cp(ncol + 1);
    colour_ptr.download(cp.data(), ncol + 1);
    colour_start_.assign(cp.begin(), cp.end());
    slot_.resize(n);
    UGX_LAUNCH(elem_slots_by_position(ne_, nv_, order_.get(), slot_local.get(), slot_.get(), c.stream));
    sync();
}

void ElementDiscretization::set_diffusion(const double* k, bool on_device) {
    if (!k) {
        has_k_ = false;
        return;
    }
    kel_.resize(ne_);
    if (on_device)
        UGX_HIP(hipMemcpyAsync(kel_.get(), k, sizeof(double) * ne_, hipMemcpyDeviceToDevice, context().stream));
    else
        kel_.upload(k, ne_);
    has_k_ = true;
}

void ElementDiscretization::set_diffusion_at_points(const double* kq, bool on_device) {
    if (type_ != HEXAHEDRON_Q2) throw UGError("ElementDiscretization: coefficients at Gauss points are for Q2");
    if (!kq) {
        has_kq_ = false;
        return;
    }
    kq_.resize(ne_ * 27);
    if (on_device)
        UGX_HIP(hipMemcpyAsync(kq_.get(), kq, sizeof(double) * ne_ * 27, hipMemcpyDeviceToDevice, context().stream));
    else
        kq_.upload(kq, ne_ * 27);
    has_kq_ = true;
}

void ElementDiscretization::set_source_dofs(const double* f_host) {
    if (type_ == HEXAHEDRON_Q1 || type_ == HEXAHEDRON_Q2)
        throw UGError("ElementDiscretization: the Q1/Q2 source is a constant");
    f_dof_.resize(ndofs_);
    f_dof_.upload(f_host, ndofs_);
    has_f_dof_ = true;
}

void ElementDiscretization::set_velocity(double vx, double vy, double vz) {
    if (type_ != TETRAHEDRON_FV1 && (vx != 0.0 || vy != 0.0 || vz != 0.0))
        throw UGError("ElementDiscretization: convection is available for tetrahedra (FV1, full upwind)");
    vel_[0] = vx;
    vel_[1] = vy;
    vel_[2] = vz;
    conv_ = vx != 0.0 || vy != 0.0 || vz != 0.0;
}

void ElementDiscretization::set_dirichlet(const int32_t* dofs_host, int64_t n) {
    for (int64_t i = 0; i < n; ++i)
        if (dofs_host[i] < 0 || dofs_host[i] >= ndofs_) throw UGError("ElementDiscretization: Dirichlet DoF out of range");
    dir_.resize(n);
    if (n > 0) dir_.upload(dofs_host, n);
    ndir_ = n;
}

void ElementDiscretization::set_dirichlet_device(const int32_t* dofs_dev, int64_t n) {
    dir_.resize(n);
    if (n > 0)
        UGX_HIP(hipMemcpyAsync(dir_.get(), dofs_dev, sizeof(int32_t) * n, hipMemcpyDeviceToDevice, context().stream));
    ndir_ = n;
}

void ElementDiscretization::run(double* val, double* rhs, const double* u, double* d) {
    auto& c = context();
    ugamd::ElemLaunch L;
    L.ne = ne_;
    L.order = order_.get();
    L.dofs = dofs_.get();
    L.coords = coords_.get();
    L.kel = has_k_ ? kel_.get() : nullptr;
    L.f_const = f_const_;
    L.f_dof = has_f_dof_ ? f_dof_.get() : nullptr;
    for (int k = 0; k < 3; ++k) L.vel[k] = vel_[k];
    L.conv = conv_ ? 1 : 0;
    L.slot = slot_.get();
    L.val = val;
    L.rhs = rhs;
    L.u = u;
    L.d = d;
    if (type_ == HEXAHEDRON_Q2) {
        for (int col = 0; col + 1 < (int)colour_start_.size(); ++col)
            UGX_LAUNCH(elem_assemble_q2_colour(ne_, order_.get(), dofs_.get(), coords_.get(), L.kel,
                                               has_kq_ ? kq_.get() : nullptr, f_const_, slot_.get(), val, rhs,
                                               colour_start_[col], colour_start_[col + 1] - colour_start_[col],
                                               c.stream));
        return;
    }
    for (int col = 0; col + 1 < (int)colour_start_.size(); ++col)
        UGX_LAUNCH(elem_assemble_colour(type_, L, colour_start_[col], colour_start_[col + 1] - colour_start_[col],
                                        c.stream));
}

bool ElementDiscretization::give_pattern(GPUSparseMatrix& A) {
    // a matrix this object assembled whose pattern is unchanged keeps it
    if (A.origin() == this && A.pattern_id() == asm_pattern_id_ && !A.building() && A.num_rows() == ndofs_ &&
        A.total_num_connections() == nnz_)
        return true;
    auto& c = context();
    A.allocate(ndofs_, ndofs_, nnz_);
    A.set_origin(nullptr);
    UGX_HIP(hipMemcpyAsync(A.rowptr(), rowptr_.get(), sizeof(int32_t) * (ndofs_ + 1), hipMemcpyDeviceToDevice,
                           c.stream));
    if (nnz_ > 0)
        UGX_HIP(hipMemcpyAsync(A.col(), col_.get(), sizeof(int32_t) * nnz_, hipMemcpyDeviceToDevice, c.stream));
    return false;
}

void ElementDiscretization::finish_matrix(GPUSparseMatrix& A, bool kept) {
    A.set_origin(this);
    if (kept) A.refresh_layout_values();
    else A.build_spmv_layout();
    asm_pattern_id_ = A.pattern_id();
}

void ElementDiscretization::assemble_linear(GPUSparseMatrix& A, GPUVector& b) {
    auto& c = context();
    const bool kept = give_pattern(A);
    if (nnz_ > 0) UGX_LAUNCH(vec_set<double>(nnz_, A.val(), 0.0, c.stream));   // fresh memory may hold NaN
    b.resize(ndofs_);
    b.set(0.0);
    run(A.val(), b.data(), nullptr, nullptr);
    if (ndir_ > 0) {
        UGX_LAUNCH(dirichlet_rows(ndir_, dir_.get(), A.rowptr(), A.col(), A.val(), c.stream));
        UGX_LAUNCH(zero_listed(ndir_, dir_.get(), b.data(), c.stream));
    }
    sync();
    finish_matrix(A, kept);
}

void ElementDiscretization::assemble_jacobian(GPUSparseMatrix& J, const GPUVector& u) {
    if (u.size() != ndofs_) throw UGError("ElementDiscretization::assemble_jacobian: size mismatch");
    auto& c = context();
    const bool kept = give_pattern(J);
    if (nnz_ > 0) UGX_LAUNCH(vec_set<double>(nnz_, J.val(), 0.0, c.stream));   // fresh memory may hold NaN
    run(J.val(), nullptr, nullptr, nullptr);
    if (ndir_ > 0) UGX_LAUNCH(dirichlet_rows(ndir_, dir_.get(), J.rowptr(), J.col(), J.val(), c.stream));
    sync();
    finish_matrix(J, kept);
}

void ElementDiscretization::assemble_defect(GPUVector& d, const GPUVector& u) {
    if (u.size() != ndofs_) throw UGError("ElementDiscretization::assemble_defect: size mismatch");
    if (&d == &u) throw UGError("ElementDiscretization::assemble_defect: d must not alias u");
    d.resize(ndofs_);
    if (type_ == HEXAHEDRON_Q2) {
        // the Q2 element routine forms matrices only: d = A u - b with the
        // assembled operator (its non-Dirichlet rows are the element sums)
        GPUSparseMatrix T;
        GPUVector b(ndofs_);
        assemble_linear(T, b);
        T.axpy(d, -1.0, b, 1.0, u);
    } else {
        d.set(0.0);
        run(nullptr, nullptr, u.data(), d.data());
    }
    if (ndir_ > 0) UGX_LAUNCH(zero_listed(ndir_, dir_.get(), d.data(), context().stream));
}

void ElementDiscretization::download_pattern(int32_t* rowptr, int32_t* col) const {
    if (rowptr) rowptr_.download(rowptr, ndofs_ + 1);
    if (col && nnz_ > 0) col_.download(col, nnz_);
}

ElementDiscretization& DomainDiscretization::element_discretization() {
    auto& c = context();
    LevelGrid& L = *levels_.back();
    const int type = kind_ == 0 ? ElementDiscretization::HEXAHEDRON_Q1
                                : (kind_ == 1 ? ElementDiscretization::TRIANGLE_FV1 : ElementDiscretization::TETRAHEDRON_FV1);
    const int nv = ElementDiscretization::vertices_of(type);
    const int64_t n = L.n;
    const int64_t ne = kind_ == 0 ? n * n * n : (kind_ == 1 ? 2 * n * n : 6 * n * n * n);
    DeviceBuffer<int32_t> dofs(ne * nv);
    DeviceBuffer<double> kel(ne);
    UGX_LAUNCH(lattice_connectivity(kind_, L.n, L.dof_of_lat.get(), kind_ == 2 ? L.cell_tets.get() : nullptr,
                                    has_theta_ ? theta_.get() : nullptr, dofs.get(), kel.get(), c.stream));
    if (!elem_ || elem_jseed_ != jseed_ || elem_jamp_ != jamp_) {
        const int dim = dim_of_kind();
        DeviceBuffer<double> X(L.N * dim);
        UGX_LAUNCH(ugamd::dof_positions(tet_disc(*this, L, top_level(), nullptr), kind_, dim, L.W, L.N,
                                        L.lat_of_dof.get(), X.get(), c.stream));
        elem_.reset(new ElementDiscretization(type, ne, dofs.get(), L.N, X.get(), true));
        std::vector<int32_t> dir((size_t)(L.N - L.n_inner));
        for (int64_t i = L.n_inner; i < L.N; ++i) dir[(size_t)(i - L.n_inner)] = (int32_t)i;
        elem_->set_dirichlet(dir.data(), (int64_t)dir.size());
        elem_jseed_ = jseed_;
        elem_jamp_ = jamp_;
    }
    elem_->set_diffusion(has_theta_ ? kel.get() : nullptr, true);
    if (kind_ == 1) {
        if (has_f_) {
            std::vector<double> f((size_t)L.N);
            f_dof_.download(f.data(), L.N);
            elem_->set_source_dofs(f.data());
        } else {
            elem_->set_source(0.0);   // the triangle discretisation takes its source per DoF only
        }
    } else {
        elem_->set_source(f_const_);
    }
    if (kind_ == 2) elem_->set_velocity(vel_[0], vel_[1], vel_[2]);
    sync();
    return *elem_;
}

void DomainDiscretization::assemble_jacobian(GPUSparseMatrix& J, const GPUVector& u) {
    if (u.size() != levels_.back()->N) throw UGError("assemble_jacobian: size mismatch");
    GPUVector b(levels_.back()->N);
    assemble_linear(J, b);
}

void DomainDiscretization::assemble_defect(GPUVector& d, const GPUVector& u) {
    if (u.size() != levels_.back()->N) throw UGError("assemble_defect: size mismatch");
    element_discretization().assemble_defect(d, u);
}

void DomainDiscretization::set_source_dofs(const double* f_host) {
    const int64_t n = levels_.back()->N;
    f_dof_.resize(n);
    f_dof_.upload(f_host, n);
    has_f_ = true;
}

void DomainDiscretization::assemble_linear(GPUSparseMatrix& A, GPUVector& b) {
    auto& c = context();
    const double t0 = wall();
    LevelGrid& L = *levels_.back();
    // the pattern is built once: a matrix this discretisation assembled whose
    // pattern has not changed since keeps it, and its SpMV layout takes the
    // new values in place
    const bool reuse = A.origin() == this && A.pattern_id() == asm_pattern_id_ && !A.building() &&
                       A.num_rows() == L.N && A.num_cols() == L.N;
    if (!reuse) {
        DeviceBuffer<int32_t> cnt(L.N);
        DeviceBuffer<uint32_t> mask;
        if (kind_ == 2) {
            mask.resize(L.N);
            mask.zero();
            UGX_LAUNCH(tet_mask(L.n, L.cell_tets.get(), mask.get(), c.stream));
            UGX_LAUNCH(pattern_count_mask(L.N, L.lat_of_dof.get(), mask.get(), cnt.get(), c.stream));
        } else {
            UGX_LAUNCH(pattern_count(L.dim, L.W, L.N, L.lat_of_dof.get(), cnt.get(), c.stream));
        }
        // row pointers scanned on the device (64-bit sums, total checked)
        DeviceBuffer<int64_t> incl(L.N);
        size_t tb = 0;
        if (rowptr_scan_temp_bytes(L.N, &tb) != 0) throw UGError("assemble_linear: scan set-up failed");
        DeviceBuffer<uint8_t> temp((int64_t)std::max<size_t>(tb, 1));
        DeviceBuffer<int32_t> rowptr(L.N + 1);
        UGX_LAUNCH(rowptr_from_counts(L.N, cnt.get(), incl.get(), temp.get(), tb, rowptr.get(), c.stream));
        int64_t nnz = 0;
        if (L.N > 0) incl.download(&nnz, 1, L.N - 1);
        if (nnz >= ((int64_t)1 << 31)) throw UGError("assemble_linear: nnz exceeds int32 range");
        A.allocate(L.N, L.N, nnz);
        UGX_HIP(hipMemcpyAsync(A.rowptr(), rowptr.get(), sizeof(int32_t) * (L.N + 1), hipMemcpyDeviceToDevice,
                               c.stream));
        if (kind_ == 2) {
            UGX_LAUNCH(pattern_fill_mask(L.W, L.N, L.lat_of_dof.get(), L.dof_of_lat.get(), mask.get(), A.rowptr(),
                                         A.col(), c.stream));
        } else {
            // row start per lattice point (the same for every matrix of this
            // pattern), then one lattice point per 32 lanes, columns by rank
            if (rp_lat_.size() != L.N) {
                rp_lat_.resize(L.N);
                UGX_LAUNCH(rowptr_of_lattice(L.N, L.dof_of_lat.get(), A.rowptr(), rp_lat_.get(), c.stream));
            }
            UGX_LAUNCH(lattice_pattern(L.dim, L.W, L.dof_of_lat.get(), rp_lat_.get(), A.col(), c.stream));
        }
    }
    const int64_t acc = A.total_num_connections();
    b.resize(L.N);
    const double hsz = 1.0 / L.n;
    const int nfine = 1 << refs_;
    if (kind_ == 0) {
        // row by row: every entry and every rhs entry written once
        if (nfine != L.n) throw UGError("assemble_linear: the top level must be the fine level");
        if (rp_lat_.size() != L.N) {
            rp_lat_.resize(L.N);
            UGX_LAUNCH(rowptr_of_lattice(L.N, L.dof_of_lat.get(), A.rowptr(), rp_lat_.get(), c.stream));
        }
        UGX_LAUNCH(q1_lattice_csr(L.n, hsz, f_const_, q1_tables_device(), exp_theta_device(), L.dof_of_lat.get(),
                                  rp_lat_.get(), A.val(), b.data(), c.stream));
    } else {
        UGX_HIP(hipMemsetAsync(A.val(), 0, sizeof(double) * acc, c.stream));
        b.set(0.0);
    }
    if (kind_ == 2) {
        UGX_LAUNCH(assemble_fv1_tet<double>(tet_disc(*this, L, top_level(), theta_device()), L.cell_tets.get(),
                                            L.dof_of_lat.get(), A.rowptr(), A.col(), A.val(), nullptr, b.data(),
                                            c.stream));
    } else if (kind_ == 1) {
        UGX_LAUNCH(assemble_fv1_tri<double>(L.n, hsz, theta_device(), nfine, has_f_ ? f_dof_.get() : nullptr,
                                            L.dof_of_lat.get(), A.rowptr(), A.col(), A.val(), nullptr, 0, b.data(),
                                            c.stream));
    }
    UGX_LAUNCH(dirichlet_csr<double>(L.N, L.n_inner, A.rowptr(), A.col(), A.val(), b.data(), c.stream));
    sync();
    A.set_origin(this);
    if (reuse) {
        A.refresh_layout_values();
    } else {
        int64_t lat_points = 1;
        for (int d = 0; d < L.dim; ++d) lat_points *= L.W;
        if (lat_points == L.N) {
            A.set_row_order_device(L.dof_of_lat.get(), L.N);
            const int64_t ns = lattice_block_slots(L.dim, L.W);
            DeviceBuffer<int32_t> so(ns);
            UGX_LAUNCH(lattice_block_order(L.dim, L.W, L.dof_of_lat.get(), so.get(), c.stream));
            A.set_slice_order_device(so.get(), ns, true);   // builds the SpMV layout
        } else {
            A.build_spmv_layout();
        }
    }
    asm_pattern_id_ = A.pattern_id();
    t_assemble_ = wall() - t0;
}

void DomainDiscretization::assemble_adjoint_operator(const GPUSparseMatrix& A, GPUSparseMatrix& AT) {
    auto& c = context();
    LevelGrid& L = *levels_.back();
    if (A.num_rows() != L.N || A.num_cols() != L.N) throw UGError("assemble_adjoint_operator: size mismatch");
    const int64_t nnz = A.total_num_connections();
    AT.allocate(L.N, L.N, nnz);
    UGX_HIP(hipMemcpyAsync(AT.rowptr(), A.rowptr(), sizeof(int32_t) * (L.N + 1), hipMemcpyDeviceToDevice, c.stream));
    UGX_HIP(hipMemcpyAsync(AT.col(), A.col(), sizeof(int32_t) * nnz, hipMemcpyDeviceToDevice, c.stream));
    DeviceBuffer<int32_t> tp(nnz);
    UGX_HIP(hipMemsetAsync(c.flag, 0, sizeof(int), c.stream));
    UGX_LAUNCH(csr_transpose_map(L.N, A.rowptr(), A.col(), tp.get(), c.flag, c.stream));
    UGX_LAUNCH(csr_adjoint_values(L.N, L.n_inner, A.rowptr(), A.col(), tp.get(), A.val(), AT.val(), c.stream));
    sync();
    int bad = 0;
    UGX_HIP(hipMemcpy(&bad, c.flag, sizeof(int), hipMemcpyDeviceToHost));
    if (bad) throw UGError("assemble_adjoint_operator: pattern is not structurally symmetric");
    AT.set_origin(nullptr);
    if (A.has_row_order()) {
        AT.set_row_order_device(L.dof_of_lat.get(), L.N);
        const int64_t ns = lattice_block_slots(L.dim, L.W);
        DeviceBuffer<int32_t> so(ns);
        UGX_LAUNCH(lattice_block_order(L.dim, L.W, L.dof_of_lat.get(), so.get(), c.stream));
        AT.set_slice_order_device(so.get(), ns, true);
    } else {
        AT.build_spmv_layout();
    }
}

void DomainDiscretization::level_dia_layout(int lev, int& nd, std::vector<int32_t>& offsets) const {
    const LevelGrid& L = *levels_[lev];
    nd = L.dim == 3 ? 27 : 9;
    offsets.assign(nd, 0);
    const int64_t W = L.W;
    for (int d = 0; d < nd; ++d) {
        const int dx = d % 3 - 1, dy = (d / 3) % 3 - 1, dz = L.dim == 3 ? d / 9 - 1 : 0;
        offsets[d] = (int32_t)(dx + dy * W + dz * W * W);
    }
}

void DomainDiscretization::assemble_level_dia(int lev, DeviceBuffer<double>& dia, int& nd,
                                              std::vector<int32_t>& offsets) {
    auto& c = context();
    LevelGrid& L = *levels_[lev];
    level_dia_layout(lev, nd, offsets);
    dia.resize(dia_tiled_size(L.N, nd));
    dia.zero();
    const double hsz = 1.0 / L.n;
    const int nfine = 1 << refs_;
    if (kind_ == 2) {
        UGX_LAUNCH(assemble_fv1_tet<double>(tet_disc(*this, L, lev, theta_device()), L.cell_tets.get(),
                                            L.dof_of_lat.get(), nullptr, nullptr, nullptr, dia.get(), nullptr,
                                            c.stream));
    } else if (kind_ == 0) {
        UGX_LAUNCH(q1_lattice_dia(L.n, nfine, hsz, q1_tables_device(), exp_theta_device(), dia.get(), c.stream));
    } else {
        UGX_LAUNCH(assemble_fv1_tri<double>(L.n, hsz, theta_device(), nfine, nullptr, L.dof_of_lat.get(), nullptr,
                                            nullptr, nullptr, dia.get(), L.N, nullptr, c.stream));
    }
    UGX_LAUNCH(dirichlet_dia<double>(L.N, nd, nd / 2, L.dir.get(), dia.get(), c.stream));
}

// ---------------------------------------------------------------------------
// AssembledMultiGridCycle
// ---------------------------------------------------------------------------
struct AssembledMultiGridCycle::Impl {
    virtual ~Impl() = default;
    virtual void init() = 0;
    virtual void apply(GPUVector& c, const GPUVector& d) = 0;
    virtual void download_level_dia(int lev, double* out) const = 0;
    virtual int num_diags(int lev) const = 0;
    virtual void smooth(int lev, int nsteps) = 0;
    virtual void apply_lattice(double* c_lat, const double* d_lat, double* cd_out, double* partial,
                               const float* d_lat_f32) = 0;
    virtual bool uses_f32_defect() const = 0;
    virtual bool surface_axpy(double* y, double alpha, const double* v, double beta, const double* x, double* yx_out,
                              double* partial) = 0;
    virtual bool surface_direction_axpy(double* q, const double* p_old, const double* z, const double* beta,
                                        double* p_new, double* qp_out, double* partial) = 0;
    virtual int64_t halo() const = 0;
};

namespace {

// dense LU with partial pivoting (no_lapack/lu_decomp.h:47-200), FP64 on the host
bool host_lu(int n, std::vector<double>& a, std::vector<int32_t>& piv) {
    piv.assign(n, 0);
    auto M = [&](int r, int c) -> double& { return a[(size_t)r * n + c]; };
    for (int k = 0; k < n; ++k) {
        int big = k;
        for (int j = k + 1; j < n; ++j)
            if (std::fabs(M(big, k)) < std::fabs(M(j, k))) big = j;
        if (big != k)
            for (int j = 0; j < n; ++j) std::swap(M(k, j), M(big, j));
        piv[k] = big;
        if (std::fabs(M(k, k)) < 1e-10) return false;
        for (int i = k + 1; i < n; ++i) {
            M(i, k) = M(i, k) / M(k, k);
            for (int j = k + 1; j < n; ++j) M(i, j) = M(i, j) - M(i, k) * M(k, j);
        }
    }
    return true;
}

// The base solve of the multigrid cycles: LU with partial pivoting of the
// dense matrix (row-major, overwritten), the row permutation of SolveLU (swap
// x[i], x[piv[i]] in order) and the inverses of the unit lower and of the
// upper triangular factor, row-major (linvT[i*n+k] = Linv(i,k)), which
// lu_solve applies as two matrix-vector products. False: singular.
bool dense_lu_factors(int m, std::vector<double>& dense, std::vector<double>& linvT, std::vector<double>& uinvT,
                      std::vector<int32_t>& perm) {
    std::vector<int32_t> piv;
    if (!host_lu(m, dense, piv)) return false;
    perm.resize(m);
    std::iota(perm.begin(), perm.end(), 0);
    for (int i = 0; i < m; ++i)
        if (i < piv[i]) std::swap(perm[i], perm[piv[i]]);
    linvT.assign((size_t)m * m, 0.0);
    uinvT.assign((size_t)m * m, 0.0);
    auto LU = [&](int r, int c) { return dense[(size_t)r * m + c]; };
    for (int k = 0; k < m; ++k) {
        // column k of L^{-1}: solve L y = e_k (unit diagonal)
        std::vector<double> y(m, 0.0);
        y[k] = 1.0;
        for (int i = k + 1; i < m; ++i) {
            double acc = 0.0;
            for (int j = k; j < i; ++j) acc += LU(i, j) * y[j];
            y[i] = -acc;
        }
        for (int i = 0; i < m; ++i) linvT[(size_t)i * m + k] = y[i];
        // column k of U^{-1}: solve U z = e_k
        std::vector<double> z(m, 0.0);
        for (int i = k; i >= 0; --i) {
            double acc = (i == k) ? 1.0 : 0.0;
            for (int j = i + 1; j <= k; ++j) acc -= LU(i, j) * z[j];
            z[i] = acc / LU(i, i);
        }
        for (int i = 0; i < m; ++i) uinvT[(size_t)i * m + k] = z[i];
    }
    return true;
}

template <typename T>
struct GmgLevel {
    int dim = 3, W = 0, nd = 0;
    int64_t N = 0, H = 0;
    std::vector<int32_t> off;
    DeviceBuffer<T> A;       // symmetric Dirichlet-eliminated storage (nd/2+1 diagonals)
    DeviceBuffer<T> Afull;   // all nd diagonals (defects with Dirichlet entries)
    DeviceBuffer<T> b_dinv, b_sc, b_sd, b_sd2, b_st;
    T *dinv = nullptr, *sc = nullptr, *sd = nullptr, *sd2 = nullptr, *st = nullptr;
    const uint8_t* dir = nullptr;
    // the level's last post-smoothing update sc += dinv .* sd is left to the
    // prolongation that reads sc next (set in a cycle, cleared by that
    // prolongation)
    bool acc_deferred = false;
};

template <typename T>
struct GmgImpl : public AssembledMultiGridCycle::Impl {
    DomainDiscretization* disc;
    int base, top, nu1, nu2, cycle;
    double omega;
    std::vector<GmgLevel<T>> L;
    DeviceBuffer<double> lu_linvT, lu_uinvT;
    DeviceBuffer<int32_t> lu_perm;
    int lu_n = 0;
    DeviceBuffer<double> surf;   // top-level operator, FP64 symmetric storage
    bool sym_ok = true;          // the discretisation's level operators admit symmetric storage
    bool transposed = false;     // level operators A_l,elim^T (adjoint)
    bool tets = false;           // edge-based transfer of the tetrahedral hierarchy
    bool mf = false;             // matrix-free top-level operator (hexahedral kind)
    DeviceBuffer<T> cell_c;      // per-element coefficients h k_e of the top level
    DeviceBuffer<double> cell_c64;
    bool sym = true;             // level products through the symmetric storage
    bool rap = false;            // Galerkin coarse operators R A P below the top level
    bool graphs = true;          // replay the lattice cycle from captured HIP graphs
    bool fuse_smoothing = true;  // matrix-free top level of the lattice entry: fused smoothing passes
    struct CycleGraph {
        const void* key[5];
        hipGraphExec_t exec;
    };
    std::vector<CycleGraph> cycle_graphs;
    void clear_graphs() {
        for (auto& g : cycle_graphs) (void)hipGraphExecDestroy(g.exec);
        cycle_graphs.clear();
    }
    ~GmgImpl() override { clear_graphs(); }

    // level products: symmetric storage when every level vector vanishes on
    // Dirichlet DoFs (the defect entering the cycle does), full otherwise
    int op_apply_sub(GmgLevel<T>& l, const T* x, const T* d_in, T* d_out) {
        auto& c = context();
        if (sym && mf && &l == &L[top]) return q1mf_apply_sub<T>(l.W, cell_c.get(), x, d_in, d_out, c.stream);
        return sym ? dias_apply_sub<T, T>(l.N, l.nd, l.off.data(), l.A.get(), x, d_in, d_out, c.stream)
                   : dia_apply_sub<T, T>(l.N, l.nd, l.off.data(), l.Afull.get(), x, d_in, d_out, c.stream);
    }
    int op_jacobi_step(GmgLevel<T>& l) {
        auto& c = context();
        if (sym && mf && &l == &L[top])
            return q1mf_jacobi_step<T>(l.W, cell_c.get(), l.dinv, l.sd, l.sd2, l.sc, c.stream);
        return sym ? dias_jacobi_step<T, T>(l.N, l.nd, l.off.data(), l.A.get(), l.dinv, l.sd, l.sd2, l.sc, c.stream)
                   : dia_jacobi_step<T, T>(l.N, l.nd, l.off.data(), l.Afull.get(), l.dinv, l.sd, l.sd2, l.sc,
                                           c.stream);
    }

    GmgImpl(DomainDiscretization* d, int b, int n1, int n2, int cy, double w, bool tr, bool matrix_free, bool galerkin,
            bool use_graphs, bool fused_smoothing)
        : disc(d), base(b), top(d->top_level()), nu1(n1), nu2(n2), cycle(cy), omega(w),
          sym_ok(d->symmetric()), transposed(tr), tets(d->kind() == 2),
          mf(matrix_free && d->kind() == 0 && d->top_level() >= 1), sym(d->symmetric()), rap(galerkin),
          graphs(use_graphs), fuse_smoothing(fused_smoothing) {
        if (rap && d->kind() == 2)
            throw UGError("GMG: Galerkin coarse operators need the P1 lattice transfer (hexahedral or triangular "
                          "hierarchy)");
    }

    void alloc_vec(DeviceBuffer<T>& buf, T*& p, int64_t N, int64_t H) {
        buf.resize(N + 2 * H);
        buf.zero();
        p = buf.get() + H;
    }

    void init() override {
        auto& c = context();
        clear_graphs();
        if (base < 0 || base > top) throw UGError("GMG: base level out of range");
        L.clear();
        L.resize(top + 1);
        if (mf) {
            // the matrix-free operator (and the fused prolongation) take every
            // lattice boundary point as an identity (Dirichlet) row: the
            // Dirichlet DoFs, all on the boundary, must number the boundary
            // points
            const LevelGrid& Gt = disc->level(top);
            const int64_t Wi = Gt.W - 2;
            if (Gt.N - Gt.n_inner != Gt.N - Wi * Wi * Wi) mf = false;
        }
        if (mf) {
            // top level: every quadrature point of element e sees k_e, so the
            // element matrix is h k_e times the unit-cube Q1 stiffness
            LevelGrid& G = disc->level(top);
            const int64_t ne = (int64_t)G.n * G.n * G.n;
            cell_c.resize(ne);
            UGX_LAUNCH(q1mf_coeff<T>(ne, 1.0 / G.n, disc->theta_device(), cell_c.get(), c.stream));
            cell_c64.resize(ne);
            UGX_LAUNCH(q1mf_coeff<double>(ne, 1.0 / G.n, disc->theta_device(), cell_c64.get(), c.stream));
            UGX_LAUNCH(q1mf_prepare(G.W));
        }
        // Galerkin hierarchy (AssembledMultiGridCycle::init_rap_operator,
        // mg_solver_impl.hpp:829-1013): the assembled top-level operator, then
        // A_{l-1} = R_l A_l P_l from the top down
        std::vector<DeviceBuffer<double>> galerkin(rap ? top + 1 : 0);
        std::vector<int> galerkin_nd(rap ? top + 1 : 0);
        std::vector<std::vector<int32_t>> galerkin_off(rap ? top + 1 : 0);
        if (rap) {
            disc->assemble_level_dia(top, galerkin[top], galerkin_nd[top], galerkin_off[top]);
            for (int lev = top; lev > base; --lev) {
                LevelGrid& Gf = disc->level(lev);
                LevelGrid& Gc = disc->level(lev - 1);
                disc->level_dia_layout(lev - 1, galerkin_nd[lev - 1], galerkin_off[lev - 1]);
                galerkin[lev - 1].resize(dia_tiled_size(Gc.N, galerkin_nd[lev - 1]));
                UGX_LAUNCH(rap_dia(Gf.dim, Gf.W, Gc.W, Gf.dir.get(), Gc.dir.get(), galerkin[lev].get(),
                                   galerkin[lev - 1].get(), c.stream));
            }
        }
        for (int lev = base; lev <= top; ++lev) {
            LevelGrid& G = disc->level(lev);
            GmgLevel<T>& l = L[lev];
            l.dim = G.dim;
            l.W = G.W;
            l.N = G.N;
            l.dir = G.dir.get();
            const int64_t W = G.W;
            l.H = G.dim == 3 ? ((W * W + W + 1 + 63) / 64) * 64 : ((W + 1 + 63) / 64) * 64;
            DeviceBuffer<double> A64;
            if (rap) {
                A64 = std::move(galerkin[lev]);
                l.nd = galerkin_nd[lev];
                l.off = galerkin_off[lev];
            } else {
                disc->assemble_level_dia(lev, A64, l.nd, l.off);
            }
            if (transposed) {
                if (l.nd != 27) throw UGError("GMG: transposed level operators need a 3D lattice");
                DeviceBuffer<double> AT(A64.size());
                UGX_LAUNCH(dia_transpose_elim<double>(l.W, l.N, l.dir, A64.get(), AT.get(), c.stream));
                A64 = std::move(AT);
            }
            alloc_vec(l.b_dinv, l.dinv, l.N, l.H);
            alloc_vec(l.b_sc, l.sc, l.N, l.H);
            alloc_vec(l.b_sd, l.sd, l.N, l.H);
            alloc_vec(l.b_sd2, l.sd2, l.N, l.H);
            alloc_vec(l.b_st, l.st, l.N, l.H);
            // Jacobi: diagInv = 1/(a_ii * (1/omega)) from the FP64 centre diagonal
            DeviceBuffer<double> dinv64(l.N);
            UGX_HIP(hipMemsetAsync(c.flag, 0, sizeof(int), c.stream));
            UGX_LAUNCH((dia_jacobi_diag<double, double>(l.N, l.nd, A64.get(), omega, dinv64.get(), c.flag, c.stream)));
            if (lev == base) {
                std::vector<double> h((size_t)dia_tiled_size(l.N, l.nd));
                A64.download(h.data(), (int64_t)h.size());
                lu_n = (int)l.N;
                if (lu_n > 4096) throw UGError("GMG: base level too large for the dense LU");
                std::vector<double> dense((size_t)lu_n * lu_n, 0.0);
                for (int64_t i = 0; i < l.N; ++i)
                    for (int d = 0; d < l.nd; ++d) {
                        const int64_t j = i + l.off[d];
                        const double v = h[(size_t)dia_tiled_index(d, i, l.nd)];
                        if (j < 0 || j >= l.N) continue;
                        dense[(size_t)i * lu_n + j] += v;
                    }
                std::vector<double> linvT, uinvT;
                std::vector<int32_t> perm;
                if (!dense_lu_factors(lu_n, dense, linvT, uinvT, perm))
                    throw UGError("GMG: base level matrix is singular");
                const int m = lu_n;
                lu_linvT.resize((int64_t)m * m);
                lu_linvT.upload(linvT.data(), (int64_t)m * m);
                lu_uinvT.resize((int64_t)m * m);
                lu_uinvT.upload(uinvT.data(), (int64_t)m * m);
                lu_perm.resize(m);
                lu_perm.upload(perm.data(), m);
            }
            if (!sym_ok) {
                // non-symmetric operator: full storage only
                if constexpr (std::is_same<T, double>::value) {
                    l.Afull = std::move(A64);
                    UGX_HIP(hipMemcpyAsync(l.dinv, dinv64.get(), sizeof(double) * l.N, hipMemcpyDeviceToDevice,
                                           c.stream));
                } else {
                    l.Afull.resize(A64.size());
                    UGX_LAUNCH((convert<double, T>(A64.size(), A64.get(), l.Afull.get(), c.stream)));
                    UGX_LAUNCH((convert<double, T>(l.N, dinv64.get(), l.dinv, c.stream)));
                }
                sync();
                int bad = 0;
                UGX_HIP(hipMemcpy(&bad, c.flag, sizeof(int), hipMemcpyDeviceToHost));
                if (bad && lev > base) throw UGError("GMG: Jacobi found a zero diagonal");
                continue;
            }
            // symmetric Dirichlet-eliminated storage of the level operator
            const int nu = l.nd / 2 + 1;
            DeviceBuffer<double> U64(dia_tiled_size(l.N, nu));
            UGX_LAUNCH(dia_symmetrize<double>(l.N, l.nd, l.off.data(), l.dir, A64.get(), U64.get(), c.stream));
            if (lev == top) {
                // the surface operator in lattice order (FP64, symmetric
                // storage) for the CG fast path
                surf.resize(U64.size());
                UGX_HIP(hipMemcpyAsync(surf.get(), U64.get(), sizeof(double) * U64.size(), hipMemcpyDeviceToDevice,
                                       c.stream));
            }
            if constexpr (std::is_same<T, double>::value) {
                l.A = std::move(U64);
                l.Afull = std::move(A64);
                UGX_HIP(hipMemcpyAsync(l.dinv, dinv64.get(), sizeof(double) * l.N, hipMemcpyDeviceToDevice, c.stream));
            } else {
                l.A.resize(U64.size());
                UGX_LAUNCH((convert<double, T>(U64.size(), U64.get(), l.A.get(), c.stream)));
                l.Afull.resize(A64.size());
                UGX_LAUNCH((convert<double, T>(A64.size(), A64.get(), l.Afull.get(), c.stream)));
                UGX_LAUNCH((convert<double, T>(l.N, dinv64.get(), l.dinv, c.stream)));
            }
            sync();
            int bad = 0;
            UGX_HIP(hipMemcpy(&bad, c.flag, sizeof(int), hipMemcpyDeviceToHost));
            if (bad && lev > base) throw UGError("GMG: Jacobi found a zero diagonal");
        }
    }

    // The reference updates the top-level defect once more after the last
    // post-smoothing step (mg_solver_impl.hpp:1954, on m_LocalFullRefLevel).
    // No caller reads that defect here: apply() returns only the correction,
    // apply_update_defect() forms d -= A c itself, and the lattice entry
    // converts the correction. The update is therefore skipped on every level,
    // and the top level ends with the same correction update as the levels
    // below it; the lattice entry fuses that update with its conversion
    // (defer_top_final).
    bool defer_top_final = false;
    // the lattice entry's outputs while its cycle runs (z, r, the partials'
    // scratch and (z, r)), and whether the fused last step has written them
    double* last_z = nullptr;
    const double* last_r = nullptr;
    double* last_partial = nullptr;
    double* last_out = nullptr;
    bool last_done = false;
    // the lattice entry's FP64 defect when the first top-level pre-smoothing
    // step reads it directly (q1mf_jacobi_first) instead of a converted copy,
    // or the caller's FP32 copy of it (q1mf_jacobi_first_t, FP32 levels)
    const double* first_defect = nullptr;
    const float* first_defect_f32 = nullptr;
    void clear_cycle_state() {
        defer_top_final = false;
        last_done = false;
        first_defect = nullptr;
        first_defect_f32 = nullptr;
        for (auto& l : L) l.acc_deferred = false;
    }
    bool first_step_fused() const { return mf && nu1 > 0 && top > base && q1mf_first_fits(L[top].W, sizeof(T)); }

    void base_solve(int lev) {
        auto& c = context();
        GmgLevel<T>& l = L[lev];
        UGX_LAUNCH(lu_solve<T>(lu_n, lu_linvT.get(), lu_uinvT.get(), lu_perm.get(), l.sd, l.sc, c.stream));
    }

    void presmooth_and_restriction(int lev) {
        auto& c = context();
        GmgLevel<T>& f = L[lev];
        GmgLevel<T>& cl = L[lev - 1];
        for (int nu = 0; nu < nu1; ++nu) {
            if (nu == 0 && lev == top && (first_defect || first_defect_f32)) {
                if constexpr (std::is_same<T, float>::value) {
                    if (first_defect_f32)
                        UGX_LAUNCH(q1mf_jacobi_first_t<T>(f.W, cell_c.get(), f.dinv, first_defect_f32, f.sd2, f.sc,
                                                          c.stream));
                    else
                        UGX_LAUNCH(
                            q1mf_jacobi_first<T>(f.W, cell_c.get(), f.dinv, first_defect, f.sd2, f.sc, c.stream));
                } else {
                    UGX_LAUNCH(q1mf_jacobi_first<T>(f.W, cell_c.get(), f.dinv, first_defect, f.sd2, f.sc, c.stream));
                }
                first_defect = nullptr;
                first_defect_f32 = nullptr;
            } else {
                UGX_LAUNCH(op_jacobi_step(f));
            }
            std::swap(f.sd, f.sd2);
        }
        if (tets) {
            UGX_LAUNCH(vec_set<T>(cl.N, cl.sc, 0.0, c.stream));
            UGX_LAUNCH(restrict_tet<T>(cl.W, f.W, cl.N, f.sd, disc->level(lev).pcode.get(), cl.dir, f.dir, cl.sd,
                                       c.stream));
        } else {
            // the restriction also clears the coarse correction
            UGX_LAUNCH(restrict_defect<T>(f.dim, cl.W, f.W, cl.N, f.sd, cl.dir, f.dir, cl.sd, c.stream, cl.sc));
        }
    }

    // last: the caller prolongates this level's correction next (the level is
    // not visited again before that)
    void prolongation_and_postsmooth(int lev, bool last) {
        auto& c = context();
        GmgLevel<T>& f = L[lev];
        GmgLevel<T>& cl = L[lev - 1];
        if (tets) {
            UGX_LAUNCH(prolong_add_tet<T>(cl.W, f.W, f.N, cl.sc, disc->level(lev).pcode.get(), f.dir, f.st, f.sc,
                                          c.stream));
        } else if (cl.acc_deferred) {
            // the coarse level's last correction update, folded in
            UGX_LAUNCH(prolong_add<T>(f.dim, cl.W, f.W, f.N, cl.sc, f.dir, f.st, f.sc, c.stream, cl.dinv, cl.sd));
        } else {
            UGX_LAUNCH(prolong_add<T>(f.dim, cl.W, f.W, f.N, cl.sc, f.dir, f.st, f.sc, c.stream));
        }
        cl.acc_deferred = false;
        if (nu2 == 0) return;
        UGX_LAUNCH(op_apply_sub(f, f.st, f.sd, f.sd));
        // matrix-free top level of the lattice entry: the last full smoothing
        // step, the deferred correction update, its conversion and (z, r) in
        // one pass (run_lattice then has nothing left to do)
        const bool fuse_last = fuse_smoothing && defer_top_final && lev == top && sym && mf && nu2 >= 2 &&
                               q1mf_last_fits(f.W, sizeof(T));
        for (int nu = 0; nu < nu2; ++nu) {
            if (fuse_last && nu == nu2 - 2) {
                if (last_out) {
                    int np = AssembledMultiGridCycle::max_partials();
                    UGX_LAUNCH(q1mf_jacobi_last<T>(f.W, cell_c.get(), f.dinv, f.sd, f.sc, last_z, last_r, c.stream,
                                                   last_partial, &np));
                    UGX_LAUNCH(sum_partials(np, last_partial, last_out, c.stream));
                } else {
                    UGX_LAUNCH(q1mf_jacobi_last<T>(f.W, cell_c.get(), f.dinv, f.sd, f.sc, last_z, last_r, c.stream,
                                                   nullptr, nullptr));
                }
                last_done = true;
            } else if (nu < nu2 - 1) {
                UGX_LAUNCH(op_jacobi_step(f));
                std::swap(f.sd, f.sd2);
            } else if (lev < top && !tets && last) {
                // folded into the prolongation to the next level (only on the
                // level's last visit: an earlier visit of a W- or F-cycle is
                // followed by another one that smooths with sc updated)
                f.acc_deferred = true;
            } else if (!(defer_top_final && lev == top)) {
                UGX_LAUNCH(jacobi_accumulate<T>(f.N, f.dinv, f.sd, f.sc, c.stream));
            }
        }
    }

    // mg_solver_impl.hpp:2090-2136: the base level is solved once; an F-cycle
    // (type -1) runs one F-cycle and then one V-cycle on the coarser level,
    // a V/W/gamma-cycle runs gamma cycles of its own type there
    void lmgc(int lev, int type, bool last) {
        if (lev == base) {
            base_solve(top);
            return;
        }
        presmooth_and_restriction(lev);
        if (lev - 1 == base) {
            base_solve(lev - 1);
        } else if (type == AssembledMultiGridCycle::_F_) {
            lmgc(lev - 1, AssembledMultiGridCycle::_F_, false);
            lmgc(lev - 1, AssembledMultiGridCycle::_V_, true);
        } else {
            for (int i = 0; i < type; ++i) lmgc(lev - 1, type, i == type - 1);
        }
        prolongation_and_postsmooth(lev, last);
    }

    void apply(GPUVector& cvec, const GPUVector& d) override {
        auto& c = context();
        LevelGrid& G = disc->level(top);
        GmgLevel<T>& t = L[top];
        if (d.size() != t.N || cvec.size() != t.N) throw UGError("GMG::apply: size mismatch");
        // Dirichlet entries of the defect (DoFs n_inner..N-1): zero in the
        // reference workflow (adjusted solution and rhs); otherwise the full
        // level storage is used
        double tail = 0.0;
        if (G.N > G.n_inner) {
            UGX_LAUNCH(dot<double>(G.N - G.n_inner, d.data() + G.n_inner, d.data() + G.n_inner, c.red_partial,
                                   c.red_out, false, c.stream));
            UGX_HIP(hipMemcpyAsync(c.red_host, c.red_out, sizeof(double), hipMemcpyDeviceToHost, c.stream));
            sync();
            tail = c.red_host[0];
        }
        sym = sym_ok && (tail == 0.0);
        UGX_LAUNCH((permute_to<double, T>(t.N, G.dof_of_lat.get(), d.data(), t.sd, c.stream)));
        cvec.set(0.0);
        UGX_LAUNCH(vec_set<T>(t.N, t.sc, 0.0, c.stream));
        try {
            lmgc(top, cycle, true);
        } catch (...) {
            clear_cycle_state();
            throw;
        }
        UGX_LAUNCH((permute_add_from<T, double>(t.N, G.lat_of_dof.get(), t.sc, cvec.data(), c.stream)));
        sym = sym_ok;
    }

    void run_lattice(double* c_lat, const double* d_lat, double* cd_out, double* partial, const float* d_lat_f32) {
        auto& c = context();
        GmgLevel<T>& t = L[top];
        sym = true;
        // with pre-smoothing on the matrix-free top level the first step reads
        // the FP64 defect itself and starts from a zero correction
        if (first_step_fused()) {
            first_defect = d_lat;
            if (std::is_same<T, float>::value) first_defect_f32 = d_lat_f32;
        } else {
            UGX_LAUNCH((convert<double, T>(t.N, d_lat, t.sd, c.stream)));
            UGX_LAUNCH(vec_set<T>(t.N, t.sc, 0.0, c.stream));
        }
        // with post-smoothing below a coarser level the cycle ends with the
        // top level's correction update, fused here with the conversion
        defer_top_final = nu2 > 0 && top > base;
        last_z = c_lat;
        last_r = d_lat;
        last_partial = partial;
        last_out = cd_out;
        last_done = false;
        try {
            lmgc(top, cycle, true);
        } catch (...) {
            clear_cycle_state();
            throw;
        }
        if (defer_top_final) {
            defer_top_final = false;
            if (!last_done)
                UGX_LAUNCH(jacobi_convert_dot<T>(t.N, t.dinv, t.sd, t.sc, c_lat, d_lat, partial, cd_out, c.stream));
        } else if (cd_out) {
            UGX_LAUNCH(convert_dot<T>(t.N, t.sc, c_lat, d_lat, partial, cd_out, c.stream));
        } else {
            UGX_LAUNCH((convert<T, double>(t.N, t.sc, c_lat, c.stream)));
        }
    }

    // One cycle is about 60 launches, most of them on the small coarse levels
    // where the launch costs more than the kernel. The cycle has no host
    // decisions and every level vector is rewritten before it is read, so it
    // is captured once per (c, d, cd_out, partial) and replayed as one graph
    // (kernel nodes only: the zeroing is a kernel, not a memset).
    // FP32 levels with the fused first step can start from the caller's FP32
    // copy of the defect (rounded as the conversion would)
    bool uses_f32_defect() const override { return std::is_same<T, float>::value && first_step_fused(); }

    void apply_lattice(double* c_lat, const double* d_lat, double* cd_out, double* partial,
                       const float* d_lat_f32) override {
        auto& c = context();
        if (!sym_ok) throw UGError("GMG::apply_lattice: needs a symmetric discretisation");
        if (!uses_f32_defect()) d_lat_f32 = nullptr;
        if (!graphs) {
            run_lattice(c_lat, d_lat, cd_out, partial, d_lat_f32);
            return;
        }
        const void* key[5] = {c_lat, d_lat, cd_out, partial, d_lat_f32};
        for (auto& g : cycle_graphs)
            if (std::equal(key, key + 5, g.key)) {
                UGX_HIP(hipGraphLaunch(g.exec, c.stream));
                return;
            }
        if (cycle_graphs.size() >= 4) clear_graphs();
        UGX_HIP(hipStreamBeginCapture(c.stream, hipStreamCaptureModeThreadLocal));
        try {
            run_lattice(c_lat, d_lat, cd_out, partial, d_lat_f32);
        } catch (...) {
            hipGraph_t broken = nullptr;
            (void)hipStreamEndCapture(c.stream, &broken);
            if (broken) (void)hipGraphDestroy(broken);
            throw;
        }
        hipGraph_t graph = nullptr;
        UGX_HIP(hipStreamEndCapture(c.stream, &graph));
        hipGraphExec_t exec = nullptr;
        const hipError_t e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        if (e != hipSuccess) throw UGError(std::string("GMG: graph instantiation failed: ") + hipGetErrorString(e));
        CycleGraph g;
        std::copy(key, key + 5, g.key);
        g.exec = exec;
        cycle_graphs.push_back(g);
        UGX_HIP(hipGraphLaunch(exec, c.stream));
    }

    bool surface_axpy(double* y, double alpha, const double* v, double beta, const double* x, double* yx_out,
                      double* partial) override {
        GmgLevel<T>& t = L[top];
        auto& c = context();
        if (mf) {
            if (yx_out) {
                int np = AssembledMultiGridCycle::max_partials();
                if (q1mf_axpy<double>(t.W, cell_c64.get(), alpha, v, beta, x, y, c.stream, partial, &np) != 0)
                    return false;
                UGX_LAUNCH(sum_partials(np, partial, yx_out, c.stream));
                return true;
            }
            UGX_LAUNCH(q1mf_axpy<double>(t.W, cell_c64.get(), alpha, v, beta, x, y, c.stream));
            return true;
        }
        if (yx_out) return false;
        UGX_LAUNCH((dias_axpy<double, double>(t.N, t.nd, t.off.data(), surf.get(), alpha, v, beta, x, y, c.stream)));
        return true;
    }

    // p_new = z + beta p_old and q = A p_new with (q, p_new) in one pass
    // (matrix-free top level only)
    bool surface_direction_axpy(double* q, const double* p_old, const double* z, const double* beta, double* p_new,
                                double* qp_out, double* partial) override {
        if (!mf) return false;
        auto& c = context();
        int np = AssembledMultiGridCycle::max_partials();
        if (q1mf_direction_axpy<double>(L[top].W, cell_c64.get(), p_old, z, beta, p_new, q, c.stream, partial, &np) !=
            0)
            return false;
        UGX_LAUNCH(sum_partials(np, partial, qp_out, c.stream));
        return true;
    }

    int64_t halo() const override { return L[top].H; }

    void download_level_dia(int lev, double* out) const override {
        // the stored operator expanded to all nd diagonals (out[d*N + i]):
        // the Dirichlet-eliminated level matrix, lower half read from the
        // mirrored upper entries
        const GmgLevel<T>& l = L[lev];
        if (!sym_ok) {
            std::vector<T> h(l.Afull.size());
            l.Afull.download(h.data(), (int64_t)h.size());
            for (int d = 0; d < l.nd; ++d)
                for (int64_t i = 0; i < l.N; ++i) out[(int64_t)d * l.N + i] = (double)h[dia_tiled_index(d, i, l.nd)];
            return;
        }
        const int nu = l.nd / 2 + 1;
        std::vector<T> h(l.A.size());
        l.A.download(h.data(), (int64_t)h.size());
        for (int d = 0; d < l.nd; ++d)
            for (int64_t i = 0; i < l.N; ++i) {
                double v = 0.0;
                if (d >= l.nd / 2) {
                    v = (double)h[dia_tiled_index(d - l.nd / 2, i, nu)];
                } else {
                    const int64_t j = i + l.off[d];
                    if (j >= 0 && j < l.N) v = (double)h[dia_tiled_index(l.nd - 1 - d - l.nd / 2, j, nu)];
                }
                out[(int64_t)d * l.N + i] = v;
            }
    }
    int num_diags(int lev) const override { return L[lev].nd; }
    void smooth(int lev, int nsteps) override {
        if (lev <= base || lev > top) throw UGError("GMG::smooth_level: level out of range");
        auto& c = context();
        GmgLevel<T>& f = L[lev];
        for (int k = 0; k < nsteps; ++k) {
            UGX_LAUNCH(op_jacobi_step(f));
            std::swap(f.sd, f.sd2);
        }
    }
};

}  // namespace

AssembledMultiGridCycle::AssembledMultiGridCycle(std::shared_ptr<DomainDiscretization> disc) : disc_(std::move(disc)) {}
AssembledMultiGridCycle::~AssembledMultiGridCycle() { delete impl_; }

bool AssembledMultiGridCycle::lattice_ready(const GPUSparseMatrix& A) const {
    return impl_ && !transposed_ && disc_->symmetric() && mat_ == &A && A.version() == mat_version_ &&
           A.origin() == disc_.get();
}
int64_t AssembledMultiGridCycle::lattice_size() const { return disc_->level(disc_->top_level()).N; }
int64_t AssembledMultiGridCycle::lattice_halo() const { return impl_ ? impl_->halo() : 0; }
const int32_t* AssembledMultiGridCycle::dof_of_lattice() const { return disc_->level(disc_->top_level()).dof_of_lat.get(); }
const int32_t* AssembledMultiGridCycle::lattice_of_dof() const { return disc_->level(disc_->top_level()).lat_of_dof.get(); }
int64_t AssembledMultiGridCycle::num_inner_dofs() const { return disc_->level(disc_->top_level()).n_inner; }
void AssembledMultiGridCycle::apply_lattice(double* c_lat, const double* d_lat, double* cd_out, double* partial,
                                            const float* d_lat_f32) {
    if (!impl_) throw UGError("GMG: not initialised");
    impl_->apply_lattice(c_lat, d_lat, cd_out, partial, d_lat_f32);
}
bool AssembledMultiGridCycle::lattice_uses_f32_defect() const { return impl_ && impl_->uses_f32_defect(); }
bool AssembledMultiGridCycle::surface_direction_axpy_lattice(double* q, const double* p_old, const double* z,
                                                             const double* beta, double* p_new, double* qp_out,
                                                             double* partial) {
    if (!impl_) throw UGError("GMG: not initialised");
    return impl_->surface_direction_axpy(q, p_old, z, beta, p_new, qp_out, partial);
}
bool AssembledMultiGridCycle::surface_axpy_lattice(double* y, double alpha, const double* v, double beta,
                                                   const double* x, double* yx_out, double* partial) {
    if (!impl_) throw UGError("GMG: not initialised");
    return impl_->surface_axpy(y, alpha, v, beta, x, yx_out, partial);
}

bool AssembledMultiGridCycle::init(std::shared_ptr<MatrixOperator> J) {
    op_ = J;
    mat_ = &J->get_matrix();
    mat_version_ = mat_->version();
    delete impl_;
    impl_ = nullptr;
    if (prec_ == Precision::F64)
        impl_ = new GmgImpl<double>(disc_.get(), base_lev_, nu1_, nu2_, cycle_, omega_, transposed_, matrix_free_,
                                    rap_, graphs_, fused_smoothing_);
    else
        impl_ = new GmgImpl<float>(disc_.get(), base_lev_, nu1_, nu2_, cycle_, omega_, transposed_, matrix_free_,
                                   rap_, graphs_, fused_smoothing_);
    impl_->init();
    return true;
}

bool AssembledMultiGridCycle::apply(GPUVector& c, const GPUVector& d) {
    if (!impl_) throw UGError("GMG::apply: not initialised");
    impl_->apply(c, d);
    return true;
}

void AssembledMultiGridCycle::download_level_dia(int lev, double* out) const {
    if (!impl_) throw UGError("GMG: not initialised");
    impl_->download_level_dia(lev, out);
}
void AssembledMultiGridCycle::smooth_level(int lev, int nsteps) {
    if (!impl_) throw UGError("GMG: not initialised");
    impl_->smooth(lev, nsteps);
}
int AssembledMultiGridCycle::level_num_diags(int lev) const { return impl_ ? impl_->num_diags(lev) : 0; }

// ---------------------------------------------------------------------------
// MatrixMultiGrid
// ---------------------------------------------------------------------------
void MatrixMultiGrid::set_levels(std::vector<std::shared_ptr<GPUSparseMatrix>> A,
                                 std::vector<std::shared_ptr<GPUSparseMatrix>> P,
                                 std::vector<std::shared_ptr<GPUSparseMatrix>> R) {
    if (A.empty() || P.size() != A.size() || R.size() != A.size())
        throw UGError("MatrixMultiGrid: one A, P and R entry per level (P[0], R[0] unused)");
    A_ = std::move(A);
    P_ = std::move(P);
    R_ = std::move(R);
    top_ = (int)A_.size() - 1;
}

bool MatrixMultiGrid::init(std::shared_ptr<MatrixOperator> J) {
    // the checks of set_cycle_type / init (mg_solver_impl.hpp:379-496)
    if (cycle_ != _F_ && cycle_ < 1) throw UGError("GMG: cycle type must be 1 (V), 2 (W), gamma > 2 or -1 (F)");
    if (nu1_ < 0 || nu2_ < 0) throw UGError("GMG: smoothing step counts must be >= 0");
    if (top_ < 0) throw UGError("MatrixMultiGrid: no levels");
    if (base_ < 0 || base_ > top_) throw UGError("MatrixMultiGrid: base level out of range");
    op_ = J;
    auto rows = [&](int l) { return Aop(l).num_rows(); };
    if (rows(top_) != A_[top_]->num_rows() || Aop(top_).num_cols() != rows(top_))
        throw UGError("MatrixMultiGrid: the operator does not match the top level");
    for (int l = base_; l <= top_; ++l) {
        if (Aop(l).building()) throw UGError("MatrixMultiGrid: a level operator is still being built");
        if (l > base_) {
            if (!P_[l] || !R_[l]) throw UGError("MatrixMultiGrid: missing transfer matrix");
            if (P_[l]->num_rows() != rows(l) || P_[l]->num_cols() != rows(l - 1) || R_[l]->num_rows() != rows(l - 1) ||
                R_[l]->num_cols() != rows(l))
                throw UGError("MatrixMultiGrid: transfer matrix sizes do not match the levels");
        }
    }
    sc_.assign(top_ + 1, GPUVector());
    sd_.assign(top_ + 1, GPUVector());
    st_.assign(top_ + 1, GPUVector());
    smoother_.clear();
    smoother_.resize(top_ + 1);
    for (int l = base_; l <= top_; ++l) {
        sc_[l].resize(rows(l));
        sd_[l].resize(rows(l));
        st_[l].resize(rows(l));
        if (l > base_) {
            smoother_[l].reset(new Jacobi(omega_));
            auto opl = l == top_ ? J : std::make_shared<MatrixOperator>(A_[l]);
            if (!smoother_[l]->init(opl)) throw UGError("GMG: Jacobi found a zero diagonal");
        }
    }
    // base level: dense LU (lu.h:236-269, dense below the sparse threshold)
    GPUSparseMatrix& B = Aop(base_);
    const int64_t m = B.num_rows();
    if (m > 4000) throw UGError("MatrixMultiGrid: the base level is too large for the dense LU (> 4000 rows)");
    std::vector<int32_t> rp(m + 1), cl(B.total_num_connections());
    std::vector<double> vl(B.total_num_connections());
    B.download_csr(rp.data(), cl.data(), vl.data());
    std::vector<double> dense((size_t)m * m, 0.0), linvT, uinvT;
    for (int64_t r = 0; r < m; ++r)
        for (int32_t k = rp[r]; k < rp[r + 1]; ++k) dense[(size_t)r * m + cl[k]] += vl[k];
    std::vector<int32_t> perm;
    if (!dense_lu_factors((int)m, dense, linvT, uinvT, perm)) throw UGError("GMG: base level matrix is singular");
    lu_n_ = (int)m;
    lu_linvT_.resize(m * m);
    lu_linvT_.upload(linvT.data(), m * m);
    lu_uinvT_.resize(m * m);
    lu_uinvT_.upload(uinvT.data(), m * m);
    lu_perm_.resize(m);
    lu_perm_.upload(perm.data(), m);
    return true;
}

bool MatrixMultiGrid::apply(GPUVector& c, const GPUVector& d) {
    if (top_ < 0 || (int)sc_.size() != top_ + 1) throw UGError("MatrixMultiGrid: not initialised");
    if (d.size() != sd_[top_].size() || c.size() != sd_[top_].size()) throw UGError("GMG::apply: size mismatch");
    // mg_solver_impl.hpp:175-275
    sd_[top_] = d;
    c.set(0.0);
    sc_[top_].set(0.0);
    lmgc(top_, cycle_);
    c += sc_[top_];
    return true;
}

void MatrixMultiGrid::lmgc(int lev, int type) {
    // mg_solver_impl.hpp:2090-2136
    if (lev == base_) {
        base_solve(top_);
        return;
    }
    presmooth_and_restriction(lev);
    if (lev - 1 == base_) {
        base_solve(lev - 1);
    } else if (type == _F_) {
        lmgc(lev - 1, _F_);
        lmgc(lev - 1, _V_);
    } else {
        for (int i = 0; i < type; ++i) lmgc(lev - 1, type);
    }
    prolongation_and_postsmooth(lev);
}

void MatrixMultiGrid::presmooth_and_restriction(int lev) {
    // mg_solver_impl.hpp:1686-1816
    GPUVector &sc = sc_[lev], &sd = sd_[lev], &st = st_[lev];
    for (int nu = 0; nu < nu1_; ++nu) {
        smoother_[lev]->apply(st, sd);
        Aop(lev).matmul_minus(sd, st);
        sc += st;
    }
    sc_[lev - 1].set(0.0);
    // StdTransfer::do_restrict (std_transfer_impl.h:774-806)
    R_[lev]->apply_ignore_zero_rows(sd_[lev - 1], 1.0, sd);
}

void MatrixMultiGrid::prolongation_and_postsmooth(int lev) {
    // mg_solver_impl.hpp:1819-1964; StdTransfer::prolongate (std_transfer_impl.h:719-772)
    GPUVector &sc = sc_[lev], &sd = sd_[lev], &st = st_[lev];
    P_[lev]->axpy(st, 0.0, st, 1.0, sc_[lev - 1]);
    sc += st;
    for (int nu = 0; nu < nu2_; ++nu) {
        Aop(lev).matmul_minus(sd, st);
        smoother_[lev]->apply(st, sd);
        sc += st;
    }
    if (lev == top_) Aop(lev).matmul_minus(sd, st);
}

void MatrixMultiGrid::base_solve(int lev) {
    // mg_solver_impl.hpp:1968-2086: the base level is solved once
    auto& c = context();
    UGX_LAUNCH(lu_solve<double>(lu_n_, lu_linvT_.get(), lu_uinvT_.get(), lu_perm_.get(), sd_[base_].data(),
                                sc_[base_].data(), c.stream));
    if (lev == top_) Aop(lev).matmul_minus(sd_[lev], sc_[lev]);
}

// ---------------------------------------------------------------------------
// Q2Discretization
// ---------------------------------------------------------------------------
Q2Discretization::Q2Discretization(int num_refs) : refs_(num_refs) {
    if (num_refs < 0 || num_refs > 6) throw UGError("Q2Discretization: num_refs must be in [0, 6]");
    auto& c = context();
    // the vertices of level l+1 of the hexahedral hierarchy are the Q2 nodes
    // of level l, and its P1 numbering orders them within each subset
    std::vector<std::unique_ptr<LevelGrid>> grids;
    build_device_levels(0, num_refs + 1, 0, grids);
    L_.resize(num_refs + 1);
    for (int l = 0; l <= num_refs; ++l) {
        LevelGrid& G = *grids[l + 1];
        Level& Q = L_[l];
        Q.n = 1 << l;
        Q.N = G.N;
        DeviceBuffer<int32_t> lat_of_dof(G.N);
        UGX_LAUNCH(invert_map(G.N, G.dof_of_lat.get(), lat_of_dof.get(), c.stream));
        size_t tb = 0;
        if (q2_numbering_temp_bytes(G.N, &tb) != 0) throw UGError("Q2Discretization: sort set-up failed");
        DeviceBuffer<uint64_t> key(G.N), key2(G.N);
        DeviceBuffer<int32_t> node(G.N);
        DeviceBuffer<uint8_t> temp((int64_t)std::max<size_t>(tb, 1));
        Q.node_of_dof.resize(G.N);
        Q.dof_of_node.resize(G.N);
        Q.dir.resize(G.N);
        UGX_LAUNCH(q2_numbering(G.N, G.n_inner, G.W, lat_of_dof.get(), key.get(), key2.get(), node.get(),
                                Q.node_of_dof.get(), temp.get(), tb, Q.dof_of_node.get(), Q.dir.get(), c.stream));
        sync();   // the scratch is released next
    }
}

void Q2Discretization::set_theta(const double* theta_host) {
    if (!theta_host) {
        has_theta_ = false;
    } else {
        const int64_t n = (int64_t)1 << refs_;
        theta_.resize(n * n * n);
        theta_.upload(theta_host, n * n * n);
        has_theta_ = true;
    }
    for (auto& Q : L_) Q.coef_valid = false;
}

void Q2Discretization::dof_positions(int lev, double* out) const {
    const Level& Q = L_.at(lev);
    std::vector<int32_t> v((size_t)Q.N);
    Q.node_of_dof.download(v.data(), Q.N);
    const int64_t W = 2 * (int64_t)Q.n + 1;
    const double s = 1.0 / (2.0 * Q.n);
    for (int64_t k = 0; k < Q.N; ++k) {
        out[3 * k] = (double)(v[(size_t)k] % W) * s;
        out[3 * k + 1] = (double)((v[(size_t)k] / W) % W) * s;
        out[3 * k + 2] = (double)(v[(size_t)k] / (W * W)) * s;
    }
}

void Q2Discretization::dirichlet_flags(int lev, uint8_t* out) const {
    const Level& Q = L_.at(lev);
    Q.dir.download(out, Q.N);
}

void Q2Discretization::assemble_linear(int lev, GPUSparseMatrix& A, GPUVector& b) {
    Level& Q = L_.at(lev);
    auto& c = context();
    const int64_t ne = (int64_t)Q.n * Q.n * Q.n;
    if (!Q.elem) {
        DeviceBuffer<int32_t> conn(ne * 27);
        DeviceBuffer<double> X(Q.N * 3);
        UGX_LAUNCH(q2_connectivity(Q.n, Q.dof_of_node.get(), conn.get(), c.stream));
        UGX_LAUNCH(q2_coords(Q.N, Q.n, Q.node_of_dof.get(), X.get(), c.stream));
        Q.elem.reset(
            new ElementDiscretization(ElementDiscretization::HEXAHEDRON_Q2, ne, conn.get(), Q.N, X.get(), true));
        std::vector<uint8_t> dir((size_t)Q.N);
        Q.dir.download(dir.data(), Q.N);
        std::vector<int32_t> dl;
        for (int64_t i = 0; i < Q.N; ++i)
            if (dir[(size_t)i]) dl.push_back((int32_t)i);
        Q.elem->set_dirichlet(dl.data(), (int64_t)dl.size());
        Q.coef_valid = false;
    }
    if (!Q.coef_valid) {
        if (has_theta_) {
            DeviceBuffer<double> kq(ne * 27);
            UGX_LAUNCH(q2_coef_points(Q.n, 1 << refs_, theta_.get(), kq.get(), c.stream));
            Q.elem->set_diffusion_at_points(kq.get(), true);
            sync();   // the copy has read kq
        } else {
            Q.elem->set_diffusion_at_points(nullptr);
            Q.elem->set_diffusion(nullptr);
        }
        Q.coef_valid = true;
    }
    Q.elem->set_source(f_);
    Q.elem->assemble_linear(A, b);
}

void Q2Discretization::prolongation(int lev, GPUSparseMatrix& P) const {
    if (lev < 1 || lev >= (int)L_.size()) throw UGError("Q2Discretization::prolongation: level out of range");
    const Level& F = L_[lev];
    const Level& C = L_[lev - 1];
    auto& c = context();
    DeviceBuffer<int32_t> cnt(F.N), rowptr(F.N + 1);
    DeviceBuffer<int64_t> incl(F.N);
    UGX_LAUNCH(q2_prolong_count(F.N, F.n, F.node_of_dof.get(), cnt.get(), c.stream));
    size_t tb = 0;
    if (rowptr_scan_temp_bytes(F.N, &tb) != 0) throw UGError("Q2Discretization: scan set-up failed");
    DeviceBuffer<uint8_t> temp((int64_t)std::max<size_t>(tb, 1));
    UGX_LAUNCH(rowptr_from_counts(F.N, cnt.get(), incl.get(), temp.get(), tb, rowptr.get(), c.stream));
    int64_t nnz = 0;
    incl.download(&nnz, 1, F.N - 1);
    P.allocate(F.N, C.N, nnz);
    P.set_origin(nullptr);
    UGX_HIP(hipMemcpyAsync(P.rowptr(), rowptr.get(), sizeof(int32_t) * (F.N + 1), hipMemcpyDeviceToDevice, c.stream));
    UGX_LAUNCH(q2_prolong_fill(F.N, F.n, F.node_of_dof.get(), F.dir.get(), C.dof_of_node.get(), P.rowptr(), P.col(),
                               P.val(), c.stream));
    sync();
    P.build_spmv_layout();
}

void Q2Discretization::restriction(int lev, const GPUSparseMatrix& P, GPUSparseMatrix& R) const {
    if (lev < 1 || lev >= (int)L_.size()) throw UGError("Q2Discretization::restriction: level out of range");
    const Level& F = L_[lev];
    const Level& C = L_[lev - 1];
    if (P.num_rows() != F.N || P.num_cols() != C.N) throw UGError("Q2Discretization::restriction: P size mismatch");
    auto& c = context();
    R.set_as_transpose_of(P);   // every stored entry, explicit zeros included
    R.values_edited_externally();
    DeviceBuffer<int32_t> last(C.N);
    UGX_HIP(hipMemsetAsync(last.get(), 0xff, sizeof(int32_t) * (size_t)C.N, c.stream));
    UGX_LAUNCH(q2_restriction_adjust(F.N, F.n, C.N, F.node_of_dof.get(), F.dir.get(), C.dof_of_node.get(), last.get(),
                                     R.rowptr(), R.col(), R.val(), c.stream));
    sync();
    R.finalize();
}

}  // namespace gpu
}  // namespace ug

