// Optimiser passes (RMSprop, SGD) over row-strided voxel data for gfx950.
//
// Update rules (the reference's, svox2/csrc/optim_kernel.cu:15-25,97-103):
//   RMSprop: rms  = (rms == 0) ? g^2 : lerp(g^2, rms, beta)
//            data = max(data - lr * g / (sqrt(rms) + eps), minval),  g = 0
//   SGD:     data = data - lr * g,                              g = 0
// BF16 data is updated in fp32 and rounded to nearest-even on store.
//
// Layout of the work.  Rows are chosen by every row, a byte mask, or an int64
// row list.  A wavefront takes 64 candidate rows at a time: one lane reads one
// candidate (a mask byte or an index, 64 consecutive entries per load), a
// ballot compacts the selected rows, and the wave then updates them two at a
// time, one 32-lane half per row, lane = column.  A row's columns are
// contiguous, so each half reads and writes its data, rms and gradient rows
// as single coalesced accesses, and unselected rows cost only their mask byte:
// the reference launches one thread per (row, column) whatever the mask.
//
// The columns of a row form two groups, [0, split) and [split, ncols), each
// with its own mask, learning rate and beta.  With the interleaved voxel row
// [SH 0..26 | density] this updates SH and density in one pass over one
// 128-byte line per array; with split = ncols - 1 it is the reference's
// lr_last for the last column.  Single-column arrays (density on its own)
// are updated lane per row.
#include "common.hpp"

#include <algorithm>
#include <climits>

namespace svox2amd {
namespace {

constexpr int kBlock = 256;

template <typename DT>
struct RowJob {
    DT* data;
    float* rms;  // NULL for SGD
    float* grad;
    int rows;
    int ncols;
    int split;
    int64_t sd, sr, sg;  // row strides of data, rms, grad
    int kind;            // SVOX2_INDEXER_*
    const uint8_t* mask_a;
    const uint8_t* mask_b;
    const int64_t* index;
    int64_t n_index;
    float lr_a, lr_b, beta_a, beta_b;
    float eps, minval;
};

template <typename DT, bool SGD>
__device__ __forceinline__ void update_one(const RowJob<DT>& J, int row, int col, float lr,
                                           float beta) {
    const int64_t og = row * J.sg + col, od = row * J.sd + col;
    const float g = J.grad[og];
    const float x = ld(J.data, od);
    if (SGD) {
        st(J.data, od, x - lr * g);
    } else {
        const int64_t orr = row * J.sr + col;
        const float g2 = g * g;
        float r = J.rms[orr];
        r = (r == 0.f) ? g2 : lerpf(g2, r, beta);
        J.rms[orr] = r;
        st(J.data, od, fmaxf(x - lr * g / (sqrtf(r) + J.eps), J.minval));
    }
    J.grad[og] = 0.f;
}

// Candidate k of the job: its row (or -1) and whether groups A / B select it.
// Written with one exit: every kind assigns the row on its own path.
template <typename DT>
__device__ __forceinline__ int candidate(const RowJob<DT>& J, int64_t k, int64_t ncand,
                                         bool& a, bool& b) {
    int row = -1;
    bool sa = false, sb = false;
    if (k < ncand) {
        if (J.kind == SVOX2_INDEXER_INDEX) {
            const int64_t r = J.index[k];
            const bool in = r >= 0 && r < J.rows;  // out of range: skipped
            row = in ? static_cast<int>(r) : -1;
            sa = sb = in;
        } else if (J.kind == SVOX2_INDEXER_MASK) {
            row = static_cast<int>(k);
            sa = J.mask_a[k] != 0;
            sb = J.mask_b[k] != 0;
        } else {
            row = static_cast<int>(k);
            sa = sb = true;
        }
    }
    a = sa;
    b = sb;
    return row;
}

template <typename DT, bool SGD>
__global__ __launch_bounds__(kBlock) void k_update_rows(RowJob<DT> J) {
    const int lane = threadIdx.x & 63;
    const int64_t wave = (static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x) >> 6;
    const int64_t nwaves = (static_cast<int64_t>(gridDim.x) * kBlock) >> 6;
    const int64_t ncand = J.kind == SVOX2_INDEXER_INDEX ? J.n_index : J.rows;
    for (int64_t base = wave * 64; base < ncand; base += nwaves * 64) {
        bool a, b;
        const int row = candidate(J, base + lane, ncand, a, b);
        if (J.ncols == 1) {  // one column, lane per row: group A if split > 0, else B
            if (row >= 0 && (J.split > 0 ? a : b))
                update_one<DT, SGD>(J, row, 0, J.split > 0 ? J.lr_a : J.lr_b,
                                    J.split > 0 ? J.beta_a : J.beta_b);
            continue;
        }
        const uint64_t ba = __ballot(row >= 0 && a);
        const uint64_t bb = __ballot(row >= 0 && b);
        uint64_t m = ba | bb;
        const int half = lane >> 5;
        while (m != 0) {  // uniform: two selected rows per pass, one per half
            const int j0 = __builtin_ctzll(m);
            m &= m - 1;
            int j1 = -1;
            if (m != 0) {
                j1 = __builtin_ctzll(m);
                m &= m - 1;
            }
            const int r0 = __builtin_amdgcn_readlane(row, j0);
            const int r1 = j1 >= 0 ? __builtin_amdgcn_readlane(row, j1) : -1;
            const int j = half ? j1 : j0;
            if (j < 0) continue;
            const int r = half ? r1 : r0;
            const bool use_a = (ba >> j) & 1, use_b = (bb >> j) & 1;
            for (int c = lane & 31; c < J.ncols; c += 32) {
                const bool in_a = c < J.split;
                if (in_a ? use_a : use_b)
                    update_one<DT, SGD>(J, r, c, in_a ? J.lr_a : J.lr_b,
                                        in_a ? J.beta_a : J.beta_b);
            }
        }
    }
}

template <typename DT>
int launch(const RowJob<DT>& J, bool sgd, hipStream_t s) {
    const int64_t ncand = J.kind == SVOX2_INDEXER_INDEX ? J.n_index : J.rows;
    const int64_t windows = (ncand + 63) / 64;
    // A wave per window of 64 candidates, up to 32 workgroups (128 waves)
    // per CU; waves then loop over windows.  Selected rows are updated two
    // at a time, so fewer windows per wave shorten each wave's chain of row
    // reads and writes.
    const int64_t blocks = std::min<int64_t>((windows + 3) / 4, 256 * 32);
    if (blocks == 0) return ok();
    if (sgd)
        k_update_rows<DT, true><<<static_cast<unsigned>(blocks), kBlock, 0, s>>>(J);
    else
        k_update_rows<DT, false><<<static_cast<unsigned>(blocks), kBlock, 0, s>>>(J);
    return check_launch(sgd ? "sgd_step" : "rmsprop_step");
}

// Validates the arrays of a row job and fills its strides and pointers.
int make_job(void* data, int32_t storage, float* rms, float* grad, int64_t rows, int32_t ncols,
             int64_t data_stride, int64_t rms_stride, int64_t grad_stride, bool need_rms,
             RowJob<float>& J) {
    SVOX2_REQUIRE(rows >= 0 && rows < INT_MAX, "rows %lld out of range", (long long)rows);
    SVOX2_REQUIRE(ncols >= 1, "bad column count %d", ncols);
    SVOX2_REQUIRE(storage == SVOX2_STORAGE_F32 || storage == SVOX2_STORAGE_BF16,
                  "unknown storage kind %d", storage);
    SVOX2_REQUIRE(data_stride >= 0 && rms_stride >= 0 && grad_stride >= 0,
                  "negative row stride");
    J.sd = data_stride > 0 ? data_stride : ncols;
    J.sr = rms_stride > 0 ? rms_stride : ncols;
    J.sg = grad_stride > 0 ? grad_stride : ncols;
    SVOX2_REQUIRE(J.sd >= ncols && J.sg >= ncols && (!need_rms || J.sr >= ncols),
                  "row stride below the column count");
    if (rows > 0) {
        SVOX2_REQUIRE(data != nullptr && grad != nullptr, "data / grad are NULL");
        SVOX2_REQUIRE(!need_rms || rms != nullptr, "rms is NULL");
    }
    J.data = static_cast<float*>(data);
    J.rms = rms;
    J.grad = grad;
    J.rows = static_cast<int>(rows);
    J.ncols = ncols;
    return ok();
}

template <typename DT>
RowJob<DT> retype(const RowJob<float>& J) {
    RowJob<DT> K;
    K.data = reinterpret_cast<DT*>(J.data);
    K.rms = J.rms;
    K.grad = J.grad;
    K.rows = J.rows;
    K.ncols = J.ncols;
    K.split = J.split;
    K.sd = J.sd;
    K.sr = J.sr;
    K.sg = J.sg;
    K.kind = J.kind;
    K.mask_a = J.mask_a;
    K.mask_b = J.mask_b;
    K.index = J.index;
    K.n_index = J.n_index;
    K.lr_a = J.lr_a;
    K.lr_b = J.lr_b;
    K.beta_a = J.beta_a;
    K.beta_b = J.beta_b;
    K.eps = J.eps;
    K.minval = J.minval;
    return K;
}

int run(const RowJob<float>& J, int32_t storage, bool sgd, void* stream) {
    hipStream_t s = static_cast<hipStream_t>(stream);
    if (storage == SVOX2_STORAGE_BF16) return launch(retype<uint16_t>(J), sgd, s);
    return launch(J, sgd, s);
}

// The reference's indexer (0-dim = all rows, bool mask, int64 indices) applied
// to both column groups, the last column taking lr_last.
int indexed_step(void* data, int32_t storage, float* rms, float* grad, int64_t rows,
                 int32_t dim, int64_t data_stride, int64_t rms_stride, int64_t grad_stride,
                 int32_t kind, const void* indexer, int64_t n_index, float beta, float lr,
                 float epsilon, float minval, float lr_last, bool sgd, void* stream) {
    SVOX2_REQUIRE(kind == SVOX2_INDEXER_ALL || kind == SVOX2_INDEXER_MASK ||
                      kind == SVOX2_INDEXER_INDEX,
                  "unknown indexer kind %d", kind);
    SVOX2_REQUIRE(n_index >= 0, "negative indexer length");
    RowJob<float> J{};
    const int st = make_job(data, storage, rms, grad, rows, dim, data_stride, rms_stride,
                            grad_stride, !sgd, J);
    if (st != SVOX2_OK) return st;
    if (lr_last < 0.f) lr_last = lr;
    if (rows == 0) return ok();
    if (kind != SVOX2_INDEXER_ALL && n_index == 0) return ok();  // empty indexer: skip
    SVOX2_REQUIRE(kind == SVOX2_INDEXER_ALL || indexer != nullptr, "indexer is NULL");
    if (kind == SVOX2_INDEXER_MASK)
        SVOX2_REQUIRE(n_index == rows, "mask has %lld entries for %lld rows",
                      (long long)n_index, (long long)rows);
    J.kind = kind;
    J.mask_a = J.mask_b = kind == SVOX2_INDEXER_MASK ? static_cast<const uint8_t*>(indexer)
                                                      : nullptr;
    J.index = kind == SVOX2_INDEXER_INDEX ? static_cast<const int64_t*>(indexer) : nullptr;
    J.n_index = n_index;
    J.split = dim - 1;
    J.lr_a = lr;
    J.lr_b = lr_last;
    J.beta_a = J.beta_b = beta;
    J.eps = epsilon;
    J.minval = minval;
    return run(J, storage, sgd, stream);
}

}  // namespace
}  // namespace svox2amd

using namespace svox2amd;

extern "C" int svox2_rmsprop_step(void* data, int32_t storage, float* rms, float* grad,
                                  int64_t rows, int32_t dim, int64_t data_stride,
                                  int64_t rms_stride, int64_t grad_stride, int32_t indexer_kind,
                                  const void* indexer, int64_t n_index, float beta, float lr,
                                  float epsilon, float minval, float lr_last, void* stream) {
    return indexed_step(data, storage, rms, grad, rows, dim, data_stride, rms_stride,
                        grad_stride, indexer_kind, indexer, n_index, beta, lr, epsilon, minval,
                        lr_last, false, stream);
}

extern "C" int svox2_sgd_step(void* data, int32_t storage, float* grad, int64_t rows,
                              int32_t dim, int64_t data_stride, int64_t grad_stride,
                              int32_t indexer_kind, const void* indexer, int64_t n_index,
                              float lr, float lr_last, void* stream) {
    return indexed_step(data, storage, nullptr, grad, rows, dim, data_stride, 0, grad_stride,
                        indexer_kind, indexer, n_index, 0.f, lr, 0.f, 0.f, lr_last, true,
                        stream);
}

extern "C" int svox2_rmsprop_step_grouped(void* data, int32_t storage, float* rms, float* grad,
                                          int64_t rows, int32_t ncols, int64_t data_stride,
                                          int64_t rms_stride, int64_t grad_stride,
                                          int32_t split, const uint8_t* mask_a,
                                          const uint8_t* mask_b, float beta_a, float lr_a,
                                          float beta_b, float lr_b, float epsilon,
                                          float minval, void* stream) {
    RowJob<float> J{};
    const int st = make_job(data, storage, rms, grad, rows, ncols, data_stride, rms_stride,
                            grad_stride, true, J);
    if (st != SVOX2_OK) return st;
    SVOX2_REQUIRE(0 <= split && split <= ncols, "split %d outside [0, %d]", split, ncols);
    SVOX2_REQUIRE((mask_a == nullptr) == (mask_b == nullptr),
                  "give both group masks or neither");
    if (rows == 0) return ok();
    J.kind = mask_a != nullptr ? SVOX2_INDEXER_MASK : SVOX2_INDEXER_ALL;
    J.mask_a = mask_a;
    J.mask_b = mask_b;
    J.index = nullptr;
    J.n_index = 0;
    J.split = split;
    J.lr_a = lr_a;
    J.lr_b = lr_b;
    J.beta_a = beta_a;
    J.beta_b = beta_b;
    J.eps = epsilon;
    J.minval = minval;
    return run(J, storage, false, stream);
}
