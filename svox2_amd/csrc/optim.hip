// RMSprop update for gfx950: a streaming elementwise pass (24 bytes per
// updated fp32 element) over every row, a byte-masked set of rows, or an
// int64 list of rows.  The gradient is zeroed where the update is applied.
//
// Semantics: rmsprop_once and the three rmsprop kernels
// (svox2/csrc/optim_kernel.cu:15-93) and the launcher (:154-218):
//   rms  = (rms == 0) ? g^2 : lerp(g^2, rms, beta)
//   data = max(data - lr * g / (sqrt(rms) + eps), minval),  g = 0,
// with lr_last used for the last channel of each row.  BF16 data is updated
// in fp32 and rounded to nearest-even on store.
#include "common.hpp"

namespace svox2amd {
namespace {

constexpr int kBlock = 256;

struct Strides {
    int64_t d, r, g;  // row strides of data, rms, grad
};

template <typename DT>
__device__ __forceinline__ void rmsprop_once(DT* data, float* rms, float* grad, int64_t row,
                                             int ch, const Strides& S, float beta, float lr,
                                             float eps, float minval) {
    const int64_t og = row * S.g + ch, orr = row * S.r + ch, od = row * S.d + ch;
    const float g = grad[og];
    const float g2 = g * g;
    float r = rms[orr];
    r = (r == 0.f) ? g2 : lerpf(g2, r, beta);
    rms[orr] = r;
    const float x = ld(data, od);
    st(data, od, fmaxf(x - lr * g / (sqrtf(r) + eps), minval));
    grad[og] = 0.f;
}

template <typename DT>
__global__ __launch_bounds__(kBlock) void k_rmsprop_all(DT* __restrict__ data,
                                                        float* __restrict__ rms,
                                                        float* __restrict__ grad, int64_t total,
                                                        int D, Strides S, float beta, float lr,
                                                        float eps, float minval, float lr_last) {
    const int64_t tid = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x;
    if (tid >= total) return;
    const int chnl = static_cast<int>(tid % D);
    rmsprop_once(data, rms, grad, tid / D, chnl, S, beta, chnl == D - 1 ? lr_last : lr, eps,
                 minval);
}

template <typename DT>
__global__ __launch_bounds__(kBlock) void k_rmsprop_mask(DT* __restrict__ data,
                                                         float* __restrict__ rms,
                                                         float* __restrict__ grad,
                                                         const uint8_t* __restrict__ mask,
                                                         int64_t total, int D, Strides S,
                                                         float beta, float lr, float eps,
                                                         float minval, float lr_last) {
    const int64_t tid = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x;
    if (tid >= total) return;
    if (!mask[tid / D]) return;
    const int chnl = static_cast<int>(tid % D);
    rmsprop_once(data, rms, grad, tid / D, chnl, S, beta, chnl == D - 1 ? lr_last : lr, eps,
                 minval);
}

template <typename DT>
__global__ __launch_bounds__(kBlock) void k_rmsprop_index(DT* __restrict__ data,
                                                          float* __restrict__ rms,
                                                          float* __restrict__ grad,
                                                          const int64_t* __restrict__ idx,
                                                          int64_t total, int D, int64_t rows,
                                                          Strides S, float beta, float lr,
                                                          float eps, float minval, float lr_last) {
    const int64_t tid = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x;
    if (tid >= total) return;
    const int64_t row = idx[tid / D];
    if (row < 0 || row >= rows) return;
    const int chnl = static_cast<int>(tid % D);
    rmsprop_once(data, rms, grad, row, chnl, S, beta, chnl == D - 1 ? lr_last : lr, eps, minval);
}

template <typename DT>
int launch(DT* data, float* rms, float* grad, int64_t rows, int D, Strides S, int kind,
           const void* indexer, int64_t n_index, float beta, float lr, float eps, float minval,
           float lr_last, hipStream_t s) {
    if (kind == SVOX2_INDEXER_ALL) {
        const int64_t total = rows * D;
        k_rmsprop_all<DT><<<blocks_for(total, kBlock), kBlock, 0, s>>>(
            data, rms, grad, total, D, S, beta, lr, eps, minval, lr_last);
    } else if (kind == SVOX2_INDEXER_MASK) {
        SVOX2_REQUIRE(n_index == rows, "mask has %lld entries for %lld rows",
                      (long long)n_index, (long long)rows);
        const int64_t total = rows * D;
        k_rmsprop_mask<DT><<<blocks_for(total, kBlock), kBlock, 0, s>>>(
            data, rms, grad, static_cast<const uint8_t*>(indexer), total, D, S, beta, lr, eps,
            minval, lr_last);
    } else {
        const int64_t total = n_index * D;
        k_rmsprop_index<DT><<<blocks_for(total, kBlock), kBlock, 0, s>>>(
            data, rms, grad, static_cast<const int64_t*>(indexer), total, D, rows, S, beta, lr,
            eps, minval, lr_last);
    }
    return check_launch("rmsprop_step");
}

}  // namespace
}  // namespace svox2amd

using namespace svox2amd;

extern "C" int svox2_rmsprop_step(void* data, int32_t storage, float* rms, float* grad,
                                  int64_t rows, int32_t dim, int64_t data_stride,
                                  int64_t rms_stride, int64_t grad_stride, int32_t indexer_kind,
                                  const void* indexer, int64_t n_index, float beta, float lr,
                                  float epsilon, float minval, float lr_last, void* stream) {
    SVOX2_REQUIRE(rows >= 0 && dim >= 1, "bad shape rows=%lld dim=%d", (long long)rows, dim);
    SVOX2_REQUIRE(indexer_kind == SVOX2_INDEXER_ALL || indexer_kind == SVOX2_INDEXER_MASK ||
                      indexer_kind == SVOX2_INDEXER_INDEX,
                  "unknown indexer kind %d", indexer_kind);
    SVOX2_REQUIRE(storage == SVOX2_STORAGE_F32 || storage == SVOX2_STORAGE_BF16,
                  "unknown storage kind %d", storage);
    SVOX2_REQUIRE(n_index >= 0, "negative indexer length");
    if (lr_last < 0.f) lr_last = lr;
    if (rows == 0) return ok();
    if (indexer_kind != SVOX2_INDEXER_ALL && n_index == 0) return ok();  // empty indexer: skip
    SVOX2_REQUIRE(data != nullptr && rms != nullptr && grad != nullptr,
                  "data / rms / grad are NULL");
    SVOX2_REQUIRE(indexer_kind == SVOX2_INDEXER_ALL || indexer != nullptr, "indexer is NULL");
    SVOX2_REQUIRE(data_stride >= 0 && rms_stride >= 0 && grad_stride >= 0, "negative row stride");
    const Strides S{data_stride > 0 ? data_stride : dim, rms_stride > 0 ? rms_stride : dim,
                    grad_stride > 0 ? grad_stride : dim};
    SVOX2_REQUIRE(S.d >= dim && S.r >= dim && S.g >= dim, "row stride below dim");
    hipStream_t s = static_cast<hipStream_t>(stream);
    if (storage == SVOX2_STORAGE_BF16)
        return launch(static_cast<uint16_t*>(data), rms, grad, rows, dim, S, indexer_kind,
                      indexer, n_index, beta, lr, epsilon, minval, lr_last, s);
    return launch(static_cast<float*>(data), rms, grad, rows, dim, S, indexer_kind, indexer,
                  n_index, beta, lr, epsilon, minval, lr_last, s);
}
