// Sparse total-variation gradient for gfx950.
//
// One lane per (cell, channel), channels fastest: a 64-lane wavefront covers
// ~2.4 cells of SH-2 data, so link reads of consecutive cells and the data and
// gradient rows of each cell are touched by adjacent lanes.
//
// Semantics of tv_grad_sparse_kernel (svox2/csrc/loss_kernel.cu:178-246) and
// its launcher (:558-604), including the reference's edge behaviour: an
// out-of-range neighbour reads link 0 (data row 0) and receives gradient,
// ignore_edge skips cells whose link is 0, and ignore_last_z drops the z+1
// neighbour of the second-to-last z slice.
#include "common.hpp"

namespace svox2amd {
namespace {

constexpr int kBlock = 256;

template <typename DT>
__global__ __launch_bounds__(kBlock) void k_tv_grad_sparse(
        const int32_t* __restrict__ links, int s0, int s1, int s2,
        const DT* __restrict__ data, int rows, int64_t dst, int64_t gst,
        const int32_t* __restrict__ cells,
        int64_t total, int start_dim, int nch, float scale, int ignore_edge,
        int ignore_last_z, uint8_t* __restrict__ mask, float* __restrict__ grad) {
    const int64_t tid = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x;
    if (tid >= total) return;
    const int idx = static_cast<int>(tid % nch) + start_dim;
    const int xyz = cells[tid / nch];
    const int z = xyz % s2;
    const int xy = xyz / s2;
    const int y = xy % s1;
    const int x = xy / s1;
    if (x >= s0 || xyz < 0) return;  // cell index outside the grid
    const int offx = s1 * s2, offy = s2;
    const int32_t* lp = links + (static_cast<int64_t>(x) * offx + y * offy + z);
    if (ignore_edge && lp[0] == 0) return;

    const float sc0 = static_cast<float>(s0) * (1.f / 256.f);
    const float sc1 = static_cast<float>(s1) * (1.f / 256.f);
    const float sc2 = static_cast<float>(s2) * (1.f / 256.f);

    const int l000 = lp[0];
    const int l001 = ((z + 1 < s2) && (!ignore_last_z || z != s2 - 2)) ? lp[1] : 0;
    const int l010 = (y + 1 < s1) ? lp[offy] : 0;
    const int l100 = (x + 1 < s0) ? lp[offx] : 0;
    if (ignore_last_z && z == s2 - 2) return;

    auto ok = [&](int l) { return static_cast<uint32_t>(l) < static_cast<uint32_t>(rows); };
    const float v000 = ok(l000) ? ld(data, static_cast<int64_t>(l000) * dst + idx) : 0.f;
    const float nul = ignore_edge ? v000 : 0.f;
    const float v001 = ok(l001) ? ld(data, static_cast<int64_t>(l001) * dst + idx) : nul;
    const float v010 = ok(l010) ? ld(data, static_cast<int64_t>(l010) * dst + idx) : nul;
    const float v100 = ok(l100) ? ld(data, static_cast<int64_t>(l100) * dst + idx) : nul;

    float dx = v100 - v000;
    float dy = v010 - v000;
    float dz = v001 - v000;
    const float idelta = scale * rsqrtf(1e-9f + dx * dx + dy * dy + dz * dz);
    dx *= sc0;
    dy *= sc1;
    dz *= sc2;
    const float sm = -(dx + dy + dz);
    auto add = [&](int l, float val) {
        if (ok(l) && val != 0.f) {
            atomic_add_f32(grad + static_cast<int64_t>(l) * gst + idx, val * idelta);
            if (mask != nullptr) mask[l] = 1;
        }
    };
    add(l000, sm);
    add(l001, dz);
    add(l010, dy);
    add(l100, dx);
}

}  // namespace
}  // namespace svox2amd

using namespace svox2amd;

extern "C" int svox2_tv_grad_sparse(const int32_t* links, const int32_t size[3], const void* data,
                                    int32_t storage, int64_t rows, int32_t data_dim,
                                    int64_t data_stride, int64_t grad_stride,
                                    const int32_t* rand_cells, int64_t n_cells,
                                    uint8_t* mask_out, int32_t start_dim, int32_t end_dim,
                                    float scale, int32_t use_logalpha, float logalpha_delta,
                                    int32_t ignore_edge, int32_t ignore_last_z,
                                    float ndc_coeffx, float ndc_coeffy, float* grad_data,
                                    void* stream) {
    (void)logalpha_delta;
    (void)ndc_coeffx;
    (void)ndc_coeffy;
    SVOX2_REQUIRE(size != nullptr && links != nullptr, "links / size are NULL");
    for (int i = 0; i < 3; ++i) SVOX2_REQUIRE(size[i] >= 1, "grid size[%d] = %d", i, size[i]);
    const int64_t cells = static_cast<int64_t>(size[0]) * size[1] * size[2];
    SVOX2_REQUIRE(cells < INT32_MAX, "too many cells");
    SVOX2_REQUIRE(!use_logalpha, "use_logalpha is no longer supported (as in the reference)");
    SVOX2_REQUIRE(storage == SVOX2_STORAGE_F32 || storage == SVOX2_STORAGE_BF16,
                  "unknown storage kind %d", storage);
    SVOX2_REQUIRE(rows >= 0 && rows < INT32_MAX, "rows out of range");
    SVOX2_REQUIRE(data_dim >= 1, "data_dim must be >= 1");
    SVOX2_REQUIRE(0 <= start_dim && start_dim < end_dim && end_dim <= data_dim,
                  "channel range [%d, %d) invalid for %d channels", start_dim, end_dim, data_dim);
    SVOX2_REQUIRE(n_cells >= 0, "negative cell count");
    if (n_cells == 0 || rows == 0) return ok();
    SVOX2_REQUIRE(data != nullptr && rand_cells != nullptr && grad_data != nullptr,
                  "data / rand_cells / grad_data are NULL");
    SVOX2_REQUIRE(data_stride >= 0 && grad_stride >= 0, "negative row stride");
    const int64_t dst = data_stride > 0 ? data_stride : data_dim;
    const int64_t gst = grad_stride > 0 ? grad_stride : data_dim;
    SVOX2_REQUIRE(dst >= data_dim && gst >= data_dim, "row stride below data_dim");
    const int nch = end_dim - start_dim;
    const int64_t total = n_cells * nch;
    // The reference divides the scale by the number of cells (loss_kernel.cu:597).
    const float sc = scale / static_cast<float>(n_cells);
    hipStream_t s = static_cast<hipStream_t>(stream);
    const unsigned blocks = blocks_for(total, kBlock);
    if (storage == SVOX2_STORAGE_BF16) {
        k_tv_grad_sparse<uint16_t><<<blocks, kBlock, 0, s>>>(
            links, size[0], size[1], size[2], static_cast<const uint16_t*>(data),
            static_cast<int>(rows), dst, gst, rand_cells, total, start_dim, nch, sc, ignore_edge,
            ignore_last_z, mask_out, grad_data);
    } else {
        k_tv_grad_sparse<float><<<blocks, kBlock, 0, s>>>(
            links, size[0], size[1], size[2], static_cast<const float*>(data),
            static_cast<int>(rows), dst, gst, rand_cells, total, start_dim, nch, sc, ignore_edge,
            ignore_last_z, mask_out, grad_data);
    }
    return check_launch("tv_grad_sparse");
}
