// Host-side pieces shared by every entry point: the per-thread last error,
// grid descriptor validation (the checks SparseGridSpec::check makes in
// svox2/csrc/include/data_spec.hpp:32-51, restated over raw pointers), and the
// version string.
#include "common.hpp"

#include <climits>

namespace svox2amd {

std::string& last_error_slot() {
    static thread_local std::string slot;
    return slot;
}

int check_grid(const svox2_grid_t* g, bool need_sh) {
    SVOX2_REQUIRE(g != nullptr, "grid is NULL");
    SVOX2_REQUIRE(g->links != nullptr, "grid.links is NULL");
    for (int i = 0; i < 3; ++i)
        SVOX2_REQUIRE(g->size[i] >= 2, "grid size[%d] = %d, must be >= 2", i, g->size[i]);
    const int64_t cells = static_cast<int64_t>(g->size[0]) * g->size[1] * g->size[2];
    SVOX2_REQUIRE(cells < INT_MAX, "grid has %lld cells; int32 indexing allows < 2^31",
                  (long long)cells);
    SVOX2_REQUIRE(g->rows >= 0 && g->rows < INT_MAX, "grid rows %lld out of range",
                  (long long)g->rows);
    SVOX2_REQUIRE(g->storage == SVOX2_STORAGE_F32 || g->storage == SVOX2_STORAGE_BF16,
                  "unknown storage kind %d", g->storage);
    SVOX2_REQUIRE(g->basis_type == SVOX2_BASIS_TYPE_SH,
                  "basis_type %d not supported (only SH = 1)", g->basis_type);
    SVOX2_REQUIRE(g->basis_dim == 1 || g->basis_dim == 4 || g->basis_dim == 9,
                  "basis_dim %d not supported (1, 4 or 9)", g->basis_dim);
    SVOX2_REQUIRE(g->sh_data_dim == 3 * g->basis_dim,
                  "sh_data_dim %d != 3 * basis_dim %d", g->sh_data_dim, g->basis_dim);
    SVOX2_REQUIRE(g->density_stride >= 0 && g->density_stride < INT_MAX,
                  "density_stride %lld out of range", (long long)g->density_stride);
    SVOX2_REQUIRE(g->sh_stride == 0 || (g->sh_stride >= g->sh_data_dim && g->sh_stride < INT_MAX),
                  "sh_stride %lld must be >= sh_data_dim %d", (long long)g->sh_stride,
                  g->sh_data_dim);
    const int64_t ss = g->sh_stride > 0 ? g->sh_stride : g->sh_data_dim;
    const int64_t ds = g->density_stride > 0 ? g->density_stride : 1;
    SVOX2_REQUIRE(g->rows * ss < (INT64_C(1) << 40) && g->rows * ds < (INT64_C(1) << 40),
                  "grid data too large");
    SVOX2_REQUIRE(g->rows == 0 || g->density_data != nullptr, "grid.density_data is NULL");
    if (need_sh) SVOX2_REQUIRE(g->rows == 0 || g->sh_data != nullptr, "grid.sh_data is NULL");
    return ok();
}

}  // namespace svox2amd

extern "C" const char* svox2_last_error(void) {
    return svox2amd::last_error_slot().c_str();
}

extern "C" const char* svox2_version(void) {
    return "svox2_amd 0.1.0 (gfx950)";
}
