/*
 * svox2_amd.h - C ABI of the MI355X (gfx950) sparse-voxel renderer.
 *
 * Every entry point takes plain device pointers and sizes, launches HIP
 * kernels on the caller's stream (a hipStream_t passed as void*; NULL is the
 * null stream), returns 0 on success and a nonzero status otherwise, and never
 * synchronises the host.  svox2_last_error() gives the message of the last
 * failure on the calling thread.  No entry point allocates device memory: any
 * scratch space is passed in by the caller (see *_workspace_bytes).
 *
 * Each function stands in for one binding of the reference's pybind11 module
 * svox2.csrc (svox2/csrc/svox2.cpp:11-72); the exact line is cited beside it.
 * Tensors that the reference takes as torch::Tensor become pointers here:
 *   links        int32  [size0][size1][size2], row-major, values >= 0 are rows,
 *                        < 0 are empty cells (-1 - k after accel_dist_prop)
 *   density_data rows x 1          (storage = F32 or BF16)
 *   sh_data      rows x sh_data_dim (storage = F32 or BF16), channel layout
 *                        [R0..R(B-1), G0..G(B-1), B0..B(B-1)], B = basis_dim
 *   gradients    always float32, same row count as the data
 *   rays         origins/dirs float32 [n][3]
 * Per-row arrays may be row-strided: a *_stride field gives the distance in
 * elements between consecutive rows (0 = contiguous: 1 for density, sh_data_dim
 * for SH).  The Python package stores each voxel as one 128-byte row
 * [SH(27) | density | pad] (stride 32) so a corner costs one aligned line, and
 * lays its float32 gradients out the same way.
 */
#ifndef SVOX2_AMD_H_
#define SVOX2_AMD_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SVOX2_OK 0
#define SVOX2_ERR_INVALID_ARGUMENT 1
#define SVOX2_ERR_UNSUPPORTED 2
#define SVOX2_ERR_HIP 3

#define SVOX2_STORAGE_F32 0
#define SVOX2_STORAGE_BF16 1

#define SVOX2_BASIS_TYPE_SH 1

/* rmsprop_step indexer kinds (svox2/csrc/optim_kernel.cu:175-215) */
#define SVOX2_INDEXER_ALL 0   /* every row (0-dim indexer in the reference) */
#define SVOX2_INDEXER_MASK 1  /* uint8/bool mask, one byte per row          */
#define SVOX2_INDEXER_INDEX 2 /* int64 row indices                          */

/* Grid descriptor; replaces SparseGridSpec (svox2/csrc/include/data_spec.hpp:18-52)
 * and its packed form (data_spec_packed.cuh:11-62).  World -> grid coordinates
 * are p_grid = p_world * scaling + offset, as SparseGrid._to_cpp computes them
 * (svox2/svox2.py:2135-2161). */
typedef struct svox2_grid {
    const int32_t* links;
    const void* density_data;
    const void* sh_data;
    int64_t rows;
    int32_t size[3];
    int32_t basis_dim;   /* 1, 4 or 9 */
    int32_t sh_data_dim; /* 3 * basis_dim */
    int32_t basis_type;  /* SVOX2_BASIS_TYPE_SH */
    int32_t storage;     /* SVOX2_STORAGE_* for density_data and sh_data */
    float offset[3];
    float scaling[3];
    int64_t density_stride; /* elements between density rows (0 = 1) */
    int64_t sh_stride;      /* elements between SH rows (0 = sh_data_dim) */
    /* Optional (NULL = none) cell map of these links and rows, as written by
     * svox2_build_cell_map; it must be rebuilt whenever links or rows change.
     * The march then decides skips and empty cells from it and reads the
     * links only for cells with an occupied corner.  Results are the same. */
    const uint16_t* cell_map;
    /* Optional MSI background (NULL background_links = none), the
     * background_links / background_data of SparseGridSpec
     * (data_spec.hpp:26-27; SparseGrid, svox2.py:497-520): links
     * [2 reso][reso] int32 (-1 = empty) into rows of background_data
     * [background_rows][nlayers][4] float32, channels (r, g, b, sigma), the
     * colours scaled by the SH DC factor.  nlayers >= 2.  The cuvol entry points
     * composite it behind the grid (render_lerp_kernel_cuvol.cu:386-607);
     * the nvol, depth and stats entry points ignore it, as the reference's do. */
    const int32_t* background_links;
    const float* background_data;
    int64_t background_rows;
    int32_t background_reso;
    int32_t background_nlayers;
} svox2_grid_t;

/* RenderOptions (svox2/csrc/include/data_spec.hpp:115-135). */
typedef struct svox2_render_options {
    float background_brightness;
    float step_size;
    float sigma_thresh;
    float stop_thresh;
    float near_clip;
    int32_t use_spheric_clip;
    int32_t last_sample_opaque;
} svox2_render_options_t;

/* CameraSpec (data_spec.hpp:84-102): c2w is the 3x4 OpenCV camera-to-world
 * matrix, row-major. */
typedef struct svox2_camera {
    float c2w[12];
    float fx, fy, cx, cy;
    int32_t width, height;
    float ndc_coeffx, ndc_coeffy;
} svox2_camera_t;

/* GridOutputGrads (data_spec.hpp:54-82).  Any pointer may be NULL. */
typedef struct svox2_grads {
    float* grad_density_out; /* rows x 1  */
    float* grad_sh_out;      /* rows x sh_data_dim */
    uint8_t* mask_out;       /* rows, set to 1 for every row a sample touched */
    int64_t grad_density_stride; /* elements between rows (0 = 1) */
    int64_t grad_sh_stride;      /* elements between rows (0 = sh_data_dim) */
    /* Background gradients (grids with a background): [background_rows]
     * [nlayers][4], and the mask [background_rows][nlayers] of the layers a
     * background density term touched. */
    float* grad_background_out;
    uint8_t* mask_background_out;
} svox2_grads_t;

/* Forward render of n_rays rays; rgb_out [n][3].  ray_clip, if not NULL, is
 * [n][2] (t_near, t_far) in world distance along the ray direction vector and
 * narrows each ray's [tmin, tmax].  log_transmit_out, if not NULL, receives the
 * final log transmittance per ray.  ray_cost_out, if not NULL, receives each
 * ray's march chunk count, which svox2_volume_render_cuvol_backward can use
 * to order its rays.  A grid with a background needs log_transmit_out: it
 * receives the grid's (foreground's) final log transmittance, behind which
 * the background is composited.
 * Replaces volume_render_cuvol (svox2.cpp:17, bound at :78; render_lerp_kernel_cuvol.cu:918). */
int svox2_volume_render_cuvol(const svox2_grid_t* grid, const float* origins,
                              const float* dirs, const float* ray_clip,
                              int64_t n_rays, const svox2_render_options_t* opt,
                              float* rgb_out, float* log_transmit_out,
                              uint32_t* ray_cost_out, void* stream);

/* svox2_volume_render_cuvol with an optional device workspace of
 * svox2_ray_order_workspace_bytes(n_rays) bytes: batches of at least four
 * times the rays the device holds resident (40960 on a 256-CU MI355X) are
 * processed in descending order of each ray's segment length in the grid
 * (same results); smaller batches, or a NULL workspace, in input order.
 * Replaces volume_render_cuvol (svox2.cpp:17, bound at :78; render_lerp_kernel_cuvol.cu:918)
 * for callers that can lend the workspace (the autograd forward). */
int svox2_volume_render_cuvol_ordered(const svox2_grid_t* grid, const float* origins,
                                      const float* dirs, const float* ray_clip,
                                      int64_t n_rays, const svox2_render_options_t* opt,
                                      float* rgb_out, float* log_transmit_out,
                                      uint32_t* ray_cost_out, void* workspace,
                                      size_t workspace_bytes, void* stream);

/* Forward render of a pinhole image; rgb_out [height][width][3].  A grid
 * with a background needs log_transmit_out ([height][width] floats).
 * Replaces volume_render_cuvol_image (svox2.cpp:18, bound at :79; render_lerp_kernel_cuvol.cu:960). */
int svox2_volume_render_cuvol_image(const svox2_grid_t* grid,
                                    const svox2_camera_t* cam,
                                    const svox2_render_options_t* opt,
                                    float* rgb_out, float* log_transmit_out,
                                    void* stream);

/* Expected termination distance of each ray, depth_out [n]: the sum over the
 * samples of svox2_volume_render_cuvol (same placement, skips, densities,
 * transmittance and stop) of weight * (t / step_size) * world_step, a length
 * along the unit world direction (not z-depth).  A ray that misses the grid
 * gets 0; last_sample_opaque and the background do not apply.
 * Replaces volume_render_expected_term (svox2.cpp:25, bound at :82;
 * render_lerp_kernel_cuvol.cu:1154). */
int svox2_volume_render_expected_term(const svox2_grid_t* grid, const float* origins,
                                      const float* dirs, const float* ray_clip,
                                      int64_t n_rays, const svox2_render_options_t* opt,
                                      float* depth_out, void* stream);

/* Distance (t / step_size) * world_step of the first sample along each ray
 * whose trilinear density strictly exceeds sigma_thresh (Dex-NeRF depth), or
 * 0 if none; the samples are those of svox2_volume_render_cuvol without the
 * transmittance stop.  Replaces volume_render_sigma_thresh (svox2.cpp:28,
 * bound at :83; render_lerp_kernel_cuvol.cu:1176). */
int svox2_volume_render_sigma_thresh(const svox2_grid_t* grid, const float* origins,
                                     const float* dirs, const float* ray_clip,
                                     int64_t n_rays, const svox2_render_options_t* opt,
                                     float sigma_thresh, float* depth_out, void* stream);

/* nvol backend forward: the samples of svox2_volume_render_cuvol (same
 * placement and densities, no transmittance stop, no opaque last sample)
 * composited by accumulated opacity A = min(A + 1 - exp(-ws sigma), 1) over
 * the samples with sigma > sigma_thresh, colour weight = the increase of A,
 * stopping where A reaches 1, background (1 - A) * background_brightness;
 * rgb_out [n][3].
 * Replaces volume_render_nvol (svox2.cpp:32, bound at :85;
 * render_lerp_kernel_nvol.cu:411). */
int svox2_volume_render_nvol(const svox2_grid_t* grid, const float* origins,
                             const float* dirs, const float* ray_clip,
                             int64_t n_rays, const svox2_render_options_t* opt,
                             float* rgb_out, void* stream);

/* Backward of svox2_volume_render_nvol given dL/drgb (grad_out [n][3]),
 * accumulated into grads (the reference's two passes, including their density
 * terms for samples at or below sigma_thresh before the saturating one).
 * color_cache is accepted and unread, as in the reference.
 * Replaces volume_render_nvol_backward (svox2.cpp:33, bound at :86;
 * render_lerp_kernel_nvol.cu:431). */
int svox2_volume_render_nvol_backward(const svox2_grid_t* grid, const float* origins,
                                      const float* dirs, const float* ray_clip,
                                      int64_t n_rays, const svox2_render_options_t* opt,
                                      const float* grad_out, const float* color_cache,
                                      const svox2_grads_t* grads, void* stream);

/* svox2_volume_render_nvol into rgb_out, then its backward for the MSE loss
 * against rgb_gt (dL/drgb = (rgb - rgb_gt) * 2 / (3 n)) plus the Cauchy
 * sparsity term; beta_loss is accepted and unused, as in the reference.
 * Replaces volume_render_nvol_fused (svox2.cpp:35, bound at :87;
 * render_lerp_kernel_nvol.cu:461). */
int svox2_volume_render_nvol_fused(const svox2_grid_t* grid, const float* origins,
                                   const float* dirs, const float* ray_clip,
                                   int64_t n_rays, const svox2_render_options_t* opt,
                                   const float* rgb_gt, float beta_loss,
                                   float sparsity_loss, float* rgb_out,
                                   const svox2_grads_t* grads, void* stream);

/* svox1 backend forward: nearest-neighbour rendering by a walk over whole
 * cells (the grid origin shifted by +0.5; each occupied cell one sample of
 * its own density and SH colour, weighted by its chord plus 1e-2; empty
 * blocks of the accelerated links skipped); a ray whose transmittance before
 * a sample is <= stop_thresh stops after it with its colour divided by 1 -
 * that transmittance.  rgb_out [n][3].  ray_clip must be NULL (the reference
 * takes none); a background is ignored.
 * Replaces volume_render_svox1 (svox2.cpp:39, bound at :89;
 * render_svox1_kernel.cu:334). */
int svox2_volume_render_svox1(const svox2_grid_t* grid, const float* origins,
                              const float* dirs, const float* ray_clip, int64_t n_rays,
                              const svox2_render_options_t* opt, float* rgb_out, void* stream);

/* svox1 backward given dL/drgb (grad_out) and the forward colours
 * (color_cache): SH and density gradients of every sample up to the one
 * after which the transmittance is <= stop_thresh, as the reference computes
 * them; grads->mask_out marks the rows of those samples.
 * Replaces volume_render_svox1_backward (svox2.cpp:40, bound at :90;
 * render_svox1_kernel.cu:351). */
int svox2_volume_render_svox1_backward(const svox2_grid_t* grid, const float* origins,
                                       const float* dirs, const float* ray_clip,
                                       int64_t n_rays, const svox2_render_options_t* opt,
                                       const float* grad_out, const float* color_cache,
                                       const svox2_grads_t* grads, void* stream);

/* svox1 forward into rgb_out, then its backward for dL/drgb = (rgb_out -
 * rgb_gt) * 2 / (3 n); beta_loss and sparsity_loss are ignored, as in the
 * reference.
 * Replaces volume_render_svox1_fused (svox2.cpp:42, bound at :91;
 * render_svox1_kernel.cu:380). */
int svox2_volume_render_svox1_fused(const svox2_grid_t* grid, const float* origins,
                                    const float* dirs, const float* ray_clip, int64_t n_rays,
                                    const svox2_render_options_t* opt, const float* rgb_gt,
                                    float beta_loss, float sparsity_loss, float* rgb_out,
                                    const svox2_grads_t* grads, void* stream);

/* Backward of svox2_volume_render_cuvol given dL/drgb (grad_out [n][3]) and
 * the forward colours (color_cache [n][3]); accumulates into grads.  With the
 * forward's ray_cost_out (ray_cost) and a device workspace of
 * svox2_ray_order_workspace_bytes(n_rays) bytes, batches of at least four
 * times the rays the device holds resident are processed in descending order
 * of cost (same results); otherwise, or with NULLs, in input order.  A grid
 * with a background needs a workspace of svox2_background_workspace_bytes(
 * n_rays) bytes (it includes the ordering workspace): the call renders the
 * grid's own colours and transmittance into it, which the background's
 * gradients start from.
 * Replaces volume_render_cuvol_backward (svox2.cpp:20, bound at :80; render_lerp_kernel_cuvol.cu:1010). */
int svox2_volume_render_cuvol_backward(const svox2_grid_t* grid,
                                       const float* origins, const float* dirs,
                                       const float* ray_clip, int64_t n_rays,
                                       const svox2_render_options_t* opt,
                                       const float* grad_out,
                                       const float* color_cache,
                                       const svox2_grads_t* grads,
                                       const uint32_t* ray_cost,
                                       void* workspace, size_t workspace_bytes,
                                       void* stream);

/* Forward render plus the gradient of mean((rgb - rgb_gt)^2) (over 3n terms),
 * with optional beta and Cauchy sparsity terms: a forward launch writing
 * rgb_out, then a backward launch that forms the MSE gradient from it and
 * accumulates into grads.  log_transmit_out (n floats, may be NULL unless
 * beta_loss > 0) receives each ray's final log transmittance, which the beta
 * term needs.  The optional device workspace (results are the same with or
 * without it):
 * - at least svox2_ray_order_workspace_bytes(n_rays) bytes: for batches of at
 *   least four times the rays the device holds resident (40960 on a 256-CU
 *   MI355X) the forward processes the rays in descending order of their
 *   segment length in the grid, records each ray's chunk count, and the
 *   backward processes the rays in descending order of that count;
 * - at least svox2_volume_render_cuvol_fused_workspace_bytes(n_rays) bytes:
 *   for batches below that threshold the forward writes every contributing
 *   sample (position, corner links, colour, transmittance) to a record
 *   stream in the workspace and the backward replays it instead of
 *   re-marching the ray (a ray whose records do not fit is re-marched);
 *   for batches at or above it the forward also logs where it placed the
 *   samples of each ray's first 16 chunks, and the backward evaluates those
 *   chunks from the log instead of placing them again.
 * A grid with a background needs a workspace of at least
 * svox2_background_workspace_bytes(n_rays) bytes (the fused workspace size
 * is never less); the call then re-marches every ray (no records, no log).
 * Replaces volume_render_cuvol_fused (svox2.cpp:22, bound at :81; render_lerp_kernel_cuvol.cu:1070). */
int svox2_volume_render_cuvol_fused(const svox2_grid_t* grid,
                                    const float* origins, const float* dirs,
                                    const float* ray_clip, int64_t n_rays,
                                    const svox2_render_options_t* opt,
                                    const float* rgb_gt, float beta_loss,
                                    float sparsity_loss, float* rgb_out,
                                    float* log_transmit_out,
                                    const svox2_grads_t* grads,
                                    void* workspace, size_t workspace_bytes,
                                    void* stream);

/* Device workspace bytes svox2_volume_render_cuvol_fused and
 * svox2_volume_render_cuvol_backward use to order the rays of an n_rays batch
 * (0 for an empty batch).  No counterpart in the reference. */
size_t svox2_ray_order_workspace_bytes(int64_t n_rays);

/* Device workspace bytes with which svox2_volume_render_cuvol_fused runs
 * fastest for an n_rays batch (0 for an empty batch): the record stream
 * and chunk list (about 4.8 KiB per ray) up to twice the ordering threshold
 * (81920 rays on a 256-CU MI355X), the ray-ordering workspace plus the chunk
 * log (about 2.3 KiB per ray) above it.  No
 * counterpart in the reference. */
size_t svox2_volume_render_cuvol_fused_workspace_bytes(int64_t n_rays);

/* Device workspace bytes svox2_volume_render_cuvol_backward and
 * svox2_volume_render_cuvol_fused need for an n_rays batch over a grid with
 * a background: the ray-ordering workspace, then 16 bytes per ray (the
 * grid's own colours or the background's, and log transmittance).  No
 * counterpart in the reference (it allocates these per call). */
size_t svox2_background_workspace_bytes(int64_t n_rays);

/* svox2_volume_render_cuvol_fused without a caller workspace and without any
 * device allocation: the workspace is a fixed 40 MiB device array inside the
 * library (one copy per device, looked up with hipGetSymbolAddress on the
 * device's first call).  Same arguments, results and errors as the fused
 * call, minus workspace / workspace_bytes.  Batches whose
 * svox2_volume_render_cuvol_fused_workspace_bytes exceeds 40 MiB (above 8474
 * rays; 8192 rays always fit) fail with SVOX2_ERR_INVALID_ARGUMENT.
 * Calls that overlap on one device share the array and race: issue them on
 * one stream.  No counterpart in the reference: for callers that cannot
 * allocate device memory (code under graph capture, allocation-free serving
 * or training loops). */
int svox2_volume_render_cuvol_fused_noalloc(const svox2_grid_t* grid,
                                            const float* origins, const float* dirs,
                                            const float* ray_clip, int64_t n_rays,
                                            const svox2_render_options_t* opt,
                                            const float* rgb_gt, float beta_loss,
                                            float sparsity_loss, float* rgb_out,
                                            float* log_transmit_out,
                                            const svox2_grads_t* grads, void* stream);

/* One tile of a multi-request render launch: n_pixels consecutive pixels
 * (row-major, starting at pixel0) of the image that camera `camera` sees of
 * grid `grid`, written to rgb_out[pixel][3] (the request's whole image). */
typedef struct svox2_tile {
    int32_t grid;     /* index into the grid array */
    int32_t camera;   /* index into the camera array */
    int32_t pixel0;
    int32_t n_pixels; /* <= tile_pixels */
    float* rgb_out;
} svox2_tile_t;

/* Renders the tiles of many image requests in one launch: grids
 * [n_grids] (device memory, validated descriptors of one storage kind and
 * basis_dim, with their cell maps built), cameras [n_cameras] and tiles
 * [max_tiles] in device memory, of which the first *n_tiles (a device int,
 * read by the kernel) are rendered.  Every pixel is computed exactly as
 * svox2_volume_render_cuvol_image computes it.  Entries naming a grid or
 * camera out of range are skipped.  The launch shape depends only on
 * max_tiles and tile_pixels, so a captured graph can be replayed with new
 * table contents.  No counterpart in the reference: the serving path. */
int svox2_volume_render_cuvol_tiles(const svox2_grid_t* grids, int32_t n_grids,
                                    int32_t storage, int32_t basis_dim,
                                    const svox2_camera_t* cameras, int32_t n_cameras,
                                    const svox2_render_options_t* opt,
                                    const svox2_tile_t* tiles, const int32_t* n_tiles,
                                    int32_t max_tiles, int32_t tile_pixels, void* stream);

/* Per-ray march counters [n][4] (int32): loop iterations, iterations not
 * skipped by the empty-space skip, samples with sigma > sigma_thresh, and a
 * FNV-1a hash of every visited cell, sample t and contribution flag plus the
 * final log transmittance.  No counterpart in the reference: it measures the
 * work of svox2_volume_render_cuvol on the same inputs and pins its sample
 * indexing bit for bit. */
int svox2_volume_render_cuvol_stats(const svox2_grid_t* grid,
                                    const float* origins, const float* dirs,
                                    const float* ray_clip, int64_t n_rays,
                                    const svox2_render_options_t* opt,
                                    int32_t* stats_out, void* stream);

/* Trilinear samples of density [n][1] and, if want_colors, sh [n][sh_data_dim]
 * at points (world coordinates through grid->offset/scaling).
 * Replaces sample_grid (svox2.cpp:11, bound at :76; svox2_kernel.cu:181). */
int svox2_sample_grid(const svox2_grid_t* grid, const float* points,
                      int64_t n_points, int32_t want_colors,
                      float* density_out, float* sh_out, void* stream);

/* Adjoint of svox2_sample_grid; accumulates into grad_density_out / grad_sh_out
 * (row strides in elements, 0 = contiguous).
 * Replaces sample_grid_backward (svox2.cpp:13, bound at :77; svox2_kernel.cu:219). */
int svox2_sample_grid_backward(const svox2_grid_t* grid, const float* points,
                               int64_t n_points,
                               const float* grad_out_density,
                               const float* grad_out_sh,
                               float* grad_density_out, float* grad_sh_out,
                               int64_t grad_density_stride,
                               int64_t grad_sh_stride,
                               int32_t want_colors, void* stream);

/* Sparse total-variation gradient over n_cells flat cell indices (rand_cells),
 * channels [start_dim, end_dim) of data (rows x data_dim, consecutive rows
 * data_stride elements apart; grad_data rows grad_stride apart; 0 =
 * contiguous), scaled by
 * scale / n_cells and accumulated into grad_data; mask_out (NULL allowed) is
 * set for every row that receives a gradient.
 * Replaces tv_grad_sparse (svox2.cpp:63, bound at :99; loss_kernel.cu:558). */
int svox2_tv_grad_sparse(const int32_t* links, const int32_t size[3],
                         const void* data, int32_t storage, int64_t rows,
                         int32_t data_dim, int64_t data_stride,
                         int64_t grad_stride, const int32_t* rand_cells,
                         int64_t n_cells, uint8_t* mask_out, int32_t start_dim,
                         int32_t end_dim, float scale, int32_t use_logalpha,
                         float logalpha_delta, int32_t ignore_edge,
                         int32_t ignore_last_z, float ndc_coeffx,
                         float ndc_coeffy, float* grad_data, void* stream);

/* svox2_tv_grad_sparse over the contiguous run of n_cells cells starting at
 * flat cell first_cell and wrapping past the last cell (the cells the
 * reference's contiguous sparse regulariser draws, svox2.py:2224-2241),
 * generated in the kernel instead of read from a cell list.  Requires
 * 0 <= first_cell < cells and n_cells <= cells.  No counterpart in the
 * reference, which materialises the run as a tensor. */
int svox2_tv_grad_sparse_run(const int32_t* links, const int32_t size[3],
                             const void* data, int32_t storage, int64_t rows,
                             int32_t data_dim, int64_t data_stride,
                             int64_t grad_stride, int64_t first_cell,
                             int64_t n_cells, uint8_t* mask_out, int32_t start_dim,
                             int32_t end_dim, float scale, int32_t ignore_edge,
                             int32_t ignore_last_z, float* grad_data, void* stream);

/* Lumisphere total variation of the SH colour: for each of the n_cells flat
 * cell indices (split by the extents size - 1) the colours A = sum_k
 * sh[c * basis_dim + k] * basis_fn[k] of the cell and its +x/+y/+z neighbours
 * and U = the same with basis_fn_u (basis_dim device floats each) give
 * per-channel differences (axis factor size / 256, dir_factor for U - A), whose
 * normalised gradient, scaled by scale / n_cells, is accumulated into
 * grads->grad_sh_out (grads->grad_sh_stride apart); grads->mask_out (NULL
 * allowed) marks every row that receives a term.  Only the first 32
 * coefficients of a row take part, as in the reference; a cell index whose x
 * part is >= size0 - 1 is skipped (the reference would read past the links).
 * ndc_coeffx / ndc_coeffy are accepted and unused, as in the reference.
 * Replaces lumisphere_tv_grad_sparse (svox2.cpp:66, bound at :101;
 * loss_kernel.cu:643). */
int svox2_lumisphere_tv_grad_sparse(const svox2_grid_t* grid,
                                    const int32_t* rand_cells, int64_t n_cells,
                                    const float* basis_fn, const float* basis_fn_u,
                                    float scale, float ndc_coeffx, float ndc_coeffy,
                                    float dir_factor, const svox2_grads_t* grads,
                                    void* stream);

/* Total variation of an MSI background: for each of the n_cells flat indices
 * ((x * reso + y) * nlayers + z over the [2 reso][reso] links and nlayers
 * layers) and each of the nchannels channels, the differences to the cell's
 * +x and +y neighbours (wrapping around) and to layer z + 1 (a missing layer
 * counts as the cell's own value for colours and as 0 for the last channel,
 * sigma), their normalised gradient (axis factors 2 reso / 256, reso / 256,
 * nlayers / 256; scale, or scale_last for the last channel, over n_cells)
 * accumulated into grad_msi [msi_rows][nlayers][nchannels]; mask_out (NULL
 * allowed, [msi_rows][nlayers]) marks every layer that receives a nonzero
 * term.  Links < 0 read as 0 and receive nothing.
 * Replaces msi_tv_grad_sparse (svox2.cpp:65, bound at :100;
 * loss_kernel.cu:606). */
int svox2_msi_tv_grad_sparse(const int32_t* links, int32_t reso, const float* msi,
                             int64_t msi_rows, int32_t nlayers, int32_t nchannels,
                             const int32_t* rand_cells, int64_t n_cells, uint8_t* mask_out,
                             float scale, float scale_last, float* grad_msi, void* stream);

/* RMSprop update of data (rows x dim, float32 or bf16 storage) with float32
 * rms and grad; grad is zeroed where updated.  Row strides (elements, 0 =
 * dim) apply to data, rms and grad separately.  indexer_kind selects every row,
 * a byte mask (n_index = rows) or int64 row indices (n_index entries; entries
 * outside [0, rows) are skipped).
 * Replaces rmsprop_step (svox2.cpp:70, bound at :110; optim_kernel.cu:154). */
int svox2_rmsprop_step(void* data, int32_t storage, float* rms, float* grad,
                       int64_t rows, int32_t dim, int64_t data_stride,
                       int64_t rms_stride, int64_t grad_stride,
                       int32_t indexer_kind,
                       const void* indexer, int64_t n_index, float beta,
                       float lr, float epsilon, float minval, float lr_last,
                       void* stream);

/* SGD update data -= lr * grad (lr_last for the last column), grad zeroed
 * where applied, over the rows the indexer selects (as svox2_rmsprop_step).
 * Replaces sgd_step (svox2.cpp:72, bound at :111; optim_kernel.cu:218). */
int svox2_sgd_step(void* data, int32_t storage, float* grad, int64_t rows,
                   int32_t dim, int64_t data_stride, int64_t grad_stride,
                   int32_t indexer_kind, const void* indexer, int64_t n_index,
                   float lr, float lr_last, void* stream);

/* RMSprop over rows of ncols columns split into two groups, [0, split) and
 * [split, ncols), each with its own byte mask (one per row), beta and
 * learning rate; both masks NULL updates every row.  With the interleaved
 * voxel row [SH | density] (ncols = sh_data_dim + 1, split = sh_data_dim)
 * this is optim_sh_step followed by optim_density_step in one pass over the
 * rows.  No counterpart in the reference: it fuses the two rmsprop_step
 * calls of SparseGrid.optim_sh_step / optim_density_step (svox2.py:1972-2041). */
int svox2_rmsprop_step_grouped(void* data, int32_t storage, float* rms, float* grad,
                               int64_t rows, int32_t ncols, int64_t data_stride,
                               int64_t rms_stride, int64_t grad_stride, int32_t split,
                               const uint8_t* mask_a, const uint8_t* mask_b,
                               float beta_a, float lr_a, float beta_b, float lr_b,
                               float epsilon, float minval, void* stream);

/* Bytes of scratch svox2_accel_dist_prop needs for a links table of this size. */
size_t svox2_accel_dist_prop_workspace_bytes(const int32_t size[3]);

/* Rewrites every empty link in place to -1 - k, k = number of consecutive
 * empty ancestor levels of the occupancy pyramid.
 * Replaces accel_dist_prop (svox2.cpp:54, bound at :105; misc_kernel.cu:356). */
int svox2_accel_dist_prop(int32_t* links, const int32_t size[3],
                          void* workspace, size_t workspace_bytes,
                          void* stream);

/* ---- Grid resampling (SparseGrid.resample, svox2.py:1223-1424) ----------- */

/* Density of every cell of a new reso[0] x reso[1] x reso[2] grid, sampled
 * trilinearly from `grid` (as sample_grid with grid coordinates) at the old
 * grid's coordinates (axes[ix], axes[reso0 + iy], axes[reso0 + reso1 + iz]);
 * density_out is the dense raster-order volume.  No counterpart of its own in
 * the reference: it is resample's first sample() pass (svox2.py:1298-1307). */
int svox2_resample_density(const svox2_grid_t* grid, const float* axes,
                           const int32_t reso[3], float* density_out, void* stream);

/* Rows of the new grid: for row r, SH channels sampled from `grid` at kept
 * cell kept_cells[r]'s point and the density density_volume[kept_cells[r]],
 * written in out_storage (SVOX2_STORAGE_*) with row strides in elements.
 * Resample's second sample() pass (svox2.py:1383-1392). */
int svox2_resample_rows(const svox2_grid_t* grid, const float* axes, const int32_t reso[3],
                        const int32_t* kept_cells, int64_t n_rows,
                        const float* density_volume, void* sh_out, int64_t sh_stride,
                        void* density_out, int64_t density_stride, int32_t out_storage,
                        void* stream);

/* Raises grid_weight_out (a dense float volume of data's size, holding
 * non-negative values) to the maximum compositing weight of any sample of the
 * camera's pixel rays through the dense density volume `data`, at the 8
 * corners of the sample's cell; offset / scaling map world to volume
 * coordinates.  Replaces grid_weight_render (svox2.cpp:55, bound at :106;
 * misc_kernel.cu:418). */
int svox2_grid_weight_render(const float* data, const int32_t size[3],
                             const svox2_camera_t* cam, float step_size,
                             float stop_thresh, int32_t last_sample_opaque,
                             const float offset[3], const float scaling[3],
                             float* grid_weight_out, void* stream);

/* mask_out[i] = values[i] >= thresh, or 1 on the last z slice when
 * keep_last_z.  Resample's thresholding (svox2.py:1349-1371). */
int svox2_threshold_mask(const float* values, const int32_t size[3], float thresh,
                         int32_t keep_last_z, uint8_t* mask_out, void* stream);

/* One 3x3x3 dilation of a byte mask (nonzero = set), out of place.
 * Replaces dilate (svox2.cpp:53, bound at :104; misc_kernel.cu:339). */
int svox2_dilate(const uint8_t* mask, const int32_t size[3], uint8_t* out, void* stream);

/* Workspace bytes of the mask compaction for a volume of this size. */
size_t svox2_mask_compact_workspace_bytes(const int32_t size[3]);

/* Compaction of a byte mask into links, pass 1: counts the set cells in order
 * (raster, or Morton when z_order: cubic power-of-two side <= 1024) into the
 * workspace; the number of set cells is then the uint32 at word
 * ceil(cells / 4096) of the workspace.  Pass 2 (svox2_mask_compact_assign)
 * gives the k-th set cell in that order link k, every other cell -1, and
 * (if kept_cells_out is not NULL) kept_cells_out[k] = its flat cell index.
 * Resample's link numbering (svox2.py:1397-1412). */
int svox2_mask_compact_count(const uint8_t* mask, const int32_t size[3], int32_t z_order,
                             void* workspace, size_t workspace_bytes, void* stream);
int svox2_mask_compact_assign(const uint8_t* mask, const int32_t size[3], int32_t z_order,
                              const void* workspace, int32_t* links_out,
                              int32_t* kept_cells_out, void* stream);

/* Bytes of the cell map of a links table of this size: one 16-bit entry per
 * 2x2x2 block of cells, ((size[i] - 2) / 2 + 1) blocks per axis.  No
 * counterpart in the reference. */
size_t svox2_cell_map_bytes(const int32_t size[3]);

/* Writes the cell map of links (rows data rows): per block of cells with
 * lower corners (2bx..2bx+1, 2by..2by+1, 2bz..2bz+1), bits 0-7 hold the
 * empty-space skip level k shared by the block (every lower-corner link is
 * -1 - k, k >= 1), 0 if no lower-corner link is below -1, or 255 if the
 * block mixes them; bit 8 + 4*(x&1) + 2*(y&1) + (z&1) is set iff the cell
 * with lower corner (x, y, z) has a corner link in [0, rows).  No
 * counterpart in the reference: an acceleration structure for
 * svox2_volume_render_cuvol and the calls built on it. */
int svox2_build_cell_map(const int32_t* links, const int32_t size[3], int64_t rows,
                         uint16_t* map_out, size_t map_bytes, void* stream);

/* Message of the last failed call on this thread ("" if none). */
const char* svox2_last_error(void);

/* Library version string. */
const char* svox2_version(void);

#ifdef __cplusplus
}
#endif

#endif /* SVOX2_AMD_H_ */
