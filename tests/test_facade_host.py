"""Host-side logic of the SparseGrid facade (no kernel launches): options,
rays, cameras, grid construction and layouts, coordinate maps, random TV
cells, indexer conversion, checkpoints, and that every compute method refuses
to run without a GPU."""
import numpy as np
import pytest
import torch

import svox2_amd
from svox2_amd import scenes, utils


def test_render_options_defaults_match_reference():
    o = svox2_amd.RenderOptions()
    # svox2/svox2.py:31-55
    assert (o.backend, o.background_brightness, o.step_size, o.sigma_thresh, o.stop_thresh,
            o.last_sample_opaque, o.near_clip, o.use_spheric_clip) == \
        ("cuvol", 1.0, 0.5, 1e-10, 1e-7, False, 0.0, False)
    c = o._to_cpp()
    assert c.step_size == 0.5 and c.stop_thresh == 1e-7


def test_rays_and_camera():
    r = svox2_amd.Rays(torch.zeros(10, 3), torch.ones(10, 3), torch.zeros(10, 2))
    s = r[2:5]
    assert s.origins.shape == (3, 3) and s.clip.shape == (3, 2) and not r.is_cuda
    cam = svox2_amd.Camera(torch.eye(4)[:3], fx=100.0, width=8, height=6)
    assert cam.fy_val == 100.0 and cam.cx_val == 4.0 and cam.cy_val == 3.0 and not cam.using_ndc
    rays = cam.gen_rays()
    assert rays.origins.shape == (48, 3)
    # centre-ish pixel looks along +z (identity rotation)
    d = rays.dirs.view(6, 8, 3)
    assert torch.allclose(d.norm(dim=-1), torch.ones(6, 8), atol=1e-6)
    assert d[3, 4, 2] > 0.99


def test_grid_construction_layouts():
    g = svox2_amd.SparseGrid(reso=8, use_z_order=True)
    assert g.links.shape == (8, 8, 8) and g.sh_data.shape == (512, 27)
    assert torch.equal(g.links, utils.gen_morton(8, dtype=torch.int32))
    g2 = svox2_amd.SparseGrid(reso=[4, 6, 5])
    assert torch.equal(g2.links.flatten(), torch.arange(120, dtype=torch.int32))
    g3 = svox2_amd.SparseGrid(reso=16, use_sphere_bound=True)
    valid = g3.links[g3.links >= 0]
    assert g3.capacity == valid.numel() < 16 ** 3
    assert torch.equal(torch.sort(valid).values, torch.arange(g3.capacity, dtype=torch.int32))
    gb = svox2_amd.SparseGrid(reso=4, dtype=torch.bfloat16)
    assert gb.sh_data.dtype == torch.bfloat16 and gb.density_data.grad_dtype == torch.float32
    with pytest.raises(NotImplementedError):
        svox2_amd.SparseGrid(reso=4, basis_type=svox2_amd.BASIS_TYPE_MLP)
    with pytest.raises(ValueError):
        svox2_amd.SparseGrid(reso=4, background_nlayers=1)
    # An MSI background: one row per pixel column in raster order, zeroed data
    # (svox2.py:499-513).
    gbg = svox2_amd.SparseGrid(reso=4, background_nlayers=3, background_reso=5)
    assert gbg.use_background and gbg.background_links.shape == (10, 5)
    assert torch.equal(gbg.background_links.flatten(), torch.arange(50, dtype=torch.int32))
    assert gbg.background_data.shape == (50, 3, 4) and not gbg.background_data.any()
    assert not svox2_amd.SparseGrid(reso=4).use_background


@pytest.mark.parametrize("size,z_order", [((8, 8, 8), False), ((8, 8, 8), True),
                                          ((16, 16, 16), True), ((7, 10, 5), False)])
def test_sphere_bound_numbering_closed_form(size, z_order):
    """use_sphere_bound keeps the cells whose lower corner (2 i / n - 1 per
    axis) is within 1 + sqrt(3) / max(n) of the centre, an exact integer test
    (sum (2 i - n)^2 / n^2 against an irrational bound), and numbers them
    0, 1, ... in ascending order of their initial link (Morton code or raster
    index); every other link is -1.  Brute force over the cells."""
    g = svox2_amd.SparseGrid(reso=list(size), use_z_order=z_order, use_sphere_bound=True)
    links = g.links.numpy()
    bound = (1 + 3 ** 0.5 / max(size)) ** 2
    kept = []
    for i in range(size[0]):
        for j in range(size[1]):
            for k in range(size[2]):
                r2 = sum(((2 * c - n) / n) ** 2 for c, n in zip((i, j, k), size))
                if r2 <= bound:
                    code = (oracle_morton(i, j, k) if z_order and size[0] == size[1] == size[2]
                            else (i * size[1] + j) * size[2] + k)
                    kept.append((code, (i, j, k)))
    want = np.full(size, -1, np.int32)
    for row, (_, cell) in enumerate(sorted(kept)):
        want[cell] = row
    assert np.array_equal(links, want)
    assert g.capacity == len(kept) and g.sh_data.shape[0] == len(kept)


def oracle_morton(x, y, z):
    import oracle.oracle as O
    return O.morton_code_3(x, y, z)


@pytest.mark.parametrize("n", [64, 256])
def test_sphere_bound_matches_float32_corner_norm(n):
    """The float64 test agrees with evaluating the corner's norm in float32
    (the corner at (i - 1/2) * 2/n + 1/n - 1, its norm and the bound all
    rounded to float32)."""
    from svox2_amd.svox2 import _sphere_bound_cells
    ax = torch.arange(n, dtype=torch.float32) - 0.5
    X, Y, Z = torch.meshgrid(ax, ax, ax, indexing="ij")
    pts = torch.stack((X, Y, Z), -1).view(-1, 3) * np.float32(2.0 / n) + np.float32(1.0 / n - 1.0)
    f32 = pts.norm(dim=-1) <= np.float32(1.0 + 3 ** 0.5 / n)
    assert torch.equal(_sphere_bound_cells((n, n, n), "cpu"), f32)


def test_world_grid_maps_inverse():
    g = svox2_amd.SparseGrid(reso=[10, 12, 14], radius=[1.5, 2.0, 0.5], center=[0.1, -0.2, 0.3])
    p = torch.rand(100, 3) * 2 - 1
    assert torch.allclose(g.grid2world(g.world2grid(p)), p, atol=1e-5)
    # voxel centres at integers: the world corner maps to -0.5
    corner = (g.center - g.radius)[None]
    assert torch.allclose(g.world2grid(corner), torch.full((1, 3), -0.5), atol=1e-5)


def test_rand_cells_and_indexer():
    g = svox2_amd.SparseGrid(reso=10)
    np.random.seed(0)
    cells = g._get_rand_cells(0.05)
    assert cells.numel() == 50 and cells.dtype == torch.int32
    assert int(cells.min()) >= 0 and int(cells.max()) < 1000
    # a run of consecutive cells from np.random.randint(0, cells), wrapping
    for seed in range(40):
        np.random.seed(seed)
        start = np.random.RandomState(seed).randint(0, 1000)
        got = g._get_rand_cells(0.05).numpy()
        assert np.array_equal(got, (start + np.arange(50)) % 1000)
        # the run the TV kernel generates in place: the same draw
        np.random.seed(seed)
        assert g._rand_cell_run(0.05) == (start, 50)
    assert g._rand_cell_run(1.0) is None
    assert g._get_rand_cells(1.0) is None
    assert g._get_rand_cells(1.0, force=True).numel() == 1000
    # the optimisers take the touched-row mask as it is (no host sync); no
    # mask means every row
    assert g._maybe_convert_sparse_grad_indexer().dim() == 0
    g.sparse_grad_indexer = torch.zeros(1000, dtype=torch.bool)
    g.sparse_grad_indexer[[3, 7]] = True
    assert g._maybe_convert_sparse_grad_indexer() is g.sparse_grad_indexer


def test_compute_methods_refuse_cpu():
    g = svox2_amd.SparseGrid(reso=8)
    rays = svox2_amd.Rays(torch.zeros(4, 3), torch.ones(4, 3))
    with pytest.raises(RuntimeError, match="GPU"):
        g.volume_render(rays)
    with pytest.raises(NotImplementedError):
        g.volume_render(rays, use_kernel=False)
    with pytest.raises(RuntimeError, match="GPU"):
        g.volume_render_fused(rays, torch.zeros(4, 3))
    with pytest.raises(RuntimeError, match="GPU"):
        g.sample(torch.zeros(4, 3))
    with pytest.raises(RuntimeError, match="GPU"):
        g.accelerate()
    with pytest.raises(RuntimeError, match="GPU"):
        g.optim_density_step(0.1)


def test_ray_lengths():
    g = svox2_amd.SparseGrid(reso=8)
    rays = svox2_amd.Rays(torch.tensor([[-2.0, 0.1, 0.2]]), torch.tensor([[1.0, 0.0, 0.0]]))
    assert float(g.volume_render(rays, return_raylen=True)[0]) == pytest.approx(8.0)


def test_save_load_roundtrip(tmp_path):
    gen = torch.Generator().manual_seed(0)
    sc = scenes.make_blob_grid(16, scenes.BlobSpec(2, (0.2, 0.2), (10.0, 10.0)), gen)
    g = scenes.to_sparse_grid(sc, device="cpu", accelerate=False)
    path = str(tmp_path / "ckpt.npz")
    g.save(path)
    z = np.load(path)
    assert z["sh_data"].dtype == np.float16 and z["links"].dtype == np.int32
    assert set(z.files) == {"radius", "center", "links", "density_data", "sh_data", "basis_type"}
    g2 = svox2_amd.SparseGrid.load(path)
    assert torch.equal(g2.links, g.links)
    assert torch.equal(g2.density_data.data, g.density_data.data)
    assert torch.allclose(g2.sh_data.data, g.sh_data.data.half().float())


def test_scene_generators_deterministic():
    a = scenes.workload_dense64(seed=0, n_rays=16)
    b = scenes.workload_dense64(seed=0, n_rays=16)
    assert torch.equal(a[0].density, b[0].density) and torch.equal(a[1], b[1])
    assert torch.allclose(a[2].norm(dim=1), torch.ones(16), atol=1e-6)
    gen = torch.Generator().manual_seed(0)
    cams = scenes.CameraSet(scenes.hemisphere_cameras(4, gen))
    o, d = cams.random_batch(100, gen)
    assert torch.allclose(o.norm(dim=1), torch.full((100,), 4.0), atol=1e-5)
    assert torch.all(o[:, 2] >= 0)
    # rays point towards the scene
    assert torch.all((-o * d).sum(1) > 0)


def test_load_legacy_data_layout(tmp_path):
    """The reference's legacy checkpoint key "data" ((rows, 1 + C): density
    first, then SH; svox2.py:1560-1564), with no radius / center (unit radius
    at the origin) loads on the CPU with the values in place."""
    rng = np.random.default_rng(0)
    links = np.full((6, 5, 4), -1, np.int32)
    links[1:4, 2, 1:3] = np.arange(6).reshape(3, 2)
    data = rng.standard_normal((6, 28)).astype(np.float32)
    path = str(tmp_path / "legacy.npz")
    np.savez(path, links=links, data=data)
    g = svox2_amd.SparseGrid.load(path)
    assert g.links.shape == (6, 5, 4) and torch.equal(g.links, torch.from_numpy(links))
    assert np.array_equal(g.density_data.data.numpy(), data[:, :1])
    assert np.array_equal(g.sh_data.data.numpy(), data[:, 1:])
    assert g.basis_dim == 9 and g.capacity == 6
    assert torch.equal(g.radius, torch.ones(3)) and torch.equal(g.center, torch.zeros(3))
    # a background model loads with its links and data (svox2.py:1570-1622)
    rng = np.random.default_rng(4)
    bl = np.arange(18, dtype=np.int32).reshape(6, 3)
    bl[2, 1] = -1
    bd = rng.normal(size=(18, 2, 4)).astype(np.float32)
    np.savez(path, links=links, data=data, background_data=bd, background_links=bl)
    g = svox2_amd.SparseGrid.load(path)
    assert g.use_background and g.background_nlayers == 2 and g.background_reso == 3
    assert np.array_equal(g.background_links.numpy(), bl)
    assert np.array_equal(g.background_data.data.numpy(), bd)
    # and saves them again
    path2 = str(tmp_path / "bg_round_trip.npz")
    g.save(path2)
    z = np.load(path2)
    assert np.array_equal(z["background_links"], bl) and np.array_equal(z["background_data"], bd)
