"""The C ABI: every function include/svox2_amd.h declares is exported by the
built library and bound by svox2_amd._C; the ctypes mirrors have the C struct
layouts; argument validation fails with a status and a message (no GPU work is
started by these calls)."""
import ctypes
import os
import re
import subprocess

import pytest

from svox2_amd import _C

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "svox2_amd.h")


def _declared():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return set(re.findall(r"\b(svox2_\w+)\s*\(", text))


def test_header_matches_binding_list():
    assert _declared() == set(_C.EXPORTED_SYMBOLS)


def test_library_exports_every_symbol():
    lib = _C.load_library()
    for name in _declared():
        assert hasattr(lib, name), name
    assert lib.svox2_version().decode().startswith("svox2_amd")


def test_struct_layouts_match_c(tmp_path):
    src = tmp_path / "layout.c"
    src.write_text(r'''
#include <stdio.h>
#include <stddef.h>
#include "svox2_amd.h"
int main(void) {
  printf("%zu %zu %zu %zu %zu %zu %zu %zu\n", sizeof(svox2_grid_t), offsetof(svox2_grid_t, rows),
         offsetof(svox2_grid_t, size), offsetof(svox2_grid_t, storage),
         offsetof(svox2_grid_t, offset), offsetof(svox2_grid_t, scaling),
         offsetof(svox2_grid_t, sh_stride), offsetof(svox2_grid_t, cell_map));
  printf("%zu %zu\n", sizeof(svox2_render_options_t), offsetof(svox2_render_options_t, last_sample_opaque));
  printf("%zu %zu %zu\n", sizeof(svox2_camera_t), offsetof(svox2_camera_t, width), offsetof(svox2_camera_t, ndc_coeffy));
  printf("%zu %zu\n", sizeof(svox2_grads_t), offsetof(svox2_grads_t, grad_sh_stride));
  printf("%zu %zu %zu %zu %zu %zu\n", offsetof(svox2_grid_t, background_links),
         offsetof(svox2_grid_t, background_rows), offsetof(svox2_grid_t, background_reso),
         offsetof(svox2_grid_t, background_nlayers),
         offsetof(svox2_grads_t, grad_background_out),
         offsetof(svox2_grads_t, mask_background_out));
  return 0;
}
''')
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c99", f"-I{os.path.dirname(HEADER)}", str(src), "-o", str(exe)],
                   check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split()
    c = [int(x) for x in out]
    G, O, Cm, Gr = _C._Grid, _C._Opt, _C._Cam, _C._Grads
    py = [ctypes.sizeof(G), G.rows.offset, G.size.offset, G.storage.offset, G.offset.offset,
          G.scaling.offset, G.sh_stride.offset, G.cell_map.offset, ctypes.sizeof(O),
          O.last_sample_opaque.offset, ctypes.sizeof(Cm),
          Cm.width.offset, Cm.ndc_coeffy.offset, ctypes.sizeof(Gr), Gr.grad_sh_stride.offset,
          G.background_links.offset, G.background_rows.offset, G.background_reso.offset,
          G.background_nlayers.offset, Gr.grad_background_out.offset,
          Gr.mask_background_out.offset]
    assert c == py


def test_invalid_arguments_return_status_and_message():
    lib = _C.load_library()
    opt = _C._Opt()
    opt.step_size = 0.5
    st = lib.svox2_volume_render_cuvol(None, None, None, None, 0, ctypes.byref(opt), None, None,
                                       None, None)
    assert st == 1
    assert b"grid is NULL" in lib.svox2_last_error()
    g = _C._Grid()
    g.links = 16
    g.density_data = 16
    g.sh_data = 16
    g.rows = 1
    g.size[0] = g.size[1] = g.size[2] = 4
    g.basis_dim, g.sh_data_dim, g.basis_type, g.storage = 9, 27, 1, 0
    bad = _C._Opt()
    bad.step_size = 0.0  # would never terminate: rejected
    st = lib.svox2_volume_render_cuvol(ctypes.byref(g), None, None, None, 0, ctypes.byref(bad),
                                       None, None, None, None)
    assert st == 1 and b"step_size" in lib.svox2_last_error()
    g.basis_dim, g.sh_data_dim = 16, 48
    st = lib.svox2_volume_render_cuvol(ctypes.byref(g), None, None, None, 0, ctypes.byref(opt),
                                       None, None, None, None)
    assert st == 1 and b"basis_dim" in lib.svox2_last_error()
    g.basis_dim, g.sh_data_dim = 9, 27
    st = lib.svox2_rmsprop_step(None, 0, None, None, 10, 1, 0, 0, 0, 7, None, 0, 0.9, 0.1,
                                1e-8, -1e9, 0.1, None)
    assert st == 1 and b"indexer kind" in lib.svox2_last_error()
    size = (ctypes.c_int32 * 3)(8, 8, 8)
    need = lib.svox2_accel_dist_prop_workspace_bytes(size)
    assert need == 4 * 4 * 4 + 2 * 2 * 2 + 1 + 3  # 73 bytes rounded up to whole words
    st = lib.svox2_accel_dist_prop(16, size, None, 0, None)
    assert st == 1 and b"workspace" in lib.svox2_last_error()


def test_python_checks_reject_cpu_tensors():
    import torch
    spec = _C.SparseGridSpec()
    spec.links = torch.zeros(4, 4, 4, dtype=torch.int32)
    spec.density_data = torch.zeros(64, 1)
    spec.sh_data = torch.zeros(64, 27)
    spec._offset = torch.zeros(3)
    spec._scaling = torch.ones(3)
    rays = _C.RaysSpec()
    rays.origins = torch.zeros(2, 3)
    rays.dirs = torch.ones(2, 3)
    with pytest.raises(RuntimeError, match="CUDA tensor"):
        _C.volume_render_cuvol(spec, rays, _C.RenderOptions())
