"""MI355X (gfx950) sparse-voxel differentiable renderer with the public
interface of svox2 (svox2/__init__.py:1-3): ``SparseGrid``, ``Camera``,
``Rays``, ``RenderOptions`` and the basis-type constants.  Rendering, its
gradients, sampling, total variation and RMSprop run as hand-written HIP
kernels in ``libsvox2_amd.so`` (C ABI in include/svox2_amd.h)."""
from .defs import *  # noqa: F401,F403
from .svox2 import SparseGrid, Camera, Rays, RenderOptions  # noqa: F401
from .version import __version__  # noqa: F401
