"""libceed_amd: an MI355X-native (gfx950) backend for the libCEED operator-apply path.

The hot path v = E^T B^T D B E u runs in hand-written HIP kernels reached
through the C ABI of ``include/ceed-mi355-kernels.h`` (library
``libceed_amd/lib/libceed-mi355.so``).  This package holds the kernels
(``csrc/``), the host C layer (``host/``), the build script and thin Python
bindings used by tests, benchmarks and the autograd wrapper.
"""

from ._lib import LIB_PATH, CeedMi355Error, lib  # noqa: F401
from .mesh import HexMesh, hex_mesh, mesh_size_for_dofs, node_coordinates  # noqa: F401
