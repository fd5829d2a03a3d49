"""lpa_amd -- MI355X-native label propagation over CGR-compressed graphs.

Python binding of the two in-tree native libraries:
  liblpa_host.so  (include/lpa_host.h)  generators, CSR, CGR encoder, file set
  liblpa_hip.so   (include/lpa_amd.h)   gfx950 decode + LP engine
"""
from . import host  # noqa: F401
from .host import Cgr, Csr, LpaError, gen_copying, gen_lattice, gen_rmat, threads  # noqa: F401
from .engine import Engine, Mlp, device_count, gen_rmat_device_bfs  # noqa: F401,E402
from . import schedule  # noqa: F401,E402
