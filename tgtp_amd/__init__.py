"""tgtp_amd -- MI355X-native TGT+/SABA spanning-edge-centrality hot path.

Layers:
  include/tgtp_hip.h, include/tgtp_graph_io.h   C ABI (the drop-in boundary)
  include/tgtp_aesc.hpp                          the reference's AESC interface (C++ header)
  tgtp_amd/csrc/*.hip, *.cpp                     HIP kernels for gfx950 + host C++
  tgtp_amd/libtgtp_hip.so                        built in-tree by tgtp_amd.build
  tgtp_amd.runtime                               Context / DeviceGraph (ctypes)
  tgtp_amd/bin/{tgtp,benchmark,rw_experiment_saba,rw_experiment,rw_experiment_simd}
                                                 command-line drop-ins
"""
from ._lib import (MODE_FAITHFUL, MODE_FAST_BF16, MODE_FAST_FP32, TAU_REFERENCE, WALK2_MT19937,  # noqa: F401,E501
                   WALK_MT_REGULAR, WALK_MT_SIMD, TgtpError)
from .graph import (Graph, chung_lu, complete, cycle, grid_rewired, load_eigens_txt,  # noqa: F401
                    load_eigens_txt_cached, load_txt, load_txt_cached, petersen)
from .runtime import Context, DeviceGraph  # noqa: F401
