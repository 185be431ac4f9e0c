// p2p_wl.hip -- point->point kernels instantiated for the WL core.
#include "p2p_kernels.hpp"
O3D_INSTANTIATE_P2P(o3d::CORE_WL, wl)
