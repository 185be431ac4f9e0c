// p2p_v2.hip -- point->point kernels instantiated for the V2 core.
#include "p2p_kernels.hpp"
O3D_INSTANTIATE_P2P(o3d::CORE_V2, v2)
