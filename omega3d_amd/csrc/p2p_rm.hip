// p2p_rm.hip -- point->point kernels instantiated for the RM core.
#include "p2p_kernels.hpp"
O3D_INSTANTIATE_P2P(o3d::CORE_RM, rm)
