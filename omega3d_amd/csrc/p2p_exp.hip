// p2p_exp.hip -- point->point kernels instantiated for the EXP core.
#include "p2p_kernels.hpp"
O3D_INSTANTIATE_P2P(o3d::CORE_EXP, exp)
