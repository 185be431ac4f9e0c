// p2p_kernels_decl.hpp -- launchers exported by the per-core kernel translation units.
#pragma once
#include "o3d_common.hpp"

namespace o3d {
#define O3D_DECL_P2P(NAME)                                                                  \
  hipError_t launch_p2p_##NAME(const P2PArgs& a, bool thick, bool grad, int prec, int tpt,  \
                               hipStream_t st);                                            \
  hipError_t launch_p2p64_##NAME(const P2PArgs64& a, bool thick, bool grad, hipStream_t st); \
  hipError_t launch_p2p_split_##NAME(const P2PArgs& a, bool thick, bool grad, int prec, int nslice, \
                                     int64_t slen, void* part, hipStream_t st);
O3D_DECL_P2P(rm)
O3D_DECL_P2P(wl)
O3D_DECL_P2P(exp)
O3D_DECL_P2P(v2)
#undef O3D_DECL_P2P
}  // namespace o3d
