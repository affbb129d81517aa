// xi_gen.hip -- the xi pass of the power-of-two lengths with the general
// per-cell epilogue (Zhang-Li with its polarisation and xi masks, with or
// without the other maps, masks, the thermal field and H_dex), in a file of
// its own so that it compiles in parallel with demag_fast.hip.
#include "fast_launch.hpp"

namespace mmx {

#define MMX_INST_XI_GEN(L) MMX_INST_XI(L, XI_GENERAL)
MMX_POW2_LENGTHS(MMX_INST_XI_GEN)
#undef MMX_INST_XI_GEN

} // namespace mmx
