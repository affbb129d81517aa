// xi_maps.hip -- the xi pass of the power-of-two lengths with the two-cell
// epilogue for maps and masks (Msat, Aex, alpha, Ku, B_ext, anisU), the
// thermal field and H_dex, in a file of its own so that it compiles in
// parallel with the other passes.
#include "fast_launch.hpp"

namespace mmx {

#define MMX_INST_XI_MAPS(L) MMX_INST_XI(L, XI_MAPS)
MMX_POW2_LENGTHS(MMX_INST_XI_MAPS)
#undef MMX_INST_XI_MAPS

} // namespace mmx
