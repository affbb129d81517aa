// fast_px.hip -- the x-pass launchers (xf, and xi with the pair epilogue) of the
// power-of-two lengths (fast_launch.hpp), in a file of its own so that it
// compiles in parallel with the other instantiations.
#include "fast_launch.hpp"

namespace mmx {

MMX_POW2_LENGTHS(MMX_INST_XF)
#define MMX_INST_XI_PLAIN(L) MMX_INST_XI(L, XI_UNIFORM)
MMX_POW2_LENGTHS(MMX_INST_XI_PLAIN)
#undef MMX_INST_XI_PLAIN

} // namespace mmx
