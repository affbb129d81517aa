// fast_py.hip -- the y-pass launchers (3-D: yf, yi; films: y2d) of the
// power-of-two lengths (fast_launch.hpp), in a file of its own so that it
// compiles in parallel with the other instantiations.
#include "fast_launch.hpp"

namespace mmx {

MMX_POW2_LENGTHS(MMX_INST_Y)
MMX_POW2_LENGTHS(MMX_INST_Y2D)

} // namespace mmx
