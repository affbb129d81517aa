// fast_py.hip -- the 3-D y-pass launchers (yf, yi) of the
// power-of-two lengths (fast_launch.hpp), in a file of its own so that it
// compiles in parallel with the other instantiations.
#include "fast_launch.hpp"

namespace mmx {

MMX_POW2_LENGTHS(MMX_INST_Y)

} // namespace mmx
