// fast_r7.hip -- the register-FFT passes of the transform lengths with a
// factor 7 (mixed-radix in-register FFTs, reg_fft.hpp): every pass for the
// lengths of at most 28 registers per component, the strided 3-D y passes for
// the longer ones.
#include "fast_launch.hpp"

namespace mmx {

#define MMX_INST_XI_BOTH(L) MMX_INST_XI(L, false) MMX_INST_XI(L, true)
MMX_R7_SHORT(MMX_INST_XF)
MMX_R7_SHORT(MMX_INST_XI_BOTH)
MMX_R7_SHORT(MMX_INST_Z)
MMX_R7_SHORT(MMX_INST_Y)
MMX_R7_LONG(MMX_INST_Y)
#undef MMX_INST_XI_BOTH

} // namespace mmx
