// fast_r3.hip -- the register-FFT passes of the transform lengths with a
// factor 3 (mixed-radix in-register FFTs, reg_fft.hpp): every pass for the
// lengths of at most 28 registers per component, the strided 3-D y passes for
// the longer ones.
#include "fast_launch.hpp"

namespace mmx {

#define MMX_INST_XI_ALL(L) MMX_INST_XI(L, XI_UNIFORM) MMX_INST_XI(L, XI_MAPS) MMX_INST_XI(L, XI_GENERAL)
MMX_R3_SHORT(MMX_INST_XF)
MMX_R3_SHORT(MMX_INST_XI_ALL)
MMX_R3_SHORT(MMX_INST_Z)
MMX_R3_SHORT(MMX_INST_Y)
MMX_R3_LONG(MMX_INST_Y)
#undef MMX_INST_XI_ALL

} // namespace mmx
