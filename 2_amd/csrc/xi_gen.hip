// xi_gen.hip -- the xi pass for power-of-two lengths with the general
// per-cell epilogue (Msat/Aex/alpha/Ku maps, the B_ext and anisU masks,
// Zhang-Li with its polarisation and xi masks, the thermal field), in a file
// of its own so that it compiles in parallel with demag_fast.hip.
#include "fast_kernels.hpp"

namespace mmx {

int fast_xi_general(const DemagPlanImpl *P, int mode, const FieldParams &fp,
                    const EpilogueIO &io, hipStream_t st)
{
    return fast_xi_modes<true>(P, mode, fp, io, st);
}

} // namespace mmx
