// fast_launch.hpp -- launchers of the register-FFT passes for one transform
// length each.  demag_fast.hip dispatches on the plan's lengths and calls
// them; they are instantiated by length in separate files (demag_fast.hip,
// xi_gen.hip, fast_r3.hip, fast_r5.hip, fast_r7.hip) that compile in parallel.
#pragma once

#include "fast_kernels.hpp"

namespace mmx {

template <int L>
int xf_len(const DemagPlanImpl *P, int pred, const float *m, const float *msat_map,
           const float *tau0, float h, const float *norm_map, float *mpred, hipStream_t st)
{
    const DemagGeom &g = P->g;
    using C = RowTile<L>;
    const int nb = div_up(g.nrows, C::R);
    const bool full = g.n2 == 2 * L;
    if (pred) {
        auto kern = full ? xf_fast<L, 1, true> : xf_fast<L, 1, false>;
        allow_lds_fast(kern, C::LDS);
        kern<<<dim3(nb), dim3(C::NT), C::LDS, st>>>(g, P->fx.tw, P->fx.rtw, m, msat_map, tau0, h,
                                                     norm_map, mpred, P->xbuf, nullptr, nullptr);
    } else {
        auto kern = full ? xf_fast<L, 0, true> : xf_fast<L, 0, false>;
        allow_lds_fast(kern, C::LDS);
        kern<<<dim3(nb), dim3(C::NT), C::LDS, st>>>(g, P->fx.tw, P->fx.rtw, m, msat_map, nullptr,
                                                     0.0f, nullptr, nullptr, P->xbuf,
                                                     P->handoff_dev, P->handoff_host);
    }
    MMX_CHECK_LAUNCH();
    return MMX_OK;
}

template <int L> int yf_len(const DemagPlanImpl *P, hipStream_t st)
{
    const DemagGeom &g = P->g;
    if constexpr (L == 2048) {
        using C = PairTile<L>;
        const int ntx = div_up(g.h2, C::W);
        auto kern = (g.n1 == g.p1) ? yf_pair<L, true> : yf_pair<L, false>;
        allow_lds_fast(kern, C::LDS);
        kern<<<dim3(ntx * g.n0 * 3), dim3(C::NT), C::LDS, st>>>(g, ntx, P->fy.tw, P->xbuf,
                                                                 P->ybuf);
        MMX_CHECK_LAUNCH();
        return MMX_OK;
    }
    using C = ColTile<L>;
    const int ntx = div_up(g.h2, C::W);
    auto kern = (g.n1 == g.p1) ? yf_fast<L, true> : yf_fast<L, false>;
    allow_lds_fast(kern, C::LDS);
    kern<<<dim3(ntx * g.n0 * 3), dim3(C::NT), C::LDS, st>>>(g, ntx, P->fy.tw, P->xbuf, P->ybuf);
    MMX_CHECK_LAUNCH();
    return MMX_OK;
}

template <int L> int yi_len(const DemagPlanImpl *P, hipStream_t st)
{
    const DemagGeom &g = P->g;
    if constexpr (L == 2048) {
        using C = PairTile<L>;
        const int ntx = div_up(g.h2, C::W);
        auto kern = (g.n1 == g.p1) ? yi_pair<L, true> : yi_pair<L, false>;
        allow_lds_fast(kern, C::LDS);
        kern<<<dim3(ntx * g.n0 * 3), dim3(C::NT), C::LDS, st>>>(g, ntx, P->fy.tw, P->ybuf,
                                                                 P->xbuf);
        MMX_CHECK_LAUNCH();
        return MMX_OK;
    }
    using C = ColTile<L>;
    const int ntx = div_up(g.h2, C::W);
    auto kern = (g.n1 == g.p1) ? yi_fast<L, true> : yi_fast<L, false>;
    allow_lds_fast(kern, C::LDS);
    kern<<<dim3(ntx * g.n0 * 3), dim3(C::NT), C::LDS, st>>>(g, ntx, P->fy.tw, P->ybuf, P->xbuf);
    MMX_CHECK_LAUNCH();
    return MMX_OK;
}

template <int L> int z_len(const DemagPlanImpl *P, hipStream_t st)
{
    const DemagGeom &g = P->g;
    using Z = ZTile<L>;
    const int ntx = div_up(g.h2, Z::W);
    auto kern = (g.n0 == g.p0) ? z_fast<L, true> : z_fast<L, false>;
    allow_lds_fast(kern, Z::LDS);
    kern<<<dim3(ntx * g.p1), dim3(Z::NT), Z::LDS, st>>>(g, ntx, P->fz.tw, P->kern, P->ybuf, 0);
    MMX_CHECK_LAUNCH();
    return MMX_OK;
}

// The z pass's kernel on a caller's geometry: a film's y transform with the
// kernel multiply, the film's rows taking the place of the z planes
// (fast_y2d_as_z in demag_fast.hip).
template <int L>
int z_len_geom(const DemagGeom &g, const float2 *tw, const float *K, float2 *buf, hipStream_t st)
{
    using Z = ZTile<L>;
    const int ntx = div_up(g.h2, Z::W);
    auto kern = (g.n0 == g.p0) ? z_fast<L, true> : z_fast<L, false>;
    allow_lds_fast(kern, Z::LDS);
    kern<<<dim3(ntx * g.p1), dim3(Z::NT), Z::LDS, st>>>(g, ntx, tw, K, buf, 1);
    MMX_CHECK_LAUNCH();
    return MMX_OK;
}

// xi of length L in `mode` with epilogue variant VAR (XiVariant; EPI_B
// always runs the uniform one, which only stores the field)
template <int L, int MODE, int VAR>
static void xi_one(const DemagPlanImpl *P, const FieldParams &fp, const EpilogueIO &io,
                   hipStream_t st)
{
    using C = RowTile<L>;
    const int nb = div_up(P->g.nrows, C::R);
    auto kern = (P->g.n2 == 2 * L) ? xi_fast<L, MODE, true, VAR> : xi_fast<L, MODE, false, VAR>;
    allow_lds_fast(kern, C::LDS);
    kern<<<dim3(nb), dim3(C::NT), C::LDS, st>>>(P->g, P->fx.tw, P->fx.rtw, P->xbuf, fp, io);
}

template <int L, int VAR>
int xi_len(const DemagPlanImpl *P, int mode, const FieldParams &fp, const EpilogueIO &io,
           hipStream_t st)
{
    switch (mode) {
    case EPI_B: xi_one<L, EPI_B, XI_UNIFORM>(P, fp, io, st); break;
    case EPI_HEFF: xi_one<L, EPI_HEFF, VAR>(P, fp, io, st); break;
    case EPI_TAU0: xi_one<L, EPI_TAU0, VAR>(P, fp, io, st); break;
    default: xi_one<L, EPI_CORR, VAR>(P, fp, io, st); break;
    }
    MMX_CHECK_LAUNCH();
    return MMX_OK;
}

// explicit instantiations of one length's launchers
#define MMX_INST_XF(L)                                                                       \
    template int xf_len<L>(const DemagPlanImpl *, int, const float *, const float *,         \
                           const float *, float, const float *, float *, hipStream_t);
#define MMX_INST_Y(L)                                                                        \
    template int yf_len<L>(const DemagPlanImpl *, hipStream_t);                              \
    template int yi_len<L>(const DemagPlanImpl *, hipStream_t);
#define MMX_INST_Z(L)                                                                        \
    template int z_len<L>(const DemagPlanImpl *, hipStream_t);                               \
    template int z_len_geom<L>(const DemagGeom &, const float2 *, const float *, float2 *,   \
                               hipStream_t);
#define MMX_INST_XI(L, VAR)                                                                  \
    template int xi_len<L, VAR>(const DemagPlanImpl *, int, const FieldParams &,             \
                                const EpilogueIO &, hipStream_t);

} // namespace mmx
