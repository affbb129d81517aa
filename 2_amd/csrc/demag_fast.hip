// demag_fast.hip -- host launchers of the demag passes for power-of-two
// transform lengths (the kernels are in fast_kernels.hpp; the general xi
// epilogue is instantiated in xi_gen.hip).
#include "fast_kernels.hpp"

namespace mmx {

// ------------------------------------------------------------ host side


int fast_yf(const DemagPlanImpl *P, hipStream_t st)
{
    const DemagGeom &g = P->g;
    MMX_REQUIRE(g.n1 == g.p1 || 2 * g.n1 == g.p1, MMX_ERR_SIZE, "fast_yf: rows %d of %d", g.n1, g.p1);
    switch (g.p1) {
#define CASE(L)                                                                              \
    case L: {                                                                                \
        using C = ColTile<L>;                                                                \
        const int ntx = div_up(g.h2, C::W);                                                  \
        auto kern = (g.n1 == g.p1) ? yf_fast<L, true> : yf_fast<L, false>;                   \
        allow_lds_fast(kern, C::LDS);                                                        \
        kern<<<dim3(ntx * g.n0 * 3), dim3(C::NT), C::LDS, st>>>(g, ntx, P->fy.tw,            \
                                                                      P->xbuf, P->ybuf);     \
        break;                                                                               \
    }
        MMX_FAST_LENGTHS(CASE)
#undef CASE
    default: set_error("fast_yf: length %d", g.p1); return MMX_ERR_SIZE;
    }
    MMX_CHECK_LAUNCH();
    return MMX_OK;
}

int fast_yi(const DemagPlanImpl *P, hipStream_t st)
{
    const DemagGeom &g = P->g;
    MMX_REQUIRE(g.n1 == g.p1 || 2 * g.n1 == g.p1, MMX_ERR_SIZE, "fast_yi: rows %d of %d", g.n1, g.p1);
    switch (g.p1) {
#define CASE(L)                                                                              \
    case L: {                                                                                \
        using C = ColTile<L>;                                                                \
        const int ntx = div_up(g.h2, C::W);                                                  \
        auto kern = (g.n1 == g.p1) ? yi_fast<L, true> : yi_fast<L, false>;                   \
        allow_lds_fast(kern, C::LDS);                                                        \
        kern<<<dim3(ntx * g.n0 * 3), dim3(C::NT), C::LDS, st>>>(g, ntx, P->fy.tw,            \
                                                                      P->ybuf, P->xbuf);     \
        break;                                                                               \
    }
        MMX_FAST_LENGTHS(CASE)
#undef CASE
    default: set_error("fast_yi: length %d", g.p1); return MMX_ERR_SIZE;
    }
    MMX_CHECK_LAUNCH();
    return MMX_OK;
}

// The z pass addresses the spectrum and the kernel octant with 32-bit byte
// offsets from wave-uniform bases: the largest lane offsets are T planes of
// the spectrum and 3 T kz-planes of the octant.
bool fast_z_addressable(const DemagGeom &g)
{
    int T = 0;
    switch (g.p0) {
#define CASE(L)                                                                              \
    case L: T = ZTile<L>::T; break;
        MMX_FAST_LENGTHS(CASE)
#undef CASE
    default: return false;
    }
    const double lim = 4294967295.0;
    const double spec = ((double)T * g.p1 + 1.0) * g.sx * sizeof(float2);
    const double kern = ((double)3 * T * g.ko1 + 1.0) * g.h2 * 6 * sizeof(float);
    return spec < lim && kern < lim && (g.n0 == g.p0 || 2 * g.n0 == g.p0);
}

int fast_z(const DemagPlanImpl *P, hipStream_t st)
{
    const DemagGeom &g = P->g;
    MMX_REQUIRE(fast_z_addressable(g), MMX_ERR_SIZE, "fast_z: unsupported geometry");
    const bool full = g.n0 == g.p0;
    switch (g.p0) {
#define CASE(L)                                                                              \
    case L: {                                                                                \
        using Z = ZTile<L>;                                                                  \
        const int ntx = div_up(g.h2, Z::W);                                                  \
        auto kern = full ? z_fast<L, true> : z_fast<L, false>;                               \
        allow_lds_fast(kern, Z::LDS);                                                        \
        kern<<<dim3(ntx * g.p1), dim3(Z::NT), Z::LDS, st>>>(g, ntx, P->fz.tw, P->kern,       \
                                                            P->ybuf);                        \
        break;                                                                               \
    }
        MMX_FAST_LENGTHS(CASE)
#undef CASE
    default: set_error("fast_z: length %d", g.p0); return MMX_ERR_SIZE;
    }
    MMX_CHECK_LAUNCH();
    return MMX_OK;
}

int fast_y2d(const DemagPlanImpl *P, hipStream_t st)
{
    const DemagGeom &g = P->g;
    MMX_REQUIRE(g.n1 == g.p1 || 2 * g.n1 == g.p1, MMX_ERR_SIZE, "fast_y2d: rows %d of %d", g.n1, g.p1);
    switch (g.p1) {
#define CASE(L)                                                                              \
    case L: {                                                                                \
        using C = ColTile<L>;                                                                \
        const int ntx = div_up(g.h2, C::W);                                                  \
        auto kern = (g.n1 == g.p1) ? y2d_fast<L, true> : y2d_fast<L, false>;                 \
        allow_lds_fast(kern, C::LDS);                                                        \
        kern<<<dim3(ntx), dim3(C::NT), C::LDS, st>>>(g, ntx, P->fy.tw, P->kern,              \
                                                            P->xbuf);                        \
        break;                                                                               \
    }
        MMX_FAST_LENGTHS(CASE)
#undef CASE
    default: set_error("fast_y2d: length %d", g.p1); return MMX_ERR_SIZE;
    }
    MMX_CHECK_LAUNCH();
    return MMX_OK;
}

// The x passes address cells with 32-bit byte offsets from the component
// bases, and spectrum rows likewise; n2 is LX (zero-padded) or 2 LX (periodic).
bool fast_x_addressable(const DemagGeom &g)
{
    const double lim = 4294967295.0;
    return (double)g.N * sizeof(float) < lim &&
           ((double)g.nrows * g.sx + g.sx) * sizeof(float2) < lim &&
           (g.n2 == g.Lx || g.n2 == 2 * g.Lx);
}

int fast_xf(const DemagPlanImpl *P, int pred, const float *m, const float *msat_map,
            const float *tau0, float h, const float *norm_map, float *mpred, hipStream_t st)
{
    const DemagGeom &g = P->g;
    MMX_REQUIRE(fast_x_addressable(g), MMX_ERR_SIZE, "fast_xf: unsupported geometry");
    switch (g.Lx) {
#define CASE(L)                                                                              \
    case L: {                                                                                \
        using C = RowTile<L>;                                                                \
        const int nb = div_up(g.nrows, C::R);                                                \
        const bool full = g.n2 == 2 * L;                                                     \
        if (pred) {                                                                          \
            auto kern = full ? xf_fast<L, 1, true> : xf_fast<L, 1, false>;                   \
            allow_lds_fast(kern, C::LDS);                                                    \
            kern<<<dim3(nb), dim3(C::NT), C::LDS, st>>>(                                     \
                g, P->fx.tw, P->fx.rtw, m, msat_map, tau0, h, norm_map, mpred, P->xbuf);     \
        } else {                                                                             \
            auto kern = full ? xf_fast<L, 0, true> : xf_fast<L, 0, false>;                   \
            allow_lds_fast(kern, C::LDS);                                                    \
            kern<<<dim3(nb), dim3(C::NT), C::LDS, st>>>(                                     \
                g, P->fx.tw, P->fx.rtw, m, msat_map, nullptr, 0.0f, nullptr, nullptr,        \
                P->xbuf);                                                                    \
        }                                                                                    \
        break;                                                                               \
    }
        MMX_FAST_LENGTHS(CASE)
#undef CASE
    default: set_error("fast_xf: length %d", g.Lx); return MMX_ERR_SIZE;
    }
    MMX_CHECK_LAUNCH();
    return MMX_OK;
}

bool fast_xi_applies(const FieldParams &fp)
{
    return !fp.msat_map && !fp.aex_map && !fp.alpha_map && !fp.ku_map && !fp.zl && !fp.htherm &&
           !fp.bext_map && !fp.u_map;
}

int fast_xi(const DemagPlanImpl *P, int mode, const FieldParams &fp, const EpilogueIO &io,
            hipStream_t st)
{
    if (mode != EPI_B && !fast_xi_applies(fp)) return fast_xi_general(P, mode, fp, io, st);
    return fast_xi_modes<false>(P, mode, fp, io, st);
}

} // namespace mmx
