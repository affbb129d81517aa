// demag_fast.hip -- dispatch of the register-FFT demag passes on the plan's
// transform lengths.  The kernels are in fast_kernels.hpp, the launchers of
// one length in fast_launch.hpp, instantiated by length in files that compile
// in parallel: fast_px.hip (x passes), fast_py.hip (y), fast_pz.hip (z) and
// xi_gen.hip (the general xi epilogue) for the powers of two, fast_r3.hip,
// fast_r5.hip and fast_r7.hip for the lengths with a factor 3, 5 or 7.  This
// file sees only the launchers' declarations, so it instantiates none.
#include "fast_kernels.hpp"

namespace mmx {

// ------------------------------------------------------------ host side

int fast_yf(const DemagPlanImpl *P, hipStream_t st)
{
    const DemagGeom &g = P->g;
    MMX_REQUIRE(g.n1 == g.p1 || 2 * g.n1 == g.p1, MMX_ERR_SIZE, "fast_yf: rows %d of %d", g.n1, g.p1);
    switch (g.p1) {
#define CASE(L)                                                                              \
    case L: return yf_len<L>(P, st);
        MMX_ALL_LENGTHS(CASE)
#undef CASE
    default: set_error("fast_yf: length %d", g.p1); return MMX_ERR_SIZE;
    }
}

int fast_yi(const DemagPlanImpl *P, hipStream_t st)
{
    const DemagGeom &g = P->g;
    MMX_REQUIRE(g.n1 == g.p1 || 2 * g.n1 == g.p1, MMX_ERR_SIZE, "fast_yi: rows %d of %d", g.n1, g.p1);
    switch (g.p1) {
#define CASE(L)                                                                              \
    case L: return yi_len<L>(P, st);
        MMX_ALL_LENGTHS(CASE)
#undef CASE
    default: set_error("fast_yi: length %d", g.p1); return MMX_ERR_SIZE;
    }
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
        MMX_SHORT_LENGTHS(CASE)
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
    switch (g.p0) {
#define CASE(L)                                                                              \
    case L: return z_len<L>(P, st);
        MMX_SHORT_LENGTHS(CASE)
#undef CASE
    default: set_error("fast_z: length %d", g.p0); return MMX_ERR_SIZE;
    }
}

// A film's y pass on the z pass's kernel: three thread groups transform the
// three components of a column tile at once (a one-group register pass that
// ran them one after the other in every thread took 11.8 against 7.7 us at
// 1x32x128, 35.8 against 16.2 at 1x256x256).  The film's rows are the planes (n0 := n1, p0 :=
// p1) of a grid with a single ky; the film's octant (ky, kx) then sits where
// the z pass reads (kz, 0, kx), with the kz stride h2 (ko1 := 1); the
// mirror of the transformed axis then negates YZ and XY (film_y), not XZ and
// XY as along internal axis 0.
bool fast_y2d_as_z_applies(const DemagGeom &g)
{
    DemagGeom f = g;
    f.n0 = g.n1;
    f.p0 = g.p1;
    f.p1 = 1;
    f.ko0 = g.ko1;
    f.ko1 = 1;
    return g.p0 == 1 && fast_length(g.p1) && fast_z_addressable(f);
}

int fast_y2d_as_z(const DemagPlanImpl *P, hipStream_t st)
{
    const DemagGeom &g = P->g;
    MMX_REQUIRE(fast_y2d_as_z_applies(g), MMX_ERR_SIZE, "fast_y2d_as_z: unsupported geometry");
    DemagGeom f = g;
    f.n0 = g.n1;
    f.p0 = g.p1;
    f.p1 = 1;
    f.ko0 = g.ko1;
    f.ko1 = 1;
    switch (f.p0) {
#define CASE(L)                                                                              \
    case L: return z_len_geom<L>(f, P->fy.tw, P->kern, P->xbuf, st);
        MMX_SHORT_LENGTHS(CASE)
#undef CASE
    default: set_error("fast_y2d_as_z: length %d", f.p0); return MMX_ERR_SIZE;
    }
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
    case L: return xf_len<L>(P, pred, m, msat_map, tau0, h, norm_map, mpred, st);
        MMX_SHORT_LENGTHS(CASE)
#undef CASE
    default: set_error("fast_xf: length %d", g.Lx); return MMX_ERR_SIZE;
    }
}

// Workgroups of the register-FFT x pass (RowTile rows per workgroup).
int fast_x_blocks(const DemagGeom &g)
{
    switch (g.Lx) {
#define CASE(L)                                                                              \
    case L: return div_up((int)g.nrows, RowTile<L>::R);
        MMX_SHORT_LENGTHS(CASE)
#undef CASE
    default: return 0;
    }
}

bool fast_xi_applies(const FieldParams &fp)
{
    return !fp.msat_map && !fp.aex_map && !fp.alpha_map && !fp.ku_map && !fp.htherm &&
           !fp.bext_map && !fp.u_map && !fp.hdex && fast_xi_maps_applies(fp);
}

bool fast_xi_maps_applies(const FieldParams &fp) { return !fp.zl; }

int fast_xi(const DemagPlanImpl *P, int mode, const FieldParams &fp, const EpilogueIO &io,
            hipStream_t st)
{
    const DemagGeom &g = P->g;
    MMX_REQUIRE(fast_x_addressable(g), MMX_ERR_SIZE, "fast_xi: unsupported geometry");
    const int var = (mode == EPI_B || fast_xi_applies(fp)) ? XI_UNIFORM
                    : fast_xi_maps_applies(fp)                ? XI_MAPS
                                                              : XI_GENERAL;
    switch (g.Lx) {
#define CASE(L)                                                                              \
    case L:                                                                                  \
        return var == XI_UNIFORM ? xi_len<L, XI_UNIFORM>(P, mode, fp, io, st)                \
               : var == XI_MAPS  ? xi_len<L, XI_MAPS>(P, mode, fp, io, st)                   \
                                 : xi_len<L, XI_GENERAL>(P, mode, fp, io, st);
        MMX_SHORT_LENGTHS(CASE)
#undef CASE
    default: set_error("fast_xi: length %d", g.Lx); return MMX_ERR_SIZE;
    }
}

} // namespace mmx
