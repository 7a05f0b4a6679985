// tos_kernels.hip -- gfx950 (CDNA4) kernels of the pair-counting estimator.
//
// Work mapping (one wavefront = 64 lanes):
//   lane   <-> one primary pixel p1 of sightline i (held in VGPRs),
//   loop   <-> every pixel p2 of every partner sightline j of i; the p2
//              record is wave-uniform and comes through the scalar cache.
// So one loop step evaluates 64 pixel pairs with one scalar load.
// Each workgroup owns one LDS histogram (u32 counts + FP64 sums) that all
// its wavefronts update with LDS atomics; at the end it is written to a
// per-workgroup partial that a deterministic tree reduction sums.
//
// Arithmetic contract (XiEstimator.cc:161-296, types.h:77-87):
//   * TOS_FP64 evaluates the reference's float/double expression graph
//     operation by operation (the file is built with -ffp-contract=off), with
//     x86 static_cast<unsigned>(double) emulated by axis_checked/unchecked.
//   * TOS_FP32C evaluates separations in FP32 and re-bins in FP64 (the same
//     functions as TOS_FP64) every pair whose FP32 coordinate lies within a
//     proven error bound of a bin edge, so bins are identical.
//   The per-pair terms f32(w1 d1 * w2 d2) and f32(w1 w2) are the reference's
//   own FP32 products in both modes; they are accumulated in FP64.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "tos_internal.h"

namespace tos {
namespace {

constexpr int kWave = 64;

// x86-64 cvttsd2si r64 then the low 32 bits: NaN / out of int64 range -> 0.
__device__ __forceinline__ uint32_t cvt_x86_slow(double t)
{
    if (!(t >= -9223372036854775808.0 && t < 9223372036854775808.0)) return 0u;
    return (uint32_t)(uint64_t)(int64_t)t;
}

// static_cast<unsigned>(t), rejected when >= n (n < 2^24).  For t in
// (-1, n) the cast is trunc(t); for t in [-4e9, -1] or [n, 4e9] the x86
// result is >= n; anything else (NaN, huge values) takes the exact path.
__device__ __forceinline__ uint32_t axis_checked(double t, uint32_t n)
{
    if (t > -1.0 && t < (double)n) return (uint32_t)(int32_t)t;
    if (fabs(t) < 4.0e9) return kReject;
    const uint32_t b = cvt_x86_slow(t);
    return b < n ? b : kReject;
}

// static_cast<unsigned>(t) without a range check (the z axis,
// XiEstimator.cc:200,253,287).
__device__ __forceinline__ uint32_t axis_unchecked(double t)
{
    if (t > -1.0 && t < 2147483647.0) return (uint32_t)(int32_t)t;
    return cvt_x86_slow(t);
}

__device__ __forceinline__ uint32_t z_bin_exact(const GridExact &g, float l1, float l2)
{
    const float lsum = l1 + l2;
    const double zc = 0.5 * (double)lsum;
    return axis_unchecked((zc - g.min[2]) * g.inv[2]);
}

// LDS slot of flat bin f < ntot (see GridExact::pad).
__device__ __forceinline__ uint32_t slot_of(const GridExact &g, uint32_t f)
{
    return f + (f / g.rowlen) * g.pad;
}

// Slot of the in-range triple (i0 < n0, i1 < n1, iz < n2): 24-bit
// multiply-adds (all factors are below 2^24).
__device__ __forceinline__ uint32_t slot3(const GridExact &g, uint32_t i0, uint32_t i1, uint32_t iz)
{
    return __umul24(i0, g.rowlen + g.pad) + __umul24(i1, g.n[2]) + iz;
}

// a * b + c for a, b < 2^24 (b wave-uniform, in an SGPR): one v_mad_u32_u24,
// where the compiler would otherwise emit two multiplies and a three-way add.
__device__ __forceinline__ uint32_t mad_u24(uint32_t a, uint32_t b, uint32_t c)
{
    uint32_t r;
    asm("v_mad_u32_u24 %0, %1, %2, %3" : "=v"(r) : "v"(a), "s"(b), "v"(c));
    return r;
}

// Flat bin (reference order, u32 arithmetic) -> LDS slot or kReject.  When
// the z bin is the sightline pair's constant izc < n_z the flat bin is in
// range and its row is i0.
__device__ __forceinline__ uint32_t finish_bin(const GridExact &g, uint32_t i0, uint32_t i1, bool zc, uint32_t izc,
                                               float l1, float l2)
{
    if (zc) return slot3(g, i0, i1, izc);
    const uint32_t f = z_bin_exact(g, l1, l2) + (i1 + i0 * g.n[1]) * g.n[2];
    return f < g.ntot ? slot_of(g, f) : kReject;
}

// sqrt(x), correctly rounded, for x = +-0, +inf, NaN, x < 0 (NaN) and every
// x >= 2^-767: the hardware reciprocal square root refined by two
// Newton-Raphson (Goldschmidt) steps on g ~ sqrt(x) and h ~ 1/(2 sqrt(x)),
// the sequence the AMDGPU backend emits for sqrt(double) without its
// rescaling of inputs below 2^-767.  The squared separations binned here
// are built from FP32 distances: a nonzero one is at least ~1e-106.
__device__ __forceinline__ double sqrt_no_tiny(double x)
{
    const double y = __builtin_amdgcn_rsq(x);
    double g = x * y;
    double h = 0.5 * y;
    const double r = __builtin_fma(-h, g, 0.5);
    g = __builtin_fma(g, r, g);
    h = __builtin_fma(h, r, h);
    double d = __builtin_fma(-g, g, x);
    g = __builtin_fma(d, h, g);
    d = __builtin_fma(-g, g, x);
    g = __builtin_fma(d, h, g);
    return (x == 0.0 || x == __builtin_inf()) ? x : g;
}

// Cartesian branch, XiEstimator.cc:226-259.
__device__ __forceinline__ uint32_t exact_bin_cart(const GridExact &g, float d1, float l1, double r1_sq,
                                                   double two_r1_cos12, float d2, float l2, bool zc,
                                                   uint32_t izc)
{
    const float adiff = fabsf(d1 - d2);
    const double r_para = (double)adiff;
    const uint32_t ipar = axis_checked((r_para - g.min[0]) * g.inv[0], g.n[0]);
    if (ipar == kReject) return kReject;
    const double d2d = (double)d2;
    const double r_sq = r1_sq + (d2d - two_r1_cos12) * d2d;
    const double rp2 = fabs(r_sq - r_para * r_para);
    const uint32_t iperp = axis_checked((sqrt_no_tiny(rp2) - g.min[1]) * g.inv[1], g.n[1]);
    if (iperp == kReject) return kReject;
    return finish_bin(g, ipar, iperp, zc, izc, l1, l2);
}

// Polar branch, XiEstimator.cc:168-225.
__device__ __forceinline__ uint32_t exact_bin_polar(const GridExact &g, float d1, float l1, double r1_sq,
                                                    double two_r1_cos12, float d2, float l2, bool zc,
                                                    uint32_t izc)
{
    const double d2d = (double)d2;
    const double r_sq = r1_sq + (d2d - two_r1_cos12) * d2d;
    const double r = sqrt_no_tiny(r_sq);
    const uint32_t ir = axis_checked((r - g.min[0]) * g.inv[0], g.n[0]);
    if (ir == kReject) return kReject;
    if (r == 0) return kReject;
    const float adiff = fabsf(d1 - d2);
    const double mu = (double)adiff / r;
    const uint32_t imu = axis_checked((mu - g.min[1]) * g.inv[1], g.n[1]);
    if (imu == kReject) return kReject;
    return finish_bin(g, ir, imu, zc, izc, l1, l2);
}

// Observing branch, XiEstimator.cc:261-294 (theta bin per sightline pair).
__device__ __forceinline__ uint32_t exact_bin_obs(const GridExact &g, float l1, float l2, uint32_t ith,
                                                  bool zc, uint32_t izc)
{
    const float adl = fabsf(l1 - l2);
    const uint32_t idl = axis_checked(((double)adl - g.min[0]) * g.inv[0], g.n[0]);
    if (idl == kReject) return kReject;
    // the theta bin is not range checked (XiEstimator.cc:273,283): it can
    // carry into the axis-0 index, so then the slot comes from the flat bin
    if (zc && ith < g.n[1]) return slot3(g, idl, ith, izc);
    const uint32_t iz = zc ? izc : z_bin_exact(g, l1, l2);
    const uint32_t f = iz + (ith + idl * g.n[1]) * g.n[2];
    return f < g.ntot ? slot_of(g, f) : kReject;
}

// u = t + g: t lies within g of a bin edge iff fract(u) < 2g.
__device__ __forceinline__ bool near_edge(float u, float guard2) { return __builtin_amdgcn_fractf(u) < guard2; }

// The workgroup's LDS histogram: NSUM double arrays and one u32 count array
// of nslots entries.  FIXED: the arrays sit at compile-time offsets sized for
// kFixedSlots, so every update needs a single address computation.
constexpr uint32_t kFixedSlots = 2560;

template <int NSUM, bool FIXED>
struct Hist {
    double *hs;
    uint32_t *hc;
    uint32_t ns;      // runtime array stride (nslots) when !FIXED

    __device__ __forceinline__ Hist(double *base, uint32_t nslots)
        : hs(base), hc(reinterpret_cast<uint32_t *>(base + (size_t)(FIXED ? kFixedSlots : nslots) * NSUM)),
          ns(FIXED ? kFixedSlots : nslots)
    {
    }
    __device__ __forceinline__ uint32_t stride() const { return FIXED ? kFixedSlots : ns; }

    // n pixel pairs with summed terms into the slot at byte offset off8.
    __device__ __forceinline__ void add8n(uint32_t off8, uint32_t n, double didj, double ww) const
    {
        const uint32_t st8 = 8u * stride();
        char *bs = reinterpret_cast<char *>(hs);
        char *bc = reinterpret_cast<char *>(hc);
        __hip_atomic_fetch_add(reinterpret_cast<uint32_t *>(bc + (off8 >> 1)), n, __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_WORKGROUP);
        unsafeAtomicAdd(reinterpret_cast<double *>(bs + off8), didj);
        unsafeAtomicAdd(reinterpret_cast<double *>(bs + off8 + st8), ww);
    }

    // One pixel pair into the slot at byte offset off8 = 8 * slot of the
    // FP64 arrays (the count sits at off8 / 2 of the u32 array), so one
    // address register serves every array through immediate offsets.
    __device__ __forceinline__ void add8(uint32_t off8, float wd1, float w1, float l1, float wd2, float w2,
                                         float l2) const
    {
        const uint32_t st8 = 8u * stride();
        char *bs = reinterpret_cast<char *>(hs);
        char *bc = reinterpret_cast<char *>(hc);
        __hip_atomic_fetch_add(reinterpret_cast<uint32_t *>(bc + (off8 >> 1)), 1u, __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_WORKGROUP);
        const float didj = wd1 * wd2;
        const float ww = w1 * w2;
        unsafeAtomicAdd(reinterpret_cast<double *>(bs + off8), (double)didj);
        unsafeAtomicAdd(reinterpret_cast<double *>(bs + off8 + st8), (double)ww);
        if constexpr (NSUM == 7) {
            const float lsum = l1 + l2;
            unsafeAtomicAdd(reinterpret_cast<double *>(bs + off8 + 2 * st8), (double)wd1);
            unsafeAtomicAdd(reinterpret_cast<double *>(bs + off8 + 3 * st8), (double)wd2);
            unsafeAtomicAdd(reinterpret_cast<double *>(bs + off8 + 4 * st8), (double)w1);
            unsafeAtomicAdd(reinterpret_cast<double *>(bs + off8 + 5 * st8), (double)w2);
            unsafeAtomicAdd(reinterpret_cast<double *>(bs + off8 + 6 * st8), 0.5 * (double)lsum);
        }
    }
};

// Bytes of the histogram arrays, rounded up to 8; the workgroup's control
// words (item counter, two u64 tallies) follow them, so the histogram
// itself starts at LDS address 0 and every slot address is a plain
// register plus an immediate offset.
template <int NSUM, bool FIXED>
__host__ __device__ constexpr uint32_t hist_bytes(uint32_t nslots)
{
    return (((FIXED ? kFixedSlots : nslots) * (8u * NSUM + 4u)) + 7u) & ~7u;
}
constexpr uint32_t kCtlBytes = 24;

struct BlockCtl {
    int32_t next;
    int32_t pad;
    unsigned long long recheck;
    unsigned long long eval;
};

// Per-lane (primary pixel) state.
struct LaneState {
    float d1, l1, w1, wd1;
    float e1;             // f32(d1*d1) - d1*d1, exact
    float two_d1;         // f32(2*d1) (XiEstimator.cc:238)
    double r1_sq;         // (double)f32(d1*d1) (XiEstimator.cc:237)
    bool active;
};

// Window of a run of pixels sorted by their axis-0 key (field 1: distance,
// 2: loglam): qa = first index with key > lo, qb = first with key >= hi.
// All 64 lanes test 64 consecutive keys per pass; the run is read only
// inside [0, n).
__device__ __forceinline__ void find_window(const float4 *__restrict__ q, int32_t n, int field, double lo,
                                            double hi, int lane, int32_t &qa, int32_t &qb)
{
    qa = n;
    qb = n;
    bool have_a = false;
    for (int32_t base = 0; base < n; base += kWave) {
        const int32_t idx = base + lane;
        const bool in = idx < n;
        float key = 0.0f;
        if (in) key = (field == 1) ? q[idx].x : q[idx].y;
        const double kd = (double)key;
        const uint64_t m_lo = __ballot(in && kd > lo);
        const uint64_t m_hi = __ballot(in && kd >= hi);
        if (!have_a && m_lo) {
            qa = base + __ffsll((unsigned long long)m_lo) - 1;
            have_a = true;
        }
        if (m_hi) {
            qb = base + __ffsll((unsigned long long)m_hi) - 1;
            break;
        }
    }
}

__device__ __forceinline__ float lane_key(const float4 &v, int field) { return field == 1 ? v.x : v.y; }

// Upper end of the window of the other run of a sightline pair with cos
// separation c, when this run's keys span [kmin, kmax]: the r_par window
// kmax + win_half and, for distances, the perpendicular cut: a pixel at
// distance >= perp_cut / (2 (1 - c) kmin) makes a rejected pair with every
// pixel of this run (XiKernelArgs::perp_cut).
__device__ __forceinline__ double window_hi(const XiKernelArgs &a, double kmax, double kmin, double c)
{
    double hi = kmax + a.win_half;
    if (a.window == 1 && a.perp_cut > 0.0) {
        const double K = 2.0 * (1.0 - c);
        if (K > 0.0 && kmin > 0.0) hi = fmin(hi, a.perp_cut / (K * kmin) * (1.0 + 1e-12));
    }
    return hi;
}

// Lane <-> pixel order inside a 64-pixel chunk.  Interleaved, lane l holds
// chunk pixel 4 (l mod 16) + {0, 2, 1, 3}[l / 16]: each quarter-wave (the
// lanes an LDS access serves together) holds pixels four apart, each
// half-wave pixels two apart, so at one partner pixel their r_par bins
// differ -- fewer same-address and bank conflicts in the histogram atomics.
__device__ __forceinline__ int chunk_pixel(int lane, bool inter)
{
    if (!inter) return lane;
    const int qw = lane >> 4;
    return ((lane & 15) << 2) + (qw == 1 ? 2 : qw == 2 ? 1 : qw);
}

// Inverse of chunk_pixel: the lane holding chunk pixel q.
__device__ __forceinline__ int chunk_lane(int q, bool inter)
{
    if (!inter) return q;
    const int m = q & 3;
    return ((m == 1 ? 2 : m == 2 ? 1 : m) << 4) + (q >> 2);
}

// One z bin for a whole sightline pair?  FP32 addition, the half, the shift
// and the scale are monotone, so the bins of the extreme loglam sums bound
// the bin of every pixel pair.  Sets izc and returns true when constant.
__device__ __forceinline__ bool pair_z_const(const GridExact &ge, const LosInfo &li, const LosInfo &lj,
                                             uint32_t &izc)
{
    const float lo = li.lmin + lj.lmin;
    const float hi = li.lmax + lj.lmax;
    const double tlo = (0.5 * (double)lo - ge.min[2]) * ge.inv[2];
    const double thi = (0.5 * (double)hi - ge.min[2]) * ge.inv[2];
    if (!(tlo > -1.0 && thi < 2147483647.0)) return false;
    const int32_t z0 = (int32_t)tlo, z1 = (int32_t)thi;
    izc = (uint32_t)z0;
    return (z0 == z1) && izc < ge.n[2];
}

// One pixel pair: primary pixel (d1, l1, w1, wd1) with partner pixel
// (d2, l2, w2, wd2) of a sightline pair with cos separation c.  The FP32
// path needs d1K = f32(d1 * f32(2(1-c))) and e1; the reference arithmetic
// needs r1_sq and two_r1_cos12 (XiEstimator.cc:237-238).
// SIMPLE: both bin minima are 0 and axis 0 is exact in FP32 (Cartesian,
// Observing) -- no guard on axis 0, no offsets, no lower-bound tests.
// ZC: the whole sightline pair sits in z bin izc < n_z, so the bin is in
// range once axes 0 and 1 are.
// BAND: the histogram holds the slots [band_lo8, band_lo8 + band_len8) / 8
// only (a band of axis-0 rows of a grid too large for one LDS histogram);
// off8 is returned relative to the band and other slots are rejected.
template <int COORD, int PREC, bool SIMPLE, bool ZC, bool BAND>
__device__ __forceinline__ void pair_bin(const XiKernelArgs &a, float d1, float l1, float e1, float d1K,
                                         double r1_sq, double two_r1_cos12, float d2, float l2, uint32_t ith,
                                         uint32_t izc, bool active, float og0, float og1, uint32_t &nre,
                                         bool &hit, uint32_t &off8_out)
{
    // og0, og1: gf.offg[0..1] held in VGPRs, so that an FMA with the
    // scalar gf.inv[k] needs no per-pair copy of the constant; a SIMPLE
    // Cartesian grid has no axis-0 offset, and og0 holds -gf.inv[0]
    const GridExact &ge = a.ge;
    const GridFast &gf = a.gf;
    uint32_t off8;
    bool ok;
    // The Observing grid bins with the reference arithmetic in both modes:
    // its input |l1 - l2| is an exact FP32 value and the FP64 scaling costs
    // less than the FP32 guard test (160,000 quasars over 10,000 deg^2:
    // 386 against 405 ms).
    if constexpr (PREC == TOS_FP64 || COORD == TOS_OBSERVING) {
        uint32_t slot;
        if constexpr (COORD == TOS_CARTESIAN)
            slot = exact_bin_cart(ge, d1, l1, r1_sq, two_r1_cos12, d2, l2, ZC, izc);
        else if constexpr (COORD == TOS_POLAR)
            slot = exact_bin_polar(ge, d1, l1, r1_sq, two_r1_cos12, d2, l2, ZC, izc);
        else
            slot = exact_bin_obs(ge, l1, l2, ith, ZC, izc);
        ok = slot != kReject;
        off8 = slot << 3;
    } else {
        // u_k = t_k + g_k (GridFast); an exact axis 0 has g_0 = 0
        bool recheck;
        float u0, u1 = 0.0f;
        if constexpr (COORD == TOS_CARTESIAN) {
            const float rpar = fabsf(d1 - d2);
            // SIMPLE: inv0 is a power of two, so one rounding of
            // (d1 - d2) inv0 gives f32(d1 - d2) inv0, the value of
            // |d1 - d2| * inv0 (values below 2^-126, where the two could
            // differ, are bin 0 either way); d1 inv0 is exact and
            // loop-invariant
            u0 = SIMPLE ? fabsf(__builtin_fmaf(d2, og0, d1 * gf.inv[0]))
                        : __builtin_fmaf(rpar, gf.inv[0], og0);
            const float rp2 = __builtin_fmaf(d2, d1K, e1);
            const float arp2 = fabsf(rp2);
            const float s = __builtin_amdgcn_sqrtf(arp2);
            u1 = __builtin_fmaf(s, gf.inv[1], og1);
            recheck = near_edge(u1, gf.guard2[1]);
            // SIMPLE Cartesian grids have small_safe set: tiny r_perp^2 is bin 0.
            // r_par == 0 (equal distances) is bin 0 in both arithmetics when
            // axis 0 starts at 0, although t_0 = 0 sits on an edge
            if constexpr (!SIMPLE) {
                const bool near0 = near_edge(u0, gf.guard2[0]) & !((rpar == 0.0f) & (bool)(gf.zero_min & 1u));
                recheck |= (bool)((int)near0 | (int)(arp2 < gf.small));
            }
        } else {
            static_assert(COORD == TOS_POLAR, "the Observing grid bins with the reference arithmetic (above)");
            const float rpar = fabsf(d1 - d2);
            const float rp2 = __builtin_fmaf(d2, d1K, e1);
            const float rsq = __builtin_fmaf(rpar, rpar, rp2);
            // one reciprocal square root serves r and mu = rpar / r
            const float ar = fabsf(rsq);
            const float ri = __builtin_amdgcn_rsqf(ar);
            const float r = ar * ri;
            u0 = __builtin_fmaf(r, gf.inv[0], og0);
            const float mu = rpar * ri;
            u1 = __builtin_fmaf(mu, gf.inv[1], og1);
            // r_par == 0 (pixels at equal distance: equal wavelengths on the
            // two sightlines, ~1/n of all pairs) gives mu == 0 exactly in both
            // arithmetics, bin 0 when the mu axis starts at 0: no re-check,
            // although t_1 = 0 sits on an edge
            const bool near1 = near_edge(u1, gf.guard2[1]) & !((rpar == 0.0f) & (bool)(gf.zero_min & 2u));
            recheck = (bool)((int)near_edge(u0, gf.guard2[0]) | (int)near1 | (int)(rsq < gf.small));
        }
        ok = u0 < gf.nf[0];
        if constexpr (COORD != TOS_OBSERVING) ok &= u1 < gf.nf[1];
        if constexpr (!SIMPLE) {
            if (!(gf.nonneg & 1u)) ok &= u0 > -1.0f;
            if constexpr (COORD != TOS_OBSERVING)
                if (!(gf.nonneg & 2u)) ok &= u1 > -1.0f;
        }
        // float -> int only feeds off8, which is used only when ok
        const uint32_t i0 = (uint32_t)(int32_t)u0;
        const uint32_t i1 = (COORD == TOS_OBSERVING) ? ith : (uint32_t)(int32_t)u1;
        if (ZC && (COORD != TOS_OBSERVING || ith < ge.n[1])) {
            // 24-bit multiply-adds: every factor is < 2^24 (i0 and i1 as
            // read here may exceed their axes only when ok is false, and the
            // slot is then unused)
            off8 = mad_u24(i0, ge.row8, mad_u24(i1, ge.nz8, izc << 3));
        } else {
            const uint32_t f = (ZC ? izc : z_bin_exact(ge, l1, l2)) + (i1 + i0 * ge.n[1]) * ge.n[2];
            ok = ok && (f < ge.ntot);
            off8 = slot_of(ge, f) << 3;
        }
        if (recheck & active) {
            ++nre;
            uint32_t slot;
            if constexpr (COORD == TOS_CARTESIAN)
                slot = exact_bin_cart(ge, d1, l1, r1_sq, two_r1_cos12, d2, l2, ZC, izc);
            else
                slot = exact_bin_polar(ge, d1, l1, r1_sq, two_r1_cos12, d2, l2, ZC, izc);
            ok = slot != kReject;
            off8 = slot << 3;
        }
    }
    if constexpr (BAND) {
        off8 -= a.band_lo8;
        ok = ok && (off8 < a.band_len8);
    }
    hit = ok & active;
    off8_out = off8;
}

template <int COORD, int PREC, int NSUM, bool FIXED, bool SIMPLE, bool ZC>
__device__ __forceinline__ void pair_body(const XiKernelArgs &a, const Hist<NSUM, FIXED> &H, float d1, float l1,
                                          float w1, float wd1, float e1, float d1K, double r1_sq,
                                          double two_r1_cos12, float d2, float l2, float w2, float wd2,
                                          uint32_t ith, uint32_t izc, bool active, float og0, float og1,
                                          uint32_t &nre)
{
    bool hit;
    uint32_t off8;
    pair_bin<COORD, PREC, SIMPLE, ZC, !FIXED>(a, d1, l1, e1, d1K, r1_sq, two_r1_cos12, d2, l2, ith, izc, active, og0, og1,
                                      nre, hit, off8);
    if (hit) H.add8(off8, wd1, w1, l1, wd2, w2, l2);
}

// Two pixel pairs of one lane (consecutive pixels of the looped run) whose
// bins are known: when both land in one slot, one set of atomics carries
// both (count 2, summed terms), halving the lanes those atomics serve.
// The per-pair terms stay the reference's FP32 products.
template <int NSUM, bool FIXED>
__device__ __forceinline__ void add_merged(const Hist<NSUM, FIXED> &H, bool ha, uint32_t oa, float dda, float wwa,
                                           bool hb, uint32_t ob, float ddb, float wwb)
{
    static_assert(NSUM == 2, "merging carries the count and the two core sums only");
    const bool same = ha & hb & (oa == ob);
    const double da = (double)dda, wa = (double)wwa, db = (double)ddb, wb = (double)wwb;
    // fma(1, b, a) = a + b (one rounding) and fma(0, b, a) = a: one select
    // of the factor instead of selects of both sums
    const double f = same ? 1.0 : 0.0;
    if (ha) H.add8n(oa, same ? 2u : 1u, __builtin_fma(f, db, da), __builtin_fma(f, wb, wa));
    if (hb & !same) H.add8n(ob, 1u, db, wb);
}

// Walks n pixel records through `fn` four at a time, issuing the next four
// scalar loads before the current four are processed.  Reads only [0, n).
template <typename F>
__device__ __forceinline__ void stream_pixels(const float4 *__restrict__ q, int32_t n, F &&fn)
{
    int32_t k = 0;
    if (n >= 4) {
        float4 A0 = q[0], A1 = q[1], A2 = q[2], A3 = q[3];
        for (; k + 8 <= n; k += 4) {
            const float4 B0 = q[k + 4], B1 = q[k + 5], B2 = q[k + 6], B3 = q[k + 7];
            fn(A0);
            fn(A1);
            fn(A2);
            fn(A3);
            A0 = B0;
            A1 = B1;
            A2 = B2;
            A3 = B3;
        }
        fn(A0);
        fn(A1);
        fn(A2);
        fn(A3);
        k += 4;
    }
    for (; k < n; ++k) fn(q[k]);
}

// The same walk handing consecutive records to fn2 in pairs (a trailing
// odd record goes to fn1).
template <typename F1, typename F2>
__device__ __forceinline__ void stream_pixel_pairs(const float4 *__restrict__ q, int32_t n, F1 &&fn1, F2 &&fn2)
{
    int32_t k = 0;
    if (n >= 4) {
        float4 A0 = q[0], A1 = q[1], A2 = q[2], A3 = q[3];
        for (; k + 8 <= n; k += 4) {
            const float4 B0 = q[k + 4], B1 = q[k + 5], B2 = q[k + 6], B3 = q[k + 7];
            fn2(A0, A1);
            fn2(A2, A3);
            A0 = B0;
            A1 = B1;
            A2 = B2;
            A3 = B3;
        }
        fn2(A0, A1);
        fn2(A2, A3);
        k += 4;
    }
    for (; k + 2 <= n; k += 2) fn2(q[k], q[k + 1]);
    if (k < n) fn1(q[k]);
}

// The same walk handing four consecutive records to fn4 (a trailing pair
// goes to fn2, a trailing single record to fn1).
template <typename F1, typename F2, typename F4>
__device__ __forceinline__ void stream_pixel_quads(const float4 *__restrict__ q, int32_t n, F1 &&fn1, F2 &&fn2,
                                                   F4 &&fn4)
{
    int32_t k = 0;
    if (n >= 4) {
        float4 A0 = q[0], A1 = q[1], A2 = q[2], A3 = q[3];
        for (; k + 8 <= n; k += 4) {
            const float4 B0 = q[k + 4], B1 = q[k + 5], B2 = q[k + 6], B3 = q[k + 7];
            fn4(A0, A1, A2, A3);
            A0 = B0;
            A1 = B1;
            A2 = B2;
            A3 = B3;
        }
        fn4(A0, A1, A2, A3);
        k += 4;
    }
    for (; k + 2 <= n; k += 2) fn2(q[k], q[k + 1]);
    if (k < n) fn1(q[k]);
}

// Four consecutive looped pixels of a lane: runs of equal slots carry one
// set of atomics each (count = run length, summed FP64 terms).
template <int NSUM, bool FIXED>
__device__ __forceinline__ void add_merged4(const Hist<NSUM, FIXED> &H, const bool (&h)[4], const uint32_t (&o)[4],
                                            const double (&d)[4], const double (&w)[4])
{
    double D = d[0], W = w[0];
    uint32_t C = 1;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        if (k > 0) {
            const bool cont = h[k - 1] & h[k] & (o[k - 1] == o[k]);
            // fma(1, D, d) = D + d (one rounding), fma(0, D, d) = d
            const double f = cont ? 1.0 : 0.0;
            D = __builtin_fma(f, D, d[k]);
            W = __builtin_fma(f, W, w[k]);
            C = cont ? C + 1u : 1u;
        }
        const bool last = (k == 3) || !(h[k] & h[k + 1] & (o[k] == o[k + 1]));
        if (h[k] & last) H.add8n(o[k], C, D, W);
    }
}

// Merge consecutive looped pixels of a lane that share a slot into one set
// of atomics (runs among four on normal items, pairs on transposed tails):
// FP32C, and FP64 on the Observing grid (both held by the LDS atomics; the
// FP64 Cartesian and Polar kernels are held by arithmetic and run faster
// without the merge).
template <int PREC, int NSUM, int COORD>
constexpr bool kMergePairs = (PREC == TOS_FP32C || COORD == TOS_OBSERVING) && NSUM == 2;

// Normal item: the lanes hold 64 primary pixels; every partner pixel is
// wave-uniform.
template <int COORD, int PREC, int NSUM, bool FIXED, bool SIMPLE, bool ZC>
__device__ __forceinline__ void partner_pixels(const XiKernelArgs &a, const Hist<NSUM, FIXED> &H,
                                               const LaneState &L, const float4 *__restrict__ q2, int32_t n2,
                                               double two_r1_cos12, float d1K, uint32_t ith, uint32_t izc_s,
                                               float og0, float og1, uint32_t &nre)
{
    // the constant z bin as a per-lane value, so that the slot's two
    // multiply-adds each read one scalar operand
    uint32_t izc;
    asm("v_mov_b32 %0, %1" : "=v"(izc) : "s"(izc_s));
    auto one = [&](const float4 Q) {
        pair_body<COORD, PREC, NSUM, FIXED, SIMPLE, ZC>(a, H, L.d1, L.l1, L.w1, L.wd1, L.e1, d1K, L.r1_sq,
                                                         two_r1_cos12, Q.x, Q.y, Q.z, Q.w, ith, izc, L.active,
                                                         og0, og1, nre);
    };
    if constexpr (kMergePairs<PREC, NSUM, COORD>) {
        auto two = [&](const float4 Qa, const float4 Qb) {
            bool ha, hb;
            uint32_t oa, ob;
            pair_bin<COORD, PREC, SIMPLE, ZC, !FIXED>(a, L.d1, L.l1, L.e1, d1K, L.r1_sq, two_r1_cos12, Qa.x, Qa.y, ith, izc,
                                              L.active, og0, og1, nre, ha, oa);
            pair_bin<COORD, PREC, SIMPLE, ZC, !FIXED>(a, L.d1, L.l1, L.e1, d1K, L.r1_sq, two_r1_cos12, Qb.x, Qb.y, ith, izc,
                                              L.active, og0, og1, nre, hb, ob);
            add_merged(H, ha, oa, L.wd1 * Qa.w, L.w1 * Qa.z, hb, ob, L.wd1 * Qb.w, L.w1 * Qb.z);
        };
        stream_pixel_quads(q2, n2, one, two, [&](const float4 Q0, const float4 Q1, const float4 Q2, const float4 Q3) {
            bool h[4];
            uint32_t o[4];
            double dd[4], ww[4];
            const float4 Q[4] = {Q0, Q1, Q2, Q3};
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                pair_bin<COORD, PREC, SIMPLE, ZC, !FIXED>(a, L.d1, L.l1, L.e1, d1K, L.r1_sq, two_r1_cos12, Q[k].x, Q[k].y,
                                                  ith, izc, L.active, og0, og1, nre, h[k], o[k]);
                dd[k] = (double)(L.wd1 * Q[k].w);
                ww[k] = (double)(L.w1 * Q[k].z);
            }
            add_merged4(H, h, o, dd, ww);
        });
    } else {
        stream_pixels(q2, n2, one);
    }
}

// Transposed item: the lanes hold 64 partner pixels (Q); the primary pixels
// [0, n1) of p1s are wave-uniform.
template <int COORD, int PREC, int NSUM, bool FIXED, bool SIMPLE, bool ZC>
__device__ __forceinline__ void primary_pixels(const XiKernelArgs &a, const Hist<NSUM, FIXED> &H,
                                               const float4 *__restrict__ p1s, int32_t n1, const float4 Q,
                                               bool active2, double c, float K, uint32_t ith, uint32_t izc_s,
                                               float og0, float og1, uint32_t &nre)
{
    uint32_t izc;
    asm("v_mov_b32 %0, %1" : "=v"(izc) : "s"(izc_s));
    auto one = [&](const float4 P) {
        const float d1 = P.x;
        const float d1sq = d1 * d1;
        const float e1 = -__builtin_fmaf(d1, d1, -d1sq);
        const float two_d1 = 2.0f * d1;
        pair_body<COORD, PREC, NSUM, FIXED, SIMPLE, ZC>(a, H, d1, P.y, P.z, P.w, e1, d1 * K, (double)d1sq,
                                                         (double)two_d1 * c, Q.x, Q.y, Q.z, Q.w, ith, izc,
                                                         active2, og0, og1, nre);
    };
    if constexpr (kMergePairs<PREC, NSUM, COORD>) {
        auto bin = [&](const float4 P, bool &h, uint32_t &o) {
            const float d1 = P.x;
            const float d1sq = d1 * d1;
            const float e1 = -__builtin_fmaf(d1, d1, -d1sq);
            const float two_d1 = 2.0f * d1;
            pair_bin<COORD, PREC, SIMPLE, ZC, !FIXED>(a, d1, P.y, e1, d1 * K, (double)d1sq, (double)two_d1 * c, Q.x, Q.y,
                                              ith, izc, active2, og0, og1, nre, h, o);
        };
        stream_pixel_pairs(p1s, n1, one, [&](const float4 Pa, const float4 Pb) {
            bool ha, hb;
            uint32_t oa, ob;
            bin(Pa, ha, oa);
            bin(Pb, hb, ob);
            add_merged(H, ha, oa, Pa.w * Q.w, Pa.z * Q.z, hb, ob, Pb.w * Q.w, Pb.z * Q.z);
        });
    } else {
        stream_pixels(p1s, n1, one);
    }
}

// One workgroup's share of a pass: items [it_begin, it_end) of job `a`,
// its partial histogram written at row `part_row` of a.part_cnt/part_sum.
// COORD: tos_coord; PREC: tos_precision; NSUM: 2 or 7.
template <int COORD, int PREC, int NSUM, bool FIXED, bool SIMPLE>
__device__ __forceinline__ void xi_block(const XiKernelArgs &a, int32_t it_begin, int32_t it_end, size_t part_row)
{
    const GridExact &ge = a.ge;
    double *lds_raw;
    if constexpr (FIXED && NSUM == 2) {
        // fixed layout: a static array, whose address is a compile-time
        // constant that folds into the LDS instructions' offsets
        __shared__ __attribute__((aligned(16))) double lds_fixed[(hist_bytes<2, true>(0) + kCtlBytes) / 8];
        lds_raw = lds_fixed;
    } else {
        extern __shared__ double lds_dyn[];
        lds_raw = lds_dyn;
    }
    // a dynamic histogram holds the launch's band of slots (the whole grid
    // unless it is too large for one LDS histogram)
    const uint32_t nslots = FIXED ? ge.nslots : a.band_len8 / 8u;
    const Hist<NSUM, FIXED> H(lds_raw, nslots);
    BlockCtl &ctl = *reinterpret_cast<BlockCtl *>(reinterpret_cast<char *>(lds_raw) +
                                                  hist_bytes<NSUM, FIXED>(nslots));
    int32_t &s_next = ctl.next;
    unsigned long long &s_recheck = ctl.recheck;
    unsigned long long &s_eval = ctl.eval;

    {
        const int ns = (int)H.stride();
        for (int b = threadIdx.x; b < ns * NSUM; b += blockDim.x) H.hs[b] = 0.0;
        for (int b = threadIdx.x; b < ns; b += blockDim.x) H.hc[b] = 0u;
    }
    if (threadIdx.x == 0) {
        s_next = it_begin;
        s_recheck = 0ull;
        s_eval = 0ull;
    }
    __syncthreads();

    const int lane = threadIdx.x & (kWave - 1);
    uint32_t nre = 0;
    uint64_t n_eval = 0;   // wave-uniform
    // the guard offsets as per-lane values (opaque to the compiler, so they
    // stay in VGPRs instead of being copied from SGPRs at every use); a
    // SIMPLE Cartesian grid's axis 0 has none and takes -inv0 instead
    float og0, og1;
    const float og0_s = (COORD == TOS_CARTESIAN && SIMPLE) ? -a.gf.inv[0] : a.gf.offg[0];
    asm("v_mov_b32 %0, %1" : "=v"(og0) : "s"(og0_s));
    asm("v_mov_b32 %0, %1" : "=v"(og1) : "s"(a.gf.offg[1]));

    for (;;) {
        int32_t k = 0;
        if (lane == 0) k = __hip_atomic_fetch_add(&s_next, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        k = __builtin_amdgcn_readfirstlane(k);
        if (k >= it_end) break;
        const XiItem it = a.items[k];
        const LosInfo li = a.los[it.los];
        if (it.p1_begin & kTransposed) {
            const int32_t p1a = it.p1_begin & ~kTransposed;
            const float4 *p1s = a.pix + li.off + p1a;
            const int32_t n1 = li.n - p1a;
            for (int32_t pr = it.pair_begin; pr < it.pair_end; ++pr) {
                const int32_t j = a.other[pr];
                const double c = a.csep[pr];
                const LosInfo lj = a.los[j];
                uint32_t ith = 0;
                if constexpr (COORD == TOS_OBSERVING) ith = a.theta_bin[pr];
                uint32_t izc = 0;
                const bool zc = a.zconst && pair_z_const(ge, li, lj, izc);
                const float K = (float)(2.0 * (1.0 - c));
                for (int32_t pc = 0; pc < lj.n; pc += kWave) {
                    const int32_t p2 = pc + chunk_pixel(lane, a.interleave);
                    const bool active2 = p2 < lj.n;
                    const int32_t n_act = min(kWave, lj.n - pc);
                    float4 Q = make_float4(0.f, 0.f, 0.f, 0.f);
                    if (active2) Q = a.pix[lj.off + p2];
                    int32_t qa = 0, qb = n1;
                    if (a.window) {
                        // partner keys ascend over the chunk: pixel 0 (lane 0)
                        // is the smallest, pixel n_act-1 the largest
                        const float kmin =
                            __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(lane_key(Q, a.window))));
                        const float kmax = __int_as_float(__builtin_amdgcn_readlane(
                            __float_as_int(lane_key(Q, a.window)), chunk_lane(n_act - 1, a.interleave)));
                        find_window(p1s, n1, a.window, (double)kmin - a.win_half,
                                    window_hi(a, (double)kmax, (double)kmin, c), lane,
                                    qa, qb);
                        if (qa >= qb) continue;
                    }
                    n_eval += (uint64_t)(qb - qa) * (uint64_t)n_act;
                    if (zc)
                        primary_pixels<COORD, PREC, NSUM, FIXED, SIMPLE, true>(a, H, p1s + qa, qb - qa, Q, active2, c,
                                                                                K, ith, izc, og0, og1, nre);
                    else
                        primary_pixels<COORD, PREC, NSUM, FIXED, SIMPLE, false>(a, H, p1s + qa, qb - qa, Q, active2,
                                                                                 c, K, ith, izc, og0, og1, nre);
                }
            }
            continue;
        }
        const int p1 = it.p1_begin + chunk_pixel(lane, a.interleave);
        LaneState L;
        L.active = p1 < li.n;
        float4 P = make_float4(0.f, 0.f, 0.f, 0.f);
        if (L.active) P = a.pix[li.off + p1];
        L.d1 = P.x;
        L.l1 = P.y;
        L.w1 = P.z;
        L.wd1 = P.w;
        const float d1sq = L.d1 * L.d1;
        L.r1_sq = (double)d1sq;
        L.two_d1 = 2.0f * L.d1;
        L.e1 = -__builtin_fmaf(L.d1, L.d1, -d1sq);
        const int32_t n_act = min(kWave, li.n - it.p1_begin);
        float kmin = 0.0f, kmax = 0.0f;
        if (a.window) {
            // primary keys ascend over the chunk
            kmin = __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(lane_key(P, a.window))));
            kmax = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(lane_key(P, a.window)),
                                                            chunk_lane(n_act - 1, a.interleave)));
        }

        for (int32_t pr = it.pair_begin; pr < it.pair_end; ++pr) {
            const int32_t j = a.other[pr];
            const double c = a.csep[pr];
            const LosInfo lj = a.los[j];
            const double two_r1_cos12 = (double)L.two_d1 * c;
            uint32_t ith = 0;
            if constexpr (COORD == TOS_OBSERVING) ith = a.theta_bin[pr];
            uint32_t izc = 0;
            const bool zc = a.zconst && pair_z_const(ge, li, lj, izc);
            const float K = (float)(2.0 * (1.0 - c));
            const float d1K = L.d1 * K;
            const float4 *q2 = a.pix + lj.off;
            int32_t qa = 0, qb = lj.n;
            if (a.window) {
                find_window(q2, lj.n, a.window, (double)kmin - a.win_half, window_hi(a, (double)kmax, (double)kmin, c),
                            lane, qa, qb);
                if (qa >= qb) continue;
            }
            n_eval += (uint64_t)(qb - qa) * (uint64_t)n_act;
            if (zc)
                partner_pixels<COORD, PREC, NSUM, FIXED, SIMPLE, true>(a, H, L, q2 + qa, qb - qa, two_r1_cos12, d1K,
                                                                        ith, izc, og0, og1, nre);
            else
                partner_pixels<COORD, PREC, NSUM, FIXED, SIMPLE, false>(a, H, L, q2 + qa, qb - qa, two_r1_cos12,
                                                                         d1K, ith, izc, og0, og1, nre);
        }
    }

    if constexpr (PREC == TOS_FP32C) {
        if (nre) __hip_atomic_fetch_add(&s_recheck, (unsigned long long)nre, __ATOMIC_RELAXED,
                                        __HIP_MEMORY_SCOPE_WORKGROUP);
    }
    if (lane == 0 && n_eval) __hip_atomic_fetch_add(&s_eval, (unsigned long long)n_eval, __ATOMIC_RELAXED,
                                                    __HIP_MEMORY_SCOPE_WORKGROUP);
    __syncthreads();
    const size_t base = part_row * (size_t)a.nbins;
    const uint32_t st = H.stride();
    // the band's bins: flat bins band_bin0 + [0, a.nbins)
    const uint32_t band_slot0 = FIXED ? 0u : a.band_lo8 / 8u;
    for (int b = threadIdx.x; b < a.nbins; b += blockDim.x) {
        const uint32_t sl = slot_of(ge, (uint32_t)((FIXED ? 0 : a.band_bin0) + b)) - band_slot0;
        a.part_cnt[base + b] = H.hc[sl];
#pragma unroll
        for (int k = 0; k < NSUM; ++k) a.part_sum[(base + b) * NSUM + k] = H.hs[(size_t)k * st + sl];
    }
    if (threadIdx.x == 0) {
        if (s_recheck) __hip_atomic_fetch_add(&a.n_recheck[0], s_recheck, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (s_eval) __hip_atomic_fetch_add(&a.n_recheck[1], s_eval, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
}

// The pixel-pair kernel of one job: workgroup w takes the item range
// wg_item_ptr[w] .. wg_item_ptr[w + 1].
template <int COORD, int PREC, int NSUM, bool FIXED, bool SIMPLE>
__global__ void __launch_bounds__(1024) xi_pairs_kernel(XiKernelArgs a)
{
    xi_block<COORD, PREC, NSUM, FIXED, SIMPLE>(a, a.wg_item_ptr[blockIdx.x], a.wg_item_ptr[blockIdx.x + 1],
                                               blockIdx.x);
}

// Many jobs in one launch: workgroup w serves job b.wg[w].job, items
// [begin, end) of that job's item list, partial row `row` of that job.
template <int COORD, int PREC, int NSUM, bool FIXED, bool SIMPLE>
__global__ void __launch_bounds__(1024) xi_pairs_batch_kernel(XiBatchArgs b)
{
    const XiWgTask t = b.wg[blockIdx.x];
    xi_block<COORD, PREC, NSUM, FIXED, SIMPLE>(b.jobs[t.job], t.begin, t.end, (size_t)t.row);
}

// Deterministic reduction of per-workgroup partials: stage 1 sums a fixed
// segment of workgroups per (bin, field), stage 2 sums the segments in order.
__global__ void reduce_stage1(const uint32_t *part_cnt, const double *part_sum, int n_wg, int nbins, int nsum,
                              int n_seg, uint32_t *seg_cnt_lo, double *seg_sum)
{
    const int64_t nfield = (int64_t)nbins * (nsum + 1);
    const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int seg = blockIdx.y;
    if (idx >= nfield) return;
    const int per = (n_wg + n_seg - 1) / n_seg;
    const int w0 = seg * per;
    const int w1 = min(n_wg, w0 + per);
    if (idx < nbins) {
        // counts: 64-bit sum kept as two u32 words (lo, hi)
        uint64_t acc = 0;
        for (int w = w0; w < w1; ++w) acc += part_cnt[(size_t)w * nbins + idx];
        seg_cnt_lo[((size_t)seg * nbins + idx) * 2 + 0] = (uint32_t)acc;
        seg_cnt_lo[((size_t)seg * nbins + idx) * 2 + 1] = (uint32_t)(acc >> 32);
    } else {
        const int64_t f = idx - nbins;   // bin * nsum + s
        double acc = 0.0;
        for (int w = w0; w < w1; ++w) acc += part_sum[(size_t)w * nbins * nsum + f];
        seg_sum[(size_t)seg * nbins * nsum + f] = acc;
    }
}

__global__ void reduce_stage2(const uint32_t *seg_cnt, const double *seg_sum, int nbins, int nsum, int n_seg,
                              uint64_t *out_cnt, double *out_sum)
{
    const int64_t nfield = (int64_t)nbins * (nsum + 1);
    const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= nfield) return;
    if (idx < nbins) {
        uint64_t acc = 0;
        for (int s = 0; s < n_seg; ++s) {
            const size_t o = ((size_t)s * nbins + idx) * 2;
            acc += (uint64_t)seg_cnt[o] | ((uint64_t)seg_cnt[o + 1] << 32);
        }
        out_cnt[idx] = acc;
    } else {
        const int64_t f = idx - nbins;
        double acc = 0.0;
        for (int s = 0; s < n_seg; ++s) acc += seg_sum[(size_t)s * nbins * nsum + f];
        out_sum[f] = acc;
    }
}

// Batched reduction: blockIdx.y is the job; each thread sums one field of
// one bin over the job's workgroups in workgroup order.
__global__ void reduce_batch(const XiReduceJob *jobs)
{
    const XiReduceJob J = jobs[blockIdx.y];
    const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t nb = J.nbins;
    if (idx >= nb * (J.nsum + 1)) return;
    if (idx < nb) {
        uint64_t acc = 0;
        for (int w = 0; w < J.n_wg; ++w) acc += J.part_cnt[(size_t)w * nb + idx];
        J.out_cnt[idx] = acc;
    } else {
        const int64_t f = idx - nb;
        const size_t stride = (size_t)nb * J.nsum;
        double acc = 0.0;
        for (int w = 0; w < J.n_wg; ++w) acc += J.part_sum[(size_t)w * stride + f];
        J.out_sum[f] = acc;
    }
}

__global__ void zero_u64(unsigned long long *p, int64_t n)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0ull;
}

__global__ void selftest_kernel(const double *a, const double *b, double *s, double *q, uint32_t *c, int64_t n)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    s[i] = sqrt_no_tiny(a[i]);
    q[i] = a[i] / b[i];
    c[i] = axis_unchecked(a[i]);
}

// Dynamic LDS bytes of a variant: none when its histogram is static.
template <int S, bool FIXED>
size_t dynamic_lds(size_t lds)
{
    return (FIXED && S == 2) ? 0 : lds;
}

template <int C, int P, int S, bool FIXED, bool SIMPLE>
hipError_t launch_one(const XiKernelArgs &a, int n_wg, int block, size_t lds, hipStream_t st)
{
    auto fn = xi_pairs_kernel<C, P, S, FIXED, SIMPLE>;
    const size_t dyn = dynamic_lds<S, FIXED>(lds);
    if (dyn > 0) {
        hipError_t e = hipFuncSetAttribute((const void *)fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)dyn);
        if (e != hipSuccess) return e;
    }
    hipLaunchKernelGGL(fn, dim3(n_wg), dim3(block), dyn, st, a);
    return hipGetLastError();
}

// SIMPLE grids: both minima 0 and axis 0 exact in FP32 (Cartesian,
// Observing) -- see partner_pixels.
inline bool simple_grid(int coord, const GridExact &ge, const GridFast &gf)
{
    if (ge.min[0] != 0.0 || (coord != TOS_OBSERVING && ge.min[1] != 0.0)) return false;
    if (coord == TOS_POLAR) return true;
    if (coord == TOS_CARTESIAN && !gf.small_safe) return false;
    return gf.exact0 != 0;
}

template <int C, int P, int S>
hipError_t launch_fs(bool fixed, bool simple, const XiKernelArgs &a, int n_wg, int block, size_t lds, hipStream_t st)
{
    if (fixed) return simple ? launch_one<C, P, S, true, true>(a, n_wg, block, lds, st)
                             : launch_one<C, P, S, true, false>(a, n_wg, block, lds, st);
    return simple ? launch_one<C, P, S, false, true>(a, n_wg, block, lds, st)
                  : launch_one<C, P, S, false, false>(a, n_wg, block, lds, st);
}

template <int C, int P>
hipError_t launch_acc(int accum, const XiKernelArgs &a, int n_wg, int block, size_t lds, hipStream_t st)
{
    const bool simple = simple_grid(C, a.ge, a.gf);
    const bool fixed = !a.banded && a.ge.nslots <= kFixedSlots;
    if (accum == TOS_ACC_ALL) return launch_fs<C, P, 7>(fixed, simple, a, n_wg, block, lds, st);
    return launch_fs<C, P, 2>(fixed, simple, a, n_wg, block, lds, st);
}

template <int C>
hipError_t launch_prec(int prec, int accum, const XiKernelArgs &a, int n_wg, int block, size_t lds, hipStream_t st)
{
    if (prec == TOS_FP32C) return launch_acc<C, TOS_FP32C>(accum, a, n_wg, block, lds, st);
    return launch_acc<C, TOS_FP64>(accum, a, n_wg, block, lds, st);
}

// n_wg == 0: only raise the kernel's dynamic LDS limit (done before any
// graph capture, so that the captured sequence holds launches only).
template <int C, int P, bool FIXED, bool SIMPLE>
hipError_t launch_batch_one(const XiBatchArgs &b, int n_wg, int block, size_t lds, hipStream_t st)
{
    auto fn = xi_pairs_batch_kernel<C, P, 2, FIXED, SIMPLE>;
    const size_t dyn = dynamic_lds<2, FIXED>(lds);
    if (n_wg == 0)
        return dyn > 0 ? hipFuncSetAttribute((const void *)fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)dyn)
                       : hipSuccess;
    hipLaunchKernelGGL(fn, dim3(n_wg), dim3(block), dyn, st, b);
    return hipGetLastError();
}

template <int C, int P>
hipError_t launch_batch_fs(bool fixed, bool simple, const XiBatchArgs &b, int n_wg, int block, size_t lds,
                           hipStream_t st)
{
    if (fixed) return simple ? launch_batch_one<C, P, true, true>(b, n_wg, block, lds, st)
                             : launch_batch_one<C, P, true, false>(b, n_wg, block, lds, st);
    return simple ? launch_batch_one<C, P, false, true>(b, n_wg, block, lds, st)
                  : launch_batch_one<C, P, false, false>(b, n_wg, block, lds, st);
}

template <int C>
hipError_t launch_batch_prec(int prec, bool fixed, bool simple, const XiBatchArgs &b, int n_wg, int block,
                             size_t lds, hipStream_t st)
{
    if (prec == TOS_FP32C) return launch_batch_fs<C, TOS_FP32C>(fixed, simple, b, n_wg, block, lds, st);
    return launch_batch_fs<C, TOS_FP64>(fixed, simple, b, n_wg, block, lds, st);
}

}  // namespace

void xi_variant(int coord, const GridExact &ge, const GridFast &gf, bool &fixed, bool &simple)
{
    fixed = ge.nslots <= kFixedSlots;
    simple = simple_grid(coord, ge, gf);
}

int launch_xi_batch(int coord, int precision, bool fixed, bool simple, const XiBatchArgs &b, int n_wg, int block,
                    size_t lds_bytes, void *stream)
{
    hipStream_t st = (hipStream_t)stream;
    switch (coord) {
    case TOS_CARTESIAN:
        return (int)launch_batch_prec<TOS_CARTESIAN>(precision, fixed, simple, b, n_wg, block, lds_bytes, st);
    case TOS_POLAR: return (int)launch_batch_prec<TOS_POLAR>(precision, fixed, simple, b, n_wg, block, lds_bytes, st);
    default: return (int)launch_batch_prec<TOS_OBSERVING>(precision, fixed, simple, b, n_wg, block, lds_bytes, st);
    }
}

int launch_reduce_batch(const XiReduceJob *jobs_dev, int n_jobs, int64_t max_fields, void *stream)
{
    hipStream_t st = (hipStream_t)stream;
    const int threads = 256;
    const int blocks = (int)((max_fields + threads - 1) / threads);
    if (n_jobs <= 0 || blocks <= 0) return 0;
    hipLaunchKernelGGL(reduce_batch, dim3(blocks, n_jobs), dim3(threads), 0, st, jobs_dev);
    return (int)hipGetLastError();
}

int launch_zero_u64(unsigned long long *p, int64_t n, void *stream)
{
    hipStream_t st = (hipStream_t)stream;
    const int threads = 256;
    const int blocks = (int)((n + threads - 1) / threads);
    if (blocks <= 0) return 0;
    hipLaunchKernelGGL(zero_u64, dim3(blocks), dim3(threads), 0, st, p, n);
    return (int)hipGetLastError();
}

int xi_nsum(int accum) { return accum == TOS_ACC_ALL ? 7 : 2; }

size_t xi_lds_bytes(int accum, int nslots)
{
    const bool fixed = (uint32_t)nslots <= kFixedSlots;
    if (accum == TOS_ACC_ALL)
        return (fixed ? hist_bytes<7, true>(nslots) : hist_bytes<7, false>(nslots)) + kCtlBytes;
    return (fixed ? hist_bytes<2, true>(nslots) : hist_bytes<2, false>(nslots)) + kCtlBytes;
}

int launch_xi(int coord, int precision, int accum, const XiKernelArgs &a, int n_wg, int block, size_t lds_bytes,
              void *stream)
{
    hipStream_t st = (hipStream_t)stream;
    switch (coord) {
    case TOS_CARTESIAN: return (int)launch_prec<TOS_CARTESIAN>(precision, accum, a, n_wg, block, lds_bytes, st);
    case TOS_POLAR: return (int)launch_prec<TOS_POLAR>(precision, accum, a, n_wg, block, lds_bytes, st);
    default: return (int)launch_prec<TOS_OBSERVING>(precision, accum, a, n_wg, block, lds_bytes, st);
    }
}

int launch_reduce(const uint32_t *part_cnt, const double *part_sum, int n_wg, int nbins, int nsum,
                  uint64_t *out_cnt, double *out_sum, uint32_t *scratch_cnt, double *scratch_sum, int n_seg,
                  void *stream)
{
    hipStream_t st = (hipStream_t)stream;
    const int64_t nfield = (int64_t)nbins * (nsum + 1);
    const int threads = 256;
    const int blocks = (int)((nfield + threads - 1) / threads);
    hipLaunchKernelGGL(reduce_stage1, dim3(blocks, n_seg), dim3(threads), 0, st, part_cnt, part_sum, n_wg, nbins,
                       nsum, n_seg, scratch_cnt, scratch_sum);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return (int)e;
    hipLaunchKernelGGL(reduce_stage2, dim3(blocks), dim3(threads), 0, st, scratch_cnt, scratch_sum, nbins, nsum,
                       n_seg, out_cnt, out_sum);
    return (int)hipGetLastError();
}

int launch_selftest(const double *a, const double *b, double *s, double *q, uint32_t *c, int64_t n, void *stream)
{
    hipStream_t st = (hipStream_t)stream;
    const int threads = 256;
    const int blocks = (int)((n + threads - 1) / threads);
    if (blocks > 0) hipLaunchKernelGGL(selftest_kernel, dim3(blocks), dim3(threads), 0, st, a, b, s, q, c, n);
    return (int)hipGetLastError();
}

int xi_max_lds(int device)
{
    int v = 0;
    if (hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, device) != hipSuccess) return 65536;
    return v;
}

}  // namespace tos
