// tos_host.cpp -- host-side algorithms of the estimator: HEALPix RING
// pixelisation, candidate sightline-pair enumeration over sky bins, work
// item construction, grid constants and the LCDM distance transform.
#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstring>
#include <map>
#include <thread>
#include <unordered_map>

#include "tos_cosmo.h"
#include "tos_internal.h"
#include "tos_parallel.h"

namespace tos {

static const double kPi = 3.141592653589793238462643383279502884;

uint32_t cvt_x86(double x)
{
    if (!(x >= -9223372036854775808.0 && x < 9223372036854775808.0)) return 0u;
    return (uint32_t)(uint64_t)(int64_t)x;
}

// ---------------------------------------------------------------------------
// HEALPix RING scheme (Gorski et al. 2005): the equatorial belt |z| <= 2/3 is
// indexed by the two families of pixel edge lines, the caps by the distance
// to the pole.  HealpixBins uses RING ordering (HealpixBins.h:23-26).
// ---------------------------------------------------------------------------
int64_t healpix_ang2pix_ring(int64_t nside, double theta, double phi)
{
    const double z = std::cos(theta);
    const double za = std::fabs(z);
    double ph = std::fmod(phi, 2.0 * kPi);
    if (ph < 0) ph += 2.0 * kPi;
    double tt = ph / (0.5 * kPi);
    if (tt >= 4.0) tt = 0.0;
    const int64_t npix = 12 * nside * nside;
    const int64_t ncap = 2 * nside * (nside - 1);
    if (za <= 2.0 / 3.0) {
        const int64_t nl4 = 4 * nside;
        const double u = (double)nside * (0.5 + tt);
        const double v = (double)nside * z * 0.75;
        const int64_t up = (int64_t)(u - v);
        const int64_t dn = (int64_t)(u + v);
        const int64_t ring = nside + 1 + up - dn;
        const int64_t shift = 1 - (ring & 1);
        int64_t ip = (up + dn - nside + shift + 1) / 2;
        ip %= nl4;
        if (ip < 0) ip += nl4;
        return ncap + (ring - 1) * nl4 + ip;
    }
    const double frac = tt - std::floor(tt);
    const double w = (double)nside * std::sqrt(3.0 * (1.0 - za));
    const int64_t a = (int64_t)(frac * w);
    const int64_t b = (int64_t)((1.0 - frac) * w);
    const int64_t ring = a + b + 1;
    int64_t ip = (int64_t)(tt * (double)ring);
    if (ip >= 4 * ring) ip -= 4 * ring;
    return z > 0 ? 2 * ring * (ring - 1) + ip : npix - 2 * ring * (ring + 1) + ip;
}

static int64_t isqrt_i64(int64_t v)
{
    int64_t r = (int64_t)std::sqrt((double)v);
    while (r * r > v) --r;
    while ((r + 1) * (r + 1) <= v) ++r;
    return r;
}

void healpix_pix2ang_ring(int64_t nside, int64_t pix, double *theta, double *phi)
{
    const int64_t npix = 12 * nside * nside;
    const int64_t ncap = 2 * nside * (nside - 1);
    const double fn = (double)nside;
    double z, ph;
    if (pix < ncap) {
        const int64_t ring = (1 + isqrt_i64(1 + 2 * pix)) / 2;
        const int64_t k = pix + 1 - 2 * ring * (ring - 1);
        z = 1.0 - (double)(ring * ring) / (3.0 * fn * fn);
        ph = ((double)k - 0.5) * kPi / (2.0 * (double)ring);
    } else if (pix < npix - ncap) {
        const int64_t p = pix - ncap;
        const int64_t ring = p / (4 * nside) + nside;
        const int64_t k = p % (4 * nside) + 1;
        const double shift = ((ring + nside) & 1) ? 1.0 : 0.5;
        z = (double)(2 * nside - ring) * 2.0 / (3.0 * fn);
        ph = ((double)k - shift) * kPi / (2.0 * fn);
    } else {
        const int64_t p = npix - pix;
        const int64_t ring = (1 + isqrt_i64(2 * p - 1)) / 2;
        const int64_t k = 4 * ring + 1 - (p - 2 * ring * (ring - 1));
        z = -1.0 + (double)(ring * ring) / (3.0 * fn * fn);
        ph = ((double)k - 0.5) * kPi / (2.0 * (double)ring);
    }
    *theta = std::acos(z);
    *phi = ph;
}

int64_t healpix_radec2pix(int64_t nside, double ra, double dec)
{
    // radec2pnt (HealpixBins.h:69-71): theta = pi/2 - dec, phi = ra
    const double pi = std::atan(1.0) * 4;   // constants.h:12
    return healpix_ang2pix_ring(nside, 0.5 * pi - dec, ra);
}

// ---------------------------------------------------------------------------
// Candidate sightline pairs (XiEstimator.cc:121-159)
// ---------------------------------------------------------------------------
Trig make_trig(double ra, double dec)
{
    // SkyObject ctor, types.h:20-26
    return Trig{std::sin(dec), std::cos(dec), std::sin(ra), std::cos(ra)};
}

namespace {

inline double cos_sep(const Trig &a, const Trig &b)
{
    // SkyObject::angularSeparation, types.h:37-39 (left to right)
    const double t1 = a.sd * b.sd;
    const double t2 = a.cd * b.cd;
    const double t3 = a.sr * b.sr;
    const double t4 = a.cr * b.cr;
    return t1 + t2 * (t3 + t4);
}

struct Vec3 {
    double x, y, z;
};

inline Vec3 unit(double ra, double dec)
{
    return Vec3{std::cos(dec) * std::cos(ra), std::cos(dec) * std::sin(ra), std::sin(dec)};
}

inline double angle(const Vec3 &a, const Vec3 &b)
{
    // atan2 form: accurate for small and large angles
    const double cx = a.y * b.z - a.z * b.y, cy = a.z * b.x - a.x * b.z, cz = a.x * b.y - a.y * b.x;
    const double s = std::sqrt(cx * cx + cy * cy + cz * cz);
    const double c = a.x * b.x + a.y * b.y + a.z * b.z;
    return std::atan2(s, c);
}

struct TaskOut {
    std::vector<int32_t> row_los, row_skybin;
    std::vector<int64_t> row_len;
    std::vector<int32_t> other;
    std::vector<double> csep;
    uint64_t considered = 0, used = 0, pixel_pairs = 0, pixels = 0;
};

}  // namespace

void sky_index(const tos_sightlines &s, const tos_skybins &sky, double max_ang, SkyIndex &ix)
{
    const int64_t n = s.n_los;
    const int64_t nside = (int64_t)1 << std::max(0, sky.order);

    std::vector<Trig> &trig = ix.trig;
    trig.assign((size_t)n, Trig{});
    std::vector<int32_t> key((size_t)n);
    parallel_for(n, 4096, [&](int64_t i) {
        trig[(size_t)i] = make_trig(s.ra[i], s.dec[i]);
        key[(size_t)i] = sky.kind == TOS_SKY_HEALPIX ? (int32_t)healpix_radec2pix(nside, s.ra[i], s.dec[i])
                                                     : sky.plate_of_los[i];
    });
    // SkyBins: occupied keys in first-insertion order, members in LOS order
    // (SkyBins.h:18,59-61; HealpixBins.h:31-43; PlateBins.h:56-70)
    std::vector<int32_t> &occupied = ix.occupied;
    std::unordered_map<int32_t, int32_t> &slot = ix.slot;
    std::vector<std::vector<int32_t>> &members = ix.members;
    for (int64_t i = 0; i < n; ++i) {
        auto it = slot.find(key[(size_t)i]);
        if (it == slot.end()) {
            slot.emplace(key[(size_t)i], (int32_t)occupied.size());
            occupied.push_back(key[(size_t)i]);
            members.emplace_back();
            members.back().push_back((int32_t)i);
        } else {
            members[(size_t)it->second].push_back((int32_t)i);
        }
    }
    const int64_t nb = (int64_t)occupied.size();
    // ascending key order of the occupied bins (std::map iteration)
    std::vector<int32_t> &by_key = ix.by_key;
    by_key.assign((size_t)nb, 0);
    for (int64_t b = 0; b < nb; ++b) by_key[(size_t)b] = (int32_t)b;
    std::sort(by_key.begin(), by_key.end(), [&](int32_t a, int32_t b) { return occupied[(size_t)a] < occupied[(size_t)b]; });

    // HEALPix: neighbour cells of each occupied cell.  A cell b can hold a
    // sightline within max_ang of a member of cell a only if
    // angle(centre_a, centre_b) <= max_ang + r_a + r_b, with r the largest
    // centre-to-member angle (triangle inequality).  This returns a superset
    // of the cells that hold accepted partners, so after the cos cut the
    // candidate set equals the reference's inclusive disc query.
    std::vector<std::vector<int32_t>> &nbr_cells = ix.nbr;
    std::vector<Trig> &plate_trig = ix.plate_trig;
    std::vector<int32_t> &plate_order = ix.plate_order;
    if (sky.kind == TOS_SKY_HEALPIX) {
        std::vector<Vec3> centre((size_t)nb);
        std::vector<double> radius((size_t)nb, 0.0);
        parallel_for(nb, 8, [&](int64_t b) {
            double th, ph;
            healpix_pix2ang_ring(nside, occupied[(size_t)b], &th, &ph);
            centre[(size_t)b] = unit(ph, 0.5 * kPi - th);
            for (int32_t i : members[(size_t)b])
                radius[(size_t)b] = std::max(radius[(size_t)b], angle(centre[(size_t)b], unit(s.ra[i], s.dec[i])));
        });
        // The angle between two centres is at least their declination
        // difference, so a cell pair whose declinations differ by more than
        // the limit (with a margin far above the rounding of either side)
        // fails the test without evaluating it.  Each cell keeps its
        // neighbours in ascending key order.
        std::vector<double> cdec((size_t)nb);
        for (int64_t b = 0; b < nb; ++b) cdec[(size_t)b] = std::asin(std::max(-1.0, std::min(1.0, centre[(size_t)b].z)));
        nbr_cells.assign((size_t)nb, {});
        parallel_for(nb, 8, [&](int64_t a) {
            for (int32_t b : by_key) {
                const double lim = max_ang + radius[(size_t)a] + radius[(size_t)b] + 1e-9;
                if (std::fabs(cdec[(size_t)a] - cdec[(size_t)b]) > lim + 1e-7) continue;
                if (angle(centre[(size_t)a], centre[(size_t)b]) <= lim) nbr_cells[(size_t)a].push_back(b);
            }
        });
    } else {
        // PlateBins: std::map<int, SkyObject> iterated in ascending plate id
        plate_order.resize((size_t)sky.n_plates);
        for (int64_t p = 0; p < sky.n_plates; ++p) plate_order[(size_t)p] = (int32_t)p;
        std::sort(plate_order.begin(), plate_order.end(), [&](int32_t a, int32_t b) { return sky.plate_id[a] < sky.plate_id[b]; });
        plate_order.erase(std::unique(plate_order.begin(), plate_order.end(),
                                      [&](int32_t a, int32_t b) { return sky.plate_id[a] == sky.plate_id[b]; }),
                          plate_order.end());
        plate_trig.resize(plate_order.size());
        for (size_t p = 0; p < plate_order.size(); ++p)
            plate_trig[p] = make_trig(sky.plate_ra[plate_order[p]], sky.plate_dec[plate_order[p]]);
    }
}

int build_pairs_host(const tos_sightlines &s, const tos_skybins &sky, double max_ang, int nthreads,
                     HostPairs &out, std::string &err)
{
    const int64_t n = s.n_los;
    if (n < 0 || (n > 0 && (!s.offsets || !s.ra || !s.dec))) {
        err = "tos_pairs_build: sightline arrays missing";
        return TOS_ERR_INVALID;
    }
    if (sky.kind != TOS_SKY_HEALPIX && sky.kind != TOS_SKY_PLATES) {
        err = "tos_pairs_build: unknown sky binning";
        return TOS_ERR_INVALID;
    }
    if (sky.kind == TOS_SKY_HEALPIX && (sky.order < 0 || sky.order > 13)) {
        err = "tos_pairs_build: HEALPix order must be in [0, 13]";
        return TOS_ERR_INVALID;
    }
    if (sky.kind == TOS_SKY_PLATES && (!sky.plate_of_los || (sky.n_plates > 0 && (!sky.plate_id || !sky.plate_ra || !sky.plate_dec)))) {
        err = "tos_pairs_build: plate arrays missing";
        return TOS_ERR_INVALID;
    }
    const double cos_max = std::cos(max_ang);   // XiEstimator.cc:25
    SkyIndex ix;
    sky_index(s, sky, max_ang, ix);
    const std::vector<Trig> &trig = ix.trig;
    const std::vector<int32_t> &occupied = ix.occupied;
    const std::unordered_map<int32_t, int32_t> &slot = ix.slot;
    const std::vector<std::vector<int32_t>> &members = ix.members;
    const std::vector<std::vector<int32_t>> &nbr_cells = ix.nbr;
    const std::vector<Trig> &plate_trig = ix.plate_trig;
    const std::vector<int32_t> &plate_order = ix.plate_order;
    const int64_t nb = (int64_t)occupied.size();
    const double deg2rad = std::atan(1.0) * 4 / 180.0;   // constants.h:12-13

    // tasks: runs of at most kTaskMembers primaries of one bin, in row order
    // (bins in first-insertion order, members in sightline order), so that
    // dense bins are shared out over the threads
    constexpr size_t kTaskMembers = 128;
    struct Task {
        int32_t bin;
        int32_t m0, m1;
    };
    std::vector<Task> tasks;
    for (int64_t b = 0; b < nb; ++b)
        for (size_t m = 0; m < members[(size_t)b].size(); m += kTaskMembers)
            tasks.push_back(Task{(int32_t)b, (int32_t)m,
                                 (int32_t)std::min(members[(size_t)b].size(), m + kTaskMembers)});
    const int64_t ntask = (int64_t)tasks.size();
    std::vector<TaskOut> task((size_t)ntask);
    std::atomic<int64_t> next{0};
    auto work = [&]() {
        std::vector<int32_t> nbr_bins;
        for (;;) {
            const int64_t tk = next.fetch_add(1);
            if (tk >= ntask) break;
            const int64_t t = tasks[(size_t)tk].bin;
            TaskOut &o = task[(size_t)tk];
            for (int32_t m = tasks[(size_t)tk].m0; m < tasks[(size_t)tk].m1; ++m) {
                const int32_t i = members[(size_t)t][(size_t)m];
                const int64_t ni = s.offsets[i + 1] - s.offsets[i];
                o.pixels += (uint64_t)ni;
                const std::vector<int32_t> *bins;
                if (sky.kind == TOS_SKY_HEALPIX) {
                    bins = &nbr_cells[(size_t)t];
                } else {
                    // PlateBins::getBinIndicesWithinRadius (PlateBins.h:74-85)
                    nbr_bins.clear();
                    for (size_t p = 0; p < plate_order.size(); ++p) {
                        const double c = cos_sep(trig[(size_t)i], plate_trig[p]);
                        const double sep = std::acos(c);
                        if (!(std::fabs(sep - 1.5 * deg2rad) < max_ang)) continue;
                        auto it = slot.find(sky.plate_id[plate_order[p]]);
                        if (it != slot.end()) nbr_bins.push_back(it->second);
                    }
                    bins = &nbr_bins;
                }
                int64_t row_n = 0;
                for (int32_t b : *bins) {
                    const std::vector<int32_t> &mem = members[(size_t)b];
                    auto first = std::upper_bound(mem.begin(), mem.end(), i);
                    for (auto it = first; it != mem.end(); ++it) {
                        const int32_t j = *it;
                        const double c = cos_sep(trig[(size_t)i], trig[(size_t)j]);
                        ++o.considered;
                        if (c <= cos_max) continue;
                        ++o.used;
                        o.pixel_pairs += (uint64_t)(ni * (s.offsets[j + 1] - s.offsets[j]));
                        o.other.push_back(j);
                        o.csep.push_back(c);
                        ++row_n;
                    }
                }
                if (row_n > 0) {
                    o.row_los.push_back(i);
                    o.row_skybin.push_back(occupied[(size_t)t]);
                    o.row_len.push_back(row_n);
                }
            }
        }
    };
    const int nt = std::max(1, std::min<int>(nthreads, (int)std::max<int64_t>(1, ntask)));
    std::vector<std::thread> th;
    for (int k = 1; k < nt; ++k) th.emplace_back(work);
    work();
    for (auto &x : th) x.join();

    // assembly: offsets of every task's rows and pairs, then the copies on
    // the same threads
    out = HostPairs();
    std::vector<size_t> row_at((size_t)ntask + 1, 0), pair_at((size_t)ntask + 1, 0);
    for (int64_t k = 0; k < ntask; ++k) {
        const TaskOut &o = task[(size_t)k];
        row_at[(size_t)k + 1] = row_at[(size_t)k] + o.row_los.size();
        pair_at[(size_t)k + 1] = pair_at[(size_t)k] + o.other.size();
        out.stats.num_sightline_pairs += o.considered;
        out.stats.num_sightline_pairs_used += o.used;
        out.stats.num_pixel_pairs += o.pixel_pairs;
        out.stats.num_pixels += o.pixels;
    }
    const size_t nrows = row_at[(size_t)ntask], npairs = pair_at[(size_t)ntask];
    out.row_los.resize(nrows);
    out.row_skybin.resize(nrows);
    out.row_ptr.resize(nrows + 1);
    out.other.resize(npairs);
    out.cos_sep.resize(npairs);
    out.row_ptr[0] = 0;
    next = 0;
    auto place = [&]() {
        for (;;) {
            const int64_t k = next.fetch_add(1);
            if (k >= ntask) break;
            const TaskOut &o = task[(size_t)k];
            const size_t r0 = row_at[(size_t)k];
            int64_t at = (int64_t)pair_at[(size_t)k];
            for (size_t r = 0; r < o.row_los.size(); ++r) {
                out.row_los[r0 + r] = o.row_los[r];
                out.row_skybin[r0 + r] = o.row_skybin[r];
                at += o.row_len[r];
                out.row_ptr[r0 + r + 1] = at;
            }
            std::copy(o.other.begin(), o.other.end(), out.other.begin() + (std::ptrdiff_t)pair_at[(size_t)k]);
            std::copy(o.csep.begin(), o.csep.end(), out.cos_sep.begin() + (std::ptrdiff_t)pair_at[(size_t)k]);
        }
    };
    th.clear();
    for (int k = 1; k < nt; ++k) th.emplace_back(place);
    place();
    for (auto &x : th) x.join();
    out.stats.n_rows = (int64_t)out.row_los.size();
    out.stats.n_skybins = nb;
    out.skybin_keys = occupied;
    std::sort(out.skybin_keys.begin(), out.skybin_keys.end());
    return TOS_OK;
}

// ---------------------------------------------------------------------------
// Workgroup balancing of the work items (built on the device, tos_search.hip)
// ---------------------------------------------------------------------------
void balance_items(const std::vector<uint64_t> &item_cost, int n_wg, std::vector<int32_t> &wg_item_ptr)
{
    const size_t n = item_cost.size();
    wg_item_ptr.assign((size_t)n_wg + 1, 0);
    uint64_t total = 0;
    for (uint64_t c : item_cost) total += c;
    uint64_t acc = 0;
    size_t k = 0;
    for (int w = 0; w < n_wg; ++w) {
        wg_item_ptr[(size_t)w] = (int32_t)k;
        const uint64_t target = (uint64_t)((long double)total * (w + 1) / n_wg);
        while (k < n && acc + item_cost[k] / 2 < target) {
            acc += item_cost[k];
            ++k;
        }
    }
    wg_item_ptr[(size_t)n_wg] = (int32_t)n;
}

// ---------------------------------------------------------------------------
// Grid constants
// ---------------------------------------------------------------------------
void grid_exact_consts(const tos_grid &g, GridExact &ge)
{
    for (int k = 0; k < 3; ++k) {
        ge.min[k] = g.min[k];
        ge.n[k] = (uint32_t)g.nbins[k];
    }
    // XiEstimator.cc:169-174, 227-232, 263-268: 1/width on axes 0 and 1,
    // 1/NBins (sic) on the z axis
    ge.inv[0] = 1.0 / g.width[0];
    ge.inv[1] = 1.0 / g.width[1];
    ge.inv[2] = 1.0 / (double)g.nbins[2];
    ge.ntot = (uint32_t)g.nbins[0] * (uint32_t)g.nbins[1] * (uint32_t)g.nbins[2];
    ge.rowlen = (uint32_t)g.nbins[1] * (uint32_t)g.nbins[2];
    ge.pad = (ge.rowlen % 2 == 0) ? 1u : 2u;
    ge.nslots = (uint32_t)g.nbins[0] * (ge.rowlen + ge.pad);
    ge.row8 = 8u * (ge.rowlen + ge.pad);
    ge.nz8 = 8u * ge.n[2];
}

static bool is_pow2(double x)
{
    if (!(x > 0) || !std::isfinite(x)) return false;
    int e;
    const double m = std::frexp(x, &e);
    return m == 0.5;
}

static double ulp32(double x)
{
    // spacing of float numbers at |x|
    x = std::fabs(x);
    if (x < 1e-30) return 1e-45;
    int e;
    std::frexp(x, &e);            // x = m 2^e, m in [0.5, 1)
    return std::ldexp(1.0, e - 24);
}

// Error bounds of the FP32 path against the reference's FP64 arithmetic.
// u = 2^-24 (FP32 unit roundoff).  For a Cartesian pair
//     r_perp^2 (reference) = 2 d1 d2 (1 - c) + e1 + delta,
// e1 = f32(d1^2) - d1^2 (|e1| <= ulp(dmax^2)/2, carried exactly by the FP32
// path), delta = the reference's own FP64 rounding (<= 7e-15 dmax^2).  The
// FP32 value fma(d2, f32(d1 f32(2(1-c))), e1) is within
//     E = 2e-7 r_perp^2 + 1.2e-7 |e1| + delta,
// so with s = sqrt (one ulp) and r_perp^2 >= small:
//     |s_f32 - s| <= 4.4e-7 s + (1.2e-7 |e1| + delta) / sqrt(small).
// The bin coordinate t = (s - min) / width adds 3u(|t| + |min/width|).
// Pairs whose t lies within `safety` times that bound of a bin edge are
// re-binned with the reference arithmetic.
void grid_fast_consts(const tos_grid &g, double dmin, double dmax, GridFast &gf)
{
    const double u = std::ldexp(1.0, -24);
    const double safety = 4.0;
    const double e1max = 0.5 * ulp32(dmax * dmax);
    double delta = 7e-15 * dmax * dmax;
    // f32(d1 - d2) is exact (Sterbenz) for every pair that can be accepted
    // when dmin >= 2 * (largest accepted |d1 - d2|); otherwise allow its error.
    const double rpar_max = g.max[0] + std::fabs(g.width[0]);
    if ((g.type == TOS_CARTESIAN || g.type == TOS_POLAR) && !(dmin >= 2.0 * rpar_max))
        delta += 2.0 * rpar_max * ulp32(rpar_max) + ulp32(rpar_max) * ulp32(rpar_max);
    const double small = 4.0;
    gf.small = (float)small;
    gf.nonneg = 0;
    gf.zero_min = 0;
    for (int k = 0; k < 2; ++k) {
        const double inv = 1.0 / g.width[k];
        gf.inv[k] = (float)inv;
        gf.nf[k] = (float)g.nbins[k];
        if (g.min[k] <= 0) gf.nonneg |= (1u << k);
    }
    auto tcomp = [&](int k) {   // rounding of t = fma(x, inv, off) in FP32
        const double inv = 1.0 / g.width[k];
        return 3.0 * u * ((double)g.nbins[k] + 1.0 + std::fabs(g.min[k] * inv)) + 2.0 * u;
    };
    double G0 = 0, G1 = 0;
    gf.exact0 = 0;
    if (g.type == TOS_CARTESIAN || g.type == TOS_OBSERVING) {
        // axis 0 is an exact FP32 value (|d1-d2| or |l1-l2|); t0 is exact when
        // the scale is a power of two and there is no offset
        if (g.min[0] == 0.0 && is_pow2(1.0 / g.width[0]) && (double)(float)(1.0 / g.width[0]) == 1.0 / g.width[0])
            gf.exact0 = 1;
        G0 = tcomp(0);
    }
    if (g.type == TOS_CARTESIAN) {
        const double smax = g.max[1] + std::fabs(g.width[1]);
        const double err_s = 4.4e-7 * smax + (1.2e-7 * e1max + delta) / std::sqrt(small);
        G1 = err_s / g.width[1] + tcomp(1);
    } else if (g.type == TOS_POLAR) {
        const double rmax = g.max[0] + std::fabs(g.width[0]);
        // r = |r^2| * rsqrt(|r^2|): half the relative error of r^2, the
        // reciprocal square root (1 ulp) and the product (half an ulp)
        const double err_r = 4.8e-7 * rmax + (1.2e-7 * e1max + delta) / std::sqrt(small);
        G0 = err_r / g.width[0] + tcomp(0);
        // mu = rpar * rsqrt(|r^2|): within the bound of r's relative error
        // plus a reciprocal (1 ulp) and the product
        const double rho = 4.2e-7 + (1.2e-7 * e1max + delta) / small;
        const double mumax = std::max(std::fabs(g.max[1]), 1.0) + 1e-6;
        G1 = mumax * (2.4e-7 + rho) / g.width[1] + tcomp(1);
    }
    G0 *= safety;
    G1 *= safety;
    const double G[2] = {G0, G1};
    for (int k = 0; k < 2; ++k) {
        const double inv = 1.0 / g.width[k];
        // an exact axis 0 needs no guard: u = t
        const double gk = (k == 0 && gf.exact0) ? 0.0 : G[k];
        gf.offg[k] = (float)(-g.min[k] * inv + gk);
        gf.guard2[k] = gk >= 0.5 ? 2.0f : (float)(2.0 * gk);   // 2.0: always re-check
        // a zero coordinate on an axis starting at 0: t = 0 in both
        // arithmetics and u = g < 0.5 truncates to bin 0
        if (g.min[k] == 0.0 && gk < 0.5) gf.zero_min |= 1u << k;
    }
    // Cartesian: a computed r_perp^2 below `small` lies within
    // small (1 + 1e-6) + E0 of the reference value, E0 the absolute error
    // term above; when even that is inside the first r_perp bin (which
    // starts at 0), both arithmetics give bin 0 and no re-check is needed
    gf.small_safe = 0;
    if (g.type == TOS_CARTESIAN && g.min[1] == 0.0) {
        const double E0 = 1.2e-7 * e1max + delta;
        const double smax_small = std::sqrt(small * (1.0 + 1e-6) + E0);
        if (smax_small / g.width[1] < 1.0 - 2.0 * G1 - 1e-6) gf.small_safe = 1;
    }
}

// Smallest f32 x >= 0 whose reference axis-0 bin cvt((x - min) * inv) is
// >= n (t is monotone in x, so bisection over float bit patterns).
static double first_rejected_f32(double min, double inv, uint32_t n)
{
    // monotone in x; the caller keeps every key difference far below the
    // 2^31-bin range where the x86 cast would wrap
    auto rejected = [&](float x) {
        const double t = ((double)x - min) * inv;
        return t >= (double)n;
    };
    uint32_t lo = 0, hi = 0x7f7fffffu;   // bit patterns of 0 and FLT_MAX
    float fhi;
    std::memcpy(&fhi, &hi, 4);
    if (!rejected(fhi)) return -1.0;
    while (lo < hi) {
        const uint32_t mid = lo + (hi - lo) / 2;
        float x;
        std::memcpy(&x, &mid, 4);
        if (rejected(x)) hi = mid; else lo = mid + 1;
    }
    float x;
    std::memcpy(&x, &lo, 4);
    return (double)x;
}

double grid_window_half(const tos_grid &g, double dmax, double kmax)
{
    const double inv0 = 1.0 / g.width[0];
    const double X = first_rejected_f32(g.min[0], inv0, (uint32_t)g.nbins[0]);
    if (X < 0) return -1.0;
    // the cast wraps beyond 2^31 bins: only use the window when every key
    // difference (at most 2 kmax) stays far below that
    if ((2.0 * kmax + std::fabs(g.min[0])) * inv0 > 1e9) return -1.0;
    if (g.type != TOS_POLAR) return X;
    // Polar: r^2 >= r_par^2 - |e1| - delta, so r_par >= sqrt(X^2 + |e1| + delta)
    // guarantees r >= X
    const double e1max = 0.5 * ulp32(dmax * dmax);
    const double delta = 7e-15 * dmax * dmax;
    const double xp = std::sqrt(X * X * (1.0 + 1e-12) + e1max + delta) * (1.0 + 1e-12) + 1e-9;
    // |f32(d1 - d2)| >= f32 value above xp whenever d1 - d2 >= that value
    float xf = (float)xp;
    if ((double)xf < xp) xf = std::nextafter(xf, INFINITY);
    return (double)xf;
}

// Smallest double s >= 0 whose reference bin cvt((s - min) * inv) is >= n
// (bisection over the bit patterns of non-negative doubles).
static double first_rejected_f64(double min, double inv, uint32_t n)
{
    auto rejected = [&](double x) { return (x - min) * inv >= (double)n; };
    uint64_t lo = 0, hi = 0x7fefffffffffffffull;   // 0 and DBL_MAX
    double dhi;
    std::memcpy(&dhi, &hi, 8);
    if (!rejected(dhi)) return -1.0;
    while (lo < hi) {
        const uint64_t mid = lo + (hi - lo) / 2;
        double x;
        std::memcpy(&x, &mid, 8);
        if (rejected(x)) hi = mid; else lo = mid + 1;
    }
    double x;
    std::memcpy(&x, &lo, 8);
    return x;
}

double grid_perp_cut(const tos_grid &g, double dmin, double dmax)
{
    // the separation whose square the bound controls: r_perp (Cartesian
    // axis 1) or r (Polar axis 0); the reference takes s = sqrt(r^2 term)
    int axis;
    if (g.type == TOS_CARTESIAN) axis = 1;
    else if (g.type == TOS_POLAR) axis = 0;
    else return -1.0;
    if (!(g.width[axis] > 0)) return -1.0;
    const double S = first_rejected_f64(g.min[axis], 1.0 / g.width[axis], (uint32_t)g.nbins[axis]);
    if (!(S > 0) || S > 1e6) return -1.0;
    // the reference's squared term is >= 2 d1 d2 (1 - c) - |e1| - delta
    // (same error terms as grid_fast_consts)
    const double e1max = 0.5 * ulp32(dmax * dmax);
    double delta = 7e-15 * dmax * dmax;
    const double rpar_max = g.max[0] + std::fabs(g.width[0]);
    if (!(dmin >= 2.0 * rpar_max)) delta += 2.0 * rpar_max * ulp32(rpar_max) + ulp32(rpar_max) * ulp32(rpar_max);
    return (S * S * (1.0 + 1e-12) + e1max + delta) * (1.0 + 1e-12);
}

// ---------------------------------------------------------------------------
// cosmo::LambdaCdmUniverse distances (h5healxi.cc:123-131,196-239), Mpc/h:
//   D_C(z) = (c/H0) int_0^z dz'/E(z'),  E^2 = Om (1+z)^3 + Ok (1+z)^2 + OL,
//   D_M = D_C (flat), (c/H0) sinh(sqrt(Ok) D_C H0/c)/sqrt(Ok) (open),
//         (c/H0) sin(sqrt(-Ok) D_C H0/c)/sqrt(-Ok) (closed).
// Integrated with composite 8-point Gauss-Legendre in ln(1+z) (tos_cosmo.h,
// the arithmetic the device's pixel transform shares).
// ---------------------------------------------------------------------------
double cosmo_los_comoving(double omega_lambda, double omega_matter, double z)
{
    return los_comoving(omega_lambda, omega_matter, z);
}

double cosmo_transverse(double omega_lambda, double omega_matter, double z)
{
    const double c_over_h0 = 2997.92458;
    const double dc = cosmo_los_comoving(omega_lambda, omega_matter, z);
    const double ok = 1.0 - omega_matter - omega_lambda;
    if (std::fabs(ok) < 1e-12) return dc;
    const double k = std::sqrt(std::fabs(ok));
    const double x = k * dc / c_over_h0;
    return ok > 0 ? c_over_h0 * std::sinh(x) / k : c_over_h0 * std::sin(x) / k;
}

}  // namespace tos
