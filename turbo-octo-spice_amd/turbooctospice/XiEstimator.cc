// XiEstimator.cc -- the turbooctospice estimator on top of the tos_amd C-ABI.
#include "XiEstimator.h"

#include <algorithm>
#include <atomic>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <iostream>
#include <thread>

#include "ThreadPool.h"

namespace turbooctospice {

namespace {

// boost::lexical_cast<std::string>(double) prints max_digits10 = 17
// significant digits; integers print plainly.
std::string lexical(double v)
{
    char buf[64];
    std::snprintf(buf, sizeof(buf), "%.17g", v);
    return buf;
}

struct Ctx {
    tos_ctx *h = nullptr;
    explicit Ctx(int device)
    {
        if (tos_ctx_create(device, &h) != TOS_OK) throw RuntimeError("XiEstimator: no usable HIP device");
    }
    ~Ctx() { tos_ctx_destroy(h); }
    void check(int st, const char *what) const
    {
        if (st != TOS_OK) throw RuntimeError(std::string(what) + ": " + tos_last_error(h));
    }
};

}  // namespace

XiEstimator::XiEstimator(double scale, AbsTwoPointGridPtr grid, BinningCoordinateType type,
                         std::vector<Forest> sightlines, SkyBinsIPtr skybins)
    : num_sightline_pairs_(0), num_sightline_pairs_used_(0), num_pixels_(0), num_pixel_pairs_(0),
      num_pixel_pairs_used_(0), max_ang_(grid->maxAngularScale(scale)),
      cos_max_ang_(std::cos(grid->maxAngularScale(scale))), coordinate_type_(type), grid_(grid),
      skybins_(skybins), sightlines_(std::move(sightlines)), num_xi_bins_((unsigned)grid->getNBinsTotal()),
      cov_good_(false)
{
}

XiEstimator::~XiEstimator() {}

// Candidate sightline pairs through the SkyBins object: one task per
// occupied sky bin (XiEstimator.cc:31-42, 121-159); rows grouped by sky bin
// in ascending key order.
void XiEstimator::enumerate_host(int nthreads, Csr &csr)
{
    const std::vector<int> occupied = skybins_->getOccupiedBins();
    struct TaskOut {
        std::vector<int32_t> row_los;
        std::vector<int64_t> row_len;
        std::vector<int32_t> other;
        std::vector<double> csep;
        unsigned long considered = 0, used = 0, pixel_pairs = 0, pixels = 0;
    };
    std::vector<TaskOut> out(occupied.size());
    {
        ThreadPool pool(nthreads);
        std::list<ThreadPool::Task> tasks;
        for (size_t t = 0; t < occupied.size(); ++t) {
            tasks.push_back([this, t, &occupied, &out]() {
                TaskOut &o = out[t];
                for (int i : skybins_->getBinContents(occupied[t])) {
                    const Forest &primary = sightlines_[(size_t)i];
                    o.pixels += primary.pixels.size();
                    int64_t row_n = 0;
                    for (int neighbor : skybins_->getBinIndicesWithinRadius(primary.ra, primary.dec, max_ang_)) {
                        if (!skybins_->checkBinExists(neighbor)) continue;
                        for (int j : skybins_->getBinContents(neighbor)) {
                            if (j <= i) continue;
                            const Forest &other = sightlines_[(size_t)j];
                            const double c = primary.angularSeparation(other);
                            ++o.considered;
                            if (c <= cos_max_ang_) continue;
                            ++o.used;
                            o.pixel_pairs += primary.pixels.size() * other.pixels.size();
                            o.other.push_back(j);
                            o.csep.push_back(c);
                            ++row_n;
                        }
                    }
                    if (row_n > 0) {
                        o.row_los.push_back(i);
                        o.row_len.push_back(row_n);
                    }
                }
                return true;
            });
        }
        pool.run(tasks);
    }
    // rows grouped by sky bin, sky bins in ascending key order (std::map)
    std::vector<size_t> order(occupied.size());
    for (size_t t = 0; t < order.size(); ++t) order[t] = t;
    std::sort(order.begin(), order.end(), [&](size_t a, size_t b) { return occupied[a] < occupied[b]; });
    std::vector<int32_t> &row_los = csr.row_los, &row_skybin = csr.row_skybin, &other = csr.other;
    std::vector<int64_t> &row_ptr = csr.row_ptr;
    std::vector<double> &csep = csr.csep;
    row_ptr.assign(1, 0);
    for (size_t t : order) {
        const TaskOut &o = out[t];
        int64_t p = 0;
        for (size_t r = 0; r < o.row_los.size(); ++r) {
            row_los.push_back(o.row_los[r]);
            row_skybin.push_back(occupied[t]);
            row_ptr.push_back(row_ptr.back() + o.row_len[r]);
            other.insert(other.end(), o.other.begin() + p, o.other.begin() + p + o.row_len[r]);
            csep.insert(csep.end(), o.csep.begin() + p, o.csep.begin() + p + o.row_len[r]);
            p += o.row_len[r];
        }
        num_sightline_pairs_ += o.considered;
        num_sightline_pairs_used_ += o.used;
        num_pixel_pairs_ += o.pixel_pairs;
        num_pixels_ += o.pixels;
    }

}

void XiEstimator::run(int nthreads, bool do_cov)
{
    std::cout << "Estimating xi..." << std::endl;
    const bool device = device_search_ && (dynamic_cast<const HealpixBinsI *>(skybins_.get()) != nullptr ||
                                           dynamic_cast<const PlateBinsI *>(skybins_.get()) != nullptr);
    if (device) {
        accumulate(nullptr, nthreads);
    } else {
        Csr csr;
        enumerate_host(nthreads, csr);
        accumulate(&csr, nthreads);
    }
    std::cout << "Xi estimation complete!" << std::endl << std::endl;
    print_stats();
    std::cout << std::endl;

    if (do_cov) {
        // weighted covariance of the sky-bin subsamples (XiEstimator.cc:60-97):
        // sample xi_b = didj_b / wgt_b where wgt_b > 0, sample weight = sum wgt;
        // computed on the device (tos_subsample_covariance)
        std::cout << "Estimating covariance..." << std::endl;
        const size_t nb = num_xi_bins_;
        const size_t ns = skybin_xis_.size();
        std::vector<double> didj(ns * nb), wgt(ns * nb);
        size_t k = 0;
        for (const auto &kv : skybin_xis_) {
            for (size_t b = 0; b < nb; ++b) {
                didj[k * nb + b] = kv.second[b].didj;
                wgt[k * nb + b] = kv.second[b].wgt;
            }
            ++k;
        }
        cov_.assign(nb * nb, 0.0);
        double wsum = 0;
        if (ns > 1) {
            Ctx ctx(device_);
            ctx.check(tos_subsample_covariance(ctx.h, (int64_t)ns, (int32_t)nb, didj.data(), wgt.data(),
                                               cov_.data(), nullptr, &wsum),
                      "tos_subsample_covariance");
        }
        if (wsum > 0) {
            // the reference's diagnostics (XiEstimator.cc:82-87)
            int32_t is_pos_def = 0;
            double log_det = 0;
            if (tos_covariance_check((int32_t)nb, cov_.data(), &is_pos_def, &log_det) != TOS_OK)
                throw RuntimeError("tos_covariance_check: invalid covariance");
            std::cout << "is pos def: " << is_pos_def << std::endl;
            if (is_pos_def) std::cout << "log(|C|): " << lexical(log_det) << std::endl;
            cov_good_ = true;
        } else {
            std::cerr << "covariance needs at least two weighted subsamples" << std::endl;
        }
        std::cout << "Covariance estimation complete!" << std::endl << std::endl;
    }
}

void XiEstimator::accumulate(const Csr *csr, int nthreads)
{
    Ctx ctx(device_);
    // device catalogue: SoA copy of the sightlines (types.h:43-66)
    const size_t nl = sightlines_.size();
    std::vector<int64_t> offsets(nl + 1, 0);
    for (size_t k = 0; k < nl; ++k) offsets[k + 1] = offsets[k] + (int64_t)sightlines_[k].pixels.size();
    std::vector<double> ra(nl), dec(nl);
    std::vector<float> val((size_t)offsets[nl]), ll(val.size()), wt(val.size()), dist(val.size());
    // sightlines in contiguous blocks, one per thread (disjoint pixel ranges)
    auto fill = [&](size_t k0, size_t k1) {
        for (size_t k = k0; k < k1; ++k) {
            ra[k] = sightlines_[k].ra;
            dec[k] = sightlines_[k].dec;
            size_t p = (size_t)offsets[k];
            for (const ForestPixel &px : sightlines_[k].pixels) {
                val[p] = px.value;
                ll[p] = px.loglam;
                wt[p] = px.weight;
                dist[p] = px.distance;
                ++p;
            }
        }
    };
    const size_t nt = std::max<size_t>(1, std::min<size_t>((size_t)std::max(1, nthreads), nl / 1024 + 1));
    std::vector<std::thread> th;
    for (size_t t = 1; t < nt; ++t) th.emplace_back(fill, nl * t / nt, nl * (t + 1) / nt);
    fill(0, nl / nt);
    for (std::thread &x : th) x.join();
    tos_sightlines s{(int64_t)nl, offsets.data(), ra.data(), dec.data(), val.data(), ll.data(), wt.data(), dist.data()};
    // the device catalogue and pair list are released on every way out,
    // exceptions included
    struct Owned {
        tos_ctx *h;
        tos_catalog *cat = nullptr;
        tos_pairs *pairs = nullptr;
        ~Owned()
        {
            if (pairs) tos_pairs_free(h, pairs);
            if (cat) tos_catalog_free(h, cat);
        }
    } own{ctx.h};
    tos_catalog *&cat = own.cat;
    ctx.check(tos_catalog_upload(ctx.h, &s, &cat), "tos_catalog_upload");
    tos_grid g = grid_->toC();
    g.type = (int32_t)coordinate_type_;   // the branch follows the coordinate type (XiEstimator.cc:168,226,262)
    tos_xi_opts opts;
    std::memset(&opts, 0, sizeof(opts));
    opts.precision = precision_;
    opts.accum = TOS_ACC_ALL;
    const size_t nb = num_xi_bins_;
    tos_pairs *&pairs = own.pairs;
    if (csr) {
        tos_pair_csr c{(int64_t)csr->row_los.size(), csr->row_los.data(), csr->row_skybin.data(),
                       csr->row_ptr.data(), csr->other.data(), csr->csep.data()};
        ctx.check(tos_pairs_upload(ctx.h, cat, &c, &pairs), "tos_pairs_upload");
    } else {
        // the GPU search over the library's own HEALPix or plate sky bins
        tos_skybins sky;
        std::memset(&sky, 0, sizeof(sky));
        std::vector<int32_t> plate_of_los, plate_id;
        std::vector<double> plate_ra, plate_dec;
        if (const HealpixBinsI *h = dynamic_cast<const HealpixBinsI *>(skybins_.get())) {
            sky.kind = TOS_SKY_HEALPIX;
            sky.order = h->order();
        } else {
            const PlateBinsI *pb = dynamic_cast<const PlateBinsI *>(skybins_.get());
            sky.kind = TOS_SKY_PLATES;
            for (const Forest &f : sightlines_) plate_of_los.push_back(f.plate);
            for (const auto &kv : pb->plates()) {
                plate_id.push_back(kv.first);
                plate_ra.push_back(kv.second.ra);
                plate_dec.push_back(kv.second.dec);
            }
            sky.plate_of_los = plate_of_los.data();
            sky.n_plates = (int64_t)plate_id.size();
            sky.plate_id = plate_id.data();
            sky.plate_ra = plate_ra.data();
            sky.plate_dec = plate_dec.data();
        }
        ctx.check(tos_pairs_build(ctx.h, cat, &sky, max_ang_, nthreads, 1, &pairs), "tos_pairs_build");
        tos_pair_stats ps;
        tos_pairs_stats(pairs, &ps);
        num_sightline_pairs_ += ps.num_sightline_pairs;
        num_sightline_pairs_used_ += ps.num_sightline_pairs_used;
        num_pixel_pairs_ += ps.num_pixel_pairs;
        num_pixels_ += ps.num_pixels;
    }
    skybin_xis_.clear();
    xi_.assign(nb, XiBin());
    int st = TOS_OK;
    if (keep_subsamples_) {
        // one pass with a histogram per sky bin (count, sum w1w2d1d2, sum
        // w1w2: the fields save_subsamples and the covariance read)
        tos_pair_stats ps;
        tos_pairs_stats(pairs, &ps);
        const size_t nk = (size_t)ps.n_skybins;
        std::vector<int32_t> keys(nk);
        std::vector<uint64_t> sn(nk * nb), n(nb);
        std::vector<double> sd(nk * nb), sw(nk * nb), d(nb), w(nb);
        tos_xi_result res;
        std::memset(&res, 0, sizeof(res));
        res.num_pairs = n.data();
        res.didj = d.data();
        res.wgt = w.data();
        res.skybin_keys = keys.data();
        res.sub_num_pairs = sn.data();
        res.sub_didj = sd.data();
        res.sub_wgt = sw.data();
        tos_xi_opts o = opts;
        o.accum = TOS_ACC_CORE;
        o.per_skybin = 1;
        st = tos_xi_run(ctx.h, cat, pairs, &g, &o, &res);
        if (st == TOS_OK) {
            for (size_t k = 0; k < nk; ++k) {
                std::vector<XiBin> &dst = skybin_xis_[keys[k]];
                dst.assign(nb, XiBin());
                for (size_t b = 0; b < nb; ++b) {
                    dst[b].didj = sd[k * nb + b];
                    dst[b].wgt = sw[k * nb + b];
                    dst[b].num_pairs = (long)sn[k * nb + b];
                }
            }
            // occupied sky bins without candidate pairs keep an empty
            // subsample, as every sky-bin task stores one (XiEstimator.cc:155)
            for (int key : skybins_->getOccupiedBins())
                if (!skybin_xis_.count(key)) skybin_xis_[key] = std::vector<XiBin>(nb);
        }
    }
    if (st == TOS_OK) {
        // the total with every XiBin field
        std::vector<uint64_t> n(nb);
        std::vector<double> f[7];
        for (auto &v : f) v.assign(nb, 0.0);
        tos_xi_result res;
        std::memset(&res, 0, sizeof(res));
        res.num_pairs = n.data();
        res.didj = f[0].data();
        res.wgt = f[1].data();
        res.di = f[2].data();
        res.dj = f[3].data();
        res.wi = f[4].data();
        res.wj = f[5].data();
        res.z = f[6].data();
        st = tos_xi_run(ctx.h, cat, pairs, &g, &opts, &res);
        for (size_t b = 0; st == TOS_OK && b < nb; ++b) {
            XiBin &x = xi_[b];
            x.didj = f[0][b];
            x.wgt = f[1][b];
            x.di = f[2][b];
            x.dj = f[3][b];
            x.wi = f[4][b];
            x.wj = f[5][b];
            x.z = f[6][b];
            x.num_pairs = (long)n[b];
        }
        num_pixel_pairs_used_ += res.num_pixel_pairs_used;
    }
    ctx.check(st, "tos_xi_run");
    for (XiBin &x : xi_) x.finalize();
}

void XiEstimator::save_results(std::string outfile) const
{
    // XiEstimator.cc:307-361
    const std::string estimator_filename(outfile + ".data");
    std::cout << "Saving correlation function to: " << estimator_filename << std::endl;
    std::ofstream est(estimator_filename.c_str());
    for (unsigned b = 0; b < num_xi_bins_; ++b) est << b << ' ' << lexical(xi_[b].didj) << std::endl;
    est.close();

    const std::string details_filename(outfile + ".details");
    std::cout << "Saving correlation function details to: " << details_filename << std::endl;
    std::ofstream det(details_filename.c_str());
    std::vector<double> c(3);
    for (unsigned b = 0; b < num_xi_bins_; ++b) {
        grid_->getBinCenters((int)b, c);
        det << b << ' ' << c[0] << ' ' << c[1] << ' ' << c[2] << ' ' << lexical(xi_[b].didj) << ' '
            << lexical(xi_[b].di) << ' ' << lexical(xi_[b].dj) << ' ' << lexical(xi_[b].wgt) << ' '
            << xi_[b].num_pairs << ' ' << lexical(xi_[b].z) << std::endl;
    }
    det.close();

    if (cov_good_) {
        const std::string cov_filename(outfile + ".cov");
        std::cout << "Saving covariance matrix to: " << cov_filename << std::endl;
        std::ofstream cf(cov_filename.c_str());
        for (unsigned col = 0; col < num_xi_bins_; ++col)
            for (unsigned row = 0; row <= col; ++row)
                cf << row << ' ' << col << ' ' << lexical(cov_[(size_t)row * num_xi_bins_ + col]) << std::endl;
        cf.close();
    }

    const std::string weight_filename(outfile + ".iwgt");
    std::cout << "Saving inverse diagonal weight matrix to: " << weight_filename << std::endl;
    std::ofstream wf(weight_filename.c_str());
    for (unsigned col = 0; col < num_xi_bins_; ++col)
        wf << col << ' ' << col << ' ' << lexical(1.0 / xi_[col].wgt) << std::endl;
    wf.close();
}

void XiEstimator::save_subsamples(std::string outfile_base) const
{
    // XiEstimator.cc:363-383
    std::cout << "Saving subsamples correlation functions to: " << outfile_base << "-<id>.data" << std::endl;
    for (const auto &kv : skybin_xis_) {
        std::ofstream f((outfile_base + "-" + std::to_string(kv.first) + ".data").c_str());
        for (unsigned b = 0; b < num_xi_bins_; ++b) {
            const double weight = kv.second[b].wgt;
            const double value = weight > 0 ? kv.second[b].didj / weight : 0.0;
            f << b << ' ' << lexical(value) << ' ' << lexical(weight) << ' ' << kv.second[b].num_pairs << std::endl;
        }
    }
}

void XiEstimator::print_stats() const
{
    // XiEstimator.cc:389-408
    const unsigned long n = sightlines_.size();
    const unsigned long total = n * (n - 1) / 2;
    std::cout << "Number of distinct los pairs " << total << std::endl;
    std::cout << "considered " << num_sightline_pairs_ << " of distinct los pairs. ("
              << (double)num_sightline_pairs_ / total << ")" << std::endl;
    std::cout << "used " << num_sightline_pairs_used_ << " of los pairs considered. ("
              << (double)num_sightline_pairs_used_ / num_sightline_pairs_ << ")" << std::endl;
    const unsigned long ptotal = num_pixels_ * (num_pixels_ - 1) / 2;
    std::cout << "Number of distinct pixel pairs " << ptotal << std::endl;
    std::cout << "considered " << num_pixel_pairs_ << " of distinct pixel pairs. ("
              << (double)num_pixel_pairs_ / ptotal << ")" << std::endl;
    std::cout << "used " << num_pixel_pairs_used_ << " of pixel pairs considered. ("
              << (double)num_pixel_pairs_used_ / num_pixel_pairs_ << ")" << std::endl;
}

void bruteGPU(std::vector<std::vector<double>> &columns, double min, double max, int nbins, std::vector<double> &xi,
              long chunksize, BruteArithmetic arithmetic)
{
    (void)chunksize;
    if (columns.size() < 5) throw RuntimeError("bruteGPU: need 5 columns (x y z d w)");
    const int64_t n = (int64_t)columns[0].size();
    for (int k = 1; k < 5; ++k)
        if ((int64_t)columns[k].size() != n) throw RuntimeError("bruteGPU: columns differ in length");
    Ctx ctx(0);
    // FP32: slot 0 underflow, slots 1..nbins the bins, nbins + 1 overflow
    // (brute.cu:39-49); FP64: nbins slots
    const bool f32 = arithmetic == BruteFP32;
    const int nslot = f32 ? nbins + 2 : nbins;
    std::vector<double> dsum(nslot), wsum(nslot);
    std::vector<uint64_t> counts(nslot);
    double ms = 0;
    ctx.check(tos_brute3d(ctx.h, f32 ? TOS_BRUTE_FP32 : TOS_BRUTE_FP64, n, columns[0].data(), columns[1].data(),
                          columns[2].data(), columns[3].data(), columns[4].data(), min, max, nbins, dsum.data(),
                          wsum.data(), counts.data(), &ms),
              "tos_brute3d");
    const int first = f32 ? 1 : 0;
    xi.assign(nbins, 0.0);
    for (int b = 0; b < nbins; ++b)   // brute.cu:184-190, 206-212
        if (wsum[b + first] > 0) xi[b] = dsum[b + first] / wsum[b + first];
}

}  // namespace turbooctospice
