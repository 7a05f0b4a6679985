// h5healxi -- correlation-function estimator for HDF5 delta fields, with the
// command-line interface of the reference's src/h5healxi.cc (same flags and
// defaults) plus --precision, --device, --cov and --device-search.  The pixel pairs run on the GPU.
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <iostream>
#include <map>
#include <string>
#include <vector>

#include "turbooctospice.h"

namespace tos = turbooctospice;

namespace {

struct Options {
    std::map<std::string, std::string> val;
    std::map<std::string, bool> flag;
};

const char *kHelp =
    "Correlation function estimator:\n"
    "  -h [ --help ]                Prints this info and exits.\n"
    "  --verbose                    Prints additional information.\n"
    "  --order arg (=5)             Healpix map order parameter\n"
    "  --nthreads arg (=3)          Number of threads to use\n"
    "  --omega-lambda arg (=0.73)   Present-day value of OmegaLambda.\n"
    "  --omega-matter arg (=0)      Present-day value of OmegaMatter or zero for 1-OmegaLambda.\n"
    "  -i [ --input ] arg           Filename to read from\n"
    "  -o [ --output ] arg (=healxi) Output filename base\n"
    "  --axis1 arg (=[0:200]*50)    Axis-1 binning, r_par (Mpc/h), r (Mpc/h), or dloglam\n"
    "  --axis2 arg (=[0:1]*1)       Axis-2 binning, r_perp (Mpc/h), mu (r_par/r), or dtheta (rad)\n"
    "  --axis3 arg (=[3.556:3.735]*1) Axis-3 binning, log10(lya*(1+z))\n"
    "  --polar                      (r,mu,z) binning\n"
    "  --cart                       (r_perp,r_par,z) binning\n"
    "  --debug                      debug flag\n"
    "  --save-subsamples            save xi estimates on individual subsamples\n"
    "  -n [ --num-sightlines ] arg (=0) number of sightlines to use.\n"
    "  --platelist-filename arg     platelist filename\n"
    "  --precision arg (=fp64)      fp64 | fp32c (pixel-pair arithmetic on the GPU)\n"
    "  --device arg (=0)            HIP device\n"
    "  --cov                        estimate the covariance of the sky-bin subsamples\n"
    "  --device-search              enumerate the candidate sightline pairs on the GPU\n";

bool parse(int argc, char **argv, Options &o, std::string &err)
{
    const std::map<std::string, std::string> alias = {{"-h", "help"}, {"-i", "input"}, {"-o", "output"},
                                                      {"-n", "num-sightlines"}};
    const std::vector<std::string> flags = {"help",  "verbose",         "polar", "cart",
                                            "debug", "save-subsamples", "cov",   "device-search"};
    for (int k = 1; k < argc; ++k) {
        std::string a = argv[k], name, value;
        bool has_value = false;
        if (alias.count(a)) {
            name = alias.at(a);
        } else if (a.rfind("--", 0) == 0) {
            name = a.substr(2);
            const size_t eq = name.find('=');
            if (eq != std::string::npos) {
                value = name.substr(eq + 1);
                name = name.substr(0, eq);
                has_value = true;
            }
        } else {
            err = "unexpected argument " + a;
            return false;
        }
        if (std::find(flags.begin(), flags.end(), name) != flags.end()) {
            o.flag[name] = true;
            continue;
        }
        if (!has_value) {
            if (k + 1 >= argc) {
                err = "missing value for " + a;
                return false;
            }
            value = argv[++k];
        }
        o.val[name] = value;
    }
    return true;
}

std::string get(const Options &o, const std::string &k, const std::string &def)
{
    auto it = o.val.find(k);
    return it == o.val.end() ? def : it->second;
}

}  // namespace

int main(int argc, char **argv)
{
    Options o;
    std::string err;
    if (!parse(argc, argv, o, err)) {
        std::cerr << "Unable to parse command line options: " << err << std::endl;
        return -1;
    }
    if (o.flag["help"]) {
        std::cout << kHelp << std::endl;
        return 1;
    }
    const bool verbose = o.flag["verbose"];
    const int order = std::atoi(get(o, "order", "5").c_str());
    const int nthreads = std::atoi(get(o, "nthreads", "3").c_str());
    const double omega_lambda = std::atof(get(o, "omega-lambda", "0.73").c_str());
    const double omega_matter = std::atof(get(o, "omega-matter", "0").c_str());
    const std::string infile = get(o, "input", "");
    const std::string outfile = get(o, "output", "healxi");
    const std::string axis1 = get(o, "axis1", "[0:200]*50"), axis2 = get(o, "axis2", "[0:1]*1"),
                      axis3 = get(o, "axis3", "[3.556:3.735]*1");
    int num_sightlines = std::atoi(get(o, "num-sightlines", "0").c_str());
    const std::string platelist = get(o, "platelist-filename", "platelist.fits");
    const std::string precision = get(o, "precision", "fp64");
    const int device = std::atoi(get(o, "device", "0").c_str());
    if (infile.empty()) {
        std::cerr << "Must specify input file." << std::endl;
        return -1;
    }
    if (precision != "fp64" && precision != "fp32c") {
        std::cerr << "--precision must be fp64 or fp32c" << std::endl;
        return -1;
    }
    try {
        tos::LambdaCdmUniverse cosmology(omega_lambda, omega_matter);
        tos::AbsTwoPointGridPtr grid;
        tos::XiEstimator::BinningCoordinateType type;
        auto b1 = tos::createBinning(axis1), b2 = tos::createBinning(axis2), b3 = tos::createBinning(axis3);
        if (o.flag["polar"]) {
            type = tos::XiEstimator::PolarCoordinates;
            grid.reset(new tos::PolarGrid(b1, b2, b3));
        } else if (o.flag["cart"]) {
            type = tos::XiEstimator::CartesianCoordinates;
            grid.reset(new tos::CartesianGrid(b1, b2, b3));
        } else {
            type = tos::XiEstimator::ObservingCoordinates;
            grid.reset(new tos::QuasarGrid(b1, b2, b3));
        }
        if (verbose) {
            std::cout << "Binning grid configuration: " << std::endl;
            for (int axis = 0; axis < 3; ++axis)
                std::cout << " axis" << axis + 1 << " : " << grid->getAxisMin(axis) << " " << grid->getAxisMax(axis)
                          << " " << grid->getAxisBinWidth(axis) << " " << grid->getAxisNBins(axis) << std::endl;
        }
        tos::SkyBinsIPtr skybins;
        if (order > 0) skybins.reset(new tos::HealpixBinsI(order));
        else skybins.reset(new tos::PlateBinsI(platelist));

        tos::HDF5Delta file(infile);
        if (verbose) std::cout << "Loading forest delta fields from " << infile << std::endl;
        std::vector<tos::Forest> sightlines(file.loadForests("lines_of_sight"));
        if (num_sightlines > 0 && num_sightlines < (int)sightlines.size())
            sightlines.erase(sightlines.begin() + num_sightlines, sightlines.end());
        num_sightlines = (int)sightlines.size();
        if (sightlines.empty()) throw tos::RuntimeError("no sightlines loaded");

        // distance transform (h5healxi.cc:196-216) on the device: every
        // pixel's FP64 comoving distance at z = 10^(loglam - logLyA) - 1
        double min_loglam = sightlines[0].pixels[0].loglam, max_loglam = sightlines[0].pixels[0].loglam;
        unsigned long num_pixels = 0;
        for (const auto &s : sightlines) num_pixels += s.pixels.size();
        {
            std::vector<float> loglam, distance(num_pixels);
            loglam.reserve(num_pixels);
            for (const auto &s : sightlines)
                for (const auto &p : s.pixels) loglam.push_back(p.loglam);
            cosmology.transformPixels((int64_t)num_pixels, loglam.data(), distance.data(), device);
            size_t k = 0;
            for (auto &s : sightlines)
                for (auto &p : s.pixels) p.distance = distance[k++];
        }
        for (auto &s : sightlines) {
            skybins->addItem(s, s.forest_id);
            if (s.pixels.front().loglam < min_loglam) min_loglam = s.pixels.front().loglam;
            if (s.pixels.back().loglam > max_loglam) max_loglam = s.pixels.back().loglam;
        }
        if (verbose) {
            std::cout << "Loaded " << num_pixels << " pixels from " << num_sightlines << " lines of sight (LOS)"
                      << std::endl;
            std::cout << "Number of sky bins occupied: " << skybins->getNBins() << std::endl;
        }
        const double zmin = std::pow(10, min_loglam - tos::logLyA) - 1;
        const double scale = cosmology.getTransverseComovingScale(zmin);
        if (verbose)
            std::cout << "Max angular scale at z = " << zmin << " : " << grid->maxAngularScale(scale) << " (rad)"
                      << std::endl;

        tos::XiEstimator xiest(scale, grid, type, sightlines, skybins);
        xiest.setDevice(device);
        xiest.setPrecision(precision == "fp32c" ? tos::XiEstimator::FP32C : tos::XiEstimator::FP64);
        xiest.setKeepSubsamples(o.flag["save-subsamples"] || o.flag["cov"]);
        xiest.setDeviceSearch(o.flag["device-search"]);
        xiest.run(nthreads, o.flag["cov"]);
        std::cout << std::endl;
        xiest.save_results(outfile);
        if (o.flag["save-subsamples"]) xiest.save_subsamples(outfile);
    } catch (tos::RuntimeError const &e) {
        std::cerr << e.what() << std::endl;
        return -1;
    } catch (std::exception const &e) {
        std::cerr << e.what() << std::endl;
        return -1;
    }
    return 0;
}
