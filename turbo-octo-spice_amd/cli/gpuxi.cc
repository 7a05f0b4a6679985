// gpuxi -- 3D brute-force correlation function of text columns "x y z d w",
// with the interface of the reference's src/gpuxi.cc (same flags, output
// "index xi").  Runs bruteGPU on the device.
#include <cstdlib>
#include <fstream>
#include <iostream>
#include <sstream>
#include <string>
#include <vector>

#include "turbooctospice.h"

namespace tos = turbooctospice;

int main(int argc, char **argv)
{
    std::string infile, outfile = "xi.dat", axis1 = "[0:200]*50";
    long chunksize = 4096;
    bool verbose = false;
    bool fp64 = false;
    for (int k = 1; k < argc; ++k) {
        const std::string a = argv[k];
        auto next = [&]() -> std::string {
            if (k + 1 >= argc) {
                std::cerr << "Unable to parse command line options: missing value for " << a << std::endl;
                std::exit(-1);
            }
            return argv[++k];
        };
        if (a == "-h" || a == "--help") {
            std::cout << "Correlation function estimator:\n  -h [ --help ]\n  --verbose\n  -i [ --input ] arg\n"
                         "  -o [ --output ] arg (=xi.dat)\n  --axis1 arg (=[0:200]*50)\n  --chunksize arg (=4096)\n"
                         "  --fp64                 bin pairs in FP64 (old/BinPair.h) instead of the\n"
                         "                         reference kernel's FP32 arithmetic\n";
            return 1;
        } else if (a == "--verbose") verbose = true;
        else if (a == "--fp64") fp64 = true;
        else if (a == "-i" || a == "--input") infile = next();
        else if (a == "-o" || a == "--output") outfile = next();
        else if (a == "--axis1") axis1 = next();
        else if (a == "--chunksize") chunksize = std::atol(next().c_str());
        else {
            std::cerr << "Unable to parse command line options: " << a << std::endl;
            return -1;
        }
    }
    if (infile.empty()) {
        std::cerr << "Missing infile parameter." << std::endl;
        return -2;
    }
    std::vector<std::vector<double>> columns(5);
    {
        std::ifstream in(infile.c_str());
        if (!in) {
            std::cerr << "Error while reading " << infile << std::endl;
            return -3;
        }
        std::string line;
        while (std::getline(in, line)) {
            std::istringstream ss(line);
            double v[5];
            if (ss >> v[0] >> v[1] >> v[2] >> v[3] >> v[4])
                for (int c = 0; c < 5; ++c) columns[c].push_back(v[c]);
        }
    }
    if (verbose) std::cout << "Read " << columns[0].size() << " rows from " << infile << std::endl;
    std::vector<double> xi;
    try {
        auto bins = tos::createBinning(axis1);
        tos::bruteGPU(columns, bins->getBinLowEdge(0), bins->getBinHighEdge(bins->getNBins() - 1), bins->getNBins(),
                      xi, chunksize, fp64 ? tos::BruteFP64 : tos::BruteFP32);
    } catch (std::exception const &e) {
        std::cerr << "Error while running the estimator: " << e.what() << std::endl;
    }
    std::ofstream out(outfile.c_str());
    for (size_t i = 0; i < xi.size(); ++i) out << i << ' ' << xi[i] << std::endl;
    return 0;
}
