// tos_mkdelta -- writes a sightline catalogue in the HDF5 delta layout read
// by HDF5Delta::loadForests (group "lines_of_sight", one subgroup per
// sightline).  Input is a raw little-endian dump:
//   int64 n_los, int64 n_pix, int64 offsets[n_los+1], double ra[n_los],
//   double dec[n_los], double z[n_los], int32 plate[n_los],
//   float delta[n_pix], float loglam[n_pix], float weight[n_pix],
//   float r_comov[n_pix]
#include <cstdint>
#include <cstdio>
#include <iostream>
#include <vector>

#include "turbooctospice.h"

namespace tos = turbooctospice;

template <typename T> bool rd(std::FILE *f, std::vector<T> &v, size_t n)
{
    v.resize(n);
    return n == 0 || std::fread(v.data(), sizeof(T), n, f) == n;
}

int main(int argc, char **argv)
{
    if (argc != 3) {
        std::cerr << "usage: tos_mkdelta <catalogue.bin> <out.h5>" << std::endl;
        return 2;
    }
    std::FILE *f = std::fopen(argv[1], "rb");
    if (!f) {
        std::cerr << "cannot open " << argv[1] << std::endl;
        return 2;
    }
    int64_t hdr[2];
    std::vector<int64_t> off;
    std::vector<double> ra, dec, z;
    std::vector<int32_t> plate;
    std::vector<float> delta, loglam, weight, rc;
    bool ok = std::fread(hdr, sizeof(int64_t), 2, f) == 2 && hdr[0] >= 0 && hdr[1] >= 0;
    ok = ok && rd(f, off, (size_t)hdr[0] + 1) && rd(f, ra, (size_t)hdr[0]) && rd(f, dec, (size_t)hdr[0]) &&
         rd(f, z, (size_t)hdr[0]) && rd(f, plate, (size_t)hdr[0]) && rd(f, delta, (size_t)hdr[1]) &&
         rd(f, loglam, (size_t)hdr[1]) && rd(f, weight, (size_t)hdr[1]) && rd(f, rc, (size_t)hdr[1]);
    std::fclose(f);
    if (!ok || off.back() != hdr[1]) {
        std::cerr << "malformed catalogue dump" << std::endl;
        return 2;
    }
    std::vector<tos::Forest> forests;
    for (int64_t k = 0; k < hdr[0]; ++k) {
        tos::Forest fr(ra[(size_t)k], dec[(size_t)k], (int)k, plate[(size_t)k]);
        for (int64_t p = off[(size_t)k]; p < off[(size_t)k + 1]; ++p)
            fr.pixels.push_back(tos::ForestPixel(delta[(size_t)p], loglam[(size_t)p], weight[(size_t)p], rc[(size_t)p]));
        forests.push_back(fr);
    }
    tos::HDF5Delta::saveForests(argv[2], "lines_of_sight", forests, z);
    std::cout << "wrote " << forests.size() << " sightlines to " << argv[2] << std::endl;
    return 0;
}
