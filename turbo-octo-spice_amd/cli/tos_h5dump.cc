// tos_h5dump -- reads an HDF5 delta file with HDF5Delta::loadForests (the
// reference's loader semantics: name order, pixel filter, empty sightlines
// dropped) and writes the raw catalogue dump that tos_mkdelta reads
// (z is written as 0; HDF5Delta does not keep it).
#include <cstdint>
#include <cstdio>
#include <iostream>
#include <vector>

#include "turbooctospice.h"

namespace tos = turbooctospice;

int main(int argc, char **argv)
{
    if (argc != 3) {
        std::cerr << "usage: tos_h5dump <in.h5> <catalogue.bin>" << std::endl;
        return 2;
    }
    std::vector<tos::Forest> f;
    try {
        f = tos::HDF5Delta(argv[1]).loadForests("lines_of_sight");
    } catch (std::exception const &e) {
        std::cerr << e.what() << std::endl;
        return 1;
    }
    std::vector<int64_t> off(1, 0);
    std::vector<double> ra, dec, z;
    std::vector<int32_t> plate;
    std::vector<float> v, l, w, d;
    for (const auto &s : f) {
        ra.push_back(s.ra);
        dec.push_back(s.dec);
        z.push_back(0.0);
        plate.push_back(s.plate);
        for (const auto &p : s.pixels) {
            v.push_back(p.value);
            l.push_back(p.loglam);
            w.push_back(p.weight);
            d.push_back(p.distance);
        }
        off.push_back((int64_t)v.size());
    }
    std::FILE *o = std::fopen(argv[2], "wb");
    if (!o) return 2;
    const int64_t hdr[2] = {(int64_t)f.size(), (int64_t)v.size()};
    std::fwrite(hdr, sizeof(int64_t), 2, o);
    std::fwrite(off.data(), sizeof(int64_t), off.size(), o);
    std::fwrite(ra.data(), sizeof(double), ra.size(), o);
    std::fwrite(dec.data(), sizeof(double), dec.size(), o);
    std::fwrite(z.data(), sizeof(double), z.size(), o);
    std::fwrite(plate.data(), sizeof(int32_t), plate.size(), o);
    for (auto *a : {&v, &l, &w, &d}) std::fwrite(a->data(), sizeof(float), a->size(), o);
    std::fclose(o);
    return 0;
}
