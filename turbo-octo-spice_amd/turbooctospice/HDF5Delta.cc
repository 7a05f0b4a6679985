// HDF5Delta.cc -- HDF5 sightline reader/writer of the turbooctospice C++ API.
#include "HDF5Delta.h"

#include <H5Cpp.h>

#include <cmath>
#include <cstdio>
#include <iostream>

namespace turbooctospice {

namespace {

int g_nforests = 0;   // persists across calls like the reference's function static

template <typename T> void read_attr(H5::Group &grp, const char *name, T &value)
{
    H5::Attribute attr = grp.openAttribute(name);
    attr.read(attr.getDataType(), &value);
}

void read_floats(H5::Group &grp, const char *name, std::vector<float> &out)
{
    H5::DataSet ds = grp.openDataSet(name);
    H5::DataSpace sp = ds.getSpace();
    hsize_t dims[1] = {0};
    sp.getSimpleExtentDims(dims, nullptr);
    out.resize((size_t)dims[0]);
    if (dims[0] > 0) ds.read(out.data(), H5::PredType::NATIVE_FLOAT);
}

struct IterState {
    std::vector<Forest> *out;
    std::string error;
};

herr_t load_one(hid_t loc, const char *name, const H5L_info_t *, void *op)
{
    IterState *st = static_cast<IterState *>(op);
    try {
        H5::Group g(H5Gopen2(loc, name, H5P_DEFAULT));
        double ra, dec, z;
        int plate;
        read_attr(g, "ra", ra);
        read_attr(g, "dec", dec);
        read_attr(g, "z", z);
        read_attr(g, "plate", plate);
        if ((g_nforests % 10000) == 0) std::cout << g_nforests << " : " << name << std::endl;
        std::vector<float> delta, weight, r_comov, loglam;
        read_floats(g, "delta", delta);
        read_floats(g, "weight", weight);
        read_floats(g, "r_comov", r_comov);
        read_floats(g, "loglam", loglam);
        if (weight.size() != delta.size() || r_comov.size() != delta.size() || loglam.size() != delta.size())
            throw RuntimeError(std::string("HDF5Delta: datasets of ") + name + " differ in length");
        Forest f(ra, dec, g_nforests, plate);
        for (size_t i = 0; i < delta.size(); ++i)
            if (std::isfinite(delta[i]) && std::isfinite(weight[i]) && weight[i] > 0)
                f.pixels.push_back(ForestPixel(delta[i], loglam[i], weight[i], r_comov[i]));
        if (!f.pixels.empty()) {
            st->out->push_back(f);
            ++g_nforests;
        }
    } catch (H5::Exception const &e) {
        st->error = std::string("HDF5Delta: ") + name + ": " + e.getDetailMsg();
        return -1;
    } catch (std::exception const &e) {
        st->error = e.what();
        return -1;
    }
    return 0;
}

}  // namespace

HDF5Delta::HDF5Delta(std::string filename) : _filename(std::move(filename)) {}

std::vector<Forest> HDF5Delta::loadForests(std::string groupname)
{
    H5::Exception::dontPrint();
    H5::Group grp;
    H5::H5File file;
    try {
        file = H5::H5File(_filename, H5F_ACC_RDONLY);
        grp = file.openGroup(groupname);
    } catch (H5::Exception const &) {
        throw RuntimeError("HDF5Delta: Could not open the specified file.");
    }
    std::vector<Forest> forests;
    IterState st{&forests, ""};
    const herr_t rc = H5Literate(grp.getId(), H5_INDEX_NAME, H5_ITER_INC, nullptr, load_one, &st);
    if (rc < 0) throw RuntimeError(st.error.empty() ? "HDF5Delta: iteration failed" : st.error);
    return forests;
}

void HDF5Delta::saveForests(std::string const &filename, std::string const &groupname,
                            std::vector<Forest> const &forests, std::vector<double> const &redshifts)
{
    H5::H5File file(filename, H5F_ACC_TRUNC);
    H5::Group grp = file.createGroup(groupname);
    const int width = (int)std::to_string(forests.size()).size();
    for (size_t k = 0; k < forests.size(); ++k) {
        char name[64];
        std::snprintf(name, sizeof(name), "%0*zu", width, k);
        H5::Group g = grp.createGroup(name);
        const Forest &f = forests[k];
        H5::DataSpace scalar;
        const double z = k < redshifts.size() ? redshifts[k] : 0.0;
        g.createAttribute("ra", H5::PredType::NATIVE_DOUBLE, scalar).write(H5::PredType::NATIVE_DOUBLE, &f.ra);
        g.createAttribute("dec", H5::PredType::NATIVE_DOUBLE, scalar).write(H5::PredType::NATIVE_DOUBLE, &f.dec);
        g.createAttribute("z", H5::PredType::NATIVE_DOUBLE, scalar).write(H5::PredType::NATIVE_DOUBLE, &z);
        g.createAttribute("plate", H5::PredType::NATIVE_INT, scalar).write(H5::PredType::NATIVE_INT, &f.plate);
        const hsize_t n = f.pixels.size();
        H5::DataSpace sp(1, &n);
        std::vector<float> col(n);
        auto put = [&](const char *dname, float ForestPixel::*m) {
            for (size_t i = 0; i < n; ++i) col[i] = f.pixels[i].*m;
            g.createDataSet(dname, H5::PredType::NATIVE_FLOAT, sp).write(col.data(), H5::PredType::NATIVE_FLOAT);
        };
        put("delta", &ForestPixel::value);
        put("weight", &ForestPixel::weight);
        put("r_comov", &ForestPixel::distance);
        put("loglam", &ForestPixel::loglam);
    }
}

}  // namespace turbooctospice
