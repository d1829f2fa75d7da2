// Drives the nni.h API the way terranni.cpp's binRun does (terranni.cpp:394-426):
// a time-sorted TPIE stream of Point4D in, performNNI per time slice, OutPoints
// appended to a TPIE stream.  Used by tests/test_gpu_nni_api.py.
//
//   nni_api_test <points.bin> <out.bin> key=value ...
// points.bin: raw nni::Point4D records; out.bin: raw nni::OutPoint records.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <string>

#include "nni.h"

int main(int argc, char **argv)
{
    if (argc < 3) {
        std::fprintf(stderr, "usage: %s points.bin out.bin key=value...\n", argv[0]);
        return 2;
    }
    std::map<std::string, std::string> kv;
    for (int a = 3; a < argc; ++a) {
        const char *eq = std::strchr(argv[a], '=');
        if (eq) kv[std::string(argv[a], eq - argv[a])] = eq + 1;
    }
    auto num = [&kv](const char *k, double dflt) { return kv.count(k) ? std::atof(kv[k].c_str()) : dflt; };

    nni::Settings s;
    s.scale = (int)num("scale", 1);
    s.cell_size = (float)num("cell_size", 0.5);
    s.nrows = (int)num("nrows", 64);
    s.ncols = (int)num("ncols", 64);
    s.site_radius = (float)num("site_radius", 5.0);
    s.query_radius = num("query_radius", 5 * s.cell_size / 2.0 - 0.01);  // terranni.cpp:663
    s.algorithm = (int)num("algorithm", 0);
    s.pointMin = (int)num("point_min", 0);
    s.multiplier = (int)num("multiplier", 1);
    nni::WorkPackage wp;
    wp.origin_x = (float)num("origin_x", 0);
    wp.origin_y = (float)num("origin_y", 0);
    wp.grid_origin_x = (float)num("grid_origin_x", 0);
    wp.grid_origin_y = (float)num("grid_origin_y", 0);
    wp.nrows = (int)num("wp_nrows", s.nrows);
    wp.ncols = (int)num("wp_ncols", s.ncols);
    const int slices = (int)num("slices", 1);
    const float time_radius = (float)num("time_radius", 0);  // space-time when > 0
    const bool delta = num("delta", 1) != 0;                  // useDelta (0: Delta^2)
    if (kv.count("accum")) nni::setAccumulationMode((int)num("accum", 0));

    tpie::ami::stream<nni::Point4D> in;
    std::FILE *f = std::fopen(argv[1], "rb");
    if (!f) return 3;
    nni::Point4D p;
    while (std::fread(&p, sizeof p, 1, f) == 1) in.write_item(p);
    std::fclose(f);
    in.seek(0);

    if (nni::init(s) == 0) return 4;
    tpie::ami::stream<nni::OutPoint> out;
    for (int t = 0; t < slices; ++t) {
        nni::performNNI(wp, out, in, time_radius, (float)t, delta);
        if (nni::lastStatus() != 0) return 5;
    }
    nni::printTimes();
    nni::cleanup();

    std::FILE *g = std::fopen(argv[2], "wb");
    if (!g) return 6;
    out.seek(0);
    nni::OutPoint *o;
    while (out.read_item(&o) == tpie::ami::NO_ERROR) std::fwrite(o, sizeof *o, 1, g);
    std::fclose(g);
    return 0;
}
