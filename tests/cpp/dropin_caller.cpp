// A caller written against the reference's stp:: API (src/stp/stp.h), built against
// include/stp/stp.h and linked with libstp_hip.so.  It reads one snapshot from a
// raw file, calls the imager the way the reduce driver does (reduce.cpp:129-139:
// ImageVisibilities from ImagerPars, then SourceFindImage on vis_grid) and the
// functor form of image_visibilities (imager.h:233-241), models the visibilities of a
// sky model (visibility.cpp:31-41), and writes every result to raw files.
//
// usage: dropin_caller <input> <output prefix>
// input: int64 n, int64 image_size, float64 cell_arcsec, int64 n_src,
//        float64 uvw[n*3] (column-major), float64 vis[2n], float64 w[n],
//        float64 skymodel[n_src*3] (column-major)
#include <cstdio>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "stp/stp.h"

template <typename T>
static void read_n(FILE* f, T* p, size_t n)
{
    if (std::fread(p, sizeof(T), n, f) != n)
        throw std::runtime_error("short input file");
}

template <typename T>
static void write_file(const std::string& fn, const T* p, size_t n)
{
    FILE* f = std::fopen(fn.c_str(), "wb");
    if (!f || std::fwrite(p, sizeof(T), n, f) != n)
        throw std::runtime_error("cannot write " + fn);
    std::fclose(f);
}

int main(int argc, char** argv)
{
    if (argc != 3) {
        std::fprintf(stderr, "usage: %s <input> <output prefix>\n", argv[0]);
        return 2;
    }
    try {
        FILE* f = std::fopen(argv[1], "rb");
        if (!f)
            throw std::runtime_error("cannot open input");
        int64_t n = 0, N = 0, n_src = 0;
        double cell = 0;
        read_n(f, &n, 1);
        read_n(f, &N, 1);
        read_n(f, &cell, 1);
        read_n(f, &n_src, 1);
        stp::mat uvw(n, 3), w(n, 1), sky(n_src, 3);
        stp::cx_mat vis(n, 1);
        read_n(f, uvw.memptr(), size_t(n) * 3);
        read_n(f, reinterpret_cast<double*>(vis.memptr()), size_t(n) * 2);
        read_n(f, w.memptr(), size_t(n));
        read_n(f, sky.memptr(), size_t(n_src) * 3);
        std::fclose(f);
        const std::string out = argv[2];

        // the reduce driver's imager call: GaussianSinc, support 3, oversampling 9, beam
        stp::ImagerPars ip(uint32_t(N), cell, 1.0, stp::KernelFunction::GaussianSinc, 3, false, 9, true, false, false);
        stp::ImageVisibilities imager(vis, w, uvw, ip);
        write_file(out + "image.f64", imager.vis_grid.memptr(), imager.vis_grid.n_elem);
        write_file(out + "beam.f64", imager.sampling_grid.memptr(), imager.sampling_grid.n_elem);

        // the functor form with an explicit kernel_creator
        auto r = stp::image_visibilities(stp::GaussianSinc(3), vis, w, uvw, ip);
        write_file(out + "image2.f64", r.first.memptr(), r.first.n_elem);

        // the source finder on the image, BINMEDIAN, 5 / 3 sigma
        stp::SourceFindImage sf(imager.vis_grid, 5.0, 3.0, 0.0, true, 5, stp::MedianMethod::BINMEDIAN);
        write_file(out + "labels.i32", sf.label_map.memptr(), sf.label_map.n_elem);
        std::vector<double> isl;
        for (const auto& i : sf.islands) {
            isl.push_back(i.label_idx);
            isl.push_back(i.sign);
            isl.push_back(i.extremum_y_idx);
            isl.push_back(i.extremum_x_idx);
            isl.push_back(i.num_samples);
            isl.push_back(i.extremum_val);
        }
        isl.push_back(sf.rms_est);
        isl.push_back(sf.bg_level);
        write_file(out + "islands.f64", isl.data(), isl.size());

        // model visibilities of the sky model
        stp::cx_mat model = stp::generate_visibilities_from_local_skymodel(sky, uvw);
        write_file(out + "model.f64", reinterpret_cast<const double*>(model.memptr()), size_t(n) * 2);
        std::printf("dropin_caller: %lld visibilities, %zu islands\n", (long long)n, sf.islands.size());
    } catch (const std::exception& e) {
        std::fprintf(stderr, "dropin_caller: %s\n", e.what());
        return 1;
    }
    return 0;
}
