// Drives libibm through the reference-shaped C++ API (include/ibm.hpp) the
// way reference code calls ib::spread / ib::interpolate / solvers::multigrid
// / ins::diffusion, for the GPU test tests/test_gpu_cpp_api.py.
//
// usage: api_check <dir>
//   reads  <dir>/x.bin, <dir>/f.bin   (3n doubles each, SoA)
//   writes <dir>/fe0.bin fe1.bin fe2.bin (spread), ul.bin (interpolate),
//          <dir>/helm.bin (multigrid solve of (I~ - 0.4 h^2 L) x = fe0),
//          <dir>/diff.bin (ins::diffusion of u = fe0 with f = 0),
//          <dir>/out.txt  (iteration counts)
// on the C1 configuration: 16 um periodic box, fd::mac{32}, cosine delta.
#include <hip/hip_runtime.h>
#include <cstdio>
#include <fstream>
#include <string>
#include <vector>

#include "ibm.hpp"

static std::vector<double> read(const std::string& path)
{
    std::ifstream f(path, std::ios::binary | std::ios::ate);
    std::streamsize bytes = f.tellg();
    f.seekg(0);
    std::vector<double> v(bytes / sizeof(double));
    f.read(reinterpret_cast<char*>(v.data()), bytes);
    return v;
}

static void write(const std::string& path, const double* dev, size_t n)
{
    std::vector<double> h(n);
    if (hipMemcpy(h.data(), dev, n * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess) throw std::runtime_error("copy");
    std::ofstream f(path, std::ios::binary);
    f.write(reinterpret_cast<const char*>(h.data()), n * sizeof(double));
}

static double* upload(const std::vector<double>& h)
{
    double* d = nullptr;
    if (hipMalloc(&d, h.size() * sizeof(double)) != hipSuccess) throw std::runtime_error("alloc");
    if (hipMemcpy(d, h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice) != hipSuccess) throw std::runtime_error("copy");
    return d;
}

static double* zeros(size_t n) { return upload(std::vector<double>(n, 0.0)); }

int main(int argc, char** argv)
{
    if (argc != 2) {
        std::fprintf(stderr, "usage: %s <dir>\n", argv[0]);
        return 2;
    }
    const std::string dir = argv[1];
    try {
        auto x = read(dir + "/x.bin");
        auto f = read(dir + "/f.bin");
        const int n = (int) (x.size() / 3);
        double* dx = upload(x);
        double* df = upload(f);

        constexpr double um = 0.1;
        fd::domain domain{fd::dimension{16 * um, fd::boundary::periodic},
                          fd::dimension{16 * um, fd::boundary::periodic},
                          fd::dimension{16 * um, fd::boundary::periodic}};
        constexpr fd::mac mac{32};
        auto grids = fd::components(mac, domain);
        std::array<double*, 3> fe;
        for (int c = 0; c < 3; ++c) fe[c] = zeros(grids[c].size());
        double* ul = zeros(3 * (size_t) n);

        ib::spread spread{mac, domain, ib::delta::cosine{}};
        spread(n, dx, df, fe);
        ib::interpolate interpolate{mac, domain, ib::delta::cosine{}};
        interpolate(n, dx, {fe[0], fe[1], fe[2]}, ul);

        const double h = 1.0 / grids[0].g.axis[0].resolution;
        const double lambda = 0.4 * h * h;
        solvers::multigrid mg{fd::helmholtz(grids[0], lambda), 1e-10, 32};
        double* xs = zeros(grids[0].size());
        int it = mg(fe[0], xs);

        ins::parameters params{1e-4, 10.0, 1.0, 1e-10};
        ins::diffusion diffusion{grids[0], params};
        double* rhs = zeros(grids[0].size());
        double* force = zeros(grids[0].size());
        double* du = zeros(grids[0].size());
        double* un = zeros(grids[0].size());
        diffusion(1.0, fe[0], rhs, force, du, un);
        ibm::context::current().synchronize();

        for (int c = 0; c < 3; ++c) write(dir + "/fe" + std::to_string(c) + ".bin", fe[c], grids[c].size());
        write(dir + "/ul.bin", ul, 3 * (size_t) n);
        write(dir + "/helm.bin", xs, grids[0].size());
        write(dir + "/diff.bin", un, grids[0].size());
        std::ofstream(dir + "/out.txt") << "multigrid_iterations " << it << "\n";
        std::printf("api_check ok: %d markers, %d V-cycles\n", n, it);
    } catch (const std::exception& e) {
        std::fprintf(stderr, "api_check failed: %s\n", e.what());
        return 1;
    }
    return 0;
}
