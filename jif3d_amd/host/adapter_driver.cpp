// Command-line driver of the C++ adapter: reads a problem from a binary file, runs
// Calculate and LQDerivative through HipGravMagImplementation (or, with --full,
// through HipFullSensitivityCalculator: a direct Calculate that stores J, then the
// cached Calculate and LQDerivative) and writes the results.
//
// input  (little-endian): int32 kind, int32 precision, int32 nx, ny, nz, nbg, nsites;
//        then doubles: xcoords[nx+1], ycoords[ny+1], zcoords[nz+1], xsizes[nx],
//        ysizes[ny], zsizes[nz], bg[nbg], sx[ns], sy[ns], sz[ns], field[3],
//        model[nx*ny*nz + nbg], misfit[ns*raw_per_site]
// output: doubles forward[ns*raw_per_site], gradient[nx*ny*nz + nbg]
#include <cstdint>
#include <cstdio>
#include <fstream>
#include <iostream>
#include <string>
#include <vector>

#include "hip_gravmag_imp.hpp"

namespace {
std::vector<double> read_doubles(std::ifstream& in, size_t n) {
  std::vector<double> v(n);
  in.read(reinterpret_cast<char*>(v.data()), static_cast<std::streamsize>(n * sizeof(double)));
  if (!in) throw jif3d_amd::FatalException("truncated input");
  return v;
}
}  // namespace

int main(int argc, char** argv) {
  const bool full = argc == 4 && std::string(argv[3]) == "--full";
  if (argc != 3 && !full) {
    std::cerr << "usage: adapter_driver <input> <output> [--full]\n";
    return 2;
  }
  try {
    std::ifstream in(argv[1], std::ios::binary);
    int32_t hdr[7];
    in.read(reinterpret_cast<char*>(hdr), sizeof(hdr));
    if (!in) throw jif3d_amd::FatalException("cannot read header");
    const auto kind = static_cast<jgm_kind>(hdr[0]);
    const auto prec = static_cast<jgm_precision>(hdr[1]);
    const size_t nx = hdr[2], ny = hdr[3], nz = hdr[4], nbg = hdr[5], ns = hdr[6];
    jif3d_amd::Mesh mesh;
    mesh.xcoords = read_doubles(in, nx + 1);
    mesh.ycoords = read_doubles(in, ny + 1);
    mesh.zcoords = read_doubles(in, nz + 1);
    mesh.xsizes = read_doubles(in, nx);
    mesh.ysizes = read_doubles(in, ny);
    mesh.zsizes = read_doubles(in, nz);
    mesh.bg_thicknesses = read_doubles(in, nbg);
    jif3d_amd::Sites sites;
    sites.x = read_doubles(in, ns);
    sites.y = read_doubles(in, ns);
    sites.z = read_doubles(in, ns);
    const std::vector<double> field = read_doubles(in, 3);
    const std::vector<double> model = read_doubles(in, nx * ny * nz + nbg);
    jif3d_amd::HipGravMagImplementation imp(kind, prec);
    if (kind == JGM_KIND_MAGNETIC) imp.SetField(field[0], field[1], field[2]);
    const std::vector<double> misfit = read_doubles(in, ns * imp.RawDataPerMeasurement());
    std::vector<double> fwd, grad;
    if (full) {
      jif3d_amd::HipFullSensitivityCalculator calc(imp);
      calc.Calculate(mesh, sites, model);         // direct; stores J
      fwd = calc.Calculate(mesh, sites, model);   // J m
      grad = calc.LQDerivative(mesh, sites, model, misfit);
    } else {
      fwd = imp.Calculate(mesh, sites, model);
      grad = imp.LQDerivative(mesh, sites, model, misfit);
    }
    std::ofstream out(argv[2], std::ios::binary);
    out.write(reinterpret_cast<const char*>(fwd.data()),
              static_cast<std::streamsize>(fwd.size() * sizeof(double)));
    out.write(reinterpret_cast<const char*>(grad.data()),
              static_cast<std::streamsize>(grad.size() * sizeof(double)));
    std::cout << "adapter_driver: " << fwd.size() << " data, " << grad.size() << " gradient\n";
    return 0;
  } catch (const std::exception& e) {
    std::cerr << "adapter_driver: " << e.what() << "\n";
    return 1;
  }
}
