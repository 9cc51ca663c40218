// Command-line driver of the C++ adapter: reads a problem from a binary file, runs
// Calculate and LQDerivative through HipGravMagImplementation (or, with --full,
// through HipFullSensitivityCalculator: a direct Calculate that stores J, then the
// cached Calculate and LQDerivative) and writes the results.
//
// --batch B: B models at once, model b = model * (1 + 0.05 b) and misfit b =
//            misfit * (b + 1), through CalculateBatch / LQDerivativeBatch.
// --objective: with --batch B, GravMagObjective over the same calculator with the
//            misfit vector as observed data (errors 1): writes B chi^2 values, then the
//            B gradients.
// --middle-sens: writes the depth profile of CalculateMiddleSens (nz values), then the
//            fitted z0 for the exponent -2 and the nz depth weights.
//
// input  (little-endian): int32 kind, int32 precision, int32 nx, ny, nz, nbg, nsites;
//        then doubles: xcoords[nx+1], ycoords[ny+1], zcoords[nz+1], xsizes[nx],
//        ysizes[ny], zsizes[nz], bg[nbg], sx[ns], sy[ns], sz[ns], field[3],
//        model[nx*ny*nz + nbg] (3*nx*ny*nz for the magnetization kind),
//        misfit[ns*raw_per_site]
// output: doubles forward[ns*raw_per_site], gradient[model length] (per model for
//         --batch, all forwards first)
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <iostream>
#include <string>
#include <vector>

#include "depth_weighting.hpp"
#include "gravmag_objective.hpp"
#include "hip_gravmag_imp.hpp"

namespace {
std::vector<double> read_doubles(std::ifstream& in, size_t n) {
  std::vector<double> v(n);
  in.read(reinterpret_cast<char*>(v.data()), static_cast<std::streamsize>(n * sizeof(double)));
  if (!in) throw jif3d_amd::FatalException("truncated input");
  return v;
}
}  // namespace

template <class Calc>
void run(Calc& calc, const jif3d_amd::Mesh& mesh, const jif3d_amd::Sites& sites,
         const std::vector<double>& model, const std::vector<double>& misfit, size_t nbatch,
         bool objective, bool prime, std::vector<double>* first, std::vector<double>* second) {
  if (prime) calc.Calculate(mesh, sites, model);  // direct; stores J
  if (nbatch == 0) {
    *first = calc.Calculate(mesh, sites, model);
    *second = calc.LQDerivative(mesh, sites, model, misfit);
    return;
  }
  std::vector<double> models, misfits;
  for (size_t b = 0; b < nbatch; ++b) {
    for (double v : model) models.push_back(v * (1.0 + 0.05 * b));
    for (double v : misfit) misfits.push_back(v * (b + 1.0));
  }
  if (objective) {
    jif3d_amd::GravMagObjective<Calc> obj(calc, mesh, sites);
    obj.SetObservedData(misfit, std::vector<double>(misfit.size(), 1.0));
    *first = obj.CalcMisfitBatch(models, nbatch);
    *second = obj.CalcGradientBatch(models, nbatch);
    return;
  }
  *first = calc.CalculateBatch(mesh, sites, models, nbatch);
  *second = calc.LQDerivativeBatch(mesh, sites, models, misfits, nbatch);
}

int main(int argc, char** argv) {
  bool full = false, objective = false, middle = false, ok = argc >= 3;
  size_t nbatch = 0;
  for (int a = 3; a < argc && ok; ++a) {
    const std::string opt(argv[a]);
    if (opt == "--full") {
      full = true;
    } else if (opt == "--objective") {
      objective = true;
    } else if (opt == "--middle-sens") {
      middle = true;
    } else if (opt == "--batch" && a + 1 < argc) {
      char* end = nullptr;
      nbatch = std::strtoul(argv[++a], &end, 10);
      ok = nbatch > 0 && end && *end == '\0';
    } else {
      ok = false;
    }
  }
  if (!ok || (objective && nbatch == 0)) {
    std::cerr << "usage: adapter_driver <input> <output> [--full] [--batch B [--objective]] "
                 "[--middle-sens]\n";
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
    const size_t nmodel =
        kind == JGM_KIND_MAGNETIZATION ? 3 * nx * ny * nz : nx * ny * nz + nbg;
    const std::vector<double> model = read_doubles(in, nmodel);
    jif3d_amd::HipGravMagImplementation imp(kind, prec);
    if (kind == JGM_KIND_MAGNETIC || kind == JGM_KIND_MAGNETIC_GRADIENT)
      imp.SetField(field[0], field[1], field[2]);
    const std::vector<double> misfit = read_doubles(in, ns * imp.RawDataPerMeasurement());
    std::vector<double> fwd, grad;
    if (middle) {
      fwd = jif3d_amd::CalculateMiddleSens(imp, mesh, model);
      const jif3d_amd::WeightingTerm f(-2.0);
      const double z0 = jif3d_amd::FitZ0(fwd, mesh.zsizes, f);
      grad = jif3d_amd::ConstructDepthWeighting(mesh.zsizes, z0, f);
      grad.insert(grad.begin(), z0);
    } else if (full) {
      jif3d_amd::HipFullSensitivityCalculator calc(imp);
      run(calc, mesh, sites, model, misfit, nbatch, objective, true, &fwd, &grad);
    } else {
      run(imp, mesh, sites, model, misfit, nbatch, objective, false, &fwd, &grad);
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
