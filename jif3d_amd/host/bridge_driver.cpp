// Command-line driver of HipGravMagImp behind the jif3D calculators (jif3d_bridge.hpp):
// the base class's call order -- per-site CalcGridded + CalcBackground, the calculator's
// data transform, HandleSensitivities into the calculator's row-major J -- for one problem
// read from a binary file.
//
// input (little-endian): int32 kind, precision, nx, ny, nz, nbg, nsites, calculator
//        (0 MinMem, 1 FullSensitivity), transform (jgm_transform codes); then doubles:
//        origin[3], xsizes[nx], ysizes[ny], zsizes[nz], bg thicknesses[nbg], bg values[nbg],
//        sx[ns], sy[ns], sz[ns], field[3], transform angles[2], cell values[nx ny nz],
//        misfit[ns * data per site]
// output: doubles: the gradient before any Calculate (grad0), the first Calculate (fwd1),
//        a second Calculate (fwd2: the cached J m for FullSensitivity), the gradient after
//        it (grad1: cached for FullSensitivity), and for FullSensitivity the stored J
//        (ns * raw per site rows of nx ny nz + nbg columns, row-major)
#include <cstdint>
#include <fstream>
#include <iostream>
#include <memory>
#include <string>
#include <vector>

#include "jif3d_bridge.hpp"
#include "hip_bridge_imp.hpp"

namespace {

std::vector<double> read_doubles(std::ifstream& in, size_t n) {
  std::vector<double> v(n);
  in.read(reinterpret_cast<char*>(v.data()), static_cast<std::streamsize>(n * sizeof(double)));
  if (!in) throw jif3D::FatalException("truncated input");
  return v;
}

void put(std::ofstream& out, const jif3D::rvec& v) {
  out.write(reinterpret_cast<const char*>(v.begin()),
            static_cast<std::streamsize>(v.size() * sizeof(double)));
}

std::shared_ptr<jif3D::VectorTransform> make_transform(int tf, double inc, double dec) {
  switch (tf) {
    case JGM_TF_FTG_INVARIANT: return std::make_shared<jif3D::FTGInvariant>();
    case JGM_TF_TOTAL_FIELD: return std::make_shared<jif3D::TotalField>();
    case JGM_TF_TOTAL_FIELD_ANOMALY: return std::make_shared<jif3D::TotalFieldAnomaly>(inc, dec);
    default: return nullptr;
  }
}

template <class DataType>
int run(std::ifstream& in, std::ofstream& out, const int32_t* h) {
  const int kind = h[0], prec = h[1], nx = h[2], ny = h[3], nz = h[4], nbg = h[5], ns = h[6];
  const int calc_type = h[7], tf = h[8];
  const std::vector<double> origin = read_doubles(in, 3);
  const std::vector<double> xs = read_doubles(in, nx), ys = read_doubles(in, ny),
                            zs = read_doubles(in, nz);
  const std::vector<double> bgt = read_doubles(in, nbg), bgv = read_doubles(in, nbg);
  const std::vector<double> sx = read_doubles(in, ns), sy = read_doubles(in, ns),
                            sz = read_doubles(in, ns);
  const std::vector<double> field = read_doubles(in, 3), angles = read_doubles(in, 2);
  jif3D::GravMagModel model;
  model.SetMeshSize(nx, ny, nz);
  model.SetOrigin(origin[0], origin[1], origin[2]);
  model.SetXCellSizes(xs);
  model.SetYCellSizes(ys);
  model.SetZCellSizes(zs);
  model.SetBackgroundThicknesses(bgt);
  model.SetBackgroundValues(bgv);
  model.SetValues() = read_doubles(in, static_cast<size_t>(nx) * ny * nz);
  DataType data;
  for (int i = 0; i < ns; ++i) data.AddMeasurementPoint(sx[i], sy[i], sz[i]);

  auto imp = std::make_shared<jif3D::HipGravMagImp<DataType>>(static_cast<jgm_kind>(kind),
                                                              static_cast<jgm_precision>(prec));
  if (kind == JGM_KIND_MAGNETIC) imp->SetField(field[0], field[1], field[2]);
  std::unique_ptr<jif3D::ThreeDGravMagCalculator<DataType>> calc;
  jif3D::FullSensitivityGravMagCalculator<DataType>* full = nullptr;
  if (calc_type == 1) {
    full = new jif3D::FullSensitivityGravMagCalculator<DataType>(imp);
    calc.reset(full);
  } else {
    calc.reset(new jif3D::MinMemGravMagCalculator<DataType>(imp));
  }
  if (auto t = make_transform(tf, angles[0], angles[1])) calc->SetDataTransform(t);
  const size_t kt = calc->GetDataPerMeasurement();
  const std::vector<double> mis = read_doubles(in, static_cast<size_t>(ns) * kt);
  jif3D::rvec misfit(mis.size());
  for (size_t i = 0; i < mis.size(); ++i) misfit(i) = mis[i];

  put(out, calc->LQDerivative(model, data, misfit));
  put(out, calc->Calculate(model, data));
  put(out, calc->Calculate(model, data));
  put(out, calc->LQDerivative(model, data, misfit));
  if (full) {
    const jif3D::rmat& J = full->GetSensitivities();
    std::vector<double> row(J.size2());
    for (size_t r = 0; r < J.size1(); ++r) {
      for (size_t c = 0; c < J.size2(); ++c) row[c] = J(r, c);
      out.write(reinterpret_cast<const char*>(row.data()),
                static_cast<std::streamsize>(row.size() * sizeof(double)));
    }
  }
  return 0;
}

}  // namespace

int main(int argc, char** argv) {
  if (argc != 3) {
    std::cerr << "usage: bridge_driver input.bin output.bin\n";
    return 2;
  }
  try {
    std::ifstream in(argv[1], std::ios::binary);
    int32_t h[9];
    in.read(reinterpret_cast<char*>(h), sizeof(h));
    if (!in) throw jif3D::FatalException("truncated header");
    std::ofstream out(argv[2], std::ios::binary);
    switch (h[0]) {
      case JGM_KIND_SCALAR: return run<jif3D::ScalarGravityData>(in, out, h);
      case JGM_KIND_TENSOR: return run<jif3D::TensorGravityData>(in, out, h);
      case JGM_KIND_MAGNETIC: return run<jif3D::TotalFieldMagneticData>(in, out, h);
      default: throw jif3D::FatalException("kind not covered by this driver");
    }
  } catch (const std::exception& e) {
    std::cerr << "bridge_driver: " << e.what() << "\n";
    return 1;
  }
}
