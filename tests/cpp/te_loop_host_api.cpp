// 2-D TE time loop written as the reference's own user loop
// (tests/testTE.cpp, with the functor calls of tests/testTM.cpp:102-113 for
// the TE components): one sweep per functor, D and B soft sources applied
// between their curl and their material step.
//   te_loop_host_api small: the golden fixture te2d_small (48 x 40, 10-cell
//     CPML on x and y, dielectric, Drude and Lorentz blocks, a Gaussian on Dx
//     and the pulse of tests/testTE.cpp on Bz, 150 steps)
//   te_loop_host_api tfsf: the golden fixture te2d_tfsf (64 x 60, an oblique
//     TE TFSF box inset 14 around a dielectric block, 160 steps)
// Prints Ex, Ey, Hz at the probe cells, one step per line.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <memory>
#include <vector>

#include "fdtd_hip.hpp"

using namespace fdtd::hip;

int main(int argc, char** argv) {
  const bool tfsf = argc > 1 && !strcmp(argv[1], "tfsf");
  const int nx = tfsf ? 64 : 48, ny = tfsf ? 60 : 40, nsteps = tfsf ? 160 : 150;
  const double dx = 1e-6, dt = (tfsf ? 0.5 : 0.9) * dx / (c0 * std::sqrt(2.0));
  YeeGrid<TE, double> grid(nx, ny, dx, dt);
  grid.loop().flush_every = 64;
  PMLParameterModel p(dx);
  grid.placePML<Dir::X>(p, 10);
  grid.placePML<Dir::Y>(p, 10);
  std::vector<uint8_t> ids(size_t(nx) * ny, 0);
  auto fill = [&](int x0, int x1, int y0, int y1, int m) {
    for (int j = y0; j < y1; j++)
      for (int i = x0; i < x1; i++) ids[size_t(i) + size_t(nx) * j] = uint8_t(m);
  };
  std::vector<std::array<int, 2>> probes;
  if (tfsf) {
    fill(28, 36, 26, 33, 1);
    probes = {{30, 30}, {15, 30}, {14, 30}, {13, 30}, {49, 20}, {50, 20}, {31, 14}, {31, 13}};
  } else {
    fill(14, 22, 13, 19, 1);
    fill(27, 33, 22, 29, 2);
    fill(15, 20, 24, 30, 3);
    fill(28, 34, 12, 17, 4);
    probes = {{24, 20}, {17, 15}, {30, 25}, {12, 30}, {3, 3}, {20, 25}};
  }
  grid.setMaterialIds(ids);
  for (auto& q : probes) {
    grid.add_probe<Ex>(q[0], q[1]);
    grid.add_probe<Ey>(q[0], q[1]);
    grid.add_probe<Hz>(q[0], q[1]);
  }
  auto gauss = [](double n, double t0, double w) { double a = (n - t0) / w; return std::exp(-a * a); };
  auto srcD = make_soft_source<Dx>([&](double t) { return eps0 * gauss(t, 25, 8); }, grid.cell(24, 20));
  auto srcB = make_soft_source<Bz>(
      [&](double t) { return std::exp(-(t - 10.0) * (t - 10.0) * c0 / (20 * dx) * dt * dt) * 1e-9; },
      grid.cell(20, 25));
  const double f0 = c0 / (20 * dx), tw = 2.0 / f0, t0 = 3.0 * tw;
  auto pulse = [f0, tw, t0](double t) {
    double a = (t - t0) / tw;
    return std::exp(-a * a) * std::sin(2 * pi * f0 * (t - t0));
  };
  const std::array<double, 3> k = {std::cos(pi / 6), std::sin(pi / 6), 0.0};
  auto box = make_tfsf_box<TE>(grid, {14, 49, 14, 45, 0, 0}, k, std::make_shared<Aux1D<double>>(100, dx, dt, pulse));
  for (int t = 0; t < nsteps; t++) {
    for_each(grid, YeeUpdate<TE, Dx>(dt, dx));
    for_each(grid, YeeUpdate<TE, Dy>(dt, dx));
    if (!tfsf) srcD.apply(t);
    for_each(grid, UpdatePML<TE, FieldType::Electric, Dir::X>(dt, dx));
    for_each(grid, UpdatePML<TE, FieldType::Electric, Dir::Y>(dt, dx));
    if (tfsf) box.updateD(dt, dx);
    for_each(grid, VacuumUpdate<TE, FieldType::Electric>());
    for_each(grid.cells(1), ConstantUpdate<TE, FieldType::Electric>(4.0));
    if (!tfsf) {
      for_each(grid.cells(2), ConstantUpdate<TE, FieldType::Electric>(2.25));
      for_each(grid.cells(3), DrudeUpdate<TE, FieldType::Electric>(1.5, 0.3 / dt, 0.02 / dt, dt));
      for_each(grid.cells(4), LorentzUpdate<TE, FieldType::Electric>(2.0, 0.7, 0.2 / dt, 0.01 / dt, dt));
    }
    for_each(grid, YeeUpdate<TE, Bz>(dt, dx));
    if (!tfsf) srcB.apply(t);
    for_each(grid, UpdatePML<TE, FieldType::Magnetic, Dir::X>(dt, dx));
    for_each(grid, UpdatePML<TE, FieldType::Magnetic, Dir::Y>(dt, dx));
    if (tfsf) box.updateB(dt, dx);
    for_each(grid, VacuumUpdate<TE, FieldType::Magnetic>());
  }
  std::vector<double> pr = grid.probes();  // runs the steps still recorded
  const size_t np = 3 * probes.size();
  for (int n = 0; n < nsteps; n++) {
    for (size_t q = 0; q < np; q++) printf("%s%.17g", q ? " " : "", pr[size_t(n) * np + q]);
    printf("\n");
  }
  return 0;
}
