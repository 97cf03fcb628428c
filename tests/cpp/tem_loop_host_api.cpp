// 1-D TEM time loop written as the reference's own user loop (tests/testTEM.cpp,
// with the functor calls of tests/testTM.cpp:102-113 for Ez and Hy): one
// sweep per functor over a YeeGrid<TEM> line, D and B soft sources applied
// between their curl and their material step, material functors per id.
// The scenario is the golden fixture tem1d_small (140 cells, 10-cell CPML at
// both ends, dielectric, Drude, Lorentz and conductive segments, 320 steps).
// Prints Ez, Hy at the seven probe cells, one step per line.
#include <cmath>
#include <cstdio>
#include <vector>

#include "fdtd_hip.hpp"

using namespace fdtd::hip;

int main() {
  const int nx = 140, nsteps = 320;
  const double dx = 1e-6, dt = 0.9 * dx / c0;
  YeeGrid<TEM, double> grid(nx, dx, dt);
  grid.loop().flush_every = 100;
  PMLParameterModel p(dx);
  grid.placePML<Dir::X>(p, 10);
  std::vector<uint8_t> ids(nx, 0);
  auto fill = [&](int x0, int x1, uint8_t m) {
    for (int i = x0; i < x1; i++) ids[i] = m;
  };
  fill(40, 52, 1);
  fill(60, 70, 2);
  fill(76, 88, 3);
  fill(96, 104, 4);
  grid.setMaterialIds(ids);
  const int probes[7] = {30, 45, 64, 80, 100, 110, 5};
  for (int q : probes) {
    grid.add_probe<Ez>(q);
    grid.add_probe<Hy>(q);
  }
  auto gauss = [](double n, double t0, double w) { double a = (n - t0) / w; return std::exp(-a * a); };
  auto srcD = make_soft_source<Dz>([&](double t) { return eps0 * gauss(t, 30, 8); }, grid.cell(30));
  auto srcB = make_soft_source<By>([&](double t) { return 1e-9 * gauss(t, 45, 9); }, grid.cell(110));
  for (int t = 0; t < nsteps; t++) {
    for_each(grid, YeeUpdate<TEM, Dz>(dt, dx));
    srcD.apply(t);
    for_each(grid, UpdatePML<TEM, FieldType::Electric, Dir::X>(dt, dx));
    for_each(grid, VacuumUpdate<TEM, FieldType::Electric>());
    for_each(grid.cells(1), ConstantUpdate<TEM, FieldType::Electric>(2.25));
    for_each(grid.cells(2), DrudeUpdate<TEM, FieldType::Electric>(1.5, 0.3 / dt, 0.02 / dt, dt));
    for_each(grid.cells(3), LorentzUpdate<TEM, FieldType::Electric>(2.0, 0.7, 0.2 / dt, 0.01 / dt, dt));
    for_each(grid.cells(4), ConductiveUpdate<TEM, FieldType::Electric>(1.2, 0.05 * 1.2 * eps0 / dt, dt));
    for_each(grid, YeeUpdate<TEM, By>(dt, dx));
    srcB.apply(t);
    for_each(grid, UpdatePML<TEM, FieldType::Magnetic, Dir::X>(dt, dx));
    for_each(grid, VacuumUpdate<TEM, FieldType::Magnetic>());
  }
  std::vector<double> pr = grid.probes();  // runs the steps still recorded
  for (int n = 0; n < nsteps; n++) {
    for (int q = 0; q < 14; q++) printf("%s%.17g", q ? " " : "", pr[size_t(n) * 14 + q]);
    printf("\n");
  }
  return 0;
}
