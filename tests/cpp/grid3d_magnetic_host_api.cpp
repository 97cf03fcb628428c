// 3-D 22 x 20 x 18 time loop with magnetic-side pole materials, written as the
// reference's own user loop (tests/test3D.cpp:146-171) against the C++ host
// surface: one for_each per functor, material functors per id on both the
// electric and the magnetic side (DrudeUpdate / LorentzUpdate /
// ConductiveUpdate on FieldType::Magnetic step B, H, M, K), Gaussian soft
// sources on Dz and Bx applied between their curl and their material step.
// The scenario is the golden fixture grid3d_magnetic.  Prints Ex, Ey, Ez, Hx,
// Hy, Hz at the five probe cells, one step per line.
#include <cmath>
#include <cstdio>
#include <vector>

#include "fdtd_hip.hpp"

using namespace fdtd::hip;

int main() {
  const int nx = 22, ny = 20, nz = 18, nsteps = 130;
  const double dx = 1e-6, dt = 0.5 * dx / (c0 * std::sqrt(3.0));
  YeeGrid<ThreeD, double> grid(nx, ny, nz, dx, dt);
  grid.loop().flush_every = 50;
  PMLParameterModel p(dx);
  grid.placePML<Dir::X>(p, 5);
  grid.placePML<Dir::Y>(p, 5);
  grid.placePML<Dir::Z>(p, 5);
  std::vector<uint8_t> ids(size_t(nx) * ny * nz, 0);
  auto fill = [&](int x0, int x1, int y0, int y1, int z0, int z1, uint8_t m) {
    for (int k = z0; k < z1; k++)
      for (int j = y0; j < y1; j++)
        for (int i = x0; i < x1; i++) ids[size_t(i) + size_t(nx) * (size_t(j) + size_t(ny) * k)] = m;
  };
  fill(7, 11, 6, 10, 6, 11, 1);
  fill(12, 16, 7, 12, 7, 10, 2);
  fill(8, 13, 11, 14, 9, 13, 3);
  fill(13, 16, 13, 16, 5, 9, 4);
  fill(5, 8, 11, 15, 11, 14, 5);
  grid.setMaterialIds(ids);
  const int probes[5][3] = {{10, 9, 8}, {14, 10, 8}, {9, 8, 8}, {14, 14, 7}, {6, 13, 12}};
  for (auto& q : probes) {
    grid.add_probe<Ex>(q[0], q[1], q[2]);
    grid.add_probe<Ey>(q[0], q[1], q[2]);
    grid.add_probe<Ez>(q[0], q[1], q[2]);
    grid.add_probe<Hx>(q[0], q[1], q[2]);
    grid.add_probe<Hy>(q[0], q[1], q[2]);
    grid.add_probe<Hz>(q[0], q[1], q[2]);
  }
  auto gauss = [](double n, double t0, double w) { double a = (n - t0) / w; return std::exp(-a * a); };
  auto srcD = make_soft_source<Dz>([&](double t) { return eps0 * gauss(t, 20, 6); }, grid.cell(10, 9, 8));
  auto srcB = make_soft_source<Bx>([&](double t) { return 1e-9 * gauss(t, 30, 7); }, grid.cell(14, 10, 8));
  for (int t = 0; t < nsteps; t++) {
    for_each(grid, YeeUpdate<ThreeD, Dx>(dt, dx));
    for_each(grid, YeeUpdate<ThreeD, Dy>(dt, dx));
    for_each(grid, YeeUpdate<ThreeD, Dz>(dt, dx));
    srcD.apply(t);
    for_each(grid, UpdatePML<ThreeD, FieldType::Electric, Dir::X>(dt, dx));
    for_each(grid, UpdatePML<ThreeD, FieldType::Electric, Dir::Y>(dt, dx));
    for_each(grid, UpdatePML<ThreeD, FieldType::Electric, Dir::Z>(dt, dx));
    for_each(grid, VacuumUpdate<ThreeD, FieldType::Electric>());
    for_each(grid.cells(1), ConstantUpdate<ThreeD, FieldType::Electric>(2.0));
    for_each(grid.cells(2), LorentzUpdate<ThreeD, FieldType::Electric>(1.5, 0.6, 0.2 / dt, 0.02 / dt, dt));
    for_each(grid.cells(3), ConstantUpdate<ThreeD, FieldType::Electric>(1.0));
    for_each(grid.cells(4), DrudeUpdate<ThreeD, FieldType::Electric>(1.2, 0.25 / dt, 0.03 / dt, dt));
    for_each(grid.cells(5), ConstantUpdate<ThreeD, FieldType::Electric>(3.0));
    for_each(grid, YeeUpdate<ThreeD, Bx>(dt, dx));
    for_each(grid, YeeUpdate<ThreeD, By>(dt, dx));
    for_each(grid, YeeUpdate<ThreeD, Bz>(dt, dx));
    srcB.apply(t);
    for_each(grid, UpdatePML<ThreeD, FieldType::Magnetic, Dir::X>(dt, dx));
    for_each(grid, UpdatePML<ThreeD, FieldType::Magnetic, Dir::Y>(dt, dx));
    for_each(grid, UpdatePML<ThreeD, FieldType::Magnetic, Dir::Z>(dt, dx));
    for_each(grid, VacuumUpdate<ThreeD, FieldType::Magnetic>());
    for_each(grid.cells(1), DrudeUpdate<ThreeD, FieldType::Magnetic>(1.0, 0.3 / dt, 0.02 / dt, dt));
    for_each(grid.cells(2), LorentzUpdate<ThreeD, FieldType::Magnetic>(1.5, 0.5, 0.15 / dt, 0.01 / dt, dt));
    for_each(grid.cells(3), ConductiveUpdate<ThreeD, FieldType::Magnetic>(1.2, 0.04 * 1.2 * mu0 / dt, dt));
    for_each(grid.cells(4), ConstantUpdate<ThreeD, FieldType::Magnetic>(2.0));
    for_each(grid.cells(5), DrudeUpdate<ThreeD, FieldType::Magnetic>(2.5, 0.2 / dt, 0.05 / dt, dt));
  }
  std::vector<double> pr = grid.probes();  // runs the steps still recorded
  for (int n = 0; n < nsteps; n++) {
    for (int q = 0; q < 30; q++) printf("%s%.17g", q ? " " : "", pr[size_t(n) * 30 + q]);
    printf("\n");
  }
  return 0;
}
