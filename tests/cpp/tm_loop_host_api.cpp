// 2-D TM time loop written as the reference's own user loop
// (tests/testTM.cpp:102-113): one sweep per functor (for_each(grid, f), or
// the functor called on the grid or on a material's cells), soft sources
// applied inside the loop (SoftSource on Ez right after the E material step
// and on Hy right after the H material step, include/Source.hpp:421-425),
// two single TFSF planes with their own Aux1D stepped by updateD / updateB.
// The scenario is the one of the golden fixture tm2d_soft_planes (56 x 50,
// 10-cell CPML on x and y, a dielectric and a Lorentz block, 170 steps).
// Prints Ez, Hx, Hy at the eight probe cells, one step per line.
#include <cmath>
#include <cstdio>
#include <memory>
#include <vector>

#include "fdtd_hip.hpp"

using namespace fdtd::hip;

int main() {
  const int nx = 56, ny = 50, nsteps = 170;
  const double dx = 1e-6, dt = 0.7 * dx / (c0 * std::sqrt(2.0));
  YeeGrid<TM, double> grid(nx, ny, dx, dt);
  grid.loop().flush_every = 64;  // the loop's steps go to the device in runs of 64
  PMLParameterModel p(dx);
  grid.placePML<Dir::X>(p, 10);
  grid.placePML<Dir::Y>(p, 10);
  std::vector<uint8_t> ids(size_t(nx) * ny, 0);
  for (int j = 26; j < 34; j++)
    for (int i = 30; i < 38; i++) ids[size_t(i) + size_t(nx) * j] = 1;
  for (int j = 30; j < 36; j++)
    for (int i = 18; i < 24; i++) ids[size_t(i) + size_t(nx) * j] = 2;
  grid.setMaterialIds(ids);
  const int probes[8][2] = {{20, 32}, {34, 18}, {14, 24}, {13, 24}, {28, 38}, {28, 39}, {40, 20}, {45, 45}};
  for (auto& q : probes) {
    grid.add_probe<Ez>(q[0], q[1]);
    grid.add_probe<Hx>(q[0], q[1]);
    grid.add_probe<Hy>(q[0], q[1]);
  }
  auto gauss = [](double n, double t0, double w) { double a = (n - t0) / w; return std::exp(-a * a); };
  const double imp = std::sqrt(mu0 / eps0);
  auto srcE = make_soft_source<Ez>([&](double t) { return gauss(t, 30, 8); }, grid.cell(20, 32));
  auto srcH = make_soft_source<Hy>([&](double t) { return gauss(t, 45, 10) / imp; }, grid.cell(34, 18));
  const double f0 = c0 / (20 * dx), tw = 2.0 / f0;
  auto pulse = [f0, tw](double t0) {
    return [f0, tw, t0](double t) {
      double a = (t - t0) / tw;
      return std::exp(-a * a) * std::sin(2 * pi * f0 * (t - t0));
    };
  };
  const std::array<double, 3> k = {std::cos(25.0 * pi / 180.0), std::sin(25.0 * pi / 180.0), 0.0};
  const std::array<double, 3> origin = {27.5, 25.5, 0.0};
  auto p0 = make_tfsf_plane<TM, Dir::X, Orientation::MIN>(grid, {14, 15, 12, 38, 0, 1}, origin, k,
                                                          std::make_shared<Aux1D<double>>(100, dx, dt, pulse(3 * tw)));
  auto p1 = make_tfsf_plane<TM, Dir::Y, Orientation::MAX>(grid, {16, 41, 38, 39, 0, 1}, origin, k,
                                                          std::make_shared<Aux1D<double>>(100, dx, dt, pulse(4 * tw)));
  for (int t = 0; t < nsteps; t++) {
    for_each(grid, YeeUpdate<TM, Dz>(dt, dx));
    for_each(grid, UpdatePML<TM, FieldType::Electric, Dir::X>(dt, dx));
    for_each(grid, UpdatePML<TM, FieldType::Electric, Dir::Y>(dt, dx));
    p0.updateD(dt, dx);
    p1.updateD(dt, dx);
    for_each(grid, VacuumUpdate<TM, FieldType::Electric>());
    ConstantUpdate<TM, FieldType::Electric>(2.5)(grid.cells(1));  // a functor called on a range: the same sweep
    for_each(grid.cells(2), LorentzUpdate<TM, FieldType::Electric>(2.0, 0.7, 0.2 / dt, 0.01 / dt, dt));
    srcE.apply(t);
    for_each(grid, YeeUpdate<TM, Bx>(dt, dx));
    YeeUpdate<TM, By>(dt, dx)(grid);
    for_each(grid, UpdatePML<TM, FieldType::Magnetic, Dir::X>(dt, dx));
    UpdatePML<TM, FieldType::Magnetic, Dir::Y>(dt, dx)(grid);
    p0.updateB(dt, dx);
    p1.updateB(dt, dx);
    for_each(grid, VacuumUpdate<TM, FieldType::Magnetic>());
    srcH.apply(t);
  }
  std::vector<double> pr = grid.probes();  // runs the steps still recorded
  for (int n = 0; n < nsteps; n++) {
    for (int q = 0; q < 24; q++) printf("%s%.17g", q ? " " : "", pr[size_t(n) * 24 + q]);
    printf("\n");
  }
  return 0;
}
