// 3-D 14 x 12 x 16 complex-field run written against the C++ host surface
// (fdtd::hip): x and y Bloch-periodic (make_bloch_periodic_boundary with phase
// e^{-/+ i theta} at the min / max face), CPML along z (placement idiom, 4
// cells), a dielectric and a Drude block, a Gaussian soft source on Dz.
// Prints, one step per line, the real then the imaginary part of the probe
// traces (Ex, Ey, Ez, Hx, Hy, Hz at four cells).  The reference's own run of
// the same scenario on std::complex<double> cells is the golden fixture
// grid3d_bloch.
#include <cmath>
#include <complex>
#include <cstdio>
#include <vector>

#include "fdtd_hip.hpp"

using namespace fdtd::hip;

int main() {
  const int nx = 14, ny = 12, nz = 16, nsteps = 80;
  const double dx = 1e-8, dt = 0.5 * dx / (c0 * std::sqrt(3.0));
  const double thx = 0.9, thy = -0.4;
  YeeGrid<ThreeD, double> grid(nx, ny, nz, dx, dt);
  PMLParameterModel p(dx);
  grid.placePML<Dir::Z>(p, 4);
  std::vector<uint8_t> ids(size_t(nx) * ny * nz, 0);
  auto id = [&](int i, int j, int k) { return size_t(i) + size_t(nx) * (size_t(j) + size_t(ny) * k); };
  for (int k = 6; k < 10; k++)
    for (int j = 3; j < 8; j++)
      for (int i = 2; i < 7; i++) ids[id(i, j, k)] = 1;
  for (int k = 9; k < 12; k++)
    for (int j = 5; j < 10; j++)
      for (int i = 8; i < 12; i++) ids[id(i, j, k)] = 2;
  grid.setMaterialIds(ids);
  const int probes[4][3] = {{7, 6, 8}, {1, 3, 5}, {12, 10, 9}, {4, 1, 11}};
  for (auto& q : probes) {
    grid.add_probe<Ex>(q[0], q[1], q[2]);
    grid.add_probe<Ey>(q[0], q[1], q[2]);
    grid.add_probe<Ez>(q[0], q[1], q[2]);
    grid.add_probe<Hx>(q[0], q[1], q[2]);
    grid.add_probe<Hy>(q[0], q[1], q[2]);
    grid.add_probe<Hz>(q[0], q[1], q[2]);
  }
  auto src = make_soft_source<Dz>([](double t) { double u = (t - 16.0) / 5.0; return eps0 * std::exp(-u * u); },
                                  grid.cell(5, 4, 7));
  Stepper<ThreeD, double> step(grid);
  step.then(YeeUpdate<ThreeD, Dx>(dt, dx))
      .then(YeeUpdate<ThreeD, Dy>(dt, dx))
      .then(YeeUpdate<ThreeD, Dz>(dt, dx))
      .source(src)
      .then(UpdatePML<ThreeD, FieldType::Electric, Dir::Z>(dt, dx))
      .material(0, VacuumUpdate<ThreeD, FieldType::Electric>())
      .material(1, ConstantUpdate<ThreeD, FieldType::Electric>(3.0))
      .material(2, DrudeUpdate<ThreeD, FieldType::Electric>(1.0, 1.37e16, 1e14, dt))
      .then(YeeUpdate<ThreeD, Bx>(dt, dx))
      .then(YeeUpdate<ThreeD, By>(dt, dx))
      .then(YeeUpdate<ThreeD, Bz>(dt, dx))
      .then(UpdatePML<ThreeD, FieldType::Magnetic, Dir::Z>(dt, dx))
      .boundary(make_bloch_periodic_boundary<ThreeD, Dir::X, Orientation::MIN>(grid, std::polar(1.0, -thx)))
      .boundary(make_bloch_periodic_boundary<ThreeD, Dir::X, Orientation::MAX>(grid, std::polar(1.0, thx)))
      .boundary(make_bloch_periodic_boundary<ThreeD, Dir::Y, Orientation::MIN>(grid, std::polar(1.0, -thy)))
      .boundary(make_bloch_periodic_boundary<ThreeD, Dir::Y, Orientation::MAX>(grid, std::polar(1.0, thy)));
  step.horizon(nsteps).run(nsteps);
  step.synchronize();
  std::vector<double> re = grid.probes(), im = grid.probes_imag();
  for (int n = 0; n < nsteps; n++) {
    for (int q = 0; q < 24; q++) printf("%s%.17g", q ? " " : "", re[size_t(n) * 24 + q]);
    for (int q = 0; q < 24; q++) printf(" %.17g", im[size_t(n) * 24 + q]);
    printf("\n");
  }
  return 0;
}
