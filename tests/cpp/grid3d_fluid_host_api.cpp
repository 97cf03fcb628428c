// 3-D 20 x 18 x 17 run with the fluid Drude plasmas, written against the C++
// host surface (fdtd::hip) in the shape of a reference user loop: CPML on
// every face (placement idiom, 4 cells), blocks of FluidDrudeUpdate,
// MagnetizedFluidDrudeUpdate and SelfMagnetizedFluidDrudeUpdate on a graded
// electron density and collision rate (the cells' ne() / nu_m()), a strong
// Gaussian soft source on Dz.  Prints the probe traces (Ex, Ey, Ez, Hx, Hy, Hz
// at four cells), one step per line.  The reference's own run of the same
// scenario is the golden fixture grid3d_fluid.
#include <cmath>
#include <cstdio>
#include <vector>

#include "fdtd_hip.hpp"

using namespace fdtd::hip;

int main() {
  const int nx = 20, ny = 18, nz = 17, nsteps = 70;
  const double dx = 1e-3, dt = 0.5 * dx / (c0 * std::sqrt(3.0));
  YeeGrid<ThreeD, double> grid(nx, ny, nz, dx, dt);
  PMLParameterModel p(dx);
  grid.placePML<Dir::X>(p, 4);
  grid.placePML<Dir::Y>(p, 4);
  grid.placePML<Dir::Z>(p, 4);
  const size_t ncell = size_t(nx) * ny * nz;
  std::vector<uint8_t> ids(ncell, 0);
  std::vector<double> ne(ncell), nu(ncell);
  for (int k = 0; k < nz; k++)
    for (int j = 0; j < ny; j++)
      for (int i = 0; i < nx; i++) {
        const size_t q = size_t(i) + size_t(nx) * (size_t(j) + size_t(ny) * k);
        ne[q] = 1.26e19 * (0.4 + 0.06 * i + 0.01 * j * (k % 4));
        nu[q] = 1e9 * (1.0 + 0.25 * ((i + 2 * j + 3 * k) % 5));
      }
  auto fill = [&](int x0, int x1, int y0, int y1, int z0, int z1, uint8_t m) {
    for (int k = z0; k < z1; k++)
      for (int j = y0; j < y1; j++)
        for (int i = x0; i < x1; i++) ids[size_t(i) + size_t(nx) * (size_t(j) + size_t(ny) * k)] = m;
  };
  fill(6, 10, 6, 12, 6, 10, 1);
  fill(11, 14, 6, 10, 8, 11, 2);
  fill(6, 9, 12, 14, 9, 12, 3);
  grid.setMaterialIds(ids);
  grid.setFluidState(ne, nu);
  const int probes[4][3] = {{10, 9, 8}, {8, 8, 8}, {12, 8, 9}, {7, 13, 10}};
  for (auto& q : probes) {
    grid.add_probe<Ex>(q[0], q[1], q[2]);
    grid.add_probe<Ey>(q[0], q[1], q[2]);
    grid.add_probe<Ez>(q[0], q[1], q[2]);
    grid.add_probe<Hx>(q[0], q[1], q[2]);
    grid.add_probe<Hy>(q[0], q[1], q[2]);
    grid.add_probe<Hz>(q[0], q[1], q[2]);
  }
  auto src = make_soft_source<Dz>([](double t) { double u = (t - 14.0) / 5.0; return eps0 * 3e9 * std::exp(-u * u); },
                                  grid.cell(10, 9, 8));
  MaterialMap mm;
  mm["vacuum"] = make_material_pair(VacuumUpdate<ThreeD, FieldType::Electric>(), VacuumUpdate<ThreeD, FieldType::Magnetic>());
  mm["fluid"] = make_material_pair(FluidDrudeUpdate<ThreeD, FieldType::Electric>(dt), VacuumUpdate<ThreeD, FieldType::Magnetic>());
  mm["biased fluid"] = make_material_pair(MagnetizedFluidDrudeUpdate<ThreeD, FieldType::Electric>(0.3, -0.4, 1.2, dt),
                                          VacuumUpdate<ThreeD, FieldType::Magnetic>());
  mm["self fluid"] = make_material_pair(SelfMagnetizedFluidDrudeUpdate<ThreeD, FieldType::Electric>(dt),
                                        VacuumUpdate<ThreeD, FieldType::Magnetic>());
  Stepper<ThreeD, double> step(grid);
  step.then(YeeUpdate<ThreeD, Dx>(dt, dx))
      .then(YeeUpdate<ThreeD, Dy>(dt, dx))
      .then(YeeUpdate<ThreeD, Dz>(dt, dx))
      .source(src)
      .then(UpdatePML<ThreeD, FieldType::Electric, Dir::X>(dt, dx))
      .then(UpdatePML<ThreeD, FieldType::Electric, Dir::Y>(dt, dx))
      .then(UpdatePML<ThreeD, FieldType::Electric, Dir::Z>(dt, dx))
      .materials(mm, {{"vacuum", 0}, {"fluid", 1}, {"biased fluid", 2}, {"self fluid", 3}})
      .then(YeeUpdate<ThreeD, Bx>(dt, dx))
      .then(YeeUpdate<ThreeD, By>(dt, dx))
      .then(YeeUpdate<ThreeD, Bz>(dt, dx))
      .then(UpdatePML<ThreeD, FieldType::Magnetic, Dir::X>(dt, dx))
      .then(UpdatePML<ThreeD, FieldType::Magnetic, Dir::Y>(dt, dx))
      .then(UpdatePML<ThreeD, FieldType::Magnetic, Dir::Z>(dt, dx));
  step.horizon(nsteps).run(nsteps);
  step.synchronize();
  std::vector<double> pr = grid.probes();
  for (int n = 0; n < nsteps; n++) {
    for (int q = 0; q < 24; q++) printf("%s%.17g", q ? " " : "", pr[size_t(n) * 24 + q]);
    printf("\n");
  }
  return 0;
}
