// 3-D 20 x 18 x 17 run with magnetized Drude plasmas, written against the C++
// host surface (fdtd::hip) in the shape of a reference user loop
// (tests/test3D.cpp:146-171): CPML on every face (placement idiom, 4 cells), a
// plasma block in a fixed bias field (MagnetizedDrudeUpdate), a plasma block
// magnetized by the wave's own field (SelfMagnetizedDrudeUpdate), a dielectric
// block, and a strong Gaussian soft source on Dz.  Prints the probe traces
// (Ex, Ey, Ez, Hx, Hy, Hz at four cells), one step per line.  The reference's
// own run of the same scenario is the golden fixture grid3d_magnetized.
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
  std::vector<uint8_t> ids(size_t(nx) * ny * nz, 0);
  auto fill = [&](int x0, int x1, int y0, int y1, int z0, int z1, uint8_t m) {
    for (int k = z0; k < z1; k++)
      for (int j = y0; j < y1; j++)
        for (int i = x0; i < x1; i++) ids[size_t(i) + size_t(nx) * (size_t(j) + size_t(ny) * k)] = m;
  };
  fill(6, 10, 6, 12, 6, 10, 1);
  fill(11, 14, 6, 10, 8, 11, 2);
  fill(6, 9, 12, 14, 9, 12, 3);
  grid.setMaterialIds(ids);
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
  SelfMagnetizedDrudeUpdate<ThreeD, FieldType::Electric> self(dt);
  self.K() = 1.5;
  self.Freq() = 1.5e11;
  self.Gamma() = 2e9;
  MaterialMap mm;
  mm["vacuum"] = make_material_pair(VacuumUpdate<ThreeD, FieldType::Electric>(), VacuumUpdate<ThreeD, FieldType::Magnetic>());
  mm["biased plasma"] = make_material_pair(MagnetizedDrudeUpdate<ThreeD, FieldType::Electric>(0.3, -0.4, 1.2, 1.0, 2e11, 1e9, dt),
                                           VacuumUpdate<ThreeD, FieldType::Magnetic>());
  mm["self plasma"] = make_material_pair(self, VacuumUpdate<ThreeD, FieldType::Magnetic>());
  mm["dielectric"] =
      make_material_pair(ConstantUpdate<ThreeD, FieldType::Electric>(2.5), VacuumUpdate<ThreeD, FieldType::Magnetic>());
  Stepper<ThreeD, double> step(grid);
  step.then(YeeUpdate<ThreeD, Dx>(dt, dx))
      .then(YeeUpdate<ThreeD, Dy>(dt, dx))
      .then(YeeUpdate<ThreeD, Dz>(dt, dx))
      .source(src)
      .then(UpdatePML<ThreeD, FieldType::Electric, Dir::X>(dt, dx))
      .then(UpdatePML<ThreeD, FieldType::Electric, Dir::Y>(dt, dx))
      .then(UpdatePML<ThreeD, FieldType::Electric, Dir::Z>(dt, dx))
      .materials(mm, {{"vacuum", 0}, {"biased plasma", 1}, {"self plasma", 2}, {"dielectric", 3}})
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
