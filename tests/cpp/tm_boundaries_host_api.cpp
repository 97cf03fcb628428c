// 2-D TM 48 x 40 with boundary conditions, written against the C++ host
// surface (fdtd::hip) in the shape of a reference user loop: x periodic,
// y PEC (min) / Symmetric (max), a dielectric block (eps_r 4) and a Gaussian
// soft source on Dz.  Prints the probe traces (Ez, Ez, Hx, Hy), one step per
// line.  The same scenario is tests/synth.py:tm2d_boundaries.
#include <cmath>
#include <cstdio>
#include <vector>

#include "fdtd_hip.hpp"

using namespace fdtd::hip;

int main() {
  const int nx = 48, ny = 40, nsteps = 200;
  const double dx = 1e-6, dt = 0.5 * dx / (c0 * std::sqrt(2.0));
  YeeGrid<TM, double> grid(nx, ny, dx, dt);
  std::vector<uint8_t> ids(size_t(nx) * ny, 0);
  for (int j = 12; j < 20; j++)
    for (int i = 10; i < 18; i++) ids[size_t(i) + size_t(nx) * j] = 1;
  grid.setMaterialIds(ids);
  grid.add_probe<Ez>(1, ny / 2);
  grid.add_probe<Ez>(nx - 2, ny / 3);
  grid.add_probe<Hx>(nx / 2, 1);
  grid.add_probe<Hy>(nx / 2, ny - 2);
  auto src = make_soft_source<Dz>([](double t) { double u = (t - 25.0) / 8.0; return eps0 * std::exp(-u * u); },
                                  grid.cell(nx / 3, ny / 2));
  BoundaryOptions bo;  // PEC on every face by default, as in the reference
  bo(Dir::X, Orientation::MIN) = Boundary::Periodic;
  bo.get<Dir::X, Orientation::MAX>() = Boundary::Periodic;
  MaterialMap mm;  // the reference's MaterialMap idiom (include/MaterialMap.hpp:30-37)
  mm["vacuum"] = make_material_pair(VacuumUpdate<TM, FieldType::Electric>(), VacuumUpdate<TM, FieldType::Magnetic>());
  mm["dielectric"] =
      make_material_pair(ConstantUpdate<TM, FieldType::Electric>(4.0), ConstantUpdate<TM, FieldType::Magnetic>(1.0));
  Stepper<TM, double> step(grid);
  step.then(YeeUpdate<TM, Dz>(dt, dx))
      .source(src)
      .materials(mm, {{"vacuum", 0}, {"dielectric", 1}})
      .then(YeeUpdate<TM, Bx>(dt, dx))
      .then(YeeUpdate<TM, By>(dt, dx))
      .boundaries(bo)
      .boundary(make_symmetric_boundary<TM, Dir::Y, Orientation::MAX>(grid));
  step.horizon(nsteps).run(77);
  step.run(nsteps - 77);
  step.synchronize();
  std::vector<double> pr = grid.probes();
  for (int n = 0; n < nsteps; n++) {
    for (int q = 0; q < 4; q++) printf("%s%.17g", q ? " " : "", pr[size_t(n) * 4 + q]);
    printf("\n");
  }
  return 0;
}
