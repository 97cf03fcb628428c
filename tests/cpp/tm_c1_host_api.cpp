// 2-D TM 256^2 run written against the C++ host surface (fdtd::hip), in the
// shape of the reference's own user loop (tests/testTM.cpp:43-113): PML
// placement idiom, vacuum material, Gaussian soft source on Dz at the centre,
// 1000 steps, Ez probes.  Prints the probe traces, one step per line.
// usage: tm_c1_host_api i0 j0 i1 j1 ...
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "fdtd_hip.hpp"

using namespace fdtd::hip;

int main(int argc, char** argv) {
  const int N = 256, nsteps = 1000;
  const double dx = 1.0, dt = 0.5 * dx / c0;
  YeeGrid<TM, double> grid(N, N, dx, dt);
  PMLParameterModel p(dx);
  grid.placePML<Dir::X>(p, 10);
  grid.placePML<Dir::Y>(p, 10);
  for (int a = 1; a + 1 < argc; a += 2) grid.add_probe<Ez>(atoi(argv[a]), atoi(argv[a + 1]));
  auto src = make_soft_source<Dz>([](double t) { double u = (t - 40.0) / 12.0; return eps0 * std::exp(-u * u); },
                                  grid.cell(128, 128));
  Stepper<TM, double> step(grid);
  step.then(YeeUpdate<TM, Dz>(dt, dx))
      .source(src)
      .then(UpdatePML<TM, FieldType::Electric, Dir::X>(dt, dx))
      .then(UpdatePML<TM, FieldType::Electric, Dir::Y>(dt, dx))
      .material(0, VacuumUpdate<TM, FieldType::Electric>())
      .then(YeeUpdate<TM, Bx>(dt, dx))
      .then(YeeUpdate<TM, By>(dt, dx))
      .then(UpdatePML<TM, FieldType::Magnetic, Dir::X>(dt, dx))
      .then(UpdatePML<TM, FieldType::Magnetic, Dir::Y>(dt, dx));
  step.horizon(nsteps).run(400);
  step.run(nsteps - 400);
  step.synchronize();
  std::vector<double> pr = grid.probes();
  int np = (argc - 1) / 2;
  for (int n = 0; n < nsteps; n++) {
    for (int q = 0; q < np; q++) printf("%s%.17g", q ? " " : "", pr[size_t(n) * np + q]);
    printf("\n");
  }
  return 0;
}
