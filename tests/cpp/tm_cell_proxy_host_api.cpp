// 2-D TM 40 x 36 run written with the reference's per-cell idiom over the C++
// host surface's cell proxies (tests/testTM.cpp:43-113): the CPML placement
// loop calls cell.setPMLParameters<ft, Dir>(K, S, A, dt) on every cell of a
// layer, the initial Ez (a Gaussian) is written through cell.Ez() on the host
// mirror and copied to the device, a dielectric block (eps_r 2) sits off
// centre.  After 60 steps the mirror is refreshed and read by walking a row
// with getNeighborMax(0).  Prints: the Ez row (i = 1 .. 38 at j = 18), then
// Hx, Hy at (20, 18), then Dz and Ez at (26, 22) (inside the block).
#include <cmath>
#include <cstdio>
#include <vector>

#include "fdtd_hip.hpp"

using namespace fdtd::hip;

int main() {
  const int nx = 40, ny = 36, nsteps = 60, n = 8;
  const double dx = 1e-6, dt = 0.5 * dx / (c0 * std::sqrt(2.0));
  YeeGrid<TM, double> grid(nx, ny, dx, dt);
  PMLParameterModel p(dx);
  // the placement idiom, cell by cell: E parameters at i and L-i, H at i and L-1-i
  for (int i = 1; i < n + 2; i++) {
    const double x = double(n - i + 1) / n, xm = x - 0.5 / n;
    for (int j = 0; j < ny; j++) {
      grid.cell(i, j).setPMLParameters<FieldType::Electric, Dir::X>(p.K(x), p.S(x), p.A(x), dt);
      grid.cell(i, j).setPMLParameters<FieldType::Magnetic, Dir::X>(p.K(xm), p.S(xm), p.A(xm), dt);
      grid.cell(nx - i, j).setPMLParameters<FieldType::Electric, Dir::X>(p.K(x), p.S(x), p.A(x), dt);
      grid.cell(nx - 1 - i, j).setPMLParameters<FieldType::Magnetic, Dir::X>(p.K(xm), p.S(xm), p.A(xm), dt);
    }
    for (int ii = 0; ii < nx; ii++) {
      grid.cell(ii, i).setPMLParameters<FieldType::Electric, Dir::Y>(p.K(x), p.S(x), p.A(x), dt);
      grid.cell(ii, i).setPMLParameters<FieldType::Magnetic, Dir::Y>(p.K(xm), p.S(xm), p.A(xm), dt);
      grid.cell(ii, ny - i).setPMLParameters<FieldType::Electric, Dir::Y>(p.K(x), p.S(x), p.A(x), dt);
      grid.cell(ii, ny - 1 - i).setPMLParameters<FieldType::Magnetic, Dir::Y>(p.K(xm), p.S(xm), p.A(xm), dt);
    }
  }
  std::vector<uint8_t> ids(size_t(nx) * ny, 0);
  for (int j = 20; j < 26; j++)
    for (int i = 24; i < 30; i++) ids[size_t(i) + size_t(nx) * j] = 1;
  grid.setMaterialIds(ids);
  grid.sync_host();
  for (int j = 1; j < ny - 1; j++)
    for (int i = 1; i < nx - 1; i++) {
      const double r2 = (i - 20.0) * (i - 20.0) + (j - 18.0) * (j - 18.0);
      grid.cell(i, j).Ez() = std::exp(-r2 / 16.0);
    }
  grid.sync_device();
  Stepper<TM, double> step(grid);
  step.then(YeeUpdate<TM, Dz>(dt, dx))
      .then(UpdatePML<TM, FieldType::Electric, Dir::X>(dt, dx))
      .then(UpdatePML<TM, FieldType::Electric, Dir::Y>(dt, dx))
      .material(0, VacuumUpdate<TM, FieldType::Electric>())
      .material(1, ConstantUpdate<TM, FieldType::Electric>(2.0))
      .then(YeeUpdate<TM, Bx>(dt, dx))
      .then(YeeUpdate<TM, By>(dt, dx))
      .then(UpdatePML<TM, FieldType::Magnetic, Dir::X>(dt, dx))
      .then(UpdatePML<TM, FieldType::Magnetic, Dir::Y>(dt, dx));
  step.run(nsteps);
  step.synchronize();
  grid.sync_host();
  auto c = grid.cell(1, 18);
  for (int i = 1; i < nx - 1; i++, c = c.getNeighborMax(0)) printf("%s%.17g", i > 1 ? " " : "", c.Ez());
  printf("\n%.17g %.17g\n", grid.cell(20, 18).Hx(), grid.cell(20, 18).Hy());
  printf("%.17g %.17g\n", grid.cell(26, 22).Dz(), grid.cell(26, 22).Ez());
  return 0;
}
