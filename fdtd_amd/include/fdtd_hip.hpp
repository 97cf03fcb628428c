// fdtd_hip.hpp — C++ host surface over the C ABI (include/fdtd_hip.h).
//
// It keeps the names and constructor signatures of the reference's functor
// API (pederson/fdtd) so that a user time loop maps onto it directly:
//
//   reference (one CPU sweep per functor)          here (one fused device step)
//   ---------------------------------------------  ------------------------------------------
//   PMLParameterModel p(dx)                         PMLParameterModel p(dx)          PML.hpp:368-393
//   cell.setPMLParameters<ft,Dir>(K,S,A,dt)         grid.setPMLParameters<ft,Dir>(i,K,S,A) PML.hpp:84-102
//   YeeUpdate<Mode,Dz>(dt,dx)                       YeeUpdate<Mode,Dz>(dt,dx)        YeeUpdates.hpp:343-477
//   UpdatePML<Mode,ft,Dir>(dt,dx)                   UpdatePML<Mode,ft,Dir>(dt,dx)    PML.hpp:296-355
//   ConstantUpdate<Mode,ft>(K) / VacuumUpdate       same names                       DispersiveMaterials.hpp:489-497
//   DrudeUpdate<Mode,ft>(K,freq,gamma,dt)           same                             DispersiveMaterials.hpp:1208-1211
//   LorentzUpdate<Mode,ft>(K,delta,freq,gamma,dt)   same                             DispersiveMaterials.hpp:881-884
//   ConductiveUpdate<Mode,ft>(K,freq,dt)            same                             DispersiveMaterials.hpp:674-677
//   MagnetizedDrudeUpdate<Mode,ft>(Bx,By,Bz,K,Freq,Gamma,dt)  same                   DispersiveMaterials.hpp:2008-2011
//   SelfMagnetizedDrudeUpdate<Mode,ft>(dt) + K()/Freq()/Gamma()  same                DispersiveMaterials.hpp:2024-2027
//   FluidDrudeUpdate<Mode,ft>(dt), cell.ne()/nu_m() same + grid.setFluidState(ne, nu) DispersiveMaterials.hpp:1214-1217
//   MagnetizedFluidDrudeUpdate<Mode,ft>(Bx,By,Bz,dt), SelfMagnetizedFluidDrudeUpdate<Mode,ft>(dt)  2014-2021
//   make_soft_source<Dz>(signal, cell)              make_soft_source<Dz>(signal, grid.cell(i,j,k))  Source.hpp:258-261
//   Aux1D<double>(ncells,dx,dt,signal)              Aux1D<double>(ncells,dx,dt,signal)               Source.hpp:280-408
//   make_tfsf_plane<Mode,Dir,Orient>(b,e,loc,k,aux) make_tfsf_box<Mode>(grid,box,k,aux)              Source.hpp:624-634
//   make_pec_boundary<Mode,Dir,Orient>(b,e)         make_pec_boundary<Mode,Dir,Orient>(grid)         BoundaryUpdates.hpp:527-565
//     (pmc, symmetric, antisymmetric, periodic)     same names; periodic image fixed (layer L-2 / 1)  BoundaryUpdates.hpp:578-832
//   BoundaryData::updateE() / updateH()             Stepper::boundary(bd): ghost writes after E / H   BoundaryUpdates.hpp:340-362
//   make_bloch_periodic_boundary<Mode,Dir,Orient>(b,e,f,factor)  make_bloch_periodic_boundary<Mode,Dir,Orient>(grid, factor)
//     (complex cells)                               complex fields; grid.download(field | FDTD_IMAG, ...)  BoundaryUpdates.hpp:845-885
//
// Instead of sweeping each functor over a user container, the functors are
// handed once to a Stepper, which checks that they form the reference's step
// composition (D curl, sources, CPML E, TFSF, E material, B curl, CPML H,
// TFSF, H material) and then runs whole steps on the GPU.  Errors from the C
// ABI surface as fdtd::hip::Error exceptions on the host side only.
#pragma once

#include <array>
#include <cmath>
#include <complex>
#include <cstdint>
#include <functional>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/fdtd_hip.h"

namespace fdtd {
namespace hip {

// --------------------------------------------------------------- constants
// include/FDTDConstants.hpp:394-398
constexpr double pi = 3.14159265358979323846264338327950288;
constexpr double eps0 = 8.854e-12;
constexpr double mu0 = pi * 4.0e-7;
constexpr double c0 = 2.99792458e+8;
inline double imp0() {
  double x = mu0 / eps0, curr = x, prev = 0;  // the reference's constexpr Newton iteration
  while (curr != prev) {
    double n = 0.5 * (curr + x / curr);
    prev = curr;
    curr = n;
  }
  return curr;
}

enum class Dir : char { X = 0, Y, Z };
enum class FieldType : char { Electric = 0, Magnetic };
enum class Orientation : char { MIN = 0, MAX };

struct TM {
  static constexpr int dim = 2, code = FDTD_TM;
};
struct ThreeD {
  static constexpr int dim = 3, code = FDTD_3D;
};

// field tags: field code of the C ABI, component, electric/magnetic
template <int F, int C>
struct FieldTag {
  static constexpr int field = F, comp = C;
};
using Dx = FieldTag<FDTD_D, 0>;
using Dy = FieldTag<FDTD_D, 1>;
using Dz = FieldTag<FDTD_D, 2>;
using Ex = FieldTag<FDTD_E, 0>;
using Ey = FieldTag<FDTD_E, 1>;
using Ez = FieldTag<FDTD_E, 2>;
using Bx = FieldTag<FDTD_B, 0>;
using By = FieldTag<FDTD_B, 1>;
using Bz = FieldTag<FDTD_B, 2>;
using Hx = FieldTag<FDTD_H, 0>;
using Hy = FieldTag<FDTD_H, 1>;
using Hz = FieldTag<FDTD_H, 2>;

struct Error : std::runtime_error {
  explicit Error(const std::string& s) : std::runtime_error(s) {}
};
inline void check(int rc, const char* what) {
  if (rc != FDTD_OK) throw Error(std::string(what) + ": " + fdtd_last_error());
}

// ------------------------------------------------------------- PML model
// PMLParameterModel, include/PML.hpp:368-393
struct PMLParameterModel {
  double mM, mMa, msMax, mkMax, maMax;
  explicit PMLParameterModel(double dx) : mM(3), mMa(1), msMax(0.8 * (mM + 1) / (imp0() * dx)), mkMax(5), maMax(0.05) {}
  PMLParameterModel(double m, double ma, double sMax, double kMax, double aMax)
      : mM(m), mMa(ma), msMax(sMax), mkMax(kMax), maMax(aMax) {}
  double K(double x) const { return x < 0.0 ? 1.0 : (x > 1.0 ? 1.0 : 1.0 + std::pow(x, mM) * (mkMax - 1.0)); }
  double S(double x) const { return x < 0.0 ? 0.0 : (x > 1.0 ? 0.0 : std::pow(x, mM) * msMax); }
  double A(double x) const { return x < 0.0 ? 0.0 : (x > 1.0 ? 0.0 : std::pow(1.0 - x, mMa) * maMax); }
};

// ------------------------------------------------------------------- grid
template <class Mode, class scalar = double>
class YeeGrid;

// One cell of a grid, with the reference's cell accessors (YeeFields.hpp:46-78,
// DispersiveMaterials.hpp:62-77, DefaultInterfaces.hpp:65-79, PML.hpp:84-102)
// over the grid's host mirror: E and H are writable (copied to the device by
// YeeGrid::sync_device), D, B, P, J (pole 0) are read-only views of the state
// YeeGrid::sync_host copied last.  Neighbours are index arithmetic, not links.
template <class Mode, class scalar>
struct CellRef {
  YeeGrid<Mode, scalar>* grid;
  int i, j, k;
  double& Ex() const;
  double& Ey() const;
  double& Ez() const;
  double& Hx() const;
  double& Hy() const;
  double& Hz() const;
  double Dx() const;
  double Dy() const;
  double Dz() const;
  double Bx() const;
  double By() const;
  double Bz() const;
  double Px() const;
  double Py() const;
  double Pz() const;
  double Jx() const;
  double Jy() const;
  double Jz() const;
  CellRef getNeighborMin(int d) const;
  CellRef getNeighborMax(int d) const;
  // NewPML::setPMLParameters of this cell: the grid keeps per-axis profiles,
  // so this sets them for every cell sharing the cell's coordinate along d
  template <FieldType ft, Dir d>
  void setPMLParameters(double K, double S, double A, double dt) const;
};

template <class Mode, class scalar>
class YeeGrid {
 public:
  YeeGrid(int nx, int ny, int nz, double dx, double dt) : nx_(nx), ny_(ny), nz_(Mode::dim == 3 ? nz : 1), dx_(dx), dt_(dt) {
    fdtd_grid_desc d{Mode::code, std::is_same<scalar, double>::value ? FDTD_F64 : FDTD_F32, nx_, ny_, nz_, dx, dt};
    fdtd_grid* g = nullptr;
    check(fdtd_grid_create(&d, &g), "fdtd_grid_create");
    g_.reset(g);
    for (int a = 0; a < 3; a++)
      for (int ft = 0; ft < 2; ft++) {
        int L = a == 0 ? nx_ : (a == 1 ? ny_ : nz_);
        ksa_[ft][a].assign(3 * L, 0.0);
        for (int i = 0; i < L; i++) ksa_[ft][a][3 * i] = 1.0;
      }
  }
  YeeGrid(int nx, int ny, double dx, double dt) : YeeGrid(nx, ny, 1, dx, dt) {}

  fdtd_grid* handle() const { return g_.get(); }
  int nx() const { return nx_; }
  int ny() const { return ny_; }
  int nz() const { return nz_; }
  double dx() const { return dx_; }
  double dt() const { return dt_; }

  CellRef<Mode, scalar> cell(int i, int j, int k = 0) { return {this, i, j, k}; }

  // NewPML::setPMLParameters for every cell whose coordinate along `d` is idx
  template <FieldType ft, Dir d>
  void setPMLParameters(int idx, double K, double S, double A) {
    auto& v = ksa_[int(ft)][int(d)];
    v[3 * idx] = K;
    v[3 * idx + 1] = S;
    v[3 * idx + 2] = A;
    pml_dirty_[int(d)] = true;
  }

  // the reference's placement idiom on one axis (tests/testTM.cpp:46-71)
  template <Dir d>
  void placePML(const PMLParameterModel& p, int n) {
    check(fdtd_set_pml_idiom(handle(), int(d), n, p.mM, p.mMa, p.msMax, p.mkMax, p.maMax), "fdtd_set_pml_idiom");
  }

  // record one E or H component at a cell after every step
  template <class F>
  void add_probe(int i, int j, int k = 0) {
    static_assert(F::field == FDTD_E || F::field == FDTD_H, "probes record E or H");
    check(fdtd_add_probe(handle(), F::field, F::comp, i, j, k), "fdtd_add_probe");
    nprobes_++;
  }
  // [steps][probes]
  std::vector<double> probes() const {
    long n = fdtd_step_count(handle());
    std::vector<double> v(size_t(std::max(1L, n)) * std::max(1, nprobes_));
    int got = 0;
    check(fdtd_get_probes(handle(), v.data(), int(n), &got), "fdtd_get_probes");
    v.resize(size_t(got) * nprobes_);
    return v;
  }

  // the imaginary part of the probes of a complex grid (a Bloch-periodic face)
  std::vector<double> probes_imag() const {
    long n = fdtd_step_count(handle());
    std::vector<double> v(size_t(std::max(1L, n)) * std::max(1, nprobes_));
    int got = 0;
    check(fdtd_get_probes_imag(handle(), v.data(), int(n), &got), "fdtd_get_probes_imag");
    v.resize(size_t(got) * nprobes_);
    return v;
  }

  void setMaterialIds(const std::vector<uint8_t>& ids) { check(fdtd_set_matid(handle(), ids.data()), "fdtd_set_matid"); }

  // Host mirror (explicit sync): sync_host copies E, H, D, B, P, J (pole 0)
  // of every existing component to the host; sync_device copies the mirror's
  // E and H back.  Cell accessors read and write the mirror only.
  void sync_host() {
    const int fields[6] = {FDTD_E, FDTD_H, FDTD_D, FDTD_B, FDTD_P, FDTD_J};
    for (int f : fields)
      for (int c = 0; c < 3; c++) {
        auto& m = mirror_[f][c];
        m.assign(size_t(nx_) * ny_ * nz_, 0.0);
        if (fdtd_download_field(handle(), f, c, 0, m.data()) != FDTD_OK) m.clear();  // component absent in this mode
      }
  }
  void sync_device() {
    for (int f : {FDTD_E, FDTD_H})
      for (int c = 0; c < 3; c++)
        if (!mirror_[f][c].empty()) upload(f, c, mirror_[f][c]);
  }
  double& mirror(int field, int comp, int i, int j, int k) {
    auto& m = mirror_[field][comp];
    if (m.empty()) throw Error("cell access: component absent in this mode, or sync_host() not called");
    return m[size_t(i) + size_t(nx_) * (size_t(j) + size_t(ny_) * k)];
  }
  // the cells' ne() and nu_m() read by the fluid Drude materials (nx*ny*nz each)
  void setFluidState(const std::vector<double>& ne, const std::vector<double>& nu_m) {
    if (ne.size() != size_t(nx_) * ny_ * nz_ || nu_m.size() != ne.size()) throw Error("setFluidState: size mismatch");
    check(fdtd_set_fluid_state(handle(), ne.data(), nu_m.data()), "fdtd_set_fluid_state");
  }

  std::vector<double> download(int field, int comp, int pole = 0) const {
    std::vector<double> v(size_t(nx_) * ny_ * nz_);
    check(fdtd_download_field(handle(), field, comp, pole, v.data()), "fdtd_download_field");
    return v;
  }
  void upload(int field, int comp, const std::vector<double>& v) {
    check(fdtd_upload_field(handle(), field, comp, 0, v.data()), "fdtd_upload_field");
  }

  void commit_pml() {
    for (int a = 0; a < Mode::dim; a++) {
      if (!pml_dirty_[a]) continue;
      for (int ft = 0; ft < 2; ft++) {
        const auto& v = ksa_[ft][a];
        size_t L = v.size() / 3;
        std::vector<double> K(L), S(L), A(L);
        for (size_t i = 0; i < L; i++) {
          K[i] = v[3 * i];
          S[i] = v[3 * i + 1];
          A[i] = v[3 * i + 2];
        }
        check(fdtd_set_pml_axis(handle(), a, ft, K.data(), S.data(), A.data()), "fdtd_set_pml_axis");
      }
      pml_dirty_[a] = false;
    }
  }

 private:
  struct Del {
    void operator()(fdtd_grid* g) const { fdtd_grid_destroy(g); }
  };
  std::unique_ptr<fdtd_grid, Del> g_;
  int nx_, ny_, nz_;
  double dx_, dt_;
  std::vector<double> ksa_[2][3];
  bool pml_dirty_[3] = {false, false, false};
  int nprobes_ = 0;
  std::vector<double> mirror_[6][3];  // [field code][component]
};

template <class Mode, class scalar>
double& CellRef<Mode, scalar>::Ex() const { return grid->mirror(FDTD_E, 0, i, j, k); }
template <class Mode, class scalar>
double& CellRef<Mode, scalar>::Ey() const { return grid->mirror(FDTD_E, 1, i, j, k); }
template <class Mode, class scalar>
double& CellRef<Mode, scalar>::Ez() const { return grid->mirror(FDTD_E, 2, i, j, k); }
template <class Mode, class scalar>
double& CellRef<Mode, scalar>::Hx() const { return grid->mirror(FDTD_H, 0, i, j, k); }
template <class Mode, class scalar>
double& CellRef<Mode, scalar>::Hy() const { return grid->mirror(FDTD_H, 1, i, j, k); }
template <class Mode, class scalar>
double& CellRef<Mode, scalar>::Hz() const { return grid->mirror(FDTD_H, 2, i, j, k); }
template <class Mode, class scalar>
double CellRef<Mode, scalar>::Dx() const { return grid->mirror(FDTD_D, 0, i, j, k); }
template <class Mode, class scalar>
double CellRef<Mode, scalar>::Dy() const { return grid->mirror(FDTD_D, 1, i, j, k); }
template <class Mode, class scalar>
double CellRef<Mode, scalar>::Dz() const { return grid->mirror(FDTD_D, 2, i, j, k); }
template <class Mode, class scalar>
double CellRef<Mode, scalar>::Bx() const { return grid->mirror(FDTD_B, 0, i, j, k); }
template <class Mode, class scalar>
double CellRef<Mode, scalar>::By() const { return grid->mirror(FDTD_B, 1, i, j, k); }
template <class Mode, class scalar>
double CellRef<Mode, scalar>::Bz() const { return grid->mirror(FDTD_B, 2, i, j, k); }
template <class Mode, class scalar>
double CellRef<Mode, scalar>::Px() const { return grid->mirror(FDTD_P, 0, i, j, k); }
template <class Mode, class scalar>
double CellRef<Mode, scalar>::Py() const { return grid->mirror(FDTD_P, 1, i, j, k); }
template <class Mode, class scalar>
double CellRef<Mode, scalar>::Pz() const { return grid->mirror(FDTD_P, 2, i, j, k); }
template <class Mode, class scalar>
double CellRef<Mode, scalar>::Jx() const { return grid->mirror(FDTD_J, 0, i, j, k); }
template <class Mode, class scalar>
double CellRef<Mode, scalar>::Jy() const { return grid->mirror(FDTD_J, 1, i, j, k); }
template <class Mode, class scalar>
double CellRef<Mode, scalar>::Jz() const { return grid->mirror(FDTD_J, 2, i, j, k); }
template <class Mode, class scalar>
CellRef<Mode, scalar> CellRef<Mode, scalar>::getNeighborMin(int d) const {
  return {grid, i - (d == 0), j - (d == 1), k - (d == 2)};
}
template <class Mode, class scalar>
CellRef<Mode, scalar> CellRef<Mode, scalar>::getNeighborMax(int d) const {
  return {grid, i + (d == 0), j + (d == 1), k + (d == 2)};
}
template <class Mode, class scalar>
template <FieldType ft, Dir d>
void CellRef<Mode, scalar>::setPMLParameters(double K, double S, double A, double) const {
  grid->template setPMLParameters<ft, d>(d == Dir::X ? i : (d == Dir::Y ? j : k), K, S, A);
}

// --------------------------------------------------------------- functors
// Value types naming one stage of the step; the Stepper checks the order.
template <class Mode, class Flux>
struct YeeUpdate {
  double dt, dx;
  YeeUpdate(double deltat, double deltax) : dt(deltat), dx(deltax) {}
  static constexpr int stage = Flux::field == FDTD_D || Flux::field == FDTD_E ? 0 : 4;
};
template <class Mode, FieldType ft, Dir d>
struct UpdatePML {
  double dt, dx;
  UpdatePML(double deltat, double deltax) : dt(deltat), dx(deltax) {}
  static constexpr int stage = ft == FieldType::Electric ? 1 : 5;
};

// material updates: one table entry each (eps_r or mu_r plus poles)
struct MaterialSpec {
  double eps = 1.0, mu = 1.0;
  std::vector<double> poles;  // kind, freq, gamma, weight per pole
  int mag = 0;                // magnetized Drude: 1 fixed bias, 2 the cell's own B
  double bias[3] = {0, 0, 0};
  bool fluid = false;         // fluid Drude: omega_p, gamma per cell (YeeGrid::setFluidState)
};
template <class Mode, FieldType ft, bool forward = false>
struct VacuumUpdate {
  MaterialSpec spec() const { return {}; }
};
template <class Mode, FieldType ft, bool forward = false>
struct ConstantUpdate {
  double K;
  explicit ConstantUpdate(double k) : K(k) {}
  MaterialSpec spec() const {
    MaterialSpec s;
    (ft == FieldType::Electric ? s.eps : s.mu) = K;
    return s;
  }
};
// Drude pole with the reference's DrudeStorage(K, Freq, Gamma, dt) arguments
template <class Mode, FieldType ft, bool forward = false>
struct DrudeUpdate {
  double K, freq, gamma, dt;
  DrudeUpdate(double k, double f, double g, double deltat) : K(k), freq(f), gamma(g), dt(deltat) {}
  MaterialSpec spec() const { return {K, 1.0, {double(FDTD_DRUDE), freq, gamma, 1.0}}; }
};
// Lorentz pole with LorentzStorage(K, Delta, Freq, Gamma, dt); like the
// reference's update, Delta does not enter (include/DispersiveMaterials.hpp:711-735)
template <class Mode, FieldType ft, bool forward = false>
struct LorentzUpdate {
  double K, delta, freq, gamma, dt;
  LorentzUpdate(double k, double de, double f, double g, double deltat) : K(k), delta(de), freq(f), gamma(g), dt(deltat) {}
  MaterialSpec spec() const { return {K, 1.0, {double(FDTD_LORENTZ), freq, gamma, 1.0}}; }
};
// Conductive medium with ConductiveStorage(K, Freq, dt), Freq being the
// conductivity (include/DispersiveMaterials.hpp:508-680): P relaxes
// implicitly towards D with factor dt Freq/(K eps0)
template <class Mode, FieldType ft, bool forward = false>
struct ConductiveUpdate {
  double K, freq, dt;
  ConductiveUpdate(double k, double f, double deltat) : K(k), freq(f), dt(deltat) {}
  MaterialSpec spec() const { return {K, 1.0, {double(FDTD_CONDUCTIVE), freq, 0.0, 1.0}}; }
};
// Magnetized Drude plasma with MagnetizedDrudeStorage(Bx, By, Bz, K, Freq,
// Gamma, dt): in 3-D the three components of P and J advance together in the
// fixed bias field B [T] (include/DispersiveMaterials.hpp:1263-1340); in TM
// the reference's update is the Drude one per component, and so is this one
template <class Mode, FieldType ft, bool forward = false>
struct MagnetizedDrudeUpdate {
  double Bx, By, Bz, K, freq, gamma, dt;
  MagnetizedDrudeUpdate(double bx, double by, double bz, double k, double f, double g, double deltat)
      : Bx(bx), By(by), Bz(bz), K(k), freq(f), gamma(g), dt(deltat) {}
  MaterialSpec spec() const {
    MaterialSpec s{K, 1.0, {double(FDTD_DRUDE), freq, gamma, 1.0}};
    s.mag = 1;
    s.bias[0] = Bx;
    s.bias[1] = By;
    s.bias[2] = Bz;
    return s;
  }
};
// The same plasma magnetized by the wave's own field B = mu_r mu0 H at the
// cell; SelfMagnetizedDrudeStorage(dt) with its K, Freq, Gamma set through
// the accessors, as in the reference (1898-1943)
template <class Mode, FieldType ft, bool forward = false>
struct SelfMagnetizedDrudeUpdate {
  double k_ = 1.0, freq_ = 0.0, gamma_ = 0.0, dt;
  explicit SelfMagnetizedDrudeUpdate(double deltat) : dt(deltat) {}
  double& K() { return k_; }
  double& Freq() { return freq_; }
  double& Gamma() { return gamma_; }
  MaterialSpec spec() const {
    MaterialSpec s{k_, 1.0, {double(FDTD_DRUDE), freq_, gamma_, 1.0}};
    s.mag = 2;
    return s;
  }
};
// The fluid forms: omega_p = sqrt(q_e^2/(m_e eps0) ne) and gamma = nu_m of
// each cell (FluidDrudeStorage, MagnetizedFluidDrudeStorage,
// SelfMagnetizedFluidDrudeStorage; include/DispersiveMaterials.hpp:1160-1199,
// 1728-1893), K = 1; the cells' ne and nu_m are given by YeeGrid::setFluidState
template <class Mode, FieldType ft, bool forward = false>
struct FluidDrudeUpdate {
  double dt;
  explicit FluidDrudeUpdate(double deltat) : dt(deltat) {}
  MaterialSpec spec() const {
    MaterialSpec s{1.0, 1.0, {double(FDTD_DRUDE), 0.0, 0.0, 1.0}};
    s.fluid = true;
    return s;
  }
};
template <class Mode, FieldType ft, bool forward = false>
struct MagnetizedFluidDrudeUpdate {
  double Bx, By, Bz, dt;
  MagnetizedFluidDrudeUpdate(double bx, double by, double bz, double deltat) : Bx(bx), By(by), Bz(bz), dt(deltat) {}
  MaterialSpec spec() const {
    MaterialSpec s{1.0, 1.0, {double(FDTD_DRUDE), 0.0, 0.0, 1.0}};
    s.fluid = true;
    s.mag = 1;
    s.bias[0] = Bx;
    s.bias[1] = By;
    s.bias[2] = Bz;
    return s;
  }
};
template <class Mode, FieldType ft, bool forward = false>
struct SelfMagnetizedFluidDrudeUpdate {
  double dt;
  explicit SelfMagnetizedFluidDrudeUpdate(double deltat) : dt(deltat) {}
  MaterialSpec spec() const {
    MaterialSpec s{1.0, 1.0, {double(FDTD_DRUDE), 0.0, 0.0, 1.0}};
    s.fluid = true;
    s.mag = 2;
    return s;
  }
};

// MaterialMap equivalent (include/MaterialMap.hpp:204-284): id -> material.
template <class Mode, class scalar>
void set_material(YeeGrid<Mode, scalar>& g, int id, const MaterialSpec& s) {
  if (s.fluid) {
    check(fdtd_set_fluid_drude(g.handle(), id, s.mu, s.mag, s.mag == 1 ? s.bias : nullptr), "fdtd_set_fluid_drude");
    return;
  }
  if (s.mag) {
    check(fdtd_set_magnetized_drude(g.handle(), id, s.eps, s.mu, s.poles[1], s.poles[2], s.mag == 1 ? s.bias : nullptr),
          "fdtd_set_magnetized_drude");
    return;
  }
  check(fdtd_set_material(g.handle(), id, s.eps, s.mu, int(s.poles.size() / 4), s.poles.empty() ? nullptr : s.poles.data()),
        "fdtd_set_material");
}

// MaterialPair / MaterialMap (include/MaterialMap.hpp:45-284): a named pair
// of electric and magnetic models.  The device takes the electric model's
// eps_r and poles and the magnetic model's mu_r (magnetic dispersion has no
// device counterpart).  The reference's XML form is read and written by
// fdtd_amd/materials.py.
struct MaterialPair {
  MaterialSpec electric, magnetic;
};
template <class E, class M>
MaterialPair make_material_pair(const E& e, const M& m) {
  return {e.spec(), m.spec()};
}
using MaterialMap = std::map<std::string, MaterialPair>;
inline MaterialSpec combine(const std::string& name, const MaterialPair& p) {
  if (!p.magnetic.poles.empty()) throw Error("material " + name + ": magnetic dispersion has no device counterpart");
  MaterialSpec s = p.electric;
  s.mu = p.magnetic.mu;
  return s;
}

// ---------------------------------------------------------------- sources
// SoftSource (include/Source.hpp:183-261): adds signal(t) to the cell's flux.
template <class Mode, class scalar>
struct SoftSource {
  std::function<double(double)> signal;
  CellRef<Mode, scalar> loc;
  int field, comp;
  double at(double t) const { return signal(t); }
};
template <class F, class Signal, class Mode, class scalar>
SoftSource<Mode, scalar> make_soft_source(Signal sig, CellRef<Mode, scalar> c) {
  static_assert(F::field == FDTD_D || F::field == FDTD_B, "soft sources act on D or B (before the material step)");
  return {std::function<double(double)>(sig), c, F::field, F::comp};
}

// Sinusoid / Ricker signals (include/Source.hpp:100-167)
struct Sinusoid {
  double freq, phase, delay;
  explicit Sinusoid(double f, double ph = 0) : freq(f), phase(ph), delay(10.0 / f * pi / 2.0) {}
  double operator()(double t) const { return std::tanh(t / delay) * std::sin(freq * t + phase); }
};
struct Ricker {
  double freq, delay;
  explicit Ricker(double f) : freq(f), delay(2.0 / f) {}
  double operator()(double t) const {
    double p = pi * freq * (t - delay);
    return (1.0 - 2.0 * p * p) * std::exp(-p * p);
  }
};

// ------------------------------------------------------------------- TFSF
// Aux1D (include/Source.hpp:280-408): here only its size and source signal;
// the line itself is stepped on the device.
template <class scalar = double>
struct Aux1D {
  int ncells;
  double dx, dt;
  std::function<double(double)> signal;
  template <class Signal>
  Aux1D(int n, double deltax, double deltat, Signal sig) : ncells(n), dx(deltax), dt(deltat), signal(sig) {}
};

// TFSF box of 2*dim make_tfsf_plane faces sharing one incident line.
// Polarisations default to the reference's rule for the mode
// (include/Source.hpp:573-584; ThreeD, which the reference leaves unset,
// gets pol = z and polH = pol x k, consistent with the line's Ez/Hy).
template <class Mode>
struct TFSFBox {
  std::array<int, 6> box;
  std::array<double, 3> k, pol, polH, origin;
  std::shared_ptr<Aux1D<double>> aux;
};
template <class Mode, class scalar>
TFSFBox<Mode> make_tfsf_box(YeeGrid<Mode, scalar>& g, std::array<int, 6> box, std::array<double, 3> kvec,
                            std::shared_ptr<Aux1D<double>> aux) {
  TFSFBox<Mode> t;
  t.box = box;
  double K = 0;
  for (int d = 0; d < Mode::dim; d++) K += kvec[d] * kvec[d];
  K = std::sqrt(K);
  for (int d = 0; d < 3; d++) t.k[d] = d < Mode::dim ? kvec[d] / K : 0.0;
  t.pol = {0, 0, 1};
  if (Mode::dim == 2)
    t.polH = {t.k[1], -t.k[0], 0};
  else
    t.polH = {-t.k[1], t.k[0], 0};  // z x k
  for (int d = 0; d < 3; d++) t.origin[d] = d < Mode::dim ? std::floor(0.5 * (box[2 * d] + box[2 * d + 1])) + 0.5 : 0.0;
  t.aux = aux;
  (void)g;
  return t;
}

// ---------------------------------------------------------------- boundaries
// fdtd::Boundary (include/BoundaryUpdates.hpp:17-26).  make_*_boundary returns
// the face description the Stepper hands to fdtd_set_boundary; the reference's
// iterator pair over the boundary cells becomes the grid itself because the
// placement is fixed: ghost layer 0 / L-1, boundary layer 1 / L-2, periodic
// image of ghost 0 layer L-2 and of ghost L-1 layer 1.
enum class Boundary : char { Periodic = 0, BlochPeriodic, PEC, PMC, Symmetric, Antisymmetric, Parallel, NONE };

struct BoundaryData {
  Boundary type;
  Dir dir;
  Orientation orient;
  std::complex<double> factor = 1.0;  // BlochPeriodic: the phase factor of E (H takes its conjugate)
};

namespace detail {
inline int boundary_kind(Boundary b) {
  switch (b) {
    case Boundary::PEC: return FDTD_BC_PEC;
    case Boundary::PMC: return FDTD_BC_PMC;
    case Boundary::Symmetric: return FDTD_BC_SYMMETRIC;
    case Boundary::Antisymmetric: return FDTD_BC_ANTISYMMETRIC;
    case Boundary::Periodic: return FDTD_BC_PERIODIC;
    case Boundary::BlochPeriodic: return FDTD_BC_BLOCH;  // set through fdtd_set_bloch_boundary
    case Boundary::NONE: return FDTD_BC_NONE;
    default: throw Error("Parallel boundaries (MPI halo exchange) have no single-GPU counterpart");
  }
}
}  // namespace detail

template <class Mode, Dir d, Orientation o, class scalar>
BoundaryData make_pec_boundary(YeeGrid<Mode, scalar>&) { return {Boundary::PEC, d, o}; }
template <class Mode, Dir d, Orientation o, class scalar>
BoundaryData make_pmc_boundary(YeeGrid<Mode, scalar>&) { return {Boundary::PMC, d, o}; }
template <class Mode, Dir d, Orientation o, class scalar>
BoundaryData make_symmetric_boundary(YeeGrid<Mode, scalar>&) { return {Boundary::Symmetric, d, o}; }
template <class Mode, Dir d, Orientation o, class scalar>
BoundaryData make_antisymmetric_boundary(YeeGrid<Mode, scalar>&) { return {Boundary::Antisymmetric, d, o}; }
template <class Mode, Dir d, Orientation o, class scalar>
BoundaryData make_periodic_boundary(YeeGrid<Mode, scalar>&) { return {Boundary::Periodic, d, o}; }
// the grid's fields become complex (the reference needs cells with a complex
// scalar for this face); the imaginary part is read with FDTD_IMAG
template <class Mode, Dir d, Orientation o, class scalar>
BoundaryData make_bloch_periodic_boundary(YeeGrid<Mode, scalar>&, std::complex<double> factor) {
  return {Boundary::BlochPeriodic, d, o, factor};
}

// Per-face kinds, PEC everywhere by default as in the reference's
// BoundaryOptions (include/BoundaryUpdates.hpp:44-69).
struct BoundaryOptions {
  Boundary conds[3][2];
  BoundaryOptions() {
    for (auto& a : conds) a[0] = a[1] = Boundary::PEC;
  }
  Boundary& operator()(Dir d, Orientation o) { return conds[int(d)][int(o)]; }
  template <Dir d, Orientation o>
  Boundary& get() { return conds[int(d)][int(o)]; }
};

// ---------------------------------------------------------------- stepper
// Collects the stages of the reference's step composition and runs whole
// steps on the device.  Material functors are given per material id.
template <class Mode, class scalar>
class Stepper {
 public:
  explicit Stepper(YeeGrid<Mode, scalar>& g) : g_(g) {}

  template <class F>
  Stepper& then(const YeeUpdate<Mode, F>& u) {
    expect(u.stage, "YeeUpdate");
    check_dt(u.dt, u.dx);
    return *this;
  }
  template <FieldType ft, Dir d>
  Stepper& then(const UpdatePML<Mode, ft, d>& u) {
    int s = u.stage;
    if (stage_ == s + 1 || stage_ == s) s = stage_;  // one UpdatePML per axis
    expect(s, "UpdatePML");
    check_dt(u.dt, u.dx);
    return *this;
  }
  Stepper& source(const SoftSource<Mode, scalar>& s) {
    sources_.push_back(s);
    return *this;
  }
  Stepper& tfsf(const TFSFBox<Mode>& t) {
    tfsf_.reset(new TFSFBox<Mode>(t));
    return *this;
  }
  // ghost-layer writes of one face after every E and H update (BoundaryData::updateE / updateH)
  Stepper& boundary(const BoundaryData& b) {
    bcs_.push_back(b);
    return *this;
  }
  Stepper& boundaries(const BoundaryOptions& bo) {
    for (int d = 0; d < Mode::dim; d++)
      for (int o = 0; o < 2; o++) bcs_.push_back({bo.conds[d][o], Dir(d), Orientation(o)});
    return *this;
  }
  template <class M>
  Stepper& material(int id, const M& m) {
    set_material(g_, id, m.spec());
    if (stage_ < 3) stage_ = 3;
    return *this;
  }
  // every named pair of `mm` listed in `ids` (name -> material id)
  Stepper& materials(const MaterialMap& mm, const std::map<std::string, int>& ids) {
    for (const auto& kv : ids) {
      auto it = mm.find(kv.first);
      if (it == mm.end()) throw Error("no material named " + kv.first);
      set_material(g_, kv.second, combine(kv.first, it->second));
    }
    if (stage_ < 3) stage_ = 3;
    return *this;
  }

  // number of steps the source and incident tables cover (default: the first run's n)
  Stepper& horizon(int total_steps) {
    horizon_ = total_steps;
    return *this;
  }

  // advance n steps; source tables are sampled at t = (step index) * time_unit
  void run(int n, double time_unit = 1.0) {
    if (!committed_) commit(std::max(n, horizon_), time_unit);
    check(fdtd_step(g_.handle(), n), "fdtd_step");
  }
  void synchronize() { check(fdtd_synchronize(g_.handle()), "fdtd_synchronize"); }

 private:
  void expect(int s, const char* what) {
    if (s < stage_) throw Error(std::string(what) + " out of the reference's step order");
    stage_ = s;
  }
  void check_dt(double dt, double dx) {
    if (dt != g_.dt() || dx != g_.dx()) throw Error("functor dt/dx differ from the grid's");
  }
  void commit(int n, double unit) {
    g_.commit_pml();
    for (auto& b : bcs_) {
      if (b.type == Boundary::BlochPeriodic)
        check(fdtd_set_bloch_boundary(g_.handle(), int(b.dir), int(b.orient), b.factor.real(), b.factor.imag()),
              "fdtd_set_bloch_boundary");
      else
        check(fdtd_set_boundary(g_.handle(), int(b.dir), int(b.orient), detail::boundary_kind(b.type)),
              "fdtd_set_boundary");
    }
    long n0 = fdtd_step_count(g_.handle());
    for (auto& s : sources_) {
      std::vector<double> tab(size_t(n0 + n));
      for (size_t q = 0; q < tab.size(); q++) tab[q] = s.at(double(q) * unit);
      check(fdtd_add_source(g_.handle(), s.field, s.comp, s.loc.i, s.loc.j, s.loc.k, tab.data(), int(tab.size())),
            "fdtd_add_source");
    }
    if (tfsf_) {
      // the aux line advances its time before sampling the signal (Source.hpp:414)
      std::vector<double> sig(size_t(n0 + n));
      double t = 0;
      for (auto& v : sig) {
        t += tfsf_->aux->dt;
        v = tfsf_->aux->signal(t);
      }
      check(fdtd_set_tfsf(g_.handle(), tfsf_->box.data(), tfsf_->k.data(), tfsf_->pol.data(), tfsf_->polH.data(),
                          tfsf_->origin.data(), tfsf_->aux->ncells, sig.data(), int(sig.size())),
            "fdtd_set_tfsf");
    }
    committed_ = true;
  }
  YeeGrid<Mode, scalar>& g_;
  int stage_ = 0;
  int horizon_ = 0;
  bool committed_ = false;
  std::vector<SoftSource<Mode, scalar>> sources_;
  std::unique_ptr<TFSFBox<Mode>> tfsf_;
  std::vector<BoundaryData> bcs_;
};

}  // namespace hip
}  // namespace fdtd
