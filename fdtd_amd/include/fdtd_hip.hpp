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
//   make_tfsf_plane<Mode,Dir,Orient>(b,e,loc,k,aux) make_tfsf_plane<Mode,Dir,Orient>(grid,cells,origin,k,aux)  Source.hpp:624-634
//     (2*dim faces, one history)                    make_tfsf_box<Mode>(grid,box,k,aux)
//   plane.updateD(dt,dx) / updateB(dt,dx)           same, inside a for_each loop     Source.hpp:595-620
//   src.apply(t)                                    same, inside a for_each loop     Source.hpp:225-229
//   std::for_each(cells.begin(), cells.end(), F)    for_each(grid, F) or F(grid); M(grid.cells(id))  testTM.cpp:102-113
//   make_pec_boundary<Mode,Dir,Orient>(b,e)         make_pec_boundary<Mode,Dir,Orient>(grid)         BoundaryUpdates.hpp:527-565
//     (pmc, symmetric, antisymmetric, periodic)     same names; periodic image fixed (layer L-2 / 1)  BoundaryUpdates.hpp:578-832
//   BoundaryData::updateE() / updateH()             Stepper::boundary(bd): ghost writes after E / H   BoundaryUpdates.hpp:340-362
//   make_bloch_periodic_boundary<Mode,Dir,Orient>(b,e,f,factor)  make_bloch_periodic_boundary<Mode,Dir,Orient>(grid, factor)
//     (complex cells)                               complex fields; grid.download(field | FDTD_IMAG, ...)  BoundaryUpdates.hpp:845-885
//
// Instead of sweeping each functor over a user container, the functors are
// handed once to a Stepper, which checks that they form the reference's step
// composition (D curl, sources, CPML E, TFSF, E material, B curl, CPML H,
// TFSF, H material) and then runs whole steps on the GPU; or the reference's
// loop is kept, one for_each(grid, F) per sweep, each recording a stage, the
// recorded steps running on the GPU many at once.  Errors from the C
// ABI surface as fdtd::hip::Error exceptions on the host side only.
#pragma once

#include <algorithm>
#include <array>
#include <cmath>
#include <complex>
#include <cstdint>
#include <cstdio>
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

struct TEM {
  static constexpr int dim = 1, code = FDTD_TEM;
};
struct TM {
  static constexpr int dim = 2, code = FDTD_TM;
};
struct TE {
  static constexpr int dim = 2, code = FDTD_TE;
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

// material updates: one table entry each (eps_r or mu_r plus poles)
struct MaterialSpec {
  double eps = 1.0, mu = 1.0;
  std::vector<double> poles;  // kind, freq, gamma, weight per pole
  int mag = 0;                // magnetized Drude: 1 fixed bias, 2 the cell's own B
  double bias[3] = {0, 0, 0};
  bool fluid = false;         // fluid Drude: omega_p, gamma per cell (YeeGrid::setFluidState)
  std::vector<double> mpoles; // magnetic side: kind, freq, gamma, weight per pole (on B, H, M, K)
};

namespace detail {
// MaterialMap equivalent (include/MaterialMap.hpp:204-284): id -> material.
inline void set_material(fdtd_grid* h, int id, const MaterialSpec& s) {
  if (s.fluid) {
    check(fdtd_set_fluid_drude(h, id, s.mu, s.mag, s.mag == 1 ? s.bias : nullptr), "fdtd_set_fluid_drude");
    return;
  }
  if (s.mag) {
    check(fdtd_set_magnetized_drude(h, id, s.eps, s.mu, s.poles[1], s.poles[2], s.mag == 1 ? s.bias : nullptr),
          "fdtd_set_magnetized_drude");
    return;
  }
  check(fdtd_set_material(h, id, s.eps, s.mu, int(s.poles.size() / 4), s.poles.empty() ? nullptr : s.poles.data()),
        "fdtd_set_material");
  if (!s.mpoles.empty())
    check(fdtd_set_material_magnetic(h, id, int(s.mpoles.size() / 4), s.mpoles.data()), "fdtd_set_material_magnetic");
}

// the spec of a pole model: K is eps_r of an electric model, mu_r of a magnetic one
inline MaterialSpec pole_spec(FieldType ft, double K, std::vector<double> pole) {
  MaterialSpec s;
  (ft == FieldType::Electric ? s.eps : s.mu) = K;
  s.poles = std::move(pole);
  return s;
}

inline std::string spec_text(const MaterialSpec& s) {
  std::string t;
  char b[32];
  for (double v : {s.eps, s.mu, double(s.mag), s.bias[0], s.bias[1], s.bias[2], double(s.fluid)}) {
    snprintf(b, sizeof b, "%.17g,", v);
    t += b;
  }
  for (double v : s.poles) {
    snprintf(b, sizeof b, "%.17g,", v);
    t += b;
  }
  for (double v : s.mpoles) {
    snprintf(b, sizeof b, "m%.17g,", v);
    t += b;
  }
  return t;
}

// Stages of the reference's step composition (tests/testTM.cpp:102-113,
// tests/test3D.cpp:146-171, include/Source.hpp:413-435), in their order.
enum Stage : int { kCurlD = 0, kPmlE = 1, kTfsfD = 2, kMatE = 3, kCurlB = 4, kPmlH = 5, kTfsfB = 6, kMatH = 7 };

// The recorded time loop of one grid (for_each).  Each call of the loop names
// one stage of the step; a step is complete when its H material stage has
// been seen and the next step starts (or the grid is read).  The first step's
// list of stages is the loop's composition: every later step must repeat it.
// The steps run on the device, many at once, when the grid is read, at
// flush(), or every flush_every steps; the per-step values of the soft
// sources the loop applies and the incident lines' signals go to the device
// with them.
class Loop {
 public:
  int flush_every = 1024;
  int pml_axes = 0;      // axes with a CPML profile (their UpdatePML stages must be in the loop)
  int nsources = 0;      // sources of the grid (the loop's and others), the next one's index
  int nplanes = 0;       // single TFSF planes of the grid, likewise
  fdtd_grid* h = nullptr;
  std::function<void()> before_run;  // the grid's pending set-up (CPML profiles)

  // one stage of the current step, named by `tag` (the functor's signature)
  void event(int stage, const std::string& tag) {
    if (stage == kCurlD && stage_ >= kMatH) {
      end_step();
      if (pending_ >= flush_every) flush();
    }
    if (stage_ < 0 && stage != kCurlD) throw Error("a step of the time loop starts with the D curl (YeeUpdate of D)");
    if (stage < stage_) throw Error(tag + " out of the reference's step order");
    stage_ = stage;
    now_.push_back(tag);
  }
  // SoftSource::apply inside the loop: v is this step's value.  D and B
  // sources belong between their curl and their material stage, E right
  // after the E material stage, H right after the H material stage.
  void source(int field, int comp, int i, int j, int k, const std::shared_ptr<int>& id, double v) {
    const bool ok = field == FDTD_D   ? stage_ >= kCurlD && stage_ < kMatE
                    : field == FDTD_B ? stage_ >= kCurlB && stage_ < kMatH
                    : field == FDTD_E ? stage_ == kMatE
                                      : stage_ == kMatH;
    if (!ok)
      throw Error("SoftSource applied outside its place in the step (D, B: after their curl, before the material "
                  "step; E, H: right after their material step)");
    if (*id < 0) {
      const double none = 0.0;
      check(fdtd_add_source(h, field, comp, i, j, k, &none, 0), "fdtd_add_source");
      *id = nsources++;
    }
    auto& q = vals_[*id];
    q.resize(size_t(pending_ + 1), 0.0);
    q[size_t(pending_)] += v;
    now_.push_back("src" + std::to_string(*id));
  }
  // a TFSF box (kind FDTD_SIGNAL_TFSF) or plane (FDTD_SIGNAL_PLANE) whose
  // updateD (stage kTfsfD) or updateB (kTfsfB) the loop calls; `add`
  // registers it on first use and returns its index
  void line(int kind, const std::shared_ptr<int>& id, const std::function<int()>& add,
            const std::function<double(double)>& sig, double dt, int stage) {
    if (*id < 0) {
      if (stage != kTfsfD) throw Error("a TFSF box or plane's updateD comes before its updateB");
      *id = add();
      lines_.push_back({kind, *id, 0.0, dt, sig});
    }
    event(stage, "tfsf" + std::to_string(kind) + ":" + std::to_string(*id));
  }
  void material(int id, int ft, const MaterialSpec& m) {
    event(ft ? kMatH : kMatE, "mat" + std::to_string(ft) + ":" + std::to_string(id) + ":" + spec_text(m));
    if (!have_first_) (ft ? hmat_ : emat_)[id] = m;
  }
  // complete steps not yet run
  long pending() const { return pending_ + (stage_ >= kMatH ? 1 : 0); }
  // run the complete steps (reading the grid in the middle of a step is an error)
  void flush() {
    if (stage_ >= 0 && stage_ < kMatH) throw Error("the grid is read in the middle of a step of the time loop");
    if (stage_ >= kMatH) end_step();
    if (pending_ == 0) return;
    if (before_run) before_run();
    if (configure_) {
      configure();
      configure_ = false;
    }
    const long n0 = fdtd_step_count(h);
    for (auto& kv : vals_) {
      kv.second.resize(size_t(pending_), 0.0);
      check(fdtd_set_signal(h, FDTD_SIGNAL_SOURCE, kv.first, n0, kv.second.data(), int(pending_)), "fdtd_set_signal");
      kv.second.clear();
    }
    for (auto& l : lines_) {
      // the line advances its time before sampling the signal (Source.hpp:414)
      std::vector<double> v(static_cast<size_t>(pending_));
      for (auto& x : v) {
        l.t += l.dt;
        x = l.sig(l.t);
      }
      check(fdtd_set_signal(h, l.kind, l.index, n0, v.data(), int(pending_)), "fdtd_set_signal");
    }
    check(fdtd_step(h, int(pending_)), "fdtd_step");
    pending_ = 0;
  }

 private:
  struct Line {
    int kind, index;
    double t, dt;
    std::function<double(double)> sig;
  };
  void end_step() {
    if (!have_first_) {
      first_ = now_;
      have_first_ = true;
      configure_ = true;
    } else if (now_ != first_) {
      throw Error("every step of the time loop must repeat the first step's functors, sources and values");
    }
    now_.clear();
    stage_ = -1;
    pending_++;
  }
  // set up the device from the first step: its materials (the electric model
  // and the magnetic model's mu_r and poles per id); CPML profiles need their
  // UpdatePML stages
  void configure() {
    for (int a = 0; a < 3; a++)
      if ((pml_axes >> a) & 1)
        for (int ft = 0; ft < 2; ft++)
          if (std::find(first_.begin(), first_.end(), "pml" + std::to_string(ft) + std::to_string(a)) == first_.end())
            throw Error("the grid has a CPML profile along an axis whose UpdatePML is not in the time loop");
    std::map<int, MaterialSpec> all = emat_;
    for (auto& kv : hmat_) {
      if (kv.second.mag || kv.second.fluid) throw Error("a magnetized or fluid model on the magnetic side has no device counterpart");
      all[kv.first].mu = kv.second.mu;
      all[kv.first].mpoles = kv.second.poles;
    }
    for (auto& kv : all) set_material(h, kv.first, kv.second);
  }
  int stage_ = -1;
  long pending_ = 0;
  bool have_first_ = false, configure_ = false;
  std::vector<std::string> now_, first_;
  std::map<int, MaterialSpec> emat_, hmat_;
  std::map<int, std::vector<double>> vals_;  // the loop's sources: values of the pending steps
  std::vector<Line> lines_;
};
}  // namespace detail

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

// the cells of one material id of a grid (YeeGrid::cells)
template <class Mode, class scalar>
struct GridCells {
  YeeGrid<Mode, scalar>* grid;
  int id;
};

template <class Mode, class scalar>
class YeeGrid {
 public:
  YeeGrid(int nx, int ny, int nz, double dx, double dt)
      : nx_(nx), ny_(Mode::dim >= 2 ? ny : 1), nz_(Mode::dim == 3 ? nz : 1), dx_(dx), dt_(dt) {
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
    loop_.h = g;
    loop_.before_run = [this]() { commit_pml(); };
  }
  YeeGrid(int nx, int ny, double dx, double dt) : YeeGrid(nx, ny, 1, dx, dt) {}
  YeeGrid(int nx, double dx, double dt) : YeeGrid(nx, 1, 1, dx, dt) {}  // TEM: a line of nx cells
  YeeGrid(const YeeGrid&) = delete;
  YeeGrid& operator=(const YeeGrid&) = delete;

  // the recorded time loop (for_each); flush() runs its complete steps,
  // which every read of the grid does first
  detail::Loop& loop() const { return loop_; }
  void flush() const { loop_.flush(); }
  // steps run so far plus the loop's complete steps not yet run
  long steps() const { return fdtd_step_count(handle()) + loop_.pending(); }
  void check_dt(double dt, double dx) const {
    if (dt != dt_ || dx != dx_) throw Error("functor dt/dx differ from the grid's");
  }

  fdtd_grid* handle() const { return g_.get(); }
  int nx() const { return nx_; }
  int ny() const { return ny_; }
  int nz() const { return nz_; }
  double dx() const { return dx_; }
  double dt() const { return dt_; }

  CellRef<Mode, scalar> cell(int i, int j, int k = 0) { return {this, i, j, k}; }
  CellRef<Mode, scalar> cell(int i) { return {this, i, 0, 0}; }  // TEM: cell i of the line
  // the cells of one material id: the range a material functor sweeps
  GridCells<Mode, scalar> cells(int id) { return {this, id}; }

  // NewPML::setPMLParameters for every cell whose coordinate along `d` is idx
  template <FieldType ft, Dir d>
  void setPMLParameters(int idx, double K, double S, double A) {
    auto& v = ksa_[int(ft)][int(d)];
    v[3 * idx] = K;
    v[3 * idx + 1] = S;
    v[3 * idx + 2] = A;
    pml_dirty_[int(d)] = true;
    if (S != 0.0 || K != 1.0 || A != 0.0) loop_.pml_axes |= 1 << int(d);
  }

  // the reference's placement idiom on one axis (tests/testTM.cpp:46-71)
  template <Dir d>
  void placePML(const PMLParameterModel& p, int n) {
    check(fdtd_set_pml_idiom(handle(), int(d), n, p.mM, p.mMa, p.msMax, p.mkMax, p.maMax), "fdtd_set_pml_idiom");
    if (n > 0) loop_.pml_axes |= 1 << int(d);
  }

  // record one E or H component at a cell after every step
  template <class F>
  void add_probe(int i) {  // TEM: cell i of the line
    add_probe<F>(i, 0, 0);
  }
  template <class F>
  void add_probe(int i, int j, int k = 0) {
    static_assert(F::field == FDTD_E || F::field == FDTD_H, "probes record E or H");
    check(fdtd_add_probe(handle(), F::field, F::comp, i, j, k), "fdtd_add_probe");
    nprobes_++;
  }
  // [steps][probes]
  std::vector<double> probes() const {
    flush();
    long n = fdtd_step_count(handle());
    std::vector<double> v(size_t(std::max(1L, n)) * std::max(1, nprobes_));
    int got = 0;
    check(fdtd_get_probes(handle(), v.data(), int(n), &got), "fdtd_get_probes");
    v.resize(size_t(got) * nprobes_);
    return v;
  }

  // the imaginary part of the probes of a complex grid (a Bloch-periodic face)
  std::vector<double> probes_imag() const {
    flush();
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
    flush();
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
    flush();
    std::vector<double> v(size_t(nx_) * ny_ * nz_);
    check(fdtd_download_field(handle(), field, comp, pole, v.data()), "fdtd_download_field");
    return v;
  }
  void upload(int field, int comp, const std::vector<double>& v) {
    flush();
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
  mutable detail::Loop loop_;
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
// Value types naming one stage of the step.  Handed to a Stepper, which
// checks their order; or called on a grid, f(grid) (a material functor also
// on grid.cells(id)), which is for_each(grid, f): the reference's sweep of
// the functor over the cells (tests/testTM.cpp:102-113).
namespace detail {
template <class F>
struct Sweep {
  template <class G>
  void operator()(G&& g) const {
    for_each(g, static_cast<const F&>(*this));
  }
};
}  // namespace detail

template <class Mode, class Flux>
struct YeeUpdate : detail::Sweep<YeeUpdate<Mode, Flux>> {
  double dt, dx;
  YeeUpdate(double deltat, double deltax) : dt(deltat), dx(deltax) {}
  static constexpr int stage = Flux::field == FDTD_D || Flux::field == FDTD_E ? 0 : 4;
};
template <class Mode, FieldType ft, Dir d>
struct UpdatePML : detail::Sweep<UpdatePML<Mode, ft, d>> {
  double dt, dx;
  UpdatePML(double deltat, double deltax) : dt(deltat), dx(deltax) {}
  static constexpr int stage = ft == FieldType::Electric ? 1 : 5;
};

// material updates: one table entry each (eps_r or mu_r plus poles), MaterialSpec above
template <class Mode, FieldType ft, bool forward = false>
struct VacuumUpdate : detail::Sweep<VacuumUpdate<Mode, ft, forward>> {
  MaterialSpec spec() const { return {}; }
};
template <class Mode, FieldType ft, bool forward = false>
struct ConstantUpdate : detail::Sweep<ConstantUpdate<Mode, ft, forward>> {
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
struct DrudeUpdate : detail::Sweep<DrudeUpdate<Mode, ft, forward>> {
  double K, freq, gamma, dt;
  DrudeUpdate(double k, double f, double g, double deltat) : K(k), freq(f), gamma(g), dt(deltat) {}
  MaterialSpec spec() const { return detail::pole_spec(ft, K, {double(FDTD_DRUDE), freq, gamma, 1.0}); }
};
// Lorentz pole with LorentzStorage(K, Delta, Freq, Gamma, dt); like the
// reference's update, Delta does not enter (include/DispersiveMaterials.hpp:711-735)
template <class Mode, FieldType ft, bool forward = false>
struct LorentzUpdate : detail::Sweep<LorentzUpdate<Mode, ft, forward>> {
  double K, delta, freq, gamma, dt;
  LorentzUpdate(double k, double de, double f, double g, double deltat) : K(k), delta(de), freq(f), gamma(g), dt(deltat) {}
  MaterialSpec spec() const { return detail::pole_spec(ft, K, {double(FDTD_LORENTZ), freq, gamma, 1.0}); }
};
// Conductive medium with ConductiveStorage(K, Freq, dt), Freq being the
// conductivity (include/DispersiveMaterials.hpp:508-680): P relaxes
// implicitly towards D with factor dt Freq/(K eps0)
template <class Mode, FieldType ft, bool forward = false>
struct ConductiveUpdate : detail::Sweep<ConductiveUpdate<Mode, ft, forward>> {
  double K, freq, dt;
  ConductiveUpdate(double k, double f, double deltat) : K(k), freq(f), dt(deltat) {}
  MaterialSpec spec() const { return detail::pole_spec(ft, K, {double(FDTD_CONDUCTIVE), freq, 0.0, 1.0}); }
};
// Magnetized Drude plasma with MagnetizedDrudeStorage(Bx, By, Bz, K, Freq,
// Gamma, dt): in 3-D the three components of P and J advance together in the
// fixed bias field B [T] (include/DispersiveMaterials.hpp:1263-1340); in TM
// the reference's update is the Drude one per component, and so is this one
template <class Mode, FieldType ft, bool forward = false>
struct MagnetizedDrudeUpdate : detail::Sweep<MagnetizedDrudeUpdate<Mode, ft, forward>> {
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
struct SelfMagnetizedDrudeUpdate : detail::Sweep<SelfMagnetizedDrudeUpdate<Mode, ft, forward>> {
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
struct FluidDrudeUpdate : detail::Sweep<FluidDrudeUpdate<Mode, ft, forward>> {
  double dt;
  explicit FluidDrudeUpdate(double deltat) : dt(deltat) {}
  MaterialSpec spec() const {
    MaterialSpec s{1.0, 1.0, {double(FDTD_DRUDE), 0.0, 0.0, 1.0}};
    s.fluid = true;
    return s;
  }
};
template <class Mode, FieldType ft, bool forward = false>
struct MagnetizedFluidDrudeUpdate : detail::Sweep<MagnetizedFluidDrudeUpdate<Mode, ft, forward>> {
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
struct SelfMagnetizedFluidDrudeUpdate : detail::Sweep<SelfMagnetizedFluidDrudeUpdate<Mode, ft, forward>> {
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
  detail::set_material(g.handle(), id, s);
}

// MaterialPair / MaterialMap (include/MaterialMap.hpp:45-284): a named pair
// of electric and magnetic models.  The device takes the electric model's
// eps_r and poles and the magnetic model's mu_r and poles (Drude, Lorentz,
// conductive on B, H, M, K).  The reference's XML form is read and written by
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
  if (p.magnetic.mag || p.magnetic.fluid)
    throw Error("material " + name + ": a magnetized or fluid model on the magnetic side has no device counterpart");
  MaterialSpec s = p.electric;
  s.mu = p.magnetic.mu;
  s.mpoles = p.magnetic.poles;
  return s;
}

// ---------------------------------------------------------------- sources
// SoftSource (include/Source.hpp:183-261): adds signal(t) to one field of a
// cell: D or B (after their curl: a lasting source) or E or H (after their
// material step: acting on one curl of the other field, Source.hpp:421-425).
// Handed to a Stepper, its signal is sampled once per step up front; called
// as apply(t) / (t) inside a time loop of for_each calls, signal(t) is that
// step's value, as in the reference's loops.
template <class Mode, class scalar>
struct SoftSource {
  std::function<double(double)> signal;
  CellRef<Mode, scalar> loc;
  int field, comp;
  std::shared_ptr<int> id = std::make_shared<int>(-1);  // the device source, once the loop has added it
  double at(double t) const { return signal(t); }
  void apply(double t) const { loc.grid->loop().source(field, comp, loc.i, loc.j, loc.k, id, signal(t)); }
  void operator()(double t) const { apply(t); }
};
template <class F, class Signal, class Mode, class scalar>
SoftSource<Mode, scalar> make_soft_source(Signal sig, CellRef<Mode, scalar> c) {
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

namespace detail {
// Polarisations by the reference's rule for the mode (include/Source.hpp:
// 573-584): TM pol = z, polH = (k_y, -k_x, 0); TE pol = (k_y, -k_x, 0),
// polH = z; TEM pol = z, polH = y.  ThreeD, which the reference leaves unset, gets pol = z and
// polH = pol x k = (-k_y, k_x, 0).  (The TM
// rule's polH is the negative of pol x k: with it the incident field the
// faces subtract is not the line's own plane wave, so the scattered-field
// region is not left empty; the rule is kept for parity with the reference,
// and the explicit-polarisation overloads take pol x k instead.)
inline void default_pol(int mode, const std::array<double, 3>& k, std::array<double, 3>& pol,
                        std::array<double, 3>& polH) {
  pol = {0, 0, 1};
  if (mode == FDTD_TE) {
    pol = {k[1], -k[0], 0};
    polH = {0, 0, 1};
  } else if (mode == FDTD_TEM) {
    polH = {0, 1, 0};
  } else if (mode == FDTD_TM) {
    polH = {k[1], -k[0], 0};
  } else {
    polH = {-k[1], k[0], 0};
  }
}
inline std::array<double, 3> unit(int dim, const std::array<double, 3>& kvec) {
  double K = 0;
  for (int d = 0; d < dim; d++) K += kvec[d] * kvec[d];
  K = std::sqrt(K);
  std::array<double, 3> k{};
  for (int d = 0; d < 3; d++) k[d] = d < dim ? kvec[d] / K : 0.0;
  return k;
}
}  // namespace detail

// TFSF box of 2*dim make_tfsf_plane faces sharing one incident line.  Handed
// to a Stepper (tfsf), or stepped by updateD / updateB inside a time loop.
template <class Mode>
struct TFSFBox {
  std::array<int, 6> box;
  std::array<double, 3> k, pol, polH, origin;
  std::shared_ptr<Aux1D<double>> aux;
  detail::Loop* loop = nullptr;
  std::shared_ptr<int> id = std::make_shared<int>(-1);
  void updateD(double, double) const { step(detail::kTfsfD); }
  void updateB(double, double) const { step(detail::kTfsfB); }

 private:
  void step(int stage) const {
    TFSFBox b = *this;
    loop->line(FDTD_SIGNAL_TFSF, id, [b]() {
      check(fdtd_set_tfsf(b.loop->h, b.box.data(), b.k.data(), b.pol.data(), b.polH.data(), b.origin.data(),
                          b.aux->ncells, nullptr, 0), "fdtd_set_tfsf");
      return 0;
    }, aux->signal, aux->dt, stage);
  }
};
template <class Mode, class scalar>
TFSFBox<Mode> make_tfsf_box(YeeGrid<Mode, scalar>& g, std::array<int, 6> box, std::array<double, 3> kvec,
                            std::shared_ptr<Aux1D<double>> aux) {
  TFSFBox<Mode> t;
  t.box = box;
  t.k = detail::unit(Mode::dim, kvec);
  detail::default_pol(Mode::code, t.k, t.pol, t.polH);
  for (int d = 0; d < 3; d++) t.origin[d] = d < Mode::dim ? std::floor(0.5 * (box[2 * d] + box[2 * d + 1])) + 0.5 : 0.0;
  t.aux = aux;
  t.loop = &g.loop();
  return t;
}

// One TFSF plane with its own incident line: make_tfsf_plane<Mode, Dir,
// Orient>(begin, end, loc, k, aux) (include/Source.hpp:624-634) with the
// iterator range given as the cells [x0,x1)x[y0,y1)x[z0,z1) of the grid (one
// layer thick along Dir) and the location functor as r = (cell - origin) dx.
// Handed to a Stepper (tfsf_plane), or stepped by updateD / updateB inside a
// time loop (TFSFPlane::updateD / updateB, 595-620).
template <class Mode>
struct TFSFPlane {
  int axis, side;
  std::array<int, 6> cells;
  std::array<double, 3> k, pol, polH, origin;
  std::shared_ptr<Aux1D<double>> aux;
  detail::Loop* loop = nullptr;
  std::shared_ptr<int> id = std::make_shared<int>(-1);
  void updateD(double, double) const { step(detail::kTfsfD); }
  void updateB(double, double) const { step(detail::kTfsfB); }
  int add(fdtd_grid* h, const double* sig, int nsig) const {
    check(fdtd_add_tfsf_plane(h, axis, side, cells.data(), k.data(), pol.data(), polH.data(), origin.data(),
                              aux->ncells, sig, nsig), "fdtd_add_tfsf_plane");
    return loop->nplanes++;
  }

 private:
  void step(int stage) const {
    TFSFPlane p = *this;
    loop->line(FDTD_SIGNAL_PLANE, id, [p]() { return p.add(p.loop->h, nullptr, 0); }, aux->signal, aux->dt, stage);
  }
};
template <class Mode, Dir d, Orientation o, class scalar>
TFSFPlane<Mode> make_tfsf_plane(YeeGrid<Mode, scalar>& g, std::array<int, 6> cells, std::array<double, 3> origin,
                                std::array<double, 3> kvec, std::shared_ptr<Aux1D<double>> aux) {
  static_assert(int(d) < Mode::dim, "the plane's normal must be an axis of the mode");
  TFSFPlane<Mode> p;
  p.axis = int(d);
  p.side = int(o);
  p.cells = cells;
  p.k = detail::unit(Mode::dim, kvec);
  detail::default_pol(Mode::code, p.k, p.pol, p.polH);
  p.origin = origin;
  p.aux = aux;
  p.loop = &g.loop();
  return p;
}
// the same with explicit polarisations (pol multiplies the line's E, polH its H)
template <class Mode, Dir d, Orientation o, class scalar>
TFSFPlane<Mode> make_tfsf_plane(YeeGrid<Mode, scalar>& g, std::array<int, 6> cells, std::array<double, 3> origin,
                                std::array<double, 3> kvec, std::array<double, 3> pol, std::array<double, 3> polH,
                                std::shared_ptr<Aux1D<double>> aux) {
  TFSFPlane<Mode> p = make_tfsf_plane<Mode, d, o>(g, cells, origin, kvec, aux);
  p.pol = pol;
  p.polH = polH;
  return p;
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
  Stepper& tfsf_plane(const TFSFPlane<Mode>& p) {
    planes_.push_back(p);
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
      g_.loop().nsources++;
    }
    for (auto& p : planes_) p.add(g_.handle(), line_signal(*p.aux, n0 + n).data(), int(n0 + n));
    if (tfsf_) {
      std::vector<double> sig = line_signal(*tfsf_->aux, n0 + n);
      check(fdtd_set_tfsf(g_.handle(), tfsf_->box.data(), tfsf_->k.data(), tfsf_->pol.data(), tfsf_->polH.data(),
                          tfsf_->origin.data(), tfsf_->aux->ncells, sig.data(), int(sig.size())),
            "fdtd_set_tfsf");
    }
    committed_ = true;
  }
  // the line's signal for steps [0, n): it advances its time before sampling (Source.hpp:414)
  static std::vector<double> line_signal(const Aux1D<double>& aux, long n) {
    std::vector<double> sig(static_cast<size_t>(n));
    double t = 0;
    for (auto& v : sig) {
      t += aux.dt;
      v = aux.signal(t);
    }
    return sig;
  }
  YeeGrid<Mode, scalar>& g_;
  int stage_ = 0;
  int horizon_ = 0;
  bool committed_ = false;
  std::vector<SoftSource<Mode, scalar>> sources_;
  std::unique_ptr<TFSFBox<Mode>> tfsf_;
  std::vector<TFSFPlane<Mode>> planes_;
  std::vector<BoundaryData> bcs_;
};

// ------------------------------------------------------------ time loop
// for_each(grid, functor): the reference's sweep of one functor over the
// grid's cells, std::for_each(cells.begin(), cells.end(), F(...))
// (tests/testTM.cpp:102-113), so that the reference's time loop reads
//
//   for (int t = 0; t < n; t++) {
//     for_each(grid, YeeUpdate<TM, Dz>(dt, dx));
//     src.apply(t);                                      // SoftSource on Dz
//     for_each(grid, UpdatePML<TM, FieldType::Electric, Dir::X>(dt, dx));
//     ...
//     plane.updateD(dt, dx);                             // TFSF plane or box
//     for_each(grid.cells(1), ConstantUpdate<TM, FieldType::Electric>(4.0));
//     for_each(grid, YeeUpdate<TM, Bx>(dt, dx));
//     ...
//     for_each(grid, VacuumUpdate<TM, FieldType::Magnetic>());
//   }
//
// Each call records one stage of the step on the grid's loop; complete steps
// run on the device as fused passes, many steps at once, when the grid is
// read (probes, download, sync_host), at flush(), or every
// loop().flush_every steps.  Every step must repeat the first one's
// functors, sources and values.  A material functor sweeps every cell
// (material id 0) or grid.cells(id); CPML runs where profiles are set, whose
// axes need their UpdatePML stages in the loop.
template <class Mode, class scalar, class Flux>
void for_each(YeeGrid<Mode, scalar>& g, const YeeUpdate<Mode, Flux>& u) {
  g.check_dt(u.dt, u.dx);
  g.loop().event(u.stage, "curl" + std::to_string(Flux::field) + std::to_string(Flux::comp));
}
template <class Mode, class scalar, FieldType ft, Dir d>
void for_each(YeeGrid<Mode, scalar>& g, const UpdatePML<Mode, ft, d>& u) {
  static_assert(int(d) < Mode::dim, "UpdatePML along an axis the mode does not have");
  g.check_dt(u.dt, u.dx);
  g.loop().event(u.stage, "pml" + std::to_string(int(ft)) + std::to_string(int(d)));
}
template <class Mode, class scalar, template <class, FieldType, bool> class M, FieldType ft, bool fwd>
void for_each(YeeGrid<Mode, scalar>& g, const M<Mode, ft, fwd>& m) {
  g.loop().material(0, int(ft), m.spec());
}
template <class Mode, class scalar, template <class, FieldType, bool> class M, FieldType ft, bool fwd>
void for_each(GridCells<Mode, scalar> c, const M<Mode, ft, fwd>& m) {
  c.grid->loop().material(c.id, int(ft), m.spec());
}

}  // namespace hip
}  // namespace fdtd
