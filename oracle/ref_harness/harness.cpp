// Oracle harness (test infrastructure only).
//
// Drives the reference's own header-only templates (pederson/fdtd, fdtd.h and
// include/*.hpp, taken with -I<reference root>) on small seeded scenarios and
// writes golden fixtures: one JSON file (parameters + array index) and one raw
// little-endian float64 .bin per scenario.  It also has a timing mode used as
// the measured CPU baseline of the reference's own update functors.
//
// Nothing here is shipped: the binary lands in oracle/_ref/ and is built only
// where the reference tree exists.  The step composition follows the
// reference's own idiom (tests/testTM.cpp:102-113, tests/test3D.cpp:146-171,
// include/Source.hpp:413-435): one sweep per functor over the interior cells
// (indices 1..N-2), outermost layer never updated.
#include <fdtd.h>

#include <array>
#include <chrono>
#include <complex>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <memory>
#include <sstream>
#include <string>
#include <vector>

namespace fdtd {
// The primary DispersiveAtomicUpdate (include/DispersiveMaterials.hpp:259-282)
// puts std::enable_if_t<implicit> on a class template parameter, which is a
// hard error as soon as the class is instantiated.  This partial
// specialisation supplies the implicit branch with exactly the argument list
// of the primary (flux, field, polarization, current, then the storage
// arguments), so the reference's own *::implicit_update arithmetic runs
// unchanged.
template <typename U, typename F>
struct DispersiveAtomicUpdate<U, true, F> {
  static constexpr Dir d = FieldDir<F>::value;
  static constexpr FieldType ft = F::field_type;
  template <typename T, typename... A>
  static void get(T&& f, A&&... a) {
    U::implicit_update(GetField<typename FluxComponent<ft, d>::type>::get(f),
                       GetField<typename FieldComponent<ft, d>::type>::get(f),
                       GetField<typename PolarizationComponent<ft, d>::type>::get(f),
                       GetField<typename CurrentComponent<ft, d>::type>::get(f),
                       std::forward<A>(a)...);
  }
};
}  // namespace fdtd

using namespace fdtd;

// ---------------------------------------------------------------- utilities
struct SplitMix64 {
  uint64_t s;
  explicit SplitMix64(uint64_t seed) : s(seed) {}
  uint64_t next() {
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
  }
  double uniform() { return (next() >> 11) * (1.0 / 9007199254740992.0); }
  int randint(int lo, int hi) { return lo + int(next() % uint64_t(hi - lo + 1)); }
};

struct Fixture {
  std::string name;
  std::string params;  // JSON object body (without braces)
  std::vector<std::pair<std::string, std::vector<size_t>>> index;
  std::vector<double> data;
  void p(const std::string& k, double v) {
    char buf[64];
    snprintf(buf, sizeof buf, "%.17g", v);
    add(k, buf);
  }
  void pi(const std::string& k, long v) { add(k, std::to_string(v)); }
  void ps(const std::string& k, const std::string& v) { add(k, "\"" + v + "\""); }
  void add(const std::string& k, const std::string& raw) {
    if (!params.empty()) params += ",\n    ";
    params += "\"" + k + "\": " + raw;
  }
  void arr(const std::string& k, const std::vector<double>& v, std::vector<size_t> shape) {
    size_t n = 1;
    for (auto s : shape) n *= s;
    if (n != v.size()) {
      fprintf(stderr, "shape mismatch for %s\n", k.c_str());
      exit(3);
    }
    index.push_back({k, shape});
    data.insert(data.end(), v.begin(), v.end());
  }
  void write(const std::string& dir) const {
    std::string js = "{\n  \"name\": \"" + name + "\",\n  \"params\": {\n    " + params + "\n  },\n  \"arrays\": [";
    size_t off = 0;
    for (size_t a = 0; a < index.size(); a++) {
      size_t n = 1;
      std::string sh;
      for (size_t s = 0; s < index[a].second.size(); s++) {
        n *= index[a].second[s];
        sh += (s ? ", " : "") + std::to_string(index[a].second[s]);
      }
      js += std::string(a ? "," : "") + "\n    {\"name\": \"" + index[a].first + "\", \"shape\": [" + sh +
            "], \"offset\": " + std::to_string(off) + "}";
      off += n;
    }
    js += "\n  ]\n}\n";
    FILE* f = fopen((dir + "/" + name + ".json").c_str(), "w");
    fputs(js.c_str(), f);
    fclose(f);
    f = fopen((dir + "/" + name + ".bin").c_str(), "wb");
    fwrite(data.data(), sizeof(double), data.size(), f);
    fclose(f);
  }
};

template <class Base>
struct Nb : public Base {
  Nb* mn[3] = {nullptr, nullptr, nullptr};
  Nb* mx[3] = {nullptr, nullptr, nullptr};
  Nb& getNeighborMin(int d) const { return *mn[d]; }
  Nb& getNeighborMax(int d) const { return *mx[d]; }
};

// the electron density and collision rate that the reference's fluid Drude
// storages read from a cell (f.ne(), f.nu_m(): DispersiveMaterials.hpp:1166-1168)
template <class Base>
struct FluidCell : public Base {
  double mne = 0, mnu = 0;
  double& ne() { return mne; }
  double& nu_m() { return mnu; }
};

typedef Nb<FluidCell<YeeCell<YeeFields<TM, double, std::array>, DoublePolarization<double>,
                             VacuumMagnetization<double>, NewPML<TM, double>>>>
    CellTM;
typedef Nb<FluidCell<YeeCell<YeeFields<ThreeD, double, std::array>, DoublePolarization<double>,
                             VacuumMagnetization<double>, NewPML<ThreeD, double>>>>
    Cell3;
typedef Nb<FluidCell<YeeCell<YeeFields<TE, double, std::array>, DoublePolarization<double>,
                             VacuumMagnetization<double>, NewPML<TE, double>>>>
    CellTE;
typedef Nb<FluidCell<YeeCell<YeeFields<TEM, double, std::array>, DoublePolarization<double>,
                             VacuumMagnetization<double>, NewPML<TEM, double>>>>
    CellTEM;
// a 3-D cell that also carries the magnetization M and magnetic current K of
// the magnetic-side pole models (DoubleMagnetization, DispersiveMaterials.hpp:86-110)
typedef Nb<FluidCell<YeeCell<YeeFields<ThreeD, double, std::array>, DoublePolarization<double>,
                             DoubleMagnetization<double>, NewPML<ThreeD, double>>>>
    Cell3M;

// material kinds: 0 vacuum, 1 constant, 2 Drude, 3 Lorentz, 4 conductive (freq = the conductivity),
// 5 magnetized Drude with the bias (bx, by, bz) [T], 6 self-magnetized Drude (the cell's own B),
// 7 fluid Drude, 8 magnetized fluid Drude (bias), 9 self-magnetized fluid Drude (the fluid
// forms read omega_p and gamma from the cell's ne() and nu_m())
struct Mat {
  int kind;
  double eps, freq, gamma, delta;
  double bx = 0, by = 0, bz = 0;
};

template <class Mode, class C>
void material_E(C& c, const Mat& m, double dt) {
  switch (m.kind) {
    case 0: VacuumUpdate<Mode, FieldType::Electric>()(c); break;
    case 1: ConstantUpdate<Mode, FieldType::Electric>(m.eps)(c); break;
    case 2: DrudeUpdate<Mode, FieldType::Electric>(m.eps, m.freq, m.gamma, dt)(c); break;
    case 3: LorentzUpdate<Mode, FieldType::Electric>(m.eps, m.delta, m.freq, m.gamma, dt)(c); break;
    case 4: ConductiveUpdate<Mode, FieldType::Electric>(m.eps, m.freq, dt)(c); break;
    case 5: MagnetizedDrudeUpdate<Mode, FieldType::Electric>(m.bx, m.by, m.bz, m.eps, m.freq, m.gamma, dt)(c); break;
    case 6: {
      SelfMagnetizedDrudeUpdate<Mode, FieldType::Electric> u(dt);
      u.K() = m.eps;
      u.Freq() = m.freq;
      u.Gamma() = m.gamma;
      u(c);
      break;
    }
    case 7: (FluidDrudeUpdate<Mode, FieldType::Electric>(dt))(c); break;
    case 8: MagnetizedFluidDrudeUpdate<Mode, FieldType::Electric>(m.bx, m.by, m.bz, dt)(c); break;
    case 9: (SelfMagnetizedFluidDrudeUpdate<Mode, FieldType::Electric>(dt))(c); break;
  }
}

// the magnetic side of a material (the reference's functors on FieldType::Magnetic,
// which run the same updates on B, H, M, K with mu0 for eps0): kinds 0 vacuum,
// 1 constant (mu = mu_r), 2 Drude, 3 Lorentz, 4 conductive (freq = the magnetic conductivity)
struct MatM {
  int kind;
  double mu, freq, gamma, delta;
};

template <class Mode, class C>
void material_H(C& c, const MatM& m, double dt) {
  switch (m.kind) {
    case 0: VacuumUpdate<Mode, FieldType::Magnetic>()(c); break;
    case 1: ConstantUpdate<Mode, FieldType::Magnetic>(m.mu)(c); break;
    case 2: DrudeUpdate<Mode, FieldType::Magnetic>(m.mu, m.freq, m.gamma, dt)(c); break;
    case 3: LorentzUpdate<Mode, FieldType::Magnetic>(m.mu, m.delta, m.freq, m.gamma, dt)(c); break;
    case 4: ConductiveUpdate<Mode, FieldType::Magnetic>(m.mu, m.freq, dt)(c); break;
  }
}

struct PmlSpec {
  int n;
  double m, ma, smax, kmax, amax;
};

// PML placement idiom of tests/testTM.cpp:46-71 and include/Source.hpp:380-389:
// for i = 1..n+1, x = (n-i+1)/n, xm = x - 0.5/n; E parameters at i and L-i,
// H parameters at i and L-1-i.  `set(idx, electric, K, S, A)` applies to every
// cell whose coordinate along the axis equals idx.
template <class Setter>
void place_pml(int L, const PmlSpec& s, Setter set) {
  PMLParameterModel p(s.m, s.ma, s.smax, s.kmax, s.amax);
  for (int i = 1; i < s.n + 2; i++) {
    double x = static_cast<double>(s.n - i + 1) / static_cast<double>(s.n);
    double xm = x - 0.5 / static_cast<double>(s.n);
    set(i, true, p.K(x), p.S(x), p.A(x));
    set(i, false, p.K(xm), p.S(xm), p.A(xm));
    set(L - i, true, p.K(x), p.S(x), p.A(x));
    set(L - 1 - i, false, p.K(xm), p.S(xm), p.A(xm));
  }
}

static void put_pml(Fixture& fx, const PmlSpec& s) {
  fx.pi("pml_n", s.n);
  fx.p("pml_m", s.m);
  fx.p("pml_ma", s.ma);
  fx.p("pml_smax", s.smax);
  fx.p("pml_kmax", s.kmax);
  fx.p("pml_amax", s.amax);
}

// ------------------------------------------------------- TFSF plane iterator
template <class C>
struct PlaneEntry {
  C* cell;
  double x, y, z;
};
template <class C>
struct PlaneIt {
  PlaneEntry<C>* p;
  C& operator*() const { return *p->cell; }
  PlaneIt& operator++() {
    ++p;
    return *this;
  }
  PlaneIt operator++(int) {
    PlaneIt t = *this;
    ++p;
    return t;
  }
  bool operator!=(const PlaneIt& o) const { return p != o.p; }
  bool operator==(const PlaneIt& o) const { return p == o.p; }
};
struct LocTE {
  std::array<double, 2> operator()(const PlaneIt<CellTE>& it) const { return {it.p->x, it.p->y}; }
};
struct LocTM {
  std::array<double, 2> operator()(const PlaneIt<CellTM>& it) const { return {it.p->x, it.p->y}; }
};
struct LocTEM {
  std::array<double, 1> operator()(const PlaneIt<CellTEM>& it) const { return {it.p->x}; }
};

// ------------------------------------------------------------ scenario: PML
static void scen_pml_coeffs(const std::string& dir) {
  Fixture fx;
  fx.name = "pml_coeffs";
  // placement idiom on a 1-D line of TM cells for three (dx, dt, n) sets
  const double dxs[3] = {1e-6, 1.0, 1e-8};
  const int ns[3] = {10, 12, 8};
  const double cfls[3] = {0.5 / std::sqrt(3.0), 0.5, 0.9 / std::sqrt(2.0)};
  const int L = 40;
  for (int t = 0; t < 3; t++) {
    double dx = dxs[t], dt = cfls[t] * dx / c0;
    PMLParameterModel dflt(dx);
    PmlSpec s{ns[t], dflt.mM, dflt.mMa, dflt.msMax, dflt.mkMax, dflt.maMax};
    std::vector<NewPML<TM, double>> line(L);
    place_pml(L, s, [&](int idx, bool e, double K, double S, double A) {
      if (e)
        line[idx].setPMLParameters<FieldType::Electric, Dir::X>(K, S, A, dt);
      else
        line[idx].setPMLParameters<FieldType::Magnetic, Dir::X>(K, S, A, dt);
    });
    std::vector<double> out;
    for (int i = 0; i < L; i++) {
      auto& c = line[i];
      double v[14] = {c.pmlK<FieldType::Electric, Dir::X>(), c.pmlS<FieldType::Electric, Dir::X>(),
                      c.pmlA<FieldType::Electric, Dir::X>(), c.pmlB<FieldType::Electric, Dir::X>(),
                      c.pmlC<FieldType::Electric, Dir::X>(), c.pmlF<FieldType::Electric, Dir::X>(),
                      c.pmlG<FieldType::Electric, Dir::X>(), c.pmlK<FieldType::Magnetic, Dir::X>(),
                      c.pmlS<FieldType::Magnetic, Dir::X>(), c.pmlA<FieldType::Magnetic, Dir::X>(),
                      c.pmlB<FieldType::Magnetic, Dir::X>(), c.pmlC<FieldType::Magnetic, Dir::X>(),
                      c.pmlF<FieldType::Magnetic, Dir::X>(), c.pmlG<FieldType::Magnetic, Dir::X>()};
      out.insert(out.end(), v, v + 14);
    }
    std::string k = "set" + std::to_string(t);
    fx.p(k + "_dx", dx);
    fx.p(k + "_dt", dt);
    put_pml(fx, s);  // last one wins for the scalar keys; per-set keys below
    fx.pi(k + "_n", s.n);
    fx.p(k + "_smax", s.smax);
    fx.arr(k + "_KSABCFG_EH", out, {size_t(L), 14});
  }
  fx.pi("L", L);
  // model evaluations at sample x (incl. x<0 and x>1 branches)
  PMLParameterModel pm(3.0, 1.0, 123.0, 5.0, 0.05);
  std::vector<double> xs = {-0.25, -0.05, 0.0, 0.05, 0.1, 0.333, 0.5, 0.95, 1.0, 1.05, 2.0};
  std::vector<double> ev;
  for (double x : xs) {
    ev.push_back(x);
    ev.push_back(pm.K(x));
    ev.push_back(pm.S(x));
    ev.push_back(pm.A(x));
  }
  fx.arr("model_x_K_S_A", ev, {xs.size(), 4});
  fx.p("imp0", imp0);
  fx.p("eps0", eps0);
  fx.p("mu0", mu0);
  fx.p("c0", c0);
  fx.write(dir);
}

// ----------------------------------------------------------- scenario: Aux1D
static void scen_aux1d(const std::string& dir) {
  Fixture fx;
  fx.name = "aux1d";
  const int ncells = 40, nsteps = 220;
  const double dx = 1e-6, dt = 0.5 * dx / c0;
  const double f0 = c0 / (20 * dx), tw = 2.0 / f0, t0 = 3.0 * tw;
  std::vector<double> st, sv;
  auto sig = [&](double t) {
    double a = (t - t0) / tw;
    double v = std::exp(-a * a) * std::sin(2 * pi * f0 * (t - t0));
    st.push_back(t);
    sv.push_back(v);
    return v;
  };
  Aux1D<double> aux(ncells, dx, dt, sig);
  const int L = ncells + 2 * 15 + 20;
  std::vector<double> eall, hall, locs = {-20.0, -7.25, 0.0, 3.5, 11.0, 19.75};
  std::vector<double> trace;
  for (int n = 0; n < nsteps; n++) {
    aux.time_step(dt);
    for (double l : locs) {
      trace.push_back(aux.interpolateE(l * dx));
      trace.push_back(aux.interpolateH(l * dx));
    }
  }
  // whole state read back through the interpolators at node positions
  for (int c = 0; c + 1 < L; c++) {
    double loc = (static_cast<double>(c) - L / 2.0) * dx;
    eall.push_back(aux.interpolateE(loc));
    hall.push_back(aux.interpolateH(loc + 0.5 * dx));
  }
  fx.pi("ncells", ncells);
  fx.pi("L", L);
  fx.pi("nsteps", nsteps);
  fx.p("dx", dx);
  fx.p("dt", dt);
  fx.p("f0", f0);
  fx.p("tw", tw);
  fx.p("t0", t0);
  fx.arr("signal_t", st, {st.size()});
  fx.arr("signal_v", sv, {sv.size()});
  fx.arr("locs", locs, {locs.size()});
  fx.arr("trace_EH", trace, {size_t(nsteps), locs.size(), 2});
  fx.arr("final_E_nodes", eall, {eall.size()});
  fx.arr("final_H_nodes", hall, {hall.size()});
  fx.write(dir);
}

// ------------------------------------------------------------ scenario: 2D TE
// YeeFields<TE> (Ex, Ey, Hz; include/YeeFields.hpp:94-119) stepped with the
// reference's functors in the composition of tests/testTE.cpp: D curl, D
// sources, CPML E along x and y, TFSF updateD, material; B curl, B source
// (tests/testTE.cpp adds its pulse to Bz), CPML H, TFSF updateB, vacuum H.
// The TFSF planes take the reference's TE polarisations (Source.hpp:573-576:
// polH = z, pol = (k_y, -k_x, 0)).
struct TEScenario {
  std::string name;
  int nx, ny, nsteps;
  double dx, dt;
  PmlSpec pml;
  std::vector<Mat> mats;
  std::vector<int> matid;  // nx*ny
  struct Src {
    int field, comp, i, j;  // field 0: D (comp 0 / 1), 2: B (comp 2)
    std::vector<double> table;
  };
  std::vector<Src> src;
  std::vector<std::pair<int, int>> probes;
  bool tfsf = false;
  int bx0 = 0, bx1 = 0, by0 = 0, by1 = 0;
  double kx = 1, ky = 0;
  int aux_ncells = 0;
  double f0 = 0, tw = 0, t0 = 0;
};

static void run_te(const TEScenario& S, const std::string& dir) {
  const int nx = S.nx, ny = S.ny;
  std::vector<CellTE> cells(size_t(nx) * ny);
  auto at = [&](int i, int j) -> CellTE& { return cells[size_t(i) + size_t(nx) * j]; };
  for (int j = 1; j < ny - 1; j++)
    for (int i = 1; i < nx - 1; i++) {
      CellTE& c = at(i, j);
      c.mn[0] = &at(i - 1, j);
      c.mx[0] = &at(i + 1, j);
      c.mn[1] = &at(i, j - 1);
      c.mx[1] = &at(i, j + 1);
    }
  const double dt = S.dt, dx = S.dx;
  place_pml(nx, S.pml, [&](int idx, bool e, double K, double Sg, double A) {
    for (int j = 0; j < ny; j++) {
      if (e)
        at(idx, j).setPMLParameters<FieldType::Electric, Dir::X>(K, Sg, A, dt);
      else
        at(idx, j).setPMLParameters<FieldType::Magnetic, Dir::X>(K, Sg, A, dt);
    }
  });
  place_pml(ny, S.pml, [&](int idx, bool e, double K, double Sg, double A) {
    for (int i = 0; i < nx; i++) {
      if (e)
        at(i, idx).setPMLParameters<FieldType::Electric, Dir::Y>(K, Sg, A, dt);
      else
        at(i, idx).setPMLParameters<FieldType::Magnetic, Dir::Y>(K, Sg, A, dt);
    }
  });
  typedef TFSFPlane<TE, LocTE, PlaneIt<CellTE>, Aux1D<double>> PlaneT;
  std::vector<std::vector<PlaneEntry<CellTE>>> faces(S.tfsf ? 4 : 0);
  std::vector<PlaneT> planes;
  std::vector<double> aux_sig;
  const double cx = std::floor(0.5 * (S.bx0 + S.bx1)) + 0.5, cy = std::floor(0.5 * (S.by0 + S.by1)) + 0.5;
  if (S.tfsf) {
    // TE fluxes (Dx, Dy, Bz) carry no offset along the tangential axis of a
    // face, so every face spans the cells lo .. hi-1 of that axis (TM's Dz,
    // Bx, By span lo+1 .. hi)
    for (int j = S.by0; j < S.by1; j++) {
      faces[0].push_back({&at(S.bx0, j), (S.bx0 - cx) * dx, (j - cy) * dx, 0});
      faces[1].push_back({&at(S.bx1, j), (S.bx1 - cx) * dx, (j - cy) * dx, 0});
    }
    for (int i = S.bx0; i < S.bx1; i++) {
      faces[2].push_back({&at(i, S.by0), (i - cx) * dx, (S.by0 - cy) * dx, 0});
      faces[3].push_back({&at(i, S.by1), (i - cx) * dx, (S.by1 - cy) * dx, 0});
    }
    const double f0 = S.f0, tw = S.tw, t0 = S.t0;
    for (int fidx = 0; fidx < 4; fidx++) {
      auto sig = [f0, tw, t0, fidx, &aux_sig](double t) {
        double a = (t - t0) / tw;
        double v = std::exp(-a * a) * std::sin(2 * pi * f0 * (t - t0));
        if (fidx == 0) aux_sig.push_back(v);
        return v;
      };
      auto aux = std::make_shared<Aux1D<double>>(S.aux_ncells, dx, dt, sig);
      std::array<double, 2> K = {S.kx, S.ky};
      PlaneIt<CellTE> b{faces[fidx].data()}, e{faces[fidx].data() + faces[fidx].size()};
      if (fidx == 0) planes.push_back(make_tfsf_plane<TE, Dir::X, Orientation::MIN>(PlaneIt<CellTE>(b), PlaneIt<CellTE>(e), LocTE(), K, aux));
      if (fidx == 1) planes.push_back(make_tfsf_plane<TE, Dir::X, Orientation::MAX>(PlaneIt<CellTE>(b), PlaneIt<CellTE>(e), LocTE(), K, aux));
      if (fidx == 2) planes.push_back(make_tfsf_plane<TE, Dir::Y, Orientation::MIN>(PlaneIt<CellTE>(b), PlaneIt<CellTE>(e), LocTE(), K, aux));
      if (fidx == 3) planes.push_back(make_tfsf_plane<TE, Dir::Y, Orientation::MAX>(PlaneIt<CellTE>(b), PlaneIt<CellTE>(e), LocTE(), K, aux));
    }
  }
  std::vector<double> ptrace;
  for (int n = 0; n < S.nsteps; n++) {
    for (int j = 1; j < ny - 1; j++)
      for (int i = 1; i < nx - 1; i++) YeeUpdate<TE, Dx>(dt, dx)(at(i, j));
    for (int j = 1; j < ny - 1; j++)
      for (int i = 1; i < nx - 1; i++) YeeUpdate<TE, Dy>(dt, dx)(at(i, j));
    for (const auto& q : S.src)
      if (q.field == 0) (q.comp == 0 ? at(q.i, q.j).Dx() : at(q.i, q.j).Dy()) += q.table[n];
    for (int j = 1; j < ny - 1; j++)
      for (int i = 1; i < nx - 1; i++) UpdatePML<TE, FieldType::Electric, Dir::X>(dt, dx)(at(i, j));
    for (int j = 1; j < ny - 1; j++)
      for (int i = 1; i < nx - 1; i++) UpdatePML<TE, FieldType::Electric, Dir::Y>(dt, dx)(at(i, j));
    for (auto& pl : planes) pl.updateD(dt, dx);
    for (int j = 1; j < ny - 1; j++)
      for (int i = 1; i < nx - 1; i++) material_E<TE>(at(i, j), S.mats[S.matid[size_t(i) + size_t(nx) * j]], dt);
    for (int j = 1; j < ny - 1; j++)
      for (int i = 1; i < nx - 1; i++) YeeUpdate<TE, Bz>(dt, dx)(at(i, j));
    for (const auto& q : S.src)
      if (q.field == 2) at(q.i, q.j).Bz() += q.table[n];
    for (int j = 1; j < ny - 1; j++)
      for (int i = 1; i < nx - 1; i++) UpdatePML<TE, FieldType::Magnetic, Dir::X>(dt, dx)(at(i, j));
    for (int j = 1; j < ny - 1; j++)
      for (int i = 1; i < nx - 1; i++) UpdatePML<TE, FieldType::Magnetic, Dir::Y>(dt, dx)(at(i, j));
    for (auto& pl : planes) pl.updateB(dt, dx);
    for (int j = 1; j < ny - 1; j++)
      for (int i = 1; i < nx - 1; i++) VacuumUpdate<TE, FieldType::Magnetic>()(at(i, j));
    for (auto& pr : S.probes) {
      ptrace.push_back(at(pr.first, pr.second).Ex());
      ptrace.push_back(at(pr.first, pr.second).Ey());
      ptrace.push_back(at(pr.first, pr.second).Hz());
    }
  }
  Fixture fx;
  fx.name = S.name;
  fx.ps("mode", "TE");
  fx.pi("nx", nx);
  fx.pi("ny", ny);
  fx.pi("nsteps", S.nsteps);
  fx.p("dx", dx);
  fx.p("dt", dt);
  put_pml(fx, S.pml);
  fx.pi("pmlx", 1);
  fx.pi("pmly", 1);
  fx.pi("tfsf", S.tfsf);
  if (S.tfsf) {
    fx.pi("bx0", S.bx0);
    fx.pi("bx1", S.bx1);
    fx.pi("by0", S.by0);
    fx.pi("by1", S.by1);
    fx.p("kx", S.kx);
    fx.p("ky", S.ky);
    fx.pi("aux_ncells", S.aux_ncells);
    fx.p("cx", cx);
    fx.p("cy", cy);
    fx.p("f0", S.f0);
    fx.p("tw", S.tw);
    fx.p("t0", S.t0);
    fx.arr("aux_signal", aux_sig, {aux_sig.size()});
  }
  if (!S.src.empty()) {
    std::vector<double> sp, st;
    for (const auto& q : S.src) {
      sp.insert(sp.end(), {double(q.field), double(q.comp), double(q.i), double(q.j)});
      st.insert(st.end(), q.table.begin(), q.table.end());
    }
    fx.arr("src_field_comp_i_j", sp, {S.src.size(), 4});
    fx.arr("src_tables", st, {S.src.size(), size_t(S.nsteps)});
  }
  std::vector<double> mt;
  for (auto& m : S.mats) mt.insert(mt.end(), {double(m.kind), m.eps, m.freq, m.gamma, m.delta});
  fx.arr("mats_kind_eps_freq_gamma_delta", mt, {S.mats.size(), 5});
  std::vector<double> mid(S.matid.begin(), S.matid.end());
  fx.arr("matid", mid, {size_t(ny), size_t(nx)});
  std::vector<double> pr;
  for (auto& p : S.probes) pr.insert(pr.end(), {double(p.first), double(p.second)});
  fx.arr("probes", pr, {S.probes.size(), 2});
  fx.arr("probe_trace_Ex_Ey_Hz", ptrace, {size_t(S.nsteps), S.probes.size(), 3});
  std::vector<double> Ex, Ey, Dxv, Dyv, Hz, Bzv;
  for (auto& c : cells) {
    Ex.push_back(c.Ex());
    Ey.push_back(c.Ey());
    Dxv.push_back(c.Dx());
    Dyv.push_back(c.Dy());
    Hz.push_back(c.Hz());
    Bzv.push_back(c.Bz());
  }
  std::vector<size_t> sh = {size_t(ny), size_t(nx)};
  fx.arr("Ex", Ex, sh);
  fx.arr("Ey", Ey, sh);
  fx.arr("Dx", Dxv, sh);
  fx.arr("Dy", Dyv, sh);
  fx.arr("Hz", Hz, sh);
  fx.arr("Bz", Bzv, sh);
  fx.write(dir);
}

// ------------------------------------------------------------ scenario: 2D TM
struct TMScenario {
  std::string name;
  int nx, ny, nsteps;
  double dx, dt;
  PmlSpec pml;
  bool pmlx, pmly;
  std::vector<Mat> mats;
  std::vector<int> matid;  // nx*ny
  std::vector<double> ne, nu;  // fluid density / collision rate per cell (empty: none)
  int src_i, src_j;
  std::vector<double> src;  // per step, added to Dz before the PML/material sweep
  std::vector<std::pair<int, int>> probes;
  bool tfsf = false;
  int bx0, bx1, by0, by1;
  double kx, ky;
  int aux_ncells;
  double f0, tw, t0;
  bool save_full = true;
  // soft sources on the derived fields, applied after the material step
  // through the reference's SoftSource (include/Source.hpp:183-261):
  // field 1 = E (comp 2: Ez), 3 = H (comp 0 / 1: Hx / Hy); table per step
  struct Soft {
    int field, comp, i, j;
    std::vector<double> table;
  };
  std::vector<Soft> soft;
  // single TFSF planes (make_tfsf_plane, Source.hpp:624-634), each with its
  // own Aux1D driven by the Gaussian-modulated sinusoid (f0, tw, t0): normal
  // along axis (0 x, 1 y), side 0 MIN / 1 MAX, on cell layer `layer`, cells
  // lo..hi (inclusive) along the other axis; kx, ky, cx, cy, aux_ncells, f0
  // and tw as for the box
  struct Plane {
    int axis, side, layer, lo, hi;
    double t0;
  };
  std::vector<Plane> planes1;
  double cx1 = 0, cy1 = 0;  // the planes' location origin (cell units) when there is no box
};

static void run_tm(const TMScenario& S, const std::string& dir) {
  const int nx = S.nx, ny = S.ny;
  std::vector<CellTM> cells(size_t(nx) * ny);
  auto at = [&](int i, int j) -> CellTM& { return cells[size_t(i) + size_t(nx) * j]; };
  for (size_t q = 0; q < S.ne.size(); q++) {
    cells[q].ne() = S.ne[q];
    cells[q].nu_m() = S.nu[q];
  }
  for (int j = 1; j < ny - 1; j++)
    for (int i = 1; i < nx - 1; i++) {
      CellTM& c = at(i, j);
      c.mn[0] = &at(i - 1, j);
      c.mx[0] = &at(i + 1, j);
      c.mn[1] = &at(i, j - 1);
      c.mx[1] = &at(i, j + 1);
    }
  const double dt = S.dt, dx = S.dx;
  if (S.pmlx)
    place_pml(nx, S.pml, [&](int idx, bool e, double K, double Sg, double A) {
      for (int j = 0; j < ny; j++) {
        if (e)
          at(idx, j).setPMLParameters<FieldType::Electric, Dir::X>(K, Sg, A, dt);
        else
          at(idx, j).setPMLParameters<FieldType::Magnetic, Dir::X>(K, Sg, A, dt);
      }
    });
  if (S.pmly)
    place_pml(ny, S.pml, [&](int idx, bool e, double K, double Sg, double A) {
      for (int i = 0; i < nx; i++) {
        if (e)
          at(i, idx).setPMLParameters<FieldType::Electric, Dir::Y>(K, Sg, A, dt);
        else
          at(i, idx).setPMLParameters<FieldType::Magnetic, Dir::Y>(K, Sg, A, dt);
      }
    });

  // TFSF: one plane per face of the box, each with its own Aux1D
  // (include/Source.hpp:595-609 steps the aux inside updateD), then the
  // single planes, each with its own Aux1D too.
  typedef TFSFPlane<TM, LocTM, PlaneIt<CellTM>, Aux1D<double>> PlaneT;
  std::vector<std::vector<PlaneEntry<CellTM>>> faces;
  std::vector<PlaneT> planes;
  std::vector<double> aux_sig;
  std::vector<std::vector<double>> plane_sig(S.planes1.size());
  const double cx = S.tfsf ? std::floor(0.5 * (S.bx0 + S.bx1)) + 0.5 : S.cx1,
               cy = S.tfsf ? std::floor(0.5 * (S.by0 + S.by1)) + 0.5 : S.cy1;
  faces.resize((S.tfsf ? 4 : 0) + S.planes1.size());
  for (size_t p = 0; p < S.planes1.size(); p++) {
    const auto& P = S.planes1[p];
    auto& F = faces[(S.tfsf ? 4 : 0) + p];
    for (int q = P.lo; q <= P.hi; q++) {
      const int i = P.axis == 0 ? P.layer : q, j = P.axis == 0 ? q : P.layer;
      F.push_back({&at(i, j), (i - cx) * dx, (j - cy) * dx, 0});
    }
  }
  if (S.tfsf) {
    for (int j = S.by0 + 1; j <= S.by1; j++) {
      faces[0].push_back({&at(S.bx0, j), (S.bx0 - cx) * dx, (j - cy) * dx, 0});
      faces[1].push_back({&at(S.bx1, j), (S.bx1 - cx) * dx, (j - cy) * dx, 0});
    }
    for (int i = S.bx0 + 1; i <= S.bx1; i++) {
      faces[2].push_back({&at(i, S.by0), (i - cx) * dx, (S.by0 - cy) * dx, 0});
      faces[3].push_back({&at(i, S.by1), (i - cx) * dx, (S.by1 - cy) * dx, 0});
    }
    const double f0 = S.f0, tw = S.tw, t0 = S.t0;
    for (int fidx = 0; fidx < 4; fidx++) {
      auto sig = [f0, tw, t0, fidx, &aux_sig](double t) {
        double a = (t - t0) / tw;
        double v = std::exp(-a * a) * std::sin(2 * pi * f0 * (t - t0));
        if (fidx == 0) aux_sig.push_back(v);
        return v;
      };
      auto aux = std::make_shared<Aux1D<double>>(S.aux_ncells, dx, dt, sig);
      std::array<double, 2> K = {S.kx, S.ky};
      PlaneIt<CellTM> b{faces[fidx].data()}, e{faces[fidx].data() + faces[fidx].size()};
      if (fidx == 0) planes.push_back(make_tfsf_plane<TM, Dir::X, Orientation::MIN>(PlaneIt<CellTM>(b), PlaneIt<CellTM>(e), LocTM(), K, aux));
      if (fidx == 1) planes.push_back(make_tfsf_plane<TM, Dir::X, Orientation::MAX>(PlaneIt<CellTM>(b), PlaneIt<CellTM>(e), LocTM(), K, aux));
      if (fidx == 2) planes.push_back(make_tfsf_plane<TM, Dir::Y, Orientation::MIN>(PlaneIt<CellTM>(b), PlaneIt<CellTM>(e), LocTM(), K, aux));
      if (fidx == 3) planes.push_back(make_tfsf_plane<TM, Dir::Y, Orientation::MAX>(PlaneIt<CellTM>(b), PlaneIt<CellTM>(e), LocTM(), K, aux));
    }
  }
  for (size_t p = 0; p < S.planes1.size(); p++) {
    const auto& P = S.planes1[p];
    const double f0 = S.f0, tw = S.tw, t0 = P.t0;
    std::vector<double>* rec = &plane_sig[p];
    auto sig = [f0, tw, t0, rec](double t) {
      double a = (t - t0) / tw;
      double v = std::exp(-a * a) * std::sin(2 * pi * f0 * (t - t0));
      rec->push_back(v);
      return v;
    };
    auto aux = std::make_shared<Aux1D<double>>(S.aux_ncells, dx, dt, sig);
    std::array<double, 2> K = {S.kx, S.ky};
    auto& F = faces[(S.tfsf ? 4 : 0) + p];
    PlaneIt<CellTM> b{F.data()}, e{F.data() + F.size()};
    const int which = 2 * P.axis + P.side;
    if (which == 0) planes.push_back(make_tfsf_plane<TM, Dir::X, Orientation::MIN>(PlaneIt<CellTM>(b), PlaneIt<CellTM>(e), LocTM(), K, aux));
    if (which == 1) planes.push_back(make_tfsf_plane<TM, Dir::X, Orientation::MAX>(PlaneIt<CellTM>(b), PlaneIt<CellTM>(e), LocTM(), K, aux));
    if (which == 2) planes.push_back(make_tfsf_plane<TM, Dir::Y, Orientation::MIN>(PlaneIt<CellTM>(b), PlaneIt<CellTM>(e), LocTM(), K, aux));
    if (which == 3) planes.push_back(make_tfsf_plane<TM, Dir::Y, Orientation::MAX>(PlaneIt<CellTM>(b), PlaneIt<CellTM>(e), LocTM(), K, aux));
  }
  // the soft sources on E and H: the reference's SoftSource objects, sampled
  // at t = step index
  std::vector<std::function<void(double)>> softE, softH;
  for (const auto& q : S.soft) {
    auto tab = std::make_shared<std::vector<double>>(q.table);
    auto sig = [tab](double t) {
      const long n = std::lround(t);
      return n < long(tab->size()) ? (*tab)[n] : 0.0;
    };
    CellTM& c = at(q.i, q.j);
    if (q.field == 1) {
      auto ss = make_soft_source<Ez>(sig, c);
      softE.push_back([ss](double t) mutable { ss.apply(t); });
    } else if (q.comp == 0) {
      auto ss = make_soft_source<Hx>(sig, c);
      softH.push_back([ss](double t) mutable { ss.apply(t); });
    } else {
      auto ss = make_soft_source<Hy>(sig, c);
      softH.push_back([ss](double t) mutable { ss.apply(t); });
    }
  }

  std::vector<double> ptrace;
  for (int n = 0; n < S.nsteps; n++) {
    for (int j = 1; j < ny - 1; j++)
      for (int i = 1; i < nx - 1; i++) YeeUpdate<TM, Dz>(dt, dx)(at(i, j));
    if (S.src_i >= 0) at(S.src_i, S.src_j).Dz() += S.src[n];
    for (int j = 1; j < ny - 1; j++)
      for (int i = 1; i < nx - 1; i++) UpdatePML<TM, FieldType::Electric, Dir::X>(dt, dx)(at(i, j));
    for (int j = 1; j < ny - 1; j++)
      for (int i = 1; i < nx - 1; i++) UpdatePML<TM, FieldType::Electric, Dir::Y>(dt, dx)(at(i, j));
    for (auto& pl : planes) pl.updateD(dt, dx);
    for (int j = 1; j < ny - 1; j++)
      for (int i = 1; i < nx - 1; i++) material_E<TM>(at(i, j), S.mats[S.matid[size_t(i) + size_t(nx) * j]], dt);
    for (auto& f : softE) f(double(n));
    for (int j = 1; j < ny - 1; j++)
      for (int i = 1; i < nx - 1; i++) YeeUpdate<TM, Bx>(dt, dx)(at(i, j));
    for (int j = 1; j < ny - 1; j++)
      for (int i = 1; i < nx - 1; i++) YeeUpdate<TM, By>(dt, dx)(at(i, j));
    for (int j = 1; j < ny - 1; j++)
      for (int i = 1; i < nx - 1; i++) UpdatePML<TM, FieldType::Magnetic, Dir::X>(dt, dx)(at(i, j));
    for (int j = 1; j < ny - 1; j++)
      for (int i = 1; i < nx - 1; i++) UpdatePML<TM, FieldType::Magnetic, Dir::Y>(dt, dx)(at(i, j));
    for (auto& pl : planes) pl.updateB(dt, dx);
    for (int j = 1; j < ny - 1; j++)
      for (int i = 1; i < nx - 1; i++) VacuumUpdate<TM, FieldType::Magnetic>()(at(i, j));
    for (auto& f : softH) f(double(n));
    for (auto& pr : S.probes) {
      ptrace.push_back(at(pr.first, pr.second).Ez());
      ptrace.push_back(at(pr.first, pr.second).Hx());
      ptrace.push_back(at(pr.first, pr.second).Hy());
    }
  }

  Fixture fx;
  fx.name = S.name;
  fx.ps("mode", "TM");
  fx.pi("nx", nx);
  fx.pi("ny", ny);
  fx.pi("nsteps", S.nsteps);
  fx.p("dx", dx);
  fx.p("dt", dt);
  put_pml(fx, S.pml);
  fx.pi("pmlx", S.pmlx);
  fx.pi("pmly", S.pmly);
  fx.pi("src_i", S.src_i);
  fx.pi("src_j", S.src_j);
  fx.pi("tfsf", S.tfsf);
  if (S.tfsf) {
    fx.pi("bx0", S.bx0);
    fx.pi("bx1", S.bx1);
    fx.pi("by0", S.by0);
    fx.pi("by1", S.by1);
    fx.p("kx", S.kx);
    fx.p("ky", S.ky);
    fx.pi("aux_ncells", S.aux_ncells);
    fx.p("cx", cx);
    fx.p("cy", cy);
    fx.p("f0", S.f0);
    fx.p("tw", S.tw);
    fx.p("t0", S.t0);
    fx.arr("aux_signal", aux_sig, {aux_sig.size()});
  }
  if (!S.planes1.empty()) {
    fx.p("kx", S.kx);
    fx.p("ky", S.ky);
    fx.pi("aux_ncells", S.aux_ncells);
    fx.p("cx", cx);
    fx.p("cy", cy);
    fx.p("f0", S.f0);
    fx.p("tw", S.tw);
    std::vector<double> pl;
    for (const auto& P : S.planes1) pl.insert(pl.end(), {double(P.axis), double(P.side), double(P.layer), double(P.lo), double(P.hi), P.t0});
    fx.arr("planes_axis_side_layer_lo_hi_t0", pl, {S.planes1.size(), 6});
    for (size_t p = 0; p < S.planes1.size(); p++) fx.arr("plane" + std::to_string(p) + "_signal", plane_sig[p], {plane_sig[p].size()});
  }
  if (!S.soft.empty()) {
    std::vector<double> sp, st;
    for (const auto& q : S.soft) {
      sp.insert(sp.end(), {double(q.field), double(q.comp), double(q.i), double(q.j)});
      st.insert(st.end(), q.table.begin(), q.table.end());
    }
    fx.arr("soft_field_comp_i_j", sp, {S.soft.size(), 4});
    fx.arr("soft_tables", st, {S.soft.size(), size_t(S.nsteps)});
  }
  std::vector<double> mt;
  for (auto& m : S.mats) {
    mt.push_back(m.kind);
    mt.push_back(m.eps);
    mt.push_back(m.freq);
    mt.push_back(m.gamma);
    mt.push_back(m.delta);
  }
  fx.arr("mats_kind_eps_freq_gamma_delta", mt, {S.mats.size(), 5});
  {
    bool any_bias = false;
    std::vector<double> bias;
    for (auto& m : S.mats) {
      any_bias = any_bias || m.kind >= 5;
      bias.insert(bias.end(), {m.bx, m.by, m.bz});
    }
    if (any_bias) fx.arr("mats_bias", bias, {S.mats.size(), 3});
  }
  if (!S.ne.empty()) {
    fx.arr("fluid_ne", S.ne, {S.ne.size()});
    fx.arr("fluid_nu", S.nu, {S.nu.size()});
  }
  std::vector<double> mid(S.matid.begin(), S.matid.end());
  if (S.save_full) fx.arr("matid", mid, {size_t(ny), size_t(nx)});
  if (S.src_i >= 0) fx.arr("src", S.src, {S.src.size()});
  std::vector<double> pr;
  for (auto& p : S.probes) {
    pr.push_back(p.first);
    pr.push_back(p.second);
  }
  if (!S.probes.empty()) {
    fx.arr("probes", pr, {S.probes.size(), 2});
    fx.arr("probe_trace_Ez_Hx_Hy", ptrace, {size_t(S.nsteps), S.probes.size(), 3});
  }
  std::vector<double> Ez, Dz, Hx, Hy, Bx, By;
  for (auto& c : cells) {
    Ez.push_back(c.Ez());
    Dz.push_back(c.Dz());
    Hx.push_back(c.Hx());
    Hy.push_back(c.Hy());
    Bx.push_back(c.Bx());
    By.push_back(c.By());
  }
  std::vector<size_t> sh = {size_t(ny), size_t(nx)};
  if (S.save_full) {
    fx.arr("Ez", Ez, sh);
    fx.arr("Dz", Dz, sh);
    fx.arr("Hx", Hx, sh);
    fx.arr("Hy", Hy, sh);
    fx.arr("Bx", Bx, sh);
    fx.arr("By", By, sh);
  } else {
    // row and column through the centre plus per-field sums of squares
    std::vector<double> row, col, ss(6, 0.0);
    for (int i = 0; i < nx; i++) row.push_back(at(i, ny / 2).Ez());
    for (int j = 0; j < ny; j++) col.push_back(at(nx / 2, j).Ez());
    for (size_t q = 0; q < cells.size(); q++) {
      ss[0] += Ez[q] * Ez[q];
      ss[1] += Dz[q] * Dz[q];
      ss[2] += Hx[q] * Hx[q];
      ss[3] += Hy[q] * Hy[q];
      ss[4] += Bx[q] * Bx[q];
      ss[5] += By[q] * By[q];
    }
    fx.arr("Ez_row_mid", row, {row.size()});
    fx.arr("Ez_col_mid", col, {col.size()});
    fx.arr("sumsq_Ez_Dz_Hx_Hy_Bx_By", ss, {6});
  }
  fx.write(dir);
}

// ------------------------------------------------------------ scenario: 3-D
struct Scenario3 {
  std::string name;
  int nx, ny, nz, nsteps;
  double dx, dt;
  PmlSpec pml;
  std::vector<Mat> mats;
  std::vector<int> matid;  // nx*ny*nz
  std::vector<double> ne, nu;  // fluid density / collision rate per cell (empty: none)
  int si, sj, sk;
  std::vector<double> src;
  std::vector<std::array<int, 3>> probes;
};

static void run_3d(const Scenario3& S, const std::string& dir, bool timing_only = false, int warm = 0,
                   double* seconds = nullptr) {
  const int nx = S.nx, ny = S.ny, nz = S.nz;
  std::vector<Cell3> cells(size_t(nx) * ny * nz);
  auto id = [&](int i, int j, int k) { return size_t(i) + size_t(nx) * (size_t(j) + size_t(ny) * k); };
  for (size_t q = 0; q < S.ne.size(); q++) {
    cells[q].ne() = S.ne[q];
    cells[q].nu_m() = S.nu[q];
  }
  auto at = [&](int i, int j, int k) -> Cell3& { return cells[id(i, j, k)]; };
  for (int k = 1; k < nz - 1; k++)
    for (int j = 1; j < ny - 1; j++)
      for (int i = 1; i < nx - 1; i++) {
        Cell3& c = at(i, j, k);
        c.mn[0] = &at(i - 1, j, k);
        c.mx[0] = &at(i + 1, j, k);
        c.mn[1] = &at(i, j - 1, k);
        c.mx[1] = &at(i, j + 1, k);
        c.mn[2] = &at(i, j, k - 1);
        c.mx[2] = &at(i, j, k + 1);
      }
  const double dt = S.dt, dx = S.dx;
  place_pml(nx, S.pml, [&](int idx, bool e, double K, double Sg, double A) {
    for (int k = 0; k < nz; k++)
      for (int j = 0; j < ny; j++) {
        if (e)
          at(idx, j, k).setPMLParameters<FieldType::Electric, Dir::X>(K, Sg, A, dt);
        else
          at(idx, j, k).setPMLParameters<FieldType::Magnetic, Dir::X>(K, Sg, A, dt);
      }
  });
  place_pml(ny, S.pml, [&](int idx, bool e, double K, double Sg, double A) {
    for (int k = 0; k < nz; k++)
      for (int i = 0; i < nx; i++) {
        if (e)
          at(i, idx, k).setPMLParameters<FieldType::Electric, Dir::Y>(K, Sg, A, dt);
        else
          at(i, idx, k).setPMLParameters<FieldType::Magnetic, Dir::Y>(K, Sg, A, dt);
      }
  });
  place_pml(nz, S.pml, [&](int idx, bool e, double K, double Sg, double A) {
    for (int j = 0; j < ny; j++)
      for (int i = 0; i < nx; i++) {
        if (e)
          at(i, j, idx).setPMLParameters<FieldType::Electric, Dir::Z>(K, Sg, A, dt);
        else
          at(i, j, idx).setPMLParameters<FieldType::Magnetic, Dir::Z>(K, Sg, A, dt);
      }
  });
  std::vector<double> ptrace;
  auto interior = [&](auto&& fn) {
    for (int k = 1; k < nz - 1; k++)
      for (int j = 1; j < ny - 1; j++)
        for (int i = 1; i < nx - 1; i++) fn(i, j, k);
  };
  auto t_start = std::chrono::steady_clock::now();
  for (int n = 0; n < S.nsteps; n++) {
    if (timing_only && n == warm) t_start = std::chrono::steady_clock::now();
    interior([&](int i, int j, int k) { YeeUpdate<ThreeD, Dx>(dt, dx)(at(i, j, k)); });
    interior([&](int i, int j, int k) { YeeUpdate<ThreeD, Dy>(dt, dx)(at(i, j, k)); });
    interior([&](int i, int j, int k) { YeeUpdate<ThreeD, Dz>(dt, dx)(at(i, j, k)); });
    if (S.si >= 0) at(S.si, S.sj, S.sk).Dz() += S.src[n];
    interior([&](int i, int j, int k) { UpdatePML<ThreeD, FieldType::Electric, Dir::X>(dt, dx)(at(i, j, k)); });
    interior([&](int i, int j, int k) { UpdatePML<ThreeD, FieldType::Electric, Dir::Y>(dt, dx)(at(i, j, k)); });
    interior([&](int i, int j, int k) { UpdatePML<ThreeD, FieldType::Electric, Dir::Z>(dt, dx)(at(i, j, k)); });
    interior([&](int i, int j, int k) { material_E<ThreeD>(at(i, j, k), S.mats[S.matid[id(i, j, k)]], dt); });
    interior([&](int i, int j, int k) { YeeUpdate<ThreeD, Bx>(dt, dx)(at(i, j, k)); });
    interior([&](int i, int j, int k) { YeeUpdate<ThreeD, By>(dt, dx)(at(i, j, k)); });
    interior([&](int i, int j, int k) { YeeUpdate<ThreeD, Bz>(dt, dx)(at(i, j, k)); });
    interior([&](int i, int j, int k) { UpdatePML<ThreeD, FieldType::Magnetic, Dir::X>(dt, dx)(at(i, j, k)); });
    interior([&](int i, int j, int k) { UpdatePML<ThreeD, FieldType::Magnetic, Dir::Y>(dt, dx)(at(i, j, k)); });
    interior([&](int i, int j, int k) { UpdatePML<ThreeD, FieldType::Magnetic, Dir::Z>(dt, dx)(at(i, j, k)); });
    interior([&](int i, int j, int k) { VacuumUpdate<ThreeD, FieldType::Magnetic>()(at(i, j, k)); });
    for (auto& p : S.probes) {
      Cell3& c = at(p[0], p[1], p[2]);
      ptrace.push_back(c.Ex());
      ptrace.push_back(c.Ey());
      ptrace.push_back(c.Ez());
      ptrace.push_back(c.Hx());
      ptrace.push_back(c.Hy());
      ptrace.push_back(c.Hz());
    }
  }
  if (timing_only) {
    *seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t_start).count();
    return;
  }
  Fixture fx;
  fx.name = S.name;
  fx.ps("mode", "ThreeD");
  fx.pi("nx", nx);
  fx.pi("ny", ny);
  fx.pi("nz", nz);
  fx.pi("nsteps", S.nsteps);
  fx.p("dx", dx);
  fx.p("dt", dt);
  put_pml(fx, S.pml);
  fx.pi("src_i", S.si);
  fx.pi("src_j", S.sj);
  fx.pi("src_k", S.sk);
  std::vector<double> mt;
  for (auto& m : S.mats) {
    mt.push_back(m.kind);
    mt.push_back(m.eps);
    mt.push_back(m.freq);
    mt.push_back(m.gamma);
    mt.push_back(m.delta);
  }
  fx.arr("mats_kind_eps_freq_gamma_delta", mt, {S.mats.size(), 5});
  if (!S.ne.empty()) {
    fx.arr("fluid_ne", S.ne, {S.ne.size()});
    fx.arr("fluid_nu", S.nu, {S.nu.size()});
  }
  bool any_bias = false;
  std::vector<double> bias;
  for (auto& m : S.mats) {
    any_bias = any_bias || m.kind >= 5;
    bias.insert(bias.end(), {m.bx, m.by, m.bz});
  }
  if (any_bias) fx.arr("mats_bias", bias, {S.mats.size(), 3});
  std::vector<double> mid(S.matid.begin(), S.matid.end());
  std::vector<size_t> sh = {size_t(nz), size_t(ny), size_t(nx)};
  fx.arr("matid", mid, sh);
  fx.arr("src", S.src, {S.src.size()});
  std::vector<double> pr;
  for (auto& p : S.probes) pr.insert(pr.end(), {double(p[0]), double(p[1]), double(p[2])});
  fx.arr("probes", pr, {S.probes.size(), 3});
  fx.arr("probe_trace_E_H", ptrace, {size_t(S.nsteps), S.probes.size(), 6});
  const char* names[6] = {"Ex", "Ey", "Ez", "Hx", "Hy", "Hz"};
  for (int f = 0; f < 6; f++) {
    std::vector<double> v;
    v.reserve(cells.size());
    for (auto& c : cells) {
      Cell3& cc = const_cast<Cell3&>(c);
      double x = f == 0 ? cc.Ex() : f == 1 ? cc.Ey() : f == 2 ? cc.Ez() : f == 3 ? cc.Hx() : f == 4 ? cc.Hy() : cc.Hz();
      v.push_back(x);
    }
    fx.arr(names[f], v, sh);
  }
  // dispersive state where the material is dispersive: D, P, J per cell, flattened in cell order
  std::vector<double> disp;
  for (size_t q = 0; q < cells.size(); q++) {
    if (S.mats[S.matid[q]].kind >= 2) {
      Cell3& c = cells[q];
      disp.insert(disp.end(), {double(q), c.Dx(), c.Dy(), c.Dz(), c.Px(), c.Py(), c.Pz(), c.Jx(), c.Jy(), c.Jz()});
    }
  }
  if (!disp.empty()) fx.arr("dispersive_q_D_P_J", disp, {disp.size() / 10, 10});
  fx.write(dir);
}

// seeded axis-aligned blocks (later blocks override earlier ones)
static void seed_blocks(std::vector<int>& matid, int nx, int ny, int nz, int nblocks, int lo, int hi, int side_lo,
                        int side_hi, SplitMix64& rng, std::vector<Mat>& mats, double eps_lo, double eps_hi,
                        std::vector<double>* boxes_out) {
  for (int b = 0; b < nblocks; b++) {
    int sx = rng.randint(side_lo, side_hi), sy = rng.randint(side_lo, side_hi), sz = rng.randint(side_lo, side_hi);
    int x0 = rng.randint(lo, std::max(lo, nx - hi - sx)), y0 = rng.randint(lo, std::max(lo, ny - hi - sy));
    int z0 = nz > 1 ? rng.randint(lo, std::max(lo, nz - hi - sz)) : 0;
    double eps = eps_lo + (eps_hi - eps_lo) * rng.uniform();
    mats.push_back({1, eps, 0, 0, 0});
    int m = int(mats.size()) - 1;
    int x1 = std::min(nx - hi, x0 + sx), y1 = std::min(ny - hi, y0 + sy), z1 = nz > 1 ? std::min(nz - hi, z0 + sz) : 1;
    for (int k = z0; k < z1; k++)
      for (int j = y0; j < y1; j++)
        for (int i = x0; i < x1; i++) matid[size_t(i) + size_t(nx) * (size_t(j) + size_t(ny) * k)] = m;
    if (boxes_out) boxes_out->insert(boxes_out->end(), {double(x0), double(x1), double(y0), double(y1), double(z0), double(z1), eps});
  }
}

static double gauss_pulse(int n, double t0, double spread) {
  double a = (n - t0) / spread;
  return std::exp(-a * a);
}

// ------------------------------------------------------------ scenario: magnetic-side poles
// 3-D grid, CPML on every face, material pairs whose magnetic side is a Drude,
// Lorentz or conductive model (the reference's functors on FieldType::Magnetic,
// stepping B, H, M, K), beside electric Drude / Lorentz / constant sides; a
// Gaussian on Dz and one on Bx.
static void scen_magnetic(const std::string& dir) {
  const int nx = 22, ny = 20, nz = 18, nsteps = 130;
  const double dx = 1e-6, dt = 0.5 * dx / (c0 * std::sqrt(3.0));
  PMLParameterModel dm(dx);
  const PmlSpec pml = {5, 3, 1, dm.msMax, 5, 0.05};
  std::vector<Mat> me = {{0, 1, 0, 0, 0}, {1, 2.0, 0, 0, 0}, {3, 1.5, 0.2 / dt, 0.02 / dt, 0.6}, {1, 1.0, 0, 0, 0},
                         {2, 1.2, 0.25 / dt, 0.03 / dt, 0}, {1, 3.0, 0, 0, 0}};
  std::vector<MatM> mh = {{0, 1, 0, 0, 0},           {2, 1.0, 0.3 / dt, 0.02 / dt, 0}, {3, 1.5, 0.15 / dt, 0.01 / dt, 0.5},
                          {4, 1.2, 0.04 * 1.2 * mu0 / dt, 0, 0}, {1, 2.0, 0, 0, 0},   {2, 2.5, 0.2 / dt, 0.05 / dt, 0}};
  std::vector<int> matid(size_t(nx) * ny * nz, 0);
  auto id = [&](int i, int j, int k) { return size_t(i) + size_t(nx) * (size_t(j) + size_t(ny) * k); };
  auto fill = [&](int x0, int x1, int y0, int y1, int z0, int z1, int m) {
    for (int k = z0; k < z1; k++)
      for (int j = y0; j < y1; j++)
        for (int i = x0; i < x1; i++) matid[id(i, j, k)] = m;
  };
  fill(7, 11, 6, 10, 6, 11, 1);
  fill(12, 16, 7, 12, 7, 10, 2);
  fill(8, 13, 11, 14, 9, 13, 3);
  fill(13, 16, 13, 16, 5, 9, 4);
  fill(5, 8, 11, 15, 11, 14, 5);
  std::vector<Cell3M> cells(size_t(nx) * ny * nz);
  auto at = [&](int i, int j, int k) -> Cell3M& { return cells[id(i, j, k)]; };
  for (int k = 1; k < nz - 1; k++)
    for (int j = 1; j < ny - 1; j++)
      for (int i = 1; i < nx - 1; i++) {
        Cell3M& c = at(i, j, k);
        c.mn[0] = &at(i - 1, j, k);
        c.mx[0] = &at(i + 1, j, k);
        c.mn[1] = &at(i, j - 1, k);
        c.mx[1] = &at(i, j + 1, k);
        c.mn[2] = &at(i, j, k - 1);
        c.mx[2] = &at(i, j, k + 1);
      }
  place_pml(nx, pml, [&](int idx, bool e, double K, double Sg, double A) {
    for (int k = 0; k < nz; k++)
      for (int j = 0; j < ny; j++) {
        if (e)
          at(idx, j, k).setPMLParameters<FieldType::Electric, Dir::X>(K, Sg, A, dt);
        else
          at(idx, j, k).setPMLParameters<FieldType::Magnetic, Dir::X>(K, Sg, A, dt);
      }
  });
  place_pml(ny, pml, [&](int idx, bool e, double K, double Sg, double A) {
    for (int k = 0; k < nz; k++)
      for (int i = 0; i < nx; i++) {
        if (e)
          at(i, idx, k).setPMLParameters<FieldType::Electric, Dir::Y>(K, Sg, A, dt);
        else
          at(i, idx, k).setPMLParameters<FieldType::Magnetic, Dir::Y>(K, Sg, A, dt);
      }
  });
  place_pml(nz, pml, [&](int idx, bool e, double K, double Sg, double A) {
    for (int j = 0; j < ny; j++)
      for (int i = 0; i < nx; i++) {
        if (e)
          at(i, j, idx).setPMLParameters<FieldType::Electric, Dir::Z>(K, Sg, A, dt);
        else
          at(i, j, idx).setPMLParameters<FieldType::Magnetic, Dir::Z>(K, Sg, A, dt);
      }
  });
  const int si = 10, sj = 9, sk = 8, bi = 14, bj = 10, bk = 8;
  std::vector<double> sD, sB;
  for (int n = 0; n < nsteps; n++) {
    sD.push_back(eps0 * gauss_pulse(n, 20, 6));
    sB.push_back(1e-9 * gauss_pulse(n, 30, 7));
  }
  const std::vector<std::array<int, 3>> probes = {{10, 9, 8}, {14, 10, 8}, {9, 8, 8}, {14, 14, 7}, {6, 13, 12}};
  std::vector<double> ptrace;
  auto interior = [&](auto&& fn) {
    for (int k = 1; k < nz - 1; k++)
      for (int j = 1; j < ny - 1; j++)
        for (int i = 1; i < nx - 1; i++) fn(i, j, k);
  };
  for (int n = 0; n < nsteps; n++) {
    interior([&](int i, int j, int k) { YeeUpdate<ThreeD, Dx>(dt, dx)(at(i, j, k)); });
    interior([&](int i, int j, int k) { YeeUpdate<ThreeD, Dy>(dt, dx)(at(i, j, k)); });
    interior([&](int i, int j, int k) { YeeUpdate<ThreeD, Dz>(dt, dx)(at(i, j, k)); });
    at(si, sj, sk).Dz() += sD[n];
    interior([&](int i, int j, int k) { UpdatePML<ThreeD, FieldType::Electric, Dir::X>(dt, dx)(at(i, j, k)); });
    interior([&](int i, int j, int k) { UpdatePML<ThreeD, FieldType::Electric, Dir::Y>(dt, dx)(at(i, j, k)); });
    interior([&](int i, int j, int k) { UpdatePML<ThreeD, FieldType::Electric, Dir::Z>(dt, dx)(at(i, j, k)); });
    interior([&](int i, int j, int k) { material_E<ThreeD>(at(i, j, k), me[matid[id(i, j, k)]], dt); });
    interior([&](int i, int j, int k) { YeeUpdate<ThreeD, Bx>(dt, dx)(at(i, j, k)); });
    interior([&](int i, int j, int k) { YeeUpdate<ThreeD, By>(dt, dx)(at(i, j, k)); });
    interior([&](int i, int j, int k) { YeeUpdate<ThreeD, Bz>(dt, dx)(at(i, j, k)); });
    at(bi, bj, bk).Bx() += sB[n];
    interior([&](int i, int j, int k) { UpdatePML<ThreeD, FieldType::Magnetic, Dir::X>(dt, dx)(at(i, j, k)); });
    interior([&](int i, int j, int k) { UpdatePML<ThreeD, FieldType::Magnetic, Dir::Y>(dt, dx)(at(i, j, k)); });
    interior([&](int i, int j, int k) { UpdatePML<ThreeD, FieldType::Magnetic, Dir::Z>(dt, dx)(at(i, j, k)); });
    interior([&](int i, int j, int k) { material_H<ThreeD>(at(i, j, k), mh[matid[id(i, j, k)]], dt); });
    for (auto& p : probes) {
      Cell3M& c = at(p[0], p[1], p[2]);
      ptrace.insert(ptrace.end(), {c.Ex(), c.Ey(), c.Ez(), c.Hx(), c.Hy(), c.Hz()});
    }
  }
  Fixture fx;
  fx.name = "grid3d_magnetic";
  fx.ps("mode", "ThreeD");
  fx.pi("nx", nx);
  fx.pi("ny", ny);
  fx.pi("nz", nz);
  fx.pi("nsteps", nsteps);
  fx.p("dx", dx);
  fx.p("dt", dt);
  put_pml(fx, pml);
  fx.pi("src_i", si);
  fx.pi("src_j", sj);
  fx.pi("src_k", sk);
  fx.arr("src", sD, {sD.size()});
  fx.arr("bsrc_comp_i_j_k", {0.0, double(bi), double(bj), double(bk)}, {4});
  fx.arr("bsrc", sB, {sB.size()});
  std::vector<double> mt, mm;
  for (auto& m : me) mt.insert(mt.end(), {double(m.kind), m.eps, m.freq, m.gamma, m.delta});
  for (auto& m : mh) mm.insert(mm.end(), {double(m.kind), m.mu, m.freq, m.gamma, m.delta});
  fx.arr("mats_kind_eps_freq_gamma_delta", mt, {me.size(), 5});
  fx.arr("mmats_kind_mu_freq_gamma_delta", mm, {mh.size(), 5});
  std::vector<size_t> sh = {size_t(nz), size_t(ny), size_t(nx)};
  std::vector<double> mid(matid.begin(), matid.end());
  fx.arr("matid", mid, sh);
  std::vector<double> pr;
  for (auto& p : probes) pr.insert(pr.end(), {double(p[0]), double(p[1]), double(p[2])});
  fx.arr("probes", pr, {probes.size(), 3});
  fx.arr("probe_trace_E_H", ptrace, {size_t(nsteps), probes.size(), 6});
  const char* names[6] = {"Ex", "Ey", "Ez", "Hx", "Hy", "Hz"};
  for (int f = 0; f < 6; f++) {
    std::vector<double> v;
    for (auto& c : cells)
      v.push_back(f == 0 ? c.Ex() : f == 1 ? c.Ey() : f == 2 ? c.Ez() : f == 3 ? c.Hx() : f == 4 ? c.Hy() : c.Hz());
    fx.arr(names[f], v, sh);
  }
  std::vector<double> de, dh;
  for (size_t q = 0; q < cells.size(); q++) {
    Cell3M& c = cells[q];
    if (me[matid[q]].kind >= 2)
      de.insert(de.end(), {double(q), c.Dx(), c.Dy(), c.Dz(), c.Px(), c.Py(), c.Pz(), c.Jx(), c.Jy(), c.Jz()});
    if (mh[matid[q]].kind >= 2)
      dh.insert(dh.end(), {double(q), c.Bx(), c.By(), c.Bz(), c.Mx(), c.My(), c.Mz(), c.Kx(), c.Ky(), c.Kz()});
  }
  fx.arr("dispersive_q_D_P_J", de, {de.size() / 10, 10});
  fx.arr("magnetic_q_B_M_K", dh, {dh.size() / 10, 10});
  fx.write(dir);
}

// ------------------------------------------------------------ scenario: 1-D TEM
// The reference's TEM mode (Ez, Hy along x; tests/testTEM.cpp) on a line of
// cells: CPML at both ends, a dielectric, a Drude, a Lorentz and a conductive
// segment, a Gaussian on Dz and one on By (variant 0); a TFSF pair of planes
// (one cell each, their own Aux1D, the reference's TEM polarisation) around a
// dielectric segment (variant 1).
static void scen_tem(const std::string& dir) {
  for (int variant = 0; variant < 2; variant++) {
    const int nx = 140, nsteps = variant == 0 ? 320 : 200;
    const double dx = 1e-6, dt = 0.9 * dx / c0;
    PMLParameterModel dm(dx);
    const PmlSpec pml = {10, 3, 1, dm.msMax, 5, 0.05};
    std::vector<Mat> mats = {{0, 1, 0, 0, 0}, {1, 2.25, 0, 0, 0}, {2, 1.5, 0.3 / dt, 0.02 / dt, 0},
                             {3, 2.0, 0.2 / dt, 0.01 / dt, 0.7}, {4, 1.2, 0.05 * 1.2 * eps0 / dt, 0, 0}};
    std::vector<int> matid(nx, 0);
    auto fill = [&](int x0, int x1, int m) {
      for (int i = x0; i < x1; i++) matid[i] = m;
    };
    if (variant == 0) {
      fill(40, 52, 1);
      fill(60, 70, 2);
      fill(76, 88, 3);
      fill(96, 104, 4);
    } else {
      fill(62, 78, 1);
    }
    std::vector<CellTEM> cells(nx);
    for (int i = 1; i < nx - 1; i++) {
      cells[i].mn[0] = &cells[i - 1];
      cells[i].mx[0] = &cells[i + 1];
    }
    place_pml(nx, pml, [&](int idx, bool e, double K, double Sg, double A) {
      if (e)
        cells[idx].setPMLParameters<FieldType::Electric, Dir::X>(K, Sg, A, dt);
      else
        cells[idx].setPMLParameters<FieldType::Magnetic, Dir::X>(K, Sg, A, dt);
    });
    std::vector<double> sD, sB;
    for (int n = 0; n < nsteps; n++) {
      sD.push_back(eps0 * gauss_pulse(n, 30, 8));
      sB.push_back(1e-9 * gauss_pulse(n, 45, 9));
    }
    const int si = 30, bi = 110;
    // TFSF: planes at x = 40 (MIN) and 100 (MAX)
    const int bx0 = 40, bx1 = 100;
    const double cx = std::floor(0.5 * (bx0 + bx1)) + 0.5;
    const double f0 = c0 / (20 * dx), tw = 2.0 / f0, t0 = 3.0 * tw;
    std::vector<double> aux_sig;
    typedef TFSFPlane<TEM, LocTEM, PlaneIt<CellTEM>, Aux1D<double>> PlaneT;
    std::vector<std::vector<PlaneEntry<CellTEM>>> faces(variant == 1 ? 2 : 0);
    std::vector<PlaneT> planes;
    if (variant == 1) {
      faces[0].push_back({&cells[bx0], (bx0 - cx) * dx, 0, 0});
      faces[1].push_back({&cells[bx1], (bx1 - cx) * dx, 0, 0});
      for (int fidx = 0; fidx < 2; fidx++) {
        auto sig = [f0, tw, t0, fidx, &aux_sig](double t) {
          double a = (t - t0) / tw;
          double v = std::exp(-a * a) * std::sin(2 * pi * f0 * (t - t0));
          if (fidx == 0) aux_sig.push_back(v);
          return v;
        };
        auto aux = std::make_shared<Aux1D<double>>(100, dx, dt, sig);
        std::array<double, 1> K = {1.0};
        PlaneIt<CellTEM> b{faces[fidx].data()}, e{faces[fidx].data() + faces[fidx].size()};
        if (fidx == 0)
          planes.push_back(make_tfsf_plane<TEM, Dir::X, Orientation::MIN>(PlaneIt<CellTEM>(b), PlaneIt<CellTEM>(e), LocTEM(), K, aux));
        else
          planes.push_back(make_tfsf_plane<TEM, Dir::X, Orientation::MAX>(PlaneIt<CellTEM>(b), PlaneIt<CellTEM>(e), LocTEM(), K, aux));
      }
    }
    const std::vector<int> probes = variant == 0 ? std::vector<int>{30, 45, 64, 80, 100, 110, 5}
                                                 : std::vector<int>{70, 50, 40, 39, 100, 101, 20};
    std::vector<double> ptrace;
    for (int n = 0; n < nsteps; n++) {
      for (int i = 1; i < nx - 1; i++) YeeUpdate<TEM, Dz>(dt, dx)(cells[i]);
      if (variant == 0) cells[si].Dz() += sD[n];
      for (int i = 1; i < nx - 1; i++) UpdatePML<TEM, FieldType::Electric, Dir::X>(dt, dx)(cells[i]);
      for (auto& pl : planes) pl.updateD(dt, dx);
      for (int i = 1; i < nx - 1; i++) material_E<TEM>(cells[i], mats[matid[i]], dt);
      for (int i = 1; i < nx - 1; i++) YeeUpdate<TEM, By>(dt, dx)(cells[i]);
      if (variant == 0) cells[bi].By() += sB[n];
      for (int i = 1; i < nx - 1; i++) UpdatePML<TEM, FieldType::Magnetic, Dir::X>(dt, dx)(cells[i]);
      for (auto& pl : planes) pl.updateB(dt, dx);
      for (int i = 1; i < nx - 1; i++) VacuumUpdate<TEM, FieldType::Magnetic>()(cells[i]);
      for (int q : probes) {
        ptrace.push_back(cells[q].Ez());
        ptrace.push_back(cells[q].Hy());
      }
    }
    Fixture fx;
    fx.name = variant == 0 ? "tem1d_small" : "tem1d_tfsf";
    fx.ps("mode", "TEM");
    fx.pi("nx", nx);
    fx.pi("ny", 1);
    fx.pi("nsteps", nsteps);
    fx.p("dx", dx);
    fx.p("dt", dt);
    put_pml(fx, pml);
    fx.pi("pmlx", 1);
    fx.pi("pmly", 0);
    if (variant == 0) {
      fx.arr("src_field_comp_i_j", {0.0, 2.0, double(si), 0.0, 2.0, 1.0, double(bi), 0.0}, {2, 4});
      std::vector<double> st(sD);
      st.insert(st.end(), sB.begin(), sB.end());
      fx.arr("src_tables", st, {2, size_t(nsteps)});
    } else {
      fx.pi("tfsf", 1);
      fx.pi("bx0", bx0);
      fx.pi("bx1", bx1);
      fx.p("kx", 1.0);
      fx.pi("aux_ncells", 100);
      fx.p("cx", cx);
      fx.p("f0", f0);
      fx.p("tw", tw);
      fx.p("t0", t0);
      fx.arr("aux_signal", aux_sig, {aux_sig.size()});
    }
    std::vector<double> mt;
    for (auto& m : mats) mt.insert(mt.end(), {double(m.kind), m.eps, m.freq, m.gamma, m.delta});
    fx.arr("mats_kind_eps_freq_gamma_delta", mt, {mats.size(), 5});
    std::vector<double> mid(matid.begin(), matid.end());
    fx.arr("matid", mid, {1, size_t(nx)});
    std::vector<double> pr;
    for (int q : probes) pr.push_back(double(q));
    fx.arr("probes", pr, {probes.size(), 1});
    fx.arr("probe_trace_Ez_Hy", ptrace, {size_t(nsteps), probes.size(), 2});
    std::vector<double> Ez, Dzv, Hy, Byv;
    for (auto& c : cells) {
      Ez.push_back(c.Ez());
      Dzv.push_back(c.Dz());
      Hy.push_back(c.Hy());
      Byv.push_back(c.By());
    }
    fx.arr("Ez", Ez, {1, size_t(nx)});
    fx.arr("Dz", Dzv, {1, size_t(nx)});
    fx.arr("Hy", Hy, {1, size_t(nx)});
    fx.arr("By", Byv, {1, size_t(nx)});
    fx.write(dir);
  }
}

// ------------------------------------------------------------ scenario: material map XML
// The reference's MaterialMap / MaterialPair (include/MaterialMap.hpp:45-284)
// holding one material of every kind the device path covers, written by its
// own print_summary (the XML form its readXML reads back).  Part of that
// output goes to std::cout rather than the given stream, so both are captured.
static void scen_material_map(const std::string& dir) {
  const double dt = 2e-17;
  typedef MaterialMap<std::string, CellTM&> MapT;
  MapT mm;
  mm.insert(std::make_pair(std::string("vacuum"), make_material_pair<CellTM&>(VacuumUpdate<TM, FieldType::Electric>(), VacuumUpdate<TM, FieldType::Magnetic>())));
  mm.insert(std::make_pair(std::string("glass"), make_material_pair<CellTM&>(ConstantUpdate<TM, FieldType::Electric>(2.25), ConstantUpdate<TM, FieldType::Magnetic>(1.0))));
  mm.insert(std::make_pair(std::string("ferrite"), make_material_pair<CellTM&>(ConstantUpdate<TM, FieldType::Electric>(4.5), ConstantUpdate<TM, FieldType::Magnetic>(3.2))));
  mm.insert(std::make_pair(std::string("conductor"), make_material_pair<CellTM&>(ConductiveUpdate<TM, FieldType::Electric>(2.0, 5e5, dt), VacuumUpdate<TM, FieldType::Magnetic>())));
  mm.insert(std::make_pair(std::string("silver"), make_material_pair<CellTM&>(DrudeUpdate<TM, FieldType::Electric>(3.7, 1.37e16, 2.73e13, dt), VacuumUpdate<TM, FieldType::Magnetic>())));
  mm.insert(std::make_pair(std::string("resonator"), make_material_pair<CellTM&>(LorentzUpdate<TM, FieldType::Electric>(1.5, 0.8, 4.1e15, 1.2e14, dt), ConstantUpdate<TM, FieldType::Magnetic>(1.0))));
  mm.insert(std::make_pair(std::string("biased plasma"), make_material_pair<CellTM&>(MagnetizedDrudeUpdate<TM, FieldType::Electric>(0.3, -0.4, 1.25, 1.0, 2e11, 1e9, dt), VacuumUpdate<TM, FieldType::Magnetic>())));
  SelfMagnetizedDrudeUpdate<TM, FieldType::Electric> self(dt);
  self.K() = 1.5;
  self.Freq() = 1.5e11;
  self.Gamma() = 2e9;
  mm.insert(std::make_pair(std::string("self plasma"), make_material_pair<CellTM&>(self, VacuumUpdate<TM, FieldType::Magnetic>())));
  mm.insert(std::make_pair(std::string("fluid plasma"), make_material_pair<CellTM&>(FluidDrudeUpdate<TM, FieldType::Electric>(dt), VacuumUpdate<TM, FieldType::Magnetic>())));
  mm.insert(std::make_pair(std::string("biased fluid"), make_material_pair<CellTM&>(MagnetizedFluidDrudeUpdate<TM, FieldType::Electric>(0.1, 0.2, 0.9, dt), ConstantUpdate<TM, FieldType::Magnetic>(1.0))));
  mm.insert(std::make_pair(std::string("self fluid"), make_material_pair<CellTM&>(SelfMagnetizedFluidDrudeUpdate<TM, FieldType::Electric>(dt), VacuumUpdate<TM, FieldType::Magnetic>())));
  std::ostringstream os;
  std::streambuf* old = std::cout.rdbuf(os.rdbuf());
  mm.print_summary(os, 0);
  std::cout.rdbuf(old);
  FILE* f = fopen((dir + "/material_map.xml").c_str(), "w");
  fputs(os.str().c_str(), f);
  fclose(f);
}

// ------------------------------------------------------------ scenario: surface quantities
// The reference's UpdateSurfaceChargeDensity / UpdateSurfaceCurrentE /
// UpdateSurfaceCurrentH (include/SurfaceQuantities.hpp:169-356) evaluated on
// every interior cell of a seeded 3-D field state and of a seeded TM state.
// The functors take any cell type with the field accessors, neighbour links
// and SurfaceQuantities storage, so a thin view over the grid cells is used.
template <class C, class SQ>
struct SQView : public SQ {
  C* c;
  double& Ex() { return c->Ex(); }
  double& Ey() { return c->Ey(); }
  double& Ez() { return c->Ez(); }
  double& Hx() { return c->Hx(); }
  double& Hy() { return c->Hy(); }
  double& Hz() { return c->Hz(); }
  C& getNeighborMin(int d) { return c->getNeighborMin(d); }
  C& getNeighborMax(int d) { return c->getNeighborMax(d); }
};

static void scen_surface(const std::string& dir) {
  Fixture fx;
  fx.name = "surface_quantities";
  const double dx = 1e-6;
  // 3-D
  {
    const int n[3] = {9, 8, 7};
    const int nx = n[0], ny = n[1], nz = n[2];
    std::vector<Cell3> cells(size_t(nx) * ny * nz);
    auto at = [&](int i, int j, int k) -> Cell3& { return cells[size_t(i) + size_t(nx) * (size_t(j) + size_t(ny) * k)]; };
    for (int k = 0; k < nz; k++)
      for (int j = 0; j < ny; j++)
        for (int i = 0; i < nx; i++) {
          Cell3& c = at(i, j, k);
          const int ci[3] = {i, j, k};
          for (int a = 0; a < 3; a++) {
            int lo[3] = {i, j, k}, hi[3] = {i, j, k};
            lo[a]--;
            hi[a]++;
            c.mn[a] = ci[a] > 0 ? &at(lo[0], lo[1], lo[2]) : nullptr;
            c.mx[a] = ci[a] < n[a] - 1 ? &at(hi[0], hi[1], hi[2]) : nullptr;
          }
        }
    SplitMix64 rng(13);
    std::vector<double> init;
    for (auto& c : cells) {
      c.Ex() = rng.uniform() - 0.5;
      c.Ey() = rng.uniform() - 0.5;
      c.Ez() = rng.uniform() - 0.5;
      c.Hx() = (rng.uniform() - 0.5) * 2e-3;
      c.Hy() = (rng.uniform() - 0.5) * 2e-3;
      c.Hz() = (rng.uniform() - 0.5) * 2e-3;
    }
    for (int f = 0; f < 6; f++)
      for (auto& c : cells) init.push_back(f == 0 ? c.Ex() : f == 1 ? c.Ey() : f == 2 ? c.Ez() : f == 3 ? c.Hx() : f == 4 ? c.Hy() : c.Hz());
    std::vector<double> out;
    UpdateSurfaceChargeDensity<ThreeD> rho(dx);
    UpdateSurfaceCurrentE<ThreeD> je(dx);
    UpdateSurfaceCurrentH<ThreeD> jh(dx);
    for (int k = 1; k < nz - 1; k++)
      for (int j = 1; j < ny - 1; j++)
        for (int i = 1; i < nx - 1; i++) {
          SQView<Cell3, SurfaceQuantities<ThreeD>> v;
          v.c = &at(i, j, k);
          rho(v);
          je(v);
          jh(v);
          out.insert(out.end(), {v.rho(), v.Jex(), v.Jey(), v.Jez(), v.Jhx(), v.Jhy(), v.Jhz()});
        }
    fx.pi("nx", nx);
    fx.pi("ny", ny);
    fx.pi("nz", nz);
    fx.p("dx", dx);
    fx.arr("init3_E_H", init, {6, size_t(nz), size_t(ny), size_t(nx)});
    fx.arr("sq3_rho_Je_Jh", out, {size_t(nz - 2), size_t(ny - 2), size_t(nx - 2), 7});
  }
  // TM: Jez, Jhx, Jhy (the TM charge density functor is a no-op)
  {
    const int nx = 10, ny = 8;
    std::vector<CellTM> cells(size_t(nx) * ny);
    auto at = [&](int i, int j) -> CellTM& { return cells[size_t(i) + size_t(nx) * j]; };
    for (int j = 0; j < ny; j++)
      for (int i = 0; i < nx; i++) {
        CellTM& c = at(i, j);
        c.mn[0] = i > 0 ? &at(i - 1, j) : nullptr;
        c.mx[0] = i < nx - 1 ? &at(i + 1, j) : nullptr;
        c.mn[1] = j > 0 ? &at(i, j - 1) : nullptr;
        c.mx[1] = j < ny - 1 ? &at(i, j + 1) : nullptr;
      }
    SplitMix64 rng(17);
    std::vector<double> init;
    for (auto& c : cells) {
      c.Ez() = rng.uniform() - 0.5;
      c.Hx() = (rng.uniform() - 0.5) * 2e-3;
      c.Hy() = (rng.uniform() - 0.5) * 2e-3;
    }
    for (int f = 0; f < 3; f++)
      for (auto& c : cells) init.push_back(f == 0 ? c.Ez() : f == 1 ? c.Hx() : c.Hy());
    std::vector<double> out;
    UpdateSurfaceCurrentE<TM> je(dx);
    UpdateSurfaceCurrentH<TM> jh(dx);
    struct TMView : public SurfaceQuantities<TM> {
      CellTM* c;
      double& Ez() { return c->Ez(); }
      double& Hx() { return c->Hx(); }
      double& Hy() { return c->Hy(); }
      CellTM& getNeighborMin(int d) { return c->getNeighborMin(d); }
      CellTM& getNeighborMax(int d) { return c->getNeighborMax(d); }
    };
    for (int j = 1; j < ny - 1; j++)
      for (int i = 1; i < nx - 1; i++) {
        TMView v;
        v.c = &at(i, j);
        je(v);
        jh(v);
        out.insert(out.end(), {v.Jez(), v.Jhx(), v.Jhy()});
      }
    fx.pi("tm_nx", nx);
    fx.pi("tm_ny", ny);
    fx.arr("initTM_Ez_Hx_Hy", init, {3, size_t(ny), size_t(nx)});
    fx.arr("sqTM_Jez_Jhx_Jhy", out, {size_t(ny - 2), size_t(nx - 2), 3});
  }
  // TE: rho, Jex, Jey, Jhz
  {
    const int nx = 9, ny = 11;
    std::vector<CellTE> cells(size_t(nx) * ny);
    auto at = [&](int i, int j) -> CellTE& { return cells[size_t(i) + size_t(nx) * j]; };
    for (int j = 0; j < ny; j++)
      for (int i = 0; i < nx; i++) {
        CellTE& c = at(i, j);
        c.mn[0] = i > 0 ? &at(i - 1, j) : nullptr;
        c.mx[0] = i < nx - 1 ? &at(i + 1, j) : nullptr;
        c.mn[1] = j > 0 ? &at(i, j - 1) : nullptr;
        c.mx[1] = j < ny - 1 ? &at(i, j + 1) : nullptr;
      }
    SplitMix64 rng(19);
    std::vector<double> init;
    for (auto& c : cells) {
      c.Ex() = rng.uniform() - 0.5;
      c.Ey() = rng.uniform() - 0.5;
      c.Hz() = (rng.uniform() - 0.5) * 2e-3;
    }
    for (int f = 0; f < 3; f++)
      for (auto& c : cells) init.push_back(f == 0 ? c.Ex() : f == 1 ? c.Ey() : c.Hz());
    std::vector<double> out;
    UpdateSurfaceChargeDensity<TE> rho(dx);
    UpdateSurfaceCurrentE<TE> je(dx);
    UpdateSurfaceCurrentH<TE> jh(dx);
    struct TEView : public SurfaceQuantities<TE> {
      CellTE* c;
      double& Ex() { return c->Ex(); }
      double& Ey() { return c->Ey(); }
      double& Hz() { return c->Hz(); }
      CellTE& getNeighborMin(int d) { return c->getNeighborMin(d); }
      CellTE& getNeighborMax(int d) { return c->getNeighborMax(d); }
    };
    for (int j = 1; j < ny - 1; j++)
      for (int i = 1; i < nx - 1; i++) {
        TEView v;
        v.c = &at(i, j);
        rho(v);
        je(v);
        jh(v);
        out.insert(out.end(), {v.rho(), v.Jex(), v.Jey(), v.Jhz()});
      }
    fx.pi("te_nx", nx);
    fx.pi("te_ny", ny);
    fx.arr("initTE_Ex_Ey_Hz", init, {3, size_t(ny), size_t(nx)});
    fx.arr("sqTE_rho_Jex_Jey_Jhz", out, {size_t(ny - 2), size_t(nx - 2), 4});
  }
  fx.write(dir);
}

// ------------------------------------------------------------ scenario: boundary updaters
// The reference's own make_{pec,pmc,symmetric,antisymmetric,periodic}_boundary
// (include/BoundaryUpdates.hpp:527-832) applied to the planes next to each face
// of a small neighbour-wired 3-D grid with seeded E and H.  Boundary cells are
// the interior cells of layer 1 (min) / L-2 (max); their outward neighbours
// (layer 0 / L-1) receive the update; the periodic image of layer 1 is layer
// L-2 and vice versa.  Bloch-periodic needs complex fields and is left out.
template <class C>
struct CellRefIt {
  typedef std::ptrdiff_t difference_type;
  typedef C value_type;
  typedef C& reference;
  typedef C* pointer;
  typedef std::forward_iterator_tag iterator_category;
  C** p;
  C& operator*() const { return **p; }
  C* operator->() const { return *p; }
  CellRefIt& operator++() {
    ++p;
    return *this;
  }
  CellRefIt operator++(int) {
    CellRefIt t = *this;
    ++p;
    return t;
  }
  bool operator!=(const CellRefIt& o) const { return p != o.p; }
  bool operator==(const CellRefIt& o) const { return p == o.p; }
};
struct PeriodicImage {
  Cell3** base;
  Cell3** img;
  Cell3& operator()(const CellRefIt<Cell3>& it) const { return *img[it.p - base]; }
};

template <Dir d, Orientation o>
static BoundaryData make_bc(int type, CellRefIt<Cell3> b, CellRefIt<Cell3> e, const PeriodicImage& pf) {
  switch (type) {
    case 1: return make_pec_boundary<ThreeD, d, o>(b, e);
    case 2: return make_pmc_boundary<ThreeD, d, o>(b, e);
    case 3: return make_symmetric_boundary<ThreeD, d, o>(b, e);
    case 4: return make_antisymmetric_boundary<ThreeD, d, o>(b, e);
    default: return make_periodic_boundary<ThreeD, d, o>(b, e, pf);
  }
}

static void scen_boundaries(const std::string& dir) {
  const int n[3] = {7, 6, 5};
  const int nx = n[0], ny = n[1], nz = n[2];
  std::vector<Cell3> cells(size_t(nx) * ny * nz);
  auto at = [&](int i, int j, int k) -> Cell3& { return cells[size_t(i) + size_t(nx) * (size_t(j) + size_t(ny) * k)]; };
  for (int k = 0; k < nz; k++)
    for (int j = 0; j < ny; j++)
      for (int i = 0; i < nx; i++) {
        Cell3& c = at(i, j, k);
        const int ci[3] = {i, j, k};
        for (int a = 0; a < 3; a++) {
          int lo[3] = {i, j, k}, hi[3] = {i, j, k};
          lo[a]--;
          hi[a]++;
          c.mn[a] = ci[a] > 0 ? &at(lo[0], lo[1], lo[2]) : nullptr;
          c.mx[a] = ci[a] < n[a] - 1 ? &at(hi[0], hi[1], hi[2]) : nullptr;
        }
      }
  std::vector<double> init;
  auto seed = [&]() {
    SplitMix64 rng(11);
    init.clear();
    for (auto& c : cells) {
      c.Ex() = rng.uniform() - 0.5;
      c.Ey() = rng.uniform() - 0.5;
      c.Ez() = rng.uniform() - 0.5;
      c.Hx() = rng.uniform() - 0.5;
      c.Hy() = rng.uniform() - 0.5;
      c.Hz() = rng.uniform() - 0.5;
    }
    // initial fields, component-major, x fastest
    for (int f = 0; f < 6; f++)
      for (auto& c : cells) init.push_back(f == 0 ? c.Ex() : f == 1 ? c.Ey() : f == 2 ? c.Ez() : f == 3 ? c.Hx() : f == 4 ? c.Hy() : c.Hz());
  };
  Fixture fx;
  fx.name = "boundaries";
  std::vector<double> cases, ghosts;
  size_t per_case = 0;
  for (int type = 1; type <= 5; type++)
    for (int a = 0; a < 3; a++)
      for (int side = 0; side < 2; side++)
        for (int ft = 0; ft < 2; ft++) {
          seed();
          const int L = n[a];
          const int src = side ? L - 2 : 1, dst = side ? L - 1 : 0, img = side ? 1 : L - 2;
          const int t1 = a == 0 ? 1 : 0, t2 = a == 2 ? 1 : 2;
          std::vector<Cell3*> bnd, per;
          for (int v = 1; v < n[t2] - 1; v++)
            for (int u = 1; u < n[t1] - 1; u++) {
              int c1[3], c2[3];
              c1[a] = src;
              c1[t1] = u;
              c1[t2] = v;
              c2[a] = img;
              c2[t1] = u;
              c2[t2] = v;
              bnd.push_back(&at(c1[0], c1[1], c1[2]));
              per.push_back(&at(c2[0], c2[1], c2[2]));
            }
          CellRefIt<Cell3> b{bnd.data()}, e{bnd.data() + bnd.size()};
          PeriodicImage pf{bnd.data(), per.data()};
          BoundaryData bd = a == 0 ? (side ? make_bc<Dir::X, Orientation::MAX>(type, b, e, pf)
                                           : make_bc<Dir::X, Orientation::MIN>(type, b, e, pf))
                            : a == 1 ? (side ? make_bc<Dir::Y, Orientation::MAX>(type, b, e, pf)
                                             : make_bc<Dir::Y, Orientation::MIN>(type, b, e, pf))
                                     : (side ? make_bc<Dir::Z, Orientation::MAX>(type, b, e, pf)
                                             : make_bc<Dir::Z, Orientation::MIN>(type, b, e, pf));
          if (ft == 0)
            bd.updateE();
          else
            bd.updateH();
          cases.insert(cases.end(), {double(type), double(a), double(side), double(ft)});
          size_t cnt = 0;
          for (int v = 1; v < n[t2] - 1; v++)
            for (int u = 1; u < n[t1] - 1; u++) {
              int c1[3];
              c1[a] = dst;
              c1[t1] = u;
              c1[t2] = v;
              const Cell3& g = at(c1[0], c1[1], c1[2]);
              if (ft == 0)
                ghosts.insert(ghosts.end(), {g.Ex(), g.Ey(), g.Ez()});
              else
                ghosts.insert(ghosts.end(), {g.Hx(), g.Hy(), g.Hz()});
              cnt += 3;
            }
          per_case = std::max(per_case, cnt);
          // pad to a common length per case (planes differ in size per axis)
          while (cnt < size_t(3 * (std::max({nx, ny, nz}) - 2) * (std::max({nx, ny, nz}) - 2))) {
            ghosts.push_back(0.0);
            cnt++;
          }
        }
  const size_t stride = size_t(3 * (std::max({nx, ny, nz}) - 2) * (std::max({nx, ny, nz}) - 2));
  fx.pi("nx", nx);
  fx.pi("ny", ny);
  fx.pi("nz", nz);
  fx.ps("ghost_layout", "per case: the ghost plane's tangential interior cells, lower tangential axis fastest, 3 comps each, zero padded");
  fx.arr("init_E_H", init, {6, size_t(nz), size_t(ny), size_t(nx)});
  fx.arr("cases_type_axis_side_ft", cases, {cases.size() / 4, 4});
  fx.arr("ghosts", ghosts, {cases.size() / 4, stride});
  fx.write(dir);
}

// ------------------------------------------------------------ scenario: Bloch-periodic faces
// make_bloch_periodic_boundary (include/BoundaryUpdates.hpp:845-885) needs
// complex fields: the reference's own functors on cells whose scalar is
// std::complex<double> (curl, CPML, materials and the Bloch ghost writes all
// take the complex scalar; sources stay real).  3-D grid: x and y
// Bloch-periodic (phase e^{-/+ i theta} at the min / max face), CPML along z
// (ghost layers there stay zero), a dielectric and a Drude block, a soft
// source on Dz.  Probes record the real and imaginary parts every step.
typedef std::complex<double> cplx;
typedef Nb<FluidCell<YeeCell<YeeFields<ThreeD, cplx, std::array>, DoublePolarization<cplx>,
                             VacuumMagnetization<cplx>, NewPML<ThreeD, cplx>>>>
    Cell3c;
struct PeriodicImageC {
  Cell3c** base;
  Cell3c** img;
  Cell3c& operator()(const CellRefIt<Cell3c>& it) const { return *img[it.p - base]; }
};

static void scen_bloch(const std::string& dir) {
  const int nx = 14, ny = 12, nz = 16, nsteps = 80;
  const double dx = 1e-8, dt = 0.5 * dx / (c0 * std::sqrt(3.0));
  const double th[2] = {0.9, -0.4};  // x, y phase across one period
  std::vector<Cell3c> cells(size_t(nx) * ny * nz);
  auto id = [&](int i, int j, int k) { return size_t(i) + size_t(nx) * (size_t(j) + size_t(ny) * k); };
  auto at = [&](int i, int j, int k) -> Cell3c& { return cells[id(i, j, k)]; };
  const int n[3] = {nx, ny, nz};
  for (int k = 0; k < nz; k++)
    for (int j = 0; j < ny; j++)
      for (int i = 0; i < nx; i++) {
        Cell3c& c = at(i, j, k);
        const int ci[3] = {i, j, k};
        for (int a = 0; a < 3; a++) {
          int lo[3] = {i, j, k}, hi[3] = {i, j, k};
          lo[a]--;
          hi[a]++;
          c.mn[a] = ci[a] > 0 ? &at(lo[0], lo[1], lo[2]) : nullptr;
          c.mx[a] = ci[a] < n[a] - 1 ? &at(hi[0], hi[1], hi[2]) : nullptr;
        }
      }
  PMLParameterModel dm(dx);
  const PmlSpec pml = {4, 3, 1, dm.msMax, 5, 0.05};
  place_pml(nz, pml, [&](int idx, bool e, double K, double Sg, double A) {
    for (int j = 0; j < ny; j++)
      for (int i = 0; i < nx; i++) {
        if (e)
          at(i, j, idx).setPMLParameters<FieldType::Electric, Dir::Z>(K, Sg, A, dt);
        else
          at(i, j, idx).setPMLParameters<FieldType::Magnetic, Dir::Z>(K, Sg, A, dt);
      }
  });
  std::vector<Mat> mats = {{0, 1, 0, 0, 0}, {1, 3.0, 0, 0, 0}, {2, 1.0, 1.37e16, 1e14, 0}};
  std::vector<int> matid(cells.size(), 0);
  for (int k = 6; k < 10; k++)
    for (int j = 3; j < 8; j++)
      for (int i = 2; i < 7; i++) matid[id(i, j, k)] = 1;
  for (int k = 9; k < 12; k++)
    for (int j = 5; j < 10; j++)
      for (int i = 8; i < 12; i++) matid[id(i, j, k)] = 2;
  // Bloch faces: the boundary layer (1 / L-2) and its periodic image (L-2 / 1)
  std::vector<std::vector<Cell3c*>> bnd(4), img(4);
  std::vector<BoundaryData> faces;
  for (int a = 0; a < 2; a++)
    for (int side = 0; side < 2; side++) {
      const int L = n[a], src = side ? L - 2 : 1, im = side ? 1 : L - 2;
      const int t = a == 0 ? 1 : 0;
      auto& B = bnd[2 * a + side];
      auto& I = img[2 * a + side];
      for (int k = 1; k < nz - 1; k++)
        for (int u = 1; u < n[t] - 1; u++) {
          int c1[3], c2[3];
          c1[a] = src;
          c1[t] = u;
          c1[2] = k;
          c2[a] = im;
          c2[t] = u;
          c2[2] = k;
          B.push_back(&at(c1[0], c1[1], c1[2]));
          I.push_back(&at(c2[0], c2[1], c2[2]));
        }
    }
  for (int a = 0; a < 2; a++)
    for (int side = 0; side < 2; side++) {
      auto& B = bnd[2 * a + side];
      auto& I = img[2 * a + side];
      CellRefIt<Cell3c> b{B.data()}, e{B.data() + B.size()};
      PeriodicImageC pf{B.data(), I.data()};
      const cplx f = std::polar(1.0, side ? th[a] : -th[a]);
      if (a == 0)
        faces.push_back(side ? make_bloch_periodic_boundary<ThreeD, Dir::X, Orientation::MAX>(b, e, pf, f)
                             : make_bloch_periodic_boundary<ThreeD, Dir::X, Orientation::MIN>(b, e, pf, f));
      else
        faces.push_back(side ? make_bloch_periodic_boundary<ThreeD, Dir::Y, Orientation::MAX>(b, e, pf, f)
                             : make_bloch_periodic_boundary<ThreeD, Dir::Y, Orientation::MIN>(b, e, pf, f));
    }
  auto interior = [&](auto&& fn) {
    for (int k = 1; k < nz - 1; k++)
      for (int j = 1; j < ny - 1; j++)
        for (int i = 1; i < nx - 1; i++) fn(i, j, k);
  };
  const int probes[4][3] = {{7, 6, 8}, {1, 3, 5}, {12, 10, 9}, {4, 1, 11}};
  std::vector<double> src, ptrace;
  for (int step = 0; step < nsteps; step++) src.push_back(eps0 * gauss_pulse(step, 16, 5));
  for (int step = 0; step < nsteps; step++) {
    interior([&](int i, int j, int k) { YeeUpdate<ThreeD, Dx>(dt, dx)(at(i, j, k)); });
    interior([&](int i, int j, int k) { YeeUpdate<ThreeD, Dy>(dt, dx)(at(i, j, k)); });
    interior([&](int i, int j, int k) { YeeUpdate<ThreeD, Dz>(dt, dx)(at(i, j, k)); });
    at(5, 4, 7).Dz() += src[step];
    interior([&](int i, int j, int k) { UpdatePML<ThreeD, FieldType::Electric, Dir::Z>(dt, dx)(at(i, j, k)); });
    interior([&](int i, int j, int k) { material_E<ThreeD>(at(i, j, k), mats[matid[id(i, j, k)]], dt); });
    for (auto& f : faces) f.updateE();
    interior([&](int i, int j, int k) { YeeUpdate<ThreeD, Bx>(dt, dx)(at(i, j, k)); });
    interior([&](int i, int j, int k) { YeeUpdate<ThreeD, By>(dt, dx)(at(i, j, k)); });
    interior([&](int i, int j, int k) { YeeUpdate<ThreeD, Bz>(dt, dx)(at(i, j, k)); });
    interior([&](int i, int j, int k) { UpdatePML<ThreeD, FieldType::Magnetic, Dir::Z>(dt, dx)(at(i, j, k)); });
    interior([&](int i, int j, int k) { VacuumUpdate<ThreeD, FieldType::Magnetic>()(at(i, j, k)); });
    for (auto& f : faces) f.updateH();
    for (auto& p : probes) {
      Cell3c& c = at(p[0], p[1], p[2]);
      for (cplx v : {c.Ex(), c.Ey(), c.Ez(), c.Hx(), c.Hy(), c.Hz()}) ptrace.insert(ptrace.end(), {v.real(), v.imag()});
    }
  }
  Fixture fx;
  fx.name = "grid3d_bloch";
  fx.ps("mode", "ThreeD");
  fx.pi("nx", nx);
  fx.pi("ny", ny);
  fx.pi("nz", nz);
  fx.pi("nsteps", nsteps);
  fx.p("dx", dx);
  fx.p("dt", dt);
  put_pml(fx, pml);
  fx.pi("pmlx", 0);
  fx.pi("pmly", 0);
  fx.p("theta_x", th[0]);
  fx.p("theta_y", th[1]);
  fx.pi("src_i", 5);
  fx.pi("src_j", 4);
  fx.pi("src_k", 7);
  std::vector<double> mt;
  for (auto& m : mats) mt.insert(mt.end(), {double(m.kind), m.eps, m.freq, m.gamma, m.delta});
  fx.arr("mats_kind_eps_freq_gamma_delta", mt, {mats.size(), 5});
  std::vector<double> mid(matid.begin(), matid.end());
  std::vector<size_t> sh = {size_t(nz), size_t(ny), size_t(nx)};
  fx.arr("matid", mid, sh);
  fx.arr("src", src, {src.size()});
  std::vector<double> pr;
  for (auto& p : probes) pr.insert(pr.end(), {double(p[0]), double(p[1]), double(p[2])});
  fx.arr("probes", pr, {4, 3});
  fx.arr("probe_trace_E_H_reim", ptrace, {size_t(nsteps), 4, 6, 2});
  const char* names[6] = {"Ex", "Ey", "Ez", "Hx", "Hy", "Hz"};
  for (int f = 0; f < 6; f++)
    for (int part = 0; part < 2; part++) {
      std::vector<double> v;
      for (auto& c : cells) {
        Cell3c& cc = const_cast<Cell3c&>(c);
        cplx x = f == 0 ? cc.Ex() : f == 1 ? cc.Ey() : f == 2 ? cc.Ez() : f == 3 ? cc.Hx() : f == 4 ? cc.Hy() : cc.Hz();
        v.push_back(part ? x.imag() : x.real());
      }
      fx.arr(std::string(names[f]) + (part ? "_im" : "_re"), v, sh);
    }
  fx.write(dir);
}

int main(int argc, char** argv) {
  if (argc >= 2 && !strcmp(argv[1], "bench3d")) {
    // timing of the reference's own functors on a bounded sample of the 3-D
    // CPML + dielectric-block workload: N^3 cells, `steps` timed steps after
    // one untimed step.
    int N = argc > 2 ? atoi(argv[2]) : 256;
    int steps = argc > 3 ? atoi(argv[3]) : 2;
    int npml = argc > 4 ? atoi(argv[4]) : 12;
    Scenario3 S;
    S.name = "bench";
    S.nx = S.ny = S.nz = N;
    S.nsteps = steps + 1;
    S.dx = 1e-6;
    S.dt = 0.5 * S.dx / (c0 * std::sqrt(3.0));
    PMLParameterModel dm(S.dx);
    S.pml = {npml, 3, 1, dm.msMax, 5, 0.05};
    S.mats = {{0, 1, 0, 0, 0}};
    S.matid.assign(size_t(N) * N * N, 0);
    SplitMix64 rng(2);
    seed_blocks(S.matid, N, N, N, 48, npml + 2, npml + 2, std::max(2, N / 32), std::max(3, N / 4), rng, S.mats, 1, 12,
                nullptr);
    S.si = S.sj = S.sk = N / 2;
    for (int n = 0; n < S.nsteps; n++) S.src.push_back(eps0 * gauss_pulse(n, 40, 12));
    double secs = 0;
    run_3d(S, "", true, 1, &secs);
    double cells = double(N - 2) * (N - 2) * (N - 2) * steps;
    printf("{\"N\": %d, \"steps\": %d, \"seconds\": %.6f, \"cell_updates_per_s\": %.6e}\n", N, steps, secs, cells / secs);
    return 0;
  }
  if (argc < 2) {
    fprintf(stderr, "usage: harness <outdir> | harness bench3d N steps [npml]\n");
    return 2;
  }
  std::string dir = argv[1];
  scen_pml_coeffs(dir);
  scen_boundaries(dir);
  scen_bloch(dir);
  scen_surface(dir);
  scen_material_map(dir);
  scen_aux1d(dir);
  scen_magnetic(dir);
  scen_tem(dir);

  // 2D TM, small, non-square, CPML x+y, two dielectric blocks, one Drude and
  // one Lorentz block, Gaussian soft source on Dz.
  {
    TMScenario S;
    S.name = "tm2d_small";
    S.nx = 48;
    S.ny = 40;
    S.nsteps = 150;
    S.dx = 1e-6;
    S.dt = 0.9 * S.dx / (c0 * std::sqrt(2.0));
    PMLParameterModel dm(S.dx);
    S.pml = {10, 3, 1, dm.msMax, 5, 0.05};
    S.pmlx = S.pmly = true;
    S.mats = {{0, 1, 0, 0, 0}, {1, 4.0, 0, 0, 0}, {1, 2.25, 0, 0, 0}};
    // Drude and Lorentz with parameters resolved at this dt
    S.mats.push_back({2, 1.5, 0.3 / S.dt, 0.02 / S.dt, 0});
    S.mats.push_back({3, 2.0, 0.2 / S.dt, 0.01 / S.dt, 0.7});
    S.matid.assign(size_t(S.nx) * S.ny, 0);
    auto fill = [&](int x0, int x1, int y0, int y1, int m) {
      for (int j = y0; j < y1; j++)
        for (int i = x0; i < x1; i++) S.matid[size_t(i) + size_t(S.nx) * j] = m;
    };
    fill(14, 22, 13, 19, 1);
    fill(27, 33, 22, 29, 2);
    fill(15, 20, 24, 30, 3);
    fill(28, 34, 12, 17, 4);
    S.src_i = 24;
    S.src_j = 20;
    for (int n = 0; n < S.nsteps; n++) S.src.push_back(eps0 * gauss_pulse(n, 25, 8));
    S.probes = {{24, 20}, {17, 15}, {30, 25}, {12, 30}, {3, 3}};
    run_tm(S, dir);
  }
  // 2D TE (Ex, Ey, Hz): CPML x+y, two dielectric blocks, a Drude and a
  // Lorentz block, a Gaussian soft source on Dx and the pulse of
  // tests/testTE.cpp on Bz; and a TE TFSF box (oblique wave, the reference's
  // TE polarisation) around a dielectric block.
  for (int variant = 0; variant < 2; variant++) {
    TEScenario S;
    S.name = variant == 0 ? "te2d_small" : "te2d_tfsf";
    S.nx = variant == 0 ? 48 : 64;
    S.ny = variant == 0 ? 40 : 60;
    S.nsteps = variant == 0 ? 150 : 160;
    S.dx = 1e-6;
    S.dt = (variant == 0 ? 0.9 : 0.5) * S.dx / (c0 * std::sqrt(2.0));
    PMLParameterModel dm(S.dx);
    S.pml = {10, 3, 1, dm.msMax, 5, 0.05};
    S.mats = {{0, 1, 0, 0, 0}, {1, 4.0, 0, 0, 0}, {1, 2.25, 0, 0, 0}};
    S.mats.push_back({2, 1.5, 0.3 / S.dt, 0.02 / S.dt, 0});
    S.mats.push_back({3, 2.0, 0.2 / S.dt, 0.01 / S.dt, 0.7});
    S.matid.assign(size_t(S.nx) * S.ny, 0);
    auto fill = [&](int x0, int x1, int y0, int y1, int m) {
      for (int j = y0; j < y1; j++)
        for (int i = x0; i < x1; i++) S.matid[size_t(i) + size_t(S.nx) * j] = m;
    };
    if (variant == 0) {
      fill(14, 22, 13, 19, 1);
      fill(27, 33, 22, 29, 2);
      fill(15, 20, 24, 30, 3);
      fill(28, 34, 12, 17, 4);
      TEScenario::Src d{0, 0, 24, 20, {}}, b{2, 2, 20, 25, {}};
      for (int n = 0; n < S.nsteps; n++) {
        d.table.push_back(eps0 * gauss_pulse(n, 25, 8));
        b.table.push_back(std::exp(-(n - 10.0) * (n - 10.0) * c0 / (20 * S.dx) * S.dt * S.dt) * 1e-9);
      }
      S.src = {d, b};
      S.probes = {{24, 20}, {17, 15}, {30, 25}, {12, 30}, {3, 3}, {20, 25}};
    } else {
      fill(28, 36, 26, 33, 1);
      S.tfsf = true;
      S.bx0 = 14;
      S.bx1 = 49;
      S.by0 = 14;
      S.by1 = 45;
      S.kx = std::cos(pi / 6);
      S.ky = std::sin(pi / 6);
      S.aux_ncells = 100;
      S.f0 = c0 / (20 * S.dx);
      S.tw = 2.0 / S.f0;
      S.t0 = 3.0 * S.tw;
      S.probes = {{30, 30}, {15, 30}, {14, 30}, {13, 30}, {49, 20}, {50, 20}, {31, 14}, {31, 13}};
    }
    run_te(S, dir);
  }
  // 2D TM at the 256^2 / 1000-step shape: dx = 1 m, c0 dt/dx = 0.5, 10-cell
  // CPML on all edges, vacuum, Dz[128,128] += eps0 exp(-((n-40)/12)^2),
  // 4 seeded probes in the non-PML interior.
  {
    TMScenario S;
    S.name = "tm2d_c1";
    S.nx = S.ny = 256;
    S.nsteps = 1000;
    S.dx = 1.0;
    S.dt = 0.5 * S.dx / c0;
    PMLParameterModel dm(S.dx);
    S.pml = {10, 3, 1, dm.msMax, 5, 0.05};
    S.pmlx = S.pmly = true;
    S.mats = {{0, 1, 0, 0, 0}};
    S.matid.assign(size_t(S.nx) * S.ny, 0);
    S.src_i = S.src_j = 128;
    for (int n = 0; n < S.nsteps; n++) S.src.push_back(eps0 * gauss_pulse(n, 40, 12));
    SplitMix64 rng(1);
    for (int p = 0; p < 4; p++) S.probes.push_back({rng.randint(12, 243), rng.randint(12, 243)});
    S.save_full = false;
    run_tm(S, dir);
  }
  // 2D TM TFSF box (inset 14) around a dielectric block, x-propagating and
  // oblique plane waves.
  for (int variant = 0; variant < 2; variant++) {
    TMScenario S;
    S.name = variant == 0 ? "tm2d_tfsf_x" : "tm2d_tfsf_oblique";
    S.nx = 64;
    S.ny = 60;
    S.nsteps = 160;
    S.dx = 1e-6;
    S.dt = 0.5 * S.dx / (c0 * std::sqrt(2.0));
    PMLParameterModel dm(S.dx);
    S.pml = {10, 3, 1, dm.msMax, 5, 0.05};
    S.pmlx = S.pmly = true;
    S.mats = {{0, 1, 0, 0, 0}, {1, 3.0, 0, 0, 0}};
    S.matid.assign(size_t(S.nx) * S.ny, 0);
    for (int j = 26; j < 33; j++)
      for (int i = 28; i < 36; i++) S.matid[size_t(i) + size_t(S.nx) * j] = variant;  // block only in variant 1
    S.src_i = -1;
    S.src_j = -1;
    S.tfsf = true;
    S.bx0 = 14;
    S.bx1 = 49;
    S.by0 = 14;
    S.by1 = 45;
    S.kx = variant == 0 ? 1.0 : std::cos(pi / 6);
    S.ky = variant == 0 ? 0.0 : std::sin(pi / 6);
    S.aux_ncells = 100;
    S.f0 = c0 / (20 * S.dx);
    S.tw = 2.0 / S.f0;
    S.t0 = 3.0 * S.tw;
    S.probes = {{30, 30}, {5 + 10, 30}, {14, 30}, {13, 30}, {49, 20}, {50, 20}, {31, 14}, {31, 13}};
    run_tm(S, dir);
  }
  // 2D TM: soft sources on Ez (inside a Lorentz block) and on Hy, added
  // after their material steps, and two single TFSF planes (an x-MIN face
  // and a y-MAX face, each with its own Aux1D and its own pulse delay) under
  // an oblique wave; CPML x+y, a dielectric block.
  {
    TMScenario S;
    S.name = "tm2d_soft_planes";
    S.nx = 56;
    S.ny = 50;
    S.nsteps = 170;
    S.dx = 1e-6;
    S.dt = 0.7 * S.dx / (c0 * std::sqrt(2.0));
    PMLParameterModel dm(S.dx);
    S.pml = {10, 3, 1, dm.msMax, 5, 0.05};
    S.pmlx = S.pmly = true;
    S.mats = {{0, 1, 0, 0, 0}, {1, 2.5, 0, 0, 0}};
    S.mats.push_back({3, 2.0, 0.2 / S.dt, 0.01 / S.dt, 0.7});
    S.matid.assign(size_t(S.nx) * S.ny, 0);
    for (int j = 26; j < 34; j++)
      for (int i = 30; i < 38; i++) S.matid[size_t(i) + size_t(S.nx) * j] = 1;
    for (int j = 30; j < 36; j++)
      for (int i = 18; i < 24; i++) S.matid[size_t(i) + size_t(S.nx) * j] = 2;
    S.src_i = -1;
    S.src_j = -1;
    TMScenario::Soft se{1, 2, 20, 32, {}}, sh{3, 1, 34, 18, {}};
    const double imp = std::sqrt(mu0 / eps0);
    for (int n = 0; n < S.nsteps; n++) {
      se.table.push_back(gauss_pulse(n, 30, 8));
      sh.table.push_back(gauss_pulse(n, 45, 10) / imp);
    }
    S.soft = {se, sh};
    S.kx = std::cos(25.0 * pi / 180.0);
    S.ky = std::sin(25.0 * pi / 180.0);
    S.aux_ncells = 100;
    S.f0 = c0 / (20 * S.dx);
    S.tw = 2.0 / S.f0;
    S.cx1 = 27.5;
    S.cy1 = 25.5;
    S.planes1 = {{0, 0, 14, 12, 37, 3.0 * S.tw}, {1, 1, 38, 16, 40, 4.0 * S.tw}};
    S.probes = {{20, 32}, {34, 18}, {14, 24}, {13, 24}, {28, 38}, {28, 39}, {40, 20}, {45, 45}};
    run_tm(S, dir);
  }
  // 3-D, small and anisotropic, CPML on all faces, two dielectric blocks,
  // Gaussian soft source on Dz.
  {
    Scenario3 S;
    S.name = "grid3d_small";
    S.nx = 22;
    S.ny = 19;
    S.nz = 17;
    S.nsteps = 50;
    S.dx = 1e-6;
    S.dt = 0.5 * S.dx / (c0 * std::sqrt(3.0));
    PMLParameterModel dm(S.dx);
    S.pml = {4, 3, 1, dm.msMax, 5, 0.05};
    S.mats = {{0, 1, 0, 0, 0}};
    S.matid.assign(size_t(S.nx) * S.ny * S.nz, 0);
    SplitMix64 rng(7);
    std::vector<double> boxes;
    seed_blocks(S.matid, S.nx, S.ny, S.nz, 3, 6, 6, 2, 5, rng, S.mats, 1, 12, &boxes);
    S.si = 11;
    S.sj = 9;
    S.sk = 8;
    for (int n = 0; n < S.nsteps; n++) S.src.push_back(eps0 * gauss_pulse(n, 15, 5));
    S.probes = {{11, 9, 8}, {7, 7, 7}, {15, 12, 10}, {2, 3, 4}, {19, 16, 14}};
    run_3d(S, dir);
  }
  // 3-D with single-pole Drude and Lorentz blocks (the reference holds one
  // P/J pair per cell, include/DispersiveMaterials.hpp:72-83).
  {
    Scenario3 S;
    S.name = "grid3d_disp";
    S.nx = 18;
    S.ny = 17;
    S.nz = 16;
    S.nsteps = 60;
    S.dx = 1e-8;
    S.dt = 0.5 * S.dx / (c0 * std::sqrt(3.0));
    PMLParameterModel dm(S.dx);
    S.pml = {4, 3, 1, dm.msMax, 5, 0.05};
    S.mats = {{0, 1, 0, 0, 0}, {2, 1.0, 1.37e16, 1e14, 0}, {3, 1.5, 4e15, 1e14, 1.5}, {1, 2.5, 0, 0, 0}};
    S.matid.assign(size_t(S.nx) * S.ny * S.nz, 0);
    auto fill = [&](int x0, int x1, int y0, int y1, int z0, int z1, int m) {
      for (int k = z0; k < z1; k++)
        for (int j = y0; j < y1; j++)
          for (int i = x0; i < x1; i++) S.matid[size_t(i) + size_t(S.nx) * (size_t(j) + size_t(S.ny) * k)] = m;
    };
    fill(6, 10, 6, 11, 6, 9, 1);
    fill(10, 13, 5, 9, 8, 11, 2);
    fill(6, 9, 9, 12, 9, 11, 3);
    S.si = 9;
    S.sj = 8;
    S.sk = 8;
    for (int n = 0; n < S.nsteps; n++) S.src.push_back(eps0 * gauss_pulse(n, 12, 4));
    S.probes = {{9, 8, 8}, {7, 8, 7}, {11, 7, 9}, {7, 10, 10}};
    run_3d(S, dir);
  }
  // 0-D material recursions on one TM cell: prescribed D_n, reference functors.
  auto cell_traces = [&](const char* name, double dt, const std::vector<Mat>& ms) {
    Fixture fx;
    fx.name = name;
    std::vector<double> out, mt;
    const int nsteps = 300;
    for (auto& m : ms) {
      CellTM c;
      for (int n = 0; n < nsteps; n++) {
        c.Dz() = eps0 * (std::sin(0.07 * n) + 0.3 * std::cos(0.011 * n * n));
        material_E<TM>(c, m, dt);
        out.insert(out.end(), {c.Dz(), c.Ez(), c.Pz(), c.Jz()});
      }
      mt.insert(mt.end(), {double(m.kind), m.eps, m.freq, m.gamma, m.delta});
    }
    fx.p("dt", dt);
    fx.pi("nsteps", nsteps);
    fx.arr("mats_kind_eps_freq_gamma_delta", mt, {ms.size(), 5});
    fx.arr("trace_D_E_P_J", out, {ms.size(), size_t(nsteps), 4});
    fx.write(dir);
  };
  cell_traces("cell_ade", 1e-17,
              {{2, 1.0, 1.37e16, 1e14, 0}, {3, 2.0, 3e15, 1e14, 0.8}, {2, 3.0, 5e15, 2e15, 0}, {3, 1.0, 6e15, 0, 1.5},
               {1, 7.5, 0, 0, 0}});
  // conductive (include/DispersiveMaterials.hpp:508-680): factor = dt sigma/(eps_r eps0) from ~0.01 to ~2
  cell_traces("cell_conductive", 1e-17, {{4, 1.0, 1e4, 0, 0}, {4, 2.5, 2e5, 0, 0}, {4, 7.0, 1.2e7, 0, 0}});
  // magnetized Drude (include/DispersiveMaterials.hpp:1263-2028), microwave
  // plasma scales (dt 1 ps: wp = dt omega_p ~ 0.2, v = dt q B/m ~ 0.2 per tesla).
  // 3-D: one cell with prescribed D_n (and, self-magnetized, B_n) per component.
  {
    Fixture fx;
    fx.name = "cell_magnetized";
    const double dt = 1e-12;
    const int nsteps = 200;
    std::vector<Mat> ms = {{5, 1.0, 2e11, 1e9, 0, 0.0, 0.0, 1.0},
                           {5, 2.0, 1.5e11, 3e10, 0, 0.6, -1.0, 1.6},
                           {5, 1.0, 3e11, 0, 0, 2.0, 0.5, 0.0},
                           {6, 1.0, 2e11, 1e9, 0},
                           {6, 1.5, 1e11, 5e9, 0}};
    std::vector<double> out, mt;
    for (auto& m : ms) {
      Cell3 c;
      for (int n = 0; n < nsteps; n++) {
        c.Dx() = eps0 * (std::sin(0.07 * n) + 0.3 * std::cos(0.011 * n * n));
        c.Dy() = eps0 * (0.5 * std::cos(0.05 * n + 0.3) - 0.2 * std::sin(0.013 * n * n));
        c.Dz() = eps0 * (0.8 * std::sin(0.031 * n + 1.1));
        c.Bx() = 0.7 * std::sin(0.02 * n);
        c.By() = -1.2 + 0.4 * std::cos(0.045 * n);
        c.Bz() = 1.5 * std::cos(0.017 * n + 0.4);
        material_E<ThreeD>(c, m, dt);
        out.insert(out.end(), {c.Dx(), c.Dy(), c.Dz(), c.Ex(), c.Ey(), c.Ez(), c.Px(), c.Py(), c.Pz(), c.Jx(),
                               c.Jy(), c.Jz(), c.Bx(), c.By(), c.Bz()});
      }
      mt.insert(mt.end(), {double(m.kind), m.eps, m.freq, m.gamma, m.bx, m.by, m.bz});
    }
    fx.p("dt", dt);
    fx.pi("nsteps", nsteps);
    fx.arr("mats_kind_eps_freq_gamma_bx_by_bz", mt, {ms.size(), 7});
    fx.arr("trace_D3_E3_P3_J3_B3", out, {ms.size(), size_t(nsteps), 15});
    fx.write(dir);
  }
  // TM: the reference runs DrudeUpdate per component for these (MagnetizedDrudeCall, 1508-1537)
  cell_traces("cell_magnetized_tm", 1e-12, {{5, 1.0, 2e11, 1e9, 0, 0.0, 0.0, 1.0}, {6, 2.0, 1.5e11, 3e10, 0}});
  // 3-D grid with a biased and a self-magnetized plasma block, CPML, a strong Dz
  // source (the self-magnetized block sees fields of ~1e8 V/m, so |v| ~ 0.1)
  {
    Scenario3 S;
    S.name = "grid3d_magnetized";
    S.nx = 20;
    S.ny = 18;
    S.nz = 17;
    S.nsteps = 70;
    S.dx = 1e-3;
    S.dt = 0.5 * S.dx / (c0 * std::sqrt(3.0));
    PMLParameterModel dm(S.dx);
    S.pml = {4, 3, 1, dm.msMax, 5, 0.05};
    S.mats = {{0, 1, 0, 0, 0}, {5, 1.0, 2e11, 1e9, 0, 0.3, -0.4, 1.2}, {6, 1.5, 1.5e11, 2e9, 0}, {1, 2.5, 0, 0, 0}};
    S.matid.assign(size_t(S.nx) * S.ny * S.nz, 0);
    auto fill = [&](int x0, int x1, int y0, int y1, int z0, int z1, int m) {
      for (int k = z0; k < z1; k++)
        for (int j = y0; j < y1; j++)
          for (int i = x0; i < x1; i++) S.matid[size_t(i) + size_t(S.nx) * (size_t(j) + size_t(S.ny) * k)] = m;
    };
    fill(6, 10, 6, 12, 6, 10, 1);
    fill(11, 14, 6, 10, 8, 11, 2);
    fill(6, 9, 12, 14, 9, 12, 3);
    S.si = 10;
    S.sj = 9;
    S.sk = 8;
    for (int n = 0; n < S.nsteps; n++) S.src.push_back(eps0 * 3e9 * gauss_pulse(n, 14, 5));
    S.probes = {{10, 9, 8}, {8, 8, 8}, {12, 8, 9}, {7, 13, 10}};
    run_3d(S, dir);
  }
  // fluid Drude family (include/DispersiveMaterials.hpp:1160-1214, 1728-1893):
  // omega_p = sqrt(q^2/(m eps0) ne) ~ 2e11 rad/s at ne ~ 1.26e19 / m^3
  {
    Fixture fx;
    fx.name = "cell_fluid";
    const double dt = 1e-12;
    const int nsteps = 200;
    std::vector<Mat> ms = {{7, 1, 0, 0, 0}, {8, 1, 0, 0, 0, 0.6, -1.0, 1.6}, {9, 1, 0, 0, 0}};
    std::vector<double> out, mt;
    for (auto& m : ms) {
      Cell3 c;
      for (int n = 0; n < nsteps; n++) {
        c.Dx() = eps0 * (std::sin(0.07 * n) + 0.3 * std::cos(0.011 * n * n));
        c.Dy() = eps0 * (0.5 * std::cos(0.05 * n + 0.3) - 0.2 * std::sin(0.013 * n * n));
        c.Dz() = eps0 * (0.8 * std::sin(0.031 * n + 1.1));
        c.Bx() = 0.7 * std::sin(0.02 * n);
        c.By() = -1.2 + 0.4 * std::cos(0.045 * n);
        c.Bz() = 1.5 * std::cos(0.017 * n + 0.4);
        c.ne() = 1.26e19 * (1.0 + 0.5 * std::sin(0.03 * n));
        c.nu_m() = 1e9 * (1.0 + 0.3 * std::cos(0.05 * n));
        material_E<ThreeD>(c, m, dt);
        out.insert(out.end(), {c.Dx(), c.Dy(), c.Dz(), c.Ex(), c.Ey(), c.Ez(), c.Px(), c.Py(), c.Pz(), c.Jx(),
                               c.Jy(), c.Jz(), c.Bx(), c.By(), c.Bz(), c.ne(), c.nu_m()});
      }
      mt.insert(mt.end(), {double(m.kind), m.bx, m.by, m.bz});
    }
    fx.p("dt", dt);
    fx.pi("nsteps", nsteps);
    fx.arr("mats_kind_bx_by_bz", mt, {ms.size(), 4});
    fx.arr("trace_D3_E3_P3_J3_B3_ne_nu", out, {ms.size(), size_t(nsteps), 17});
    fx.write(dir);
  }
  {
    Fixture fx;
    fx.name = "cell_fluid_tm";
    const double dt = 1e-12;
    const int nsteps = 200;
    std::vector<Mat> ms = {{7, 1, 0, 0, 0}, {8, 1, 0, 0, 0, 0.6, -1.0, 1.6}, {9, 1, 0, 0, 0}};
    std::vector<double> out, mt;
    for (auto& m : ms) {
      CellTM c;
      for (int n = 0; n < nsteps; n++) {
        c.Dz() = eps0 * (std::sin(0.07 * n) + 0.3 * std::cos(0.011 * n * n));
        c.ne() = 1.26e19 * (1.0 + 0.5 * std::sin(0.03 * n));
        c.nu_m() = 1e9 * (1.0 + 0.3 * std::cos(0.05 * n));
        material_E<TM>(c, m, dt);
        out.insert(out.end(), {c.Dz(), c.Ez(), c.Pz(), c.Jz(), c.ne(), c.nu_m()});
      }
      mt.insert(mt.end(), {double(m.kind), m.bx, m.by, m.bz});
    }
    fx.p("dt", dt);
    fx.pi("nsteps", nsteps);
    fx.arr("mats_kind_bx_by_bz", mt, {ms.size(), 4});
    fx.arr("trace_D_E_P_J_ne_nu", out, {ms.size(), size_t(nsteps), 6});
    fx.write(dir);
  }
  // 3-D grid with the three fluid forms on a graded density, CPML, strong Dz source
  {
    Scenario3 S;
    S.name = "grid3d_fluid";
    S.nx = 20;
    S.ny = 18;
    S.nz = 17;
    S.nsteps = 70;
    S.dx = 1e-3;
    S.dt = 0.5 * S.dx / (c0 * std::sqrt(3.0));
    PMLParameterModel dm(S.dx);
    S.pml = {4, 3, 1, dm.msMax, 5, 0.05};
    S.mats = {{0, 1, 0, 0, 0}, {7, 1, 0, 0, 0}, {8, 1, 0, 0, 0, 0.3, -0.4, 1.2}, {9, 1, 0, 0, 0}};
    const size_t ncell = size_t(S.nx) * S.ny * S.nz;
    S.matid.assign(ncell, 0);
    S.ne.assign(ncell, 0.0);
    S.nu.assign(ncell, 0.0);
    for (int k = 0; k < S.nz; k++)
      for (int j = 0; j < S.ny; j++)
        for (int i = 0; i < S.nx; i++) {
          const size_t q = size_t(i) + size_t(S.nx) * (size_t(j) + size_t(S.ny) * k);
          S.ne[q] = 1.26e19 * (0.4 + 0.06 * i + 0.01 * j * (k % 4));
          S.nu[q] = 1e9 * (1.0 + 0.25 * ((i + 2 * j + 3 * k) % 5));
        }
    auto fill = [&](int x0, int x1, int y0, int y1, int z0, int z1, int m) {
      for (int k = z0; k < z1; k++)
        for (int j = y0; j < y1; j++)
          for (int i = x0; i < x1; i++) S.matid[size_t(i) + size_t(S.nx) * (size_t(j) + size_t(S.ny) * k)] = m;
    };
    fill(6, 10, 6, 12, 6, 10, 1);
    fill(11, 14, 6, 10, 8, 11, 2);
    fill(6, 9, 12, 14, 9, 12, 3);
    S.si = 10;
    S.sj = 9;
    S.sk = 8;
    for (int n = 0; n < S.nsteps; n++) S.src.push_back(eps0 * 3e9 * gauss_pulse(n, 14, 5));
    S.probes = {{10, 9, 8}, {8, 8, 8}, {12, 8, 9}, {7, 13, 10}};
    run_3d(S, dir);
  }
  // 2D TM with the fluid forms (all the Drude update per component in TM)
  {
    TMScenario S;
    S.name = "tm2d_fluid";
    S.nx = 40;
    S.ny = 36;
    S.nsteps = 120;
    S.dx = 1e-3;
    S.dt = 0.9 * S.dx / (c0 * std::sqrt(2.0));
    PMLParameterModel dm(S.dx);
    S.pml = {8, 3, 1, dm.msMax, 5, 0.05};
    S.pmlx = S.pmly = true;
    S.mats = {{0, 1, 0, 0, 0}, {7, 1, 0, 0, 0}, {8, 1, 0, 0, 0, 0.3, -0.4, 1.2}, {9, 1, 0, 0, 0}};
    S.matid.assign(size_t(S.nx) * S.ny, 0);
    S.ne.assign(size_t(S.nx) * S.ny, 0.0);
    S.nu.assign(size_t(S.nx) * S.ny, 0.0);
    for (int j = 0; j < S.ny; j++)
      for (int i = 0; i < S.nx; i++) {
        S.ne[size_t(i) + size_t(S.nx) * j] = 1.26e19 * (0.3 + 0.05 * i + 0.02 * (j % 7));
        S.nu[size_t(i) + size_t(S.nx) * j] = 2e9 * (1.0 + 0.5 * ((i * j) % 3));
      }
    auto fill = [&](int x0, int x1, int y0, int y1, int m) {
      for (int j = y0; j < y1; j++)
        for (int i = x0; i < x1; i++) S.matid[size_t(i) + size_t(S.nx) * j] = m;
    };
    fill(12, 18, 11, 17, 1);
    fill(22, 28, 19, 25, 2);
    fill(12, 17, 21, 26, 3);
    S.src_i = 20;
    S.src_j = 18;
    for (int n = 0; n < S.nsteps; n++) S.src.push_back(eps0 * gauss_pulse(n, 20, 6));
    S.probes = {{20, 18}, {15, 14}, {25, 22}, {10, 28}};
    run_tm(S, dir);
  }
  // 2D TM with conductive blocks next to a dielectric one, CPML x+y, Dz source
  {
    TMScenario S;
    S.name = "tm2d_conductive";
    S.nx = 46;
    S.ny = 42;
    S.nsteps = 160;
    S.dx = 1e-6;
    S.dt = 0.9 * S.dx / (c0 * std::sqrt(2.0));
    PMLParameterModel dm(S.dx);
    S.pml = {10, 3, 1, dm.msMax, 5, 0.05};
    S.pmlx = S.pmly = true;
    S.mats = {{0, 1, 0, 0, 0}, {4, 2.0, 1e6, 0, 0}, {4, 1.0, 5e4, 0, 0}, {1, 3.0, 0, 0, 0}};
    S.matid.assign(size_t(S.nx) * S.ny, 0);
    auto fill = [&](int x0, int x1, int y0, int y1, int m) {
      for (int j = y0; j < y1; j++)
        for (int i = x0; i < x1; i++) S.matid[size_t(i) + size_t(S.nx) * j] = m;
    };
    fill(14, 21, 13, 20, 1);
    fill(26, 32, 22, 29, 2);
    fill(14, 19, 24, 30, 3);
    S.src_i = 23;
    S.src_j = 21;
    for (int n = 0; n < S.nsteps; n++) S.src.push_back(eps0 * gauss_pulse(n, 25, 8));
    S.probes = {{23, 21}, {17, 16}, {29, 25}, {12, 30}};
    run_tm(S, dir);
  }
  printf("fixtures written to %s\n", dir.c_str());
  return 0;
}
