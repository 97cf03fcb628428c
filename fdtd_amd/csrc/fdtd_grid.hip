// Host side of the C ABI declared in include/fdtd_hip.h.
//
// A grid handle owns its device memory; set-up calls stage small tables on the
// host and upload them, fdtd_step() launches the per-step passes on the
// grid's stream.  The reference has no device layer at all: its "grid" is a
// user container of YeeCell objects linked by neighbour pointers
// (tests/test3D.cpp:33-49), swept once per functor on one CPU thread.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>
#include <map>
#include <chrono>
#include <mutex>
#include <thread>
#include <tuple>
#include <vector>

#include "../../include/fdtd_hip.h"
#include "fdtd_adjoint.hpp"
#include "fdtd_kernels.hpp"

namespace {

thread_local std::string g_err;

int fail(int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_err = buf;
  return code;
}

#define HIP_TRY(expr)                                                                     \
  do {                                                                                    \
    hipError_t e_ = (expr);                                                               \
    if (e_ != hipSuccess) return fail(FDTD_EDEVICE, "%s: %s", #expr, hipGetErrorString(e_)); \
  } while (0)

// include/FDTDConstants.hpp:394-398 (imp0 via the same Newton iteration)
constexpr double kPi = 3.14159265358979323846264338327950288;
constexpr double kEps0 = 8.854e-12;
// electron charge and mass of the reference's Drude models (DispersiveMaterials.hpp:899-900)
constexpr double kQe = 1.6e-19;
constexpr double kMe = 9.11e-31;
const double kMu0 = kPi * 4.0e-7;
double newton_sqrt(double x) {
  double curr = x, prev = 0;
  while (curr != prev) {
    double nxt = 0.5 * (curr + x / curr);
    prev = curr;
    curr = nxt;
  }
  return curr;
}
const double kImp0 = newton_sqrt(kMu0 / kEps0);

// NewPML::setPMLParameters, include/PML.hpp:85-102: {B, C, F, G}
void pml_bcfg(double K, double S, double A, double dt, double out[4]) {
  double nu = 1.0 / kEps0 * (S / K + A);
  double B = std::exp(-dt * nu);
  out[0] = B;
  out[1] = (nu == 0) ? 0.0 : S / K * 1.0 / (S + K * A) * (B - 1.0);
  double u = -S / (kEps0 * K * K);
  out[2] = (nu == 0) ? 1.0 / K : 1.0 / K - 1.0 / dt * (1.0 - B - nu * dt) * u / (nu * nu);
  out[3] = (nu == 0) ? 0.0 : 1.0 / dt * (1.0 - B) / nu;
}

// PMLParameterModel K/S/A, include/PML.hpp:378-391
void pml_ksa(double x, double m, double ma, double smax, double kmax, double amax, double out[3]) {
  if (x < 0.0) {
    out[0] = 1.0;
    out[1] = 0.0;
    out[2] = 0.0;
    return;
  }
  out[0] = x > 1.0 ? 1.0 : 1.0 + std::pow(x, m) * (kmax - 1.0);
  out[1] = x > 1.0 ? 0.0 : std::pow(x, m) * smax;
  out[2] = x > 1.0 ? 0.0 : std::pow(1.0 - x, ma) * amax;
}

// placement idiom (tests/testTM.cpp:46-71, include/Source.hpp:380-389): per-index
// {B,C,F,G} rows for E and H along an axis of length L; unplaced rows keep the
// NewPML() defaults B=C=G=0, F=1.
void pml_idiom_rows(int L, int n, double m, double ma, double smax, double kmax, double amax, double dt,
                    std::vector<double>& e, std::vector<double>& h) {
  e.assign(4 * L, 0.0);
  h.assign(4 * L, 0.0);
  for (int i = 0; i < L; i++) e[4 * i + 2] = h[4 * i + 2] = 1.0;
  for (int i = 1; i < n + 2; i++) {
    double x = double(n - i + 1) / double(n);
    double xm = x - 0.5 / double(n);
    double ke[3], kh[3], r[4];
    pml_ksa(x, m, ma, smax, kmax, amax, ke);
    pml_ksa(xm, m, ma, smax, kmax, amax, kh);
    auto put = [&](std::vector<double>& v, int idx, const double* ksa) {
      if (idx < 0 || idx >= L) return;
      pml_bcfg(ksa[0], ksa[1], ksa[2], dt, r);
      std::copy(r, r + 4, v.begin() + 4 * idx);
    };
    put(e, i, ke);
    put(h, i, kh);
    put(e, L - i, ke);
    put(h, L - 1 - i, kh);
  }
}

constexpr int kChunk = 256;    // steps per incident-table chunk
constexpr int kIncBlock = 64;  // steps per launch of the adjoint's precomputed incident rows
// timed kernels: 0 E pass, 1 H pass, 2 aux line, 3 fused pass, 4 DFT,
// 5 adjoint E pass, 6 adjoint H pass
constexpr int kNumKernels = 8;

// tile of the per-step passes: threads per block in x and y, z planes per thread
// bx == 0 selects the streamed fused pass (3-D FP32) with E regions of `by`
// rows and `kc` workgroups per compute unit, bx == 1 the z-marching fused
// pass with workgroups of 64 x `by` columns and `kc` planes, bx == 2 the
// time-blocked pass (two steps per launch, 64 x `by` columns, `kc` planes);
// bx >= 8 the two passes with bx x by threads per workgroup.  TM grids:
// bx == 0 selects the time-blocked pass, `by` steps per launch on cores of
// kc x kc cells.
struct Tile {
  int bx, by, kc;
};
constexpr Tile kDefaultTile3D = {256, 1, 1};
constexpr Tile kDefaultTileTM = {64, 4, 1};

}  // namespace

// what a captured chunk depends on beyond the staged arguments: the argument
// slot, the chunk length, the step phase within a DFT block, the schedule,
// the E-pass instance (a material set after stepping started can change it)
// and the DFT launch sizes
struct GraphKey {
  int slot, chunk, phase, bx, by, kc, disp;
  long dft_work, dft_cells;
  bool operator<(const GraphKey& o) const {
    if (disp != o.disp) return disp < o.disp;
    if (slot != o.slot) return slot < o.slot;
    if (chunk != o.chunk) return chunk < o.chunk;
    if (phase != o.phase) return phase < o.phase;
    if (bx != o.bx) return bx < o.bx;
    if (by != o.by) return by < o.by;
    if (kc != o.kc) return kc < o.kc;
    if (dft_work != o.dft_work) return dft_work < o.dft_work;
    return dft_cells < o.dft_cells;
  }
};

struct MatHost {
  double eps = 1.0, mu = 1.0;
  int np = 0;
  int kind[FDTD_MAX_POLES] = {0, 0, 0};
  double freq[FDTD_MAX_POLES] = {0, 0, 0}, gamma[FDTD_MAX_POLES] = {0, 0, 0}, weight[FDTD_MAX_POLES] = {0, 0, 0};
  // magnetized Drude: 0 none, 1 fixed bias field `bias` [T], 2 the cell's own B
  int mag = 0;
  double bias[3] = {0, 0, 0};
  // fluid Drude: omega_p and gamma of each cell from fdtd_set_fluid_state
  int fluid = 0;
  // magnetic side: poles on B, H, M, K (fdtd_set_material_magnetic)
  int mnp = 0;
  int mkind[FDTD_MAX_POLES] = {0, 0, 0};
  double mfreq[FDTD_MAX_POLES] = {0, 0, 0}, mgamma[FDTD_MAX_POLES] = {0, 0, 0}, mweight[FDTD_MAX_POLES] = {0, 0, 0};
};

// One Aux1D incident line on the device (include/Source.hpp:280-458): its
// state, CPML rows and source signal, and the rows it writes for the current
// chunk (Ez after each step's advance, Hy before it).
struct AuxLine {
  int L = 0;
  double *state = nullptr, *pE = nullptr, *pH = nullptr, *sig = nullptr;
  int nsig = 0, sig_cap = 0;  // signal entries in use / allocated
  double *incE = nullptr, *incH = nullptr;
};

struct fdtd_grid {
  fdtd_grid_desc d{};
  int ndim = 3;
  bool f64 = false;
  size_t esz = 4;
  long ncell = 0;
  int len[3] = {1, 1, 1};
  hipStream_t stream = nullptr;
  bool actE[3] = {false, false, false}, actH[3] = {false, false, false};
  std::vector<void*> allocs;
  std::vector<size_t> alloc_bytes;

  void* E[3] = {nullptr, nullptr, nullptr};
  void* H[3] = {nullptr, nullptr, nullptr};
  uint8_t* mat = nullptr;
  std::vector<uint8_t> mat_host;
  void* mats_dev = nullptr;
  void* mie_dev = nullptr;  // T[256]: 1/(eps_r eps0), negated for dispersive ids
  void* mim_dev = nullptr;  // T[256]: 1/(mu_r mu0)
  MatHost mats[FDTD_MAX_MATERIALS];

  // CPML rows {B,C,F,G} per [ft][axis]
  std::vector<double> pml[2][3];
  void* coef[2][3][4] = {};  // f, g, b, c device arrays of T
  void* psi[2][3][3] = {};
  int slab_a[3] = {0, 0, 0}, slab_b[3] = {1, 1, 1}, slab_S[3] = {0, 0, 0};

  // dispersive box
  int dbox[6] = {0, 0, 0, 0, 0, 0};
  int max_np = 0;
  void* Dd[3] = {};
  // fluid Drude: per-cell electron density and collision rate (host, whole
  // grid) and the per-cell coefficients over the dispersive box (device)
  std::vector<double> fluid_ne, fluid_nu;
  void* fl[2] = {};
  bool fluid_dirty = true;
  void* Pd[FDTD_MAX_POLES][3] = {};
  void* Jd[FDTD_MAX_POLES][3] = {};
  // second copies of the stepped state for the fused (ping-pong) schedule
  bool pingpong = false;
  void* E2[3] = {};
  void* H2[3] = {};
  void* psi2[2][3][3] = {};
  void* Dd2[3] = {};
  void* Pd2[FDTD_MAX_POLES][3] = {};
  void* Jd2[FDTD_MAX_POLES][3] = {};
  // magnetic-pole box: bounding box of the interior cells whose material has
  // magnetic poles, and B, M, K over it (in place: such grids run the
  // general two-pass instances)
  int mbox[6] = {0, 0, 0, 0, 0, 0};
  int max_mnp = 0;
  void* Bd[3] = {};
  void* Md[FDTD_MAX_POLES][3] = {};
  void* Kd[FDTD_MAX_POLES][3] = {};

  // sources (nsrc_eh of them on E or H); their tables and allocated lengths
  int nsrc = 0, nsrc_eh = 0;
  fdtdk::SrcDev src[FDTD_MAX_SOURCES];
  double* src_table[FDTD_MAX_SOURCES] = {};
  int src_cap[FDTD_MAX_SOURCES] = {};

  // TFSF box
  bool tfsf = false;
  int box[6] = {};
  double kvec[3] = {}, pol[3] = {}, polH[3] = {}, origin[3] = {};
  AuxLine aux;

  // single TFSF planes, each with its own line (dev: all but the row pointers)
  int nplane = 0;
  fdtdk::PlaneDev plane[FDTD_MAX_PLANES];
  AuxLine plane_aux[FDTD_MAX_PLANES];

  // DFT
  int ndft = 0;
  struct Dft {
    int r[6];
    int mask, nfreq, ncomp, nphase;
    long ncells;
    void* acc;
    void* smp;  // samples of the steps not yet folded into acc
    double *phE, *phH;
    void *acc_im, *smp_im;  // the imaginary part's (complex grids)
  } dft[FDTD_MAX_DFT];
  long dft_done = 0;  // steps folded into the DFT accumulators

  // probes
  int nprobe = 0;
  fdtdk::ProbeDev probe[FDTD_MAX_PROBES];
  double* trace = nullptr;
  double* trace_im = nullptr;  // the imaginary part's (complex grids)
  long trace_cap = 0;

  // complex fields (a Bloch-periodic face): the imaginary part lives in the
  // second copies of the stepped state (E2, H2, psi2, Dd2, Pd2, Jd2)
  bool cplx = false;
  double bloch[3][2][2] = {};  // phase factor (re, im) of E per [axis][side]

  // design box: per-cell eps_r replacing the material table's value there
  bool design = false;
  int dsg[6] = {0, 0, 0, 0, 0, 0};
  long ndes = 0;
  std::vector<double> dsg_eps;
  void* dsg_ie_dev = nullptr;  // T[ndes] = 1/(eps_r eps0)

  // forward-step controls used by the adjoint
  bool force_two_pass = false;  // the design increments are recorded by the two-pass E pass
  int bc[3][2] = {};            // boundary condition per [axis][side] (FDTD_BC_*)
  int nbc = 0;                  // faces with a condition (they need the two-pass schedule)
  int nmag = 0;                 // magnetized Drude materials in the table (3-D: two-pass schedule, own E instance)
  bool recompute = false;       // segment recompute: no DFT accumulation, no probes
  void* hist = nullptr;         // T[hist_rows][3][ndes]
  long hist_step0 = 0, hist_rows = 0;
  double *inc_allE = nullptr, *inc_allH = nullptr;  // incident rows of steps [0, inc_rows)
  long inc_rows = 0;
  std::vector<hipEvent_t> inc_ev;  // block b (kIncBlock steps) of those rows is written

  // the incident lines' own stream (default priority): the rows of a call's
  // next chunk are computed there while the current chunk's passes run, into
  // the other of two row buffers (line_ev: rows of buffer b written; pass_ev:
  // the passes reading buffer b are done)
  hipStream_t line_stream = nullptr;
  hipEvent_t line_ev[2] = {}, pass_ev[2] = {}, start_ev = nullptr;

  int ncu = 256;  // compute units of the grid's device (the streamed pass launches a multiple of them)
  Tile tile = {0, 0, 0};
  Tile rtile = {0, 0, 0};  // the adjoint gradient's reverse sweep and recompute (0: the grid's tile)
  // Light-cone recompute of an adjoint segment ending at step cone_last: the
  // passes of step cone_last - r run over the tiles cone_tab[cone_off[r] ..
  // cone_off[r+1]), built for tile shape cone_tile (device table owned by the
  // adjoint call; null: whole-grid passes)
  const int4* cone_tab = nullptr;
  std::vector<int> cone_off;
  Tile cone_tile = {0, 0, 0};
  long cone_last = -1;
  bool finalized = false;
  bool coef_dirty = true;
  bool mats_dirty = true;
  long step = 0;
  long nswap = 0;  // role swaps of the two state copies

  // kernel arguments live in device memory (two variants per chunk: the
  // fused pass alternates read/write roles); a ring of slots lets the host
  // fill the next chunk while earlier chunks still run
  static constexpr int kRing = 8;
  static constexpr size_t kArgsBytes = (sizeof(fdtdk::StepArgs<double>) + 255) & ~size_t(255);
  char* args_dev = nullptr;
  char* args_host = nullptr;
  hipEvent_t args_ev[kRing] = {};
  bool args_ev_used[kRing] = {};
  int args_slot = 0;

  // captured chunks (fdtd_set_graph)
  bool use_graph = false;
  hipStream_t own_stream = nullptr;  // made for graph capture when the grid was on the default stream
  std::map<GraphKey, hipGraphExec_t> graphs;

  // profiling
  bool prof = false;
  std::vector<hipEvent_t> ev_pool;
  std::vector<std::pair<int, int>> ev_pairs;  // (kernel id, index of start event)
  size_t ev_used = 0;
  double kms[kNumKernels] = {};
  long kcount[kNumKernels] = {};

  ~fdtd_grid() {
    if (line_stream) {
      (void)hipStreamSynchronize(line_stream);
      (void)hipStreamDestroy(line_stream);
    }
    for (hipEvent_t e : {line_ev[0], line_ev[1], pass_ev[0], pass_ev[1], start_ev})
      if (e) (void)hipEventDestroy(e);
    for (auto& kv : graphs) (void)hipGraphExecDestroy(kv.second);
    if (own_stream) (void)hipStreamDestroy(own_stream);
    for (void* p : allocs) (void)hipFree(p);
    for (auto e : ev_pool) (void)hipEventDestroy(e);
    for (auto e : args_ev)
      if (e) (void)hipEventDestroy(e);
    if (args_host) (void)hipHostFree(args_host);
  }

  int alloc(void** p, size_t bytes, bool zero = true) {
    if (bytes == 0) bytes = 16;
    if (hipMalloc(p, bytes) != hipSuccess) {
      *p = nullptr;
      return fail(FDTD_ENOMEM, "hipMalloc(%zu) failed", bytes);
    }
    allocs.push_back(*p);
    alloc_bytes.push_back(bytes);
    if (zero && hipMemset(*p, 0, bytes) != hipSuccess) return fail(FDTD_EDEVICE, "hipMemset failed");
    return 0;
  }
  void release(void* p) {
    if (!p) return;
    auto it = std::find(allocs.begin(), allocs.end(), p);
    if (it != allocs.end()) {
      alloc_bytes.erase(alloc_bytes.begin() + (it - allocs.begin()));
      allocs.erase(it);
    }
    (void)hipFree(p);
  }
  long q(int i, int j, int k) const { return long(i) + long(d.nx) * (long(j) + long(d.ny) * k); }
};

namespace {

// Store a host-side device pointer into a kernel-argument field (those fields
// are global-address-space pointers in the device compile).
template <typename D, typename S>
inline void gput(D& dst, S src) {
  dst = (D)src;
}

// host -> device array of T from doubles
int upload_T(fdtd_grid* g, void* dst, const double* src, size_t n) {
  if (g->f64) {
    HIP_TRY(hipMemcpy(dst, src, n * 8, hipMemcpyHostToDevice));
  } else {
    std::vector<float> tmp(src, src + n);
    HIP_TRY(hipMemcpy(dst, tmp.data(), n * 4, hipMemcpyHostToDevice));
  }
  return 0;
}
int download_T(fdtd_grid* g, double* dst, const void* src, size_t n) {
  if (g->f64) {
    HIP_TRY(hipMemcpy(dst, src, n * 8, hipMemcpyDeviceToHost));
  } else {
    std::vector<float> tmp(n);
    HIP_TRY(hipMemcpy(tmp.data(), src, n * 4, hipMemcpyDeviceToHost));
    for (size_t i = 0; i < n; i++) dst[i] = tmp[i];
  }
  return 0;
}

// a 3-D grid whose table holds a magnetized or fluid Drude material (nmag is
// counted by the material setters)
static bool has_magnetized(const fdtd_grid* g) { return g->ndim == 3 && g->nmag > 0; }
static void count_magnetized(fdtd_grid* g) {
  g->nmag = 0;
  for (int id = 0; id < FDTD_MAX_MATERIALS; id++)
    g->nmag += g->mats[id].np > 0 && (g->mats[id].mag != 0 || g->mats[id].fluid != 0);
}

// Per-cell coefficients of the fluid Drude cells of the dispersive box:
// omega_p = K sqrt(ne) with K = sqrt(q_e^2/(m_e eps0)), gamma = nu_m
// (FluidDrudeStorage, DispersiveMaterials.hpp:1160-1181).  Magnetized in 3-D:
// {wp^2/dt, 1 + dt gamma + wp^2} (1297-1305); otherwise the Drude recursion's
// {(1 - b)/(1 + b), w^2/dt eps0} with b = gamma dt/2, w = dt omega_p (986-996).
template <typename T>
int upload_fluid(fdtd_grid* g) {
  g->fluid_dirty = false;
  if (g->max_np == 0 || !g->fl[0]) return 0;
  const long nb = long(g->dbox[1] - g->dbox[0]) * (g->dbox[3] - g->dbox[2]) * (g->dbox[5] - g->dbox[4]);
  std::vector<double> c0(nb, 0.0), c1(nb, 0.0);
  const double dt = g->d.dt, kval = newton_sqrt(kQe * kQe / (kMe * kEps0));
  bool any = false;
  long t = 0;
  for (int k = g->dbox[4]; k < g->dbox[5]; k++)
    for (int j = g->dbox[2]; j < g->dbox[3]; j++)
      for (int i = g->dbox[0]; i < g->dbox[1]; i++, t++) {
        const long qq = g->q(i, j, k);
        const MatHost& h = g->mats[g->mat_host[qq]];
        if (!h.np || !h.fluid) continue;
        if (g->fluid_ne.empty()) {
          g->fluid_dirty = true;
          return fail(FDTD_ESTATE, "a fluid Drude material needs fdtd_set_fluid_state");
        }
        any = true;
        const double freq = kval * std::sqrt(g->fluid_ne[qq]), gamma = g->fluid_nu[qq];
        if (h.mag && g->ndim == 3) {
          const double wp = dt * freq, vc = dt * gamma;
          c0[t] = wp * wp / dt;
          c1[t] = 1.0 + vc + wp * wp;
        } else {
          const double b = gamma * dt * 0.5, w = dt * freq;
          c0[t] = (1.0 - b) / (1.0 + b);
          c1[t] = w * w / dt * kEps0;
        }
      }
  if (!any) return 0;
  int rc = upload_T(g, g->fl[0], c0.data(), nb);
  if (!rc) rc = upload_T(g, g->fl[1], c1.data(), nb);
  return rc;
}

template <typename T>
int upload_materials(fdtd_grid* g) {
  std::vector<fdtdk::MatDev<T>> m(FDTD_MAX_MATERIALS);
  const double dt = g->d.dt;
  for (int id = 0; id < FDTD_MAX_MATERIALS; id++) {
    const MatHost& h = g->mats[id];
    m[id].inv_eps = T(1.0 / (h.eps * kEps0));
    m[id].inv_mu = T(1.0 / (h.mu * kMu0));
    m[id].eps_e0 = T(h.eps * kEps0);
    m[id].cG = T(1.0);
    m[id].np = h.np;
    m[id].mag = 0;
    m[id].fluid = h.np > 0 && h.fluid;
    m[id].w2dt = m[id].op = m[id].vq = T(0);
    for (int c = 0; c < 3; c++) m[id].v[c] = T(0);
    if (h.mag && g->ndim == 3) {
      // MagnetizedDrudeUpdate3D::implicit_update, DispersiveMaterials.hpp:1297-1305
      // (a fluid material's wp and gamma are per cell: upload_fluid)
      const double wp = dt * h.freq[0], vc = dt * h.gamma[0], op = 1.0 + vc + wp * wp;
      m[id].mag = h.mag;
      m[id].w2dt = T(wp * wp / dt);
      m[id].op = T(op);
      m[id].vq = T(dt * kQe / kMe);
      for (int c = 0; c < 3; c++) m[id].v[c] = T(dt * h.bias[c] * kQe / kMe);
    }
    for (int p = 0; p < fdtdk::kMaxPoles; p++) {
      double cJ = 0, cE = 0, cP = 0, cD = 0;
      if (p < h.np && h.kind[p] == FDTD_CONDUCTIVE) {
        // ConductiveUpdate::implicit_update, DispersiveMaterials.hpp:549-560:
        // factor = dt cond_freq/(eps_rel eps0); J is not touched
        cD = dt * h.freq[p] / (h.eps * kEps0);
        cJ = 1.0;
      } else if (p < h.np && h.kind[p] == FDTD_DRUDE) {
        // DrudeUpdate::implicit_update, DispersiveMaterials.hpp:986-996 (also a
        // magnetized Drude material in TM, whose update the reference routes
        // through DrudeUpdate per component: MagnetizedDrudeCall, 1508-1537)
        double b = h.gamma[p] * dt * 0.5, w = dt * h.freq[p];
        cJ = (1.0 - b) / (1.0 + b);
        cE = w * w / dt * (h.weight[p] * kEps0);
      } else if (p < h.np) {
        // LorentzUpdate::implicit_update, DispersiveMaterials.hpp:719-732
        double b = 1.0 + h.gamma[p] * dt, bm = 1.0 - h.gamma[p] * dt, w = h.freq[p] * dt;
        cJ = bm / b;
        cE = h.freq[p] * w / b * (h.weight[p] * kEps0);
        cP = -(h.freq[p] * w / b);
      }
      m[id].cJ[p] = T(cJ);
      m[id].cJm[p] = T(p < h.np ? (h.kind[p] == FDTD_DRUDE ? 2.0 * (h.gamma[p] * dt * 0.5) / (1.0 + h.gamma[p] * dt * 0.5)
                                   : h.kind[p] == FDTD_CONDUCTIVE ? 0.0
                                                                  : 2.0 * h.gamma[p] * dt / (1.0 + h.gamma[p] * dt))
                                : 1.0 - cJ);
      m[id].cE[p] = T(cE);
      m[id].cP[p] = T(cP);
      m[id].cD[p] = T(cD);
      if (p == 0) m[id].cG = T(1.0 / (1.0 + cD));
    }
    // the magnetic side: the same recursions on B, H, M, K with mu_r mu0
    m[id].mnp = h.mnp;
    m[id].mu_m0 = T(h.mu * kMu0);
    m[id].mcG = T(1.0);
    for (int p = 0; p < fdtdk::kMaxPoles; p++) {
      double cJ = 0, cE = 0, cP = 0, cD = 0, cJm = 1.0;
      if (p < h.mnp && h.mkind[p] == FDTD_CONDUCTIVE) {
        cD = dt * h.mfreq[p] / (h.mu * kMu0);
        cJ = 1.0;
        cJm = 0.0;
      } else if (p < h.mnp && h.mkind[p] == FDTD_DRUDE) {
        const double b = h.mgamma[p] * dt * 0.5, w = dt * h.mfreq[p];
        cJ = (1.0 - b) / (1.0 + b);
        cJm = 2.0 * b / (1.0 + b);
        cE = w * w / dt * (h.mweight[p] * kMu0);
      } else if (p < h.mnp) {
        const double b = 1.0 + h.mgamma[p] * dt, bm = 1.0 - h.mgamma[p] * dt, w = h.mfreq[p] * dt;
        cJ = bm / b;
        cJm = 2.0 * h.mgamma[p] * dt / b;
        cE = h.mfreq[p] * w / b * (h.mweight[p] * kMu0);
        cP = -(h.mfreq[p] * w / b);
      }
      m[id].mcJ[p] = T(cJ);
      m[id].mcJm[p] = T(cJm);
      m[id].mcE[p] = T(cE);
      m[id].mcP[p] = T(cP);
      m[id].mcD[p] = T(cD);
      if (p == 0) m[id].mcG = T(1.0 / (1.0 + cD));
    }
  }
  HIP_TRY(hipMemcpy(g->mats_dev, m.data(), sizeof(fdtdk::MatDev<T>) * m.size(), hipMemcpyHostToDevice));
  std::vector<T> ie(FDTD_MAX_MATERIALS), im(FDTD_MAX_MATERIALS);
  for (int id = 0; id < FDTD_MAX_MATERIALS; id++) {
    ie[id] = g->mats[id].np > 0 ? -m[id].inv_eps : m[id].inv_eps;
    im[id] = g->mats[id].mnp > 0 ? -m[id].inv_mu : m[id].inv_mu;  // negative: a magnetic-pole material
  }
  HIP_TRY(hipMemcpy(g->mie_dev, ie.data(), sizeof(T) * ie.size(), hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(g->mim_dev, im.data(), sizeof(T) * im.size(), hipMemcpyHostToDevice));
  g->mats_dirty = false;
  return 0;
}

// per-axis device coefficients: f = dt F/dx, g = dt G, b = B, c = C/dx
int upload_coefs(fdtd_grid* g) {
  const double dt = g->d.dt, dx = g->d.dx;
  for (int ft = 0; ft < 2; ft++)
    for (int a = 0; a < 3; a++) {
      const int L = g->len[a];
      std::vector<double> f(L), gg(L), b(L), c(L);
      for (int i = 0; i < L; i++) {
        const double* r = &g->pml[ft][a][4 * i];
        f[i] = dt * r[2] / dx;
        gg[i] = dt * r[3];
        b[i] = r[0];
        c[i] = r[1] / dx;
      }
      const std::vector<double>* src[4] = {&f, &gg, &b, &c};
      for (int w = 0; w < 4; w++) {
        int rc = upload_T(g, g->coef[ft][a][w], src[w]->data(), L);
        if (rc) return rc;
      }
    }
  g->coef_dirty = false;
  return 0;
}

int ensure_pingpong(fdtd_grid* g);
int ensure_dft_im(fdtd_grid* g);

// CPML slabs: psi is kept only where some C coefficient of the axis is non-zero
// (elsewhere psi stays exactly zero in the reference as well).
int finalize(fdtd_grid* g) {
  for (int a = 0; a < g->ndim; a++) {
    const int L = g->len[a];
    int lo = 0, hi = L;
    for (int i = 0; i < L; i++) {
      bool nz = g->pml[0][a][4 * i + 1] != 0.0 || g->pml[1][a][4 * i + 1] != 0.0;
      if (!nz) continue;
      if (i < L / 2)
        lo = std::max(lo, i + 1);
      else
        hi = std::min(hi, i);
    }
    if (lo > hi) lo = hi;
    g->slab_a[a] = lo;
    g->slab_b[a] = hi;
    g->slab_S[a] = lo + (L - hi);
  }
  for (int ft = 0; ft < 2; ft++)
    for (int I = 0; I < 3; I++)
      for (int J = 0; J < g->ndim; J++) {
        bool act = ft == 0 ? g->actE[I] : g->actH[I];
        if (!act || I == J || g->slab_S[J] == 0) continue;
        long n = long(g->slab_S[J]);
        for (int a = 0; a < 3; a++)
          if (a != J) n *= g->len[a];
        int rc = g->alloc(&g->psi[ft][I][J], size_t(n) * g->esz);
        if (rc) return rc;
      }
  // dispersive box: bounding box of interior cells whose material has poles
  int mx = 0;
  for (int id = 0; id < FDTD_MAX_MATERIALS; id++) mx = std::max(mx, g->mats[id].np);
  g->max_np = mx;
  if (mx > 0) {
    int b[6] = {g->len[0], -1, g->len[1], -1, g->len[2], -1};
    const int nz = g->ndim == 3 ? g->d.nz : 1;
    for (int k = 0; k < nz; k++)
      for (int j = 0; j < g->d.ny; j++)
        for (int i = 0; i < g->d.nx; i++)
          if (g->mats[g->mat_host[g->q(i, j, k)]].np > 0) {
            b[0] = std::min(b[0], i);
            b[1] = std::max(b[1], i);
            b[2] = std::min(b[2], j);
            b[3] = std::max(b[3], j);
            b[4] = std::min(b[4], k);
            b[5] = std::max(b[5], k);
          }
    if (b[1] < 0) {
      g->max_np = 0;
    } else {
      g->dbox[0] = b[0];
      g->dbox[1] = b[1] + 1;
      g->dbox[2] = b[2];
      g->dbox[3] = b[3] + 1;
      g->dbox[4] = b[4];
      g->dbox[5] = b[5] + 1;
      long nb = long(g->dbox[1] - g->dbox[0]) * (g->dbox[3] - g->dbox[2]) * (g->dbox[5] - g->dbox[4]);
      // D = eps_r eps0 E from the current E (zero unless uploaded)
      for (int c = 0; c < 3; c++) {
        if (!g->actE[c]) continue;
        int rc = g->alloc(&g->Dd[c], size_t(nb) * g->esz);
        if (rc) return rc;
        std::vector<double> e(g->ncell), dbx(nb);
        rc = download_T(g, e.data(), g->E[c], g->ncell);
        if (rc) return rc;
        long t = 0;
        for (int k = g->dbox[4]; k < g->dbox[5]; k++)
          for (int j = g->dbox[2]; j < g->dbox[3]; j++)
            for (int i = g->dbox[0]; i < g->dbox[1]; i++, t++) {
              long qq = g->q(i, j, k);
              dbx[t] = g->mats[g->mat_host[qq]].eps * kEps0 * e[qq];
            }
        rc = upload_T(g, g->Dd[c], dbx.data(), nb);
        if (rc) return rc;
        for (int p = 0; p < mx; p++) {
          if ((rc = g->alloc(&g->Pd[p][c], size_t(nb) * g->esz))) return rc;
          if ((rc = g->alloc(&g->Jd[p][c], size_t(nb) * g->esz))) return rc;
        }
      }
      for (int w = 0; w < 2; w++) {
        int rc = g->alloc(&g->fl[w], size_t(nb) * g->esz);
        if (rc) return rc;
      }
      g->fluid_dirty = true;
    }
  }
  // magnetic-pole box: bounding box of interior cells whose material has
  // magnetic poles; B = mu_r mu0 H there from the current H
  int mmx = 0;
  for (int id = 0; id < FDTD_MAX_MATERIALS; id++) mmx = std::max(mmx, g->mats[id].mnp);
  g->max_mnp = 0;
  if (mmx > 0) {
    int b[6] = {g->len[0], -1, g->len[1], -1, g->len[2], -1};
    const int nz = g->ndim == 3 ? g->d.nz : 1;
    for (int k = 0; k < nz; k++)
      for (int j = 0; j < g->d.ny; j++)
        for (int i = 0; i < g->d.nx; i++)
          if (g->mats[g->mat_host[g->q(i, j, k)]].mnp > 0) {
            b[0] = std::min(b[0], i);
            b[1] = std::max(b[1], i);
            b[2] = std::min(b[2], j);
            b[3] = std::max(b[3], j);
            b[4] = std::min(b[4], k);
            b[5] = std::max(b[5], k);
          }
    if (b[1] >= 0) {
      if (g->cplx) return fail(FDTD_EINVAL, "magnetic poles are not stepped on complex (Bloch-periodic) grids");
      for (int id = 0; id < FDTD_MAX_MATERIALS; id++)
        if (g->mats[id].np > 0 && g->mats[id].mag == 2 && g->ndim == 3)
          return fail(FDTD_EINVAL, "a self-magnetized Drude material (the cell's own B = mu_r mu0 H) cannot share a "
                                   "grid with magnetic poles");
      g->max_mnp = mmx;
      for (int x = 0; x < 3; x++) {
        g->mbox[2 * x] = b[2 * x];
        g->mbox[2 * x + 1] = b[2 * x + 1] + 1;
      }
      const long nb = long(g->mbox[1] - g->mbox[0]) * (g->mbox[3] - g->mbox[2]) * (g->mbox[5] - g->mbox[4]);
      for (int c = 0; c < 3; c++) {
        if (!g->actH[c]) continue;
        int rc = g->alloc(&g->Bd[c], size_t(nb) * g->esz);
        if (rc) return rc;
        std::vector<double> h(g->ncell), bb(nb);
        if ((rc = download_T(g, h.data(), g->H[c], g->ncell))) return rc;
        long t = 0;
        for (int k = g->mbox[4]; k < g->mbox[5]; k++)
          for (int j = g->mbox[2]; j < g->mbox[3]; j++)
            for (int i = g->mbox[0]; i < g->mbox[1]; i++, t++) {
              const long qq = g->q(i, j, k);
              bb[t] = g->mats[g->mat_host[qq]].mu * kMu0 * h[qq];
            }
        if ((rc = upload_T(g, g->Bd[c], bb.data(), nb))) return rc;
        for (int p = 0; p < mmx; p++) {
          if ((rc = g->alloc(&g->Md[p][c], size_t(nb) * g->esz))) return rc;
          if ((rc = g->alloc(&g->Kd[p][c], size_t(nb) * g->esz))) return rc;
        }
      }
    }
  }
  if (g->cplx) {
    int rc = ensure_pingpong(g);  // the imaginary part's state
    if (!rc) rc = ensure_dft_im(g);
    if (rc) return rc;
  }
  g->finalized = true;
  return 0;
}

// the state pointers: read from the current buffers, write to the same ones
// (in-place passes) or to the second copies (fused pass)
template <typename T>
void set_state(fdtd_grid* g, fdtdk::StepArgs<T>& a, bool pingpong) {
  for (int c = 0; c < 3; c++) {
    gput(a.E[c], (T*)g->E[c]);
    gput(a.H[c], (T*)g->H[c]);
    gput(a.Eo[c], (T*)(pingpong ? g->E2[c] : g->E[c]));
    gput(a.Ho[c], (T*)(pingpong ? g->H2[c] : g->H[c]));
    gput(a.Dd[c], (T*)g->Dd[c]);
    gput(a.Ddo[c], (T*)(pingpong ? g->Dd2[c] : g->Dd[c]));
    for (int p = 0; p < fdtdk::kMaxPoles; p++) {
      gput(a.Pd[p][c], (T*)g->Pd[p][c]);
      gput(a.Jd[p][c], (T*)g->Jd[p][c]);
      gput(a.Pdo[p][c], (T*)(pingpong ? g->Pd2[p][c] : g->Pd[p][c]));
      gput(a.Jdo[p][c], (T*)(pingpong ? g->Jd2[p][c] : g->Jd[p][c]));
      gput(a.Md[p][c], (T*)g->Md[p][c]);
      gput(a.Kd[p][c], (T*)g->Kd[p][c]);
    }
    gput(a.Bd[c], (T*)g->Bd[c]);
  }
  for (int x = 0; x < 6; x++) a.mbox[x] = g->mbox[x];
  for (int ft = 0; ft < 2; ft++)
    for (int I = 0; I < 3; I++)
      for (int J = 0; J < 3; J++) {
        gput(a.psi[ft][I][J], (T*)g->psi[ft][I][J]);
        gput(a.psio[ft][I][J], (T*)(pingpong ? g->psi2[ft][I][J] : g->psi[ft][I][J]));
      }
}

void swap_state(fdtd_grid* g) {
  g->nswap++;
  for (int c = 0; c < 3; c++) {
    std::swap(g->E[c], g->E2[c]);
    std::swap(g->H[c], g->H2[c]);
    std::swap(g->Dd[c], g->Dd2[c]);
    for (int p = 0; p < FDTD_MAX_POLES; p++) {
      std::swap(g->Pd[p][c], g->Pd2[p][c]);
      std::swap(g->Jd[p][c], g->Jd2[p][c]);
    }
  }
  for (int ft = 0; ft < 2; ft++)
    for (int I = 0; I < 3; I++)
      for (int J = 0; J < 3; J++) std::swap(g->psi[ft][I][J], g->psi2[ft][I][J]);
}

// allocate the second copies (zero: the fused pass writes every interior cell,
// the outer layer stays zero in both copies)
int ensure_pingpong(fdtd_grid* g) {
  if (g->pingpong) return 0;
  auto twin = [&](void* src, void** dst) -> int {
    if (!src) return 0;
    auto it = std::find(g->allocs.begin(), g->allocs.end(), src);
    if (it == g->allocs.end()) return fail(FDTD_ESTATE, "untracked allocation");
    return g->alloc(dst, g->alloc_bytes[it - g->allocs.begin()]);
  };
  int rc = 0;
  for (int c = 0; c < 3 && !rc; c++) {
    if (!rc) rc = twin(g->E[c], &g->E2[c]);
    if (!rc) rc = twin(g->H[c], &g->H2[c]);
    if (!rc) rc = twin(g->Dd[c], &g->Dd2[c]);
    for (int p = 0; p < FDTD_MAX_POLES && !rc; p++) {
      rc = twin(g->Pd[p][c], &g->Pd2[p][c]);
      if (!rc) rc = twin(g->Jd[p][c], &g->Jd2[p][c]);
    }
  }
  for (int ft = 0; ft < 2 && !rc; ft++)
    for (int I = 0; I < 3 && !rc; I++)
      for (int J = 0; J < 3 && !rc; J++) rc = twin(g->psi[ft][I][J], &g->psi2[ft][I][J]);
  if (rc) return rc;
  g->pingpong = true;
  return 0;
}

// Boundary faces for the passes (BoundaryUpdates.hpp:258-300 and the masks of
// 527-722): source layer 1 / L-2 (periodic: L-2 for the min ghost, 1 for the
// max ghost), ghost layer 0 / L-1, and the reference's per-component signs
// (+-1 for the component along the face normal / across it): PEC E +-, H -+;
// PMC E -+, H +-; Symmetric -+ for both; Antisymmetric +- for both; Periodic ++.
template <typename T>
void bc_args(const fdtd_grid* g, fdtdk::StepArgs<T>& a) {
  const long stride[3] = {1, long(g->d.nx), long(g->d.nx) * g->d.ny};
  a.bc_any = g->nbc > 0;
  for (int ax = 0; ax < 3; ax++)
    for (int sd = 0; sd < 2; sd++) {
      const int kind = ax < g->ndim ? g->bc[ax][sd] : 0;
      a.bc_src[ax][sd] = -1;
      a.bc_off[ax][sd] = 0;
      a.bc_neg[0][ax][sd] = a.bc_neg[1][ax][sd] = 0;
      if (!kind || kind == FDTD_BC_BLOCH) continue;  // Bloch faces: bloch_ghosts
      const int L = g->len[ax];
      const int src = kind == FDTD_BC_PERIODIC ? (sd ? 1 : L - 2) : (sd ? L - 2 : 1);
      const int dst = sd ? L - 1 : 0;
      a.bc_src[ax][sd] = src;
      a.bc_off[ax][sd] = long(dst - src) * stride[ax];
      for (int ft = 0; ft < 2; ft++)
        for (int q = 0; q < 3; q++) {
          const bool along = q == ax;
          bool neg = false;
          if (kind == FDTD_BC_PEC) neg = ft == 0 ? !along : along;
          if (kind == FDTD_BC_PMC) neg = ft == 0 ? along : !along;
          if (kind == FDTD_BC_SYMMETRIC) neg = along;
          if (kind == FDTD_BC_ANTISYMMETRIC) neg = !along;
          if (neg) a.bc_neg[ft][ax][sd] |= 1 << q;
        }
    }
}

// The imaginary part's arguments (complex grids): the real part's with the
// state, DFT and probe buffers of the imaginary part, and no sources or TFSF
// (the excitation is real).
template <typename T>
void im_args(const fdtd_grid* g, const fdtdk::StepArgs<T>& re, fdtdk::StepArgs<T>& a) {
  a = re;
  for (int c = 0; c < 3; c++) {
    gput(a.E[c], (T*)g->E2[c]);
    gput(a.Eo[c], (T*)g->E2[c]);
    gput(a.H[c], (T*)g->H2[c]);
    gput(a.Ho[c], (T*)g->H2[c]);
    gput(a.Dd[c], (T*)g->Dd2[c]);
    gput(a.Ddo[c], (T*)g->Dd2[c]);
    for (int p = 0; p < fdtdk::kMaxPoles; p++) {
      gput(a.Pd[p][c], (T*)g->Pd2[p][c]);
      gput(a.Pdo[p][c], (T*)g->Pd2[p][c]);
      gput(a.Jd[p][c], (T*)g->Jd2[p][c]);
      gput(a.Jdo[p][c], (T*)g->Jd2[p][c]);
    }
  }
  for (int ft = 0; ft < 2; ft++)
    for (int I = 0; I < 3; I++)
      for (int J = 0; J < 3; J++) {
        gput(a.psi[ft][I][J], (T*)g->psi2[ft][I][J]);
        gput(a.psio[ft][I][J], (T*)g->psi2[ft][I][J]);
      }
  a.nsrc = 0;
  a.nsrc_eh = 0;
  a.tfsf = 0;
  a.nplane = 0;
  for (int q = 0; q < g->ndft; q++) {
    gput(a.dft[q].acc, (double*)g->dft[q].acc_im);
    gput(a.dft[q].smp, (T*)g->dft[q].smp_im);
  }
  gput(a.trace, g->trace_im);
}

// the Bloch faces of a complex grid
fdtdk::BlochFaces bloch_faces(const fdtd_grid* g, long* maxcells) {
  fdtdk::BlochFaces bf{};
  bf.len[0] = g->d.nx;
  bf.len[1] = g->d.ny;
  bf.len[2] = g->ndim == 3 ? g->d.nz : 1;
  *maxcells = 0;
  for (int ax = 0; ax < g->ndim; ax++)
    for (int sd = 0; sd < 2; sd++) {
      if (g->bc[ax][sd] != FDTD_BC_BLOCH) continue;
      const int L = g->len[ax];
      fdtdk::BlochFace& f = bf.f[bf.n++];
      f.axis = ax;
      f.src = sd ? 1 : L - 2;  // the periodic image
      f.dst = sd ? L - 1 : 0;
      f.fr = g->bloch[ax][sd][0];
      f.fi = g->bloch[ax][sd][1];
      const int t1 = ax == 0 ? 1 : 0, t2 = ax == 2 ? 1 : 2;
      const long n = long(bf.len[t1] - 2) * (bf.len[t2] > 1 ? bf.len[t2] - 2 : 1);
      *maxcells = std::max(*maxcells, n);
    }
  return bf;
}

template <typename T>
void fill_args(fdtd_grid* g, fdtdk::StepArgs<T>& a) {
  memset(&a, 0, sizeof a);
  set_state(g, a, false);
  gput(a.mat, g->mat);
  gput(a.mats, (const fdtdk::MatDev<T>*)g->mats_dev);
  gput(a.mie, (const T*)g->mie_dev);
  gput(a.mim, (const T*)g->mim_dev);
  for (int ft = 0; ft < 2; ft++)
    for (int I = 0; I < 3; I++) {
      gput(a.cf[ft][I], (const T*)g->coef[ft][I][0]);
      gput(a.cg[ft][I], (const T*)g->coef[ft][I][1]);
      gput(a.cb[ft][I], (const T*)g->coef[ft][I][2]);
      gput(a.cc[ft][I], (const T*)g->coef[ft][I][3]);
    }
  for (int x = 0; x < 3; x++) {
    a.slab_a[x] = g->ndim > x ? g->slab_a[x] : 0;
    a.slab_b[x] = g->ndim > x ? g->slab_b[x] : 1 << 30;
    a.slab_S[x] = g->slab_S[x];
  }
  a.nx = g->d.nx;
  a.ny = g->d.ny;
  a.nz = g->ndim == 3 ? g->d.nz : 1;
  a.sy = g->d.nx;
  a.sz = long(g->d.nx) * g->d.ny;
  for (int x = 0; x < 6; x++) a.dbox[x] = g->dbox[x];
  gput(a.fl[0], (const T*)g->fl[0]);
  gput(a.fl[1], (const T*)g->fl[1]);
  a.dt = T(g->d.dt);
  a.nsrc = g->nsrc;
  a.nsrc_eh = g->nsrc_eh;
  for (int s = 0; s < g->nsrc; s++) a.src[s] = g->src[s];
  a.tfsf = g->tfsf;
  for (int x = 0; x < 6; x++) a.box[x] = g->box[x];
  bc_args(g, a);
  for (int x = 0; x < 3; x++) {
    a.kvec[x] = g->kvec[x];
    a.pol[x] = g->pol[x];
    a.polH[x] = g->polH[x];
    a.origin[x] = g->origin[x];
  }
  a.dx = g->d.dx;
  a.auxL = g->aux.L;
  a.nplane = g->nplane;
  for (int p = 0; p < g->nplane; p++) {
    a.plane[p] = g->plane[p];
    gput(a.plane[p].incE, g->plane_aux[p].incE);
    gput(a.plane[p].incH, g->plane_aux[p].incH);
  }
  a.ndft = g->ndft;
  for (int q = 0; q < g->ndft; q++) {
    const auto& h = g->dft[q];
    auto& dd = a.dft[q];
    for (int x = 0; x < 6; x++) dd.r[x] = h.r[x];
    dd.mask = h.mask;
    dd.nfreq = h.nfreq;
    dd.ncomp = h.ncomp;
    dd.ncells = h.ncells;
    gput(dd.acc, (double*)h.acc);
    gput(dd.phE, h.phE);
    gput(dd.phH, h.phH);
    dd.nphase = h.nphase;
    dd.base = a.dft_work;
    a.dft_work += long(h.nfreq) * h.ncells;
    dd.cbase = a.dft_cells;
    a.dft_cells += fdtdk::fold_segment(h.ncells);  // whole 256-thread blocks per monitor
    dd.sbase = a.dft_slots;
    a.dft_slots += long(h.ncomp) * fdtdk::fold_segment(h.ncells);
    gput(dd.smp, (T*)h.smp);
  }
  a.nprobe = g->nprobe;
  for (int p = 0; p < g->nprobe; p++) a.probe[p] = g->probe[p];
  gput(a.trace, g->trace);
  // byte sizes of one (padded) field array and of the id volume, for the
  // buffer-addressed streamed pass (0 when they do not fit 31 bits)
  const size_t fb = size_t(g->ncell + fdtdk::field_pad(g->d.nx)) * sizeof(T);
  a.field_bytes = fb < (size_t(1) << 31) ? unsigned(fb) : 0u;
  a.mat_bytes = a.field_bytes ? unsigned(g->ncell + fdtdk::field_pad(g->d.nx)) : 0u;
  gput(a.dsg_ie, (const T*)g->dsg_ie_dev);
  for (int x = 0; x < 6; x++) a.dsg[x] = g->dsg[x];
  gput(a.hist, (T*)g->hist);
  a.hist_step0 = g->hist_step0;
  a.hist_rows = g->hist_rows;
  if (g->recompute) {
    a.ndft = 0;
    a.dft_work = 0;
    a.dft_cells = 0;
    a.dft_slots = 0;
    a.nprobe = 0;
  }
}

hipEvent_t next_event(fdtd_grid* g) {
  if (g->ev_used == g->ev_pool.size()) {
    hipEvent_t e;
    (void)hipEventCreate(&e);
    g->ev_pool.push_back(e);
  }
  return g->ev_pool[g->ev_used++];
}

int collect_events(fdtd_grid* g) {
  if (!g->prof || g->ev_pairs.empty()) return 0;
  for (auto& pr : g->ev_pairs) {
    float ms = 0;
    // the pair may sit on another stream than the grid's (the adjoint's recompute stream)
    HIP_TRY(hipEventSynchronize(g->ev_pool[pr.second + 1]));
    HIP_TRY(hipEventElapsedTime(&ms, g->ev_pool[pr.second], g->ev_pool[pr.second + 1]));
    g->kms[pr.first] += ms;
    g->kcount[pr.first] += 1;
  }
  g->ev_pairs.clear();
  g->ev_used = 0;
  return 0;
}

template <typename K, typename... A>
int launch(fdtd_grid* g, int kid, K kern, dim3 grid, dim3 block, const A&... args) {
  if (g->prof) {
    if (g->ev_used + 2 > 4096) {
      int rc = collect_events(g);
      if (rc) return rc;
    }
    size_t s = g->ev_used;
    hipEvent_t e0 = next_event(g), e1 = next_event(g);
    HIP_TRY(hipEventRecord(e0, g->stream));
    hipLaunchKernelGGL(kern, grid, block, 0, g->stream, args...);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipEventRecord(e1, g->stream));
    g->ev_pairs.push_back({kid, int(s)});
    return 0;
  }
  hipLaunchKernelGGL(kern, grid, block, 0, g->stream, args...);
  HIP_TRY(hipGetLastError());
  return 0;
}

// grow the probe trace so that it holds `steps` rows
int ensure_trace(fdtd_grid* g, long steps) {
  if (!g->nprobe || g->trace_cap >= steps) return 0;
  long cap = std::max<long>(steps, 2 * g->trace_cap);
  for (int part = 0; part < (g->cplx ? 2 : 1); part++) {
    double*& tr = part ? g->trace_im : g->trace;
    double* nt = nullptr;
    int rc = g->alloc((void**)&nt, size_t(cap) * g->nprobe * 8);
    if (rc) return rc;
    if (tr) {
      HIP_TRY(hipMemcpyAsync(nt, tr, size_t(g->step) * g->nprobe * 8, hipMemcpyDeviceToDevice, g->stream));
      HIP_TRY(hipStreamSynchronize(g->stream));
      g->release(tr);
    }
    tr = nt;
  }
  g->trace_cap = cap;
  return 0;
}

// the imaginary part's DFT accumulators and sample rings (complex grids)
int ensure_dft_im(fdtd_grid* g) {
  if (!g->cplx) return 0;
  for (int q = 0; q < g->ndft; q++) {
    auto& d = g->dft[q];
    if (d.acc_im) continue;
    int rc = g->alloc(&d.acc_im, size_t(d.ncomp) * d.nfreq * 2 * d.ncells * 8);
    if (!rc) rc = g->alloc(&d.smp_im, size_t(fdtdk::kDftBlock) * d.ncomp * d.ncells * g->esz);
    if (rc) return rc;
  }
  return 0;
}

// advance one incident line nsteps from step0, its rows into outE / outH, on
// stream s (the grid's stream by default)
int run_aux_line(fdtd_grid* g, const AuxLine& l, double* outE, double* outH, long step0, int nsteps,
                 hipStream_t s = nullptr, bool on_grid = true) {
  fdtdk::AuxArgs a;
  a.L = l.L;
  a.npml = 15;
  a.dt = g->d.dt;
  a.dx = g->d.dx;
  a.eps0 = kEps0;
  a.mu0 = kMu0;
  double* st = l.state;
  a.Dz = st;
  a.Ez = st + l.L;
  a.By = st + 2 * l.L;
  a.Hy = st + 3 * l.L;
  a.psiE = st + 4 * l.L;
  a.psiH = st + 5 * l.L;
  a.pE = l.pE;
  a.pH = l.pH;
  a.signal = l.sig;
  a.nsignal = l.nsig;
  a.step0 = step0;
  a.nsteps = nsteps;
  a.outE = outE;
  a.outH = outH;
  // cells per thread: the fewest that fit the line in one workgroup
  int cpt = 1;
  while (cpt < 8 && (l.L + cpt - 1) / cpt > fdtdk::kAuxThreads) cpt *= 2;
  if ((l.L + cpt - 1) / cpt > fdtdk::kAuxThreads || 2 * a.npml + 5 > fdtdk::kAuxPmlMax)
    return fail(FDTD_EINVAL, "incident line of %d cells exceeds the line kernel", l.L);
  const dim3 block(unsigned(((l.L + cpt - 1) / cpt + 63) / 64 * 64));
  auto go = [&](auto kern) -> int {
    if (on_grid) return launch(g, 2, kern, dim3(1), block, a);  // timed with the grid's kernels
    hipLaunchKernelGGL(kern, dim3(1), block, 0, s, a);
    HIP_TRY(hipGetLastError());
    return 0;
  };
  switch (cpt) {
    case 1: return go(fdtdk::aux_line<1>);
    case 2: return go(fdtdk::aux_line<2>);
    case 4: return go(fdtdk::aux_line<4>);
    default: return go(fdtdk::aux_line<8>);
  }
}

// the incident lines' stream and its events, made on first use
int ensure_line_stream(fdtd_grid* g) {
  if (g->line_stream) return 0;
  HIP_TRY(hipStreamCreateWithFlags(&g->line_stream, hipStreamNonBlocking));
  for (int b = 0; b < 2; b++) {
    HIP_TRY(hipEventCreateWithFlags(&g->line_ev[b], hipEventDisableTiming));
    HIP_TRY(hipEventCreateWithFlags(&g->pass_ev[b], hipEventDisableTiming));
  }
  HIP_TRY(hipEventCreateWithFlags(&g->start_ev, hipEventDisableTiming));
  return 0;
}

// The stream the lines run on: their own, or the grid's while kernels are
// timed (fdtd_set_profiling), so that every launch is attributed in order.
hipStream_t lines_on(const fdtd_grid* g) { return g->prof ? g->stream : g->line_stream; }

// rows of steps [step0, step0 + n) of every incident line into row buffer b,
// on the lines' stream; line_ev[b] marks them written
int line_rows(fdtd_grid* g, int b, long step0, int n) {
  hipStream_t ls = lines_on(g);
  const bool timed = ls == g->stream;
  int rc = 0;
  if (g->tfsf)
    rc = run_aux_line(g, g->aux, g->aux.incE + size_t(b) * kChunk * g->aux.L,
                      g->aux.incH + size_t(b) * kChunk * g->aux.L, step0, n, ls, timed);
  for (int p = 0; p < g->nplane && !rc; p++) {
    const AuxLine& l = g->plane_aux[p];
    rc = run_aux_line(g, l, l.incE + size_t(b) * kChunk * l.L, l.incH + size_t(b) * kChunk * l.L, step0, n, ls,
                      timed);
  }
  if (rc) return rc;
  HIP_TRY(hipEventRecord(g->line_ev[b], ls));
  return 0;
}

// the chunk's rows of every incident line of the grid (the box's and the planes')
int run_aux(fdtd_grid* g, long step0, int nsteps) {
  int rc = g->tfsf ? run_aux_line(g, g->aux, g->aux.incE, g->aux.incH, step0, nsteps) : 0;
  for (int p = 0; p < g->nplane && !rc; p++)
    rc = run_aux_line(g, g->plane_aux[p], g->plane_aux[p].incE, g->plane_aux[p].incH, step0, nsteps);
  return rc;
}

template <typename T>
int step_impl(fdtd_grid* g, int nsteps) {
  if (g->cplx)
    for (int id = 0; id < FDTD_MAX_MATERIALS; id++)
      if (g->mats[id].np > 0 && g->mats[id].mag == 2 && g->ndim == 3)
        return fail(FDTD_EINVAL,
                    "a self-magnetized material is nonlinear in the field: complex (Bloch) grids cannot hold one");
  if (g->mats_dirty) {
    int rc = upload_materials<T>(g);
    if (rc) return rc;
  }
  if (g->fluid_dirty) {
    int rc = upload_fluid<T>(g);
    if (rc) return rc;
  }
  if (g->coef_dirty) {
    int rc = upload_coefs(g);
    if (rc) return rc;
  }
  int rc0 = ensure_trace(g, g->step + nsteps);
  if (rc0) return rc0;
  fdtdk::StepArgs<T> a;
  fill_args<T>(g, a);
  Tile t = g->tile.by ? g->tile : (g->ndim == 3 ? kDefaultTile3D : kDefaultTileTM);
  // design increments and boundary ghost writes are made by the two-pass kernels
  // (E / H sources and single planes too: the general instance of the passes)
  // (and the H update of magnetic-pole cells)
  const bool general = g->nbc || g->nsrc_eh || g->nplane || g->max_mnp > 0;
  // the time-blocked TM pass covers non-dispersive grids without those, complex
  // fields or DFT monitors
  const bool te = g->d.mode == FDTD_TE, tem = g->ndim == 1;
  const bool tmb = g->ndim == 2 && !te && t.bx == 0 && !general && !g->cplx && g->max_np == 0 && g->ndft == 0 &&
                   !g->force_two_pass;
  if (g->ndim == 2 && !tmb) {
    if (t.bx == 0) t = kDefaultTileTM;
    t.kc = 1;
  }
  if ((g->force_two_pass || general || has_magnetized(g)) && t.bx <= 2 && g->ndim == 3) t = kDefaultTile3D;
  // the time-blocked 3-D pass covers grids without dispersive cells or complex
  // fields whose DFT monitors lie inside the interior (it samples the cells
  // it updates)
  if (g->ndim == 3 && t.bx == 2) {
    bool inner = true;
    for (int q = 0; q < g->ndft; q++) {
      const int* r = g->dft[q].r;
      inner = inner && r[0] >= 1 && r[2] >= 1 && r[4] >= 1 && r[1] <= g->d.nx - 1 && r[3] <= g->d.ny - 1 &&
              r[5] <= g->d.nz - 1;
    }
    if (g->cplx || g->max_np > 0 || !inner || g->design || (g->tfsf && g->aux.L > fdtdk::kTbRowMax))
      t = kDefaultTile3D;
    else if (sizeof(T) == 8) {
      t.by = 8;  // FP64: 256 registers per lane at 8 rows (at 12 they would spill)
      t.kc = std::min(t.kc, fdtdk::kTbMaxKc64);  // the chunk's TFSF tables in LDS
    } else if (t.by == 16) {
      t.by = 12;  // 16 rows would spill at 128 registers per lane
    }
  }
  // the streamed pass is FP32, addresses fields with 32-bit buffer offsets and
  // keeps the z coefficients in LDS
  if (g->ndim == 3 && t.bx == 0 && (sizeof(T) != 4 || a.field_bytes == 0 || g->d.nz > fdtdk::kSMaxZ || g->design))
    t = kDefaultTile3D;

  const bool wavef = g->ndim == 3 && t.bx == 1;
  const bool fused = g->ndim == 3 && t.bx <= 1;
  const bool tb3 = g->ndim == 3 && t.bx == 2;
  // the fused and time-blocked passes write the second copy of the state
  const bool pp = fused || tmb || tb3;
  if (pp) {
    int rc = ensure_pingpong(g);
    if (rc) return rc;
  }
  const int nzi = g->ndim == 3 ? g->d.nz - 2 : 1;
  long units = 0;  // streamed pass: (tile, plane) units
  const bool graphs = g->use_graph && !g->prof && !g->cone_tab;
  const bool cone = g->cone_tab && g->recompute && !fused && !g->nbc && !g->cplx && !has_magnetized(g) &&
                    t.bx == g->cone_tile.bx && t.by == g->cone_tile.by && t.kc == g->cone_tile.kc;
  dim3 block, grid;
  int ntx = 1, nty = 1, nkc = 1;
  if (tb3) {
    // time-blocked pass: cores of 60 x (by - 4) columns, kc planes
    ntx = (g->d.nx - 2 + fdtdk::kTbCore - 1) / fdtdk::kTbCore;
    nty = (g->d.ny - 2 + (t.by - 4) - 1) / (t.by - 4);
    nkc = (nzi + t.kc - 1) / t.kc;
    block = dim3(fdtdk::kTbCols, t.by, 1);
    grid = dim3(ntx * nty * nkc, 1, 1);
  } else if (wavef) {
    // z-marching fused pass: tiles of 63 x (by - 1) owned columns, kc planes
    ntx = (g->d.nx - 2 + fdtdk::kZCols - 2) / (fdtdk::kZCols - 1);
    nty = (g->d.ny - 2 + t.by - 2) / (t.by - 1);
    nkc = (nzi + t.kc - 1) / t.kc;
    block = dim3(fdtdk::kZCols, t.by, 1);
    grid = dim3(ntx * nty * nkc, 1, 1);
  } else if (fused) {
    // streamed pass: E regions of kSCols x kSRows cells own one column and one
    // row less; ncu * kc persistent workgroups split the (tile, plane) units
    ntx = (g->d.nx - 2 + fdtdk::kSCols - 2) / (fdtdk::kSCols - 1);
    nty = (g->d.ny - 2 + fdtdk::kSRows - 2) / (fdtdk::kSRows - 1);
    units = long(ntx) * nty * nzi;
    block = dim3(fdtdk::kSThreads, 1, 1);
    grid = dim3(unsigned(std::min<long>(long(g->ncu) * t.kc, units)), 1, 1);
  } else if (g->ndim == 3) {
    ntx = (g->d.nx + t.bx - 1) / t.bx;
    nty = (g->d.ny - 2 + t.by - 1) / t.by;
    nkc = (nzi + t.kc - 1) / t.kc;
    block = dim3(t.bx, t.by, 1);
    grid = dim3(ntx * nty * nkc, 1, 1);
  } else if (tmb) {
    // cores of kc x kc cells over the whole grid (the outer layer is copied)
    ntx = (g->d.nx + t.kc - 1) / t.kc;
    nty = (g->d.ny + t.kc - 1) / t.kc;
    const int R = t.kc + 2 * t.by;
    block = dim3(unsigned((R * R + 63) / 64 * 64), 1, 1);
    grid = dim3(unsigned(ntx * nty), 1, 1);
  } else if (tem) {
    block = dim3(256, 1, 1);  // a thread per cell of the line
    grid = dim3(unsigned((g->d.nx + 255) / 256), 1, 1);
  } else {
    block = dim3(t.bx, t.by, 1);
    grid = dim3((g->d.nx + t.bx - 1) / t.bx, (g->d.ny - 2 + t.by - 1) / t.by, 1);
  }
  // incident lines stepped here (not precomputed): chunk c's rows go to row
  // buffer c % 2, written on the lines' stream while chunk c - 1 runs
  const bool lines = (g->tfsf || g->nplane) && !(g->tfsf && g->inc_rows);
  if (lines && nsteps > 0) {
    int rc = ensure_line_stream(g);
    if (rc) return rc;
    HIP_TRY(hipEventRecord(g->start_ev, g->stream));  // earlier chunks are done with both buffers
    HIP_TRY(hipStreamWaitEvent(lines_on(g), g->start_ev, 0));
    if ((rc = line_rows(g, 0, g->step, std::min(kChunk, nsteps)))) return rc;
  }
  int done = 0;
  for (int cidx = 0; done < nsteps; cidx++) {
    const int chunk = std::min(kChunk, nsteps - done);
    const int lb = cidx % 2;
    a.step0 = g->step;
    gput(a.incE, g->aux.incE);
    gput(a.incH, g->aux.incH);
    if (g->tfsf && g->inc_rows) {
      // incident rows computed ahead for every step (adjoint forward and recompute)
      if (g->step + chunk > g->inc_rows) return fail(FDTD_ESTATE, "step beyond the precomputed incident rows");
      gput(a.incE, g->inc_allE + g->step * g->aux.L);
      gput(a.incH, g->inc_allH + g->step * g->aux.L);
      const size_t blk = size_t((g->step + chunk - 1) / kIncBlock);
      if (blk < g->inc_ev.size()) HIP_TRY(hipStreamWaitEvent(g->stream, g->inc_ev[blk], 0));
    } else if (lines) {
      HIP_TRY(hipStreamWaitEvent(g->stream, g->line_ev[lb], 0));
      gput(a.incE, g->aux.incE + size_t(lb) * kChunk * g->aux.L);
      gput(a.incH, g->aux.incH + size_t(lb) * kChunk * g->aux.L);
      for (int p = 0; p < g->nplane; p++) {
        gput(a.plane[p].incE, g->plane_aux[p].incE + size_t(lb) * kChunk * g->plane_aux[p].L);
        gput(a.plane[p].incH, g->plane_aux[p].incH + size_t(lb) * kChunk * g->plane_aux[p].L);
      }
    }
    // stage this chunk's arguments: variant 0 reads the current state,
    // variant 1 (fused pass only) the second copy; captured chunks alternate
    // between two slots so that their graphs are replayed
    const int slot = graphs ? g->args_slot % 2 : g->args_slot;
    g->args_slot = graphs ? (slot + 1) % 2 : (slot + 1) % fdtd_grid::kRing;
    a.dft_done0 = g->dft_done;
    if (g->args_ev_used[slot]) HIP_TRY(hipEventSynchronize(g->args_ev[slot]));
    auto* host = reinterpret_cast<fdtdk::StepArgs<T>*>(g->args_host + size_t(slot) * 2 * fdtd_grid::kArgsBytes);
    auto* host1 = reinterpret_cast<fdtdk::StepArgs<T>*>(g->args_host + (size_t(slot) * 2 + 1) * fdtd_grid::kArgsBytes);
    const auto* dev0 = reinterpret_cast<const fdtdk::StepArgs<T>*>(g->args_dev + size_t(slot) * 2 * fdtd_grid::kArgsBytes);
    const auto* dev1 =
        reinterpret_cast<const fdtdk::StepArgs<T>*>(g->args_dev + (size_t(slot) * 2 + 1) * fdtd_grid::kArgsBytes);
    set_state<T>(g, a, pp);
    *host = a;
    gput(host->self, dev0);
    if (pp) {
      swap_state(g);
      set_state<T>(g, a, true);
      *host1 = a;
      swap_state(g);
    } else if (g->cplx) {
      im_args<T>(g, a, *host1);  // variant 1: the imaginary part
    }
    gput(host1->self, dev1);
    HIP_TRY(hipMemcpyAsync((void*)dev0, host, 2 * fdtd_grid::kArgsBytes, hipMemcpyHostToDevice, g->stream));
    const long dft_work = host->dft_work, dft_cells = host->dft_cells, dft_slots = host->dft_slots;
    const long step0 = g->step;
    long bloch_cells = 0;
    const fdtdk::BlochFaces bfaces = bloch_faces(g, &bloch_cells);
    // the chunk's launches; every step-dependent value comes from the staged
    // arguments and the step index s, so the sequence can be captured once
    // and replayed for later chunks of the same shape
    // the time-blocked 3-D pass advances two steps per launch; with DFT
    // monitors a launch starts on an even step, so that a DFT block never
    // ends on a launch's first step (its fold follows the launch)
    auto tb_steps = [&](int s) { return (chunk - s >= 2 && (dft_work == 0 || (step0 + s) % 2 == 0)) ? 2 : 1; };
    int ntb = 0;
    if (tb3)
      for (int s = 0; s < chunk; s += tb_steps(s)) ntb++;
    const int nlaunch = tmb ? (chunk + t.by - 1) / t.by : (tb3 ? ntb : chunk);  // launches that swap the state roles
    auto enqueue = [&]() -> int {
      if (tb3) {
        // launch l reads copy l % 2 and writes the other
        int l = 0;
        for (int s = 0; s < chunk; l++) {
          const auto* ap = (l % 2) ? dev1 : dev0;
          const int nst = tb_steps(s);
          int rc;
          if (sizeof(T) == 8 || t.by == 8)
            rc = launch(g, 3, fdtdk::tb_3d<T, 8>, grid, block, ap, t.kc, s, nst, ntx, nty, nkc);
          else
            rc = launch(g, 3, fdtdk::tb_3d<float, 12>, grid, block, (const fdtdk::StepArgs<float>*)ap, t.kc, s, nst,
                        ntx, nty, nkc);
          if (!rc && dft_work > 0 && (step0 + s + nst) % fdtdk::kDftBlock == 0)
            rc = launch(g, 4, fdtdk::dft_fold_step<T>, dim3(unsigned((dft_slots + 255) / 256)), dim3(256), dev0,
                        s + nst - 1);
          if (rc) return rc;
          s += nst;
        }
        return 0;
      }
      if (tmb) {
        // launch l reads copy l % 2 and writes the other
        for (int l = 0; l < nlaunch; l++) {
          const auto* ap = (l % 2) ? dev1 : dev0;
          const int s = l * t.by, nst = std::min(t.by, chunk - s);
          int rc;
          switch (t.by) {
            case 2: rc = launch(g, 3, fdtdk::tm_block<T, 2>, grid, block, ap, t.kc, s, nst, ntx); break;
            case 4: rc = launch(g, 3, fdtdk::tm_block<T, 4>, grid, block, ap, t.kc, s, nst, ntx); break;
            case 6: rc = launch(g, 3, fdtdk::tm_block<T, 6>, grid, block, ap, t.kc, s, nst, ntx); break;
            default: rc = launch(g, 3, fdtdk::tm_block<T, 8>, grid, block, ap, t.kc, s, nst, ntx); break;
          }
          if (rc) return rc;
        }
        return 0;
      }
      for (int s = 0; s < chunk; s++) {
        int rc;
        const auto* ap = (fused && (s % 2)) ? dev1 : dev0;
        if (wavef) {
          const bool disp = g->max_np > 0;
          if (t.by == 4)
            rc = disp ? launch(g, 3, fdtdk::fusedz_3d<T, 4, 1>, grid, block, ap, t.kc, s, ntx, nty, nkc)
                      : launch(g, 3, fdtdk::fusedz_3d<T, 4, 0>, grid, block, ap, t.kc, s, ntx, nty, nkc);
          else if (t.by == 8)
            rc = disp ? launch(g, 3, fdtdk::fusedz_3d<T, 8, 1>, grid, block, ap, t.kc, s, ntx, nty, nkc)
                      : launch(g, 3, fdtdk::fusedz_3d<T, 8, 0>, grid, block, ap, t.kc, s, ntx, nty, nkc);
          else
            rc = disp ? launch(g, 3, fdtdk::fusedz_3d<T, 16, 1>, grid, block, ap, t.kc, s, ntx, nty, nkc)
                      : launch(g, 3, fdtdk::fusedz_3d<T, 16, 0>, grid, block, ap, t.kc, s, ntx, nty, nkc);
        } else if (fused) {
          if constexpr (sizeof(T) == 4) {
            if (g->max_np > 0)
              rc = launch(g, 3, fdtdk::fused3d_lds<1>, grid, block, (const fdtdk::StepArgs<float>*)ap, s, ntx, units);
            else
              rc = launch(g, 3, fdtdk::fused3d_lds<0>, grid, block, (const fdtdk::StepArgs<float>*)ap, s, ntx, units);
          } else {
            rc = fail(FDTD_ESTATE, "the streamed pass is FP32 only");
          }
        } else if (g->cplx) {
          // complex fields: each pass over both parts, then the Bloch ghosts
          // from both parts' new values
          const int disp = g->max_np > 0 ? (has_magnetized(g) ? 2 : 1) : 0;
          for (int ft = 0; ft < 2; ft++) {
            for (int part = 0; part < 2; part++) {
              const auto* p = part ? dev1 : dev0;
              if (g->ndim == 2 && te)
                rc = ft ? launch(g, 1, fdtdk::h_pass_te<T, true>, grid, block, p, 1, s)
                        : launch(g, 0, fdtdk::e_pass_te<T, true>, grid, block, p, 1, s);
              else if (g->ndim == 2)
                rc = ft ? launch(g, 1, fdtdk::h_pass_tm<T, true>, grid, block, p, 1, s)
                        : launch(g, 0, fdtdk::e_pass_tm<T, true>, grid, block, p, 1, s);
              else if (ft)
                rc = launch(g, 1, fdtdk::h_pass_3d<T, true>, grid, block, p, t.kc, s, ntx, nty, nkc);
              else if (disp == 2)
                rc = launch(g, 0, fdtdk::e_pass_3d<T, true, 2>, grid, block, p, t.kc, s, ntx, nty, nkc);
              else if (disp)
                rc = launch(g, 0, fdtdk::e_pass_3d<T, true, 1>, grid, block, p, t.kc, s, ntx, nty, nkc);
              else
                rc = launch(g, 0, fdtdk::e_pass_3d<T, true, 0>, grid, block, p, t.kc, s, ntx, nty, nkc);
              if (rc) return rc;
            }
            rc = launch(g, 7, fdtdk::bloch_ghosts<T>, dim3(unsigned((bloch_cells + 255) / 256), unsigned(bfaces.n)),
                        dim3(256), dev0, dev1, bfaces, ft);
            if (rc) return rc;
          }
        } else if (g->ndim == 3 && cone && g->cone_last - (step0 + s) >= 0 &&
                   g->cone_last - (step0 + s) + 1 < (long)g->cone_off.size()) {
          // light-cone recompute: the batch passes (the same cell code) over
          // the step's tile table, the grid as a batch of one
          const long r = g->cone_last - (step0 + s);
          const int o0 = g->cone_off[r], cnt = g->cone_off[r + 1] - o0;
          const int disp = g->max_np > 0 ? 1 : 0;
          rc = disp ? launch(g, 0, fdtdk::e_pass_batch<T, 1>, dim3(unsigned(cnt)), block, dev0, g->cone_tab + o0, t.kc, s)
                    : launch(g, 0, fdtdk::e_pass_batch<T, 0>, dim3(unsigned(cnt)), block, dev0, g->cone_tab + o0, t.kc, s);
          if (rc) return rc;
          rc = launch(g, 1, fdtdk::h_pass_batch<T>, dim3(unsigned(cnt)), block, dev0, g->cone_tab + o0, t.kc, s);
        } else if (g->ndim == 3) {
          // the general instances (boundary ghost writes, E / H sources,
          // single TFSF planes) and the E-pass instance without the ADE code when no material is
          // dispersive (with the magnetized Drude code only when one is magnetized)
          const int disp = g->max_np > 0 ? (has_magnetized(g) ? 2 : 1) : 0;
          if (general && disp == 2)
            rc = launch(g, 0, fdtdk::e_pass_3d<T, true, 2>, grid, block, dev0, t.kc, s, ntx, nty, nkc);
          else if (general && disp)
            rc = launch(g, 0, fdtdk::e_pass_3d<T, true, 1>, grid, block, dev0, t.kc, s, ntx, nty, nkc);
          else if (general)
            rc = launch(g, 0, fdtdk::e_pass_3d<T, true, 0>, grid, block, dev0, t.kc, s, ntx, nty, nkc);
          else if (disp == 2)
            rc = launch(g, 0, fdtdk::e_pass_3d<T, false, 2>, grid, block, dev0, t.kc, s, ntx, nty, nkc);
          else if (disp)
            rc = launch(g, 0, fdtdk::e_pass_3d<T, false, 1>, grid, block, dev0, t.kc, s, ntx, nty, nkc);
          else
            rc = launch(g, 0, fdtdk::e_pass_3d<T, false, 0>, grid, block, dev0, t.kc, s, ntx, nty, nkc);
          if (rc) return rc;
          if (general)
            rc = launch(g, 1, fdtdk::h_pass_3d<T, true>, grid, block, dev0, t.kc, s, ntx, nty, nkc);
          else
            rc = launch(g, 1, fdtdk::h_pass_3d<T>, grid, block, dev0, t.kc, s, ntx, nty, nkc);
        } else if (tem && general) {
          rc = launch(g, 0, fdtdk::e_pass_tem<T, true>, grid, block, dev0, 1, s);
          if (!rc) rc = launch(g, 1, fdtdk::h_pass_tem<T, true>, grid, block, dev0, 1, s);
        } else if (tem) {
          rc = launch(g, 0, fdtdk::e_pass_tem<T>, grid, block, dev0, 1, s);
          if (!rc) rc = launch(g, 1, fdtdk::h_pass_tem<T>, grid, block, dev0, 1, s);
        } else if (te && general) {
          rc = launch(g, 0, fdtdk::e_pass_te<T, true>, grid, block, dev0, 1, s);
          if (!rc) rc = launch(g, 1, fdtdk::h_pass_te<T, true>, grid, block, dev0, 1, s);
        } else if (te) {
          rc = launch(g, 0, fdtdk::e_pass_te<T>, grid, block, dev0, 1, s);
          if (!rc) rc = launch(g, 1, fdtdk::h_pass_te<T>, grid, block, dev0, 1, s);
        } else if (general) {
          rc = launch(g, 0, fdtdk::e_pass_tm<T, true>, grid, block, dev0, 1, s);
          if (!rc) rc = launch(g, 1, fdtdk::h_pass_tm<T, true>, grid, block, dev0, 1, s);
        } else {
          rc = launch(g, 0, fdtdk::e_pass_tm<T>, grid, block, dev0, 1, s);
          if (!rc) rc = launch(g, 1, fdtdk::h_pass_tm<T>, grid, block, dev0, 1, s);
        }
        for (int part = 0; part < (g->cplx ? 2 : 1) && !rc && dft_work > 0; part++) {
          const auto* dp = part ? dev1 : ap;
          rc = launch(g, 4, fdtdk::dft_sample<T>, dim3(unsigned((dft_cells + 255) / 256)), dim3(256), dp, s);
          if (!rc && (step0 + s + 1) % fdtdk::kDftBlock == 0)
            rc = launch(g, 4, fdtdk::dft_fold_step<T>, dim3(unsigned((dft_slots + 255) / 256)), dim3(256), dp, s);
        }
        if (rc) return rc;
      }
      return 0;
    };
    if (graphs) {
      const int disp = g->max_np > 0 ? (has_magnetized(g) ? 2 : 1) : 0;
      const GraphKey key{slot, chunk, int(step0 % fdtdk::kDftBlock), t.bx, t.by, t.kc, disp, dft_work, dft_cells};
      auto it = g->graphs.find(key);
      if (it == g->graphs.end()) {
        hipGraph_t graph;
        HIP_TRY(hipStreamBeginCapture(g->stream, hipStreamCaptureModeThreadLocal));
        int rc = enqueue();
        hipError_t e = hipStreamEndCapture(g->stream, &graph);
        if (rc) return rc;
        if (e != hipSuccess) return fail(FDTD_EDEVICE, "hipStreamEndCapture: %s", hipGetErrorString(e));
        hipGraphExec_t exec;
        e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        if (e != hipSuccess) return fail(FDTD_EDEVICE, "hipGraphInstantiate: %s", hipGetErrorString(e));
        it = g->graphs.emplace(key, exec).first;
      }
      HIP_TRY(hipGraphLaunch(it->second, g->stream));
    } else {
      int rc = enqueue();
      if (rc) return rc;
    }
    // host-side effects of the chunk
    if (pp && (nlaunch % 2)) swap_state(g);
    if (pp) g->nswap += nlaunch - (nlaunch % 2);
    if (dft_work > 0) {
      const long last = ((step0 + chunk) / fdtdk::kDftBlock) * fdtdk::kDftBlock;  // last block end in the chunk
      if (last > step0) g->dft_done = last;
    }
    g->step += chunk;
    HIP_TRY(hipEventRecord(g->args_ev[slot], g->stream));
    g->args_ev_used[slot] = true;
    done += chunk;
    if (lines && done < nsteps) {
      // the next chunk's rows, into the other buffer once the chunk before
      // this one has finished reading it
      HIP_TRY(hipEventRecord(g->pass_ev[lb], g->stream));
      if (cidx >= 1) HIP_TRY(hipStreamWaitEvent(lines_on(g), g->pass_ev[1 - lb], 0));
      int rc = line_rows(g, 1 - lb, g->step, std::min(kChunk, nsteps - done));
      if (rc) return rc;
    }
  }
  return 0;
}

// Fold the samples of steps [dft_done, step) into the DFT accumulators: done
// before the accumulators are read and before the monitors or the way the
// grid is stepped change.
template <typename T>
int dft_flush_impl(fdtd_grid* g) {
  if (g->ndft == 0 || g->dft_done >= g->step) {
    g->dft_done = g->step;
    return 0;
  }
  fdtdk::StepArgs<T> a;
  fill_args<T>(g, a);
  const int slot = g->args_slot;
  g->args_slot = (slot + 1) % fdtd_grid::kRing;
  if (g->args_ev_used[slot]) HIP_TRY(hipEventSynchronize(g->args_ev[slot]));
  auto* host = reinterpret_cast<fdtdk::StepArgs<T>*>(g->args_host + size_t(slot) * 2 * fdtd_grid::kArgsBytes);
  const auto* dev = reinterpret_cast<const fdtdk::StepArgs<T>*>(g->args_dev + size_t(slot) * 2 * fdtd_grid::kArgsBytes);
  const auto* dev1 =
      reinterpret_cast<const fdtdk::StepArgs<T>*>(g->args_dev + (size_t(slot) * 2 + 1) * fdtd_grid::kArgsBytes);
  *host = a;
  gput(host->self, dev);
  if (g->cplx) {
    auto* host1 = reinterpret_cast<fdtdk::StepArgs<T>*>(g->args_host + (size_t(slot) * 2 + 1) * fdtd_grid::kArgsBytes);
    im_args<T>(g, a, *host1);
    gput(host1->self, dev1);
  }
  HIP_TRY(hipMemcpyAsync((void*)dev, host, 2 * fdtd_grid::kArgsBytes, hipMemcpyHostToDevice, g->stream));
  for (int part = 0; part < (g->cplx ? 2 : 1); part++) {
    int rc = launch(g, 4, fdtdk::dft_fold<T>, dim3(unsigned((a.dft_slots + 255) / 256)), dim3(256), part ? dev1 : dev,
                    g->dft_done, g->step);
    if (rc) return rc;
  }
  HIP_TRY(hipEventRecord(g->args_ev[slot], g->stream));
  g->args_ev_used[slot] = true;
  g->dft_done = g->step;
  return 0;
}

int dft_flush(fdtd_grid* g) { return g->f64 ? dft_flush_impl<double>(g) : dft_flush_impl<float>(g); }

// Copy of every device allocation of a grid (fields, psi, dispersive state,
// DFT accumulators and samples, aux line, probe trace) and of the host-side
// step state, so that trial steps can be undone.
struct GridSnapshot {
  char* buf = nullptr;
  std::vector<void*> ptrs;
  std::vector<size_t> sizes;
  long step = 0, dft_done = 0, nswap = 0;
};

int snapshot_take(fdtd_grid* g, GridSnapshot& s) {
  int rc = 0;
  // pending table uploads go first: the snapshot must hold the tables the
  // trials will run with, or the restore would roll them back to stale contents
  if (g->mats_dirty && (rc = g->f64 ? upload_materials<double>(g) : upload_materials<float>(g))) return rc;
  if (g->fluid_dirty && (rc = g->f64 ? upload_fluid<double>(g) : upload_fluid<float>(g))) return rc;
  if (g->coef_dirty && (rc = upload_coefs(g))) return rc;
  HIP_TRY(hipStreamSynchronize(g->stream));
  s.ptrs = g->allocs;
  s.sizes = g->alloc_bytes;
  size_t total = 0;
  for (size_t b : s.sizes) total += (b + 255) & ~size_t(255);
  if (hipMalloc((void**)&s.buf, total) != hipSuccess) return fail(FDTD_ENOMEM, "snapshot (%zu bytes)", total);
  size_t off = 0;
  for (size_t q = 0; q < s.ptrs.size(); q++) {
    HIP_TRY(hipMemcpyAsync(s.buf + off, s.ptrs[q], s.sizes[q], hipMemcpyDeviceToDevice, g->stream));
    off += (s.sizes[q] + 255) & ~size_t(255);
  }
  s.step = g->step;
  s.dft_done = g->dft_done;
  s.nswap = g->nswap;
  return 0;
}

void snapshot_restore(fdtd_grid* g, GridSnapshot& s) {
  if (!s.buf) return;
  size_t off = 0;
  for (size_t q = 0; q < s.ptrs.size(); q++) {
    (void)hipMemcpyAsync(s.ptrs[q], s.buf + off, s.sizes[q], hipMemcpyDeviceToDevice, g->stream);
    off += (s.sizes[q] + 255) & ~size_t(255);
  }
  (void)hipStreamSynchronize(g->stream);
  (void)hipFree(s.buf);
  s.buf = nullptr;
  // fused trials swap the roles of the two state copies; put them back
  if ((g->nswap - s.nswap) % 2) swap_state(g);
  g->step = s.step;
  g->dft_done = s.dft_done;
}

// (i, j, k) is an interior cell of the mode's axes, with coordinate 0 on the others
static bool interior_cell(const fdtd_grid* g, int i, int j, int k) {
  const int c[3] = {i, j, k};
  for (int a = 0; a < 3; a++)
    if (a < g->ndim ? (c[a] < 1 || c[a] >= g->len[a] - 1) : c[a] != 0) return false;
  return true;
}

bool valid_comp(const fdtd_grid* g, int field, int comp) {
  if (comp < 0 || comp > 2) return false;
  if (field == FDTD_E || field == FDTD_D || field == FDTD_P || field == FDTD_J) return g->actE[comp];
  if (field == FDTD_H || field == FDTD_B || field == FDTD_M || field == FDTD_K) return g->actH[comp];
  return false;
}

// eps_r of a cell: the design value inside the design box, else the material's
double cell_eps(const fdtd_grid* g, int i, int j, int k) {
  if (g->design && i >= g->dsg[0] && i < g->dsg[1] && j >= g->dsg[2] && j < g->dsg[3] && k >= g->dsg[4] &&
      k < g->dsg[5]) {
    const long wx = g->dsg[1] - g->dsg[0], wy = g->dsg[3] - g->dsg[2];
    return g->dsg_eps[(i - g->dsg[0]) + wx * ((j - g->dsg[2]) + wy * (k - g->dsg[4]))];
  }
  return g->mats[g->mat_host[g->q(i, j, k)]].eps;
}

// ---------------------------------------------------------------- adjoint
// The stepped state of a non-dispersive grid (E, H, CPML psi), in a fixed order.
void state_arrays(fdtd_grid* g, std::vector<std::pair<void*, size_t>>& out) {
  out.clear();
  auto add = [&](void* p) {
    if (!p) return;
    auto it = std::find(g->allocs.begin(), g->allocs.end(), p);
    if (it != g->allocs.end()) out.push_back({p, g->alloc_bytes[it - g->allocs.begin()]});
  };
  for (int c = 0; c < 3; c++) {
    add(g->E[c]);
    add(g->H[c]);
    add(g->Dd[c]);  // pole cells: D, P_p, J_p over the dispersive box
    for (int p = 0; p < FDTD_MAX_POLES; p++) {
      add(g->Pd[p][c]);
      add(g->Jd[p][c]);
    }
  }
  for (int ft = 0; ft < 2; ft++)
    for (int I = 0; I < 3; I++)
      for (int J = 0; J < 3; J++) add(g->psi[ft][I][J]);
}

// Temporaries and forward-step controls of one adjoint call, undone on every exit.
struct AdjointScope {
  fdtd_grid* g;
  hipStream_t main;
  std::vector<void*> ptrs;
  hipStream_t rs = nullptr;  // recompute stream
  std::vector<hipEvent_t> evs;
  explicit AdjointScope(fdtd_grid* gg) : g(gg), main(gg->stream) {}
  ~AdjointScope() {
    g->stream = main;
    if (g->line_stream) (void)hipStreamSynchronize(g->line_stream);  // precomputed rows still being written
    g->inc_ev.clear();
    if (rs) {
      (void)hipStreamSynchronize(rs);
      (void)hipStreamDestroy(rs);
    }
    for (void* p : ptrs) (void)hipFreeAsync(p, g->stream);
    (void)hipStreamSynchronize(g->stream);
    for (auto e : evs) (void)hipEventDestroy(e);
    g->force_two_pass = false;
    g->recompute = false;
    g->hist = nullptr;
    g->hist_rows = 0;
    g->inc_allE = g->inc_allH = nullptr;
    g->inc_rows = 0;
    g->cone_tab = nullptr;
    g->cone_off.clear();
    g->cone_last = -1;
  }
  // Temporaries come from the stream-ordered allocator: hipMalloc / hipFree
  // synchronise the whole device and would serialise adjoints of several
  // designs running side by side on their own streams.
  int get(void** p, size_t bytes) {
    if (bytes == 0) bytes = 16;
    if (hipMallocAsync(p, bytes, g->stream) != hipSuccess) {
      *p = nullptr;
      return fail(FDTD_ENOMEM, "hipMallocAsync(%zu) failed", bytes);
    }
    ptrs.push_back(*p);
    HIP_TRY(hipMemsetAsync(*p, 0, bytes, g->stream));
    return 0;
  }
  int event(hipEvent_t* e) {
    if (hipEventCreateWithFlags(e, hipEventDisableTiming) != hipSuccess) return fail(FDTD_EDEVICE, "hipEventCreate");
    evs.push_back(*e);
    return 0;
  }
};

// Gradient of J = sum |F|^2 (monitor `didx`, E components in `mask`, all its
// frequencies) over steps [0, nsteps) with respect to the design cells' eps_r.
// Segments of K steps: the design increments of the last R segments (as many
// as `history_mb` holds, at least one) are recorded during the forward pass;
// every earlier segment is checkpointed (E, H, psi) and recomputed from its
// checkpoint, into the second state copy, while the next later segment's
// adjoint steps run.
template <typename T>
int adjoint_impl(fdtd_grid* g, int nsteps, int K, double history_mb, int didx, int mask, double* objective,
                 double* grad) {
  AdjointScope sc(g);
  int rc = 0;
  g->force_two_pass = true;
  if (g->tfsf) {
    // every step's incident rows, in blocks of kIncBlock steps on the lines'
    // stream: the forward's chunks wait only for the blocks they read
    if ((rc = sc.get((void**)&g->inc_allE, size_t(nsteps) * g->aux.L * 8))) return rc;
    if ((rc = sc.get((void**)&g->inc_allH, size_t(nsteps) * g->aux.L * 8))) return rc;
    if ((rc = ensure_line_stream(g))) return rc;
    hipStream_t ls = lines_on(g);
    HIP_TRY(hipEventRecord(g->start_ev, g->stream));
    HIP_TRY(hipStreamWaitEvent(ls, g->start_ev, 0));
    for (long b0 = 0; b0 < nsteps; b0 += kIncBlock) {
      const int n = int(std::min<long>(kIncBlock, nsteps - b0));
      const size_t off = size_t(b0) * g->aux.L;
      if ((rc = run_aux_line(g, g->aux, g->inc_allE + off, g->inc_allH + off, b0, n, ls, ls == g->stream))) return rc;
      hipEvent_t e;
      HIP_TRY(hipEventCreateWithFlags(&e, hipEventDisableTiming));
      sc.evs.push_back(e);
      HIP_TRY(hipEventRecord(e, ls));
      g->inc_ev.push_back(e);
    }
    g->inc_rows = nsteps;
  }
  const int S = (nsteps + K - 1) / K;
  const long nd = g->ndes;
  const size_t row_bytes = size_t(3) * nd * g->esz;
  const size_t seg_bytes = size_t(K) * row_bytes;
  long R = long(history_mb * 1e6 / double(seg_bytes));
  R = std::max(1L, std::min<long>(R, S));
  const int nck = S - int(R);
  char* hist = nullptr;
  if ((rc = sc.get((void**)&hist, size_t(R) * seg_bytes))) return rc;
  std::vector<std::pair<void*, size_t>> st;
  state_arrays(g, st);
  size_t st_bytes = 0;
  for (auto& p : st) st_bytes += (p.second + 255) & ~size_t(255);
  char* ck = nullptr;
  if (nck && (rc = sc.get((void**)&ck, size_t(nck) * st_bytes))) return rc;
  if (nck && (rc = ensure_pingpong(g))) return rc;
  auto copy_state = [&](char* base, bool save) -> int {
    state_arrays(g, st);
    size_t off = 0;
    for (auto& p : st) {
      void* dst = save ? (void*)(base + off) : p.first;
      const void* src = save ? (const void*)p.first : (const void*)(base + off);
      HIP_TRY(hipMemcpyAsync(dst, src, p.second, hipMemcpyDeviceToDevice, g->stream));
      off += (p.second + 255) & ~size_t(255);
    }
    return 0;
  };

  // forward pass with checkpoints / recorded increments
  for (int s = 0; s < S && !rc; s++) {
    const long n0 = long(s) * K;
    const int len = int(std::min<long>(K, nsteps - n0));
    if (s < nck) {
      rc = copy_state(ck + size_t(s) * st_bytes, true);
    } else {
      g->hist = hist + size_t(s - nck) * seg_bytes;
      g->hist_step0 = n0;
      g->hist_rows = len;
    }
    if (!rc) rc = step_impl<T>(g, len);
    g->hist = nullptr;
    g->hist_rows = 0;
  }
  if (rc) return rc;

  // objective and its adjoint source weights 2 F
  if ((rc = dft_flush_impl<T>(g))) return rc;
  HIP_TRY(hipStreamSynchronize(g->stream));
  const auto& d = g->dft[didx];
  const size_t per_slot = size_t(d.nfreq) * 2 * d.ncells;
  std::vector<double> acc(size_t(d.ncomp) * per_slot);
  HIP_TRY(hipMemcpy(acc.data(), d.acc, acc.size() * 8, hipMemcpyDeviceToHost));
  fdtdk::AdjArgs<T> q;
  memset(&q, 0, sizeof q);
  q.eps0 = kEps0;
  for (int x = 0; x < 6; x++) q.obj_r[x] = d.r[x];
  q.obj_nfreq = d.nfreq;
  q.obj_ncells = d.ncells;
  gput(q.obj_ph, d.phE);
  q.obj_nphase = d.nphase;
  double J = 0;
  std::vector<double> w;
  int ns = 0;
  for (int c = 0; c < 3; c++) q.obj_comp[c] = -1;
  for (int c = 0; c < 3; c++) {
    if (!((mask >> c) & 1)) continue;
    const int sm = __builtin_popcount(unsigned(d.mask) & ((1u << c) - 1));
    const double* src = acc.data() + size_t(sm) * per_slot;
    for (size_t x = 0; x < per_slot; x++) {
      J += src[x] * src[x];
      w.push_back(2.0 * src[x]);
    }
    q.obj_comp[ns++] = c;
  }
  void* wdev = nullptr;
  if ((rc = sc.get(&wdev, w.size() * g->esz))) return rc;
  HIP_TRY(hipStreamSynchronize(g->stream));  // the stream-ordered allocation exists before the copy
  if ((rc = upload_T(g, wdev, w.data(), w.size()))) return rc;
  gput(q.obj_w, (const T*)wdev);

  // adjoint state (zero: the adjoint of the state after the last step)
  for (int c = 0; c < 3; c++) {
    if ((rc = sc.get((void**)&q.nu[c], size_t(g->ncell) * g->esz))) return rc;
    if ((rc = sc.get((void**)&q.mu[c], size_t(g->ncell) * g->esz))) return rc;
  }
  for (int ft = 0; ft < 2; ft++)
    for (int I = 0; I < 3; I++)
      for (int Jx = 0; Jx < 3; Jx++) {
        void* p = g->psi[ft][I][Jx];
        if (!p) continue;
        auto it = std::find(g->allocs.begin(), g->allocs.end(), p);
        if ((rc = sc.get((void**)&q.lpsi[ft][I][Jx], g->alloc_bytes[it - g->allocs.begin()]))) return rc;
      }
  if (g->max_np > 0) {  // adjoints of the pole cells' D, P_p, J_p
    const long nb = long(g->dbox[1] - g->dbox[0]) * (g->dbox[3] - g->dbox[2]) * (g->dbox[5] - g->dbox[4]);
    for (int c = 0; c < 3; c++) {
      if ((rc = sc.get((void**)&q.aD[c], size_t(nb) * g->esz))) return rc;
      for (int p = 0; p < g->max_np; p++) {
        if ((rc = sc.get((void**)&q.aP[p][c], size_t(nb) * g->esz))) return rc;
        if ((rc = sc.get((void**)&q.aJ[p][c], size_t(nb) * g->esz))) return rc;
      }
    }
  }
  if ((rc = sc.get((void**)&q.grad, size_t(nd) * 8))) return rc;
  fdtdk::StepArgs<T> a;
  fill_args<T>(g, a);
  void *adev = nullptr, *qdev = nullptr;
  if ((rc = sc.get(&adev, sizeof a))) return rc;
  if ((rc = sc.get(&qdev, sizeof q))) return rc;
  HIP_TRY(hipStreamSynchronize(g->stream));  // the stream-ordered allocations exist before the copies
  gput(a.self, (const fdtdk::StepArgs<T>*)adev);
  gput(q.self, (const fdtdk::AdjArgs<T>*)qdev);
  HIP_TRY(hipMemcpy(adev, &a, sizeof a, hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(qdev, &q, sizeof q, hipMemcpyHostToDevice));
  const auto* ap = (const fdtdk::StepArgs<T>*)adev;
  const auto* qp = (const fdtdk::AdjArgs<T>*)qdev;

  // a two-pass tile (the fused and time-blocked tiles are not reverse-sweep
  // shapes), at most kAdjMaxKc planes per block (the z coefficients in LDS)
  Tile t = g->rtile.by ? g->rtile : ((g->tile.by && g->tile.bx >= 8) ? g->tile : kDefaultTile3D);
  t.kc = std::min(t.kc, fdtdk::kAdjMaxKc);
  const int nzi = g->d.nz - 2;
  const dim3 block(t.bx, t.by, 1);
  const int ntx = (g->d.nx + t.bx - 1) / t.bx, nty = (g->d.ny - 2 + t.by - 1) / t.by, nkc = (nzi + t.kc - 1) / t.kc;
  const dim3 grd(ntx * nty * nkc, 1, 1);

  // Checkpointed segments are recomputed on a second stream, one segment
  // ahead of the reverse steps on the grid's stream, into two alternating
  // increment buffers: segment s-1 is recomputed while segment s is reversed.
  // ev_r[b]: buffer b holds its segment; ev_a[b]: the reverse steps that read
  // buffer b are done.
  hipStream_t M = g->stream;
  char* X[2] = {nullptr, nullptr};
  hipEvent_t ev_fwd = nullptr, ev_r[2] = {nullptr, nullptr}, ev_a[2] = {nullptr, nullptr};
  if (nck) {
    if (hipStreamCreateWithFlags(&sc.rs, hipStreamNonBlocking) != hipSuccess)
      return fail(FDTD_EDEVICE, "hipStreamCreateWithFlags failed");
    for (int b = 0; b < 2; b++) {
      if ((rc = sc.get((void**)&X[b], seg_bytes))) return rc;
      if ((rc = sc.event(&ev_r[b])) || (rc = sc.event(&ev_a[b]))) return rc;
    }
    if ((rc = sc.event(&ev_fwd))) return rc;
    HIP_TRY(hipEventRecord(ev_fwd, M));
    HIP_TRY(hipStreamWaitEvent(sc.rs, ev_fwd, 0));
  }
  // Light-cone tables for the recompute.  A segment's recompute only has to
  // reproduce the design cells' E increments: E(i) reads H(i-1), H(i) and
  // H(i) reads E(i), E(i+1).  If the passes of step m run over the design box
  // grown by D = L - m cells on every side (L: the segment's last step), then
  // after step m E is exact on the box grown by D and H on it less its upper
  // layer, which is all that step m + 1's E reads on the box grown by D - 1;
  // at step L the box itself is exact.  Cells outside the region keep stale
  // values that no exact cell reads.  Table r holds the tiles (of the
  // recompute's tile shape) meeting the box grown by r.  Grids with boundary
  // ghost writes, complex fields or magnetized materials recompute whole.
  if (nck && !g->nbc && !g->cplx && !has_magnetized(g)) {
    Tile tr = g->rtile.by ? g->rtile : g->tile;
    if (!tr.by) tr = kDefaultTile3D;
    if (tr.bx < 8) tr = kDefaultTile3D;
    std::vector<int4> tab;
    std::vector<int> off(1, 0);
    // tables stop where the cone covers every tile: later steps of a segment
    // take the whole-grid passes (so the tables hold at most a grid edge's
    // worth of distances whatever checkpoint_every is)
    const int nx = g->d.nx, ny = g->d.ny, nz = g->d.nz;
    const long all = long((nx + tr.bx - 1) / tr.bx) * ((ny - 2 + tr.by - 1) / tr.by) * ((nz - 2 + tr.kc - 1) / tr.kc);
    for (int r = 0; r < K; r++) {
      const int d = r;
      const int i0 = std::max(1, g->dsg[0] - d), i1 = std::min(nx - 2, g->dsg[1] - 1 + d);
      const int j0 = std::max(1, g->dsg[2] - d), j1 = std::min(ny - 2, g->dsg[3] - 1 + d);
      const int k0 = std::max(1, g->dsg[4] - d), k1 = std::min(nz - 2, g->dsg[5] - 1 + d);
      const long n = long(i1 / tr.bx - i0 / tr.bx + 1) * ((j1 - 1) / tr.by - (j0 - 1) / tr.by + 1) *
                     ((k1 - 1) / tr.kc - (k0 - 1) / tr.kc + 1);
      if (n >= all) break;
      for (int z = (k0 - 1) / tr.kc; z <= (k1 - 1) / tr.kc; z++)
        for (int y = (j0 - 1) / tr.by; y <= (j1 - 1) / tr.by; y++)
          for (int x = i0 / tr.bx; x <= i1 / tr.bx; x++) tab.push_back(make_int4(0, x, y, z));
      off.push_back(int(tab.size()));
    }
    if (!tab.empty()) {
      void* tdev = nullptr;
      if ((rc = sc.get(&tdev, tab.size() * sizeof(int4)))) return rc;
      HIP_TRY(hipStreamSynchronize(g->stream));  // the stream-ordered allocation exists before the copy
      HIP_TRY(hipMemcpy(tdev, tab.data(), tab.size() * sizeof(int4), hipMemcpyHostToDevice));
      g->cone_tab = (const int4*)tdev;
      g->cone_off = off;
      g->cone_tile = tr;
    }
  }
  auto recompute = [&](int s) -> int {
    const long n0 = long(s) * K;
    const int len = int(std::min<long>(K, nsteps - n0));
    const int b = s % 2;
    if (s + 2 < nck) HIP_TRY(hipStreamWaitEvent(sc.rs, ev_a[b], 0));
    g->stream = sc.rs;
    swap_state(g);
    int r = copy_state(ck + size_t(s) * st_bytes, false);
    const long step_saved = g->step;
    g->step = n0;
    g->recompute = true;
    g->hist = X[b];
    g->hist_step0 = n0;
    g->hist_rows = len;
    g->cone_last = n0 + len - 1;
    const Tile tile_saved = g->tile;
    if (g->rtile.by) g->tile = g->rtile;
    if (!r) r = step_impl<T>(g, len);
    g->tile = tile_saved;
    g->step = step_saved;
    g->recompute = false;
    g->hist = nullptr;
    g->hist_rows = 0;
    swap_state(g);
    g->stream = M;
    if (r) return r;
    HIP_TRY(hipEventRecord(ev_r[b], sc.rs));
    return 0;
  };
  if (nck && (rc = recompute(nck - 1))) return rc;
  for (int s = S - 1; s >= 0; s--) {
    const long n0 = long(s) * K;
    const int len = int(std::min<long>(K, nsteps - n0));
    char* hb;
    if (s >= nck) {
      hb = hist + size_t(s - nck) * seg_bytes;
    } else {
      hb = X[s % 2];
      HIP_TRY(hipStreamWaitEvent(M, ev_r[s % 2], 0));
    }
    if (s >= 1 && s < nck && (rc = recompute(s - 1))) return rc;
    for (long n = n0 + len - 1; n >= n0 && !rc; n--) {
      const T* hrow = (const T*)(hb + size_t(n - n0) * row_bytes);
      rc = g->max_np > 0 ? launch(g, 5, fdtdk::adj_e_pass_3d<T, true>, grd, block, ap, qp, t.kc, ntx, nty, nkc, n, hrow)
                         : launch(g, 5, fdtdk::adj_e_pass_3d<T>, grd, block, ap, qp, t.kc, ntx, nty, nkc, n, hrow);
      if (!rc) rc = launch(g, 6, fdtdk::adj_h_pass_3d<T>, grd, block, ap, qp, t.kc, ntx, nty, nkc);
    }
    if (rc) return rc;
    if (s < nck) HIP_TRY(hipEventRecord(ev_a[s % 2], M));
  }
  if (sc.rs) HIP_TRY(hipStreamSynchronize(sc.rs));
  HIP_TRY(hipStreamSynchronize(g->stream));
  HIP_TRY(hipMemcpy(grad, q.grad, size_t(nd) * 8, hipMemcpyDeviceToHost));
  *objective = J;
  return collect_events(g);
}

}  // namespace

// quantities per cell of fdtd_surface_quantities: 3-D 7, TM 3, TE 4
static int surface_count(const fdtd_grid* g) { return g->d.mode == FDTD_3D ? 7 : (g->d.mode == FDTD_TE ? 4 : 3); }

template <typename T>
int surface_impl(fdtd_grid* g, const fdtdk::SurfRegion& r, long cells, double* out) {
  const int nq = surface_count(g);
  double* dev = nullptr;
  HIP_TRY(hipMallocAsync((void**)&dev, size_t(nq) * cells * sizeof(double), g->stream));
  const long sy = g->d.nx, sz = long(g->d.nx) * g->d.ny;
  auto F = [](void* p) { return (const T*)p; };
  hipLaunchKernelGGL(fdtdk::surface_quantities<T>, dim3(unsigned((cells + 255) / 256)), dim3(256), 0, g->stream,
                     F(g->E[0]), F(g->E[1]), F(g->E[2]), F(g->H[0]), F(g->H[1]), F(g->H[2]), sy, sz, r, g->d.mode,
                     g->d.dx, kEps0, dev);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) e = hipMemcpyAsync(out, dev, size_t(nq) * cells * sizeof(double), hipMemcpyDeviceToHost, g->stream);
  (void)hipFreeAsync(dev, g->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(g->stream);
  if (e != hipSuccess) return fail(FDTD_EDEVICE, "surface quantities: %s", hipGetErrorString(e));
  return 0;
}

// Device staging of fdtd_surface_quantities_static: one copy per GPU, held
// until the process ends.  Host code uses the address hipGetSymbolAddress
// returns for the current device (scratch_ptr), never the array's name, which
// on the host is only a stand-in symbol.
constexpr long kSurfStaticMax = FDTD_SURFACE_STATIC_MAX_CELLS;
static __device__ __align__(256) double scratch[7 * kSurfStaticMax];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double* g_scratch_ptr[kMaxDevices] = {};

// the scratch array's address on the current device, resolved once per device
int scratch_ptr(double** out) {
  int dev = 0;
  HIP_TRY(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) return fail(FDTD_EDEVICE, "device index %d out of range", dev);
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) return fail(FDTD_EDEVICE, "hipGetSymbolAddress(scratch): %s", hipGetErrorString(e));
    g_scratch_ptr[dev] = static_cast<double*>(p);
  }
  *out = g_scratch_ptr[dev];
  return 0;
}

template <typename T>
int surface_static_impl(fdtd_grid* g, const fdtdk::SurfRegion& r, long cells, double* out) {
  const int nq = surface_count(g);
  double* dev = nullptr;
  int rc = scratch_ptr(&dev);
  if (rc) return rc;
  const long sy = g->d.nx, sz = long(g->d.nx) * g->d.ny;
  auto F = [](void* p) { return (const T*)p; };
  // every output element is written by the kernel: nothing to clear
  hipLaunchKernelGGL(fdtdk::surface_quantities<T>, dim3(unsigned((cells + 255) / 256)), dim3(256), 0, g->stream,
                     F(g->E[0]), F(g->E[1]), F(g->E[2]), F(g->H[0]), F(g->H[1]), F(g->H[2]), sy, sz, r, g->d.mode,
                     g->d.dx, kEps0, dev);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) e = hipMemcpyAsync(out, dev, size_t(nq) * cells * sizeof(double), hipMemcpyDeviceToHost, g->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(g->stream);
  if (e != hipSuccess) return fail(FDTD_EDEVICE, "surface quantities: %s", hipGetErrorString(e));
  return 0;
}

extern "C" {

const char* fdtd_last_error(void) { return g_err.c_str(); }

int fdtd_grid_create(const fdtd_grid_desc* desc, fdtd_grid** out) {
  if (!desc || !out) return fail(FDTD_EINVAL, "null argument");
  *out = nullptr;
  if (desc->mode != FDTD_TM && desc->mode != FDTD_TE && desc->mode != FDTD_3D && desc->mode != FDTD_TEM)
    return fail(FDTD_EINVAL, "mode must be FDTD_TEM, FDTD_TM, FDTD_TE or FDTD_3D");
  if (desc->precision != FDTD_F32 && desc->precision != FDTD_F64) return fail(FDTD_EINVAL, "bad precision");
  const int nz = desc->mode == FDTD_3D ? desc->nz : 1;
  const int ny = desc->mode == FDTD_TEM ? 1 : desc->ny;
  if (desc->nx < 3 || (desc->mode != FDTD_TEM && desc->ny < 3) || (desc->mode == FDTD_3D && nz < 3))
    return fail(FDTD_EINVAL, "grid must have at least 3 cells per axis");
  if (!(desc->dx > 0) || !(desc->dt > 0)) return fail(FDTD_EINVAL, "dx and dt must be positive");
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return fail(FDTD_EDEVICE, "no HIP device");
  fdtd_grid* g = new fdtd_grid();
  if (hipDeviceGetAttribute(&g->ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || g->ncu < 1) g->ncu = 256;
  g->d = *desc;
  g->d.nz = nz;
  g->d.ny = ny;
  g->ndim = desc->mode == FDTD_3D ? 3 : (desc->mode == FDTD_TEM ? 1 : 2);
  g->f64 = desc->precision == FDTD_F64;
  g->esz = g->f64 ? 8 : 4;
  g->len[0] = desc->nx;
  g->len[1] = ny;
  g->len[2] = nz;
  g->ncell = long(desc->nx) * ny * nz;
  if (g->ndim == 3) {
    for (int c = 0; c < 3; c++) g->actE[c] = g->actH[c] = true;
  } else if (desc->mode == FDTD_TE) {
    g->actE[0] = g->actE[1] = true;
    g->actH[2] = true;
  } else if (desc->mode == FDTD_TEM) {
    g->actE[2] = true;
    g->actH[1] = true;
  } else {
    g->actE[2] = true;
    g->actH[0] = g->actH[1] = true;
  }
  int rc = 0;
  for (int c = 0; c < 3 && !rc; c++) {
    if (g->actE[c]) rc = g->alloc(&g->E[c], (g->ncell + fdtdk::field_pad(desc->nx)) * g->esz);
    if (!rc && g->actH[c]) rc = g->alloc(&g->H[c], (g->ncell + fdtdk::field_pad(desc->nx)) * g->esz);
  }
  if (!rc) rc = g->alloc((void**)&g->mat, g->ncell + fdtdk::field_pad(desc->nx));
  g->mat_host.assign(g->ncell, 0);
  if (!rc) rc = g->alloc(&g->mats_dev, sizeof(fdtdk::MatDev<double>) * FDTD_MAX_MATERIALS);
  if (!rc) rc = g->alloc(&g->mie_dev, 8 * FDTD_MAX_MATERIALS);
  if (!rc) rc = g->alloc(&g->mim_dev, 8 * FDTD_MAX_MATERIALS);
  if (!rc) rc = g->alloc((void**)&g->args_dev, fdtd_grid::kRing * 2 * fdtd_grid::kArgsBytes);
  if (!rc && hipHostMalloc((void**)&g->args_host, fdtd_grid::kRing * 2 * fdtd_grid::kArgsBytes) != hipSuccess)
    rc = fail(FDTD_ENOMEM, "pinned staging for kernel arguments");
  for (int r = 0; r < fdtd_grid::kRing && !rc; r++)
    if (hipEventCreateWithFlags(&g->args_ev[r], hipEventDisableTiming) != hipSuccess)
      rc = fail(FDTD_EDEVICE, "hipEventCreateWithFlags failed");
  for (int ft = 0; ft < 2 && !rc; ft++)
    for (int a = 0; a < 3 && !rc; a++) {
      g->pml[ft][a].assign(4 * g->len[a], 0.0);
      for (int i = 0; i < g->len[a]; i++) g->pml[ft][a][4 * i + 2] = 1.0;
      for (int w = 0; w < 4 && !rc; w++) rc = g->alloc(&g->coef[ft][a][w], size_t(g->len[a]) * g->esz);
    }
  if (rc) {
    delete g;
    return rc;
  }
  *out = g;
  return 0;
}

int fdtd_grid_destroy(fdtd_grid* g) {
  if (!g) return 0;
  (void)hipStreamSynchronize(g->stream);
  delete g;
  return 0;
}

int fdtd_set_stream(fdtd_grid* g, void* s) {
  if (!g) return fail(FDTD_EINVAL, "null grid");
  g->stream = (hipStream_t)s;
  return 0;
}

int fdtd_set_pml_axis(fdtd_grid* g, int axis, int ft, const double* K, const double* S, const double* A) {
  if (!g || !K || !S || !A) return fail(FDTD_EINVAL, "null argument");
  if (axis < 0 || axis >= g->ndim || ft < 0 || ft > 1) return fail(FDTD_EINVAL, "bad axis/ft");
  if (g->finalized) return fail(FDTD_ESTATE, "CPML profiles are fixed once stepping has started");
  for (int i = 0; i < g->len[axis]; i++) pml_bcfg(K[i], S[i], A[i], g->d.dt, &g->pml[ft][axis][4 * i]);
  g->coef_dirty = true;
  return 0;
}

int fdtd_set_pml_idiom(fdtd_grid* g, int axis, int n, double m, double ma, double smax, double kmax, double amax) {
  if (!g) return fail(FDTD_EINVAL, "null grid");
  if (axis < 0 || axis >= g->ndim || n < 1 || 2 * n + 4 > g->len[axis]) return fail(FDTD_EINVAL, "bad axis or layer count");
  if (g->finalized) return fail(FDTD_ESTATE, "CPML profiles are fixed once stepping has started");
  pml_idiom_rows(g->len[axis], n, m, ma, smax, kmax, amax, g->d.dt, g->pml[0][axis], g->pml[1][axis]);
  g->coef_dirty = true;
  return 0;
}

int fdtd_set_material(fdtd_grid* g, int id, double eps_r, double mu_r, int npoles, const double* poles) {
  if (!g) return fail(FDTD_EINVAL, "null grid");
  if (id < 0 || id >= FDTD_MAX_MATERIALS || npoles < 0 || npoles > FDTD_MAX_POLES || (npoles && !poles))
    return fail(FDTD_EINVAL, "bad material id or pole count");
  if (!(eps_r > 0) || !(mu_r > 0)) return fail(FDTD_EINVAL, "eps_r and mu_r must be positive");
  MatHost& m = g->mats[id];
  if (g->finalized && (npoles > 0) != (m.np > 0)) return fail(FDTD_ESTATE, "dispersive layout is fixed once stepping has started");
  if (g->finalized && npoles > g->max_np) return fail(FDTD_ESTATE, "pole count exceeds the allocated dispersive state");
  for (int p = 0; p < npoles; p++)
    if (int(poles[4 * p]) != FDTD_DRUDE && int(poles[4 * p]) != FDTD_LORENTZ && int(poles[4 * p]) != FDTD_CONDUCTIVE)
      return fail(FDTD_EINVAL, "bad pole kind");
  m.eps = eps_r;
  m.mu = mu_r;
  m.np = npoles;
  m.mag = 0;
  m.fluid = 0;
  g->fluid_dirty = true;
  for (int p = 0; p < npoles; p++) {
    m.kind[p] = int(poles[4 * p]);
    m.freq[p] = poles[4 * p + 1];
    m.gamma[p] = poles[4 * p + 2];
    m.weight[p] = poles[4 * p + 3];
  }
  count_magnetized(g);
  g->mats_dirty = true;
  return 0;
}

int fdtd_set_material_magnetic(fdtd_grid* g, int id, int npoles, const double* poles) {
  if (!g) return fail(FDTD_EINVAL, "null grid");
  if (id < 0 || id >= FDTD_MAX_MATERIALS || npoles < 0 || npoles > FDTD_MAX_POLES || (npoles && !poles))
    return fail(FDTD_EINVAL, "bad material id or pole count");
  MatHost& m = g->mats[id];
  if (g->finalized && (npoles > 0) != (m.mnp > 0))
    return fail(FDTD_ESTATE, "the magnetic-pole layout is fixed once stepping has started");
  if (g->finalized && npoles > g->max_mnp) return fail(FDTD_ESTATE, "pole count exceeds the allocated magnetic state");
  for (int p = 0; p < npoles; p++)
    if (int(poles[4 * p]) != FDTD_DRUDE && int(poles[4 * p]) != FDTD_LORENTZ && int(poles[4 * p]) != FDTD_CONDUCTIVE)
      return fail(FDTD_EINVAL, "bad pole kind");
  m.mnp = npoles;
  for (int p = 0; p < npoles; p++) {
    m.mkind[p] = int(poles[4 * p]);
    m.mfreq[p] = poles[4 * p + 1];
    m.mgamma[p] = poles[4 * p + 2];
    m.mweight[p] = poles[4 * p + 3];
  }
  g->mats_dirty = true;
  return 0;
}

int fdtd_set_magnetized_drude(fdtd_grid* g, int id, double eps_r, double mu_r, double freq, double gamma,
                              const double* bias) {
  if (!g) return fail(FDTD_EINVAL, "null grid");
  const double pole[4] = {double(FDTD_DRUDE), freq, gamma, 1.0};
  int rc = fdtd_set_material(g, id, eps_r, mu_r, 1, pole);
  if (rc) return rc;
  MatHost& m = g->mats[id];
  m.mag = bias ? 1 : 2;
  for (int c = 0; c < 3; c++) m.bias[c] = bias ? bias[c] : 0.0;
  count_magnetized(g);
  return 0;
}

int fdtd_set_fluid_drude(fdtd_grid* g, int id, double mu_r, int magnetization, const double* bias) {
  if (!g) return fail(FDTD_EINVAL, "null grid");
  if (magnetization < 0 || magnetization > 2 || (magnetization == 1 && !bias))
    return fail(FDTD_EINVAL, "magnetization must be 0 (none), 1 (fixed bias, given) or 2 (the cell's own field)");
  const double pole[4] = {double(FDTD_DRUDE), 0.0, 0.0, 1.0};
  int rc = fdtd_set_material(g, id, 1.0, mu_r, 1, pole);  // K = 1 (FluidDrudeStorage::getK)
  if (rc) return rc;
  MatHost& m = g->mats[id];
  m.fluid = 1;
  m.mag = magnetization;
  for (int c = 0; c < 3; c++) m.bias[c] = magnetization == 1 ? bias[c] : 0.0;
  count_magnetized(g);
  return 0;
}

int fdtd_set_fluid_state(fdtd_grid* g, const double* ne, const double* nu_m) {
  if (!g || !ne || !nu_m) return fail(FDTD_EINVAL, "null argument");
  for (long i = 0; i < g->ncell; i++)
    if (!(ne[i] >= 0) || !(nu_m[i] >= 0)) return fail(FDTD_EINVAL, "densities and collision rates must be >= 0");
  g->fluid_ne.assign(ne, ne + g->ncell);
  g->fluid_nu.assign(nu_m, nu_m + g->ncell);
  g->fluid_dirty = true;
  return 0;
}

int fdtd_set_matid(fdtd_grid* g, const uint8_t* ids) {
  if (!g || !ids) return fail(FDTD_EINVAL, "null argument");
  if (g->finalized && g->max_np > 0) {
    for (long i = 0; i < g->ncell; i++)
      if ((g->mats[ids[i]].np > 0) != (g->mats[g->mat_host[i]].np > 0))
        return fail(FDTD_ESTATE, "dispersive layout is fixed once stepping has started");
  }
  std::copy(ids, ids + g->ncell, g->mat_host.begin());
  g->fluid_dirty = true;
  HIP_TRY(hipMemcpy(g->mat, ids, g->ncell, hipMemcpyHostToDevice));
  return 0;
}

int fdtd_add_source(fdtd_grid* g, int field, int comp, int i, int j, int k, const double* table, int ntable) {
  if (!g || !table || ntable < 0) return fail(FDTD_EINVAL, "null argument");
  if (g->nsrc >= FDTD_MAX_SOURCES) return fail(FDTD_EINVAL, "too many sources (max %d)", FDTD_MAX_SOURCES);
  if (field < FDTD_D || field > FDTD_H || !valid_comp(g, field, comp)) return fail(FDTD_EINVAL, "bad source field/component");
  if (!interior_cell(g, i, j, k)) return fail(FDTD_EINVAL, "source cell must be an interior cell");
  double* t = nullptr;
  int rc = g->alloc((void**)&t, size_t(std::max(ntable, 1)) * 8);
  if (rc) return rc;
  if (ntable) HIP_TRY(hipMemcpy(t, table, size_t(ntable) * 8, hipMemcpyHostToDevice));
  g->src_table[g->nsrc] = t;
  g->src_cap[g->nsrc] = std::max(ntable, 1);
  fdtdk::SrcDev& s = g->src[g->nsrc++];
  s.cell = g->q(i, j, k);
  s.field = field;
  s.comp = comp;
  gput(s.table, t);
  s.n = ntable;
  if (field == FDTD_E || field == FDTD_H) g->nsrc_eh++;
  return 0;
}

// Allocate and set up one incident line of aux_ncells cells (Aux1D's length
// ncells + 2*15 + 20, Source.hpp:401-408) driven by signal[n] at its n-th step.
int aux_line_init(fdtd_grid* g, int aux_ncells, const double* signal, int nsignal, AuxLine& l) {
  l.L = aux_ncells + 2 * 15 + 20;
  const int L = l.L;
  int rc = g->alloc((void**)&l.state, size_t(6) * L * 8);
  if (!rc) rc = g->alloc((void**)&l.pE, size_t(4) * L * 8);
  if (!rc) rc = g->alloc((void**)&l.pH, size_t(4) * L * 8);
  if (!rc) rc = g->alloc((void**)&l.sig, size_t(std::max(nsignal, 1)) * 8);
  // two chunk buffers of rows (a chunk's rows are written while the chunk before runs)
  if (!rc) rc = g->alloc((void**)&l.incE, size_t(2) * kChunk * L * 8);
  if (!rc) rc = g->alloc((void**)&l.incH, size_t(2) * kChunk * L * 8);
  if (rc) return rc;
  // Aux1D's own CPML: PMLParameterModel(dx) defaults, 15 layers (Source.hpp:362, 380-389)
  std::vector<double> e, h;
  const double m = 3;
  pml_idiom_rows(L, 15, m, 1, 0.8 * (m + 1) / (kImp0 * g->d.dx), 5, 0.05, g->d.dt, e, h);
  HIP_TRY(hipMemcpy(l.pE, e.data(), e.size() * 8, hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(l.pH, h.data(), h.size() * 8, hipMemcpyHostToDevice));
  if (nsignal) HIP_TRY(hipMemcpy(l.sig, signal, size_t(nsignal) * 8, hipMemcpyHostToDevice));
  l.nsig = nsignal;
  l.sig_cap = std::max(nsignal, 1);
  return 0;
}

// Write values[0..n) as entries step0.. of a device table of *len entries in
// use and *cap allocated, growing it (zero-filled, contents kept) when needed.
int table_write(fdtd_grid* g, double** table, int* len, int* cap, long step0, const double* values, int n) {
  const long end = step0 + n;
  if (end > (1L << 30)) return fail(FDTD_EINVAL, "table longer than 2^30 steps");
  if (end > *cap) {
    const int ncap = int(std::max<long>(end, std::min<long>(2L * *cap, 1L << 30)));
    double* t = nullptr;
    int rc = g->alloc((void**)&t, size_t(ncap) * 8);
    if (rc) return rc;
    if (*len) HIP_TRY(hipMemcpy(t, *table, size_t(*len) * 8, hipMemcpyDeviceToDevice));
    g->release(*table);
    *table = t;
    *cap = ncap;
  }
  if (n) HIP_TRY(hipMemcpy(*table + step0, values, size_t(n) * 8, hipMemcpyHostToDevice));
  *len = int(std::max<long>(*len, end));
  return 0;
}

int fdtd_set_tfsf(fdtd_grid* g, const int box[6], const double kvec[3], const double pol[3], const double polH[3],
                  const double origin[3], int aux_ncells, const double* signal, int nsignal) {
  if (!g || !box || !kvec || !pol || !polH || !origin || (!signal && nsignal)) return fail(FDTD_EINVAL, "null argument");
  if (g->tfsf) return fail(FDTD_ESTATE, "TFSF already set");
  for (int a = 0; a < g->ndim; a++)
    if (box[2 * a] < 1 || box[2 * a + 1] > g->len[a] - 2 || box[2 * a] >= box[2 * a + 1])
      return fail(FDTD_EINVAL, "TFSF box must lie on interior cells with lo < hi");
  if (aux_ncells < 2 || (aux_ncells + 50) > 8192) return fail(FDTD_EINVAL, "bad aux line length");
  g->tfsf = true;
  for (int x = 0; x < 6; x++) g->box[x] = x < 2 * g->ndim ? box[x] : 0;
  // TFSFPlane normalises the wave vector (include/Source.hpp:566-570)
  double K = 0;
  for (int d = 0; d < g->ndim; d++) K += kvec[d] * kvec[d];
  K = std::sqrt(K);
  if (!(K > 0)) return fail(FDTD_EINVAL, "zero wave vector");
  for (int d = 0; d < 3; d++) {
    g->kvec[d] = d < g->ndim ? kvec[d] / K : 0.0;
    g->pol[d] = pol[d];
    g->polH[d] = polH[d];
    g->origin[d] = origin[d];
  }
  return aux_line_init(g, aux_ncells, signal, nsignal, g->aux);
}

int fdtd_set_signal(fdtd_grid* g, int kind, int index, long step0, const double* values, int n) {
  if (!g || (!values && n) || n < 0 || step0 < 0) return fail(FDTD_EINVAL, "bad argument");
  // queued steps may still read the table being rewritten or replaced
  HIP_TRY(hipStreamSynchronize(g->stream));
  if (kind == FDTD_SIGNAL_SOURCE) {
    if (index < 0 || index >= g->nsrc) return fail(FDTD_EINVAL, "no source %d", index);
    fdtdk::SrcDev& s = g->src[index];
    int rc = table_write(g, &g->src_table[index], &s.n, &g->src_cap[index], step0, values, n);
    gput(s.table, g->src_table[index]);
    return rc;
  }
  AuxLine* l = nullptr;
  if (kind == FDTD_SIGNAL_TFSF && index == 0 && g->tfsf) l = &g->aux;
  if (kind == FDTD_SIGNAL_PLANE && index >= 0 && index < g->nplane) l = &g->plane_aux[index];
  if (!l) return fail(FDTD_EINVAL, "no incident line of kind %d, index %d", kind, index);
  return table_write(g, &l->sig, &l->nsig, &l->sig_cap, step0, values, n);
}

int fdtd_add_tfsf_plane(fdtd_grid* g, int axis, int side, const int cells[6], const double kvec[3],
                        const double pol[3], const double polH[3], const double origin[3], int aux_ncells,
                        const double* signal, int nsignal) {
  if (!g || !cells || !kvec || !pol || !polH || !origin || (!signal && nsignal)) return fail(FDTD_EINVAL, "null argument");
  if (g->nplane >= FDTD_MAX_PLANES) return fail(FDTD_EINVAL, "too many TFSF planes (max %d)", FDTD_MAX_PLANES);
  if (axis < 0 || axis >= g->ndim || (side != FDTD_MIN && side != FDTD_MAX))
    return fail(FDTD_EINVAL, "bad plane axis or side");
  int c[6];
  for (int x = 0; x < 6; x++) c[x] = cells[x];
  for (int a = g->ndim; a < 3; a++) {  // axes the mode does not have: the one layer 0
    c[2 * a] = 0;
    c[2 * a + 1] = 1;
  }
  for (int a = 0; a < g->ndim; a++)
    if (c[2 * a] < 1 || c[2 * a + 1] > g->len[a] - 1 || c[2 * a] >= c[2 * a + 1])
      return fail(FDTD_EINVAL, "TFSF plane cells must be interior cells with lo < hi");
  if (c[2 * axis + 1] != c[2 * axis] + 1) return fail(FDTD_EINVAL, "a TFSF plane is one cell layer thick along its axis");
  if (aux_ncells < 2 || (aux_ncells + 50) > 8192) return fail(FDTD_EINVAL, "bad aux line length");
  double K = 0;
  for (int d = 0; d < g->ndim; d++) K += kvec[d] * kvec[d];
  K = std::sqrt(K);
  if (!(K > 0)) return fail(FDTD_EINVAL, "zero wave vector");
  fdtdk::PlaneDev P{};
  P.axis = axis;
  P.nrm = side == FDTD_MAX ? 1.0 : -1.0;
  for (int d = 0; d < 3; d++) {
    P.lo[d] = c[2 * d];
    P.hi[d] = c[2 * d + 1] - 1;
    // TFSFPlane normalises the wave vector (include/Source.hpp:566-570)
    P.kvec[d] = d < g->ndim ? kvec[d] / K : 0.0;
    P.pol[d] = pol[d];
    P.polH[d] = polH[d];
    P.origin[d] = origin[d];
  }
  AuxLine l;
  int rc = aux_line_init(g, aux_ncells, signal, nsignal, l);
  if (rc) return rc;
  P.auxL = l.L;
  g->plane[g->nplane] = P;
  g->plane_aux[g->nplane] = l;
  g->nplane++;
  return 0;
}

int fdtd_add_dft(fdtd_grid* g, const int region[6], int comp_mask, int nfreq, const double* phaseE,
                 const double* phaseH, int nphase, int* idx) {
  if (!g || !region || !phaseE || !phaseH || nfreq < 1 || nphase < 0) return fail(FDTD_EINVAL, "bad argument");
  if (g->ndft >= FDTD_MAX_DFT) return fail(FDTD_EINVAL, "too many DFT monitors (max %d)", FDTD_MAX_DFT);
  int r[6];
  for (int x = 0; x < 6; x++) r[x] = region[x];
  for (int a = g->ndim; a < 3; a++) {  // axes the mode does not have: the one layer 0
    r[2 * a] = 0;
    r[2 * a + 1] = 1;
  }
  for (int a = 0; a < 3; a++)
    if (r[2 * a] < 0 || r[2 * a + 1] > g->len[a] || r[2 * a] >= r[2 * a + 1]) return fail(FDTD_EINVAL, "bad DFT region");
  for (int c = 0; c < 6; c++)
    if (((comp_mask >> c) & 1) && !(c < 3 ? g->actE[c] : g->actH[c - 3]))
      return fail(FDTD_EINVAL, "DFT component %d does not exist in this mode", c);
  if (comp_mask <= 0 || comp_mask >= 64) return fail(FDTD_EINVAL, "bad component mask");
  int rc0 = dft_flush(g);
  if (rc0) return rc0;
  auto& d = g->dft[g->ndft];
  for (int x = 0; x < 6; x++) d.r[x] = r[x];
  d.mask = comp_mask;
  d.nfreq = nfreq;
  d.nphase = nphase;
  d.ncomp = __builtin_popcount(comp_mask);
  d.ncells = long(r[1] - r[0]) * (r[3] - r[2]) * (r[5] - r[4]);
  if (d.ncells * d.ncomp * fdtdk::kDftBlock >= (1L << 31)) return fail(FDTD_EINVAL, "DFT monitor too large");
  int rc = g->alloc(&d.acc, size_t(d.ncomp) * nfreq * 2 * d.ncells * 8);  // FP64 accumulators
  if (!rc) rc = g->alloc(&d.smp, size_t(fdtdk::kDftBlock) * d.ncomp * d.ncells * g->esz);
  if (!rc) rc = g->alloc((void**)&d.phE, size_t(std::max(nphase, 1)) * nfreq * 2 * 8);
  if (!rc) rc = g->alloc((void**)&d.phH, size_t(std::max(nphase, 1)) * nfreq * 2 * 8);
  if (rc) return rc;
  if (nphase) {
    HIP_TRY(hipMemcpy(d.phE, phaseE, size_t(nphase) * nfreq * 2 * 8, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(d.phH, phaseH, size_t(nphase) * nfreq * 2 * 8, hipMemcpyHostToDevice));
  }
  if (idx) *idx = g->ndft;
  g->ndft++;
  return g->finalized ? ensure_dft_im(g) : 0;
}

int fdtd_get_dft(fdtd_grid* g, int idx, double* out) {
  if (!g || !out || idx < 0 || idx >= g->ndft) return fail(FDTD_EINVAL, "bad DFT index");
  int rc = dft_flush(g);
  if (rc) return rc;
  HIP_TRY(hipStreamSynchronize(g->stream));
  const auto& d = g->dft[idx];
  const size_t n = size_t(d.ncomp) * d.nfreq * 2 * d.ncells;
  HIP_TRY(hipMemcpy(out, d.acc, n * 8, hipMemcpyDeviceToHost));
  if (!g->cplx || !d.acc_im) return 0;  // before the first step the imaginary part is zero
  // complex field: F = F_re + i F_im, each [comp][freq][re, im][cell]
  std::vector<double> fi(n);
  HIP_TRY(hipMemcpy(fi.data(), d.acc_im, n * 8, hipMemcpyDeviceToHost));
  for (size_t b = 0; b < size_t(d.ncomp) * d.nfreq; b++) {
    double* o = out + b * 2 * d.ncells;
    const double* q = fi.data() + b * 2 * d.ncells;
    for (long c = 0; c < d.ncells; c++) {
      o[c] -= q[d.ncells + c];
      o[d.ncells + c] += q[c];
    }
  }
  return 0;
}

int fdtd_add_probe(fdtd_grid* g, int field, int comp, int i, int j, int k) {
  if (!g) return fail(FDTD_EINVAL, "null grid");
  if (g->nprobe >= FDTD_MAX_PROBES) return fail(FDTD_EINVAL, "too many probes (max %d)", FDTD_MAX_PROBES);
  if ((field != FDTD_E && field != FDTD_H) || !valid_comp(g, field, comp)) return fail(FDTD_EINVAL, "probes record E or H components");
  if (!interior_cell(g, i, j, k)) return fail(FDTD_EINVAL, "probe cell must be an interior cell");
  if (g->step > 0) return fail(FDTD_ESTATE, "probes must be added before stepping");
  g->probe[g->nprobe].cell = g->q(i, j, k);
  g->probe[g->nprobe].field = field;
  g->probe[g->nprobe].comp = comp;
  g->nprobe++;
  return 0;
}

int fdtd_get_probes(fdtd_grid* g, double* out, int max_steps, int* nsteps_out) {
  if (!g || !out) return fail(FDTD_EINVAL, "null argument");
  HIP_TRY(hipStreamSynchronize(g->stream));
  long n = std::min<long>(g->step, max_steps);
  if (n > 0 && g->nprobe) HIP_TRY(hipMemcpy(out, g->trace, size_t(n) * g->nprobe * 8, hipMemcpyDeviceToHost));
  if (nsteps_out) *nsteps_out = int(n);
  return 0;
}

int fdtd_get_probes_imag(fdtd_grid* g, double* out, int max_steps, int* nsteps_out) {
  if (!g || !out) return fail(FDTD_EINVAL, "null argument");
  if (!g->cplx) return fail(FDTD_EINVAL, "the grid has real fields (no Bloch-periodic face)");
  HIP_TRY(hipStreamSynchronize(g->stream));
  long n = std::min<long>(g->step, max_steps);
  if (n > 0 && g->nprobe) HIP_TRY(hipMemcpy(out, g->trace_im, size_t(n) * g->nprobe * 8, hipMemcpyDeviceToHost));
  if (nsteps_out) *nsteps_out = int(n);
  return 0;
}

int fdtd_upload_field(fdtd_grid* g, int field, int comp, int pole, const double* host) {
  if (!g || !host) return fail(FDTD_EINVAL, "null argument");
  if (field & FDTD_IMAG) {
    field &= ~FDTD_IMAG;
    if (!g->cplx || !g->finalized) return fail(FDTD_ESTATE, "the imaginary part exists once a grid with a Bloch face is stepped");
    if (!valid_comp(g, field, comp) || (field != FDTD_E && field != FDTD_H))
      return fail(FDTD_EINVAL, "only E and H can be uploaded");
    HIP_TRY(hipStreamSynchronize(g->stream));
    return upload_T(g, field == FDTD_E ? g->E2[comp] : g->H2[comp], host, g->ncell);
  }
  if (!valid_comp(g, field, comp)) return fail(FDTD_EINVAL, "component does not exist in this mode");
  HIP_TRY(hipStreamSynchronize(g->stream));
  if (field == FDTD_E) return upload_T(g, g->E[comp], host, g->ncell);
  if (field == FDTD_H) return upload_T(g, g->H[comp], host, g->ncell);
  return fail(FDTD_EINVAL, "only E and H can be uploaded");
}

int fdtd_download_field(fdtd_grid* g, int field, int comp, int pole, double* host) {
  if (!g || !host) return fail(FDTD_EINVAL, "null argument");
  if (field & FDTD_IMAG) {
    field &= ~FDTD_IMAG;
    if (!g->cplx) return fail(FDTD_EINVAL, "the grid has real fields (no Bloch-periodic face)");
    if (!valid_comp(g, field, comp) || (field != FDTD_E && field != FDTD_H))
      return fail(FDTD_EINVAL, "the imaginary part is read for E and H");
    if (!g->finalized) {
      std::fill(host, host + g->ncell, 0.0);
      return 0;
    }
    HIP_TRY(hipStreamSynchronize(g->stream));
    return download_T(g, host, field == FDTD_E ? g->E2[comp] : g->H2[comp], g->ncell);
  }
  if (!valid_comp(g, field, comp)) return fail(FDTD_EINVAL, "component does not exist in this mode");
  HIP_TRY(hipStreamSynchronize(g->stream));
  if (field == FDTD_E) return download_T(g, host, g->E[comp], g->ncell);
  if (field == FDTD_H) return download_T(g, host, g->H[comp], g->ncell);
  if (field == FDTD_B || field == FDTD_M || field == FDTD_K) {
    // mu_r mu0 H (B) or 0 (M, K), and the magnetic-pole box's stored values on its pole cells
    if (field == FDTD_B) {
      int rc = download_T(g, host, g->H[comp], g->ncell);
      if (rc) return rc;
      for (long i = 0; i < g->ncell; i++) host[i] *= g->mats[g->mat_host[i]].mu * kMu0;
    } else {
      std::fill(host, host + g->ncell, 0.0);
      if (pole < 0 || pole >= FDTD_MAX_POLES) return fail(FDTD_EINVAL, "bad pole index");
    }
    void* src = field == FDTD_B ? g->Bd[comp] : (field == FDTD_M ? g->Md[pole][comp] : g->Kd[pole][comp]);
    if (!src || g->max_mnp == 0) return 0;
    const long nb = long(g->mbox[1] - g->mbox[0]) * (g->mbox[3] - g->mbox[2]) * (g->mbox[5] - g->mbox[4]);
    std::vector<double> bx(nb);
    int rc = download_T(g, bx.data(), src, nb);
    if (rc) return rc;
    long t = 0;
    for (int k = g->mbox[4]; k < g->mbox[5]; k++)
      for (int j = g->mbox[2]; j < g->mbox[3]; j++)
        for (int i = g->mbox[0]; i < g->mbox[1]; i++, t++) {
          const long qq = g->q(i, j, k);
          if (g->mats[g->mat_host[qq]].mnp > 0) host[qq] = bx[t];
        }
    return 0;
  }
  // D, P, J: dispersive box values where stored, eps_r eps0 E (D) or 0 elsewhere
  if (field == FDTD_D) {
    int rc = download_T(g, host, g->E[comp], g->ncell);
    if (rc) return rc;
    for (int k = 0; k < g->d.nz; k++)
      for (int j = 0; j < g->d.ny; j++)
        for (int i = 0; i < g->d.nx; i++) host[g->q(i, j, k)] *= cell_eps(g, i, j, k) * kEps0;
  } else {
    std::fill(host, host + g->ncell, 0.0);
    if (pole < 0 || pole >= FDTD_MAX_POLES) return fail(FDTD_EINVAL, "bad pole index");
  }
  void* src = field == FDTD_D ? g->Dd[comp] : (field == FDTD_P ? g->Pd[pole][comp] : g->Jd[pole][comp]);
  if (!src || g->max_np == 0) return 0;
  long nb = long(g->dbox[1] - g->dbox[0]) * (g->dbox[3] - g->dbox[2]) * (g->dbox[5] - g->dbox[4]);
  std::vector<double> bx(nb);
  int rc = download_T(g, bx.data(), src, nb);
  if (rc) return rc;
  long t = 0;
  for (int k = g->dbox[4]; k < g->dbox[5]; k++)
    for (int j = g->dbox[2]; j < g->dbox[3]; j++)
      for (int i = g->dbox[0]; i < g->dbox[1]; i++, t++) {
        long qq = g->q(i, j, k);
        if (g->mats[g->mat_host[qq]].np > 0) host[qq] = bx[t];
      }
  return 0;
}

int fdtd_step(fdtd_grid* g, int nsteps) {
  if (!g || nsteps < 0) return fail(FDTD_EINVAL, "bad argument");
  if (nsteps == 0) return 0;
  if (!g->finalized) {
    int rc = finalize(g);
    if (rc) return rc;
  }
  return g->f64 ? step_impl<double>(g, nsteps) : step_impl<float>(g, nsteps);
}

int fdtd_synchronize(fdtd_grid* g) {
  if (!g) return fail(FDTD_EINVAL, "null grid");
  HIP_TRY(hipStreamSynchronize(g->stream));
  return collect_events(g);
}

long fdtd_step_count(const fdtd_grid* g) { return g ? g->step : -1; }

int fdtd_set_tile(fdtd_grid* g, int bx, int by, int kc) {
  if (!g) return fail(FDTD_EINVAL, "null grid");
  if (bx == 0 && g->ndim == 2) {
    if ((by != 2 && by != 4 && by != 6 && by != 8) || kc < 1 || kc + 2 * by > fdtdk::kTmMaxRegion)
      return fail(FDTD_EINVAL, "the time-blocked TM schedule (bx = 0) takes by in {2, 4, 6, 8} steps per launch and "
                               "cores of kc x kc cells with kc + 2 by <= %d", fdtdk::kTmMaxRegion);
  } else if (bx == 0) {
    if (g->ndim != 3 || by != fdtdk::kSRows || kc < 1 || kc > 2)
      return fail(FDTD_EINVAL,
                  "the streamed fused schedule (bx = 0) is 3-D only, with by = %d E-region rows and kc in 1..2 "
                  "workgroups per compute unit (FP64 grids, and grids of more than %d planes, run it as the default "
                  "two-pass tile)", fdtdk::kSRows, fdtdk::kSMaxZ);
  } else if (bx == 1) {
    if (g->ndim != 3 || (by != 4 && by != 8 && by != 16) || kc <= 0)
      return fail(FDTD_EINVAL, "the z-marching fused schedule (bx = 1) is 3-D only, with by in {4, 8, 16} rows of 64 "
                               "columns and kc >= 1 planes per workgroup");
  } else if (bx == 2) {
    if (g->ndim != 3 || (by != 8 && by != 12 && by != 16) || kc <= 0 || kc > fdtdk::kTbMaxKc)
      return fail(FDTD_EINVAL, "the time-blocked schedule (bx = 2) is 3-D only, with by in {8, 12, 16} rows of 64 "
                               "columns (FP32 runs 16 as 12, FP64 every by as 8 with kc <= 32) and 1 <= kc <= %d planes per workgroup",
                  fdtdk::kTbMaxKc);
  } else if (bx < 8 || by <= 0 || kc <= 0 || bx * by > 256 || (bx * by) % 64 != 0) {
    return fail(FDTD_EINVAL, "a two-pass tile has bx >= 8, a multiple of 64 threads, at most 256, and kc >= 1");
  }
  g->tile = Tile{bx, by, kc};
  return 0;
}

int fdtd_set_adjoint_tile(fdtd_grid* g, int bx, int by, int kc) {
  if (!g) return fail(FDTD_EINVAL, "null grid");
  if (bx == 0 && by == 0 && kc == 0) {
    g->rtile = Tile{0, 0, 0};
    return 0;
  }
  if (bx < 8 || by <= 0 || kc <= 0 || bx * by > 256 || (bx * by) % 64 != 0)
    return fail(FDTD_EINVAL,
                "the reverse sweep takes a two-pass tile: bx >= 8, a multiple of 64 threads, at most 256, kc >= 1");
  g->rtile = Tile{bx, by, kc};
  return 0;
}

int fdtd_autotune(fdtd_grid* g, int trial_steps, int* bx, int* by, int* kc) {
  if (!g || trial_steps < 1) return fail(FDTD_EINVAL, "bad argument");
  int rc = 0;
  if (!g->finalized && (rc = finalize(g))) return rc;
  if (g->ndim == 1) {  // TEM: one schedule (a thread per cell), nothing to choose
    const Tile t = g->tile.by ? g->tile : kDefaultTileTM;
    if (bx) *bx = t.bx;
    if (by) *by = t.by;
    if (kc) *kc = t.kc;
    return 0;
  }
  static const Tile cand3[] = {{0, 5, 1},  {0, 5, 2},  {256, 1, 1}, {256, 1, 2}, {256, 1, 4}, {256, 1, 8},
                               {128, 2, 2}, {128, 2, 4}, {64, 4, 1}, {64, 4, 2}, {1, 8, 16}, {1, 8, 32},
                               {1, 16, 16}, {1, 4, 32}, {2, 12, 16}, {2, 12, 32}};
  static const Tile cand2[] = {{256, 1, 1}, {128, 2, 1}, {64, 4, 1}, {64, 1, 1}, {128, 1, 1},
                               {0, 2, 16},  {0, 4, 16},  {0, 4, 24}, {0, 8, 16}, {0, 4, 8}};
  const Tile* cand = g->ndim == 3 ? cand3 : cand2;
  const int ncand = g->ndim == 3 ? int(sizeof cand3 / sizeof cand3[0]) : int(sizeof cand2 / sizeof cand2[0]);
  if ((rc = ensure_trace(g, g->step + long(ncand) * (2 * trial_steps + 1)))) return rc;
  if (g->ndim == 3 && (rc = ensure_pingpong(g))) return rc;
  GridSnapshot snap;
  if ((rc = snapshot_take(g, snap))) return rc;
  const bool prof = g->prof;
  double kms[kNumKernels];
  long kcount[kNumKernels];
  std::copy(g->kms, g->kms + kNumKernels, kms);
  std::copy(g->kcount, g->kcount + kNumKernels, kcount);
  const Tile saved = g->tile;
  double best = 1e300;
  Tile bt = g->ndim == 3 ? kDefaultTile3D : cand[0];
  for (int c = 0; c < ncand && !rc; c++) {
    if (g->ndim == 3 && cand[c].bx == 0 && g->f64) continue;  // the streamed pass is FP32 only
    g->tile = cand[c];
    double t = 0;  // time per step
    if (g->ndim == 2) {
      // TM steps are bound by launches: wall time of a run after one of the
      // same length (which captures its graphs, if any)
      if ((rc = fdtd_step(g, trial_steps))) break;
      if ((rc = fdtd_synchronize(g))) break;
      const auto t0 = std::chrono::steady_clock::now();
      if ((rc = fdtd_step(g, trial_steps))) break;
      if ((rc = fdtd_synchronize(g))) break;
      t = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count() / trial_steps;
    } else {
      if ((rc = fdtd_step(g, 1))) break;
      if ((rc = fdtd_set_profiling(g, 1))) break;
      if ((rc = fdtd_step(g, trial_steps))) break;
      for (int kid : {0, 1, 3, 4}) {  // device time of the step's kernels
        double ms = 0;
        long nl = 0;
        if ((rc = fdtd_kernel_time(g, kid, &ms, &nl))) break;
        t += ms / trial_steps;
      }
    }
    if (rc) break;
    if (t < best) {
      best = t;
      bt = cand[c];
    }
  }
  (void)fdtd_set_profiling(g, 0);
  snapshot_restore(g, snap);
  g->prof = prof;
  std::copy(kms, kms + kNumKernels, g->kms);
  std::copy(kcount, kcount + kNumKernels, g->kcount);
  if (rc) {
    g->tile = saved;
    return rc;
  }
  g->tile = bt;
  if (bx) *bx = bt.bx;
  if (by) *by = bt.by;
  if (kc) *kc = bt.kc;
  return 0;
}

int fdtd_get_tile(fdtd_grid* g, int* bx, int* by, int* kc) {
  if (!g) return fail(FDTD_EINVAL, "null grid");
  Tile t = g->tile.by ? g->tile : (g->ndim == 3 ? kDefaultTile3D : kDefaultTileTM);
  if (bx) *bx = t.bx;
  if (by) *by = t.by;
  if (kc) *kc = t.kc;
  return 0;
}

int fdtd_set_profiling(fdtd_grid* g, int on) {
  if (!g) return fail(FDTD_EINVAL, "null grid");
  int rc = collect_events(g);
  g->prof = on != 0;
  for (int k = 0; k < kNumKernels; k++) {
    g->kms[k] = 0;
    g->kcount[k] = 0;
  }
  return rc;
}

int fdtd_kernel_time(fdtd_grid* g, int kernel, double* ms, long* launches) {
  if (!g || kernel < 0 || kernel >= kNumKernels) return fail(FDTD_EINVAL, "bad argument");
  int rc = collect_events(g);
  if (rc) return rc;
  if (ms) *ms = g->kms[kernel];
  if (launches) *launches = g->kcount[kernel];
  return 0;
}

int fdtd_set_boundary(fdtd_grid* g, int axis, int side, int kind) {
  if (!g) return fail(FDTD_EINVAL, "null grid");
  if (axis < 0 || axis >= g->ndim || side < 0 || side > 1) return fail(FDTD_EINVAL, "bad axis or side");
  if (kind == FDTD_BC_BLOCH)
    return fail(FDTD_EINVAL, "Bloch-periodic faces take a complex phase factor: use fdtd_set_bloch_boundary");
  if (kind < FDTD_BC_NONE || kind > FDTD_BC_PERIODIC) return fail(FDTD_EINVAL, "bad boundary kind");
  if (g->finalized) return fail(FDTD_ESTATE, "boundary conditions are fixed once stepping has started");
  g->nbc += (kind != FDTD_BC_NONE) - (g->bc[axis][side] != FDTD_BC_NONE);
  g->bc[axis][side] = kind;
  g->cplx = false;
  for (int a = 0; a < g->ndim; a++) g->cplx = g->cplx || g->bc[a][0] == FDTD_BC_BLOCH || g->bc[a][1] == FDTD_BC_BLOCH;
  return 0;
}

int fdtd_set_bloch_boundary(fdtd_grid* g, int axis, int side, double factor_re, double factor_im) {
  if (!g) return fail(FDTD_EINVAL, "null grid");
  if (g->ndim == 1) return fail(FDTD_EINVAL, "Bloch-periodic faces are for 2-D and 3-D grids");
  if (axis < 0 || axis >= g->ndim || side < 0 || side > 1) return fail(FDTD_EINVAL, "bad axis or side");
  if (g->finalized) return fail(FDTD_ESTATE, "boundary conditions are fixed once stepping has started");
  g->nbc += (g->bc[axis][side] == FDTD_BC_NONE);
  g->bc[axis][side] = FDTD_BC_BLOCH;
  g->bloch[axis][side][0] = factor_re;
  g->bloch[axis][side][1] = factor_im;
  g->cplx = true;
  return 0;
}

int fdtd_surface_quantities(fdtd_grid* g, const int region[6], double* out) {
  if (!g || !region || !out) return fail(FDTD_EINVAL, "null argument");
  if (g->ndim == 1) return fail(FDTD_EINVAL, "the reference defines surface quantities for TE, TM and 3-D");
  for (int a = 0; a < 3; a++) {
    const int lo = a < g->ndim ? 1 : 0, hi = a < g->ndim ? g->len[a] - 1 : 1;
    if (region[2 * a] < lo || region[2 * a + 1] > hi || region[2 * a] >= region[2 * a + 1])
      return fail(FDTD_EINVAL, "region must be a non-empty box of interior cells");
  }
  int rc = 0;
  if (!g->finalized && (rc = finalize(g))) return rc;
  const fdtdk::SurfRegion r{region[0], region[1], region[2], region[3], region[4], region[5]};
  const long cells = long(r.x1 - r.x0) * (r.y1 - r.y0) * (r.z1 - r.z0);
  return g->f64 ? surface_impl<double>(g, r, cells, out) : surface_impl<float>(g, r, cells, out);
}

int fdtd_surface_quantities_static(fdtd_grid* g, const int region[6], double* out) {
  if (!g || !region || !out) return fail(FDTD_EINVAL, "null argument");
  if (g->ndim == 1) return fail(FDTD_EINVAL, "the reference defines surface quantities for TE, TM and 3-D");
  for (int a = 0; a < 3; a++) {
    const int lo = a < g->ndim ? 1 : 0, hi = a < g->ndim ? g->len[a] - 1 : 1;
    if (region[2 * a] < lo || region[2 * a + 1] > hi || region[2 * a] >= region[2 * a + 1])
      return fail(FDTD_EINVAL, "region must be a non-empty box of interior cells");
  }
  const fdtdk::SurfRegion r{region[0], region[1], region[2], region[3], region[4], region[5]};
  const long cells = long(r.x1 - r.x0) * (r.y1 - r.y0) * (r.z1 - r.z0);
  if (cells > kSurfStaticMax)
    return fail(FDTD_EINVAL, "region of %ld cells exceeds FDTD_SURFACE_STATIC_MAX_CELLS (%ld)", cells, kSurfStaticMax);
  // the device set-up allocates: it must have run already
  if (!g->finalized) return fail(FDTD_ESTATE, "grid not set up on the device yet: step it (or call fdtd_surface_quantities) once first");
  return g->f64 ? surface_static_impl<double>(g, r, cells, out) : surface_static_impl<float>(g, r, cells, out);
}

int fdtd_set_graph(fdtd_grid* g, int on) {
  if (!g) return fail(FDTD_EINVAL, "null grid");
  if (on && !g->stream) {
    // capture needs a stream other than the legacy default one; a blocking
    // stream still orders itself with work on the default stream
    if (!g->own_stream && hipStreamCreate(&g->own_stream) != hipSuccess)
      return fail(FDTD_EDEVICE, "hipStreamCreate failed");
    g->stream = g->own_stream;
  }
  g->use_graph = on != 0;
  return 0;
}

int fdtd_set_design(fdtd_grid* g, const int box[6], const double* eps_r) {
  if (!g || !box || !eps_r) return fail(FDTD_EINVAL, "null argument");
  int b[6];
  for (int x = 0; x < 6; x++) b[x] = box[x];
  for (int a = g->ndim; a < 3; a++) {  // axes the mode does not have: the one layer 0
    b[2 * a] = 0;
    b[2 * a + 1] = 1;
  }
  for (int a = 0; a < g->ndim; a++)
    if (b[2 * a] < 1 || b[2 * a + 1] > g->len[a] - 1 || b[2 * a] >= b[2 * a + 1])
      return fail(FDTD_EINVAL, "design box must cover interior cells only, with lo < hi");
  const long n = long(b[1] - b[0]) * (b[3] - b[2]) * (b[5] - b[4]);
  for (long i = 0; i < n; i++)
    if (!(eps_r[i] > 0)) return fail(FDTD_EINVAL, "design eps_r must be positive");
  HIP_TRY(hipStreamSynchronize(g->stream));
  if (g->design && n != g->ndes) {
    g->release(g->dsg_ie_dev);
    g->dsg_ie_dev = nullptr;
  }
  if (!g->dsg_ie_dev) {
    int rc = g->alloc(&g->dsg_ie_dev, size_t(n) * g->esz);
    if (rc) return rc;
  }
  std::vector<double> ie(n);
  for (long i = 0; i < n; i++) ie[i] = 1.0 / (eps_r[i] * kEps0);
  int rc = upload_T(g, g->dsg_ie_dev, ie.data(), n);
  if (rc) return rc;
  g->dsg_eps.assign(eps_r, eps_r + n);
  for (int x = 0; x < 6; x++) g->dsg[x] = b[x];
  g->ndes = n;
  g->design = true;
  return 0;
}

int fdtd_adjoint_eps_gradient(fdtd_grid* g, int nsteps, int checkpoint_every, double history_mb, int dft,
                              int comp_mask, double* objective, double* grad) {
  if (!g || !objective || !grad) return fail(FDTD_EINVAL, "null argument");
  if (nsteps < 1 || checkpoint_every < 1 || !(history_mb >= 0)) return fail(FDTD_EINVAL, "bad step counts or budget");
  if (g->ndim != 3) return fail(FDTD_EINVAL, "the adjoint gradient is implemented for 3-D grids");
  if (!g->design) return fail(FDTD_ESTATE, "no design box (fdtd_set_design)");
  if (g->step != 0) return fail(FDTD_ESTATE, "the adjoint gradient runs from the initial state (step 0)");
  if (dft < 0 || dft >= g->ndft) return fail(FDTD_EINVAL, "bad DFT monitor index");
  if (comp_mask < 1 || comp_mask > 7 || (comp_mask & ~g->dft[dft].mask))
    return fail(FDTD_EINVAL, "objective components must be E components recorded by the monitor");
  int rc = 0;
  if (!g->finalized && (rc = finalize(g))) return rc;
  for (long c = 0; g->max_np > 0 && c < g->ncell; c++) {
    const MatHost& m = g->mats[g->mat_host[c]];
    if (m.np > 0 && (m.mag || m.fluid))
      return fail(FDTD_EINVAL, "the adjoint gradient covers Drude, Lorentz and conductive poles, not magnetized or "
                               "fluid Drude materials");
  }
  if (g->max_mnp > 0) return fail(FDTD_EINVAL, "the adjoint gradient does not cover magnetic poles");
  if (g->nbc) return fail(FDTD_EINVAL, "the adjoint gradient does not cover boundary conditions");
  if (g->nsrc_eh || g->nplane)
    return fail(FDTD_EINVAL, "the adjoint gradient covers D / B sources and the TFSF box, not E / H sources or single planes");
  return g->f64 ? adjoint_impl<double>(g, nsteps, checkpoint_every, history_mb, dft, comp_mask, objective, grad)
                : adjoint_impl<float>(g, nsteps, checkpoint_every, history_mb, dft, comp_mask, objective, grad);
}


int fdtd_adjoint_eps_gradient_batch(fdtd_grid* const* grids, int n, int nsteps, int checkpoint_every,
                                    double history_mb, int dft, int comp_mask, double* objectives,
                                    double* const* grads) {
  if (!grids || n < 1 || !objectives || !grads) return fail(FDTD_EINVAL, "null argument");
  for (int q = 0; q < n; q++) {
    if (!grids[q] || !grads[q]) return fail(FDTD_EINVAL, "null grid or gradient buffer in batch");
    for (int r = 0; r < q; r++)
      if (grids[r] == grids[q]) return fail(FDTD_EINVAL, "a grid appears twice in the batch");
  }
  int dev = 0;
  HIP_TRY(hipGetDevice(&dev));
  // one host thread and one non-blocking stream per design, so the designs'
  // kernels share the GPU (a 128^3 design leaves most of it idle)
  std::vector<hipStream_t> own(n, nullptr), saved(n, nullptr);
  int rc = 0;
  for (int q = 0; q < n && !rc; q++) {
    saved[q] = grids[q]->stream;
    if (!saved[q]) {
      if (hipStreamCreateWithFlags(&own[q], hipStreamNonBlocking) != hipSuccess)
        rc = fail(FDTD_EDEVICE, "hipStreamCreateWithFlags failed");
      else
        grids[q]->stream = own[q];
    }
  }
  // device allocations that synchronise the device happen here, before the
  // designs run side by side
  for (int q = 0; q < n && !rc; q++) {
    if (!grids[q]->finalized) rc = finalize(grids[q]);
    if (!rc && grids[q]->ndim == 3) rc = ensure_pingpong(grids[q]);
  }
  std::vector<int> rcs(n, 0);
  std::vector<std::string> errs(n);
  if (!rc) {
    std::vector<std::thread> th;
    for (int q = 0; q < n; q++)
      th.emplace_back([&, q]() {
        if (hipSetDevice(dev) != hipSuccess) {
          rcs[q] = FDTD_EDEVICE;
          errs[q] = "hipSetDevice failed";
          return;
        }
        rcs[q] = fdtd_adjoint_eps_gradient(grids[q], nsteps, checkpoint_every, history_mb, dft, comp_mask,
                                           &objectives[q], grads[q]);
        if (rcs[q]) errs[q] = g_err;
      });
    for (auto& t : th) t.join();
  }
  for (int q = 0; q < n; q++) {
    if (own[q]) {
      (void)hipStreamSynchronize(own[q]);
      (void)hipStreamDestroy(own[q]);
    }
    grids[q]->stream = saved[q];
  }
  if (rc) return rc;
  for (int q = 0; q < n; q++)
    if (rcs[q]) return fail(rcs[q], "design %d: %s", q, errs[q].c_str());
  return 0;
}

}  // extern "C"

// ---------------------------------------------------------------- batches
// Independent grids of one precision stepped together: one launch per pass
// covers the tiles of every grid (a tile table maps workgroups to grids), and
// each chunk of steps is captured once in a hipGraph and replayed.
struct fdtd_batch {
  std::vector<fdtd_grid*> grids;
  std::vector<hipStream_t> saved_streams;
  bool f64 = false;
  Tile tile = {64, 4, 2};
  hipStream_t stream = nullptr;
  int4* tiles = nullptr;
  int ntiles = 0;
  long* dft_offs = nullptr;  // n+1 prefix sums of each grid's DFT items
  long dft_total = 0;
  // the time-blocked DFT: n+1 prefix sums of each grid's sampler and fold blocks
  long *smp_offs = nullptr, *fold_offs = nullptr;
  long smp_total = 0, fold_total = 0;
  static constexpr int kRing = 2;
  char* args_dev = nullptr;
  char* args_host = nullptr;
  hipEvent_t ev[kRing] = {};
  bool ev_used[kRing] = {};
  int slot = 0;
  bool use_graph = true;
  // persistent passes: pwg workgroups per compute unit walk the tile table
  // (0: one workgroup per tile); ncu compute units
  int pwg = 0, ncu = 256;
  // (slot, steps, DFT block phase or -1) -> executable graph
  std::map<std::tuple<int, int, int>, hipGraphExec_t> graphs;

  ~fdtd_batch() {
    if (stream) (void)hipStreamSynchronize(stream);
    for (auto& kv : graphs) (void)hipGraphExecDestroy(kv.second);
    for (auto e : ev)
      if (e) (void)hipEventDestroy(e);
    if (tiles) (void)hipFree(tiles);
    if (dft_offs) (void)hipFree(dft_offs);
    if (smp_offs) (void)hipFree(smp_offs);
    if (fold_offs) (void)hipFree(fold_offs);
    if (args_dev) (void)hipFree(args_dev);
    if (args_host) (void)hipHostFree(args_host);
    for (size_t q = 0; q < grids.size(); q++) grids[q]->stream = saved_streams[q];
    if (stream) (void)hipStreamDestroy(stream);
  }
};

namespace {

// (Re)build the tile table of a batch for tile t: every grid's two-pass
// tiles, grid by grid; captured graphs refer to the old table and are dropped.
int batch_tiles(fdtd_batch* b, Tile t) {
  std::vector<int4> tab;
  for (size_t q = 0; q < b->grids.size(); q++) {
    const fdtd_grid* g = b->grids[q];
    const int tx = (g->d.nx + t.bx - 1) / t.bx, ty = (g->d.ny - 2 + t.by - 1) / t.by;
    const int tz = (g->d.nz - 2 + t.kc - 1) / t.kc;
    for (int z = 0; z < tz; z++)
      for (int y = 0; y < ty; y++)
        for (int x = 0; x < tx; x++) tab.push_back(make_int4(int(q), x, y, z));
  }
  HIP_TRY(hipStreamSynchronize(b->stream));
  for (auto& kv : b->graphs) (void)hipGraphExecDestroy(kv.second);
  b->graphs.clear();
  if (b->tiles) (void)hipFree(b->tiles);
  b->tiles = nullptr;
  if (hipMalloc((void**)&b->tiles, tab.size() * sizeof(int4)) != hipSuccess) return fail(FDTD_ENOMEM, "tile table");
  HIP_TRY(hipMemcpy(b->tiles, tab.data(), tab.size() * sizeof(int4), hipMemcpyHostToDevice));
  b->ntiles = int(tab.size());
  b->tile = t;
  return 0;
}

template <typename T>
int batch_step_impl(fdtd_batch* b, int nsteps) {
  const size_t n = b->grids.size();
  for (fdtd_grid* g : b->grids) {
    int rc = 0;
    if (has_magnetized(g)) return fail(FDTD_EINVAL, "batched grids cannot hold magnetized Drude materials (step them alone)");
    if (g->nbc || g->nsrc_eh || g->nplane)
      return fail(FDTD_EINVAL, "batched grids cannot have boundary conditions, E / H sources or single TFSF planes");
    if (g->mats_dirty && (rc = upload_materials<T>(g))) return rc;
    if (g->coef_dirty && (rc = upload_coefs(g))) return rc;
    if ((rc = ensure_trace(g, g->step + nsteps))) return rc;
  }
  {
    long items = 0, sblocks = 0;
    for (fdtd_grid* g : b->grids)
      for (int m = 0; m < g->ndft; m++) {
        items += long(g->dft[m].nfreq) * g->dft[m].ncells;
        sblocks += fdtdk::fold_segment(g->dft[m].ncells) / 256;
      }
    if (items != b->dft_total || sblocks != b->smp_total)
      return fail(FDTD_ESTATE, "a grid's DFT monitors changed after the batch was made (make a new batch)");
  }
  // the time-blocked DFT (samples every step, folded once per kDftBlock
  // steps) when every grid is at the same phase of its block; otherwise the
  // grids' pending samples are folded and the sums run step by step
  int phase = b->smp_total ? int(b->grids[0]->step % fdtdk::kDftBlock) : -1;
  for (fdtd_grid* g : b->grids)
    if (g->step % fdtdk::kDftBlock != phase) phase = -1;
  if (phase < 0)
    for (fdtd_grid* g : b->grids) {
      int rc = dft_flush(g);
      if (rc) return rc;
    }
  int done = 0;
  while (done < nsteps) {
    const int chunk = std::min(kChunk, nsteps - done);
    for (fdtd_grid* g : b->grids)
      if (g->tfsf || g->nplane) {
        int rc = run_aux(g, g->step, chunk);
        if (rc) return rc;
      }
    const int slot = b->slot;
    b->slot = (slot + 1) % fdtd_batch::kRing;
    if (b->ev_used[slot]) HIP_TRY(hipEventSynchronize(b->ev[slot]));
    auto* host = reinterpret_cast<fdtdk::StepArgs<T>*>(b->args_host + slot * n * sizeof(fdtdk::StepArgs<T>));
    for (size_t q = 0; q < n; q++) {
      fdtd_grid* g = b->grids[q];
      fdtdk::StepArgs<T> a;
      fill_args<T>(g, a);
      a.step0 = g->step;
      a.dft_done0 = g->dft_done;
      gput(a.incE, g->aux.incE);
      gput(a.incH, g->aux.incH);
      gput(a.self, reinterpret_cast<const fdtdk::StepArgs<T>*>(b->args_dev + (slot * n + q) * sizeof(fdtdk::StepArgs<T>)));
      host[q] = a;
    }
    auto* dev = reinterpret_cast<const fdtdk::StepArgs<T>*>(b->args_dev + slot * n * sizeof(fdtdk::StepArgs<T>));
    HIP_TRY(hipMemcpyAsync((void*)dev, host, n * sizeof(fdtdk::StepArgs<T>), hipMemcpyHostToDevice, b->stream));
    const dim3 block(b->tile.bx, b->tile.by, 1), grid(b->ntiles, 1, 1);
    // persistent passes: a multiple of 8 workgroups (one set per XCD), at most one per tile
    const dim3 pgrid(unsigned(std::max(8, std::min(b->ntiles / 8 * 8, b->pwg * b->ncu / 8 * 8))), 1, 1);
    // the E-pass instance without the ADE code when no grid is dispersive
    bool disp = false;
    for (fdtd_grid* g : b->grids) disp = disp || g->max_np > 0;
    // one step's launches: the two passes over every grid's tiles, then the DFT
    auto enqueue_step = [&](int s) {
      if (b->pwg && disp)
        hipLaunchKernelGGL(fdtdk::e_pass_batch_p<T>, pgrid, block, 0, b->stream, dev, b->tiles, b->ntiles, b->tile.kc, s);
      else if (b->pwg)
        hipLaunchKernelGGL((fdtdk::e_pass_batch_p<T, 0>), pgrid, block, 0, b->stream, dev, b->tiles, b->ntiles,
                           b->tile.kc, s);
      else if (disp)
        hipLaunchKernelGGL(fdtdk::e_pass_batch<T>, grid, block, 0, b->stream, dev, b->tiles, b->tile.kc, s);
      else
        hipLaunchKernelGGL((fdtdk::e_pass_batch<T, 0>), grid, block, 0, b->stream, dev, b->tiles, b->tile.kc, s);
      if (b->pwg)
        hipLaunchKernelGGL(fdtdk::h_pass_batch_p<T>, pgrid, block, 0, b->stream, dev, b->tiles, b->ntiles, b->tile.kc, s);
      else
        hipLaunchKernelGGL(fdtdk::h_pass_batch<T>, grid, block, 0, b->stream, dev, b->tiles, b->tile.kc, s);
      if (phase >= 0) {
        hipLaunchKernelGGL(fdtdk::dft_sample_batch<T>, dim3(unsigned(b->smp_total)), dim3(256), 0, b->stream, dev,
                           b->smp_offs, int(n), s);
        if ((phase + s + 1) % fdtdk::kDftBlock == 0)
          hipLaunchKernelGGL(fdtdk::dft_fold_batch<T>, dim3(unsigned(b->fold_total)), dim3(256), 0, b->stream, dev,
                             b->fold_offs, int(n), s);
      } else if (b->dft_total) {
        hipLaunchKernelGGL(fdtdk::dft_accumulate_batch<T>, dim3(unsigned((b->dft_total + 255) / 256)), dim3(256), 0,
                           b->stream, dev, b->dft_offs, int(n), s);
      }
    };
    if (b->use_graph) {
      auto key = std::make_tuple(slot, chunk, phase);
      auto it = b->graphs.find(key);
      if (it == b->graphs.end()) {
        hipGraph_t graph;
        HIP_TRY(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
        for (int s = 0; s < chunk; s++) enqueue_step(s);
        HIP_TRY(hipStreamEndCapture(b->stream, &graph));
        hipGraphExec_t exec;
        hipError_t e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        if (e != hipSuccess) return fail(FDTD_EDEVICE, "hipGraphInstantiate: %s", hipGetErrorString(e));
        it = b->graphs.emplace(key, exec).first;
      }
      HIP_TRY(hipGraphLaunch(it->second, b->stream));
    } else {
      for (int s = 0; s < chunk; s++) enqueue_step(s);
      HIP_TRY(hipGetLastError());
    }
    HIP_TRY(hipEventRecord(b->ev[slot], b->stream));
    b->ev_used[slot] = true;
    for (fdtd_grid* g : b->grids) {
      const long last = ((g->step + chunk) / fdtdk::kDftBlock) * fdtdk::kDftBlock;  // last block end in the chunk
      if (phase < 0)
        g->dft_done = g->step + chunk;  // summed step by step
      else if (last > g->step)
        g->dft_done = last;
      g->step += chunk;
    }
    if (phase >= 0) phase = int((phase + chunk) % fdtdk::kDftBlock);
    done += chunk;
  }
  return 0;
}

}  // namespace

extern "C" {

int fdtd_batch_create(fdtd_grid* const* grids, int n, fdtd_batch** out) {
  if (!grids || n < 1 || !out) return fail(FDTD_EINVAL, "bad argument");
  *out = nullptr;
  for (int q = 0; q < n; q++) {
    if (!grids[q]) return fail(FDTD_EINVAL, "null grid in batch");
    if (grids[q]->ndim != 3) return fail(FDTD_EINVAL, "batches hold 3-D grids");
    if (grids[q]->f64 != grids[0]->f64) return fail(FDTD_EINVAL, "batched grids must share one precision");
    if (grids[q]->nbc) return fail(FDTD_EINVAL, "batched grids cannot have boundary conditions (step them alone)");
    if (grids[q]->nsrc_eh || grids[q]->nplane)
      return fail(FDTD_EINVAL, "batched grids cannot have E / H sources or single TFSF planes (step them alone)");
    if (has_magnetized(grids[q]))
      return fail(FDTD_EINVAL, "batched grids cannot hold magnetized Drude materials (step them alone)");
    for (long c = 0; c < grids[q]->ncell; c++)
      if (grids[q]->mats[grids[q]->mat_host[c]].mnp > 0)
        return fail(FDTD_EINVAL, "batched grids cannot hold magnetic poles (step them alone)");
    for (int r = 0; r < q; r++)
      if (grids[r] == grids[q]) return fail(FDTD_EINVAL, "a grid appears twice in the batch");
  }
  fdtd_batch* b = new fdtd_batch();
  b->f64 = grids[0]->f64;
  b->ncu = grids[0]->ncu;
  int rc = 0;
  if (hipStreamCreate(&b->stream) != hipSuccess) rc = fail(FDTD_EDEVICE, "hipStreamCreate failed");
  for (int q = 0; q < n && !rc; q++) {
    fdtd_grid* g = grids[q];
    if (!rc) rc = dft_flush(g);  // on the grid's own stream, finished before the batch's stream takes over
    (void)hipStreamSynchronize(g->stream);
    b->grids.push_back(g);
    b->saved_streams.push_back(g->stream);
    g->stream = b->stream;
    if (!g->finalized) rc = finalize(g);
  }
  if (!rc) rc = batch_tiles(b, b->tile);
  std::vector<long> offs(1, 0);
  for (int q = 0; q < n; q++) {
    long w = 0;
    for (int m = 0; m < grids[q]->ndft; m++) w += long(grids[q]->dft[m].nfreq) * grids[q]->dft[m].ncells;
    offs.push_back(offs.back() + w);
  }
  b->dft_total = offs.back();
  if (!rc && hipMalloc((void**)&b->dft_offs, offs.size() * sizeof(long)) != hipSuccess) rc = fail(FDTD_ENOMEM, "offsets");
  if (!rc && hipMemcpy(b->dft_offs, offs.data(), offs.size() * sizeof(long), hipMemcpyHostToDevice) != hipSuccess)
    rc = fail(FDTD_EDEVICE, "offset upload");
  // sampler and fold blocks per grid: as the grid's own launches count them
  // (each monitor's cells, and each (monitor, component) segment, padded to
  // whole blocks)
  std::vector<long> so(1, 0), fo(1, 0);
  for (int q = 0; q < n; q++) {
    long sb = 0, fb = 0;
    for (int m = 0; m < grids[q]->ndft; m++) {
      sb += fdtdk::fold_segment(grids[q]->dft[m].ncells) / 256;
      fb += long(grids[q]->dft[m].ncomp) * fdtdk::fold_segment(grids[q]->dft[m].ncells) / 256;
    }
    so.push_back(so.back() + sb);
    fo.push_back(fo.back() + fb);
  }
  b->smp_total = so.back();
  b->fold_total = fo.back();
  if (!rc && hipMalloc((void**)&b->smp_offs, so.size() * sizeof(long)) != hipSuccess) rc = fail(FDTD_ENOMEM, "offsets");
  if (!rc && hipMalloc((void**)&b->fold_offs, fo.size() * sizeof(long)) != hipSuccess) rc = fail(FDTD_ENOMEM, "offsets");
  if (!rc && hipMemcpy(b->smp_offs, so.data(), so.size() * sizeof(long), hipMemcpyHostToDevice) != hipSuccess)
    rc = fail(FDTD_EDEVICE, "offset upload");
  if (!rc && hipMemcpy(b->fold_offs, fo.data(), fo.size() * sizeof(long), hipMemcpyHostToDevice) != hipSuccess)
    rc = fail(FDTD_EDEVICE, "offset upload");
  const size_t abytes = size_t(fdtd_batch::kRing) * n * sizeof(fdtdk::StepArgs<double>);
  if (!rc && hipMalloc((void**)&b->args_dev, abytes) != hipSuccess) rc = fail(FDTD_ENOMEM, "batch arguments");
  if (!rc && hipHostMalloc((void**)&b->args_host, abytes) != hipSuccess) rc = fail(FDTD_ENOMEM, "batch staging");
  for (int r = 0; r < fdtd_batch::kRing && !rc; r++)
    if (hipEventCreateWithFlags(&b->ev[r], hipEventDisableTiming) != hipSuccess) rc = fail(FDTD_EDEVICE, "event");
  if (rc) {
    delete b;
    return rc;
  }
  *out = b;
  return 0;
}

int fdtd_batch_destroy(fdtd_batch* b) {
  delete b;
  return 0;
}

int fdtd_batch_set_tile(fdtd_batch* b, int bx, int by, int kc) {
  if (!b) return fail(FDTD_EINVAL, "null batch");
  if (bx < 8 || by <= 0 || kc <= 0 || bx * by > 256 || (bx * by) % 64 != 0)
    return fail(FDTD_EINVAL, "batch tiles are two-pass tiles: bx >= 8, bx*by a multiple of 64 up to 256, kc >= 1");
  return batch_tiles(b, Tile{bx, by, kc});
}

int fdtd_batch_autotune(fdtd_batch* b, int trial_steps, int* bx, int* by, int* kc) {
  if (!b || trial_steps < 1) return fail(FDTD_EINVAL, "bad argument");
  // 32-wide rows waste fewer lanes on the edges of small grids
  static const Tile cand[] = {{64, 4, 1}, {64, 4, 2}, {64, 2, 4}, {128, 2, 2}, {128, 2, 4}, {128, 1, 4},
                              {256, 1, 2}, {256, 1, 4}, {32, 8, 2}, {32, 8, 4}, {32, 4, 4}, {32, 2, 8}};
  int rc = 0;
  std::vector<GridSnapshot> snaps(b->grids.size());
  for (size_t q = 0; q < b->grids.size() && !rc; q++) {
    fdtd_grid* g = b->grids[q];
    if (!rc) rc = ensure_trace(g, g->step + long(sizeof cand / sizeof cand[0]) * (trial_steps + 1));
    if (!rc) rc = snapshot_take(g, snaps[q]);
  }
  const Tile saved = b->tile;
  const bool graph = b->use_graph;
  b->use_graph = false;  // no graphs for one-off trial tiles
  hipEvent_t e0 = nullptr, e1 = nullptr;
  if (!rc && (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess))
    rc = fail(FDTD_EDEVICE, "hipEventCreate failed");
  double best = 1e300;
  Tile bt = saved;
  const int saved_pwg = b->pwg;
  int bp = 0;
  // one trial: a step to settle, then trial_steps timed
  auto trial = [&](const Tile& t, int pwg) -> int {
    b->pwg = pwg;
    int r = batch_tiles(b, t);
    if (!r) r = fdtd_batch_step(b, 1);
    if (!r && hipEventRecord(e0, b->stream) != hipSuccess) r = fail(FDTD_EDEVICE, "hipEventRecord");
    if (!r) r = fdtd_batch_step(b, trial_steps);
    if (!r && hipEventRecord(e1, b->stream) != hipSuccess) r = fail(FDTD_EDEVICE, "hipEventRecord");
    if (!r && hipEventSynchronize(e1) != hipSuccess) r = fail(FDTD_EDEVICE, "hipEventSynchronize");
    float ms = 0;
    if (!r && hipEventElapsedTime(&ms, e0, e1) != hipSuccess) r = fail(FDTD_EDEVICE, "hipEventElapsedTime");
    if (!r && ms < best) {
      best = ms;
      bt = t;
      bp = pwg;
    }
    return r;
  };
  for (const Tile& t : cand) {
    if (rc) break;
    rc = trial(t, 0);
  }
  // the best tile's persistent forms
  const Tile tile_best = bt;
  for (int pwg : {2, 4, 6, 8}) {
    if (rc) break;
    rc = trial(tile_best, pwg);
  }
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  for (size_t q = 0; q < b->grids.size(); q++) snapshot_restore(b->grids[q], snaps[q]);
  b->use_graph = graph;
  b->pwg = rc ? saved_pwg : bp;
  int rc2 = batch_tiles(b, rc ? saved : bt);
  if (rc) return rc;
  if (rc2) return rc2;
  if (bx) *bx = bt.bx;
  if (by) *by = bt.by;
  if (kc) *kc = bt.kc;
  return 0;
}

int fdtd_batch_set_persistent(fdtd_batch* b, int wg_per_cu) {
  if (!b) return fail(FDTD_EINVAL, "null batch");
  if (wg_per_cu < 0 || wg_per_cu > 32) return fail(FDTD_EINVAL, "workgroups per compute unit must be 0..32");
  b->pwg = wg_per_cu;
  return batch_tiles(b, b->tile);  // drops the graphs captured with the other passes
}

int fdtd_batch_set_graph(fdtd_batch* b, int on) {
  if (!b) return fail(FDTD_EINVAL, "null batch");
  b->use_graph = on != 0;
  return 0;
}

int fdtd_batch_step(fdtd_batch* b, int nsteps) {
  if (!b || nsteps < 0) return fail(FDTD_EINVAL, "bad argument");
  if (nsteps == 0) return 0;
  return b->f64 ? batch_step_impl<double>(b, nsteps) : batch_step_impl<float>(b, nsteps);
}

int fdtd_batch_synchronize(fdtd_batch* b) {
  if (!b) return fail(FDTD_EINVAL, "null batch");
  HIP_TRY(hipStreamSynchronize(b->stream));
  return 0;
}

}  // extern "C"
