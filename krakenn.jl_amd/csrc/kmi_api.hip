// kmi_api.hip — host implementation of the kraken_mi355x C ABI.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "kmi_internal.h"

using namespace kmi;

static constexpr int kGraphSteps = 32;  // steps per captured graph (even: parity returns)

struct kmi_lattice {
  kmi_ctx* ctx = nullptr;
  kmi_desc desc{};
  int Q = 9;
  long long ncells = 0, plane = 0;
  size_t esize = 8;  // bytes per flow population
  size_t gsize = 8;  // bytes per thermal population (thermal_storage of the storage kind)
  void* f[2] = {nullptr, nullptr};
  void* g[2] = {nullptr, nullptr};
  uint8_t* mask = nullptr;
  double* uprev = nullptr;
  double* prof[4][2] = {};  // device per-node values of *_PROFILE faces
  std::vector<uint8_t> solid_h;  // host copy of the solid mask (empty: none)
  std::vector<int> cut_cells;    // cells holding cut links (sorted)
  int* cut_slot = nullptr;       // dense per-cell slot of cut cells (device)
  double* cut_rec = nullptr;     // per cut cell: Q wall fractions, Q × 3 wall velocities
  long long nlinks = 0;          // cut-link list for the drag reductions (device)
  int* link_cell = nullptr;
  int* link_q = nullptr;
  double* link_qw = nullptr;
  double* link_uw = nullptr;
  void* ffield = nullptr;        // per-node body force (KMI_FORCE_GUO_FIELD), Store::real planes
  void* stage = nullptr;         // device staging of host transfers (grown on demand, kept)
  size_t stage_bytes = 0;
  int cur = 0;
  int64_t steps = 0;
  StepArgs args{};
  bool bnd = true;
  int bx = 0, by = 1, nts = 0;  // launch shape and store policy (0: not chosen yet)
  int tb_tile = 0, tb_zchunks = 1;  // two-step launches (0: off), see kmi_step2.hip
  // work-group → tile table of the walled multi-step launch (device), built for order_key
  int* order = nullptr;
  size_t order_cap = 0;
  int order_key[2] = {0, 0};
  dim3 grid, block;
  hipGraphExec_t graph[2] = {nullptr, nullptr};
  hipEvent_t ev[2] = {nullptr, nullptr};
};

static thread_local std::string g_err = "no error";

namespace kmi {
int fail(kmi_ctx* ctx, int code, const std::string& msg) {
  if (ctx) ctx->err = msg;
  else g_err = msg;
  return code;
}

int hip_fail(kmi_ctx* ctx, hipError_t e, const char* what) {
  return fail(ctx, KMI_ERR_DEVICE, std::string(what) + ": " + hipGetErrorString(e));
}
}  // namespace kmi

static const char* kFaceName[6] = {"west", "east", "south", "north", "bottom", "top"};

extern "C" {

int kmi_abi_version(void) { return KMI_ABI_VERSION; }

int kmi_device_count(int* n) {
  if (!n) return fail(nullptr, KMI_ERR_ARG, "kmi_device_count: n is NULL");
  int c = 0;
  if (hipGetDeviceCount(&c) != hipSuccess) c = 0;
  *n = c;
  return KMI_OK;
}

const char* kmi_last_error(const kmi_ctx* ctx) { return ctx ? ctx->err.c_str() : g_err.c_str(); }

int kmi_ctx_create(int device, kmi_ctx** out) {
  if (!out) return fail(nullptr, KMI_ERR_ARG, "kmi_ctx_create: out is NULL");
  *out = nullptr;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n == 0)
    return fail(nullptr, KMI_ERR_DEVICE,
                "kmi_ctx_create: no HIP device is visible; the MI355X backend is unavailable");
  if (device < 0 || device >= n)
    return fail(nullptr, KMI_ERR_ARG, "kmi_ctx_create: device " + std::to_string(device) +
                                          " out of range (" + std::to_string(n) + " visible)");
  hipError_t e = hipSetDevice(device);
  if (e != hipSuccess) return hip_fail(nullptr, e, "hipSetDevice");
  kmi_ctx* c = new (std::nothrow) kmi_ctx();
  if (!c) return fail(nullptr, KMI_ERR_NOMEM, "kmi_ctx_create: out of host memory");
  c->device = device;
  if ((e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking)) != hipSuccess ||
      (e = hipMalloc(&c->partials, 2048 * 3 * 2 * sizeof(double))) != hipSuccess ||
      (e = hipMalloc(&c->out, 8 * sizeof(double))) != hipSuccess ||
      (e = resolve_noalloc_workspace(device)) != hipSuccess) {
    if (c->stream) (void)hipStreamDestroy(c->stream);
    if (c->partials) (void)hipFree(c->partials);
    if (c->out) (void)hipFree(c->out);
    delete c;
    return hip_fail(nullptr, e, "kmi_ctx_create");
  }
  c->err = "no error";
  *out = c;
  return KMI_OK;
}

int kmi_ctx_destroy(kmi_ctx* ctx) {
  if (!ctx) return KMI_OK;
  (void)hipSetDevice(ctx->device);
  (void)hipStreamSynchronize(ctx->stream);
  (void)hipFree(ctx->partials);
  (void)hipFree(ctx->out);
  (void)hipStreamDestroy(ctx->stream);
  delete ctx;
  return KMI_OK;
}

int kmi_ctx_synchronize(kmi_ctx* ctx) {
  if (!ctx) return fail(nullptr, KMI_ERR_ARG, "kmi_ctx_synchronize: ctx is NULL");
  KMI_HIP(ctx, hipSetDevice(ctx->device));
  KMI_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return KMI_OK;
}

void kmi_trt_rates(double nu, double Lambda, double* s_plus, double* s_minus) {
  // fused_trt_2d.jl:129-133
  if (s_plus) *s_plus = 1.0 / (3 * nu + 0.5);
  if (s_minus) *s_minus = 1.0 / (Lambda / (3 * nu) + 0.5);
}

void kmi_desc_init(kmi_desc* d, int32_t stencil, int64_t nx, int64_t ny, int64_t nz, double nu) {
  if (!d) return;
  std::memset(d, 0, sizeof(*d));
  d->stencil = stencil;
  d->storage = KMI_STORE_F64;
  d->collision = KMI_BGK;
  d->nx = nx;
  d->ny = ny;
  d->nz = stencil == KMI_D2Q9 ? 1 : nz;
  d->omega = 1.0 / (3.0 * nu + 0.5);  // omega(config), basic.jl:30
  kmi_trt_rates(nu, 3.0 / 16.0, &d->s_plus, &d->s_minus);
  d->s_e = 1.4;  // collide_mrt_2d! defaults, collide_mrt_2d.jl:118
  d->s_eps = 1.4;
  d->s_q = 1.2;
  d->s_nu = 1.0 / (3.0 * nu + 0.5);
  d->omega_T = d->omega;
  d->coupling = KMI_COUPLING_SPLIT;
  for (int f = 0; f < 6; ++f) {
    d->faces[f].kind = KMI_FACE_WALL;
    d->faces[f].thermal = KMI_THERMAL_ADIABATIC;
    d->faces[f].rho = 1.0;
  }
}

}  // extern "C"

// ---- validation -----------------------------------------------------------------------
static int validate(const kmi_desc& d, std::string& why, int& code) {
  code = KMI_ERR_ARG;
  if (d.stencil != KMI_D2Q9 && d.stencil != KMI_D3Q19) return why = "unknown stencil", 1;
  if (d.storage < KMI_STORE_F64 || d.storage > KMI_STORE_F16S) return why = "unknown storage", 1;
  if (d.collision < KMI_BGK || d.collision > KMI_MRT) return why = "unknown collision", 1;
  const int64_t nz = d.stencil == KMI_D2Q9 ? 1 : d.nz;
  if (d.nx < 1 || d.ny < 1 || nz < 1) return why = "grid sizes must be positive", 1;
  if (d.ny > 65535 || nz > 65535) return why = "ny and nz must be at most 65535", 1;
  if (double(d.nx) * double(d.ny) * double(nz) * (d.stencil == KMI_D2Q9 ? 9 : 19) >= 2147483647.0)
    return code = KMI_ERR_UNSUPPORTED, why = "more than 2^31 populations per field", 1;
  const int nf = d.stencil == KMI_D2Q9 ? 4 : 6;
  for (int f = 0; f < nf; ++f)
    if (d.faces[f].kind < KMI_FACE_PERIODIC || d.faces[f].kind > KMI_FACE_ZOUHE_PRESSURE_REBUILD)
      return why = std::string("unknown face kind on ") + kFaceName[f], 1;
  for (int a = 0; a < nf / 2; ++a) {
    const bool p0 = d.faces[2 * a].kind == KMI_FACE_PERIODIC;
    const bool p1 = d.faces[2 * a + 1].kind == KMI_FACE_PERIODIC;
    if (p0 != p1)
      return why = std::string("periodic face ") + kFaceName[p0 ? 2 * a : 2 * a + 1] +
                   " needs a periodic opposite face",
             1;
  }
  if (d.collision == KMI_MRT && d.stencil != KMI_D2Q9)
    return code = KMI_ERR_UNSUPPORTED, why = "MRT is defined for D2Q9 only", 1;
  if (d.thermal && d.stencil == KMI_D3Q19 && d.coupling != KMI_COUPLING_SPLIT)
    return code = KMI_ERR_UNSUPPORTED,
           why = "the D3Q19 thermal chain couples through the unforced velocity only "
                 "(run_natural_convection_3d, thermal.jl:500-515)",
           1;
  if (d.thermal && d.collision != KMI_BGK)
    return code = KMI_ERR_UNSUPPORTED,
           why = "the thermal DDF uses BGK + Guo for the flow (thermal_2d.jl:421)", 1;
  if (d.solid_mode != KMI_SOLID_BOUNCE_BACK && d.solid_mode != KMI_SOLID_INERT)
    return why = "unknown solid mode", 1;
  if (d.solid_mode == KMI_SOLID_INERT && d.thermal)
    return code = KMI_ERR_UNSUPPORTED, why = "inert solids are for the flow populations only", 1;
  if (d.forcing < KMI_FORCE_NONE || d.forcing > KMI_FORCE_GUO_FIELD)
    return why = "unknown forcing", 1;
  if (d.forcing != KMI_FORCE_NONE && (d.collision != KMI_BGK || d.thermal))
    return code = KMI_ERR_UNSUPPORTED,
           why = "Guo body force is a BGK collision (collide_guo_2d.jl:5-140, "
                 "collide_guo_3d.jl:5-325) and not combined with the thermal buoyancy",
           1;
  for (int f = 0; f < nf; ++f) {
    const kmi_face& F = d.faces[f];
    if (F.kind == KMI_FACE_ZOUHE_VELOCITY_REBUILD || F.kind == KMI_FACE_ZOUHE_PRESSURE_REBUILD) {
      const bool vel = F.kind == KMI_FACE_ZOUHE_VELOCITY_REBUILD;
      if (f != (vel ? KMI_WEST : KMI_EAST))
        return code = KMI_ERR_UNSUPPORTED,
               why = std::string(vel ? "the velocity rebuild is a west face" : "the pressure rebuild is an east face") +
                     " (apply_bc_rebuild_2d!/_3d!: west inlet, east outlet)",
               1;
      if (d.collision != KMI_TRT || d.thermal)
        return code = KMI_ERR_UNSUPPORTED,
               why = "rebuilt faces collide with TRT rates (src/bc/specs.jl:7-108): TRT, isothermal lattices only", 1;
      if (!vel && !(F.rho > 0.0)) return why = "rebuild pressure density must be positive", 1;
    }
    if (F.kind == KMI_FACE_ZOUHE_VELOCITY_PROFILE || F.kind == KMI_FACE_ZOUHE_PRESSURE_PROFILE) {
      if (d.stencil != KMI_D2Q9)
        return code = KMI_ERR_UNSUPPORTED, why = "profile faces are D2Q9 only (boundary_spatial_2d.jl)", 1;
      if (F.kind == KMI_FACE_ZOUHE_VELOCITY_PROFILE && f == KMI_EAST)
        return code = KMI_ERR_UNSUPPORTED,
               why = "Zou-He velocity profile on the east face is not defined by the reference "
                     "(boundary_spatial_2d.jl: north, south, west)",
               1;
      if (F.kind == KMI_FACE_ZOUHE_PRESSURE_PROFILE && f != KMI_EAST)
        return code = KMI_ERR_UNSUPPORTED,
               why = std::string("Pressure profile on face ") + kFaceName[f] +
                     " is not supported (boundary_spatial_2d.jl: east)",
               1;
    }
    if ((F.kind == KMI_FACE_ZOUHE_PRESSURE || F.kind == KMI_FACE_ZOUHE_PRESSURE_PROFILE) &&
        !(F.rho > 0.0))
      return why = "Zou-He pressure density must be positive", 1;
    if (F.kind == KMI_FACE_ZOUHE_VELOCITY && d.stencil == KMI_D3Q19) {
      // ρ_w = (Σ_par + 2 Σ_out) / (1 + sign u_n) (boundary_3d.jl:186-187)
      const double un = F.u[f / 2], sign = (f % 2) ? 1.0 : -1.0;
      if (!(1.0 + sign * un > 0.0))
        return why = std::string("Zou-He normal velocity on ") + kFaceName[f] +
                     " makes 1 + sign*u_n non-positive",
               1;
    }
    if ((F.kind == KMI_FACE_ZOUHE_VELOCITY || F.kind == KMI_FACE_ZOUHE_PRESSURE) &&
        d.stencil == KMI_D2Q9) {
      if (F.kind == KMI_FACE_ZOUHE_VELOCITY) {
        if (f == KMI_EAST)
          return code = KMI_ERR_UNSUPPORTED,
                 why = "Zou-He velocity on the east face is not defined by the reference "
                       "(handlers.jl:228-237 support north, south, west)",
                 1;
        if (F.u[1] != 0.0 || F.u[2] != 0.0)
          return code = KMI_ERR_UNSUPPORTED,
                 why = std::string("Zou-He velocity on ") + kFaceName[f] +
                       " takes an x velocity only (boundary_2d.jl:13-185)",
                 1;
      } else {
        if (f != KMI_EAST && f != KMI_WEST)
          return code = KMI_ERR_UNSUPPORTED,
                 why = std::string("Pressure boundary condition on face ") + kFaceName[f] +
                       " is not supported (supported: west, east; handlers.jl:256-269)",
                 1;
      }
    }
    if (d.thermal && F.thermal == KMI_THERMAL_DIRICHLET && F.kind == KMI_FACE_PERIODIC)
      return why = std::string("Dirichlet temperature on periodic face ") + kFaceName[f], 1;
  }
  if (d.n_bc_order < 0 || d.n_bc_order > 6) return why = "n_bc_order must be in 0..6", 1;
  for (int n = 0; n < d.n_bc_order; ++n)
    if (d.bc_order[n] < 0 || d.bc_order[n] >= nf) return why = "bc_order names an unknown face", 1;
  return 0;
}

// Nodes of a face: nx / ny along a D2Q9 face; the product of the two tangent extents in 3D.
static int64_t face_length(const kmi_lattice* L, int f) {
  const kmi_desc& d = L->desc;
  if (L->Q == 9) return (f == KMI_SOUTH || f == KMI_NORTH) ? d.nx : d.ny;
  const int64_t n[3] = {d.nx, d.ny, d.nz};
  const int ax = f / 2;
  return n[ax == 0 ? 1 : 0] * n[ax == 2 ? 1 : 2];
}

static void build_args(kmi_lattice* L) {
  const kmi_desc& d = L->desc;
  StepArgs& a = L->args;
  std::memset(&a, 0, sizeof(a));
  a.nx = int(d.nx);
  a.ny = int(d.ny);
  a.nz = int(d.stencil == KMI_D2Q9 ? 1 : d.nz);
  a.plane = L->plane;
  a.per[0] = d.faces[KMI_WEST].kind == KMI_FACE_PERIODIC;
  a.per[1] = d.faces[KMI_SOUTH].kind == KMI_FACE_PERIODIC;
  a.per[2] = L->Q == 19 && d.faces[KMI_BOTTOM].kind == KMI_FACE_PERIODIC;
  const int nf = L->Q == 9 ? 4 : 6;
  a.any_moving = 0;
  for (int f = 0; f < nf; ++f) a.any_moving |= d.faces[f].kind == KMI_FACE_MOVING_WALL;
  int order[6] = {0, 1, 2, 3, 4, 5};
  int n = nf;
  if (d.n_bc_order > 0) {
    n = d.n_bc_order;
    for (int i = 0; i < n; ++i) order[i] = d.bc_order[i];
  }
  a.n_zh = 0;
  for (int i = 0; i < n; ++i) {
    const int k = d.faces[order[i]].kind;
    if (k >= KMI_FACE_ZOUHE_VELOCITY && k <= KMI_FACE_ZOUHE_PRESSURE_PROFILE) a.zh_face[a.n_zh++] = order[i];
  }
  a.omega = d.omega;
  a.s_plus = d.s_plus;
  a.s_minus = d.s_minus;
  a.s_e = d.s_e;
  a.s_eps = d.s_eps;
  a.s_q = d.s_q;
  a.s_nu = d.s_nu;
  a.omega_T = d.omega_T;
  a.beta_g = d.beta_g;
  a.T_ref = d.T_ref;
  a.coupling = d.coupling;
  a.forcing = d.forcing != KMI_FORCE_NONE;
  a.solid_inert = d.solid_mode == KMI_SOLID_INERT;
  a.cut_slot = L->cut_slot;
  a.cut_rec = L->cut_rec;
  a.fx = d.force[0];
  a.fy = d.force[1];
  a.fz = L->Q == 19 ? d.force[2] : 0.0;
  a.ffield = d.forcing == KMI_FORCE_GUO_FIELD ? L->ffield : nullptr;
  bool dir[6];
  for (int f = 0; f < 6; ++f)
    dir[f] = d.thermal && f < nf && d.faces[f].kind != KMI_FACE_PERIODIC &&
             d.faces[f].thermal == KMI_THERMAL_DIRICHLET;
  // Corner / edge link temperature, in the storage arithmetic (_corner_temp,
  // thermal_2d.jl:356-357).
  auto corner = [&](int fa, int fb) -> double {
    if (!dir[fb]) return d.faces[fa].T;
    if (d.storage == KMI_STORE_F32) return double((float(d.faces[fa].T) + float(d.faces[fb].T)) / 2.0f);
    return (d.faces[fa].T + d.faces[fb].T) / 2.0;
  };
  for (int f = 0; f < 6; ++f) {
    FaceArgs& F = a.face[f];
    F.kind = d.faces[f].kind;
    F.dirichlet = dir[f];
    for (int c = 0; c < 3; ++c) F.u[c] = d.faces[f].u[c];
    F.rho = d.faces[f].rho;
    F.T = d.faces[f].T;
    for (int e = 0; e < 4; ++e) F.Tc[e] = d.faces[f].T, F.skip[e] = 0;
    F.prof[0] = f < 4 ? L->prof[f][0] : nullptr;
    F.prof[1] = f < 4 ? L->prof[f][1] : nullptr;
    if (!dir[f]) continue;
    // Edge slots in the order of the face's tangents (y then z for x-faces, x then z for
    // y-faces, x then y for z-faces; low index first): x-faces own every edge, y-faces
    // leave the edges with Dirichlet x-faces alone, z-faces every edge with a Dirichlet x-
    // or y-face (_apply_fixed_temp_faces_3d!, thermal_3d.jl:494-535; 2D: the first two).
    if (f == KMI_WEST || f == KMI_EAST) {
      F.Tc[0] = corner(f, KMI_SOUTH);
      F.Tc[1] = corner(f, KMI_NORTH);
      F.Tc[2] = corner(f, KMI_BOTTOM);
      F.Tc[3] = corner(f, KMI_TOP);
    } else if (f == KMI_SOUTH || f == KMI_NORTH) {
      F.skip[0] = dir[KMI_WEST];
      F.skip[1] = dir[KMI_EAST];
      F.Tc[2] = corner(f, KMI_BOTTOM);
      F.Tc[3] = corner(f, KMI_TOP);
    } else {
      F.skip[0] = dir[KMI_WEST];
      F.skip[1] = dir[KMI_EAST];
      F.skip[2] = dir[KMI_SOUTH];
      F.skip[3] = dir[KMI_NORTH];
    }
  }
  bool any_nonperiodic = false;
  for (int f = 0; f < nf; ++f) any_nonperiodic |= d.faces[f].kind != KMI_FACE_PERIODIC;
  L->bnd = any_nonperiodic || L->mask != nullptr || d.forcing != KMI_FORCE_NONE;
}

static void set_geometry(kmi_lattice* L, int bx, int by, int nts) {
  L->bx = bx;
  L->by = by;
  L->nts = nts;
  L->block = dim3(unsigned(bx), unsigned(by), 1);
  L->grid = dim3(unsigned((L->desc.nx + bx - 1) / bx), unsigned((L->desc.ny + by - 1) / by),
                 unsigned(L->args.nz));
}

// Default launch: one x-row segment of up to 256 cells per work-group (64-wide groups on
// small grids so the launch still spreads over the 256 CUs); non-temporal stores once a
// population field outgrows the 256 MB last-level cache.
static void default_geometry(kmi_lattice* L) {
  const kmi_desc& d = L->desc;
  int bx = int(std::min<int64_t>(256, ((d.nx + 63) / 64) * 64));
  const long long groups = ((d.nx + bx - 1) / bx) * d.ny * L->args.nz;
  if (groups < 2048 && bx > 64) bx = 64;
  const double field_bytes =
      double(L->plane) * (L->Q * L->esize + (L->desc.thermal ? L->Q * L->gsize : 0));
  set_geometry(L, bx, 1, field_bytes > 512e6 ? 1 : 0);
}

// Whether multi-step launches with `tile` apply.  D3Q19: BGK/TRT without thermal populations,
// the x–y extent a multiple of the tile and at least 4 planes in z; fully periodic lattices
// without mask on any tile, walled ones (walls, moving walls, constant Zou–He faces, solid
// mask; no forcing, no cut links) on the walled tiles (12, 14; 9 except FP64; not F16S).  D2Q9 (k_step_tb2d):
// no cut links, faces periodic / wall / moving wall / constant Zou–He only (thermal lattices
// included), and a lattice of at least the tile plus its ring on each axis.
static bool tb_eligible(const kmi_lattice* L, int tile) {
  if (L->Q == 9) {
    int TW, TH, K;
    if (!tb2d_variant(tile, TW, TH, K) || !L->cut_cells.empty()) return false;
    if (L->desc.thermal && L->desc.storage == KMI_STORE_F64) return false;  // see k_step_tb2d
    for (int f = 0; f < 4; ++f)
      if (L->desc.faces[f].kind > KMI_FACE_ZOUHE_PRESSURE) return false;
    return L->desc.nx >= TW + 2 * K && L->desc.ny >= TH + 2 * K;
  }
  int TX, TY;
  if (!step2_tile(tile, L->desc.storage, TX, TY)) return false;
  if (L->Q != 19 || L->desc.thermal || (L->desc.collision != KMI_BGK && L->desc.collision != KMI_TRT) ||
      L->desc.nx % TX != 0 || L->desc.ny % TY != 0 || L->args.nz < 4)
    return false;
  if (!L->bnd) return true;
  // walled lattices: halfway / moving-wall bounce-back, constant Zou–He faces and solid masks,
  // without forcing or cut links, on the walled tiles
  if (!step2_tile_walled(tile, L->desc.storage) || L->desc.forcing != KMI_FORCE_NONE || !L->cut_cells.empty())
    return false;
  for (int f = 0; f < 6; ++f)
    if (L->desc.faces[f].kind > KMI_FACE_ZOUHE_PRESSURE) return false;
  return true;
}

// Steps advanced by one launch.
static int launch_span(const kmi_lattice* L) {
  if (!L->tb_tile || !tb_eligible(L, L->tb_tile)) return 1;
  if (L->Q == 9) {
    int TW, TH, K;
    tb2d_variant(L->tb_tile, TW, TH, K);
    return K;
  }
  return temporal_span(L->tb_tile);
}

// Steps per captured graph: a multiple of the span with an even number of launches, so the
// buffer parity returns to where the capture started.
static int graph_steps(const kmi_lattice* L) { return launch_span(L) == 3 ? 30 : kGraphSteps; }

// One launch: a single step, or the temporal-blocking launch (multi = true: two or three
// steps).  Either way the result lands in the other buffer.
static int launch_one(kmi_lattice* L, hipStream_t s, bool multi = false) {
  StepArgs a = L->args;
  a.fi = L->f[L->cur];
  a.fo = L->f[L->cur ^ 1];
  a.gi = L->g[L->cur];
  a.go = L->g[L->cur ^ 1];
  a.mask = L->mask;
  // the table only for the tile and z-split it was built for (any other launch keeps the
  // default order, which covers every tile as well)
  a.tile_order = multi && L->bnd && L->Q == 19 && L->order && L->order_key[0] == L->tb_tile &&
                         L->order_key[1] == L->tb_zchunks
                     ? L->order
                     : nullptr;
  hipError_t e =
      multi ? (L->Q == 9 ? launch_step_tb2d(L->desc.storage, L->desc.collision, L->tb_tile,
                                            L->desc.thermal != 0, a, s)
                         : launch_step2(L->desc.storage, L->desc.collision, L->tb_tile, L->tb_zchunks,
                                        L->nts != 0, L->bnd, a, s))
            : launch_step(L->Q, L->desc.storage, L->desc.collision, L->desc.thermal != 0, L->bnd,
                          L->nts != 0, a, L->grid, L->block, s);
  if (e != hipSuccess) return hip_fail(L->ctx, e, "step kernel launch");
  L->cur ^= 1;
  return KMI_OK;
}

static void drop_graphs(kmi_lattice* L) {
  for (int p = 0; p < 2; ++p)
    if (L->graph[p]) {
      (void)hipGraphExecDestroy(L->graph[p]);
      L->graph[p] = nullptr;
    }
}

// graph_steps(L) steps captured once per buffer parity and replayed.
// The walled D3Q19 multi-step launch's work-group order (walled_tile_order) for the current
// tile and z-split, rebuilt when those change — before any graph capture, since the upload
// is an ordinary copy on the lattice stream.
static int refresh_tile_order(kmi_lattice* L) {
  if (!(L->bnd && L->Q == 19 && L->tb_tile)) return KMI_OK;
  if (L->order && L->order_key[0] == L->tb_tile && L->order_key[1] == L->tb_zchunks) return KMI_OK;
  std::vector<int> h;
  walled_tile_order(L->args, L->tb_tile, L->tb_zchunks, h);
  if (h.empty()) return fail(L->ctx, KMI_ERR_ARG, "walled multi-step launch: unknown tile");
  const size_t bytes = h.size() * sizeof(int);
  if (bytes > L->order_cap) {
    if (L->order) {
      KMI_HIP(L->ctx, hipStreamSynchronize(L->ctx->stream));
      (void)hipFree(L->order);
      L->order = nullptr;
      L->order_cap = 0;
    }
    KMI_HIP(L->ctx, hipMalloc(&L->order, bytes));
    L->order_cap = bytes;
  }
  KMI_HIP(L->ctx, hipMemcpyAsync(L->order, h.data(), bytes, hipMemcpyHostToDevice, L->ctx->stream));
  KMI_HIP(L->ctx, hipStreamSynchronize(L->ctx->stream));  // h leaves scope
  L->order_key[0] = L->tb_tile;
  L->order_key[1] = L->tb_zchunks;
  return KMI_OK;
}

static int run_steps(kmi_lattice* L, int64_t n) {
  kmi_ctx* c = L->ctx;
  hipStream_t s = c->stream;
  if (int rc = refresh_tile_order(L); rc != KMI_OK) return rc;
  const int span = launch_span(L);
  const int gsteps = graph_steps(L);
  while (n >= gsteps) {
    const int p = L->cur;
    if (!L->graph[p]) {
      hipGraph_t gph = nullptr;
      KMI_HIP(c, hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
      for (int k = 0; k < gsteps; k += span) {
        int rc = launch_one(L, s, span > 1);
        if (rc != KMI_OK) {
          (void)hipStreamEndCapture(s, &gph);
          if (gph) (void)hipGraphDestroy(gph);
          return rc;
        }
      }
      KMI_HIP(c, hipStreamEndCapture(s, &gph));
      hipError_t e = hipGraphInstantiate(&L->graph[p], gph, nullptr, nullptr, 0);
      (void)hipGraphDestroy(gph);
      if (e != hipSuccess) return hip_fail(c, e, "hipGraphInstantiate");
      // capture only recorded the launches; parity is back where it started
    }
    KMI_HIP(c, hipGraphLaunch(L->graph[p], s));
    n -= gsteps;
    L->steps += gsteps;
  }
  for (; span > 1 && n >= span; n -= span) {
    int rc = launch_one(L, s, true);
    if (rc != KMI_OK) return rc;
    L->steps += span;
  }
  for (; n > 0; --n) {
    int rc = launch_one(L, s);
    if (rc != KMI_OK) return rc;
    L->steps += 1;
  }
  return KMI_OK;
}

static double storage_shift_f() { return 1.0; }

// Bytes of the per-node force field: 2 (D2Q9) or 3 (D3Q19) planes in the collision arithmetic.
static size_t ffield_bytes(const kmi_lattice* L) {
  return size_t(L->Q == 19 ? 3 : 2) * size_t(L->plane) *
         (L->desc.storage == KMI_STORE_F64 ? sizeof(double) : sizeof(float));
}


extern "C" {

int kmi_desc_validate(const kmi_desc* desc) {
  if (!desc) return fail(nullptr, KMI_ERR_ARG, "kmi_desc_validate: desc is NULL");
  std::string why;
  int code;
  if (validate(*desc, why, code)) return fail(nullptr, code, "kmi_desc_validate: " + why);
  return KMI_OK;
}

size_t kmi_desc_sizeof(void) { return sizeof(kmi_desc); }

int kmi_lattice_create(kmi_ctx* ctx, const kmi_desc* desc, kmi_lattice** out) {
  if (!ctx) return fail(nullptr, KMI_ERR_ARG, "kmi_lattice_create: ctx is NULL");
  if (!desc || !out) return fail(ctx, KMI_ERR_ARG, "kmi_lattice_create: NULL argument");
  *out = nullptr;
  std::string why;
  int code;
  if (validate(*desc, why, code)) return fail(ctx, code, "kmi_lattice_create: " + why);
  KMI_HIP(ctx, hipSetDevice(ctx->device));
  kmi_lattice* L = new (std::nothrow) kmi_lattice();
  if (!L) return fail(ctx, KMI_ERR_NOMEM, "kmi_lattice_create: out of host memory");
  L->ctx = ctx;
  L->desc = *desc;
  if (L->desc.stencil == KMI_D2Q9) L->desc.nz = 1;
  L->Q = desc->stencil == KMI_D2Q9 ? 9 : 19;
  L->ncells = (long long)desc->nx * desc->ny * L->desc.nz;
  L->plane = L->ncells;
  L->esize = storage_bytes(desc->storage);
  L->gsize = storage_bytes(thermal_storage(desc->storage));
  const size_t bytes = size_t(L->plane) * L->Q * L->esize;
  const size_t gbytes = size_t(L->plane) * L->Q * L->gsize;
  hipError_t e = hipSuccess;
  for (int p = 0; p < 2 && e == hipSuccess; ++p) e = hipMalloc(&L->f[p], bytes);
  for (int p = 0; p < 2 && e == hipSuccess && desc->thermal; ++p) e = hipMalloc(&L->g[p], gbytes);
  // Profile faces: velocity profiles start at zero, pressure profiles at the face's rho.
  for (int f = 0; f < 4 && e == hipSuccess; ++f) {
    const int kind = desc->faces[f].kind;
    if (kind != KMI_FACE_ZOUHE_VELOCITY_PROFILE && kind != KMI_FACE_ZOUHE_PRESSURE_PROFILE &&
        kind != KMI_FACE_ZOUHE_VELOCITY_REBUILD)
      continue;
    const size_t len = size_t(face_length(L, f));
    const std::vector<double> init(len, kind == KMI_FACE_ZOUHE_PRESSURE_PROFILE ? desc->faces[f].rho : 0.0);
    for (int k = 0; k < (kind == KMI_FACE_ZOUHE_VELOCITY_PROFILE ? 2 : 1) && e == hipSuccess; ++k) {
      e = hipMalloc(&L->prof[f][k], len * sizeof(double));
      if (e == hipSuccess)
        e = hipMemcpy(L->prof[f][k], init.data(), len * sizeof(double), hipMemcpyHostToDevice);
    }
  }
  // Per-node force field, zero until kmi_set_force_field.
  if (e == hipSuccess && desc->forcing == KMI_FORCE_GUO_FIELD) {
    e = hipMalloc(&L->ffield, ffield_bytes(L));
    if (e == hipSuccess) e = hipMemset(L->ffield, 0, ffield_bytes(L));
  }
  if (e == hipSuccess) e = hipEventCreate(&L->ev[0]);
  if (e == hipSuccess) e = hipEventCreate(&L->ev[1]);
  if (e != hipSuccess) {
    kmi_lattice_destroy(L);
    return fail(ctx, e == hipErrorOutOfMemory ? KMI_ERR_NOMEM : KMI_ERR_DEVICE,
                std::string("kmi_lattice_create: ") + hipGetErrorString(e));
  }
  build_args(L);
  default_geometry(L);
  // rest equilibrium f = w_q, g = w_q T_ref (initialize_2d/3d, basic.jl:64-70)
  int rc = kmi_init_equilibrium(L, nullptr, nullptr, nullptr, nullptr);
  if (rc == KMI_OK && desc->thermal) rc = kmi_init_temperature(L, nullptr);
  if (rc != KMI_OK) {
    kmi_lattice_destroy(L);
    return rc;
  }
  *out = L;
  return KMI_OK;
}

int kmi_lattice_destroy(kmi_lattice* L) {
  if (!L) return KMI_OK;
  (void)hipSetDevice(L->ctx->device);
  (void)hipStreamSynchronize(L->ctx->stream);
  drop_graphs(L);
  for (int p = 0; p < 2; ++p) {
    if (L->f[p]) (void)hipFree(L->f[p]);
    if (L->g[p]) (void)hipFree(L->g[p]);
    if (L->ev[p]) (void)hipEventDestroy(L->ev[p]);
  }
  if (L->mask) (void)hipFree(L->mask);
  if (L->uprev) (void)hipFree(L->uprev);
  for (auto& pf : L->prof)
    for (double* p : pf)
      if (p) (void)hipFree(p);
  for (void* p : {(void*)L->cut_slot, (void*)L->cut_rec, (void*)L->link_cell, (void*)L->link_q,
                  (void*)L->link_qw, (void*)L->link_uw})
    if (p) (void)hipFree(p);
  if (L->ffield) (void)hipFree(L->ffield);
  if (L->stage) (void)hipFree(L->stage);
  if (L->order) (void)hipFree(L->order);
  delete L;
  return KMI_OK;
}

// Device mask = solid flags combined with the cut-cell class; none when both are empty.
static hipError_t upload_mask(kmi_lattice* L) {
  if (L->solid_h.empty() && L->cut_cells.empty()) {
    if (L->mask) (void)hipFree(L->mask);
    L->mask = nullptr;
    return hipSuccess;
  }
  std::vector<uint8_t> m(size_t(L->ncells), kCellFluid);
  if (!L->solid_h.empty())
    for (size_t c = 0; c < m.size(); ++c) m[c] = L->solid_h[c] ? kCellSolid : kCellFluid;
  for (int c : L->cut_cells) m[size_t(c)] = kCellCut;
  hipError_t e = hipSuccess;
  if (!L->mask) e = hipMalloc(&L->mask, m.size());
  if (e == hipSuccess) e = hipMemcpy(L->mask, m.data(), m.size(), hipMemcpyHostToDevice);
  return e;
}

int kmi_set_mask(kmi_lattice* L, const uint8_t* is_solid) {
  if (!L) return fail(nullptr, KMI_ERR_ARG, "kmi_set_mask: lattice is NULL");
  kmi_ctx* c = L->ctx;
  if (is_solid)
    for (int cell : L->cut_cells)
      if (is_solid[cell])
        return fail(c, KMI_ERR_ARG, "kmi_set_mask: a cell holding cut links cannot be solid");
  KMI_HIP(c, hipSetDevice(c->device));
  KMI_HIP(c, hipStreamSynchronize(c->stream));
  drop_graphs(L);
  if (!is_solid) L->solid_h.clear();
  else L->solid_h.assign(is_solid, is_solid + L->ncells);
  KMI_HIP(c, upload_mask(L));
  build_args(L);
  return KMI_OK;
}

int kmi_set_force_field(kmi_lattice* L, const double* fx, const double* fy, const double* fz) {
  if (!L) return fail(nullptr, KMI_ERR_ARG, "kmi_set_force_field: lattice is NULL");
  kmi_ctx* c = L->ctx;
  if (L->desc.forcing != KMI_FORCE_GUO_FIELD)
    return fail(c, KMI_ERR_STATE, "kmi_set_force_field: the lattice's forcing is not KMI_FORCE_GUO_FIELD");
  if (!fx || !fy || (L->Q == 19 && !fz))
    return fail(c, KMI_ERR_ARG, "kmi_set_force_field: fx and fy (and fz for D3Q19) are required");
  const double* src[3] = {fx, fy, fz};
  const int nc = L->Q == 19 ? 3 : 2;
  const bool dbl = L->desc.storage == KMI_STORE_F64;
  std::vector<unsigned char> h(ffield_bytes(L));
  for (int k = 0; k < nc; ++k)
    for (long long i = 0; i < L->ncells; ++i) {
      const size_t at = size_t(k) * size_t(L->plane) + size_t(i);
      if (dbl) {
        std::memcpy(h.data() + at * sizeof(double), &src[k][i], sizeof(double));
      } else {
        const float v = float(src[k][i]);  // rounded once, as a Float32 field array
        std::memcpy(h.data() + at * sizeof(float), &v, sizeof(float));
      }
    }
  KMI_HIP(c, hipSetDevice(c->device));
  KMI_HIP(c, hipStreamSynchronize(c->stream));
  KMI_HIP(c, hipMemcpy(L->ffield, h.data(), h.size(), hipMemcpyHostToDevice));
  return KMI_OK;
}

int kmi_set_cut_links(kmi_lattice* L, int64_t n, const int32_t* cell, const int32_t* q,
                      const double* qw, const double* uw) {
  if (!L) return fail(nullptr, KMI_ERR_ARG, "kmi_set_cut_links: lattice is NULL");
  kmi_ctx* c = L->ctx;
  const int Q = L->Q;
  if (n < 0 || (n > 0 && (!cell || !q || !qw || !uw)))
    return fail(c, KMI_ERR_ARG, "kmi_set_cut_links: bad count or NULL array");
  if (L->desc.thermal)
    return fail(c, KMI_ERR_UNSUPPORTED, "kmi_set_cut_links: cut links act on the flow populations of "
                                        "non-thermal lattices");
  if (L->desc.collision == KMI_MRT)
    return fail(c, KMI_ERR_UNSUPPORTED, "kmi_set_cut_links: the LI-BB step collides with BGK or TRT");
  for (int64_t k = 0; k < n; ++k) {
    if (cell[k] < 0 || cell[k] >= L->ncells || q[k] < 1 || q[k] >= Q)
      return fail(c, KMI_ERR_ARG, "kmi_set_cut_links: link " + std::to_string(k) +
                                      " has a cell or direction out of range");
    if (!(qw[k] > 0.0 && qw[k] <= 1.0))
      return fail(c, KMI_ERR_ARG, "kmi_set_cut_links: wall fraction must be in (0, 1]");
    if (!L->solid_h.empty() && L->solid_h[size_t(cell[k])])
      return fail(c, KMI_ERR_ARG, "kmi_set_cut_links: link " + std::to_string(k) +
                                      " leaves a solid cell");
  }
  // per-cell records (duplicates: the last entry of a (cell, q) pair wins)
  std::vector<int> cells(cell, cell + n);
  std::sort(cells.begin(), cells.end());
  cells.erase(std::unique(cells.begin(), cells.end()), cells.end());
  std::vector<double> rec(cells.size() * 4 * Q, 0.0);
  for (int64_t k = 0; k < n; ++k) {
    const size_t s = size_t(std::lower_bound(cells.begin(), cells.end(), cell[k]) - cells.begin());
    double* r = rec.data() + s * 4 * Q;
    r[q[k]] = qw[k];
    for (int a = 0; a < 3; ++a) r[Q + 3 * q[k] + a] = uw[3 * k + a];
  }
  KMI_HIP(c, hipSetDevice(c->device));
  KMI_HIP(c, hipStreamSynchronize(c->stream));
  drop_graphs(L);
  for (void** p : {(void**)&L->cut_rec, (void**)&L->link_cell, (void**)&L->link_q,
                   (void**)&L->link_qw, (void**)&L->link_uw}) {
    if (*p) (void)hipFree(*p);
    *p = nullptr;
  }
  L->cut_cells = cells;
  L->nlinks = n;
  hipError_t e = hipSuccess;
  if (!cells.empty()) {
    std::vector<int> slot_h(size_t(L->ncells), -1);  // dense slots, read only for cut cells
    for (size_t s = 0; s < cells.size(); ++s) slot_h[size_t(cells[s])] = int(s);
    if (!L->cut_slot) e = hipMalloc(&L->cut_slot, size_t(L->ncells) * sizeof(int));
    if (e == hipSuccess)
      e = hipMemcpy(L->cut_slot, slot_h.data(), slot_h.size() * sizeof(int), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMalloc(&L->cut_rec, rec.size() * sizeof(double));
    if (e == hipSuccess) e = hipMemcpy(L->cut_rec, rec.data(), rec.size() * sizeof(double), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMalloc(&L->link_cell, size_t(n) * sizeof(int));
    if (e == hipSuccess) e = hipMalloc(&L->link_q, size_t(n) * sizeof(int));
    if (e == hipSuccess) e = hipMalloc(&L->link_qw, size_t(n) * sizeof(double));
    if (e == hipSuccess) e = hipMalloc(&L->link_uw, size_t(n) * 3 * sizeof(double));
    if (e == hipSuccess) e = hipMemcpy(L->link_cell, cell, size_t(n) * sizeof(int), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(L->link_q, q, size_t(n) * sizeof(int), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(L->link_qw, qw, size_t(n) * sizeof(double), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(L->link_uw, uw, size_t(n) * 3 * sizeof(double), hipMemcpyHostToDevice);
  }
  if (e == hipSuccess) e = upload_mask(L);
  build_args(L);
  if (e != hipSuccess) return hip_fail(c, e, "kmi_set_cut_links");
  return KMI_OK;
}

int kmi_set_face_profile(kmi_lattice* L, int32_t face, const double* a, const double* b) {
  if (!L) return fail(nullptr, KMI_ERR_ARG, "kmi_set_face_profile: lattice is NULL");
  kmi_ctx* c = L->ctx;
  if (face < 0 || face >= (L->Q == 9 ? 4 : 6))
    return fail(c, KMI_ERR_ARG, "kmi_set_face_profile: unknown face");
  const int kind = L->desc.faces[face].kind;
  if (kind != KMI_FACE_ZOUHE_VELOCITY_PROFILE && kind != KMI_FACE_ZOUHE_PRESSURE_PROFILE &&
      kind != KMI_FACE_ZOUHE_VELOCITY_REBUILD)
    return fail(c, KMI_ERR_STATE, std::string("kmi_set_face_profile: face ") + kFaceName[face] +
                                      " is not a profile face");
  if (!a || (kind == KMI_FACE_ZOUHE_VELOCITY_PROFILE && !b))
    return fail(c, KMI_ERR_ARG, "kmi_set_face_profile: NULL profile array");
  const size_t len = size_t(face_length(L, face));
  if (kind == KMI_FACE_ZOUHE_PRESSURE_PROFILE)
    for (size_t n = 1; n + 1 < len; ++n)
      if (!(a[n] > 0.0)) return fail(c, KMI_ERR_ARG, "kmi_set_face_profile: density must be positive");
  KMI_HIP(c, hipSetDevice(c->device));
  const double* src[2] = {a, b};
  for (int k = 0; k < (kind == KMI_FACE_ZOUHE_VELOCITY_PROFILE ? 2 : 1); ++k)
    KMI_HIP(c, hipMemcpyAsync(L->prof[face][k], src[k], len * sizeof(double),
                              hipMemcpyHostToDevice, c->stream));
  KMI_HIP(c, hipStreamSynchronize(c->stream));
  return KMI_OK;
}

int kmi_set_face(kmi_lattice* L, int32_t face, const kmi_face* v) {
  if (!L) return fail(nullptr, KMI_ERR_ARG, "kmi_set_face: lattice is NULL");
  kmi_ctx* c = L->ctx;
  const int nf = L->Q == 9 ? 4 : 6;
  if (!v || face < 0 || face >= nf) return fail(c, KMI_ERR_ARG, "kmi_set_face: bad face or NULL values");
  const kmi_face& old = L->desc.faces[face];
  if (v->kind != old.kind || v->thermal != old.thermal)
    return fail(c, KMI_ERR_ARG, "kmi_set_face: the face kind and thermal kind cannot change");
  kmi_desc d = L->desc;
  d.faces[face] = *v;
  std::string why;
  int code;
  if (validate(d, why, code)) return fail(c, code, "kmi_set_face: " + why);
  KMI_HIP(c, hipSetDevice(c->device));
  KMI_HIP(c, hipStreamSynchronize(c->stream));
  drop_graphs(L);
  L->desc = d;
  build_args(L);
  return KMI_OK;
}

static int field_ptrs(kmi_lattice* L, int field, void**& buf, double& shift, int& st) {
  if (field == KMI_FIELD_F) {
    buf = L->f;
    shift = storage_shift_f();
    st = L->desc.storage;
    return KMI_OK;
  }
  if (field == KMI_FIELD_G && L->desc.thermal) {
    buf = L->g;
    shift = L->desc.T_ref;
    st = thermal_storage(L->desc.storage);
    return KMI_OK;
  }
  return fail(L->ctx, KMI_ERR_ARG, "unknown field or lattice without thermal populations");
}

// Device staging for host transfers, owned by the lattice: grown when a transfer needs more,
// then reused, so repeated uploads, downloads, moments and initialisations (a run_simulation
// callback every few steps) allocate nothing.  The stream is idle when it is replaced.
static hipError_t stage_reserve(kmi_lattice* L, size_t bytes, double** p) {
  if (L->stage_bytes < bytes) {
    hipError_t e = hipStreamSynchronize(L->ctx->stream);
    if (e != hipSuccess) return e;
    if (L->stage) (void)hipFree(L->stage);
    L->stage = nullptr;
    L->stage_bytes = 0;
    if ((e = hipMalloc(&L->stage, bytes)) != hipSuccess) return e;
    L->stage_bytes = bytes;
  }
  *p = static_cast<double*>(L->stage);
  return hipSuccess;
}

int kmi_upload(kmi_lattice* L, int32_t field, const void* host, int32_t host_dtype) {
  if (!L) return fail(nullptr, KMI_ERR_ARG, "kmi_upload: lattice is NULL");
  kmi_ctx* c = L->ctx;
  if (!host) return fail(c, KMI_ERR_ARG, "kmi_upload: host buffer is NULL");
  if (host_dtype != KMI_HOST_F64 && host_dtype != KMI_HOST_F32 && host_dtype != KMI_HOST_STORAGE)
    return fail(c, KMI_ERR_ARG, "kmi_upload: unknown host dtype");
  void** buf;
  double shift;
  int st;
  if (int rc = field_ptrs(L, field, buf, shift, st)) return rc;
  KMI_HIP(c, hipSetDevice(c->device));
  const size_t n = size_t(L->ncells) * L->Q;
  if (host_dtype == KMI_HOST_STORAGE) {  // storage elements verbatim (planes are unpadded)
    KMI_HIP(c, hipStreamSynchronize(c->stream));
    KMI_HIP(c, hipMemcpy(buf[L->cur], host, n * storage_bytes(st), hipMemcpyHostToDevice));
    if (field == KMI_FIELD_F) L->steps = 0;
    return KMI_OK;
  }
  std::vector<double> tmp;
  const double* src = static_cast<const double*>(host);
  if (host_dtype == KMI_HOST_F32) {
    tmp.resize(n);
    const float* h = static_cast<const float*>(host);
    for (size_t i = 0; i < n; ++i) tmp[i] = h[i];
    src = tmp.data();
  }
  double* stage = nullptr;
  KMI_HIP(c, stage_reserve(L, n * sizeof(double), &stage));
  hipError_t e = hipMemcpyAsync(stage, src, n * sizeof(double), hipMemcpyHostToDevice, c->stream);
  if (e == hipSuccess)
    e = launch_to_storage(L->Q, st, stage, buf[L->cur], L->ncells, L->plane, shift, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  if (e != hipSuccess) return hip_fail(c, e, "kmi_upload");
  if (field == KMI_FIELD_F) L->steps = 0;
  return KMI_OK;
}

int kmi_download(kmi_lattice* L, int32_t field, void* host, int32_t host_dtype) {
  if (!L) return fail(nullptr, KMI_ERR_ARG, "kmi_download: lattice is NULL");
  kmi_ctx* c = L->ctx;
  if (!host) return fail(c, KMI_ERR_ARG, "kmi_download: host buffer is NULL");
  if (host_dtype != KMI_HOST_F64 && host_dtype != KMI_HOST_F32 && host_dtype != KMI_HOST_STORAGE)
    return fail(c, KMI_ERR_ARG, "kmi_download: unknown host dtype");
  void** buf;
  double shift;
  int st;
  if (int rc = field_ptrs(L, field, buf, shift, st)) return rc;
  KMI_HIP(c, hipSetDevice(c->device));
  const size_t n = size_t(L->ncells) * L->Q;
  if (host_dtype == KMI_HOST_STORAGE) {
    KMI_HIP(c, hipStreamSynchronize(c->stream));
    KMI_HIP(c, hipMemcpy(host, buf[L->cur], n * storage_bytes(st), hipMemcpyDeviceToHost));
    return KMI_OK;
  }
  double* stage = nullptr;
  KMI_HIP(c, stage_reserve(L, n * sizeof(double), &stage));
  std::vector<double> tmp;
  double* dst = static_cast<double*>(host);
  if (host_dtype == KMI_HOST_F32) {
    tmp.resize(n);
    dst = tmp.data();
  }
  hipError_t e = launch_from_storage(L->Q, st, buf[L->cur], stage, L->ncells, L->plane, shift,
                                     c->stream);
  if (e == hipSuccess)
    e = hipMemcpyAsync(dst, stage, n * sizeof(double), hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  if (e != hipSuccess) return hip_fail(c, e, "kmi_download");
  if (host_dtype == KMI_HOST_F32) {
    float* h = static_cast<float*>(host);
    for (size_t i = 0; i < n; ++i) h[i] = float(tmp[i]);
  }
  return KMI_OK;
}

// Copies up to `count` host arrays of ncells doubles to the lattice's staging block (NULL
// stays NULL).
static int stage_fields(kmi_lattice* L, const double* const* host, int count, double** dev,
                        double*& block) {
  kmi_ctx* c = L->ctx;
  int used = 0;
  for (int k = 0; k < count; ++k) used += host[k] != nullptr;
  block = nullptr;
  for (int k = 0; k < count; ++k) dev[k] = nullptr;
  if (used == 0) return KMI_OK;
  KMI_HIP(c, stage_reserve(L, size_t(used) * L->ncells * sizeof(double), &block));
  int slot = 0;
  for (int k = 0; k < count; ++k) {
    if (!host[k]) continue;
    dev[k] = block + size_t(slot++) * L->ncells;
    hipError_t e = hipMemcpyAsync(dev[k], host[k], size_t(L->ncells) * sizeof(double),
                                  hipMemcpyHostToDevice, c->stream);
    if (e != hipSuccess) return hip_fail(c, e, "staging host field");
  }
  return KMI_OK;
}

int kmi_init_equilibrium(kmi_lattice* L, const double* rho, const double* ux, const double* uy,
                         const double* uz) {
  if (!L) return fail(nullptr, KMI_ERR_ARG, "kmi_init_equilibrium: lattice is NULL");
  kmi_ctx* c = L->ctx;
  KMI_HIP(c, hipSetDevice(c->device));
  const double* host[4] = {rho, ux, uy, L->Q == 19 ? uz : nullptr};
  double* dev[4];
  double* block;
  if (int rc = stage_fields(L, host, 4, dev, block)) return rc;
  hipError_t e = launch_init_eq(L->Q, L->desc.storage, L->f[L->cur], dev[0], dev[1], dev[2],
                                dev[3], L->ncells, L->plane, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  if (e != hipSuccess) return hip_fail(c, e, "kmi_init_equilibrium");
  L->steps = 0;
  return KMI_OK;
}

int kmi_init_temperature(kmi_lattice* L, const double* T) {
  if (!L) return fail(nullptr, KMI_ERR_ARG, "kmi_init_temperature: lattice is NULL");
  kmi_ctx* c = L->ctx;
  if (!L->desc.thermal) return fail(c, KMI_ERR_STATE, "kmi_init_temperature: lattice is not thermal");
  KMI_HIP(c, hipSetDevice(c->device));
  const double* host[1] = {T};
  double* dev[1];
  double* block;
  if (int rc = stage_fields(L, host, 1, dev, block)) return rc;
  hipError_t e = launch_init_temp(L->Q, L->desc.storage, L->g[L->cur], dev[0], L->ncells,
                                  L->plane, L->desc.T_ref, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  if (e != hipSuccess) return hip_fail(c, e, "kmi_init_temperature");
  return KMI_OK;
}

int kmi_step(kmi_lattice* L, int64_t nsteps) {
  if (!L) return fail(nullptr, KMI_ERR_ARG, "kmi_step: lattice is NULL");
  if (nsteps < 0) return fail(L->ctx, KMI_ERR_ARG, "kmi_step: nsteps must be >= 0");
  KMI_HIP(L->ctx, hipSetDevice(L->ctx->device));
  if (int rc = run_steps(L, nsteps)) return rc;
  KMI_HIP(L->ctx, hipStreamSynchronize(L->ctx->stream));
  return KMI_OK;
}

int kmi_step_timed(kmi_lattice* L, int64_t nsteps, double* ms) {
  if (!L) return fail(nullptr, KMI_ERR_ARG, "kmi_step_timed: lattice is NULL");
  kmi_ctx* c = L->ctx;
  if (nsteps < 0 || !ms) return fail(c, KMI_ERR_ARG, "kmi_step_timed: bad arguments");
  KMI_HIP(c, hipSetDevice(c->device));
  KMI_HIP(c, hipEventRecord(L->ev[0], c->stream));
  if (int rc = run_steps(L, nsteps)) return rc;
  KMI_HIP(c, hipEventRecord(L->ev[1], c->stream));
  KMI_HIP(c, hipEventSynchronize(L->ev[1]));
  float t = 0.0f;
  KMI_HIP(c, hipEventElapsedTime(&t, L->ev[0], L->ev[1]));
  *ms = t;
  return KMI_OK;
}

// Body force seen by the moments (u = (Σcf + F/2)/ρ under Guo forcing): none for
// GUO_PLAIN_MOMENTS, the per-node field for GUO_FIELD.
static MomForce mom_force(const kmi_lattice* L) {
  MomForce F{{0.0, 0.0, 0.0}, nullptr};
  if (L->desc.forcing == KMI_FORCE_GUO)
    for (int k = 0; k < 3; ++k) F.f[k] = (k < 2 || L->Q == 19) ? L->desc.force[k] : 0.0;
  if (L->desc.forcing == KMI_FORCE_GUO_FIELD) F.field = L->ffield;
  return F;
}

static int reduce_cells(kmi_lattice* L, int kind, double* out3) {
  kmi_ctx* c = L->ctx;
  hipError_t e = launch_reduce_cells(L->Q, L->desc.storage, L->f[L->cur],
                                     L->desc.thermal ? L->g[L->cur] : nullptr, L->mask, L->uprev,
                                     L->ncells, L->plane, L->desc.T_ref, mom_force(L), kind,
                                     c->partials, c->out, c->stream);
  if (e != hipSuccess) return hip_fail(c, e, "reduction kernel");
  if (kind == 101) return KMI_OK;
  KMI_HIP(c, hipMemcpyAsync(out3, c->out, 3 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  KMI_HIP(c, hipStreamSynchronize(c->stream));
  return KMI_OK;
}

int kmi_step_residual(kmi_lattice* L, int64_t nsteps, double* r_u) {
  if (!L) return fail(nullptr, KMI_ERR_ARG, "kmi_step_residual: lattice is NULL");
  kmi_ctx* c = L->ctx;
  if (nsteps < 1 || !r_u) return fail(c, KMI_ERR_ARG, "kmi_step_residual: nsteps >= 1 and r_u required");
  KMI_HIP(c, hipSetDevice(c->device));
  if (!L->uprev) KMI_HIP(c, hipMalloc(&L->uprev, size_t(3) * L->ncells * sizeof(double)));
  if (int rc = run_steps(L, nsteps - 1)) return rc;
  double tmp[3];
  if (int rc = reduce_cells(L, 101, tmp)) return rc;  // u^(k-1) → uprev
  if (int rc = run_steps(L, 1)) return rc;
  if (int rc = reduce_cells(L, 100, tmp)) return rc;
  *r_u = std::sqrt(tmp[0] / tmp[1]);  // gnf_periodic_steady.jl:192
  return KMI_OK;
}

int kmi_step_residual_window(kmi_lattice* L, int64_t nsteps, int64_t window, double* r_u) {
  if (!L) return fail(nullptr, KMI_ERR_ARG, "kmi_step_residual_window: lattice is NULL");
  kmi_ctx* c = L->ctx;
  if (window < 1 || nsteps < window || !r_u)
    return fail(c, KMI_ERR_ARG, "kmi_step_residual_window: need 1 <= window <= nsteps and r_u");
  KMI_HIP(c, hipSetDevice(c->device));
  if (!L->uprev) KMI_HIP(c, hipMalloc(&L->uprev, size_t(3) * L->ncells * sizeof(double)));
  if (int rc = run_steps(L, nsteps - window)) return rc;
  double tmp[3];
  if (int rc = reduce_cells(L, 101, tmp)) return rc;  // u^(k-m) → uprev
  if (int rc = run_steps(L, window)) return rc;
  if (int rc = reduce_cells(L, 100, tmp)) return rc;
  *r_u = std::sqrt(tmp[0] / tmp[1]) / double(window);
  return KMI_OK;
}

int kmi_moments(kmi_lattice* L, double* rho, double* ux, double* uy, double* uz, double* T) {
  if (!L) return fail(nullptr, KMI_ERR_ARG, "kmi_moments: lattice is NULL");
  kmi_ctx* c = L->ctx;
  if (T && !L->desc.thermal) return fail(c, KMI_ERR_STATE, "kmi_moments: lattice is not thermal");
  KMI_HIP(c, hipSetDevice(c->device));
  double* host[5] = {rho, ux, uy, L->Q == 19 ? uz : nullptr, T};
  int used = 0;
  for (double* h : host) used += h != nullptr;
  if (uz && L->Q == 9)
    for (long long i = 0; i < L->ncells; ++i) uz[i] = 0.0;
  if (used == 0) return KMI_OK;
  double* block = nullptr;
  KMI_HIP(c, stage_reserve(L, size_t(used) * L->ncells * sizeof(double), &block));
  double* dev[5];
  int slot = 0;
  for (int k = 0; k < 5; ++k) dev[k] = host[k] ? block + size_t(slot++) * L->ncells : nullptr;
  hipError_t e = launch_moments(L->Q, L->desc.storage, L->f[L->cur],
                                L->desc.thermal ? L->g[L->cur] : nullptr, dev[0], dev[1], dev[2],
                                dev[3], dev[4], L->ncells, L->plane, L->desc.T_ref, mom_force(L),
                                c->stream);
  for (int k = 0; k < 5 && e == hipSuccess; ++k)
    if (host[k])
      e = hipMemcpyAsync(host[k], dev[k], size_t(L->ncells) * sizeof(double),
                         hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  if (e != hipSuccess) return hip_fail(c, e, "kmi_moments");
  return KMI_OK;
}

int kmi_reduce(kmi_lattice* L, int32_t kind, int32_t arg, double* out) {
  if (!L) return fail(nullptr, KMI_ERR_ARG, "kmi_reduce: lattice is NULL");
  kmi_ctx* c = L->ctx;
  if (!out) return fail(c, KMI_ERR_ARG, "kmi_reduce: out is NULL");
  KMI_HIP(c, hipSetDevice(c->device));
  double r[3] = {0, 0, 0};
  switch (kind) {
    case KMI_RED_MASS:
    case KMI_RED_KINETIC_ENERGY:
    case KMI_RED_MAX_SPEED:
      if (int rc = reduce_cells(L, kind, r)) return rc;
      out[0] = r[0];
      return KMI_OK;
    case KMI_RED_MOMENTUM:
      if (int rc = reduce_cells(L, kind, r)) return rc;
      out[0] = r[0];
      out[1] = r[1];
      out[2] = r[2];
      return KMI_OK;
    case KMI_RED_THERMAL:
      if (!L->desc.thermal) return fail(c, KMI_ERR_STATE, "kmi_reduce: lattice is not thermal");
      if (int rc = reduce_cells(L, kind, r)) return rc;
      out[0] = r[0];
      return KMI_OK;
    case KMI_RED_NUSSELT: {
      const kmi_desc& d = L->desc;
      if (!d.thermal) return fail(c, KMI_ERR_STATE, "kmi_reduce: lattice is not thermal");
      const int nf = L->Q == 9 ? 4 : 6;
      if (arg < 0 || arg >= nf)
        return fail(c, KMI_ERR_ARG, L->Q == 9 ? "kmi_reduce: Nusselt face must be W, E, S or N"
                                              : "kmi_reduce: Nusselt face must be W, E, S, N, B or T");
      const int opp = arg ^ 1;
      auto dirichlet = [&](int f) {
        return d.faces[f].kind != KMI_FACE_PERIODIC && d.faces[f].thermal == KMI_THERMAL_DIRICHLET;
      };
      if (!dirichlet(arg) || !dirichlet(opp))
        return fail(c, KMI_ERR_STATE, "kmi_reduce: Nusselt needs Dirichlet temperatures on the named face and its opposite");
      // Wall nodes: the face's tangent indices without the edge rows of a non-periodic
      // tangent axis (thermal.jl:230-233, 530-533).
      const int64_t n[3] = {d.nx, d.ny, L->Q == 9 ? 1 : d.nz};
      const int ax = arg / 2, t1 = ax == 0 ? 1 : 0, t2 = ax == 2 ? 1 : 2;
      auto range = [&](int t, int& lo, int& hi) {
        if (t >= nf / 2) { lo = 0, hi = 1; return; }  // 2D: no second tangent
        const bool per = d.faces[2 * t].kind == KMI_FACE_PERIODIC;
        lo = per ? 0 : 1;
        hi = int(per ? n[t] : n[t] - 1);
      };
      int lo1, hi1, lo2, hi2;
      range(t1, lo1, hi1);
      range(t2, lo2, hi2);
      if (hi1 - lo1 < 1 || hi2 - lo2 < 1 || n[ax] < 2)
        return fail(c, KMI_ERR_STATE, "kmi_reduce: grid too small for the Nusselt estimator");
      const double Th = d.faces[arg].T, dT = d.faces[arg].T - d.faces[opp].T;
      hipError_t e = launch_nusselt(L->Q, d.storage, L->g[L->cur], int(d.nx), int(d.ny),
                                    int(n[2]), L->plane, d.T_ref, arg, Th, dT, lo1, hi1, lo2, hi2,
                                    c->partials, c->out, c->stream);
      if (e != hipSuccess) return hip_fail(c, e, "Nusselt kernel");
      KMI_HIP(c, hipMemcpyAsync(r, c->out, sizeof(double), hipMemcpyDeviceToHost, c->stream));
      KMI_HIP(c, hipStreamSynchronize(c->stream));
      out[0] = r[0] / double(int64_t(hi1 - lo1) * (hi2 - lo2));
      return KMI_OK;
    }
    case KMI_RED_DRAG_MEA:
    case KMI_RED_DRAG_HALFWAY: {
      if (L->nlinks > 0) {
        const kmi_desc& d = L->desc;
        hipError_t e = launch_drag(L->Q, d.storage, L->f[L->cur], L->nlinks, L->link_cell,
                                   L->link_q, L->link_qw, L->link_uw, int(d.nx), int(d.ny),
                                   L->args.nz, L->plane, kind == KMI_RED_DRAG_MEA, c->partials,
                                   c->out, c->stream);
        if (e != hipSuccess) return hip_fail(c, e, "drag kernel");
        KMI_HIP(c, hipMemcpyAsync(r, c->out, 3 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
        KMI_HIP(c, hipStreamSynchronize(c->stream));
      }
      out[0] = r[0];
      out[1] = r[1];
      out[2] = r[2];
      return KMI_OK;
    }
    default:
      return fail(c, KMI_ERR_ARG, "kmi_reduce: unknown reduction");
  }
}

int kmi_lattice_steps(const kmi_lattice* L, int64_t* steps) {
  if (!L || !steps) return fail(nullptr, KMI_ERR_ARG, "kmi_lattice_steps: NULL argument");
  *steps = L->steps;
  return KMI_OK;
}

int kmi_lattice_set_steps(kmi_lattice* L, int64_t steps) {
  if (!L) return fail(nullptr, KMI_ERR_ARG, "kmi_lattice_set_steps: lattice is NULL");
  if (steps < 0) return fail(L->ctx, KMI_ERR_ARG, "kmi_lattice_set_steps: steps must be >= 0");
  L->steps = steps;
  return KMI_OK;
}

int kmi_lattice_set_launch(kmi_lattice* L, int32_t bx, int32_t by, int32_t nontemporal) {
  if (!L) return fail(nullptr, KMI_ERR_ARG, "kmi_lattice_set_launch: lattice is NULL");
  if (bx < 1 || by < 1 || bx * by > 256 || (bx * by) % 64 != 0 || (L->desc.ny + by - 1) / by > 65535)
    return fail(L->ctx, KMI_ERR_ARG,
                "kmi_lattice_set_launch: need bx*by a multiple of 64 and at most 256");
  KMI_HIP(L->ctx, hipSetDevice(L->ctx->device));
  KMI_HIP(L->ctx, hipStreamSynchronize(L->ctx->stream));
  drop_graphs(L);
  set_geometry(L, bx, by, nontemporal ? 1 : 0);
  return KMI_OK;
}

int kmi_lattice_get_launch(const kmi_lattice* L, int32_t* bx, int32_t* by, int32_t* nontemporal) {
  if (!L || !bx || !by || !nontemporal)
    return fail(nullptr, KMI_ERR_ARG, "kmi_lattice_get_launch: NULL argument");
  *bx = L->bx;
  *by = L->by;
  *nontemporal = L->nts;
  return KMI_OK;
}

int kmi_lattice_set_temporal(kmi_lattice* L, int32_t tile, int32_t zchunks) {
  if (!L) return fail(nullptr, KMI_ERR_ARG, "kmi_lattice_set_temporal: lattice is NULL");
  if (tile != 0 && !tb_eligible(L, tile))
    return fail(L->ctx, KMI_ERR_UNSUPPORTED,
                "kmi_lattice_set_temporal: multi-step launches need a D3Q19 BGK/TRT lattice "
                "without thermal populations, nz >= 4, whose nx, ny are multiples of the tile, "
                "fully periodic without mask, or with walls, moving walls, constant Zou-He faces "
                "or a solid mask (no forcing, no cut links; tiles 12, 14 and 9 (not FP64), not "
                "F16S) "
                "(two steps 1: 32x8, 2: 64x4, 3: 64x8, 4: 32x16, 5: 64x16 (not FP64), "
                "12: 64x8, 13: 32x16, 14: 32x8; three steps 6: 64x8 (not FP64), 7: 32x8, "
                "8: 32x16 (not FP64), 9: 32x8, 10: 32x16 (not FP64), 11: 64x8 (not FP64)), "
                "or a D2Q9 lattice without cut links or per-node faces (thermal: not FP64), "
                "at least the tile plus its ring on each axis (1: 28x28 2 steps, 2: 24x24 4 "
                "steps, 3: 56x8 4 steps, 4: 16x16 8 steps, 5: 60x12 2 steps, 6: 8x8 4 steps, "
                "7: 16x16 4 steps, 8: 8x8 8 steps)");
  if (tile != 0 && (zchunks < 1 || zchunks > L->args.nz))
    return fail(L->ctx, KMI_ERR_ARG, "kmi_lattice_set_temporal: zchunks must be in 1..nz");
  KMI_HIP(L->ctx, hipSetDevice(L->ctx->device));
  KMI_HIP(L->ctx, hipStreamSynchronize(L->ctx->stream));
  drop_graphs(L);
  L->tb_tile = tile;
  L->tb_zchunks = tile ? zchunks : 1;
  return KMI_OK;
}

int kmi_lattice_get_temporal(const kmi_lattice* L, int32_t* tile, int32_t* zchunks) {
  if (!L || !tile || !zchunks)
    return fail(nullptr, KMI_ERR_ARG, "kmi_lattice_get_temporal: NULL argument");
  *tile = L->tb_tile;
  *zchunks = L->tb_zchunks;
  return KMI_OK;
}

// Times each candidate launch shape × store policy on the lattice itself (graph-replayed
// steps between HIP events, best of two), then the two-step variants where they apply,
// and keeps the fastest.  The populations, buffer
// parity and step counter are restored afterwards, so tuning leaves the run unchanged.
int kmi_lattice_autotune(kmi_lattice* L, int64_t steps_per_trial, double* best_ms_per_step) {
  if (!L) return fail(nullptr, KMI_ERR_ARG, "kmi_lattice_autotune: lattice is NULL");
  kmi_ctx* c = L->ctx;
  KMI_HIP(c, hipSetDevice(c->device));
  KMI_HIP(c, hipStreamSynchronize(c->stream));
  const int64_t trial = std::max<int64_t>(kGraphSteps, (steps_per_trial + kGraphSteps - 1) /
                                                           kGraphSteps * kGraphSteps);
  const size_t bytes = size_t(L->plane) * L->Q * L->esize;
  const size_t gbytes = size_t(L->plane) * L->Q * L->gsize;
  void* save_f = nullptr;
  void* save_g = nullptr;
  KMI_HIP(c, hipMalloc(&save_f, bytes));
  hipError_t e = hipMemcpyAsync(save_f, L->f[L->cur], bytes, hipMemcpyDeviceToDevice, c->stream);
  if (e == hipSuccess && L->desc.thermal) {
    e = hipMalloc(&save_g, gbytes);
    if (e == hipSuccess)
      e = hipMemcpyAsync(save_g, L->g[L->cur], gbytes, hipMemcpyDeviceToDevice, c->stream);
  }
  const int cur0 = L->cur;
  const int64_t steps0 = L->steps;
  // Time the current configuration: one warm-up graph, then the best of two timed runs.
  auto time_current = [&]() -> double {
    double t = 1e300;
    drop_graphs(L);
    if (run_steps(L, kGraphSteps) != KMI_OK) { e = hipErrorUnknown; return t; }
    for (int rep = 0; rep < 2 && e == hipSuccess; ++rep) {
      e = hipEventRecord(L->ev[0], c->stream);
      if (e == hipSuccess && run_steps(L, trial) != KMI_OK) e = hipErrorUnknown;
      if (e == hipSuccess) e = hipEventRecord(L->ev[1], c->stream);
      if (e == hipSuccess) e = hipEventSynchronize(L->ev[1]);
      float ms = 0.f;
      if (e == hipSuccess) e = hipEventElapsedTime(&ms, L->ev[0], L->ev[1]);
      t = std::min(t, double(ms) / double(trial));
    }
    return t;
  };
  const int shapes[][2] = {{256, 1}, {128, 1}, {128, 2}, {64, 1}, {64, 2}, {64, 4}, {32, 2},
                           {32, 4}, {32, 8}, {16, 4}, {16, 8}, {16, 16}};
  double best = 1e300;
  int bb[5] = {L->bx, L->by, L->nts, L->tb_tile, L->tb_zchunks};
  L->tb_tile = 0;
  for (const auto& sh : shapes) {
    if (e != hipSuccess) break;
    const int bx = sh[0], by = sh[1];
    if (bx > ((L->desc.nx + 63) / 64) * 64 && bx > 64) continue;  // mostly idle lanes
    if (by > 1 && by > L->desc.ny) continue;
    if ((L->desc.ny + by - 1) / by > 65535) continue;
    for (int nts = 0; nts < 2 && e == hipSuccess; ++nts) {
      set_geometry(L, bx, by, nts);
      const double t = time_current();
      if (e == hipSuccess && t < best) {
        best = t;
        bb[0] = bx, bb[1] = by, bb[2] = nts, bb[3] = 0, bb[4] = 1;
      }
    }
  }
  // Two- and three-step launches (the single-step shape found above serves remainders).
  const int zc_list[] = {1, 2, 3, 4, 6, 8, 12, 16};
  for (int tile = 1; tile <= 14 && e == hipSuccess; ++tile) {
    if (!tb_eligible(L, tile)) continue;
    for (int zc : zc_list) {
      if (zc > L->args.nz || e != hipSuccess) continue;
      for (int nts = 0; nts < 2 && e == hipSuccess; ++nts) {
        set_geometry(L, bb[0], bb[1], nts);
        L->tb_tile = tile;
        L->tb_zchunks = zc;
        const double t = time_current();
        if (e == hipSuccess && t < best) {
          best = t;
          bb[2] = nts, bb[3] = tile, bb[4] = zc;
        }
      }
    }
  }
  drop_graphs(L);
  set_geometry(L, bb[0], bb[1], bb[2]);
  L->tb_tile = bb[3];
  L->tb_zchunks = bb[4];
  hipError_t e2 = hipMemcpyAsync(L->f[cur0], save_f, bytes, hipMemcpyDeviceToDevice, c->stream);
  if (e2 == hipSuccess && save_g)
    e2 = hipMemcpyAsync(L->g[cur0], save_g, gbytes, hipMemcpyDeviceToDevice, c->stream);
  if (e2 == hipSuccess) e2 = hipStreamSynchronize(c->stream);
  (void)hipFree(save_f);
  if (save_g) (void)hipFree(save_g);
  L->cur = cur0;
  L->steps = steps0;
  if (e != hipSuccess) return hip_fail(c, e, "kmi_lattice_autotune");
  if (e2 != hipSuccess) return hip_fail(c, e2, "kmi_lattice_autotune (restore)");
  if (best_ms_per_step) *best_ms_per_step = best;
  return KMI_OK;
}

int kmi_measure_copy_bandwidth(kmi_ctx* ctx, int64_t bytes, int32_t reps, double* gbps) {
  if (!ctx || !gbps || bytes < 1024 || reps < 1)
    return fail(ctx, KMI_ERR_ARG, "kmi_measure_copy_bandwidth: bad arguments");
  KMI_HIP(ctx, hipSetDevice(ctx->device));
  bytes = bytes / 256 * 256;
  void *a = nullptr, *b = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  hipError_t e = hipMalloc(&a, size_t(bytes));
  if (e == hipSuccess) e = hipMalloc(&b, size_t(bytes));
  if (e == hipSuccess) e = hipMemsetAsync(a, 0, size_t(bytes), ctx->stream);
  if (e == hipSuccess) e = hipEventCreate(&e0);
  if (e == hipSuccess) e = hipEventCreate(&e1);
  if (e == hipSuccess) e = launch_copy_probe(a, b, bytes, ctx->stream);  // warm-up
  if (e == hipSuccess) e = hipEventRecord(e0, ctx->stream);
  for (int r = 0; r < reps && e == hipSuccess; ++r) e = launch_copy_probe(a, b, bytes, ctx->stream);
  if (e == hipSuccess) e = hipEventRecord(e1, ctx->stream);
  if (e == hipSuccess) e = hipEventSynchronize(e1);
  float ms = 0.f;
  if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  if (a) (void)hipFree(a);
  if (b) (void)hipFree(b);
  if (e != hipSuccess) return hip_fail(ctx, e, "kmi_measure_copy_bandwidth");
  *gbps = 2.0 * double(bytes) * reps / (double(ms) * 1e-3) / 1e9;
  return KMI_OK;
}

int kmi_lattice_device_bytes(const kmi_lattice* L, int64_t* bytes) {
  if (!L || !bytes) return fail(nullptr, KMI_ERR_ARG, "kmi_lattice_device_bytes: NULL argument");
  int64_t b = 2 * int64_t(L->plane) * L->Q * int64_t(L->esize);
  if (L->desc.thermal) b += 2 * int64_t(L->plane) * L->Q * int64_t(L->gsize);
  if (L->mask) b += L->ncells;
  if (L->uprev) b += 3 * L->ncells * int64_t(sizeof(double));
  b += int64_t(L->stage_bytes);
  *bytes = b;
  return KMI_OK;
}

}  // extern "C"
