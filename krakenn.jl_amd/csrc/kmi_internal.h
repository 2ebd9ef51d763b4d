// kmi_internal.h — host-side internals shared by the C-ABI translation units.
#pragma once

#include <string>
#include <vector>

#include "kmi_common.h"

struct kmi_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  std::string err;
  double* partials = nullptr;  // 2048 groups × 3 values × (sum, error)
  double* out = nullptr;       // 8 values
};

namespace kmi {

// Record `msg` as the last error of `ctx` (or of the calling thread when ctx is NULL).
int fail(kmi_ctx* ctx, int code, const std::string& msg);
int hip_fail(kmi_ctx* ctx, hipError_t e, const char* what);

hipError_t launch_step(int q, int st, int coll, bool th, bool bnd, bool nts, const StepArgs& a,
                       dim3 g, dim3 b, hipStream_t s);
// Two or three steps per launch on a fully periodic D3Q19 BGK/TRT lattice (kmi_step2.hip);
// `tile` selects the work-group tile and the steps per launch (step2_tile, temporal_span),
// `zchunks` the z-ranges per column.
bool step2_tile(int tile, int st, int& TX, int& TY);
// Steps one launch of `tile` advances: 2 (tiles 1–5, 12–14), 3 (tiles 6–11), 1 (tile 0).
int temporal_span(int tile);
// bnd: a lattice with walls, moving walls, Zou–He faces or a solid mask (step2_tile_walled
// tiles only; non-temporal stores whatever `nts` says).
bool step2_tile_walled(int tile, int st);
// Work-group → tile table of a walled multi-step launch (ring tiles first in each XCD band).
void walled_tile_order(const StepArgs& a, int tile, int zchunks, std::vector<int>& order);
hipError_t launch_step2(int st, int coll, int tile, int zchunks, bool nts, bool bnd, const StepArgs& a,
                        hipStream_t s);
// K steps per launch on a D2Q9 lattice (kmi_step_tb2d.hip): `tile` selects tile and K
// (tb2d_variant).
bool tb2d_variant(int tile, int& TW, int& TH, int& K);
hipError_t launch_step_tb2d(int st, int coll, int tile, bool th, const StepArgs& a, hipStream_t s);
hipError_t launch_copy_probe(const void* src, void* dst, long long bytes, hipStream_t s);
hipError_t launch_to_storage(int Q, int st, const double* src, void* dst, long long ncells,
                             long long plane, double shift, hipStream_t s);
hipError_t launch_from_storage(int Q, int st, const void* src, double* dst, long long ncells,
                               long long plane, double shift, hipStream_t s);
hipError_t launch_init_eq(int Q, int st, void* dst, const double* rho, const double* ux,
                          const double* uy, const double* uz, long long ncells, long long plane,
                          hipStream_t s);
hipError_t launch_init_temp(int Q, int st, void* dst, const double* T, long long ncells,
                            long long plane, double T0, hipStream_t s);
hipError_t launch_moments(int Q, int st, const void* f, const void* g, double* rho, double* ux,
                          double* uy, double* uz, double* T, long long ncells, long long plane,
                          double T0, const MomForce& F, hipStream_t s);
int reduce_groups(long long ncells);
hipError_t launch_reduce_cells(int Q, int st, const void* f, const void* g, const uint8_t* mask,
                               double* uprev, long long ncells, long long plane, double T0,
                               const MomForce& F, int kind, double* partials, double* out,
                               hipStream_t s);
hipError_t launch_drag(int Q, int st, const void* f, long long nlinks, const int* cell,
                       const int* q, const double* qw, const double* uw, int nx, int ny, int nz,
                       long long plane, int mea, double* partials, double* out, hipStream_t s);
// Static device workspace of kmi_gemm_bf16_noalloc (kmi_nn.hip): its address on `device`,
// taken with hipGetSymbolAddress while that device is current (once per device, at context
// creation); nullptr until then.
hipError_t resolve_noalloc_workspace(int device);
void* noalloc_workspace(int device);
hipError_t launch_nusselt(int Q, int st, const void* g, int nx, int ny, int nz, long long plane,
                          double T0, int face, double Th, double dT, int lo1, int hi1, int lo2,
                          int hi2, double* partials, double* out, hipStream_t s);

}  // namespace kmi

#define KMI_HIP(ctx, call)                                         \
  do {                                                             \
    hipError_t _e = (call);                                        \
    if (_e != hipSuccess) return ::kmi::hip_fail(ctx, _e, #call);  \
  } while (0)
