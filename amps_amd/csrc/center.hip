// center.hip — cell-centred linear interpolation of B on the AMR blocks
// (PIC::InterpolationRoutines::CellCentered::Linear::InitStencil,
// pic_interpolation_routines.cpp:224-1070, AMPS branch of an AMR mesh).
//
// B is given at cell centres with one ghost layer per block, as the reference
// stages it for the mover (SetBlock_B, pic_mover.cpp:84-186): cell (i,j,k) in
// [-1,8]^3 at (i+1) + 10((j+1) + 10(k+1)); the caller fills the ghosts.  A
// ghost cell exists when its centre lies inside the domain.  The stencil is
// never stored: each support point is accumulated into B as it is produced,
// in the reference's stencil order, with the coarse/fine blend merging
// duplicate cells exactly as cStencilGeneric::Add does (pic.h:7273-7291).
#include <hip/hip_runtime.h>

#include <cstring>
#include <vector>

#include "amps_common.h"
#include "center_device.h"
#include "pic_device.h"

using namespace amps;

namespace {

__global__ __launch_bounds__(256) void k_interp_center(MeshView m, const double* __restrict__ Bc, int64_t n,
                                                        const double* __restrict__ x,
                                                        const int* __restrict__ blk, double* __restrict__ B) {
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  const double xp[3] = {x[p], x[n + p], x[2 * n + p]};
  double b[3];
  center_B(m, Bc, xp, blk[p], b);
  B[p] = b[0];
  B[n + p] = b[1];
  B[2 * n + p] = b[2];
}

__global__ __launch_bounds__(256) void k_mode3d_field(MeshView m, const double* __restrict__ Bm, int64_t n,
                                                      const double* __restrict__ x, double* __restrict__ B) {
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  const double xp[3] = {x[3 * p], x[3 * p + 1], x[3 * p + 2]};
  double b[3];
  mode3d_B(m, Bm, xp, b);
  B[3 * p] = b[0];
  B[3 * p + 1] = b[1];
  B[3 * p + 2] = b[2];
}

}  // namespace

// The compact Mode3D field of every leaf's interior cells
// (GlobalMagneticField.h:18-30: globalCell = Temp_ID*512 + i + 8(j + 8k),
// block id = Temp_ID), resident for the tracer's mesh field.
extern "C" int amps_gpu_set_mode3d_field(amps_ctx* c, const double* B_cells) {
  if (!c || !B_cells) return fail(AMPS_E_ARG, "amps_gpu_set_mode3d_field: null argument");
  if (!c->has_mesh) return fail(AMPS_E_STATE, "amps_gpu_set_mode3d_field: set the mesh first");
  (void)hipSetDevice(c->device);
  const size_t bytes = sizeof(double) * 3 * NCELL * (size_t)c->mv.nb;
  if (!c->d_Bm) HIP_TRY(hipMalloc(&c->d_Bm, bytes));
  HIP_TRY(hipMemcpy(c->d_Bm, B_cells, bytes, hipMemcpyHostToDevice));
  c->has_mode3d = true;
  return AMPS_OK;
}

extern "C" int amps_gpu_mode3d_field(amps_ctx* c, int64_t n, const double* x, double* B) {
  if (!c) return fail(AMPS_E_ARG, "null context");
  if (!c->has_mode3d) return fail(AMPS_E_STATE, "amps_gpu_mode3d_field: set the Mode3D field first");
  if (n < 0 || n > ((int64_t)1 << 31)) return fail(AMPS_E_RANGE, "amps_gpu_mode3d_field: n out of range");
  if (n == 0) return AMPS_OK;
  if (!x || !B) return fail(AMPS_E_ARG, "amps_gpu_mode3d_field: null array");
  (void)hipSetDevice(c->device);
  double *dx = nullptr, *dB = nullptr;
  HIP_TRY(hipMalloc(&dx, sizeof(double) * 3 * n));
  HIP_TRY(hipMalloc(&dB, sizeof(double) * 3 * n));
  HIP_TRY(hipMemcpyAsync(dx, x, sizeof(double) * 3 * n, hipMemcpyHostToDevice, c->stream));
  hipLaunchKernelGGL(k_mode3d_field, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c->stream, c->mv, c->d_Bm, n,
                     dx, dB);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipMemcpyAsync(B, dB, sizeof(double) * 3 * n, hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  (void)hipFree(dx);
  (void)hipFree(dB);
  return AMPS_OK;
}

extern "C" int amps_gpu_set_center_fields(amps_ctx* c, const double* Bc) {
  if (!c || !Bc) return fail(AMPS_E_ARG, "amps_gpu_set_center_fields: null argument");
  if (!c->has_mesh) return fail(AMPS_E_STATE, "amps_gpu_set_center_fields: set the mesh first");
  (void)hipSetDevice(c->device);
  const size_t bytes = sizeof(double) * 3 * CC_CELLS * (size_t)c->mv.nb;
  if (!c->d_Bc) HIP_TRY(hipMalloc(&c->d_Bc, bytes));
  HIP_TRY(hipMemcpy(c->d_Bc, Bc, bytes, hipMemcpyHostToDevice));
  c->has_center = true;
  return AMPS_OK;
}

extern "C" int amps_gpu_interp_center(amps_ctx* c, int64_t n, const double* x, const int32_t* blk,
                                      double* B) {
  if (!c) return fail(AMPS_E_ARG, "null context");
  if (!c->has_center) return fail(AMPS_E_STATE, "amps_gpu_interp_center: set the cell-centred field first");
  if (n < 0 || n > ((int64_t)1 << 31)) return fail(AMPS_E_RANGE, "amps_gpu_interp_center: n out of range");
  if (n == 0) return AMPS_OK;
  if (!x || !blk || !B) return fail(AMPS_E_ARG, "amps_gpu_interp_center: null array");
  for (int64_t i = 0; i < n; i++)
    if (blk[i] < 0 || blk[i] >= c->mv.nb) return fail(AMPS_E_ARG, "amps_gpu_interp_center: block id out of range");
  (void)hipSetDevice(c->device);
  double *dx = nullptr, *dB = nullptr;
  int* db = nullptr;
  HIP_TRY(hipMalloc(&dx, sizeof(double) * 3 * n));
  HIP_TRY(hipMalloc(&dB, sizeof(double) * 3 * n));
  HIP_TRY(hipMalloc(&db, sizeof(int) * n));
  HIP_TRY(hipMemcpyAsync(dx, x, sizeof(double) * 3 * n, hipMemcpyHostToDevice, c->stream));
  HIP_TRY(hipMemcpyAsync(db, blk, sizeof(int) * n, hipMemcpyHostToDevice, c->stream));
  hipLaunchKernelGGL(k_interp_center, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c->stream, c->mv, c->d_Bc,
                     n, dx, db, dB);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipMemcpyAsync(B, dB, sizeof(double) * 3 * n, hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  (void)hipFree(dx);
  (void)hipFree(dB);
  (void)hipFree(db);
  return AMPS_OK;
}
