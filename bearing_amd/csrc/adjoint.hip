// The adjoint form on the MI355X: one workgroup (one wave) per world steps its held
// trajectory forward, keeping checkpoints, and backward, each segment stepped again with
// its tape and walked in reverse (the arithmetic is adjoint.h's). A world's colours and
// substeps are dependent sweeps ended by workgroup barriers, as in the soft step; the
// worlds of a batch fill the chip.
#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "adjoint_dev.h"
#include "held_layout.h"

namespace bm {

constexpr int HELD_THREADS = 64;

struct HeldDevice {
  int device = 0;
  char* mem = nullptr;
  size_t bytes = 0;
  adj::Trajectory T{};
  adj::HeldIO io{};
  hipEvent_t begin = nullptr, end = nullptr;
  float forward_ms = 0, backward_ms = 0;
  bool forwarded = false;
  size_t rows = 0;
  std::vector<int32_t> body_offset, row_offset;  // the host's copies, for branch records
};

namespace {

struct Barrier {
  __device__ void operator()() const { __syncthreads(); }
};

__global__ void __launch_bounds__(HELD_THREADS) held_forward_kernel(adj::Trajectory T, adj::HeldIO io) {
  const int wld = blockIdx.x;
  const adj::Ctx c{(int)threadIdx.x, (int)blockDim.x, nullptr};
  adj::start_world(T, io, wld, c.tid, c.threads);
  __syncthreads();
  adj::forward_world(c, Barrier{}, T, wld);
  __syncthreads();
  adj::end_world(T, io, wld, c.tid, c.threads);
}

__global__ void __launch_bounds__(HELD_THREADS) held_backward_kernel(adj::Trajectory T, adj::HeldIO io) {
  // Each thread's row adjoints while a row's reverse runs (72 KB a workgroup: two worlds a CU).
  __shared__ double slots[HELD_THREADS * adj::ROW_SLOT];
  const int wld = blockIdx.x;
  const adj::Ctx c{(int)threadIdx.x, (int)blockDim.x, slots + threadIdx.x * adj::ROW_SLOT};
  adj::start_bars(T, io, wld, c.tid, c.threads);
  __syncthreads();
  adj::backward_world(c, Barrier{}, T, wld);
  __syncthreads();
  adj::end_bars(T, io, wld, c.tid, c.threads);
  if (c.tid == 0) adj::parameters_world(T, wld, io.mass_bar, io.friction_bar, io.restitution_bar);
}

int hip_fail(hipError_t e, const char* what, std::string& err) {
  err = std::string(what) + ": " + hipGetErrorString(e);
  return BM_ERR_DEVICE;
}

}  // namespace

void held_free(HeldDevice* h) {
  if (!h) return;
  (void)hipSetDevice(h->device);
  if (h->mem) (void)hipFree(h->mem);
  if (h->begin) (void)hipEventDestroy(h->begin);
  if (h->end) (void)hipEventDestroy(h->end);
  delete h;
}

int held_load(int device, const bm_held* held, HeldDevice** out, std::string& err) {
  if (!held || held->worlds <= 0 || held->steps <= 0 || held->every <= 0 || held->bodies <= 0) {
    err = "bm_held_load: an empty or malformed trajectory";
    return BM_ERR_ARGUMENT;
  }
  held_free(*out);
  *out = nullptr;
  auto* h = new HeldDevice();
  h->device = device;
  hipError_t e;
  if ((e = hipSetDevice(device)) != hipSuccess) { held_free(h); return hip_fail(e, "hipSetDevice", err); }
  std::vector<int32_t> work_at;
  h->bytes = adj::lay_out(*held, nullptr, h->T, h->io, nullptr, work_at);
  if ((e = hipMalloc(&h->mem, h->bytes)) != hipSuccess) { held_free(h); return hip_fail(e, "hipMalloc", err); }
  std::vector<adj::Copy> copies;
  adj::lay_out(*held, h->mem, h->T, h->io, &copies, work_at);
  for (const adj::Copy& c : copies)
    if ((e = hipMemcpy(c.to, c.from, c.bytes, hipMemcpyHostToDevice)) != hipSuccess) {
      held_free(h);
      return hip_fail(e, "hipMemcpy", err);
    }
  if ((e = hipEventCreate(&h->begin)) != hipSuccess || (e = hipEventCreate(&h->end)) != hipSuccess) {
    held_free(h);
    return hip_fail(e, "hipEventCreate", err);
  }
  h->rows = held->rows;
  h->body_offset.assign(held->body_offset, held->body_offset + held->worlds + 1);
  h->row_offset.assign(held->row_offset, held->row_offset + (size_t)held->steps * held->worlds + 1);
  *out = h;
  return BM_OK;
}

int held_forward(HeldDevice* h, hipStream_t stream, const double* state0, const double* mass, const double* friction,
                 const double* restitution, double* state_out, std::string& err) {
  if (!h) { err = "bm_held_forward: nothing loaded (bm_held_load)"; return BM_ERR_ARGUMENT; }
  const size_t nb = h->T.bodies;
  hipError_t e;
  (void)hipSetDevice(h->device);
  const struct { double* to; const double* from; size_t n; } in[] = {
      {h->io.state0, state0, BM_HELD_STATE * nb}, {h->io.mass_in, mass, nb}, {h->io.friction_in, friction, nb},
      {h->io.restitution_in, restitution, nb}};
  for (const auto& c : in)
    if ((e = hipMemcpyAsync(c.to, c.from, c.n * 8, hipMemcpyHostToDevice, stream)) != hipSuccess)
      return hip_fail(e, "hipMemcpyAsync", err);
  (void)hipEventRecord(h->begin, stream);
  hipLaunchKernelGGL(held_forward_kernel, dim3(h->T.worlds), dim3(HELD_THREADS), 0, stream, h->T, h->io);
  if ((e = hipGetLastError()) != hipSuccess) return hip_fail(e, "held_forward_kernel", err);
  (void)hipEventRecord(h->end, stream);
  if ((e = hipMemcpyAsync(state_out, h->io.state_out, BM_HELD_STATE * nb * 8, hipMemcpyDeviceToHost, stream)) != hipSuccess)
    return hip_fail(e, "hipMemcpyAsync", err);
  if ((e = hipStreamSynchronize(stream)) != hipSuccess) return hip_fail(e, "held_forward_kernel", err);
  (void)hipEventElapsedTime(&h->forward_ms, h->begin, h->end);
  h->forwarded = true;
  return BM_OK;
}

int held_backward(HeldDevice* h, hipStream_t stream, const double* state_bar, double* state0_bar, double* mass_bar,
                  double* friction_bar, double* restitution_bar, std::string& err) {
  if (!h || !h->forwarded) { err = "bm_held_backward: no forward pass to go back over"; return BM_ERR_ARGUMENT; }
  const size_t nb = h->T.bodies;
  hipError_t e;
  (void)hipSetDevice(h->device);
  if ((e = hipMemcpyAsync(h->io.state_bar, state_bar, BM_HELD_STATE * nb * 8, hipMemcpyHostToDevice, stream)) != hipSuccess)
    return hip_fail(e, "hipMemcpyAsync", err);
  // the rows' adjoints start at 0: the stores', mu's and e's columns lie one after another
  const adj::RowsBar& RB = h->T.RB;
  const size_t span = (size_t)((char*)(RB.e + h->rows) - (char*)RB.store_impulse);
  if ((e = hipMemsetAsync(RB.store_impulse, 0, span, stream)) != hipSuccess)
    return hip_fail(e, "hipMemsetAsync", err);
  (void)hipEventRecord(h->begin, stream);
  hipLaunchKernelGGL(held_backward_kernel, dim3(h->T.worlds), dim3(HELD_THREADS), 0, stream, h->T, h->io);
  if ((e = hipGetLastError()) != hipSuccess) return hip_fail(e, "held_backward_kernel", err);
  (void)hipEventRecord(h->end, stream);
  const struct { double* to; const double* from; size_t n; } out[] = {
      {state0_bar, h->io.state0_bar, BM_HELD_STATE * nb}, {mass_bar, h->io.mass_bar, nb},
      {friction_bar, h->io.friction_bar, nb}, {restitution_bar, h->io.restitution_bar, nb}};
  for (const auto& c : out)
    if ((e = hipMemcpyAsync(c.to, c.from, c.n * 8, hipMemcpyDeviceToHost, stream)) != hipSuccess)
      return hip_fail(e, "hipMemcpyAsync", err);
  if ((e = hipStreamSynchronize(stream)) != hipSuccess) return hip_fail(e, "held_backward_kernel", err);
  (void)hipEventElapsedTime(&h->backward_ms, h->begin, h->end);
  return BM_OK;
}

int held_branches(HeldDevice* h, hipStream_t stream, uint64_t* per_world, std::string& err) {
  if (!h || !h->forwarded) { err = "bm_held_branches: no forward pass"; return BM_ERR_ARGUMENT; }
  (void)hipSetDevice(h->device);
  std::vector<uint64_t> body(h->T.bodies), row(h->rows ? h->rows : 1);
  hipError_t e;
  if ((e = hipMemcpyAsync(body.data(), h->T.body_branch, body.size() * 8, hipMemcpyDeviceToHost, stream)) != hipSuccess ||
      (e = hipMemcpyAsync(row.data(), h->T.row_branch, row.size() * 8, hipMemcpyDeviceToHost, stream)) != hipSuccess ||
      (e = hipStreamSynchronize(stream)) != hipSuccess)
    return hip_fail(e, "bm_held_branches: copy", err);
  adj::fold_branches(h->T.worlds, h->T.steps, h->body_offset.data(), h->row_offset.data(), body.data(), row.data(),
                     per_world);
  return BM_OK;
}

void held_times(const HeldDevice* h, double* forward_ms, double* backward_ms) {
  if (forward_ms) *forward_ms = h ? h->forward_ms : 0;
  if (backward_ms) *backward_ms = h ? h->backward_ms : 0;
}

}  // namespace bm
