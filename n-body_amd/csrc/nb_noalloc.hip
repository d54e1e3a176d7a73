// nb_noalloc.hip — nb_bh_forces_noalloc: the Barnes-Hut force pass of a context
// (nb_bh_build + nb_bh_traverse, the kernels behind nb_forces) on caller-owned
// device buffers, with every intermediate array carved out of one static device
// array instead of hipMalloc'd memory. No counterpart in the reference; it is
// for callers that cannot allocate (stream capture, serving loops).
#include <cmath>
#include <mutex>

#include "nb_internal.h"

constexpr int64_t kScratchBytes = nb_bh_layout_for(kNoAllocMaxBodies).off[kBhBufCount];

// The scratch of every call on a device (22.2 MB for 65536 bodies). In host
// code this name is only a handle for hipGetSymbolAddress: the device address of
// each GPU's copy is looked up once per device and kept in g_scratch.
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch[kMaxDevices] = {};

// Device address of this GPU's scratch (the current device must be `dev`).
int scratch_address(int dev, char** out) {
  if (dev < 0 || dev >= kMaxDevices)
    return nb_fail(NB_E_ARG, "nb_bh_forces_noalloc: device index out of range");
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return nb_fail(NB_E_HIP, std::string("nb_bh_forces_noalloc: hipGetSymbolAddress(scratch): ") +
                                   hipGetErrorString(e));
    g_scratch[dev] = p;
  }
  *out = static_cast<char*>(g_scratch[dev]);
  return NB_OK;
}

}  // namespace

extern "C" int nb_bh_forces_noalloc(const float* pm4, int64_t n, float eps2, float G, float theta,
                                    int32_t leaf_size, int32_t flags, double* force3,
                                    int32_t* status, void* stream) {
  NB_TRY
  NB_RANGE("nb_bh_forces_noalloc");
  if (!pm4 || !force3) return nb_fail(NB_E_ARG, "nb_bh_forces_noalloc: null buffer");
  if (n < 1 || n > kNoAllocMaxBodies)
    return nb_fail(NB_E_ARG, "nb_bh_forces_noalloc: n must be in [1, NB_NOALLOC_MAX_BODIES = 65536]");
  if (!(eps2 >= 0.f) || !std::isfinite(eps2))
    return nb_fail(NB_E_ARG, "nb_bh_forces_noalloc: eps2 must be finite and >= 0");
  if (!(theta >= 0.f) || !std::isfinite(theta))
    return nb_fail(NB_E_ARG, "nb_bh_forces_noalloc: theta must be finite and >= 0");
  if (leaf_size < 0 || leaf_size > 64)
    return nb_fail(NB_E_ARG, "nb_bh_forces_noalloc: leaf_size must be in [0, 64]");
  if (flags & ~(NB_F_BH_AXIS_BOX | NB_F_BH_PER_BODY))
    return nb_fail(NB_E_ARG, "nb_bh_forces_noalloc: only NB_F_BH_AXIS_BOX and NB_F_BH_PER_BODY apply");
  int dev = 0;
  NB_HIP(hipGetDevice(&dev));
  char* base = nullptr;
  const int rc = scratch_address(dev, &base);
  if (rc != NB_OK) return rc;

  // A context on the host stack whose arrays are views: the caller's bodies and
  // forces, and the scratch laid out exactly as nb_bh_alloc lays out its block.
  nb_ctx c;
  c.cfg.n = n;
  c.cfg.eps2 = eps2;
  c.cfg.G = G;
  c.cfg.theta = theta == 0.f ? 0.5f : theta;
  c.cfg.solver = NB_SOLVER_BH;
  c.cfg.leaf_size = leaf_size == 0 ? 8 : leaf_size;
  c.cfg.flags = flags;
  c.n = (int)n;
  c.n_pad = (int)n;
  c.device = dev;
  c.stream = static_cast<hipStream_t>(stream);
  c.theta2 = c.cfg.theta * c.cfg.theta;
  c.G = G == 0.f ? 1.0 : (double)G;
  c.guard = !(eps2 >= 1e-20f);
  c.pm = reinterpret_cast<float4*>(const_cast<float*>(pm4));
  c.force = force3;
  nb_bh_carve(&c, base, kBhBufCount);
  c.flags = status ? status : reinterpret_cast<int*>(base + nb_bh_layout_for(n).off[kBhFlags]);

  // What must start at zero: the status word, and the zero-mass bodies after
  // the sorted ones (read by the batched leaf loop). Fill kernels, no memset.
  NB_HIP(nb_launch_fill_i32(c.flags, 0, 1, c.stream));
  NB_HIP(nb_launch_fill_i32(reinterpret_cast<int*>(c.bh.pms + n), 0, 4 * kSpare, c.stream));
  NB_HIP(nb_bh_build(&c));
  NB_HIP(nb_bh_traverse(&c, false));
  c.bh = nb_bh_state{};  // views only: nothing to free
  return NB_OK;
  NB_CATCH("nb_bh_forces_noalloc")
}
