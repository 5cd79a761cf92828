// The entry point that allocates nothing: vfvm_residual_norm on device memory, with the
// residual it reduces kept in a file-scope device array instead of a handle or call
// buffer.  For callers that cannot call hipMalloc (code under HIP graph capture, tight
// serving loops, harnesses that forbid allocation).
//
// The array is `scratch` below.  In host code its name is only a stand-in for the
// device symbol, so the host never passes it to a kernel or a copy: noalloc_setup looks
// up its device address once per device (with that device current) and every use goes
// through that address.  Each GPU holds its own copy of the array until the process
// ends, and every call on a GPU shares it: calls that overlap on one GPU race.
#include <mutex>

#include "vfvm_sys.h"

namespace {

// residual of up to VFVM_NOALLOC_MAX_DOF dofs (32 MiB), after a head of 32 doubles
// (256 bytes, so the residual keeps the array's 256-byte alignment) whose first word
// takes the assembly's NaN flag
constexpr int64_t kHead = 32;
static __device__ __align__(256) double scratch[kHead + VFVM_NOALLOC_MAX_DOF];

constexpr int kMaxDevices = 64;
std::mutex g_mtx;
double *g_scratch[kMaxDevices] = {};

}  // namespace

namespace vfvm {

// Looks up this device's copy of `scratch`; `device` must be current.  Called by
// vfvm_ctx_create, so every device a context is created on is set up once.
void noalloc_setup(int device) {
  if (device < 0 || device >= kMaxDevices) throw Error(VFVM_ERR_ARG, "device index beyond the no-alloc table");
  std::lock_guard<std::mutex> lk(g_mtx);
  if (g_scratch[device]) return;
  void *p = nullptr;
  hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !p)
    throw Error(VFVM_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
  g_scratch[device] = static_cast<double *>(p);
}

}  // namespace vfvm

extern "C" int vfvm_residual_norm_device_noalloc(vfvm_sys *s, const double *dU, const double *dUold, double time,
                                                 double tstep, double lambda, double *dnorm) {
  (void)time;
  if (!s || !dU || !dnorm) return VFVM_ERR_ARG;
  vfvm_ctx *ctx = s->ctx;
  try {
    if (s->ndof > VFVM_NOALLOC_MAX_DOF)
      throw Error(VFVM_ERR_ARG, "vfvm_residual_norm_device_noalloc: " + std::to_string(s->ndof) +
                                    " dofs exceed the no-alloc maximum of " + std::to_string(VFVM_NOALLOC_MAX_DOF));
    double *base = nullptr;
    {
      std::lock_guard<std::mutex> lk(g_mtx);
      if (ctx->device >= 0 && ctx->device < kMaxDevices) base = g_scratch[ctx->device];
    }
    if (!base) throw Error(VFVM_ERR_HIP, "no-alloc scratch of this device not set up");
    hipStream_t st = ctx->stream;
    double *F = base + kHead;
    AsmArgs a = asm_args(*s, dU, dUold ? dUold : dU, F, nullptr, tstep, lambda);
    a.flags = reinterpret_cast<int32_t *>(base);  // the residual-only assembly never sets it
    run_assembly(*s, a, s->res_group, st);
    // the norm kernel max-reduces the bit patterns of non-negative doubles into dnorm
    launch_fill_u64(1, reinterpret_cast<unsigned long long *>(dnorm), st);
    launch_penalty_free_norm(s->nnodes, s->nspec, F, s->d_bn_ptr.p, s->d_bn_reg.p, nullptr, s->d_inst.p, dnorm, 1,
                             st);
    return VFVM_OK;
  } catch (const Error &e) {
    ctx->err = e.what();
    return e.code;
  } catch (const std::exception &e) {
    ctx->err = e.what();
    return VFVM_ERR_HIP;
  }
}
