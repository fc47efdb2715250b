// noalloc.hip — kukan_scf_hartree_noalloc: the Hartree CG solve of kukan_scf_hartree on a
// fixed device workspace, for callers that cannot allocate device memory (code under HIP
// graph capture, tight serving loops, harnesses that forbid hipMalloc).
//
// The workspace is the file-scope array `scratch` below: 5 grid vectors of at most
// KUKAN_HARTREE_NOALLOC_MAX_POINTS points plus the small arrays of one solve (partial sums,
// CG scalars, boundary weights, the done flag).  Each GPU the library runs on holds its own
// copy until the process ends.  Its device address is looked up once per device at set-up
// (noalloc_setup, called from the library's initialisation with that device current); host
// code only ever passes that looked-up address, never the array's name (in host code the
// name is a host-side stand-in, not a device address).
//
// Every call on one GPU shares that GPU's array and leaves the last call's contents in it;
// a solve initialises everything it reads (right-hand side, residual, direction, scalars,
// flag) with kernels before use, so stale contents never matter.  Calls that overlap on one
// GPU (two host threads, or a call while another is still running on the library stream)
// race on the array: the library stream serialises calls made through this library, other
// streams do not.
#include "kukan_hip.h"
#include "kukan_internal.hpp"
#include "solvers.hpp"

// 5 vectors of the largest grid + 1 MiB for the small arrays (partials: at most a few
// thousand doubles each)
static constexpr size_t kNoallocBytes = 5ull * 8ull * KUKAN_HARTREE_NOALLOC_MAX_POINTS + (1ull << 20);
static __device__ __align__(256) char scratch[kNoallocBytes];

namespace kukan {

namespace {
constexpr int kMaxDevices = 64;
void *g_scratch_ptr[kMaxDevices] = {nullptr};
}  // namespace

// With `device` current: the device address of this device's copy of `scratch`.
void noalloc_setup(int device) {
  if (device < 0 || device >= kMaxDevices) throw Error(1, "noalloc: device index out of range");
  if (g_scratch_ptr[device]) return;
  void *ptr = nullptr;
  // ::scratch, the array above: inside namespace kukan the bare name is the library's
  // allocator function kukan::scratch
  HIPCK(hipGetSymbolAddress(&ptr, HIP_SYMBOL(::scratch)));
  if (!ptr) throw Error(1, "noalloc: no device address for the workspace");
  g_scratch_ptr[device] = ptr;
}

}  // namespace kukan

using namespace kukan;

extern "C" int kukan_scf_hartree_noalloc(int nd, int ncpx_d, int ncpy_d, int ncpz_d, int nfh,
                                         int ndisp, int nperi, int ncgres, int ncgmin, int ncgmax,
                                         double xmax, double ymax, double zmax, double epsvh,
                                         const double *rho_aug_dense, const double *boundx,
                                         const double *boundy, const double *boundz,
                                         double *vh_dense) {
  (void)ndisp;
  try {
    ensure_init();
    noalloc_setup(lib().device);
    if (nfh != 4 && nfh != 5 && nfh != 6)
      throw Error(6, "error in scf_hartree! This subroutine is for nfh ne 4, 5 or 6.");
    if (ncpx_d <= 0 || ncpy_d <= 0 || ncpz_d <= 0)
      throw Error(1, "scf_hartree_noalloc: grid extents must be positive");
    const long n = (long)ncpx_d * ncpy_d * ncpz_d;
    if (n > (long)KUKAN_HARTREE_NOALLOC_MAX_POINTS)
      throw Error(7, "scf_hartree_noalloc: grid of " + std::to_string(n) +
                         " points exceeds the workspace limit of " +
                         std::to_string((long)KUKAN_HARTREE_NOALLOC_MAX_POINTS) + " points");
    // inputs and outputs must already be device memory: staging host arrays would allocate
    const bool need_bounds = nd == 0 && nperi < 3;
    if (!is_device_ptr(rho_aug_dense) || !is_device_ptr(vh_dense) ||
        (need_bounds && (!is_device_ptr(boundx) || !is_device_ptr(boundy) || !is_device_ptr(boundz))))
      throw Error(1, "scf_hartree_noalloc: rho_aug_dense, vh_dense (and boundx/y/z when nd = 0) "
                     "must be device pointers");
    HartreeArgs a{nd, ncpx_d, ncpy_d, ncpz_d, nfh, nperi, ncgres, ncgmin, ncgmax, xmax, ymax, zmax,
                  epsvh, rho_aug_dense, boundx, boundy, boundz, vh_dense};
    HartreeWork ws{(char *)g_scratch_ptr[lib().device], kNoallocBytes};
    const int loops = hartree(a, &ws);
    lib().params["hartree_iterations"] = loops;
    sync();
    return 0;
  } catch (const Error &e) {
    lib().last_error = e.what();
    (void)hipStreamSynchronize(lib().stream);
    return e.code;
  } catch (const std::exception &e) {
    lib().last_error = e.what();
    return 99;
  }
}
