// abi_common.hpp — shared by the extern "C" entry point files: the error-to-return-code
// wrapper and the per-call set-up of the grid, projector plan and dij / sss tables.
#pragma once
#include <vector>

#include "kukan_internal.hpp"
#include "solvers.hpp"

#define API_BEGIN \
  try {           \
    ensure_init();
#define API_END                                  \
  sync();                                        \
  return 0;                                      \
  }                                              \
  catch (const kukan::Error &e) {                \
    kukan::lib().last_error = e.what();          \
    (void)hipStreamSynchronize(kukan::lib().stream); \
    return e.code;                               \
  }                                              \
  catch (const std::exception &e) {              \
    kukan::lib().last_error = e.what();          \
    return 99;                                   \
  }

namespace kukan {
namespace abi {

constexpr size_t D = sizeof(double), I = sizeof(int), Z = 2 * sizeof(double);

inline bool sss_is_zero(const double *sss, size_t count) {
  std::vector<double> h = host_doubles(sss, count);
  for (double v : h)
    if (v != 0.0) return false;
  return true;
}

// Grid + nonlocal plan + tables for one call.
inline void setup(Ctx &c, int natom, int nprjmx, int num_spe, int num_ppcell, int num_list,
                  int nums, int key_in, int key_inps, const int *nprj, const int *indspe,
                  const int *natinf, const int *natpri, const int *naps, const Arr &vnl,
                  const Arr &lst, const double *dij, const double *sss_dev,
                  const double *sss_host) {
  build_plan(c.pl, c.g, natom, nprjmx, num_spe, num_ppcell, num_list, nums, key_in, key_inps,
             nprj, indspe, natinf, natpri, naps, vnl.d(), lst.i(), "plan");
  c.dij = dij;
  c.sss = sss_dev;
  c.q_zero = sss_host ? sss_is_zero(sss_host, (size_t)nprjmx * nprjmx * num_spe) : true;
}

}  // namespace abi
}  // namespace kukan
