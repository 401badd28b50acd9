// sum(a, dim), cumsum(a, dim), cumprod, cummin and cummax on device tables:
// the C ABI over nt2/hip/dimwise.hpp (NT2's order, or for sums the parallel
// fast order).
#include <hip/hip_runtime.h>

#include <cstddef>

#include "common.hpp"
#include "nt2/hip/dimwise.hpp"
#include "nt2h.h"

namespace nt2h {

namespace dw = nt2::hip::dw;

template <class T>
static int run_dim(const char* name, bool scan, const T* x, const size_t* ext, int dim, T* out, int policy,
                   void* stream) {
  if (!ext || dim < 1) return fail(NT2H_EINVAL, name);
  if (policy != NT2H_EXACT && policy != NT2H_FAST) return fail(NT2H_EINVAL, name);
  size_t e[4];
  size_t numel = 1;
  for (int k = 0; k < 4; ++k) {
    e[k] = ext[k];
    numel *= e[k];
  }
  const dw::lines L = dw::lines_of(e, dim);
  const size_t nout = scan ? numel : L.inner * L.outer;
  if (nout == 0) return NT2H_OK;  // nothing to write (x and out may be null)
  if (!out || (numel && !x)) return fail(NT2H_EINVAL, name);
  const bool fast = policy == NT2H_FAST;
  hipStream_t s = as_stream(stream);
  double* scratch = nullptr;
  const size_t need = dw::scratch_doubles(L, fast, scan);
  if (need) {
    Scratch sc;
    if (int rc = get_scratch(need, &sc, s)) return rc;
    scratch = sc.partial;
  }
  const dw::ptr_src<T> src{x};
  if (scan) dw::launch_scan<T>(src, L, fast, scratch, out, s);
  else dw::launch_sum<T>(src, L, fast, scratch, out, s);
  NT2H_CHECK_LAUNCH(name);
  return NT2H_OK;
}

// cumprod, cummin, cummax: NT2's sequential order under both policies (the
// chunked fast form is a sum's: products and NT2's min/max have no exact
// regrouping)
template <class T, class Op>
static int run_cum(const char* name, const T* x, const size_t* ext, int dim, T* out, int policy, void* stream) {
  if (!ext || dim < 1) return fail(NT2H_EINVAL, name);
  if (policy != NT2H_EXACT && policy != NT2H_FAST) return fail(NT2H_EINVAL, name);
  size_t e[4];
  size_t numel = 1;
  for (int k = 0; k < 4; ++k) {
    e[k] = ext[k];
    numel *= e[k];
  }
  if (numel == 0) return NT2H_OK;
  if (!out || !x) return fail(NT2H_EINVAL, name);
  dw::launch_scan_ordered<T, Op>(dw::ptr_src<T>{x}, dw::lines_of(e, dim), out, as_stream(stream));
  NT2H_CHECK_LAUNCH(name);
  return NT2H_OK;
}

}  // namespace nt2h

using namespace nt2h;

extern "C" {

int nt2h_sum_dim_f32(const float* x, const size_t* ext, int dim, float* out, int policy, void* stream) {
  return run_dim<float>("nt2h_sum_dim_f32: bad argument", false, x, ext, dim, out, policy, stream);
}
int nt2h_sum_dim_f64(const double* x, const size_t* ext, int dim, double* out, int policy, void* stream) {
  return run_dim<double>("nt2h_sum_dim_f64: bad argument", false, x, ext, dim, out, policy, stream);
}
int nt2h_cumsum_dim_f32(const float* x, const size_t* ext, int dim, float* out, int policy, void* stream) {
  return run_dim<float>("nt2h_cumsum_dim_f32: bad argument", true, x, ext, dim, out, policy, stream);
}
int nt2h_cumsum_dim_f64(const double* x, const size_t* ext, int dim, double* out, int policy, void* stream) {
  return run_dim<double>("nt2h_cumsum_dim_f64: bad argument", true, x, ext, dim, out, policy, stream);
}
int nt2h_cumprod_dim_f32(const float* x, const size_t* ext, int dim, float* out, int policy, void* stream) {
  return run_cum<float, dw::op_prod>("nt2h_cumprod_dim_f32: bad argument", x, ext, dim, out, policy, stream);
}
int nt2h_cumprod_dim_f64(const double* x, const size_t* ext, int dim, double* out, int policy, void* stream) {
  return run_cum<double, dw::op_prod>("nt2h_cumprod_dim_f64: bad argument", x, ext, dim, out, policy, stream);
}
int nt2h_cummin_dim_f32(const float* x, const size_t* ext, int dim, float* out, int policy, void* stream) {
  return run_cum<float, dw::op_min>("nt2h_cummin_dim_f32: bad argument", x, ext, dim, out, policy, stream);
}
int nt2h_cummin_dim_f64(const double* x, const size_t* ext, int dim, double* out, int policy, void* stream) {
  return run_cum<double, dw::op_min>("nt2h_cummin_dim_f64: bad argument", x, ext, dim, out, policy, stream);
}
int nt2h_cummax_dim_f32(const float* x, const size_t* ext, int dim, float* out, int policy, void* stream) {
  return run_cum<float, dw::op_max>("nt2h_cummax_dim_f32: bad argument", x, ext, dim, out, policy, stream);
}
int nt2h_cummax_dim_f64(const double* x, const size_t* ext, int dim, double* out, int policy, void* stream) {
  return run_cum<double, dw::op_max>("nt2h_cummax_dim_f64: bad argument", x, ext, dim, out, policy, stream);
}

}  // extern "C"
