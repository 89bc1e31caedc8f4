// Context lifetime, basis upload and error plumbing of libgigpu.
#include "gi_context.h"
#include <cstring>

namespace gi {

static thread_local std::string g_last_error;
void set_error(const std::string &m) { g_last_error = m; }
const char *last_error() { return g_last_error.c_str(); }

Context::Context(int dev) : device(dev) {
  int n = 0;
  GI_HIP(hipGetDeviceCount(&n));
  if (dev < 0 || dev >= n) throw Error{1, "invalid HIP device index " + std::to_string(dev)};
  GI_HIP(hipSetDevice(dev));
  GI_HIP(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
  if (rocblas_create_handle(&blas) != rocblas_status_success) throw Error{2, "rocblas_create_handle failed"};
  rocblas_set_stream(blas, stream);
  upload_constants(*this);
}

Context::~Context() {
  (void)hipSetDevice(device);
  if (stream) (void)hipStreamSynchronize(stream);
  if (blas) rocblas_destroy_handle(blas);
  dshells.release();
  dexps.release();
  dcntr.release();
  dnorms.release();
  dboys.release();
  md_hdr.release();
  md_jobs.release();
  md_prim.release();
  md_et.release();
  md_rdesc.release();
  md_cdesc.release();
  md_idx.release();
  md_side.release();
  md_rec.release();
  chol.inv_l.release();
  chol.rkpq.release();
  if (stream) (void)hipStreamDestroy(stream);
}

}  // namespace gi
