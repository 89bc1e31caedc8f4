// Context lifetime, basis upload and error plumbing of libgigpu.
#include "gi_context.h"
#include <cstring>
#include <algorithm>

namespace gi {

static thread_local std::string g_last_error;
thread_local hipStream_t g_alloc_stream = nullptr;
void set_error(const std::string &m) { g_last_error = m; }
const char *last_error() { return g_last_error.c_str(); }

Context::Context(int dev) : device(dev) {
  int n = 0;
  GI_HIP(hipGetDeviceCount(&n));
  if (dev < 0 || dev >= n) throw Error{1, "invalid HIP device index " + std::to_string(dev)};
  GI_HIP(hipSetDevice(dev));
  GI_HIP(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
  g_alloc_stream = stream;
  // keep freed stream-ordered memory in the device pool (reused by the next
  // build instead of going back to the driver at every synchronization)
  hipMemPool_t pool = nullptr;
  if (hipDeviceGetDefaultMemPool(&pool, dev) == hipSuccess && pool) {
    uint64_t keep = UINT64_MAX;
    (void)hipMemPoolSetAttribute(pool, hipMemPoolAttrReleaseThreshold, &keep);
  }
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
  md_prim.release();
  md_et.release();
  md_blob.release();
  md_rec.release();
  chol.inv_l.release();
  chol.rkpq.release();
  chol.mo_adj_off.release();
  chol.mo_adj_col.release();
  chol.mo_adj_p.release();
  release_chol_work(*this);
  release_thc_work(*this);
  for (int i = 0; i < 2; ++i) {
    spl_slices[i].release();
    spl_exp[i].release();
  }
  // every buffer is released (stream-ordered frees on `stream`) before the
  // stream goes away
  if (stream) (void)hipStreamSynchronize(stream);
  for (int i = 0; i < ERI_MAX_STREAMS - 1; ++i) {
    if (eri_side[i]) (void)hipStreamDestroy(eri_side[i]);
    if (eri_join[i]) (void)hipEventDestroy(eri_join[i]);
  }
  if (eri_fork) (void)hipEventDestroy(eri_fork);
  if (stream) (void)hipStreamDestroy(stream);
  if (md_stage) (void)hipHostFree(md_stage);
  if (md_stage_ev) (void)hipEventDestroy(md_stage_ev);
  if (wait_ev) (void)hipEventDestroy(wait_ev);
  if (xfer) (void)hipHostFree(xfer);
}

void Context::wait() {
  if (!blocking_waits) {
    GI_HIP(hipStreamSynchronize(stream));
  } else {
    if (!wait_ev) GI_HIP(hipEventCreateWithFlags(&wait_ev, hipEventBlockingSync | hipEventDisableTiming));
    GI_HIP(hipEventRecord(wait_ev, stream));
    GI_HIP(hipEventSynchronize(wait_ev));
  }
  xfer_used = 0;   // every staged copy has completed
}

unsigned char *Context::xfer_slot(size_t bytes) {
  const size_t need = (bytes + 255) & ~(size_t)255;
  if (xfer_used + need > xfer_cap) {
    if (xfer_used > 0) wait();   // copies out of the arena still in flight
    if (need > xfer_cap) {
      if (xfer) GI_HIP(hipHostFree(xfer));
      xfer = nullptr;
      const size_t cap = std::max(need, std::max<size_t>(1 << 20, 2 * xfer_cap));
      GI_HIP(hipHostMalloc((void **)&xfer, cap, hipHostMallocDefault));
      xfer_cap = cap;
    }
  }
  unsigned char *p = xfer + xfer_used;
  xfer_used += need;
  return p;
}

void Context::h2d(void *dst, const void *src, size_t bytes) {
  if (!bytes) return;
  unsigned char *p = xfer_slot(bytes);
  memcpy(p, src, bytes);
  GI_HIP(hipMemcpyAsync(dst, p, bytes, hipMemcpyHostToDevice, stream));
}

void Context::read(void *dst, const void *src, size_t bytes) {
  if (!bytes) return;
  unsigned char *p = xfer_slot(bytes);
  GI_HIP(hipMemcpyAsync(p, src, bytes, hipMemcpyDeviceToHost, stream));
  wait();
  memcpy(dst, p, bytes);
}

unsigned char *Context::stage_host(size_t n) {
  if (!md_stage_ev) GI_HIP(hipEventCreateWithFlags(&md_stage_ev, hipEventDisableTiming));
  else GI_HIP(hipEventSynchronize(md_stage_ev));
  if (n <= md_stage_cap && md_stage) return md_stage;
  if (md_stage) {
    GI_HIP(hipHostFree(md_stage));
    md_stage = nullptr;
  }
  const size_t cap = std::max(n, md_stage_cap + md_stage_cap / 2);
  GI_HIP(hipHostMalloc((void **)&md_stage, cap, hipHostMallocDefault));
  md_stage_cap = cap;
  return md_stage;
}

}  // namespace gi
