// Error reporting, plan lifetime and small helpers of libvamx_hip.so.
#include <atomic>

#include "vamx_internal.h"

namespace vamx {

static thread_local std::string g_last_error;

int set_error(int code, const std::string& msg) {
  g_last_error = msg;
  return code;
}

int hip_check(hipError_t e, const char* what) {
  std::string m = std::string(what) + ": " + hipGetErrorString(e);
  return set_error(VAMX_ERR_HIP, m);
}

bool check_plan(vamx_plan* p, int kind) { return p != nullptr && p->kind == kind; }

unsigned long long next_plan_serial() {
  static std::atomic<unsigned long long> next{1};
  return next.fetch_add(1);
}

unsigned blocks_for(long long total, int block) {
  long long b = (total + block - 1) / block;
  return (unsigned)(b < 1 ? 1 : b);
}

}  // namespace vamx

extern "C" void vamx_destroy(vamx_plan* p) {
  if (!p) return;
  if (p->kind == vamx::PLAN_PARALLEL) vamx::batch_forget_plan(p);
  int cur = 0;
  (void)hipGetDevice(&cur);
  if (cur != p->device) (void)hipSetDevice(p->device);
  (void)hipDeviceSynchronize();
  if (p->d_mats) (void)hipFree(p->d_mats);
  if (p->d_cs) (void)hipFree(p->d_cs);
  if (p->d_outside) (void)hipFree(p->d_outside);
  if (p->d_minmax_part) (void)hipFree(p->d_minmax_part);
  if (p->d_minmax) (void)hipFree(p->d_minmax);
  if (p->d_occ) (void)hipFree(p->d_occ);
  if (p->d_occ_active) (void)hipFree(p->d_occ_active);
  if (p->d_units) (void)hipFree(p->d_units);
  if (p->d_rt_tables) (void)hipFree(p->d_rt_tables);
  if (p->d_accum) (void)hipFree(p->d_accum);
  if (p->d_gmax) (void)hipFree(p->d_gmax);
  if (p->d_stats) (void)hipFree(p->d_stats);
  if (cur != p->device) (void)hipSetDevice(cur);
  delete p;
}

extern "C" const char* vamx_last_error(void) { return vamx::g_last_error.c_str(); }

extern "C" int vamx_version(void) { return 1; }
