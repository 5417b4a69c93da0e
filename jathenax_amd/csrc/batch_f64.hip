// Batched ensemble engine instantiated for double state.
#include "jx_batch_host.hpp"

jx::BatchBase* jx_make_batch_f64(int device, int max_inst, long long pool_bytes, std::string* err) {
  auto* b = new jx::Batch<double>(device, max_inst, pool_bytes);
  if (b->init() != JX_OK) {
    if (err) *err = b->err;
    delete b;
    return nullptr;
  }
  return b;
}
