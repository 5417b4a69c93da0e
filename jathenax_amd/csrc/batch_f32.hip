// Batched ensemble engine instantiated for float state.
#include "jx_batch_host.hpp"

jx::BatchBase* jx_make_batch_f32(int device, int max_inst, long long pool_bytes, std::string* err) {
  auto* b = new jx::Batch<float>(device, max_inst, pool_bytes);
  if (b->init() != JX_OK) {
    if (err) *err = b->err;
    delete b;
    return nullptr;
  }
  return b;
}
