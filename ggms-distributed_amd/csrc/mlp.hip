// Learned splitter weights (placeholder until the MLP kernels land).
#include <hip/hip_runtime.h>
namespace ggms {
const float* learned_weights_device(int device) { (void)device; return nullptr; }
}
