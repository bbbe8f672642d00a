#pragma once
// Drop-in name of the reference's synchronous transform class
// (include/ntt/cuda_ntt.hpp:34-79): callers written against
// ntt::cuda::CudaNttTransform get the gfx950 implementation.
#include "ntt/hip_ntt.hpp"

namespace ntt::cuda {
using CudaNttTransform = ::ntt::hip::HipNttTransform;
}
