// Layout of the aclTensor stand-in (opaque to callers; see include/aclnn/acl_meta.h).
#pragma once
#include <cstdint>

namespace sip {
constexpr int kMaxTensorDims = 8;

inline int64_t dtype_size(int32_t dtype)
{
    switch (dtype) {
        case 0: return 4;    // ACL_FLOAT
        case 1: return 2;    // ACL_FLOAT16
        case 11: return 8;   // ACL_DOUBLE
        case 16: return 8;   // ACL_COMPLEX64
        case 17: return 16;  // ACL_COMPLEX128
        case 27: return 2;   // ACL_BF16
        case 33: return 4;   // ACL_COMPLEX32
        case 9: case 10: return 8;
        case 3: case 8: return 4;
        case 6: case 7: return 2;
        default: return 1;
    }
}
}  // namespace sip

struct aclTensor {
    int32_t dtype;
    int32_t format;
    int32_t ndim;
    int32_t pad_;
    int64_t offset;  // in elements
    int64_t dims[sip::kMaxTensorDims];
    int64_t strides[sip::kMaxTensorDims];  // in elements
    void *data;
};
