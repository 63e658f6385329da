// aclTensor stand-in: a plain descriptor of a device buffer (view shape, strides, offset, dtype).
// Keeps the names and argument meaning of CANN's aclCreateTensor / aclDestroyTensor /
// aclGetDataType / aclGetViewShape, which are all the reference's FFT path uses
// (core/fft/fft_api.cpp:1016-1025, 1047-1167; example/A2/FFT/1d/example_acl_fft_c2c_1d.cpp:74-83).
#include <cstring>
#include <new>

#include "aclnn/acl_meta.h"
#include "tensor.h"

extern "C" {

aclTensor *aclCreateTensor(const int64_t *viewDims, uint64_t viewDimsNum, aclDataType dataType,
                           const int64_t *stride, int64_t offset, aclFormat format, const int64_t *storageDims,
                           uint64_t storageDimsNum, void *tensorData)
{
    (void)storageDims;
    (void)storageDimsNum;
    if (viewDimsNum > sip::kMaxTensorDims || (viewDimsNum > 0 && viewDims == nullptr)) return nullptr;
    aclTensor *t = new (std::nothrow) aclTensor;
    if (!t) return nullptr;
    std::memset(t, 0, sizeof(*t));
    t->dtype = (int32_t)dataType;
    t->format = (int32_t)format;
    t->ndim = (int32_t)viewDimsNum;
    t->offset = offset;
    t->data = tensorData;
    // contiguous row-major strides when none are given
    int64_t acc = 1;
    for (int64_t i = (int64_t)viewDimsNum - 1; i >= 0; i--) {
        t->dims[i] = viewDims[i];
        t->strides[i] = stride ? stride[i] : acc;
        acc *= viewDims[i];
    }
    return t;
}

int aclDestroyTensor(const aclTensor *tensor)
{
    delete tensor;
    return 0;
}

int aclGetDataType(const aclTensor *tensor, aclDataType *dataType)
{
    if (!tensor || !dataType) return 1;
    *dataType = (aclDataType)tensor->dtype;
    return 0;
}

int aclGetViewShape(const aclTensor *tensor, int64_t **viewDims, uint64_t *viewDimsNum)
{
    if (!tensor || !viewDims || !viewDimsNum) return 1;
    int64_t *d = new (std::nothrow) int64_t[tensor->ndim > 0 ? tensor->ndim : 1];
    if (!d) return 1;
    for (int i = 0; i < tensor->ndim; i++) d[i] = tensor->dims[i];
    *viewDims = d;
    *viewDimsNum = (uint64_t)tensor->ndim;
    return 0;
}

void aclFreeViewShape(int64_t *viewDims) { delete[] viewDims; }

void *aclGetTensorData(const aclTensor *tensor)
{
    if (!tensor || !tensor->data) return nullptr;
    return static_cast<char *>(tensor->data) + tensor->offset * sip::dtype_size(tensor->dtype);
}

}  // extern "C"
