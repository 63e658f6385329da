/*
 * Minimal device-tensor descriptor standing in for CANN's aclTensor on the FFT path.
 *
 * The reference reads tensors only through aclGetDataType, aclGetViewShape and the storage
 * address (core/fft/fft_api.cpp:1016-1025, 1047-1167), and its examples build them with
 * aclCreateTensor / aclDestroyTensor (example/A2/FFT/1d/example_acl_fft_c2c_1d.cpp:74-83, 170-171).
 * These five functions keep CANN's names and argument meaning (aclnn/acl_meta.h in CANN), so a
 * caller of the reference compiles unchanged. The data pointer is a HIP device address.
 * Only contiguous row-major views are accepted by the FFT entry points.
 */
#ifndef SIP_ACLNN_ACL_META_H
#define SIP_ACLNN_ACL_META_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct aclTensor aclTensor;

/* Values as in CANN's aclDataType (ACL_FLOAT = 0 and ACL_COMPLEX64 = 16 are the ones the FFT
 * path accepts; see the comment at core/fft/fftplan/include/fftplan/fft_plan.h:26). */
typedef enum {
    ACL_DT_UNDEFINED = -1,
    ACL_FLOAT = 0,
    ACL_FLOAT16 = 1,
    ACL_INT8 = 2,
    ACL_INT32 = 3,
    ACL_UINT8 = 4,
    ACL_INT16 = 6,
    ACL_UINT16 = 7,
    ACL_UINT32 = 8,
    ACL_INT64 = 9,
    ACL_UINT64 = 10,
    ACL_DOUBLE = 11,
    ACL_BOOL = 12,
    ACL_STRING = 13,
    ACL_COMPLEX64 = 16,
    ACL_COMPLEX128 = 17,
    ACL_BF16 = 27,
    ACL_COMPLEX32 = 33
} aclDataType;

typedef enum {
    ACL_FORMAT_UNDEFINED = -1,
    ACL_FORMAT_NCHW = 0,
    ACL_FORMAT_NHWC = 1,
    ACL_FORMAT_ND = 2
} aclFormat;

/* Creates a descriptor (does not copy data). Returns NULL on invalid arguments. */
aclTensor *aclCreateTensor(const int64_t *viewDims, uint64_t viewDimsNum, aclDataType dataType,
                           const int64_t *stride, int64_t offset, aclFormat format, const int64_t *storageDims,
                           uint64_t storageDimsNum, void *tensorData);

int aclDestroyTensor(const aclTensor *tensor);

int aclGetDataType(const aclTensor *tensor, aclDataType *dataType);

/* *viewDims is allocated with new[]; the caller releases it with delete[] (C++), as in CANN.
 * C callers release it with aclFreeViewShape. */
int aclGetViewShape(const aclTensor *tensor, int64_t **viewDims, uint64_t *viewDimsNum);

void aclFreeViewShape(int64_t *viewDims);

/* Device address of element 0 of the view (storage address + offset * element size). */
void *aclGetTensorData(const aclTensor *tensor);

#ifdef __cplusplus
}
#endif

#endif
