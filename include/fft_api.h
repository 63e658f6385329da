/*
 * asdFft* public API of the MI355X FFT library (C++ view).
 *
 * The declarations reproduce the interface of AscendSiPBoost's include/fft_api.h
 * (Copyright (c) 2025 Huawei Technologies Co., Ltd.; CANN Open Software License Agreement
 * Version 2.0) so that this library is a drop-in for it; the comments and everything behind the
 * declarations are this library's own.
 *
 * Same entry points, argument meaning, enum values and status codes as the reference's
 * include/fft_api.h:15-101 (declared there inside extern "C" { namespace AsdSip { ... } } with
 * C++ references and one default argument). The symbols are unmangled C names; a reference is
 * passed as a pointer, so include/sip_fft.h declares the very same symbols for plain-C and FFI
 * callers with pointer arguments.
 *
 * Semantics kept from the reference (core/fft/fft_api.cpp): forward uses e^{-2 pi i kn/N}; the
 * inverse is unnormalised; R2C writes N/2+1 bins of the last dimension and C2R reads them;
 * 1D shapes are (batch, N) or, with ASCEND_FFT_VERTICAL, (N, batch); 2D (..., X, Y) and 3D
 * (..., X, Y, Z) are row-major with the last dimension fastest; 0 < N <= 2^27. Exec only enqueues
 * work on the handle's stream (a hipStream_t; NULL = the default stream) and asdFftSynchronize
 * waits for it. The input is never modified.
 */
#ifndef ASDSIP_FFT_API_H
#define ASDSIP_FFT_API_H

#include "utils/aspb_status.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct aclTensor aclTensor;

namespace AsdSip {
typedef void *asdFftHandle;

enum asdFftStatus { SUCCESS = 0, FAILED = 1 };

enum asdFftType {
    ASCEND_FFT_C2C = 0x10,
    ASCEND_FFT_C2R = 0x11,
    ASCEND_FFT_R2C = 0x12,
    ASCEND_STFT_C2C = 0x20,
    ASCEND_STFT_C2R = 0x21,
    ASCEND_STFT_R2C = 0x22,
    ASCEND_FFT_C2C_SEP = 0x30
};

enum asdFftDirection {
    ASCEND_FFT_FORWARD = 0x10,
    ASCEND_FFT_INVERSE = 0x11,
};

enum asdFft1dDimType {
    ASCEND_FFT_HORIZONTAL = 0x10,
    ASCEND_FFT_VERTICAL = 0x11,
};

// Creates an opaque handle (fft_api.h:47-48).
AsdSip::AspbStatus asdFftCreate(asdFftHandle &handle);

// Binds a hipStream_t to the handle (fft_api.h:50-51).
AsdSip::AspbStatus asdFftSetStream(asdFftHandle handle, void *stream);

// 1D plan (fft_api.h:53-57).
AsdSip::AspbStatus asdFftMakePlan1D(asdFftHandle handle, int64_t fftSize, asdFftType fftType,
                                    asdFftDirection direction, int64_t batchSize,
                                    asdFft1dDimType dimType = asdFft1dDimType::ASCEND_FFT_HORIZONTAL);

// ISTFT plan (fft_api.h:59-62): the reference's parameter checks and status codes; the inverse DFTs
// and the windowed overlap-add run on the GPU.
AspbStatus asdFftIstftMakePlan(asdFftHandle handle, const aclTensor *input, const int64_t nFft,
                               const int64_t hopLengthOpt, const int64_t winLengthOpt, const bool center = true,
                               const bool normalized = false, const bool onesidedOpt = false,
                               int64_t lengthOpt = 0, const bool returnComplex = true);

// C2C execution (fft_api.h:64-65).
AsdSip::AspbStatus asdFftExecC2C(asdFftHandle handle, const aclTensor *input, const aclTensor *output);

// Split-plane C2C execution (fft_api.h:67-69).
AsdSip::AspbStatus asdFftExecC2CSeparated(asdFftHandle handle, const aclTensor *inputReal, const aclTensor *inputImag,
    const aclTensor *outputReal, const aclTensor *outputImag);

// C2R execution (fft_api.h:72-73).
AsdSip::AspbStatus asdFftExecC2R(asdFftHandle handle, const aclTensor *input, const aclTensor *output);

// R2C execution (fft_api.h:75-76).
AsdSip::AspbStatus asdFftExecR2C(asdFftHandle handle, const aclTensor *input, const aclTensor *output);

// ISTFT execution (fft_api.h:78-79) of a plan made by asdFftIstftMakePlan.
AspbStatus asdFftExecIstft(asdFftHandle handle, const aclTensor *input,
                           const aclTensor *windowOpt, const aclTensor *output);

// Destroys a handle and everything it owns (fft_api.h:81-82).
AsdSip::AspbStatus asdFftDestroy(asdFftHandle handle);

// 2D plan (fft_api.h:84-85).
AsdSip::AspbStatus asdFftMakePlan2D(asdFftHandle handle, int64_t fftSizeX, int64_t fftSizeY, asdFftType fftType,
                                    asdFftDirection direction, int32_t batchSize);

// 3D plan (fft_api.h:87-88).
AspbStatus asdFftMakePlan3D(asdFftHandle handle, int64_t fftSizeX, int64_t fftSizeY, int64_t fftSizeZ,
    asdFftType fftType, asdFftDirection direction, int32_t batchSize);

// Workspace the caller must allocate (device memory) before exec (fft_api.h:90).
AsdSip::AspbStatus asdFftGetWorkspaceSize(asdFftHandle handle, size_t &workspaceSize);

AsdSip::AspbStatus asdFftSetWorkspace(asdFftHandle handle, void *workspace);

// Waits for the handle's stream (fft_api.h:94).
AsdSip::AspbStatus asdFftSynchronize(asdFftHandle handle);

AsdSip::AspbStatus asdFftGetType(asdFftHandle handle, asdFftType &fftType);
}

#ifdef __cplusplus
}
#endif

#endif
