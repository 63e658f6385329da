/*
 * Status codes returned by every asdFft* entry point.
 * Same names and values as AscendSiPBoost's core/utils/include/utils/aspb_status.h:16-27
 * (Copyright (c) 2025 Huawei Technologies Co., Ltd.; CANN Open Software License Agreement
 * Version 2.0), reproduced so that callers see the reference's codes.
 */
#ifndef ASPB_STATUS_H
#define ASPB_STATUS_H

#include <stdint.h>

#ifdef __cplusplus
namespace AsdSip {
using AspbStatus = int32_t;

enum ErrorType : int {
    ACL_SUCCESS = 0,
    ACL_ERROR_INVALID_PARAM = 100000,
    ACL_ERROR_OP_INPUT_NOT_MATCH = 100021,
    ACL_ERROR_OP_OUTPUT_NOT_MATCH = 100022,
    ACL_ERROR_UNSUPPORTED_DATA_TYPE = 100026,
    ACL_ERROR_FORMAT_NOT_MATCH = 100027,
    ACL_ERROR_API_NOT_SUPPORT = 200001,
    ACL_ERROR_INTERNAL_ERROR = 500000
};
}  // namespace AsdSip
#endif

#endif
