// Launch descriptors of the persistent multi-length Stockham kernel (batching form).
//
// One launch transforms the rows of many length groups. A group is every row of one length L; its
// radix sequence is runtime data (chosen per group on the host, by the tuner when asked), so one
// compiled kernel serves every 2,3,5-smooth length up to kPersistMaxL. The rows of a group are cut
// into tiles of F transforms (T threads each, F * T <= kPersistThreads); the workgroups of the
// launch walk the tile list with a grid stride, largest groups first.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "stockham_args.h"

namespace sip {

constexpr int kPersistThreads = 256;  // threads per workgroup
constexpr int kPersistMaxE = 24;      // complex values one thread holds in one pass
constexpr int kPersistMaxPass = 8;
constexpr int kPersistMaxL = 4096;

// Radices the kernel has butterflies for (products of 2, 3, 5 up to 20).
constexpr int kPersistRadices[] = {2, 3, 4, 5, 6, 8, 9, 10, 12, 15, 16, 20};

// Butterflies of radix R one thread can hold in one pass.
constexpr int persist_max_nb(int R) { return kPersistMaxE / R > 0 ? kPersistMaxE / R : 1; }

struct PersistGroup {
    int32_t L;                        // transform length
    int32_t T;                        // threads per transform
    int32_t F;                        // transforms per tile
    int32_t LP;                       // LDS pitch (floats) of one transform in a plane
    int32_t npass;
    int32_t radix[kPersistMaxPass];
    int32_t ns[kPersistMaxPass];      // product of the radices before each pass
    int32_t nbf[kPersistMaxPass];     // butterflies per transform in each pass (L / radix)
    int32_t tw_off[kPersistMaxPass];  // offset (complex entries) of each pass's twiddle rows
    FastDiv div_ns[kPersistMaxPass];
    FastDiv div_T;
    const float2 *tw;                 // pass twiddles, the layout of the leaf kernels' tables
    const int64_t *in_rows;           // element offset of each row from PersistArgs::in
    const int64_t *out_rows;          // element offset of each row from PersistArgs::out
    int64_t nrows;
    int64_t tile0;                    // index of the group's first tile in the launch
};

struct PersistArgs {
    const float2 *in;
    float2 *out;
    const PersistGroup *groups;
    const int32_t *tile_group;        // group of each tile
    int64_t ntiles;
};

}  // namespace sip
