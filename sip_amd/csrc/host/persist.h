// Host side of the persistent multi-length Stockham kernel (batching form): per-group radix
// sequence and thread shape, the tile list, and the launch.
#pragma once
#include <memory>
#include <string>
#include <vector>

#include "../kernels/persist_args.h"
#include "plan.h"

namespace sip {

extern const void *const g_persist_kernel[2];  // forward, inverse (persist.hip)

// One length group run by the persistent kernel: rows at element offsets from the launch's bases.
struct PersistSpec {
    int64_t n = 0;
    std::vector<int64_t> in_rows, out_rows;
    std::vector<int> radices;  // empty: the default sequence
};

// The default radix sequence of length n for the persistent kernel (the leaf kernel's sequence),
// or empty when n is not 2,3,5-smooth, exceeds kPersistMaxL or has no sequence within the limits.
std::vector<int> persist_default_radices(int64_t n);
// Radix sequences the tuner tries for n (the default first), at most `max_count`.
std::vector<std::vector<int>> persist_candidate_radices(int64_t n, size_t max_count);
// Thread shape for a sequence: T threads per transform, F transforms per tile. False when the
// sequence does not fit the kernel (a radix it lacks, too many passes or values per thread).
bool persist_shape(int64_t n, const std::vector<int> &radices, int &T, int &F);

class PersistLaunch {
public:
    // in_base / out_base: device addresses the row offsets (complex elements) are relative to.
    PersistLaunch(const std::vector<PersistSpec> &specs, const void *in_base, void *out_base, int sign);
    ~PersistLaunch();
    PersistLaunch(const PersistLaunch &) = delete;
    PersistLaunch &operator=(const PersistLaunch &) = delete;

    void launch(hipStream_t stream) const;
    int64_t tiles() const { return args_.ntiles; }
    size_t group_count() const { return host_.size(); }
    unsigned grid() const { return grid_; }
    size_t lds() const { return lds_; }
    // JSON fragment: the groups with their radices and shapes.
    std::string describe() const;

private:
    std::vector<PersistGroup> host_;
    std::vector<std::vector<int>> radices_;
    std::vector<TablePtr> tables_;
    void *dev_ = nullptr;
    PersistArgs args_{};
    size_t lds_ = 0;
    unsigned grid_ = 1;
    int sign_ = -1;
};

}  // namespace sip
