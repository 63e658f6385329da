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

// Device image of one persistent launch: groups | tile -> group | input rows | output rows, with
// the launch arguments and shape. Built on the host (persist_build_image), given its device address
// (persist_place, which writes the image's own pointers), then copied there by the caller.
struct PersistImage {
    std::vector<char> bytes;
    std::vector<PersistGroup> groups;        // host view (rows pointers are element indices)
    std::vector<std::vector<int>> radices;
    std::vector<TablePtr> tables;            // pass twiddles, held while the image is in use
    size_t gbytes = 0, tbytes = 0, rbytes = 0;
    PersistArgs args{};
    size_t lds = 0;
    unsigned grid = 1;
    int sign = -1;
};

// in_base / out_base: device addresses the row offsets (complex elements) are relative to.
void persist_build_image(const std::vector<PersistSpec> &specs, const void *in_base, void *out_base, int sign,
                         PersistImage &img);
void persist_place(PersistImage &img, void *dev_base);
void persist_launch(const PersistImage &img, hipStream_t stream);

// A persistent launch that owns its device image (batching form).
class PersistLaunch {
public:
    PersistLaunch(const std::vector<PersistSpec> &specs, const void *in_base, void *out_base, int sign);
    ~PersistLaunch();
    PersistLaunch(const PersistLaunch &) = delete;
    PersistLaunch &operator=(const PersistLaunch &) = delete;

    void launch(hipStream_t stream) const;
    int64_t tiles() const { return img_.args.ntiles; }
    size_t group_count() const { return img_.groups.size(); }
    // JSON fragment: the groups with their radices and shapes.
    std::string describe() const;

private:
    PersistImage img_;
    void *dev_ = nullptr;
};

}  // namespace sip
