// Persistent multi-length Stockham kernel of the batching form: one launch transforms the rows of
// every short length group of a batch (lengths up to kPersistMaxL), whatever their lengths.
//
// The reference batches a whole bulk inside one launch by looping bulks over its cores
// (ops/fft/common/include/kernel/fft_all_common.cceh:102-119, ops/fft/c2c/fft_b/fft_b/kernel/
// fft_b.cce:2318-2338); here a fixed grid of workgroups walks a tile list that mixes lengths. The
// radix sequence is runtime data of each group: every pass switches once (uniformly across the
// workgroup) to the butterfly of its radix, whose register shape is compile-time. The first pass
// reads each transform straight from HBM, the last writes straight to HBM, the passes between
// exchange through two LDS planes (re, im). A pass computes what a leaf kernel's pass of the same
// radix computes (same twiddle table layout, same butterflies).
#include <hip/hip_runtime.h>

#include "persist_args.h"
#include "stockham_device.h"

namespace sip {

namespace {

struct TileCtx {
    float *re, *im;      // LDS planes
    int lbase;           // this transform's offset in a plane
    int t;               // thread index within the transform
    bool active;         // the transform exists (row < nrows)
    const float2 *gin;   // row input
    float2 *gout;        // row output
};

// One radix-R pass of group g: inputs b + r L/R (HBM for the first pass, LDS otherwise), twiddle
// W^{(b mod NS) r}, R-point DFT, outputs (b / NS) NS R + b mod NS + r NS (HBM for the last pass).
template <int R, int S>
__device__ __forceinline__ void persist_pass(const PersistGroup &g, int p, const TileCtx &c, bool first, bool last)
{
    constexpr int MAXNB = persist_max_nb(R);
    const int T = g.T;
    const int NBF = g.nbf[p];
    const int NS = g.ns[p];
    const float2 *tw = g.tw + g.tw_off[p];
    float2 v[MAXNB][R];
#pragma unroll
    for (int q = 0; q < MAXNB; q++) {
        const int b = c.t + q * T;
        if (c.active && b < NBF) {
#pragma unroll
            for (int r = 0; r < R; r++) {
                const int j = b + r * NBF;
                if (first) {
                    v[q][r] = gload(c.gin + j);
                } else {
                    const int l = c.lbase + lds_pad(j);
                    v[q][r] = make_float2(c.re[l], c.im[l]);
                }
            }
        }
    }
    if (!first) __syncthreads();  // every read of the planes is done before they are overwritten
#pragma unroll
    for (int q = 0; q < MAXNB; q++) {
        const int b = c.t + q * T;
        if (c.active && b < NBF) {
            const int hi = (int)fdiv((uint32_t)b, g.div_ns[p]);
            const int k = b - hi * NS;
            if (NS > 1) {
                float2 w[R];
                tw_load<R, 2>(tw, k, w);
                tw_expand<R>(w);
#pragma unroll
                for (int r = 1; r < R; r++) v[q][r] = cmul(v[q][r], w[r]);
            }
            dft<R, S>(v[q]);
            const int d = hi * NS * R + k;
#pragma unroll
            for (int r = 0; r < R; r++) {
                const int j = d + r * NS;
                if (last) {
                    gstore(c.gout + j, v[q][r]);
                } else {
                    const int l = c.lbase + lds_pad(j);
                    c.re[l] = v[q][r].x;
                    c.im[l] = v[q][r].y;
                }
            }
        }
    }
    if (!last) __syncthreads();
}

template <int S>
__device__ __forceinline__ void persist_tile(const PersistArgs &a, const PersistGroup &g, int64_t tile, float *lds)
{
    const int tid = threadIdx.x;
    const int fl = (int)fdiv((uint32_t)tid, g.div_T);
    const int t = tid - fl * g.T;
    const int64_t row = (tile - g.tile0) * g.F + fl;
    TileCtx c;
    c.re = lds;
    c.im = lds + g.F * g.LP;
    c.lbase = fl * g.LP;
    c.t = t;
    c.active = fl < g.F && row < g.nrows;
    c.gin = a.in;
    c.gout = a.out;
    if (c.active) {
        c.gin = a.in + g.in_rows[row];
        c.gout = a.out + g.out_rows[row];
    }
    for (int p = 0; p < g.npass; p++) {
        const bool first = p == 0, last = p == g.npass - 1;
        switch (g.radix[p]) {
            case 2: persist_pass<2, S>(g, p, c, first, last); break;
            case 3: persist_pass<3, S>(g, p, c, first, last); break;
            case 4: persist_pass<4, S>(g, p, c, first, last); break;
            case 5: persist_pass<5, S>(g, p, c, first, last); break;
            case 6: persist_pass<6, S>(g, p, c, first, last); break;
            case 8: persist_pass<8, S>(g, p, c, first, last); break;
            case 9: persist_pass<9, S>(g, p, c, first, last); break;
            case 10: persist_pass<10, S>(g, p, c, first, last); break;
            case 12: persist_pass<12, S>(g, p, c, first, last); break;
            case 15: persist_pass<15, S>(g, p, c, first, last); break;
            case 16: persist_pass<16, S>(g, p, c, first, last); break;
            default: persist_pass<20, S>(g, p, c, first, last); break;
        }
    }
}

}  // namespace

// Grid-stride walk over the tile list; the host orders the groups largest first and checks that
// every group's radices are among kPersistRadices and its passes fit kPersistMaxE.
template <int S>
__global__ __launch_bounds__(kPersistThreads) void persist_kernel(const PersistArgs a)
{
    extern __shared__ float lds[];
    for (int64_t tile = blockIdx.x; tile < a.ntiles; tile += gridDim.x) {
        const PersistGroup &g = a.groups[a.tile_group[tile]];
        persist_tile<S>(a, g, tile, lds);
    }
}

extern const void *const g_persist_kernel[2] = {(const void *)&persist_kernel<-1>, (const void *)&persist_kernel<1>};

}  // namespace sip
