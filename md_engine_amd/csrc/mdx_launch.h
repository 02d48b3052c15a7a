// Launch layer for the per-step kernels.
//
// A per-step kernel is written as a __device__ body
//     body(blk, nblk, args...)
// where blk is the workgroup's index within the context's own grid and nblk
// that grid's size (what blockIdx.x and gridDim.x are for one context).  Two
// generic kernels run a body:
//   * k_one: one context, one launch (blk = blockIdx.x, nblk = gridDim.x);
//   * k_many: many contexts in ONE launch.  Workgroup b reads (context, blk,
//     nblk) from a block map and its context's arguments from an argument
//     table, so a batch of replicas of uneven size takes one launch per
//     kernel instead of one per replica.
// Both run the same body, so a batched replica computes bit for bit what it
// computes alone.
#pragma once

#include <hip/hip_runtime.h>

#include <cstring>
#include <functional>
#include <utility>
#include <vector>

namespace mdx {

// Plain (trivially copyable) argument pack: one leaf per argument, so the
// host can copy packs into a device table byte for byte.
template <size_t I, class T>
struct ArgLeaf {
    T v;
};
template <class Seq, class... A>
struct ArgPackImpl;
template <size_t... I, class... A>
struct ArgPackImpl<std::index_sequence<I...>, A...> : ArgLeaf<I, A>... {};
template <class... A>
using ArgPack = ArgPackImpl<std::index_sequence_for<A...>, A...>;

template <auto BODY, size_t... I, class... A>
__device__ __forceinline__ void call_body(int blk, int nblk, const ArgPackImpl<std::index_sequence<I...>, A...> &p) {
    BODY(blk, nblk, static_cast<const ArgLeaf<I, A> &>(p).v...);
}

template <size_t... I, class... A>
inline void pack_fill(ArgPackImpl<std::index_sequence<I...>, A...> &p, const A &...a) {
    ((static_cast<ArgLeaf<I, A> &>(p).v = a), ...);
}

template <auto BODY, int NT, class... A>
__global__ void __launch_bounds__(NT) k_one(A... a) {
    BODY((int)blockIdx.x, (int)gridDim.x, a...);
}

// map[b] = (context, workgroup within its grid, its grid size, 0)
template <auto BODY, int NT, class... A>
__global__ void __launch_bounds__(NT) k_many(const ArgPack<A...> *__restrict__ tab, const int4 *__restrict__ map) {
    const int4 m = map[blockIdx.x];
    const ArgPack<A...> p = tab[m.x];
    call_body<BODY>(m.y, m.z, p);
}

template <auto BODY, int NT, class... A>
void launch_many(hipStream_t s, const void *tab, const int4 *map, int total, size_t shm) {
    hipLaunchKernelGGL((k_many<BODY, NT, A...>), dim3(total), dim3(NT), shm, s,
                       static_cast<const ArgPack<A...> *>(tab), map);
}

// One recorded step of a context while a batch is being put together: a
// kernel of this layer (identity, grid, dynamic LDS, packed arguments, the
// batched launcher) or another operation on the stream (an FFT).
struct LaunchRec {
    const void *kid = nullptr;
    int blocks = 0;
    size_t shm = 0;
    std::vector<unsigned char> args;
    size_t align = 1;
    void (*many)(hipStream_t, const void *, const int4 *, int, size_t) = nullptr;
    std::function<void(hipStream_t)> host;
    int fft = 0;  // host operation that is a P3M transform (1 forward, 2 inverse)
};

}  // namespace mdx
