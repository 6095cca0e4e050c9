// ntt_launch.hpp — host launchers of the pass kernel.  The explicit
// instantiations live in pass_*.hip (one translation unit per storage type,
// compiled in parallel); ntt_lib.hip only declares them.
#pragma once
#include <hip/hip_runtime.h>

#include <mutex>

#include "ntt_kernels.hpp"

namespace nttk {

constexpr int kPassThreads = (int)kNT;
constexpr uint32_t kPassLdsMax = 81920;  // LDS bytes a pass workgroup may use

// Launches pass_kernel<M, Sin, Sout, 4, kPassThreads, P4, LR, LC, MF> with one
// workgroup per tile (LR = 0: the generic kernel; MF: matrix-core rounds).
template <class M, class Sin, class Sout, bool P4, int LR = 0, int LC = 0, bool MF = false, bool IP = false>
hipError_t launch_pass_kernel(const PassJob* djobs, uint32_t njobs, const PassJob& inl, uint32_t ntiles, uint32_t lds,
                              hipStream_t stream) {
    auto fn = pass_kernel<M, Sin, Sout, 4, kPassThreads, P4, LR, LC, MF, IP>;
    static std::once_flag once;
    static hipError_t ready = hipSuccess;
    std::call_once(once, [&] {
        (void)hipFuncSetAttribute((const void*)fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)kPassLdsMax);
        // The kernel addresses its dynamic LDS from address 0 (lds_ld/lds_st),
        // which holds only while it has no static LDS.
        hipFuncAttributes at{};
        ready = hipFuncGetAttributes(&at, (const void*)fn);
        if (ready == hipSuccess && at.sharedSizeBytes != 0) ready = hipErrorInvalidKernelFile;
    });
    if (ready != hipSuccess) return ready;
    hipLaunchKernelGGL(fn, dim3(ntiles), dim3(kPassThreads), lds, stream, djobs, njobs, inl);
    return hipGetLastError();
}

// Launches queue_kernel<M, Sin, Sout, 4, kPassThreads, P4, MF> (one launch
// for every pass of a batch; ntt_kernels.hpp) on nblocks workgroups.
template <class M, class Sin, class Sout, bool P4, bool MF>
hipError_t launch_queue_kernel(const PassJob* djobs, const uint32_t* tjob, const uint32_t* qbeg, uint32_t nq,
                               unsigned long long* ctl, uint32_t tslot, uint32_t spin_limit, uint32_t proto,
                               uint32_t nblocks, uint32_t lds, hipStream_t stream) {
    auto fn = queue_kernel<M, Sin, Sout, 4, kPassThreads, P4, MF>;
    static std::once_flag once;
    static hipError_t ready = hipSuccess;
    std::call_once(once, [&] {
        (void)hipFuncSetAttribute((const void*)fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)kPassLdsMax);
        hipFuncAttributes at{};
        ready = hipFuncGetAttributes(&at, (const void*)fn);
        if (ready == hipSuccess && at.sharedSizeBytes != 0) ready = hipErrorInvalidKernelFile;
    });
    if (ready != hipSuccess) return ready;
    hipLaunchKernelGGL(fn, dim3(nblocks), dim3(kPassThreads), lds, stream, djobs, tjob, qbeg, nq, ctl, tslot,
                       spin_limit, proto);
    return hipGetLastError();
}

// Launches pipe_kernel<M, Sin, Sout, 4, kPassThreads, P4, LR, LC, MF> (one
// job, several tiles per workgroup with the next tile's loads in flight).
template <class M, class Sin, class Sout, bool P4, int LR, int LC, bool MF>
hipError_t launch_pipe_kernel(const PassJob& J, uint32_t nblocks, uint32_t lds, hipStream_t stream) {
    auto fn = pipe_kernel<M, Sin, Sout, 4, kPassThreads, P4, LR, LC, MF>;
    static std::once_flag once;
    static hipError_t ready = hipSuccess;
    std::call_once(once, [&] {
        (void)hipFuncSetAttribute((const void*)fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)kPassLdsMax);
        hipFuncAttributes at{};
        ready = hipFuncGetAttributes(&at, (const void*)fn);
        if (ready == hipSuccess && at.sharedSizeBytes != 0) ready = hipErrorInvalidKernelFile;
    });
    if (ready != hipSuccess) return ready;
    hipLaunchKernelGGL(fn, dim3(nblocks), dim3(kPassThreads), lds, stream, J);
    return hipGetLastError();
}

// Workgroups of pipe_kernel<...> that fit on one CU with `lds` bytes of LDS.
template <class M, class Sin, class Sout, bool P4, int LR, int LC, bool MF>
int pipe_kernel_per_cu(uint32_t lds) {
    auto fn = pipe_kernel<M, Sin, Sout, 4, kPassThreads, P4, LR, LC, MF>;
    (void)hipFuncSetAttribute((const void*)fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)kPassLdsMax);
    int nb = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, fn, kPassThreads, lds) != hipSuccess) return 1;
    return nb > 0 ? nb : 1;
}

// Tile shapes with a pipelined instance (32-bit coefficients, matrix-core
// rounds): tiles of 2^12 and 2^13 words (A: the [8, 8, 8]-style plans and the
// [10, 10] two-pass plan) and of 2^14 words (B: [11, 9], [10, 10]), one
// translation unit each.
#define NTT_PIPE_SHAPES_A(X) X(8, 4) X(8, 5) X(10, 3)
#define NTT_PIPE_SHAPES_B(X) X(11, 3) X(9, 5) X(10, 4)
#define NTT_PIPE_SHAPES(X) NTT_PIPE_SHAPES_A(X) NTT_PIPE_SHAPES_B(X)
#define NTT_PIPE_INSTANCE_1(EXT, LR, LC, P4)                                                                   \
    EXT template hipError_t launch_pipe_kernel<Mod32, uint32_t, uint32_t, P4, LR, LC, true>(                    \
        const PassJob&, uint32_t, uint32_t, hipStream_t);                                                       \
    EXT template int pipe_kernel_per_cu<Mod32, uint32_t, uint32_t, P4, LR, LC, true>(uint32_t);
#define NTT_PIPE_INSTANCE_EXTERN(LR, LC) NTT_PIPE_INSTANCE_1(extern, LR, LC, false) NTT_PIPE_INSTANCE_1(extern, LR, LC, true)
#define NTT_PIPE_INSTANCE(LR, LC) NTT_PIPE_INSTANCE_1(, LR, LC, false) NTT_PIPE_INSTANCE_1(, LR, LC, true)

// Workgroups of queue_kernel<...> that fit on one CU with `lds` bytes of LDS.
template <class M, class Sin, class Sout, bool P4, bool MF>
int queue_kernel_per_cu(uint32_t lds) {
    auto fn = queue_kernel<M, Sin, Sout, 4, kPassThreads, P4, MF>;
    (void)hipFuncSetAttribute((const void*)fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)kPassLdsMax);
    int nb = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, fn, kPassThreads, lds) != hipSuccess) return 1;
    return nb > 0 ? nb : 1;
}

#define NTT_QUEUE_INSTANCE_1(EXT, MF)                                                                          \
    EXT template hipError_t launch_queue_kernel<Mod32, uint32_t, uint32_t, true, MF>(                           \
        const PassJob*, const uint32_t*, const uint32_t*, uint32_t, unsigned long long*, uint32_t, uint32_t,    \
        uint32_t, uint32_t, uint32_t, hipStream_t);                                                              \
    EXT template int queue_kernel_per_cu<Mod32, uint32_t, uint32_t, true, MF>(uint32_t);
#define NTT_QUEUE_INSTANCES(EXT) NTT_QUEUE_INSTANCE_1(EXT, false) NTT_QUEUE_INSTANCE_1(EXT, true)

#define NTT_PASS_INSTANCE_1(EXT, M, SIN, SOUT, P4, MF)                                                         \
    EXT template hipError_t launch_pass_kernel<M, SIN, SOUT, P4, 0, 0, MF>(const PassJob*, uint32_t,           \
                                                                          const PassJob&, uint32_t, uint32_t,  \
                                                                          hipStream_t);
// VALU-round kernels; Mod32 also has the matrix-core-round (MF) kernels.
#define NTT_PASS_INSTANCES(EXT, M, SIN, SOUT)                                                                  \
    NTT_PASS_INSTANCE_1(EXT, M, SIN, SOUT, false, false) NTT_PASS_INSTANCE_1(EXT, M, SIN, SOUT, true, false)
#define NTT_PASS_INSTANCES_MF(EXT, SIN, SOUT)                                                                  \
    NTT_PASS_INSTANCES(EXT, Mod32, SIN, SOUT)                                                                  \
    NTT_PASS_INSTANCE_1(EXT, Mod32, SIN, SOUT, false, true) NTT_PASS_INSTANCE_1(EXT, Mod32, SIN, SOUT, true, true)

// Tile shapes (log2 R, log2 C) with a specialised kernel for 32-bit
// coefficients: radix 2^6..2^12 at tiles of 2^12, 2^13 and 2^14 elements
// (at least 4 columns, for vector loads).
#define NTT_SHAPES_A(X) X(6, 7) X(7, 6) X(8, 5)
#define NTT_SHAPES_B(X) X(9, 4) X(10, 3) X(11, 2)
#define NTT_SHAPES_C(X) X(8, 6) X(9, 5) X(10, 4)
#define NTT_SHAPES_D(X) X(11, 3) X(12, 2)
#define NTT_SHAPES_E(X) X(6, 6) X(7, 5) X(8, 4)
#define NTT_SHAPES_F(X) X(9, 3) X(10, 2)
#define NTT_SHAPES(X) \
    NTT_SHAPES_A(X) NTT_SHAPES_B(X) NTT_SHAPES_C(X) NTT_SHAPES_D(X) NTT_SHAPES_E(X) NTT_SHAPES_F(X)

#define NTT_SHAPE_INSTANCE_1(EXT, LR, LC, P4, MF)                                                               \
    EXT template hipError_t launch_pass_kernel<Mod32, uint32_t, uint32_t, P4, LR, LC, MF>(                      \
        const PassJob*, uint32_t, const PassJob&, uint32_t, uint32_t, hipStream_t);
#define NTT_SHAPE_INSTANCE_ALL(EXT, LR, LC)                                                                     \
    NTT_SHAPE_INSTANCE_1(EXT, LR, LC, false, false) NTT_SHAPE_INSTANCE_1(EXT, LR, LC, true, false)              \
    NTT_SHAPE_INSTANCE_1(EXT, LR, LC, false, true) NTT_SHAPE_INSTANCE_1(EXT, LR, LC, true, true)
#define NTT_SHAPE_INSTANCE_EXTERN(LR, LC) NTT_SHAPE_INSTANCE_ALL(extern, LR, LC)
#define NTT_SHAPE_INSTANCE(LR, LC) NTT_SHAPE_INSTANCE_ALL(, LR, LC)

// Shapes with an instance for first passes whose load forms a product or an
// inner product (32-bit storage, matrix-core rounds, four-step post): the
// tiles of the two-pass plans of 2^14..2^18 (pass_ip.hip).
#define NTT_IP_SHAPES(X) X(7, 5) X(8, 4) X(8, 5) X(9, 4) X(10, 3)
#define NTT_IP_INSTANCE_1(EXT, LR, LC)                                                                          \
    EXT template hipError_t launch_pass_kernel<Mod32, uint32_t, uint32_t, true, LR, LC, true, true>(            \
        const PassJob*, uint32_t, const PassJob&, uint32_t, uint32_t, hipStream_t);
#define NTT_IP_INSTANCE_EXTERN(LR, LC) NTT_IP_INSTANCE_1(extern, LR, LC)
#define NTT_IP_INSTANCE(LR, LC) NTT_IP_INSTANCE_1(, LR, LC)

// Shapes with matrix-core-round instances for mixed storage: the first pass
// of a transform on uint64 data (uint64 in, 32-bit out, with post) and its
// last pass (32-bit in, uint64 out, with or without post).  These are the
// shapes the tuned plans of single transforms use; other shapes run the
// generic kernel.
#define NTT_SHAPES_U64_A(X) X(8, 5) X(8, 4) X(7, 5) X(9, 4) X(7, 6) X(10, 4)
#define NTT_SHAPES_U64_B(X) X(10, 2) X(10, 3) X(9, 3) X(11, 2) X(11, 3) X(12, 2)
#define NTT_SHAPES_U64(X) NTT_SHAPES_U64_A(X) NTT_SHAPES_U64_B(X)
#define NTT_SHAPE_U64_INSTANCE_ALL(EXT, LR, LC)                                                                 \
    EXT template hipError_t launch_pass_kernel<Mod32, uint64_t, uint32_t, true, LR, LC, true>(                  \
        const PassJob*, uint32_t, const PassJob&, uint32_t, uint32_t, hipStream_t);                             \
    EXT template hipError_t launch_pass_kernel<Mod32, uint32_t, uint64_t, true, LR, LC, true>(                  \
        const PassJob*, uint32_t, const PassJob&, uint32_t, uint32_t, hipStream_t);                             \
    EXT template hipError_t launch_pass_kernel<Mod32, uint32_t, uint64_t, false, LR, LC, true>(                 \
        const PassJob*, uint32_t, const PassJob&, uint32_t, uint32_t, hipStream_t);
#define NTT_SHAPE_U64_INSTANCE_EXTERN(LR, LC) NTT_SHAPE_U64_INSTANCE_ALL(extern, LR, LC)
#define NTT_SHAPE_U64_INSTANCE(LR, LC) NTT_SHAPE_U64_INSTANCE_ALL(, LR, LC)

// Instantiated in pass_*.hip.
NTT_PASS_INSTANCES_MF(extern, uint32_t, uint32_t)
NTT_PASS_INSTANCES_MF(extern, uint32_t, uint64_t)
NTT_PASS_INSTANCES_MF(extern, uint64_t, uint32_t)
NTT_PASS_INSTANCES_MF(extern, uint64_t, uint64_t)
NTT_PASS_INSTANCES(extern, Mod64, uint64_t, uint64_t)
NTT_PASS_INSTANCES(extern, ModGold, uint64_t, uint64_t)
NTT_SHAPES(NTT_SHAPE_INSTANCE_EXTERN)
NTT_SHAPES_U64(NTT_SHAPE_U64_INSTANCE_EXTERN)
NTT_IP_SHAPES(NTT_IP_INSTANCE_EXTERN)
NTT_QUEUE_INSTANCES(extern)
NTT_PIPE_SHAPES(NTT_PIPE_INSTANCE_EXTERN)

}  // namespace nttk
